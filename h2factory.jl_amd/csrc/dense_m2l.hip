// Dense (and symmetry-reduced dense) M2L on MFMA.
//
// For one level and receiving box t the coupling stage is
//     loc[t][r] = sum_{(s, e) in list(t)} A_e mom[s][r]
// with A_e the K x K coupling of the pair (src/farinteractions/i2otranslations/
// i2otranslationcollection.jl:367-409: out (+)= op(M[couplingkey]) in), one product per partner.
// Complex maps run the same kernel on the real form: a complex moment of K coefficients is stored
// interleaved (re, im), i.e. as a real vector of Kr = 2K values, and the complex K x K coupling
// becomes the 2K x 2K real matrix of multiplication by it ([[Re, -Im], [Im, Re]] per entry).
//
// Operators.  At core creation every distinct (translation id, symmetry) of the pairs is expanded
// on the device into a "fragment" matrix in MFMA operand order, one per orientation (N; T; and C
// for complex maps): the symmetry-reduced kinds keep 16 canonical matrices per level in the
// reference (symmetryreduceddense.jl:190-220: gather in[pi_c], multiply, scatter out[pi_r]);
// here the gather  A_e[k][j] = C[pi^-1(k)][pi^-1(j)]  runs once in the build kernel, so the
// coupling stage reads one plain matrix per pair.  The fragment order makes every operand load a
// 16-byte load: for the 16 x 16 x 4 MFMAs of row tile rt and k group g (16 reduction indices)
//     frag[e][rt][g][lane][kk] = A_e[16 rt + lane % 16][16 g + 4 (lane / 16) + kk],   kk = 0..3,
// so lane (m + 16 k') holds A[m][16 g + 4 k' + kk] for the kk-th MFMA of the group; the moments
// are read the same way ([box][r][k], k contiguous: 4 consecutive k per 16-byte load) and each
// group of 4 MFMAs reduces 16 k.  Rows and columns beyond Kr are zero.
//
// Kernel.  One workgroup (4 waves) per (receiving box, block of 16 or 32 columns); wave w owns the
// row tiles w, w + 4, ... and all column tiles, and runs every partner of the box with no
// barrier: per k group it loads its A fragments (L2 / MALL: the level's operators) and the
// partner's moment columns (L1: the 4 waves read the same block), then issues the MFMAs.
//   * FP64 (plain form): v_mfma_f64_16x16x4f64, FP64 accumulators;
//   * FP32 (mixed form): v_mfma_f32_16x16x4f32 on FP32 operands; every 4 k groups (64 products
//     per accumulator) the FP32 partial sums are added into FP64 accumulators, so the error stays
//     at the FP32 rounding of the operands instead of growing with the ~190 partners x K terms;
//   * split BF16 (precisions BF16X2 / BF16X3): the operands as 2 or 3 BF16 planes on
//     v_mfma_f32_16x16x32_bf16 (3 or 6 plane products), FP64 accumulation as for FP32.
#include <algorithm>
#include <cstdint>
#include <map>
#include <vector>

#include "h2g_internal.h"

namespace h2g {

namespace {

typedef double double4_d __attribute__((ext_vector_type(4)));
typedef float float4_d __attribute__((ext_vector_type(4)));

constexpr int DNT = 256;  // threads per workgroup (4 waves)

// Element (i, j) of the real-form operator e for orientation `orient` (0: A, 1: A^T, 2: A^H,
// 3: conj(A)), from
// the canonical couplings `ops` (core type, column-major K x K, doubles; complex interleaved):
// A_e[p][q] = C[pi^-1(p)][pi^-1(q)] for the symmetry of e (none: the identity).
template <bool CPLX>
__device__ inline double dense_value(const double *__restrict__ ops, int K, int Kr, int64_t e, int i, int j,
                                     const int32_t *__restrict__ fop_op, const int32_t *__restrict__ fop_sym,
                                     const int32_t *__restrict__ perms_inv, int orient) {
    if (i >= Kr || j >= Kr) return 0.0;
    const int ii = CPLX ? i >> 1 : i, jj = CPLX ? j >> 1 : j;
    const bool tr = orient == 1 || orient == 2;
    const int p = tr ? jj : ii, q = tr ? ii : jj;
    const int sym = fop_sym ? fop_sym[e] : -1;
    const int pp = sym >= 0 ? perms_inv[(int64_t)sym * K + p] : p;
    const int qq = sym >= 0 ? perms_inv[(int64_t)sym * K + q] : q;
    const int64_t at = (int64_t)fop_op[e] * K * K + pp + (int64_t)K * qq;
    if (!CPLX) return ops[at];
    const double re = ops[2 * at], im = orient >= 2 ? -ops[2 * at + 1] : ops[2 * at + 1];
    const int a = i & 1, b = j & 1;
    return a == b ? re : (a == 0 ? -im : im);
}

// Fragment matrices for the 16 x 16 x 4 MFMAs (FP64 / FP32): [e][rt][g][lane][kk].
template <class OUT, bool CPLX>
__global__ void dense_frag_kernel(const double *__restrict__ ops, int K, int Kr, int KrT, int64_t n_fops,
                                  const int32_t *__restrict__ fop_op, const int32_t *__restrict__ fop_sym,
                                  const int32_t *__restrict__ perms_inv, int orient, OUT *__restrict__ frag) {
    const int64_t per = (int64_t)(16 * KrT) * (16 * KrT);
    const int64_t total = n_fops * per;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / per;
        const int rem = (int)(idx - e * per);
        const int kk = rem & 3, lane = (rem >> 2) & 63, tg = rem >> 8;
        const int rt = tg / KrT, g = tg - rt * KrT;
        const int i = 16 * rt + (lane & 15), j = 16 * g + 4 * (lane >> 4) + kk;
        frag[idx] = (OUT)dense_value<CPLX>(ops, K, Kr, e, i, j, fop_op, fop_sym, perms_inv, orient);
    }
}

// Complex fragments (complex maps, FP64 and FP32): the real form repeats every complex entry
// twice (rows 2t and 2t + 1 of the real form read the same Re, Im), so the operator is stored
// once, as complex values in the MFMA order of the real form's row pairs:
//     cfrag[e][rt][g][tp + 8 kq][0..3] = (Re, Im) of M[t][s0], M[t][s1],
//     t = 8 rt + tp (tp < 8), s0 = 8 g + 2 kq, s1 = s0 + 1,
// half the bytes of the real-form fragments.  Lane (m + 16 kq) of the kernel loads entry
// (m / 2 + 8 kq) -- lanes m and m ^ 1 the same 16 or 32 bytes -- and forms its four real-form
// values A[16 rt + m][16 g + 4 kq + kk]: (Re0, -Im0, Re1, -Im1) for even m, (Im0, Re0, Im1, Re1)
// for odd m.
template <class OUT>
__global__ void dense_cfrag_kernel(const double *__restrict__ ops, int K, int Kr, int KrT, int64_t n_fops,
                                   const int32_t *__restrict__ fop_op, const int32_t *__restrict__ fop_sym,
                                   const int32_t *__restrict__ perms_inv, int orient, OUT *__restrict__ frag) {
    const int64_t per = (int64_t)KrT * KrT * 128;
    const int64_t total = n_fops * per;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / per;
        const int rem = (int)(idx - e * per);
        const int v = rem & 3, slot = (rem >> 2) & 31, tg = rem >> 7;
        const int rt = tg / KrT, g = tg - rt * KrT;
        const int t = 8 * rt + (slot & 7), s = 8 * g + 2 * (slot >> 3) + (v >> 1);
        // real-form rows 2t (a = 0) and 2t + 1 (a = 1), column 2s (b = 0): Re and Im
        frag[idx] = (OUT)dense_value<true>(ops, K, Kr, e, 2 * t + (v & 1), 2 * s, fop_op, fop_sym, perms_inv, orient);
    }
}

// Fragment matrices for the 16 x 16 x 32 BF16 MFMAs: three planes (hi, mid, lo) of
// [e][rt][g][lane][8] with G2 = ceil(KrT / 2) groups of 32 reductions; lane (m + 16 k'') holds
// A[16 rt + m][32 g + 8 k'' + u], u = 0..7.  The FP32 value x is split as hi = bf16(x),
// mid = bf16(x - hi), lo = bf16(x - hi - mid) (each difference exact in FP32).
template <bool CPLX>
__global__ void dense_frag_bf16_kernel(const double *__restrict__ ops, int K, int Kr, int KrT, int64_t n_fops,
                                       const int32_t *__restrict__ fop_op, const int32_t *__restrict__ fop_sym,
                                       const int32_t *__restrict__ perms_inv, int orient, __bf16 *__restrict__ frag) {
    const int G2 = (KrT + 1) / 2;
    const int64_t per = (int64_t)KrT * G2 * 512;
    const int64_t total = n_fops * per;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / per;
        const int rem = (int)(idx - e * per);
        const int u = rem & 7, lane = (rem >> 3) & 63, tg = rem >> 9;
        const int rt = tg / G2, g = tg - rt * G2;
        const int i = 16 * rt + (lane & 15), j = 32 * g + 8 * (lane >> 4) + u;
        const float x = (float)dense_value<CPLX>(ops, K, Kr, e, i, j, fop_op, fop_sym, perms_inv, orient);
        const __bf16 hi = (__bf16)x;
        const float r1 = x - (float)hi;
        const __bf16 mid = (__bf16)r1;
        const __bf16 lo = (__bf16)(r1 - (float)mid);
        frag[idx] = hi;
        frag[total + idx] = mid;
        frag[2 * total + idx] = lo;
    }
}

template <class V> struct dvec;
template <> struct dvec<double> { using type = double4_d; };
template <> struct dvec<float> { using type = float4_d; };

// 4 consecutive k of column vector `col` (Kr values), zero beyond Kr
template <class V>
__device__ inline typename dvec<V>::type load_k4(const V *__restrict__ col, int k0, int Kr, bool valid) {
    using V4 = typename dvec<V>::type;
    V4 v = V4{0, 0, 0, 0};
    if (!valid) return v;
    if ((Kr & 3) == 0 && k0 + 4 <= Kr) return *reinterpret_cast<const V4 *>(col + k0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (k0 + u < Kr) v[u] = col[k0 + u];
    return v;
}

// Workgroup b of nb runs on XCD b % 8 (the dispatcher's round robin); this maps the workgroups of
// each XCD to a contiguous range of [0, nb), so that boxes adjacent in the receiving list (Morton
// order: neighbours, which share most of their operators) run on one XCD and share its L2
// (Laplace dense, N = 2M, p = 6, 32 columns: coupling stage 416 -> 389 ms).  A bijection for any nb.
__device__ inline int xcd_order(int b, int nb) {
    const int per = nb >> 3, rem = nb & 7, x = b & 7, k = b >> 3;
    return x * per + min(x, rem) + k;
}

// Accumulators started from the current output (the accumulate form, e.g. L2L into a box that
// already holds couplings), in the D layout of the 16 x 16 MFMAs: FP32 -- D[m][n] in lane
// n + 16 (m / 4), element m % 4; FP64 -- D[m][n] in lane n + 16 (m % 4), element m / 4.
template <bool F64LAYOUT, int RTW, int NCT, class V>
__device__ inline void init_from_out(double (&acc)[RTW][NCT][4], const V *__restrict__ out, int t, int R, int Kr,
                                     int KrT, int r0, int wave, int ntw, int n, int kq) {
#pragma unroll
    for (int j = 0; j < RTW; ++j) {
        if (j >= ntw) continue;
        const int rt = wave + 4 * j;
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int col = r0 + 16 * c + n;
            if (col >= R) continue;
            const V *o = out + ((int64_t)t * R + col) * Kr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * rt + (F64LAYOUT ? kq + 4 * v : 4 * kq + v);
                if (row < Kr) acc[j][c][v] = (double)o[row];
            }
        }
    }
}

// MODE 0: FP64 operands on FP64 MFMA; MODE 1: FP32 operands on FP32 MFMA with FP64 accumulation
// of the partial sums.  RTW: row tiles per wave (16 rows each), NCT: column tiles (16 columns).
template <int MODE, int RTW, int NCT, bool CF>
__global__ void __launch_bounds__(DNT, 2)
    dense_m2l_kernel(const int32_t *__restrict__ recv, const int64_t *__restrict__ item_ptr,
                     const int32_t *__restrict__ item_src, const int32_t *__restrict__ item_fop,
                     const typename std::conditional<MODE == 0, double, float>::type *__restrict__ frag, int Kr,
                     int KrT, const typename std::conditional<MODE == 0, double, float>::type *__restrict__ mom,
                     typename std::conditional<MODE == 0, double, float>::type *__restrict__ loc, int R,
                     const uint8_t *__restrict__ acc_flags) {
    using V = typename std::conditional<MODE == 0, double, float>::type;
    using V4 = typename dvec<V>::type;
    const int ti = xcd_order(blockIdx.x, gridDim.x);
    const int t = recv[ti];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kq = lane >> 4;
    const int r0 = blockIdx.y * 16 * NCT;
    const int G = KrT;
    const int ntw = (KrT - wave + 3) >> 2;  // row tiles of this wave: wave, wave + 4, ...
    // real-form fragments: (16 KrT)^2 values per operator; complex fragments: half of that
    const int64_t fsz = CF ? (int64_t)KrT * KrT * 128 : (int64_t)(16 * KrT) * (16 * KrT);
    double acc[RTW][NCT][4];
#pragma unroll
    for (int j = 0; j < RTW; ++j)
#pragma unroll
        for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[j][c][v] = 0.0;
    V4 part[RTW][NCT];
#pragma unroll
    for (int j = 0; j < RTW; ++j)
#pragma unroll
        for (int c = 0; c < NCT; ++c) part[j][c] = V4{0, 0, 0, 0};
    const int64_t i0 = item_ptr[ti], i1 = item_ptr[ti + 1];
    if (acc_flags && i1 > i0 && acc_flags[i0])
        init_from_out<MODE == 0>(acc, loc, t, R, Kr, KrT, r0, wave, ntw, n, kq);
    for (int64_t it = i0; it < i1; ++it) {
        const V *A0 = frag + (int64_t)item_fop[it] * fsz + (CF ? (int64_t)((n >> 1) + 8 * kq) : (int64_t)lane) * 4;
        const int64_t s = item_src[it];
        const V *Bc[NCT];
        bool bval[NCT];
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int col = r0 + 16 * c + n;
            bval[c] = col < R;
            Bc[c] = mom + (s * R + (bval[c] ? col : 0)) * (int64_t)Kr;
        }
        for (int g = 0; g < G; ++g) {
            V4 a[RTW], b[NCT];
#pragma unroll
            for (int j = 0; j < RTW; ++j) {
                if constexpr (CF) {
                    const V4 c = j < ntw ? *reinterpret_cast<const V4 *>(A0 + ((int64_t)(wave + 4 * j) * G + g) * 128)
                                         : V4{0, 0, 0, 0};
                    a[j] = (n & 1) ? V4{c[1], c[0], c[3], c[2]} : V4{c[0], -c[1], c[2], -c[3]};
                } else {
                    a[j] = j < ntw ? *reinterpret_cast<const V4 *>(A0 + ((int64_t)(wave + 4 * j) * G + g) * 256)
                                   : V4{0, 0, 0, 0};
                }
            }
#pragma unroll
            for (int c = 0; c < NCT; ++c) b[c] = load_k4<V>(Bc[c], 16 * g + 4 * kq, Kr, bval[c]);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int j = 0; j < RTW; ++j) {
                    if (j >= ntw) continue;
#pragma unroll
                    for (int c = 0; c < NCT; ++c) {
                        if (MODE == 0) {
                            double4_d d = {acc[j][c][0], acc[j][c][1], acc[j][c][2], acc[j][c][3]};
                            d = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[j][kk], (double)b[c][kk], d, 0, 0, 0);
#pragma unroll
                            for (int v = 0; v < 4; ++v) acc[j][c][v] = d[v];
                        } else {
                            float4_d d = {(float)part[j][c][0], (float)part[j][c][1], (float)part[j][c][2],
                                          (float)part[j][c][3]};
                            d = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a[j][kk], (float)b[c][kk], d, 0, 0, 0);
#pragma unroll
                            for (int v = 0; v < 4; ++v) part[j][c][v] = d[v];
                        }
                    }
                }
            if (MODE == 1 && ((g & 3) == 3 || g == G - 1)) {
#pragma unroll
                for (int j = 0; j < RTW; ++j)
#pragma unroll
                    for (int c = 0; c < NCT; ++c) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[j][c][v] += (double)part[j][c][v];
                        part[j][c] = V4{0, 0, 0, 0};
                    }
            }
        }
    }
    // D layouts: FP32 16x16x4 -- D[m][n] in lane n + 16 (m / 4), element m % 4;
    //            FP64 16x16x4 -- D[m][n] in lane n + 16 (m % 4), element m / 4
#pragma unroll
    for (int j = 0; j < RTW; ++j) {
        if (j >= ntw) continue;
        const int rt = wave + 4 * j;
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int col = r0 + 16 * c + n;
            if (col >= R) continue;
            V *o = loc + ((int64_t)t * R + col) * Kr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * rt + (MODE == 0 ? kq + 4 * v : 4 * kq + v);
                if (row < Kr) o[row] = (V)acc[j][c][v];
            }
        }
    }
}

// Two receiving boxes per workgroup (M2L, real operators): the pair's items are merged by operator,
// so an operator that both receivers apply (to different sources) is read once and feeds both
// receivers' MFMAs.  Item = (operator, source of receiver 0, source of receiver 1), -1 where a
// receiver has no partner with that operator (its MFMAs are skipped: the source index is
// wave-uniform).  NW waves share the row tiles (wave, wave + NW, ...), each holding RTW row tiles x
// 2 receivers of 16-column accumulators.  Workgroup index i (in XCD order) = pair i / nchunk,
// columns 16 (i % nchunk) .. + 15: the chunks of one pair run on one XCD and share its operators
// in that XCD's L2.
template <int MODE, int NW, int RTW>
__global__ void __launch_bounds__(64 * NW, 1)
    dense_m2l_pair_kernel(const int32_t *__restrict__ recv2, const int64_t *__restrict__ item_ptr,
                          const int2 *__restrict__ item_src, const int32_t *__restrict__ item_fop,
                          const typename std::conditional<MODE == 0, double, float>::type *__restrict__ frag,
                          int Kr, int KrT,
                          const typename std::conditional<MODE == 0, double, float>::type *__restrict__ mom,
                          typename std::conditional<MODE == 0, double, float>::type *__restrict__ loc, int R,
                          int nchunk) {
    using V = typename std::conditional<MODE == 0, double, float>::type;
    using V4 = typename dvec<V>::type;
    const int bi = xcd_order(blockIdx.x, gridDim.x);
    const int ti = bi / nchunk;
    const int r0 = (bi - ti * nchunk) * 16;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n = lane & 15, kq = lane >> 4;
    const int G = KrT;
    const int ntw = (KrT - wave + NW - 1) / NW;  // row tiles of this wave: wave, wave + NW, ...
    const int64_t fsz = (int64_t)(16 * KrT) * (16 * KrT);
    const int col = r0 + n;
    double acc[2][RTW][4];
    V4 part[2][RTW];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int j = 0; j < RTW; ++j) {
            part[u][j] = V4{0, 0, 0, 0};
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][j][v] = 0.0;
        }
    const int64_t i0 = item_ptr[ti], i1 = item_ptr[ti + 1];
    for (int64_t it = i0; it < i1; ++it) {
        const V *A0 = frag + (int64_t)item_fop[it] * fsz + (int64_t)lane * 4;
        const int2 sp = item_src[it];
        const int src[2] = {__builtin_amdgcn_readfirstlane(sp.x), __builtin_amdgcn_readfirstlane(sp.y)};
        const V *Bc[2];
        bool bval[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            bval[u] = src[u] >= 0 && col < R;
            Bc[u] = mom + ((int64_t)(src[u] >= 0 ? src[u] : 0) * R + (col < R ? col : 0)) * (int64_t)Kr;
        }
        for (int g = 0; g < G; ++g) {
            V4 a[RTW], b[2];
#pragma unroll
            for (int j = 0; j < RTW; ++j)
                a[j] = j < ntw ? *reinterpret_cast<const V4 *>(A0 + ((int64_t)(wave + NW * j) * G + g) * 256)
                               : V4{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < 2; ++u) b[u] = load_k4<V>(Bc[u], 16 * g + 4 * kq, Kr, bval[u]);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (src[u] < 0) continue;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                    for (int j = 0; j < RTW; ++j) {
                        if (j >= ntw) continue;
                        if (MODE == 0) {
                            double4_d d = {acc[u][j][0], acc[u][j][1], acc[u][j][2], acc[u][j][3]};
                            d = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[j][kk], (double)b[u][kk], d, 0, 0, 0);
#pragma unroll
                            for (int v = 0; v < 4; ++v) acc[u][j][v] = d[v];
                        } else {
                            float4_d d = {(float)part[u][j][0], (float)part[u][j][1], (float)part[u][j][2],
                                          (float)part[u][j][3]};
                            d = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a[j][kk], (float)b[u][kk], d, 0, 0, 0);
#pragma unroll
                            for (int v = 0; v < 4; ++v) part[u][j][v] = d[v];
                        }
                    }
            }
            if (MODE == 1 && ((g & 3) == 3 || g == G - 1)) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int j = 0; j < RTW; ++j) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][j][v] += (double)part[u][j][v];
                        part[u][j] = V4{0, 0, 0, 0};
                    }
            }
        }
    }
    // D layouts as in dense_m2l_kernel
    if (col >= R) return;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t = recv2[2 * ti + u];
        if (t < 0) continue;
        V *o = loc + ((int64_t)t * R + col) * Kr;
#pragma unroll
        for (int j = 0; j < RTW; ++j) {
            if (j >= ntw) continue;
            const int rt = wave + NW * j;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * rt + (MODE == 0 ? kq + 4 * v : 4 * kq + v);
                if (row < Kr) o[row] = (V)acc[u][j][v];
            }
        }
    }
}

typedef __bf16 bf16x8_d __attribute__((ext_vector_type(8)));

// 8 consecutive k of an FP32 column (zero beyond Kr), split into NPL BF16 planes
template <int NPL>
__device__ inline void load_split8(const float *__restrict__ col, int k0, int Kr, bool valid, bf16x8_d (&out)[NPL]) {
    float x[8];
    if (valid && (Kr & 3) == 0 && k0 + 8 <= Kr) {
        const float4_d v0 = *reinterpret_cast<const float4_d *>(col + k0);
        const float4_d v1 = *reinterpret_cast<const float4_d *>(col + k0 + 4);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] = v0[u];
            x[u + 4] = v1[u];
        }
    } else {
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = valid && k0 + u < Kr ? col[k0 + u] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        float r = x[u];
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
            const __bf16 h = (__bf16)r;
            out[pl][u] = h;
            r -= (float)h;
        }
    }
}

// Split BF16 form of the mixed precision: A and B split into NPL BF16 planes; the products of
// plane pairs (pa, pb) with pa + pb < NPL (3 for hi + lo, 6 for hi + mid + lo) on
// v_mfma_f32_16x16x32_bf16 into FP32 partial sums, added into FP64 every 2 groups (64 reductions).
template <int NPL, int RTW, int NCT>
__global__ void __launch_bounds__(DNT, 2)
    dense_m2l_bf16_kernel(const int32_t *__restrict__ recv, const int64_t *__restrict__ item_ptr,
                          const int32_t *__restrict__ item_src, const int32_t *__restrict__ item_fop,
                          const bf16x8_d *__restrict__ frag, int64_t plane, int Kr, int KrT,
                          const float *__restrict__ mom, float *__restrict__ loc, int R,
                          const uint8_t *__restrict__ acc_flags) {
    const int ti = xcd_order(blockIdx.x, gridDim.x);
    const int t = recv[ti];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kq = lane >> 4;
    const int r0 = blockIdx.y * 16 * NCT;
    const int G2 = (KrT + 1) / 2;
    const int ntw = (KrT - wave + 3) >> 2;
    const int64_t fsz = (int64_t)KrT * G2 * 64;  // bf16x8 per operator and plane
    double acc[RTW][NCT][4];
    float4_d part[RTW][NCT];
#pragma unroll
    for (int j = 0; j < RTW; ++j)
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            part[j][c] = float4_d{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[j][c][v] = 0.0;
        }
    const int64_t i0 = item_ptr[ti], i1 = item_ptr[ti + 1];
    if (acc_flags && i1 > i0 && acc_flags[i0]) init_from_out<false>(acc, loc, t, R, Kr, KrT, r0, wave, ntw, n, kq);
    for (int64_t it = i0; it < i1; ++it) {
        const bf16x8_d *A0 = frag + (int64_t)item_fop[it] * fsz + lane;
        const int64_t s = item_src[it];
        const float *Bc[NCT];
        bool bval[NCT];
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int col = r0 + 16 * c + n;
            bval[c] = col < R;
            Bc[c] = mom + (s * R + (bval[c] ? col : 0)) * (int64_t)Kr;
        }
        for (int g = 0; g < G2; ++g) {
            bf16x8_d a[RTW][NPL], b[NCT][NPL];
#pragma unroll
            for (int j = 0; j < RTW; ++j)
#pragma unroll
                for (int pl = 0; pl < NPL; ++pl)
                    if (j < ntw) a[j][pl] = A0[pl * plane + ((int64_t)(wave + 4 * j) * G2 + g) * 64];
#pragma unroll
            for (int c = 0; c < NCT; ++c) load_split8<NPL>(Bc[c], 32 * g + 8 * kq, Kr, bval[c], b[c]);
#pragma unroll
            for (int j = 0; j < RTW; ++j) {
                if (j >= ntw) continue;
#pragma unroll
                for (int c = 0; c < NCT; ++c) {
                    float4_d d = part[j][c];
                    // the small products first
#pragma unroll
                    for (int ord = NPL - 1; ord >= 0; --ord)
#pragma unroll
                        for (int pa = 0; pa <= ord; ++pa)
                            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j][pa], b[c][ord - pa], d, 0, 0, 0);
                    part[j][c] = d;
                }
            }
            if ((g & 1) == 1 || g == G2 - 1) {
#pragma unroll
                for (int j = 0; j < RTW; ++j)
#pragma unroll
                    for (int c = 0; c < NCT; ++c) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[j][c][v] += (double)part[j][c][v];
                        part[j][c] = float4_d{0.f, 0.f, 0.f, 0.f};
                    }
            }
        }
    }
    // D layout of the 16 x 16 f32 MFMAs: D[m][n] in lane n + 16 (m / 4), element m % 4
#pragma unroll
    for (int j = 0; j < RTW; ++j) {
        if (j >= ntw) continue;
        const int rt = wave + 4 * j;
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int col = r0 + 16 * c + n;
            if (col >= R) continue;
            float *o = loc + ((int64_t)t * R + col) * Kr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * rt + 4 * kq + v;
                if (row < Kr) o[row] = (float)acc[j][c][v];
            }
        }
    }
}

template <int NPL, int RTW>
void dense_bf16_go_nct(int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
                       const int32_t *item_fop, const void *frag, int64_t plane, int Kr, int KrT, const void *mom,
                       void *loc, int R, const uint8_t *acc, hipStream_t s) {
    const dim3 grid((unsigned)n_recv, (unsigned)((R + 16 * nct - 1) / (16 * nct)));
    if (nct == 1)
        hipLaunchKernelGGL((dense_m2l_bf16_kernel<NPL, RTW, 1>), grid, DNT, 0, s, recv, item_ptr, item_src, item_fop,
                           (const bf16x8_d *)frag, plane, Kr, KrT, (const float *)mom, (float *)loc, R, acc);
    else
        hipLaunchKernelGGL((dense_m2l_bf16_kernel<NPL, RTW, 2>), grid, DNT, 0, s, recv, item_ptr, item_src, item_fop,
                           (const bf16x8_d *)frag, plane, Kr, KrT, (const float *)mom, (float *)loc, R, acc);
}

template <int NPL>
void dense_bf16_go(int rtw, int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr,
                   const int32_t *item_src, const int32_t *item_fop, const void *frag, int64_t plane, int Kr, int KrT,
                   const void *mom, void *loc, int R, const uint8_t *acc, hipStream_t s) {
    switch (rtw) {
    case 1: dense_bf16_go_nct<NPL, 1>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    case 2: dense_bf16_go_nct<NPL, 2>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    case 3: dense_bf16_go_nct<NPL, 3>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    case 4: dense_bf16_go_nct<NPL, 4>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    case 5: dense_bf16_go_nct<NPL, 5>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    case 6: dense_bf16_go_nct<NPL, 6>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    default: dense_bf16_go_nct<NPL, 7>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, plane, Kr, KrT, mom, loc, R, acc, s); break;
    }
}

template <int MODE, int RTW, bool CF>
void dense_go_nct(int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
                  const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom, void *loc, int R,
                  const uint8_t *acc, hipStream_t s) {
    using V = typename std::conditional<MODE == 0, double, float>::type;
    const dim3 grid((unsigned)n_recv, (unsigned)((R + 16 * nct - 1) / (16 * nct)));
    if (nct == 1)
        hipLaunchKernelGGL((dense_m2l_kernel<MODE, RTW, 1, CF>), grid, DNT, 0, s, recv, item_ptr, item_src, item_fop,
                           (const V *)frag, Kr, KrT, (const V *)mom, (V *)loc, R, acc);
    else
        hipLaunchKernelGGL((dense_m2l_kernel<MODE, RTW, 2, CF>), grid, DNT, 0, s, recv, item_ptr, item_src, item_fop,
                           (const V *)frag, Kr, KrT, (const V *)mom, (V *)loc, R, acc);
}

template <int MODE, bool CF>
void dense_go_cf(int rtw, int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
                 const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom, void *loc, int R,
                 const uint8_t *acc, hipStream_t s) {
    switch (rtw) {
    case 1: dense_go_nct<MODE, 1, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    case 2: dense_go_nct<MODE, 2, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    case 3: dense_go_nct<MODE, 3, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    case 4: dense_go_nct<MODE, 4, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    case 5: dense_go_nct<MODE, 5, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    case 6: dense_go_nct<MODE, 6, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    default: dense_go_nct<MODE, 7, CF>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s); break;
    }
}

// cf: the operators are complex fragments (complex maps)
template <int MODE>
void dense_go(bool cf, int rtw, int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr,
              const int32_t *item_src, const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom,
              void *loc, int R, const uint8_t *acc, hipStream_t s) {
    if (cf)
        dense_go_cf<MODE, true>(rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s);
    else
        dense_go_cf<MODE, false>(rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, acc, s);
}


template <int MODE, int NW, int RTW>
void dense_pair_go3(const int32_t *recv2, int n_pairs, const int64_t *item_ptr, const int2 *item_src,
                    const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom, void *loc, int R,
                    hipStream_t s) {
    using V = typename std::conditional<MODE == 0, double, float>::type;
    const int nchunk = (R + 15) / 16;
    hipLaunchKernelGGL((dense_m2l_pair_kernel<MODE, NW, RTW>), dim3((unsigned)((int64_t)n_pairs * nchunk)), 64 * NW, 0,
                       s, recv2, item_ptr, item_src, item_fop, (const V *)frag, Kr, KrT, (const V *)mom, (V *)loc, R,
                       nchunk);
}

// NW = 4 waves for at most 4 row tiles, else 8; RTW = row tiles per wave (at most 4)
template <int MODE>
void dense_pair_go(int nw, int rtw, const int32_t *recv2, int n_pairs, const int64_t *item_ptr, const int2 *item_src,
                   const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom, void *loc, int R,
                   hipStream_t s) {
#define H2G_PAIR_GO(NW_, RTW_) \
    dense_pair_go3<MODE, NW_, RTW_>(recv2, n_pairs, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s)
    switch (nw * 10 + rtw) {
    case 41: H2G_PAIR_GO(4, 1); break;
    case 81: H2G_PAIR_GO(8, 1); break;
    case 82: H2G_PAIR_GO(8, 2); break;
    case 83: H2G_PAIR_GO(8, 3); break;
    case 84: H2G_PAIR_GO(8, 4); break;
    default: throw Error{H2G_EINVAL, "dense pair kernel: no instantiation for this tile shape"};
    }
#undef H2G_PAIR_GO
}

}  // namespace

void build_dense_ops(DenseOps &D, int dtype, int K, const void *ops, const int32_t *fop_op, const int32_t *fop_sym,
                     const int32_t *perms_inv, int64_t n, int orient_mask, int reps) {
    D = DenseOps{};
    const bool cplx = dtype == H2G_C64;
    const int Kr = cplx ? 2 * K : K;
    const int KrT = (Kr + 15) / 16;
    if (KrT > 4 * DENSE_MAX_RTW || n <= 0) return;
    D.Kr = Kr;
    D.KrT = KrT;
    D.n = n;
    DevBuf d_op, d_sym;
    d_op.upload(fop_op, (size_t)n * sizeof(int32_t));
    if (fop_sym) d_sym.upload(fop_sym, (size_t)n * sizeof(int32_t));
    const int32_t *dsym = fop_sym ? d_sym.as<int32_t>() : nullptr;
    const size_t per = (size_t)(16 * KrT) * (16 * KrT);
    const int64_t total = n * (int64_t)per;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 20);
    for (int o = 0; o < 4; ++o) {
        if (!(orient_mask & (1 << o))) continue;
        auto build = [&](auto tag, DevBuf &dst) {
            using OUT = decltype(tag);
            if (cplx) {
                // complex fragments: half the real form's values
                const int64_t ctot = n * (int64_t)KrT * KrT * 128;
                dst.alloc((size_t)ctot * sizeof(OUT));
                const unsigned cb = (unsigned)std::min<int64_t>((ctot + 255) / 256, 1 << 20);
                hipLaunchKernelGGL((dense_cfrag_kernel<OUT>), cb, 256, 0, 0, (const double *)ops, K, Kr, KrT, n,
                                   d_op.as<int32_t>(), dsym, perms_inv, o, dst.as<OUT>());
                H2G_CHECK_LAUNCH();
                return;
            }
            dst.alloc(per * n * sizeof(OUT));
            if (cplx)
                hipLaunchKernelGGL((dense_frag_kernel<OUT, true>), blocks, 256, 0, 0, (const double *)ops, K, Kr, KrT, n,
                                   d_op.as<int32_t>(), dsym, perms_inv, o, dst.as<OUT>());
            else
                hipLaunchKernelGGL((dense_frag_kernel<OUT, false>), blocks, 256, 0, 0, (const double *)ops, K, Kr, KrT,
                                   n, d_op.as<int32_t>(), dsym, perms_inv, o, dst.as<OUT>());
            H2G_CHECK_LAUNCH();
        };
        if (reps & DENSE_REP_F64) build(0.0, D.f64[o]);
        if (reps & DENSE_REP_F32) build(0.0f, D.f32[o]);
        if (reps & DENSE_REP_BF16) {
            const int64_t tb = n * (int64_t)KrT * ((KrT + 1) / 2) * 512;
            D.bf[o].alloc(3 * (size_t)tb * sizeof(__bf16));
            const unsigned bb = (unsigned)std::min<int64_t>((tb + 255) / 256, 1 << 20);
            if (cplx)
                hipLaunchKernelGGL((dense_frag_bf16_kernel<true>), bb, 256, 0, 0, (const double *)ops, K, Kr, KrT, n,
                                   d_op.as<int32_t>(), dsym, perms_inv, o, D.bf[o].as<__bf16>());
            else
                hipLaunchKernelGGL((dense_frag_bf16_kernel<false>), bb, 256, 0, 0, (const double *)ops, K, Kr, KrT, n,
                                   d_op.as<int32_t>(), dsym, perms_inv, o, D.bf[o].as<__bf16>());
            H2G_CHECK_LAUNCH();
        }
    }
    H2G_HIP(hipDeviceSynchronize());
    D.cf = cplx;
    D.ok = true;
}

void launch_dense(const DenseOps &D, int orient, int rep, const int32_t *recv, int n_recv, const int64_t *item_ptr,
                  const int32_t *item_src, const int32_t *item_fop, const void *in, void *out, int R,
                  const uint8_t *acc_flags, hipStream_t s) {
    if (n_recv == 0) return;
    const int rtw = (D.KrT + 3) / 4;
    const int nct = R > 16 ? 2 : 1;
    if (rep == DENSE_REP_BF16X2 || rep == DENSE_REP_BF16X3) {
        const int64_t plane = D.n * (int64_t)D.KrT * ((D.KrT + 1) / 2) * 64;  // in bf16x8
        if (!D.bf[orient].ptr) throw Error{H2G_EINVAL, "split-BF16 operators were not built for this core"};
        if (rep == DENSE_REP_BF16X2)
            dense_bf16_go<2>(rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, D.bf[orient].ptr, plane, D.Kr,
                             D.KrT, in, out, R, acc_flags, s);
        else
            dense_bf16_go<3>(rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, D.bf[orient].ptr, plane, D.Kr,
                             D.KrT, in, out, R, acc_flags, s);
    } else if (rep == DENSE_REP_F32) {
        if (!D.f32[orient].ptr) throw Error{H2G_EINVAL, "FP32 operators were not built for this core"};
        dense_go<1>(D.cf, rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, D.f32[orient].ptr, D.Kr, D.KrT, in,
                    out, R, acc_flags, s);
    } else {
        if (!D.f64[orient].ptr) throw Error{H2G_EINVAL, "FP64 operators were not built for this core"};
        dense_go<0>(D.cf, rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, D.f64[orient].ptr, D.Kr, D.KrT, in,
                    out, R, acc_flags, s);
    }
    H2G_CHECK_LAUNCH();
}

bool launch_dense_pair(const DenseOps &D, int orient, int rep, int want, const int32_t *recv2, int n_pairs,
                       const int64_t *item_ptr, const int32_t *item_src2, const int32_t *item_fop, const void *in,
                       void *out, int R, hipStream_t s) {
    if (want < 0 || D.cf || (rep != DENSE_REP_F32 && rep != DENSE_REP_F64)) return false;
    const DevBuf &F = rep == DENSE_REP_F32 ? D.f32[orient] : D.f64[orient];
    const void *frag = F.ptr;
    if (!frag) return false;
    // measured on dense Laplace maps at N = 2M: p = 6 (1.6k operators of 216 x 216, 0.33 GB in
    // FP32, 0.66 GB in FP64) coupling stage from 1 to 16 columns 513-520 -> 222-259 ms (FP32),
    // 553-691 -> 366-610 ms (FP64), at 32 columns 391 -> 519 ms (FP32); p = 4 (operators in the
    // caches) 10-25 % slower in pairs
    if (want == 0 && (R > 16 || (int64_t)F.bytes <= DENSE_PAIR_MIN_BYTES)) return false;
    const int nw = D.KrT <= 4 ? 4 : 8, rtw = (D.KrT + nw - 1) / nw;
    if (rtw > 4) return false;
    if (n_pairs == 0) return true;
    const int2 *src2 = reinterpret_cast<const int2 *>(item_src2);
    if (rep == DENSE_REP_F32)
        dense_pair_go<1>(nw, rtw, recv2, n_pairs, item_ptr, src2, item_fop, frag, D.Kr, D.KrT, in, out, R, s);
    else
        dense_pair_go<0>(nw, rtw, recv2, n_pairs, item_ptr, src2, item_fop, frag, D.Kr, D.KrT, in, out, R, s);
    H2G_CHECK_LAUNCH();
    return true;
}

}  // namespace h2g

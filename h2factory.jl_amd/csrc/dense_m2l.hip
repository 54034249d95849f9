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
//     at the FP32 rounding of the operands instead of growing with the ~190 partners x K terms.
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

// Fragment matrix of expanded operator e for orientation `orient` (0: A, 1: A^T, 2: A^H), from the
// canonical couplings `ops` (core type, column-major K x K, doubles; complex interleaved).
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
        double v = 0.0;
        if (i < Kr && j < Kr) {
            const int ii = CPLX ? i >> 1 : i, jj = CPLX ? j >> 1 : j;
            const int p = orient == 0 ? ii : jj, q = orient == 0 ? jj : ii;
            const int sym = fop_sym ? fop_sym[e] : -1;
            const int pp = sym >= 0 ? perms_inv[(int64_t)sym * K + p] : p;
            const int qq = sym >= 0 ? perms_inv[(int64_t)sym * K + q] : q;
            const int64_t at = (int64_t)fop_op[e] * K * K + pp + (int64_t)K * qq;
            if (CPLX) {
                const double re = ops[2 * at], im = orient == 2 ? -ops[2 * at + 1] : ops[2 * at + 1];
                const int a = i & 1, b = j & 1;
                v = a == b ? re : (a == 0 ? -im : im);
            } else {
                v = ops[at];
            }
        }
        frag[idx] = (OUT)v;
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

// MODE 0: FP64 operands on FP64 MFMA; MODE 1: FP32 operands on FP32 MFMA with FP64 accumulation
// of the partial sums.  RTW: row tiles per wave (16 rows each), NCT: column tiles (16 columns).
template <int MODE, int RTW, int NCT>
__global__ void __launch_bounds__(DNT, 2)
    dense_m2l_kernel(const int32_t *__restrict__ recv, const int64_t *__restrict__ item_ptr,
                     const int32_t *__restrict__ item_src, const int32_t *__restrict__ item_fop,
                     const typename std::conditional<MODE == 0, double, float>::type *__restrict__ frag, int Kr,
                     int KrT, const typename std::conditional<MODE == 0, double, float>::type *__restrict__ mom,
                     typename std::conditional<MODE == 0, double, float>::type *__restrict__ loc, int R) {
    using V = typename std::conditional<MODE == 0, double, float>::type;
    using V4 = typename dvec<V>::type;
    const int ti = blockIdx.x;
    const int t = recv[ti];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kq = lane >> 4;
    const int r0 = blockIdx.y * 16 * NCT;
    const int G = KrT;
    const int ntw = (KrT - wave + 3) >> 2;  // row tiles of this wave: wave, wave + 4, ...
    const int64_t fsz = (int64_t)(16 * KrT) * (16 * KrT);
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
    for (int64_t it = i0; it < i1; ++it) {
        const V *A0 = frag + (int64_t)item_fop[it] * fsz + (int64_t)lane * 4;
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
            for (int j = 0; j < RTW; ++j)
                a[j] = j < ntw ? *reinterpret_cast<const V4 *>(A0 + ((int64_t)(wave + 4 * j) * G + g) * 256)
                               : V4{0, 0, 0, 0};
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

template <int MODE, int RTW>
void dense_go_nct(int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
                  const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom, void *loc, int R,
                  hipStream_t s) {
    using V = typename std::conditional<MODE == 0, double, float>::type;
    const dim3 grid((unsigned)n_recv, (unsigned)((R + 16 * nct - 1) / (16 * nct)));
    if (nct == 1)
        hipLaunchKernelGGL((dense_m2l_kernel<MODE, RTW, 1>), grid, DNT, 0, s, recv, item_ptr, item_src, item_fop,
                           (const V *)frag, Kr, KrT, (const V *)mom, (V *)loc, R);
    else
        hipLaunchKernelGGL((dense_m2l_kernel<MODE, RTW, 2>), grid, DNT, 0, s, recv, item_ptr, item_src, item_fop,
                           (const V *)frag, Kr, KrT, (const V *)mom, (V *)loc, R);
}

template <int MODE>
void dense_go(int rtw, int nct, const int32_t *recv, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
              const int32_t *item_fop, const void *frag, int Kr, int KrT, const void *mom, void *loc, int R,
              hipStream_t s) {
    switch (rtw) {
    case 1: dense_go_nct<MODE, 1>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    case 2: dense_go_nct<MODE, 2>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    case 3: dense_go_nct<MODE, 3>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    case 4: dense_go_nct<MODE, 4>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    case 5: dense_go_nct<MODE, 5>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    case 6: dense_go_nct<MODE, 6>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    default: dense_go_nct<MODE, 7>(nct, recv, n_recv, item_ptr, item_src, item_fop, frag, Kr, KrT, mom, loc, R, s); break;
    }
}

}  // namespace

void build_dense_frags(Core &c, const int32_t *fop_op, const int32_t *fop_sym, int64_t n_fops) {
    c.dense_ok = false;
    const bool cplx = c.dtype == H2G_C64;
    const int Kr = cplx ? 2 * c.K : c.K;
    const int KrT = (Kr + 15) / 16;
    if (KrT > 4 * DENSE_MAX_RTW || n_fops <= 0) return;
    c.dense_Kr = Kr;
    c.dense_KrT = KrT;
    c.n_fops = n_fops;
    DevBuf d_op, d_sym;
    d_op.upload(fop_op, (size_t)n_fops * sizeof(int32_t));
    if (fop_sym) d_sym.upload(fop_sym, (size_t)n_fops * sizeof(int32_t));
    const size_t per = (size_t)(16 * KrT) * (16 * KrT);
    const int norient = cplx ? 3 : 2;
    const int64_t total = n_fops * (int64_t)per;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 20);
    for (int o = 0; o < norient; ++o) {
        auto build = [&](auto tag, DevBuf &dst) {
            using OUT = decltype(tag);
            dst.alloc(per * n_fops * sizeof(OUT));
            if (cplx)
                hipLaunchKernelGGL((dense_frag_kernel<OUT, true>), blocks, 256, 0, 0, c.ops.as<double>(), c.K, Kr, KrT,
                                   n_fops, d_op.as<int32_t>(), fop_sym ? d_sym.as<int32_t>() : nullptr,
                                   c.perms_inv.as<int32_t>(), o, dst.as<OUT>());
            else
                hipLaunchKernelGGL((dense_frag_kernel<OUT, false>), blocks, 256, 0, 0, c.ops.as<double>(), c.K, Kr,
                                   KrT, n_fops, d_op.as<int32_t>(), fop_sym ? d_sym.as<int32_t>() : nullptr,
                                   c.perms_inv.as<int32_t>(), o, dst.as<OUT>());
            H2G_CHECK_LAUNCH();
        };
        if (c.precisions & (1 << H2G_PREC_PLAIN)) build(0.0, c.dfrag64[o]);
        if (c.precisions & (1 << H2G_PREC_MIXED)) build(0.0f, c.dfrag32[o]);
    }
    H2G_HIP(hipDeviceSynchronize());
    c.dense_ok = true;
}

void launch_dense_m2l(const Core &c, bool mixed, int op, const int32_t *recv, int n_recv, const int64_t *item_ptr,
                      const int32_t *item_src, const int32_t *item_fop, const void *mom, void *loc, int R,
                      hipStream_t s) {
    if (n_recv == 0) return;
    const int o = op == H2G_OP_N ? 0 : (op == H2G_OP_T || c.dtype != H2G_C64 ? 1 : 2);
    const int rtw = (c.dense_KrT + 3) / 4;
    const int nct = R > 16 ? 2 : 1;
    if (mixed)
        dense_go<1>(rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, c.dfrag32[o].ptr, c.dense_Kr, c.dense_KrT,
                    mom, loc, R, s);
    else
        dense_go<0>(rtw, nct, recv, n_recv, item_ptr, item_src, item_fop, c.dfrag64[o].ptr, c.dense_Kr, c.dense_KrT,
                    mom, loc, R, s);
    H2G_CHECK_LAUNCH();
}

}  // namespace h2g

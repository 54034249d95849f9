// Device kernels of libh2g.
//
// Buffers in tree order, R = right-hand sides of the current pass:
//   xs, ys   [dof][R]        FP64 / ComplexF64 (near accumulation and final sum stay FP64)
//   mom, loc [box][R][K]     compute type (FP64/C64 plain, FP32/C32 mixed)
//   spectra  [slot][R][L^3]  complex compute type, frequency x fastest
#include <algorithm>
#include <type_traits>

#include "h2g_internal.h"

namespace h2g {

namespace {

constexpr int TPB = 256;

template <class T>
using store_t = typename std::conditional<is_complex<T>::value, cplx<double>, double>::type;

template <class T>
__device__ inline T maybe_conj(T v, bool c) {
    return c ? conjv(v) : v;
}

inline unsigned grid1(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// ------------------------------------------------------------------------------------------
// gather: xs[i][r] = x[perm[i] + r*ldx]
// ------------------------------------------------------------------------------------------
template <class S, class X>
__global__ void gather_kernel(const X *__restrict__ x, int64_t ldx, const int64_t *__restrict__ perm, int64_t n,
                              int R, S *__restrict__ xs) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * R) return;
    int64_t i = idx / R;
    int r = (int)(idx - i * R);
    xs[idx] = cvt<S>::from(x[perm[i] + (int64_t)r * ldx]);
}

// ------------------------------------------------------------------------------------------
// scatter: y[perm[i] + r*ldy] = a*ys[i][r] + b*y (b == 0: y is not read); masked rows skipped
// ------------------------------------------------------------------------------------------
template <class S, class Y>
__global__ void scatter_kernel(const S *__restrict__ ys, const int64_t *__restrict__ perm, int64_t n, int R,
                               Y *__restrict__ y, int64_t ldy, cplx<double> a, cplx<double> b,
                               const uint8_t *__restrict__ mask) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * R) return;
    int64_t i = idx / R;
    int r = (int)(idx - i * R);
    if (mask && !mask[i]) return;
    int64_t o = perm[i] + (int64_t)r * ldy;
    S v = scale_by<S>(a, ys[idx]);
    if (b.re != 0.0 || b.im != 0.0) v = v + scale_by<S>(b, cvt<S>::from(y[o]));
    y[o] = cvt<Y>::from(v);
}

// ------------------------------------------------------------------------------------------
// Tiled forms for batches (R >= 8): a workgroup takes GT consecutive points in input order and a
// chunk of up to GR columns, with the inverse permutation (input index -> tree index).  The
// column-major x / y side is read / written coalesced along the points of each column, the
// tree-ordered xs / ys side row by row (R contiguous values per point); LDS holds the tile.
// (The plain forms above touch x / y at random rows for every (point, column).)
// ------------------------------------------------------------------------------------------
constexpr int GT = 64, GR = 32;

template <class S, class X>
__global__ void __launch_bounds__(TPB) gather_tiled_kernel(const X *__restrict__ x, int64_t ldx,
                                                          const int64_t *__restrict__ iperm, int64_t n, int R,
                                                          S *__restrict__ xs) {
    __shared__ S t[GT][GR + 1];
    const int64_t j0 = (int64_t)blockIdx.x * GT;
    const int r0 = blockIdx.y * GR;
    const int nr = min(GR, R - r0);
    for (int e = threadIdx.x; e < GT * nr; e += TPB) {
        const int c = e / GT, j = e - c * GT;
        if (j0 + j < n) t[j][c] = cvt<S>::from(x[j0 + j + (int64_t)(r0 + c) * ldx]);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < GT * nr; e += TPB) {
        const int j = e / nr, c = e - j * nr;
        if (j0 + j < n) xs[iperm[j0 + j] * R + r0 + c] = t[j][c];
    }
}

template <class S, class Y>
__global__ void __launch_bounds__(TPB) scatter_tiled_kernel(const S *__restrict__ ys, const int64_t *__restrict__ iperm,
                                                           int64_t n, int R, Y *__restrict__ y, int64_t ldy,
                                                           cplx<double> a, cplx<double> b,
                                                           const uint8_t *__restrict__ mask) {
    __shared__ S t[GT][GR + 1];
    const int64_t j0 = (int64_t)blockIdx.x * GT;
    const int r0 = blockIdx.y * GR;
    const int nr = min(GR, R - r0);
    for (int e = threadIdx.x; e < GT * nr; e += TPB) {
        const int j = e / nr, c = e - j * nr;
        if (j0 + j < n) t[j][c] = ys[iperm[j0 + j] * R + r0 + c];
    }
    __syncthreads();
    const bool beta = b.re != 0.0 || b.im != 0.0;
    for (int e = threadIdx.x; e < GT * nr; e += TPB) {
        const int c = e / GT, j = e - c * GT;
        if (j0 + j >= n || (mask && !mask[iperm[j0 + j]])) continue;
        const int64_t o = j0 + j + (int64_t)(r0 + c) * ldy;
        S v = scale_by<S>(a, t[j][c]);
        if (beta) v = v + scale_by<S>(b, cvt<S>::from(y[o]));
        y[o] = cvt<Y>::from(v);
    }
}

// ------------------------------------------------------------------------------------------
// near field: one workgroup per target leaf; blocks are column-major nr x nc.
//   forward:   ys[rb + i] += B[i, j] * xs[cb + j]
//   transpose: ys[cb + j] += op(B[i, j]) * xs[rb + i]
// Wide products (R >= 8): a thread owns (row, 8 consecutive RHS); lanes of a wave run over rows
// so each B load is coalesced and each x load is a broadcast.
// Narrow products: a thread owns (row, column group); partial sums are reduced through LDS.
// ------------------------------------------------------------------------------------------
constexpr int NRC = 8;  // RHS at which the wide kernel takes over

constexpr int NRG = 4;  // rows per thread in the wide kernel

template <class S, bool TRANS>
__global__ void near_wide_kernel(const int32_t *__restrict__ ptr_tgt, const int64_t *__restrict__ ptr,
                                 const int32_t *__restrict__ blk, const int64_t *__restrict__ rb,
                                 const int64_t *__restrict__ cb, const int32_t *__restrict__ nr,
                                 const int32_t *__restrict__ nc, const int64_t *__restrict__ off,
                                 const S *__restrict__ vals, const S *__restrict__ xs, S *__restrict__ ys, int R,
                                 bool conj) {
    // a thread owns NRG rows x one RHS; lanes run over r, so x loads are coalesced and the
    // NRG block entries of a column are wave-uniform broadcasts
    const int t = blockIdx.x;
    const int64_t b0 = ptr[t], b1 = ptr[t + 1];
    if (b0 == b1) return;
    const int first = blk[b0];
    const int n_out = TRANS ? nc[first] : nr[first];
    const int64_t out0 = TRANS ? cb[first] : rb[first];
    const int ngroups = (n_out + NRG - 1) / NRG;
    for (int o = threadIdx.x; o < ngroups * R; o += blockDim.x) {
        const int r = o % R;
        const int g0 = (o / R) * NRG;
        const int ng = min(NRG, n_out - g0);
        S acc[NRG];
#pragma unroll
        for (int q = 0; q < NRG; ++q) acc[q] = zero_v<S>();
        for (int64_t qq = b0; qq < b1; ++qq) {
            const int bl = blk[qq];
            const int m = nr[bl], n = nc[bl];
            const S *B = vals + off[bl];
            if (!TRANS) {
                const S *xin = xs + cb[bl] * R + r;
                int j = 0;
                for (; j + 4 <= n; j += 4) {
                    S xv[4], bv[4][NRG];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        xv[u] = xin[(int64_t)(j + u) * R];
                        const S *Bc = B + (int64_t)m * (j + u) + g0;
#pragma unroll
                        for (int q = 0; q < NRG; ++q) bv[u][q] = q < ng ? Bc[q] : zero_v<S>();
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int q = 0; q < NRG; ++q) fmacc(acc[q], bv[u][q], xv[u]);
                }
                for (; j < n; ++j) {
                    const S xv = xin[(int64_t)j * R];
                    const S *Bc = B + (int64_t)m * j + g0;
#pragma unroll
                    for (int q = 0; q < NRG; ++q)
                        if (q < ng) fmacc(acc[q], Bc[q], xv);
                }
            } else {
                const S *xin = xs + rb[bl] * R + r;
                int i = 0;
                for (; i + 4 <= m; i += 4) {
                    S xv[4], bv[4][NRG];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        xv[u] = xin[(int64_t)(i + u) * R];
#pragma unroll
                        for (int q = 0; q < NRG; ++q)
                            bv[u][q] = q < ng ? maybe_conj(B[i + u + (int64_t)m * (g0 + q)], conj) : zero_v<S>();
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int q = 0; q < NRG; ++q) fmacc(acc[q], bv[u][q], xv[u]);
                }
                for (; i < m; ++i) {
                    const S xv = xin[(int64_t)i * R];
#pragma unroll
                    for (int q = 0; q < NRG; ++q)
                        if (q < ng) fmacc(acc[q], maybe_conj(B[i + (int64_t)m * (g0 + q)], conj), xv);
                }
            }
        }
        for (int q = 0; q < ng; ++q) ys[(out0 + g0 + q) * R + r] = acc[q];
    }
}

template <class S, bool TRANS>
__global__ void near_narrow_kernel(const int32_t *__restrict__ ptr_tgt, const int64_t *__restrict__ ptr,
                                   const int32_t *__restrict__ blk, const int64_t *__restrict__ rb,
                                   const int64_t *__restrict__ cb, const int32_t *__restrict__ nr,
                                   const int32_t *__restrict__ nc, const int64_t *__restrict__ off,
                                   const S *__restrict__ vals, const S *__restrict__ xs, S *__restrict__ ys, int R,
                                   bool conj) {
    __shared__ S part[256];
    const int t = blockIdx.x;
    const int64_t b0 = ptr[t], b1 = ptr[t + 1];
    if (b0 == b1) return;
    const int first = blk[b0];
    const int n_out = TRANS ? nc[first] : nr[first];
    const int64_t out0 = TRANS ? cb[first] : rb[first];
    const int G = max(1, (int)blockDim.x / max(n_out, 1));  // column groups per row
    for (int r = 0; r < R; ++r) {
        for (int rbase = 0; rbase < n_out; rbase += (int)blockDim.x) {
            const int rows = min(n_out - rbase, (int)blockDim.x);
            const int g = min(G, max(1, (int)blockDim.x / rows));
            const int row = rbase + threadIdx.x % rows;
            const int jg = threadIdx.x / rows;
            S acc = zero_v<S>();
            if (jg < g) {
                for (int64_t qq = b0; qq < b1; ++qq) {
                    const int bl = blk[qq];
                    const int m = nr[bl], n = nc[bl];
                    const S *B = vals + off[bl];
                    if (!TRANS) {
                        const S *xin = xs + cb[bl] * R + r;
                        for (int j = jg; j < n; j += g) fmacc(acc, B[row + (int64_t)m * j], xin[(int64_t)j * R]);
                    } else {
                        const S *xin = xs + rb[bl] * R + r;
                        const S *col = B + (int64_t)m * row;
                        for (int i = jg; i < m; i += g) fmacc(acc, maybe_conj(col[i], conj), xin[(int64_t)i * R]);
                    }
                }
            }
            part[threadIdx.x] = acc;
            __syncthreads();
            if (threadIdx.x < rows) {
                S sum = part[threadIdx.x];
                for (int q = 1; q < g; ++q) sum = sum + part[threadIdx.x + q * rows];
                ys[(out0 + rbase + threadIdx.x) * R + r] = sum;
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------------------------------
// P2M: mom[leaf][r][k] = sum_j opC(C[k, b0 + j]) * xs[b0 + j][r]
// A thread owns (k, 8 RHS); the leaf's x block (converted to the compute type) is staged in
// LDS in chunks of 64 points, so each C load (coalesced over k) feeds 8 FMAs.
// ------------------------------------------------------------------------------------------
constexpr int TRC = 8;  // RHS per thread

template <class T>
__global__ void p2m_kernel(const int32_t *__restrict__ leaves, const int64_t *__restrict__ begin,
                           const int64_t *__restrict__ end, const T *__restrict__ coll, int K,
                           const store_t<T> *__restrict__ xs, int R, T *__restrict__ mom, bool conj, int PCH) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *xb = reinterpret_cast<T *>(smem_raw);  // [PCH][R]
    const int b = leaves[blockIdx.x];
    const int64_t j0 = begin[b], n = end[b] - j0;
    const int nch = (R + TRC - 1) / TRC;
    const int ntask = K * nch;
    T *m = mom + (int64_t)b * K * R;
    for (int task0 = 0; task0 < ntask; task0 += blockDim.x) {
        const int task = task0 + threadIdx.x;
        const int k = task % K, r0 = (task / K) * TRC;
        const int nr = min(TRC, R - r0);
        T acc[TRC];
#pragma unroll
        for (int q = 0; q < TRC; ++q) acc[q] = zero_v<T>();
        for (int64_t c0 = 0; c0 < n; c0 += PCH) {
            const int cn = (int)min((int64_t)PCH, n - c0);
            __syncthreads();
            for (int e = threadIdx.x; e < cn * R; e += blockDim.x) xb[e] = cvt<T>::from(xs[(j0 + c0) * R + e]);
            __syncthreads();
            if (task < ntask) {
                const T *c = coll + (j0 + c0) * K + k;
                for (int j = 0; j < cn; ++j) {
                    const T a = maybe_conj(c[(int64_t)j * K], conj);
                    const T *xr = xb + j * R + r0;
#pragma unroll
                    for (int q = 0; q < TRC; ++q)
                        if (q < nr) fmacc(acc[q], a, xr[q]);
                }
            }
        }
        if (task < ntask)
            for (int q = 0; q < nr; ++q) m[(int64_t)(r0 + q) * K + k] = acc[q];
    }
}

// ------------------------------------------------------------------------------------------
// L2P: ys[b0 + j][r] += alpha * sum_k opC(C[k, b0 + j]) * loc[leaf][r][k]
// The leaf's local moments are staged in LDS transposed, [k][r]; a thread owns one (point, r)
// pair, so a wave spans a few points x consecutive RHS: collection loads are wave-broadcasts
// and the LDS reads are conflict-free.
// ------------------------------------------------------------------------------------------
template <class T>
__global__ void l2p_kernel(const int32_t *__restrict__ leaves, const int64_t *__restrict__ begin,
                           const int64_t *__restrict__ end, const T *__restrict__ coll, int K,
                           const T *__restrict__ loc, int R, store_t<T> *__restrict__ ys, cplx<double> alpha,
                           bool conj, int RC) {
    using S = store_t<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lb = reinterpret_cast<T *>(smem_raw);  // [K][RC]
    const int b = leaves[blockIdx.x];
    const int64_t j0 = begin[b], n = end[b] - j0;
    const T *l = loc + (int64_t)b * K * R;
    for (int r0 = 0; r0 < R; r0 += RC) {
        const int nr = min(RC, R - r0);
        __syncthreads();
        for (int e = threadIdx.x; e < K * nr; e += blockDim.x) {
            const int q = e / K, k = e - q * K;  // coalesced read of loc[r0+q][k]
            lb[k * RC + q] = l[(int64_t)(r0 + q) * K + k];
        }
        __syncthreads();
        for (int64_t o = threadIdx.x; o < n * nr; o += blockDim.x) {
            const int64_t j = o / nr;
            const int q = (int)(o - j * nr);
            const T *c = coll + (j0 + j) * K;
            T acc0 = zero_v<T>(), acc1 = zero_v<T>();
            int k = 0;
            for (; k + 2 <= K; k += 2) {
                fmacc(acc0, maybe_conj(c[k], conj), lb[k * RC + q]);
                fmacc(acc1, maybe_conj(c[k + 1], conj), lb[(k + 1) * RC + q]);
            }
            if (k < K) fmacc(acc0, maybe_conj(c[k], conj), lb[k * RC + q]);
            const T acc = acc0 + acc1;
            S *yo = ys + (j0 + j) * R + r0 + q;
            *yo = *yo + scale_by<S>(alpha, cvt<S>::from(acc));
        }
    }
}

// ------------------------------------------------------------------------------------------
// box GEMM: out[t] = (acc ? out[t] : 0) + sum_items opA(A_item) * in[src_item]
// opA: 0 A, 1 A^T, 2 A^H, 3 conj(A).  With a permutation pi (symmetry-reduced dense):
//   out[k] = sum_j opA(canonical)[pi_inv[k], j] * in[pi[j]]
// One workgroup per (target, RHS chunk of RB); threads over rows k.
// ------------------------------------------------------------------------------------------
constexpr int RB = 8;

template <class T>
__device__ inline T load_op(const T *A, int K, int row, int col, int opA) {
    switch (opA) {
    case 0: return A[row + (int64_t)K * col];
    case 1: return A[col + (int64_t)K * row];
    case 2: return conjv(A[col + (int64_t)K * row]);
    default: return conjv(A[row + (int64_t)K * col]);
    }
}

template <class T>
__global__ void boxgemm_kernel(int opA, const int32_t *__restrict__ targets, const int64_t *__restrict__ item_ptr,
                               const int32_t *__restrict__ item_src, const int32_t *__restrict__ item_mat,
                               const int32_t *__restrict__ item_sym, const int32_t *__restrict__ perms,
                               const int32_t *__restrict__ perms_inv, const T *__restrict__ mats, int K,
                               const T *__restrict__ in, T *__restrict__ out, int R, int acc_mode,
                               const uint8_t *__restrict__ acc_flags) {
    const int ti = blockIdx.x;
    const int t = targets[ti];
    const int r0 = blockIdx.y * RB;
    const int nr = min(RB, R - r0);
    const int64_t i0 = item_ptr[ti], i1 = item_ptr[ti + 1];
    const bool acc_in = acc_mode == 1 || (acc_mode == 2 && acc_flags[i0]);
    T *o = out + (int64_t)t * K * R;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        T acc[RB];
        for (int q = 0; q < RB; ++q) acc[q] = zero_v<T>();
        if (acc_in)
            for (int q = 0; q < nr; ++q) acc[q] = o[(int64_t)(r0 + q) * K + k];
        for (int64_t it = i0; it < i1; ++it) {
            const T *A = mats + (int64_t)item_mat[it] * K * K;
            const T *x = in + (int64_t)item_src[it] * K * R + (int64_t)r0 * K;
            const int sym = item_sym ? item_sym[it] : -1;
            const int32_t *pg = sym >= 0 ? perms + (int64_t)sym * K : nullptr;
            const int row = sym >= 0 ? perms_inv[(int64_t)sym * K + k] : k;
            for (int j = 0; j < K; ++j) {
                const T a = load_op(A, K, row, j, opA);
                const int jj = pg ? pg[j] : j;
                for (int q = 0; q < RB; ++q)
                    if (q < nr) fmacc(acc[q], a, x[(int64_t)q * K + jj]);
            }
        }
        for (int q = 0; q < nr; ++q) o[(int64_t)(r0 + q) * K + k] = acc[q];
    }
}

// ------------------------------------------------------------------------------------------
// Toeplitz modulation: pad -> 3-D DFT (fft_in) and 3-D inverse DFT -> crop (fft_out).
// Separable stages over LDS; lengths L = 2p-1 are odd, so each stage is a direct DFT of the
// non-zero / kept entries only.  `inverse` selects exp(+2 pi i/L) and the 1/L^3 scale.
// ------------------------------------------------------------------------------------------
template <class T>
__global__ void fft_in_kernel(int box0, const T *__restrict__ mom, int K, int p, int L, int R,
                              typename complex_of<T>::type *__restrict__ spec,
                              const cplx<double> *__restrict__ twiddle, bool inverse) {
    using C = typename complex_of<T>::type;
    using Rl = typename real_of<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *tw = reinterpret_cast<C *>(smem_raw);
    C *m = tw + L;
    C *b1 = m + p * p * p;
    C *b2 = b1 + L * p * p;
    const int slot = blockIdx.x;
    const int r = blockIdx.y;
    const int box = box0 + slot;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        cplx<double> w = twiddle[i];
        tw[i] = C{(Rl)w.re, (Rl)(inverse ? -w.im : w.im)};
    }
    const T *src = mom + ((int64_t)box * R + r) * K;
    for (int i = threadIdx.x; i < p * p * p; i += blockDim.x) m[i] = cvt<C>::from(src[i]);
    __syncthreads();
    // x: b1[fx + L*(iy + p*iz)] = sum_ix m[ix + p*(iy + p*iz)] w^(fx*ix)
    for (int i = threadIdx.x; i < L * p * p; i += blockDim.x) {
        const int fx = i % L, rest = i / L;
        C acc = zero_v<C>();
        for (int ix = 0; ix < p; ++ix) fmacc(acc, tw[(fx * ix) % L], m[ix + p * rest]);
        b1[i] = acc;
    }
    __syncthreads();
    // y: b2[fx + L*(fy + L*iz)] = sum_iy b1[fx + L*(iy + p*iz)] w^(fy*iy)
    for (int i = threadIdx.x; i < L * L * p; i += blockDim.x) {
        const int fx = i % L, fy = (i / L) % L, iz = i / (L * L);
        C acc = zero_v<C>();
        for (int iy = 0; iy < p; ++iy) fmacc(acc, tw[(fy * iy) % L], b1[fx + L * (iy + p * iz)]);
        b2[i] = acc;
    }
    __syncthreads();
    const int L3 = L * L * L;
    const Rl scale = inverse ? (Rl)(1.0 / (double)L3) : (Rl)1;
    C *dst = spec + ((int64_t)slot * R + r) * L3;
    for (int i = threadIdx.x; i < L3; i += blockDim.x) {
        const int fxy = i % (L * L), fz = i / (L * L);
        C acc = zero_v<C>();
        for (int iz = 0; iz < p; ++iz) fmacc(acc, tw[(fz * iz) % L], b2[fxy + L * L * iz]);
        dst[i] = acc * scale;
    }
}

template <class T>
__global__ void fft_out_kernel(const int32_t *__restrict__ boxes, const typename complex_of<T>::type *__restrict__ spec,
                               int K, int p, int L, int R, T *__restrict__ loc,
                               const cplx<double> *__restrict__ twiddle, bool inverse) {
    using C = typename complex_of<T>::type;
    using Rl = typename real_of<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *tw = reinterpret_cast<C *>(smem_raw);
    const int L3 = L * L * L;
    C *s = tw + L;
    C *a1 = s + L3;
    C *a2 = a1 + L * L * p;
    const int idx = blockIdx.x;
    const int r = blockIdx.y;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        cplx<double> w = twiddle[i];
        tw[i] = C{(Rl)w.re, (Rl)(inverse ? -w.im : w.im)};
    }
    const C *src = spec + ((int64_t)idx * R + r) * L3;
    for (int i = threadIdx.x; i < L3; i += blockDim.x) s[i] = src[i];
    __syncthreads();
    // z: a1[fxy + L*L*iz] = sum_fz s[fxy + L*L*fz] w^(fz*iz)
    for (int i = threadIdx.x; i < L * L * p; i += blockDim.x) {
        const int fxy = i % (L * L), iz = i / (L * L);
        C acc = zero_v<C>();
        for (int fz = 0; fz < L; ++fz) fmacc(acc, tw[(fz * iz) % L], s[fxy + L * L * fz]);
        a1[i] = acc;
    }
    __syncthreads();
    // y: a2[fx + L*(iy + p*iz)] = sum_fy a1[fx + L*(fy + L*iz)] w^(fy*iy)
    for (int i = threadIdx.x; i < L * p * p; i += blockDim.x) {
        const int fx = i % L, iy = (i / L) % p, iz = i / (L * p);
        C acc = zero_v<C>();
        for (int fy = 0; fy < L; ++fy) fmacc(acc, tw[(fy * iy) % L], a1[fx + L * (fy + L * iz)]);
        a2[i] = acc;
    }
    __syncthreads();
    const Rl scale = inverse ? (Rl)(1.0 / (double)L3) : (Rl)1;
    T *dst = loc + ((int64_t)boxes[idx] * R + r) * K;
    for (int i = threadIdx.x; i < p * p * p; i += blockDim.x) {
        const int ix = i % p, rest = i / p;
        C acc = zero_v<C>();
        for (int fx = 0; fx < L; ++fx) fmacc(acc, tw[(fx * ix) % L], a2[fx + L * rest]);
        dst[i] = cvt<T>::from(acc * scale);
    }
}

// ------------------------------------------------------------------------------------------
// Hadamard accumulation over a receiving box's partners:
//   spec_out[recv][r][f] = sum_items transform(D_op[perm? perm[f] : f]) * spec_in[src][r][f]
// ------------------------------------------------------------------------------------------
template <class C, class D>
__global__ void hadamard_kernel(const int64_t *__restrict__ item_ptr, const int32_t *__restrict__ item_src,
                                const int32_t *__restrict__ item_op, const int32_t *__restrict__ item_sym,
                                const int32_t *__restrict__ perms, const D *__restrict__ ops, int L3,
                                int src_slot_base, const C *__restrict__ spec_in, C *__restrict__ spec_out, int R,
                                bool conj) {
    const int recv = blockIdx.x;
    const int r = blockIdx.y;
    const int64_t i0 = item_ptr[recv], i1 = item_ptr[recv + 1];
    for (int f = threadIdx.x; f < L3; f += blockDim.x) {
        C acc = zero_v<C>();
        for (int64_t it = i0; it < i1; ++it) {
            const int sym = item_sym ? item_sym[it] : -1;
            const int ff = sym >= 0 ? perms[(int64_t)sym * L3 + f] : f;
            const C d = cvt<C>::from(ops[(int64_t)item_op[it] * L3 + ff]);
            const C x = spec_in[((int64_t)(item_src[it] - src_slot_base) * R + r) * L3 + f];
            fmacc(acc, maybe_conj(d, conj), x);
        }
        spec_out[((int64_t)recv * R + r) * L3 + f] = acc;
    }
}

__global__ void fill_zero_kernel(uint32_t *__restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------
#define DISPATCH_T(cd, ...)                                                   \
    switch (cd) {                                                             \
    case H2G_F64: { using T = double; __VA_ARGS__; } break;                   \
    case H2G_C64: { using T = cplx<double>; __VA_ARGS__; } break;             \
    case H2G_F32: { using T = float; __VA_ARGS__; } break;                    \
    case H2G_C32: { using T = cplx<float>; __VA_ARGS__; } break;             \
    default: throw Error{H2G_EINVAL, "unknown element type"};                 \
    }

#define DISPATCH_S(cd, ...)                                                   \
    switch (cd) {                                                             \
    case H2G_F64: { using S = double; __VA_ARGS__; } break;                   \
    case H2G_C64: { using S = cplx<double>; __VA_ARGS__; } break;             \
    default: throw Error{H2G_EINVAL, "core element type must be F64 or C64"}; \
    }

#define DISPATCH_X(dt, ...)                                                   \
    switch (dt) {                                                             \
    case H2G_F64: { using X = double; __VA_ARGS__; } break;                   \
    case H2G_C64: { using X = cplx<double>; __VA_ARGS__; } break;             \
    case H2G_F32: { using X = float; __VA_ARGS__; } break;                    \
    case H2G_C32: { using X = cplx<float>; __VA_ARGS__; } break;             \
    default: throw Error{H2G_EINVAL, "unknown vector element type"};          \
    }

// Diagonal (elementwise) couplings, one workgroup per receiving box, threads over (r, k): the
// coefficients of a pair are read once per column block from L2, the moments stream from HBM.
template <class T>
__global__ void diag_m2l_kernel(const int32_t *__restrict__ recv, const int64_t *__restrict__ item_ptr,
                                const int32_t *__restrict__ item_src, const int32_t *__restrict__ item_op,
                                const T *__restrict__ ops, int K, const T *__restrict__ in, T *__restrict__ out,
                                int R, bool conj) {
    const int ti = blockIdx.x;
    const int t = recv[ti];
    const int64_t i0 = item_ptr[ti], i1 = item_ptr[ti + 1];
    for (int e = threadIdx.x; e < R * K; e += blockDim.x) {
        const int r = e / K, k = e - r * K;
        T acc = zero_v<T>();
        for (int64_t it = i0; it < i1; ++it) {
            const T c = maybe_conj(ops[(int64_t)item_op[it] * K + k], conj);
            fmacc(acc, c, in[((int64_t)item_src[it] * R + r) * K + k]);
        }
        out[((int64_t)t * R + r) * K + k] = acc;
    }
}

// Zero fill by kernel (usable under stream capture, unlike a memset node).
void launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
    const int64_t n = (int64_t)(bytes / 4);
    if (n == 0) return;
    hipLaunchKernelGGL((fill_zero_kernel), grid1(n), TPB, 0, s, (uint32_t *)p, n);
    H2G_CHECK_LAUNCH();
}

void launch_diag_m2l(int cdtype, bool conj, int n_recv, const int32_t *recv, const int64_t *item_ptr,
                     const int32_t *item_src, const int32_t *item_op, const void *ops, int K, const void *in, void *out,
                     int R, hipStream_t s) {
    if (n_recv == 0) return;
    DISPATCH_T(cdtype, {
        hipLaunchKernelGGL((diag_m2l_kernel<T>), n_recv, TPB, 0, s, recv, item_ptr, item_src, item_op, (const T *)ops, K,
                           (const T *)in, (T *)out, R, conj);
    });
    H2G_CHECK_LAUNCH();
}

void launch_gather(int dtype_core, int dtype_x, const void *x, int64_t ldx, const int64_t *perm,
                   const int64_t *iperm, int64_t n, int R, void *xs, hipStream_t s) {
    if (n * R == 0) return;
    const dim3 tg((unsigned)((n + GT - 1) / GT), (unsigned)((R + GR - 1) / GR));
    DISPATCH_S(dtype_core, DISPATCH_X(dtype_x, {
        if (R >= 8 && iperm)
            hipLaunchKernelGGL((gather_tiled_kernel<S, X>), tg, TPB, 0, s, (const X *)x, ldx, iperm, n, R, (S *)xs);
        else
            hipLaunchKernelGGL((gather_kernel<S, X>), grid1(n * R), TPB, 0, s, (const X *)x, ldx, perm, n, R, (S *)xs);
    }));
    H2G_CHECK_LAUNCH();
}

void launch_scatter(int dtype_core, int dtype_y, const void *ys, const int64_t *perm, const int64_t *iperm, int64_t n,
                    int R, void *y, int64_t ldy, cplx<double> a, cplx<double> b, const uint8_t *mask, hipStream_t s) {
    if (n * R == 0) return;
    const dim3 tg((unsigned)((n + GT - 1) / GT), (unsigned)((R + GR - 1) / GR));
    DISPATCH_S(dtype_core, DISPATCH_X(dtype_y, {
        if (R >= 8 && iperm)
            hipLaunchKernelGGL((scatter_tiled_kernel<S, X>), tg, TPB, 0, s, (const S *)ys, iperm, n, R, (X *)y, ldy, a, b, mask);
        else
            hipLaunchKernelGGL((scatter_kernel<S, X>), grid1(n * R), TPB, 0, s, (const S *)ys, perm, n, R, (X *)y, ldy, a, b, mask);
    }));
    H2G_CHECK_LAUNCH();
}

bool launch_near_mfma(int op, const Core &c, const void *xs, void *ys, int R, float *x32, int near_cols,
                      hipStream_t s);

void launch_near(int dtype_core, int op, const Core &c, const void *xs, void *ys, int R, float *x32, int near_cols,
                 hipStream_t s) {
    if (launch_near_mfma(op, c, xs, ys, R, x32, near_cols, s)) return;
    const bool trans = op != H2G_OP_N;
    const int ntgt = trans ? c.n_near_rev_tgt : c.n_near_fwd_tgt;
    if (ntgt == 0) return;
    const int32_t *tgt = trans ? c.near_rev_tgt.as<int32_t>() : c.near_fwd_tgt.as<int32_t>();
    const int64_t *ptr = trans ? c.near_rev_ptr.as<int64_t>() : c.near_fwd_ptr.as<int64_t>();
    const int32_t *blk = trans ? c.near_rev_blk.as<int32_t>() : c.near_fwd_blk.as<int32_t>();
    const bool wide = R >= NRC;
    DISPATCH_S(dtype_core, {
        auto args = [&](auto kern) {
            hipLaunchKernelGGL((kern), ntgt, TPB, 0, s, tgt, ptr, blk, c.near_blk_rb.as<int64_t>(), c.near_blk_cb.as<int64_t>(),
                                      c.near_blk_nr.as<int32_t>(), c.near_blk_nc.as<int32_t>(),
                                      c.near_blk_off.as<int64_t>(), c.near_vals.as<S>(), (const S *)xs, (S *)ys, R,
                                      op == H2G_OP_C);
        };
        if (wide) {
            if (trans) args(near_wide_kernel<S, true>);
            else args(near_wide_kernel<S, false>);
        } else {
            if (trans) args(near_narrow_kernel<S, true>);
            else args(near_narrow_kernel<S, false>);
        }
    });
    H2G_CHECK_LAUNCH();
}

bool launch_p2m_mfma(int cdtype, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                     const void *coll, int K, const void *xs, int R, void *mom, hipStream_t s);
bool launch_l2p_mfma(int cdtype, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                     const void *coll, int K, const void *loc, int R, void *ys, double alpha, bool stage,
                     hipStream_t s);

void launch_p2m(int cdtype, bool conj, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                const void *coll, int K, const void *xs, int R, void *mom, hipStream_t s) {
    if (n == 0) return;
    if (launch_p2m_mfma(cdtype, leaves, n, begin, end, coll, K, xs, R, mom, s)) return;
    DISPATCH_T(cdtype, {
        const int pch = std::max(1, std::min(64, (int)(49152 / (sizeof(T) * (size_t)R))));
        hipLaunchKernelGGL((p2m_kernel<T>), n, TPB, sizeof(T) * pch * (size_t)R, s, leaves, begin, end, (const T *)coll, K,
                                                                 (const store_t<T> *)xs, R, (T *)mom, conj, pch);
    });
    H2G_CHECK_LAUNCH();
}

void launch_l2p(int cdtype, bool conj, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                const void *coll, int K, const void *loc, int R, void *ys, cplx<double> alpha, int avg_points,
                hipStream_t s) {
    if (n == 0) return;
    if (alpha.im == 0.0 &&
        launch_l2p_mfma(cdtype, leaves, n, begin, end, coll, K, loc, R, ys, alpha.re, avg_points > 16, s))
        return;
    DISPATCH_T(cdtype, {
        // RHS per LDS pass: all of them when K x R fits in 48 KB
        const int rc = std::max(1, std::min(R, (int)(49152 / (sizeof(T) * (size_t)K))));
        hipLaunchKernelGGL((l2p_kernel<T>), n, TPB, sizeof(T) * rc * (size_t)K, s, leaves, begin, end, (const T *)coll, K,
                                                                (const T *)loc, R, (store_t<T> *)ys, alpha, conj, rc);
    });
    H2G_CHECK_LAUNCH();
}

void launch_boxgemm(int cdtype, int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                    const int32_t *item_src, const int32_t *item_mat, const int32_t *item_sym,
                    const int32_t *perms, const int32_t *perms_inv, const void *mats, int K, const void *in,
                    void *out, int R, int acc_mode, const uint8_t *acc_flags, hipStream_t s) {
    if (n_targets == 0) return;
    dim3 grid(n_targets, (R + RB - 1) / RB);
    DISPATCH_T(cdtype, {
        hipLaunchKernelGGL((boxgemm_kernel<T>), grid, TPB, 0, s, opA, targets, item_ptr, item_src, item_mat, item_sym, perms,
                                               perms_inv, (const T *)mats, K, (const T *)in, (T *)out, R,
                                               acc_mode, acc_flags);
    });
    H2G_CHECK_LAUNCH();
}

void launch_fft_in(int cdtype, bool inverse, const int32_t *, int n, int box0, const void *mom, int K, int p,
                   int L, int R, void *spec, const void *twiddle, hipStream_t s) {
    if (n == 0) return;
    dim3 grid(n, R);
    DISPATCH_T(cdtype, {
        using C = typename complex_of<T>::type;
        size_t sh = sizeof(C) * (size_t)(L + p * p * p + L * p * p + L * L * p);
        hipLaunchKernelGGL((fft_in_kernel<T>), grid, TPB, sh, s, box0, (const T *)mom, K, p, L, R, (C *)spec,
                                               (const cplx<double> *)twiddle, inverse);
    });
    H2G_CHECK_LAUNCH();
}

void launch_hadamard(int cdtype, bool conj, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
                     const int32_t *item_op, const int32_t *item_sym, const int32_t *perms, const void *ops,
                     int L3, int src_slot_base, const void *spec_in, void *spec_out, int R, hipStream_t s) {
    if (n_recv == 0) return;
    dim3 grid(n_recv, R);
    switch (cdtype) {
    case H2G_F64:
    case H2G_C64:
        hipLaunchKernelGGL((hadamard_kernel<cplx<double>, cplx<double>>), grid, TPB, 0, s,
            item_ptr, item_src, item_op, item_sym, perms, (const cplx<double> *)ops, L3, src_slot_base,
            (const cplx<double> *)spec_in, (cplx<double> *)spec_out, R, conj);
        break;
    default:
        hipLaunchKernelGGL((hadamard_kernel<cplx<float>, cplx<float>>), grid, TPB, 0, s,
            item_ptr, item_src, item_op, item_sym, perms, (const cplx<float> *)ops, L3, src_slot_base,
            (const cplx<float> *)spec_in, (cplx<float> *)spec_out, R, conj);
    }
    H2G_CHECK_LAUNCH();
}

void launch_fft_out(int cdtype, bool inverse, const int32_t *boxes, int n, const void *spec, int K, int p, int L,
                    int R, void *loc, const void *twiddle, hipStream_t s) {
    if (n == 0) return;
    dim3 grid(n, R);
    DISPATCH_T(cdtype, {
        using C = typename complex_of<T>::type;
        size_t sh = sizeof(C) * (size_t)(L + L * L * L + L * L * p + L * p * p);
        hipLaunchKernelGGL((fft_out_kernel<T>), grid, TPB, sh, s, boxes, (const C *)spec, K, p, L, R, (T *)loc,
                                                (const cplx<double> *)twiddle, inverse);
    });
    H2G_CHECK_LAUNCH();
}


// ------------------------------------------------------------------------------------------
// Separable level transfers (tensor-product M2M / L2L, the reference's SeparableI2I,
// src/maps/leveltranslator.jl:168-265): T = Az (x) Ay (x) Ax with p x p factors, applied as
// three contractions in LDS.  factors: [n_levels*8][3][p][p] column-major per factor (x, y, z).
// opA: 0 T, 1 T^T, 2 T^H, 3 conj(T); out[t] = (acc ? out[t] : 0) + sum_items op(T_item) in[src].
// ------------------------------------------------------------------------------------------
namespace {

template <class T>
__device__ inline T sep_factor(const T *F, int p, int row, int col, int opA) {
    // F column-major p x p
    switch (opA) {
    case 0: return F[row + p * col];
    case 1: return F[col + p * row];
    case 2: return conjv(F[col + p * row]);
    default: return conjv(F[row + p * col]);
    }
}

// Jobs are (target, rhs) pairs, JB per workgroup (several targets per workgroup when R is
// small).  Each stage is a batch of length-p line transforms: a thread loads one line into
// registers and writes the transformed line (6 LDS reads per 36 FMAs at p = 6).
template <class T, int JB, int P>
__global__ void sep_transfer_kernel(int opA, int n_targets, const int32_t *__restrict__ targets,
                                    const int64_t *__restrict__ item_ptr, const int32_t *__restrict__ item_src,
                                    const int32_t *__restrict__ item_mat, const T *__restrict__ factors, int mat_base,
                                    const T *__restrict__ in, T *__restrict__ out, int R, int acc_mode,
                                    const uint8_t *__restrict__ acc_flags) {
    constexpr int p = P, K = P * P * P, LINES = P * P;
    // one dynamic LDS region, element buffers first (16-byte aligned base), then the factors of
    // the launch's 8 sectors with opA applied, then the job tables
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *t1 = reinterpret_cast<T *>(smem_raw);  // [JB][K]
    T *t2 = t1 + JB * K;                      // [JB][K]
    T *acc = t2 + JB * K;                     // [JB][K]
    T *fs = acc + JB * K;                     // [8 sectors][3 axes][p out][p in]
    int *s_item0 = reinterpret_cast<int *>(fs + 8 * 3 * p * p);
    int *s_nitem = s_item0 + JB, *s_tgt = s_nitem + JB, *s_r = s_tgt + JB, *s_maxp = s_r + JB;
    const int64_t job0 = (int64_t)blockIdx.x * JB;
    const int64_t njobs = (int64_t)n_targets * R;
    if (threadIdx.x < JB) {
        const int64_t job = job0 + threadIdx.x;
        if (job < njobs) {
            const int ti = (int)(job / R);
            s_tgt[threadIdx.x] = ti;
            s_r[threadIdx.x] = (int)(job - (int64_t)ti * R);
            s_item0[threadIdx.x] = (int)item_ptr[ti];
            s_nitem[threadIdx.x] = (int)(item_ptr[ti + 1] - item_ptr[ti]);
        } else {
            s_tgt[threadIdx.x] = -1;
            s_nitem[threadIdx.x] = 0;
        }
    }
    // the items of one launch use the factors of one level: matrices mat_base .. mat_base + 7
    for (int e = threadIdx.x; e < 8 * 3 * p * p; e += blockDim.x) {
        const int sa = e / (p * p), oi = e - sa * p * p;
        const int o = oi / p, i = oi - o * p;
        fs[e] = sep_factor(factors + ((int64_t)mat_base * 3 + sa) * p * p, p, o, i, opA);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int mx = 0;
        for (int j = 0; j < JB; ++j) mx = max(mx, s_nitem[j]);
        *s_maxp = mx;
    }
    for (int e = threadIdx.x; e < JB * K; e += blockDim.x) {
        const int j = e / K, k = e - j * K;
        T v = zero_v<T>();
        const int ti = s_tgt[j];
        if (ti >= 0) {
            const bool acc_in = acc_mode == 1 || (acc_mode == 2 && acc_flags[s_item0[j]]);
            if (acc_in) v = out[((int64_t)targets[ti] * R + s_r[j]) * K + k];
        }
        acc[e] = v;
    }
    __syncthreads();
    const int nmax = *s_maxp;
    for (int it = 0; it < nmax; ++it) {
        // x: input line (jy, jz) -> t1[kx + p*(jy + p*jz)]
        for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
            const int j = e / LINES, line = e - j * LINES;
            if (it >= s_nitem[j]) continue;
            const int64_t itg = s_item0[j] + it;
            const T *F = fs + (item_mat[itg] - mat_base) * 3 * p * p;
            const T *x = in + ((int64_t)item_src[itg] * R + s_r[j]) * K + p * line;
            T v[p];
#pragma unroll
            for (int i = 0; i < p; ++i) v[i] = x[i];
#pragma unroll
            for (int o = 0; o < p; ++o) {
                T a = zero_v<T>();
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
                t1[j * K + o + p * line] = a;
            }
        }
        __syncthreads();
        // y: line (kx, jz) -> t2[kx + p*(ky + p*jz)]
        for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
            const int j = e / LINES, line = e - j * LINES;
            if (it >= s_nitem[j]) continue;
            const int kx = line % p, jz = line / p;
            const T *F = fs + ((item_mat[s_item0[j] + it] - mat_base) * 3 + 1) * p * p;
            T v[p];
#pragma unroll
            for (int i = 0; i < p; ++i) v[i] = t1[j * K + kx + p * (i + p * jz)];
#pragma unroll
            for (int o = 0; o < p; ++o) {
                T a = zero_v<T>();
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
                t2[j * K + kx + p * (o + p * jz)] = a;
            }
        }
        __syncthreads();
        // z: line (kx, ky) -> acc[kx + p*(ky + p*kz)] +=
        for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
            const int j = e / LINES, line = e - j * LINES;
            if (it >= s_nitem[j]) continue;
            const T *F = fs + ((item_mat[s_item0[j] + it] - mat_base) * 3 + 2) * p * p;
            T v[p];
#pragma unroll
            for (int i = 0; i < p; ++i) v[i] = t2[j * K + line + LINES * i];
#pragma unroll
            for (int o = 0; o < p; ++o) {
                T a = acc[j * K + line + LINES * o];
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
                acc[j * K + line + LINES * o] = a;
            }
        }
        // the next x stage writes t1 only; the y stage after it is behind that stage's barrier
    }
    __syncthreads();
    for (int e = threadIdx.x; e < JB * K; e += blockDim.x) {
        const int j = e / K, k = e - j * K;
        const int ti = s_tgt[j];
        if (ti >= 0) out[((int64_t)targets[ti] * R + s_r[j]) * K + k] = acc[e];
    }
}
}  // namespace

// L2L grouped by parent: a job is (parent, RHS); the parent's local expansion is read once into
// LDS and every child c of the parent gets loc[c] = (acc_c ? loc[c] : 0) + T_{sector(c)}^op loc[p],
// the three line stages per child as in sep_transfer_kernel.
template <class T, int JB, int P>
__global__ void sep_l2l_kernel(int opA, int n_par, const int32_t *__restrict__ parents,
                               const int64_t *__restrict__ child_ptr, const int32_t *__restrict__ child_box,
                               const int32_t *__restrict__ child_mat, const uint8_t *__restrict__ child_acc,
                               const T *__restrict__ factors, int mat_base, T *__restrict__ loc, int R) {
    constexpr int p = P, K = P * P * P, LINES = P * P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *pin = reinterpret_cast<T *>(smem_raw);  // [JB][K] parent
    T *t1 = pin + JB * K;                      // [JB][K]
    T *t2 = t1 + JB * K;                       // [JB][K]
    T *fs = t2 + JB * K;                       // [8 sectors][3 axes][p out][p in]
    int *s_c0 = reinterpret_cast<int *>(fs + 8 * 3 * p * p);
    int *s_nc = s_c0 + JB, *s_r = s_nc + JB, *s_maxc = s_r + JB;
    const int64_t job0 = (int64_t)blockIdx.x * JB;
    const int64_t njobs = (int64_t)n_par * R;
    if (threadIdx.x < JB) {
        const int64_t job = job0 + threadIdx.x;
        if (job < njobs) {
            const int pi = (int)(job / R);
            s_r[threadIdx.x] = (int)(job - (int64_t)pi * R);
            s_c0[threadIdx.x] = (int)child_ptr[pi];
            s_nc[threadIdx.x] = (int)(child_ptr[pi + 1] - child_ptr[pi]);
        } else {
            s_r[threadIdx.x] = 0;
            s_c0[threadIdx.x] = 0;
            s_nc[threadIdx.x] = 0;
        }
    }
    for (int e = threadIdx.x; e < 8 * 3 * p * p; e += blockDim.x) {
        const int sa = e / (p * p), oi = e - sa * p * p;
        const int o = oi / p, i = oi - o * p;
        fs[e] = sep_factor(factors + ((int64_t)mat_base * 3 + sa) * p * p, p, o, i, opA);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < JB * K; e += blockDim.x) {
        const int j = e / K, k = e - j * K;
        const int64_t job = job0 + j;
        pin[e] = job < njobs ? loc[((int64_t)parents[job / R] * R + s_r[j]) * K + k] : zero_v<T>();
    }
    if (threadIdx.x == 0) {
        int mx = 0;
        for (int j = 0; j < JB; ++j) mx = max(mx, s_nc[j]);
        *s_maxc = mx;
    }
    __syncthreads();
    const int nmax = *s_maxc;
    for (int it = 0; it < nmax; ++it) {
        // x: parent line (jy, jz) -> t1[kx + p*(jy + p*jz)]
        for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
            const int j = e / LINES, line = e - j * LINES;
            if (it >= s_nc[j]) continue;
            const T *F = fs + (child_mat[s_c0[j] + it] - mat_base) * 3 * p * p;
            T v[p];
#pragma unroll
            for (int i = 0; i < p; ++i) v[i] = pin[j * K + p * line + i];
#pragma unroll
            for (int o = 0; o < p; ++o) {
                T a = zero_v<T>();
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
                t1[j * K + o + p * line] = a;
            }
        }
        __syncthreads();
        // y: line (kx, jz) -> t2[kx + p*(ky + p*jz)]
        for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
            const int j = e / LINES, line = e - j * LINES;
            if (it >= s_nc[j]) continue;
            const int kx = line % p, jz = line / p;
            const T *F = fs + ((child_mat[s_c0[j] + it] - mat_base) * 3 + 1) * p * p;
            T v[p];
#pragma unroll
            for (int i = 0; i < p; ++i) v[i] = t1[j * K + kx + p * (i + p * jz)];
#pragma unroll
            for (int o = 0; o < p; ++o) {
                T a = zero_v<T>();
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
                t2[j * K + kx + p * (o + p * jz)] = a;
            }
        }
        __syncthreads();
        // z: line (kx, ky) -> the child's expansion, accumulated when the child already holds
        // couplings
        for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
            const int j = e / LINES, line = e - j * LINES;
            if (it >= s_nc[j]) continue;
            const int item = s_c0[j] + it;
            const T *F = fs + ((child_mat[item] - mat_base) * 3 + 2) * p * p;
            const bool acc = child_acc[item] != 0;
            T *out = loc + ((int64_t)child_box[item] * R + s_r[j]) * K + line;
            T v[p];
#pragma unroll
            for (int i = 0; i < p; ++i) v[i] = t2[j * K + line + LINES * i];
#pragma unroll
            for (int o = 0; o < p; ++o) {
                T a = acc ? out[LINES * o] : zero_v<T>();
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
                out[LINES * o] = a;
            }
        }
        // the next x stage writes t1 only; the y stage after it is behind that stage's barrier
    }
}

// threads per workgroup of the shared-factor L2L for FP32 (8 jobs per workgroup), measured on the
// N = 2M, p = 6 map (levels 5 + 6): 256 threads 1.97 ms, 512 threads 1.57 ms; 16 jobs on 512 or
// 1024 threads and 4 jobs on 256 threads were slower.  The shared-factor M2M is fastest at 256.
constexpr int SEP_L2L_NT4 = 512;
// L2L grouped by parent when the Kronecker factors are shared between sectors (the x factor
// depends on the child's x bit only, and so on: interpolation transfers).  The children's
// transforms then share stages: 2 x-line variants (child x bit), 4 y variants (x, y bits) and 8
// z variants, 14 line batches instead of 24, all in one pass with two barriers.  The children's
// current expansions (accumulated when they already hold couplings) are loaded at the start and
// used by the last stage, so their latency overlaps the first two stages.
template <class T, int JB, int P, int NT>
__global__ void __launch_bounds__(NT) sep_l2l_shared_kernel(int opA, int n_par, const int32_t *__restrict__ parents,
                                                            const int64_t *__restrict__ child_ptr,
                                                            const int32_t *__restrict__ child_box,
                                                            const int32_t *__restrict__ child_mat,
                                                            const uint8_t *__restrict__ child_acc,
                                                            const T *__restrict__ factors, int mat_base,
                                                            T *__restrict__ loc, int R) {
    constexpr int p = P, K = P * P * P, LINES = P * P;
    constexpr int ZL = JB * 8 * LINES, ZPT = (ZL + NT - 1) / NT;  // z lines, per thread
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *pin = reinterpret_cast<T *>(smem_raw);  // [JB][K] parent
    T *t1 = pin + JB * K;                      // [JB][2][K] x stage per child x bit
    T *t2 = t1 + JB * 2 * K;                   // [JB][4][K] y stage per child (x, y) bits
    T *fs = t2 + JB * 4 * K;                   // [8 sectors][3 axes][p out][p in]
    int *s_box = reinterpret_cast<int *>(fs + 8 * 3 * p * p);  // [JB][8] child box of a sector, or -1
    int *s_acc = s_box + JB * 8;                                // [JB][8]
    int *s_r = s_acc + JB * 8;                                  // [JB]
    const int64_t job0 = (int64_t)blockIdx.x * JB;
    const int64_t njobs = (int64_t)n_par * R;
    for (int e = threadIdx.x; e < JB * 8; e += blockDim.x) {
        s_box[e] = -1;
        s_acc[e] = 0;
    }
    for (int e = threadIdx.x; e < 8 * 3 * p * p; e += blockDim.x) {
        const int sa = e / (p * p), oi = e - sa * p * p;
        const int o = oi / p, i = oi - o * p;
        fs[e] = sep_factor(factors + ((int64_t)mat_base * 3 + sa) * p * p, p, o, i, opA);
    }
    __syncthreads();
    if (threadIdx.x < JB) {
        const int64_t job = job0 + threadIdx.x;
        s_r[threadIdx.x] = 0;
        if (job < njobs) {
            const int pi = (int)(job / R);
            s_r[threadIdx.x] = (int)(job - (int64_t)pi * R);
            for (int64_t c = child_ptr[pi]; c < child_ptr[pi + 1]; ++c) {
                const int sec = child_mat[c] - mat_base;
                s_box[threadIdx.x * 8 + sec] = child_box[c];
                s_acc[threadIdx.x * 8 + sec] = child_acc[c];
            }
        }
    }
    // (the job tables above are written by threads 0 .. JB-1 in this same phase: recompute r here)
    for (int e = threadIdx.x; e < JB * K; e += blockDim.x) {
        const int j = e / K, k = e - j * K;
        const int64_t job = job0 + j;
        pin[e] = job < njobs ? loc[(int64_t)parents[job / R] * R * K + (job % R) * K + k] : zero_v<T>();
    }
    __syncthreads();
    // the children's current values for the last stage, loaded now
    T pre[ZPT][p];
#pragma unroll
    for (int u = 0; u < ZPT; ++u) {
        const int e = threadIdx.x + u * NT;
        const int js = e / LINES, line = e - js * LINES;  // js = j * 8 + sector
        const int box = e < ZL ? s_box[js] : -1;
        const bool acc = box >= 0 && s_acc[js] != 0;
        // (the column table is read only for lines of the workgroup's jobs: past ZL, js / 8 >= JB)
        const T *src = loc + ((int64_t)(box < 0 ? 0 : box) * R + (e < ZL ? s_r[js >> 3] : 0)) * K + line;
#pragma unroll
        for (int o = 0; o < p; ++o) pre[u][o] = acc ? src[LINES * o] : zero_v<T>();
    }
    // x: parent line (jy, jz) -> t1[j][bx][kx + p*(jy + p*jz)]
    for (int e = threadIdx.x; e < JB * 2 * LINES; e += blockDim.x) {
        const int jb = e / LINES, line = e - jb * LINES;  // jb = j * 2 + bx
        const int j = jb >> 1, bx = jb & 1;
        const T *F = fs + (bx * 3 + 0) * p * p;
        T v[p];
#pragma unroll
        for (int i = 0; i < p; ++i) v[i] = pin[j * K + p * line + i];
#pragma unroll
        for (int o = 0; o < p; ++o) {
            T a = zero_v<T>();
#pragma unroll
            for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
            t1[jb * K + o + p * line] = a;
        }
    }
    __syncthreads();
    // y: line (kx, jz) of t1[j][bx] -> t2[j][bxy][kx + p*(ky + p*jz)], bxy = bx + 2 by
    for (int e = threadIdx.x; e < JB * 4 * LINES; e += blockDim.x) {
        const int jb = e / LINES, line = e - jb * LINES;  // jb = j * 4 + bxy
        const int j = jb >> 2, bxy = jb & 3;
        const int kx = line % p, jz = line / p;
        const T *F = fs + (bxy * 3 + 1) * p * p;
        const T *src = t1 + (j * 2 + (bxy & 1)) * K;
        T v[p];
#pragma unroll
        for (int i = 0; i < p; ++i) v[i] = src[kx + p * (i + p * jz)];
#pragma unroll
        for (int o = 0; o < p; ++o) {
            T a = zero_v<T>();
#pragma unroll
            for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
            t2[jb * K + kx + p * (o + p * jz)] = a;
        }
    }
    __syncthreads();
    // z: line (kx, ky) of t2[j][s % 4] -> child expansion of sector s
#pragma unroll
    for (int u = 0; u < ZPT; ++u) {
        const int e = threadIdx.x + u * NT;
        if (e >= ZL) break;
        const int js = e / LINES, line = e - js * LINES;
        const int box = s_box[js];
        if (box < 0) continue;
        const int j = js >> 3, sec = js & 7;
        const T *F = fs + (sec * 3 + 2) * p * p;
        const T *src = t2 + (j * 4 + (sec & 3)) * K + line;
        T v[p];
#pragma unroll
        for (int i = 0; i < p; ++i) v[i] = src[LINES * i];
        T *out = loc + ((int64_t)box * R + s_r[j]) * K + line;
#pragma unroll
        for (int o = 0; o < p; ++o) {
            T a = pre[u][o];
#pragma unroll
            for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], v[i]);
            out[LINES * o] = a;
        }
    }
}

// M2M of all children of a parent in one pass, for shared factors: the parent's moment is
// sum_s Cz[sz] (x) Cy[sy] (x) Cx[sx] m_s = sum_sz Cz[sz] (sum_sy Cy[sy] (sum_sx Cx[sx] m_s)), so the
// x stage reads the two children of each (y, z) bit pair and sums them (4 results per parent),
// the y stage sums over y bits (2 results), the z stage over z bits: 14 line batches, two
// barriers.  The children's lines are loaded at the start, all at once.  A missing child (empty
// sector) contributes zero.  The parent's moment is overwritten.
template <class T, int JB, int P, int NT>
__global__ void __launch_bounds__(NT) sep_m2m_shared_kernel(int opA, int n_targets, const int32_t *__restrict__ targets,
                                                            const int64_t *__restrict__ item_ptr,
                                                            const int32_t *__restrict__ item_src,
                                                            const int32_t *__restrict__ item_mat,
                                                            const T *__restrict__ factors, int mat_base,
                                                            T *__restrict__ mom, int R) {
    constexpr int p = P, K = P * P * P, LINES = P * P;
    constexpr int XL = JB * 4 * LINES, XPT = (XL + NT - 1) / NT;  // x items, per thread
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *w = reinterpret_cast<T *>(smem_raw);  // [JB][4 (y, z bits)][K]
    T *v = w + JB * 4 * K;                   // [JB][2 (z bit)][K]
    T *fs = v + JB * 2 * K;                  // [8 sectors][3 axes][p out][p in]
    int *s_box = reinterpret_cast<int *>(fs + 8 * 3 * p * p);  // [JB][8] child box of a sector, or -1
    const int64_t job0 = (int64_t)blockIdx.x * JB;
    const int64_t njobs = (int64_t)n_targets * R;
    for (int e = threadIdx.x; e < JB * 8; e += blockDim.x) s_box[e] = -1;
    for (int e = threadIdx.x; e < 8 * 3 * p * p; e += blockDim.x) {
        const int sa = e / (p * p), oi = e - sa * p * p;
        const int o = oi / p, i = oi - o * p;
        fs[e] = sep_factor(factors + ((int64_t)mat_base * 3 + sa) * p * p, p, o, i, opA);
    }
    __syncthreads();
    if (threadIdx.x < JB) {
        const int64_t job = job0 + threadIdx.x;
        if (job < njobs) {
            const int ti = (int)(job / R);
            for (int64_t it = item_ptr[ti]; it < item_ptr[ti + 1]; ++it)
                s_box[threadIdx.x * 8 + (item_mat[it] - mat_base)] = item_src[it];
        }
    }
    __syncthreads();
    // x: child lines (jy, jz) of sectors (0, syz) and (1, syz) -> w[j][syz][kx + p*(jy + p*jz)]
    T in[XPT][2][p];
#pragma unroll
    for (int u = 0; u < XPT; ++u) {
        const int e = threadIdx.x + u * NT;
        const int jq = e / LINES, line = e - jq * LINES;  // jq = j * 4 + syz
        const int j = jq >> 2, syz = jq & 3;
        const int r = (int)((job0 + j) % R);
#pragma unroll
        for (int sx = 0; sx < 2; ++sx) {
            const int box = e < XL ? s_box[j * 8 + sx + 2 * syz] : -1;
            const T *src = mom + ((int64_t)(box < 0 ? 0 : box) * R + r) * K + p * line;
#pragma unroll
            for (int i = 0; i < p; ++i) in[u][sx][i] = box >= 0 ? src[i] : zero_v<T>();
        }
    }
#pragma unroll
    for (int u = 0; u < XPT; ++u) {
        const int e = threadIdx.x + u * NT;
        if (e >= XL) break;
        const int jq = e / LINES, line = e - jq * LINES;
#pragma unroll
        for (int o = 0; o < p; ++o) {
            T a = zero_v<T>();
#pragma unroll
            for (int sx = 0; sx < 2; ++sx) {
                const T *F = fs + (sx * 3 + 0) * p * p;
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a, F[o * p + i], in[u][sx][i]);
            }
            w[jq * K + o + p * line] = a;
        }
    }
    __syncthreads();
    // y: lines (kx, jz) of w[j][sy + 2 sz] -> v[j][sz][kx + p*(ky + p*jz)], summed over sy
    for (int e = threadIdx.x; e < JB * 2 * LINES; e += blockDim.x) {
        const int jz2 = e / LINES, line = e - jz2 * LINES;  // jz2 = j * 2 + sz
        const int j = jz2 >> 1, sz = jz2 & 1;
        const int kx = line % p, jz = line / p;
        T a[p];
#pragma unroll
        for (int o = 0; o < p; ++o) a[o] = zero_v<T>();
#pragma unroll
        for (int sy = 0; sy < 2; ++sy) {
            const T *F = fs + ((2 * sy) * 3 + 1) * p * p;
            const T *src = w + (j * 4 + sy + 2 * sz) * K;
            T x[p];
#pragma unroll
            for (int i = 0; i < p; ++i) x[i] = src[kx + p * (i + p * jz)];
#pragma unroll
            for (int o = 0; o < p; ++o)
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a[o], F[o * p + i], x[i]);
        }
#pragma unroll
        for (int o = 0; o < p; ++o) v[jz2 * K + kx + p * (o + p * jz)] = a[o];
    }
    __syncthreads();
    // z: lines (kx, ky) of v[j][sz] -> the parent's moment, summed over sz
    for (int e = threadIdx.x; e < JB * LINES; e += blockDim.x) {
        const int j = e / LINES, line = e - j * LINES;
        const int64_t job = job0 + j;
        if (job >= njobs) continue;
        T a[p];
#pragma unroll
        for (int o = 0; o < p; ++o) a[o] = zero_v<T>();
#pragma unroll
        for (int sz = 0; sz < 2; ++sz) {
            const T *F = fs + ((4 * sz) * 3 + 2) * p * p;
            const T *src = v + (j * 2 + sz) * K + line;
            T x[p];
#pragma unroll
            for (int i = 0; i < p; ++i) x[i] = src[LINES * i];
#pragma unroll
            for (int o = 0; o < p; ++o)
#pragma unroll
                for (int i = 0; i < p; ++i) fmacc(a[o], F[o * p + i], x[i]);
        }
        T *out = mom + ((int64_t)targets[job / R] * R + job % R) * K + line;
#pragma unroll
        for (int o = 0; o < p; ++o) out[LINES * o] = a[o];
    }
}

template <class T, int P>
static void sep_m2m_p(int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr, const int32_t *item_src,
                      const int32_t *item_mat, const void *factors, int mat_base, void *mom, int R, hipStream_t s) {
    constexpr int JS = 32 / (int)sizeof(T);  // jobs per workgroup: 8 (FP32), 4 (FP64)
    constexpr int NT = TPB;
    constexpr int K = P * P * P;
    const int64_t njobs = (int64_t)n_targets * R;
    const size_t sh = sizeof(T) * (size_t)(6 * JS * K + 8 * 3 * P * P) + sizeof(int) * (size_t)(8 * JS);
    hipLaunchKernelGGL((sep_m2m_shared_kernel<T, JS, P, NT>), (unsigned)((njobs + JS - 1) / JS), NT, sh, s,
        opA, n_targets, targets, item_ptr, item_src, item_mat, (const T *)factors, mat_base, (T *)mom, R);
}

template <class T, int P>
static void sep_l2l_p(int opA, int n_par, const int32_t *parents, const int64_t *child_ptr, const int32_t *child_box,
                      const int32_t *child_mat, const uint8_t *child_acc, const void *factors, int mat_base, void *loc,
                      int R, bool shared, hipStream_t s) {
    constexpr int JB = 64 / (int)sizeof(T);
    constexpr int K = P * P * P;
    const int64_t njobs = (int64_t)n_par * R;
    if (shared) {
        constexpr int JS = 32 / (int)sizeof(T);  // jobs per workgroup: 8 (FP32), 4 (FP64)
        constexpr int NT = sizeof(T) == 4 ? SEP_L2L_NT4 : TPB;
        const size_t shs = sizeof(T) * (size_t)(7 * JS * K + 8 * 3 * P * P) + sizeof(int) * (size_t)(17 * JS);
        hipLaunchKernelGGL((sep_l2l_shared_kernel<T, JS, P, NT>), (unsigned)((njobs + JS - 1) / JS), NT, shs, s,
            opA, n_par, parents, child_ptr, child_box, child_mat, child_acc, (const T *)factors, mat_base, (T *)loc, R);
        return;
    }
    const size_t sh = sizeof(T) * (size_t)(3 * JB * K + 8 * 3 * P * P) + sizeof(int) * (size_t)(3 * JB + 1);
    hipLaunchKernelGGL((sep_l2l_kernel<T, JB, P>), (unsigned)((njobs + JB - 1) / JB), TPB, sh, s,
        opA, n_par, parents, child_ptr, child_box, child_mat, child_acc, (const T *)factors, mat_base, (T *)loc, R);
}

// One workgroup per box of the list: RB rows at a time pass through LDS and return permuted.
template <class Rl>
__global__ void cplx_planar_kernel(const int32_t *__restrict__ boxes, Rl *__restrict__ pool, int R, int K, int RB,
                                   bool to_planar) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Rl *buf = reinterpret_cast<Rl *>(smem_raw);
    const int K2 = 2 * K;
    Rl *base = pool + (int64_t)boxes[blockIdx.x] * R * K2;
    for (int r0 = 0; r0 < R; r0 += RB) {
        const int n2 = min(RB, R - r0) * K2;
        Rl *rows = base + (int64_t)r0 * K2;
        for (int e = threadIdx.x; e < n2; e += (int)blockDim.x) buf[e] = rows[e];
        __syncthreads();
        for (int e = threadIdx.x; e < n2; e += (int)blockDim.x) {
            const int row = e / K2, q = e - row * K2;
            // planar q = (part, k) <- interleaved 2 k + part; interleaved q = (k, part) <- part K + k
            const int src = to_planar ? 2 * (q % K) + q / K : (q & 1) * K + (q >> 1);
            rows[e] = buf[row * K2 + src];
        }
        __syncthreads();
    }
}

__global__ void real_part_kernel(const double *__restrict__ c, double *__restrict__ re, float *__restrict__ re32,
                                 int64_t n, int *__restrict__ nonreal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = c[2 * i];
    re[i] = r;
    if (re32) re32[i] = (float)r;
    if (c[2 * i + 1] != 0.0) atomicOr(nonreal, 1);
}

void make_real_collection(Side &S, const Core &c) {
    if (c.dtype != H2G_C64 || !S.sep_cplx || c.alpha.im != 0.0 || S.coll_real || !S.coll.ptr) return;
    const int64_t n = (int64_t)(S.coll.bytes / (2 * sizeof(double)));
    DevBuf flag;
    flag.alloc(sizeof(int));
    H2G_HIP(hipMemset(flag.ptr, 0, sizeof(int)));
    S.coll_re.alloc((size_t)n * sizeof(double));
    if (S.coll32.ptr) S.coll32_re.alloc((size_t)n * sizeof(float));
    if (n > 0)
        hipLaunchKernelGGL(real_part_kernel, (unsigned)((n + TPB - 1) / TPB), TPB, 0, 0, S.coll.as<double>(),
                           S.coll_re.as<double>(), S.coll32_re.ptr ? S.coll32_re.as<float>() : nullptr, n,
                           flag.as<int>());
    H2G_CHECK_LAUNCH();
    int nonreal = 0;
    H2G_HIP(hipMemcpy(&nonreal, flag.ptr, sizeof(int), hipMemcpyDeviceToHost));
    if (nonreal) {
        S.coll_re.release();
        S.coll32_re.release();
        return;
    }
    S.coll.release();
    S.coll32.release();
    S.coll_real = true;
}

void launch_cplx_planar(int cdtype, const int32_t *boxes, int n, void *pool, int R, int K, bool to_planar,
                        hipStream_t s) {
    if (n == 0) return;
    auto go = [&](auto tv) {
        using Rl = decltype(tv);
        const int RB = std::max(1, std::min(R, (int)(16384 / (2 * (size_t)K * sizeof(Rl)))));
        hipLaunchKernelGGL((cplx_planar_kernel<Rl>), n, TPB, sizeof(Rl) * (size_t)RB * 2 * K, s, boxes, (Rl *)pool, R,
                           K, RB, to_planar);
    };
    if (cdtype == H2G_C64)
        go(0.0);
    else if (cdtype == H2G_C32)
        go(0.0f);
    else
        throw Error{H2G_EINVAL, "planar rows are for complex expansions"};
    H2G_CHECK_LAUNCH();
}

void launch_sep_m2m(int cdtype, int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                    const int32_t *item_src, const int32_t *item_mat, const void *factors, int mat_base, int p,
                    void *mom, int R, hipStream_t s) {
    if (n_targets == 0) return;
    DISPATCH_T(cdtype, {
        switch (p) {
        case 2: sep_m2m_p<T, 2>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, mom, R, s); break;
        case 3: sep_m2m_p<T, 3>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, mom, R, s); break;
        case 4: sep_m2m_p<T, 4>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, mom, R, s); break;
        case 5: sep_m2m_p<T, 5>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, mom, R, s); break;
        case 6: sep_m2m_p<T, 6>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, mom, R, s); break;
        default: throw Error{H2G_EUNSUPPORTED, "separable transfers support 2 <= p <= 6"};
        }
    });
    H2G_CHECK_LAUNCH();
}

void launch_sep_l2l(int cdtype, int opA, int n_parents, const int32_t *parents, const int64_t *child_ptr,
                    const int32_t *child_box, const int32_t *child_mat, const uint8_t *child_acc, const void *factors,
                    int mat_base, int p, bool shared, void *loc, int R, hipStream_t s) {
    if (n_parents == 0) return;
    DISPATCH_T(cdtype, {
        switch (p) {
        case 2: sep_l2l_p<T, 2>(opA, n_parents, parents, child_ptr, child_box, child_mat, child_acc, factors, mat_base, loc, R, shared, s); break;
        case 3: sep_l2l_p<T, 3>(opA, n_parents, parents, child_ptr, child_box, child_mat, child_acc, factors, mat_base, loc, R, shared, s); break;
        case 4: sep_l2l_p<T, 4>(opA, n_parents, parents, child_ptr, child_box, child_mat, child_acc, factors, mat_base, loc, R, shared, s); break;
        case 5: sep_l2l_p<T, 5>(opA, n_parents, parents, child_ptr, child_box, child_mat, child_acc, factors, mat_base, loc, R, shared, s); break;
        case 6: sep_l2l_p<T, 6>(opA, n_parents, parents, child_ptr, child_box, child_mat, child_acc, factors, mat_base, loc, R, shared, s); break;
        default: throw Error{H2G_EUNSUPPORTED, "separable transfers support 2 <= p <= 6"};
        }
    });
    H2G_CHECK_LAUNCH();
}

template <class T, int P>
static void sep_transfer_p(int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                           const int32_t *item_src, const int32_t *item_mat, const void *factors, int mat_base,
                           const void *in, void *out, int R, int acc_mode, const uint8_t *acc_flags, hipStream_t s) {
    // jobs per workgroup: three K-long LDS buffers per job, 41 KB at p = 6 for every type
    constexpr int JB = 64 / (int)sizeof(T);
    constexpr int K = P * P * P;
    const int64_t njobs = (int64_t)n_targets * R;
    const size_t sh = sizeof(T) * (size_t)(3 * JB * K + 8 * 3 * P * P) + sizeof(int) * (size_t)(4 * JB + 1);
    hipLaunchKernelGGL((sep_transfer_kernel<T, JB, P>), (unsigned)((njobs + JB - 1) / JB), TPB, sh, s,
        opA, n_targets, targets, item_ptr, item_src, item_mat, (const T *)factors, mat_base, (const T *)in,
        (T *)out, R, acc_mode, acc_flags);
}

void launch_sep_transfer(int cdtype, int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                         const int32_t *item_src, const int32_t *item_mat, const void *factors, int mat_base, int p,
                         const void *in, void *out, int R, int acc_mode, const uint8_t *acc_flags, hipStream_t s) {
    if (n_targets == 0) return;
    DISPATCH_T(cdtype, {
        switch (p) {
        case 2: sep_transfer_p<T, 2>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, in, out, R, acc_mode, acc_flags, s); break;
        case 3: sep_transfer_p<T, 3>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, in, out, R, acc_mode, acc_flags, s); break;
        case 4: sep_transfer_p<T, 4>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, in, out, R, acc_mode, acc_flags, s); break;
        case 5: sep_transfer_p<T, 5>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, in, out, R, acc_mode, acc_flags, s); break;
        case 6: sep_transfer_p<T, 6>(opA, n_targets, targets, item_ptr, item_src, item_mat, factors, mat_base, in, out, R, acc_mode, acc_flags, s); break;
        default: throw Error{H2G_EUNSUPPORTED, "separable transfers support 2 <= p <= 6"};
        }
    });
    H2G_CHECK_LAUNCH();
}

}  // namespace h2g

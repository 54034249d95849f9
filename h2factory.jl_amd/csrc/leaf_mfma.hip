// MFMA kernels of libh2g for real maps: the near field as row-block products on FP64 MFMA, the
// leaf expansions (P2M, L2P) on FP64 MFMA, and the lane-layout probes of the MFMA instructions.
#include <algorithm>
#include <type_traits>

#include "h2g_internal.h"

namespace h2g {

constexpr int TPB = 256;
// threads per workgroup of the MFMA leaf kernels (the waves of a workgroup take its row tiles):
// 512 measured 0.05-0.1 ms faster than 256 per level on the N = 2M, p = 6 map, 1024 slower for P2M
constexpr int L2P_NT = 512, P2M_NT = 512;

// ------------------------------------------------------------------------------------------
// Near field on FP64 MFMA (real maps, R >= MFMA_MIN_R), one row-block product per task:
//   Y (nrow x R) = A (nrow x ncol) X (ncol x R)
// with v_mfma_f64_16x16x4f64.  A's columns are the source points of all near blocks of the
// output leaf, concatenated (one NearCol entry each), so the contraction is padded to a multiple
// of 4 KU once per leaf instead of once per block.  Operand lane layouts of the 16x16x4 FP64
// instruction (checked by the layout probe below on gfx950): A[m][k] in lane m + 16k,
// B[k][n] in lane n + 16k, D[m][n] in lane n + 16(m % 4), element m / 4.
// ------------------------------------------------------------------------------------------
// narrowest batch for the MFMA near-field and leaf kernels (16-column tiles, unused columns
// masked): they beat the vector kernels at every width (N = 2M, p = 6, mixed products: 1 column
// 14.5 -> 11.6 ms, 2 columns 22.9 -> 15.1 ms, 3 columns 37.2 -> 24.4 ms)
constexpr int MFMA_MIN_R = 1;

namespace {
typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ inline int mf_a_row(int lane) { return lane & 15; }
__device__ inline int mf_a_k(int lane) { return lane >> 4; }
__device__ inline int mf_d_row(int lane, int v) { return (lane >> 4) + 4 * v; }

// One wave per task, 4 independent waves per workgroup; tasks [t0, t1) all have RT row tiles.
// The wave holds RT x NCT accumulator tiles, so every source value x[j][c] it loads feeds RT
// MFMAs and every matrix value NCT.  A batch is KU k-steps (KB = 4 KU columns).  Lane l holds the
// column-table entry l % KB of a batch, loaded two batches ahead of its use; a lane takes the
// entries of its own columns from the owning lanes (ds_bpermute), so the table costs 2 x 4
// registers rather than 4 per column in flight.  Operands are loaded one batch ahead of the
// MFMAs.  Every load reads a clamped (valid) address and out-of-range rows, columns and RHS are
// multiplied by zero, so the loop has no branches and the accumulators stay in the MFMA
// registers.  Workgroups are spread so that each XCD (dispatched round robin) works on a
// contiguous range of tasks, whose source points overlap in its L2.
template <bool TRANS, int RT, int KU, int NCT, class V>
__global__ void __launch_bounds__(TPB, 2) near_task_kernel(const NearTask *__restrict__ tasks, int64_t t0, int64_t t1,
                                                        const NearCol *__restrict__ cols, const V *__restrict__ vals,
                                                        const double *__restrict__ xs, double *__restrict__ ys, int R) {
    constexpr int KB = 4 * KU;
    static_assert(KB <= 64, "a batch's table entries fit one wave");
    int wg = blockIdx.x;
    const int n8 = gridDim.x & ~7;
    if (wg < n8) wg = (wg & 7) * (n8 >> 3) + (wg >> 3);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t ti = t0 + (int64_t)wg * (TPB / 64) + wave;
    if (ti >= t1) return;
    const NearTask tk = tasks[ti];
    const int nrow = tk.nrow, ncol = tk.ncol;
    const NearCol *__restrict__ cl = cols + tk.col0;
    const int lane = threadIdx.x & 63;
    const int ar = mf_a_row(lane), ak = mf_a_k(lane);
    const int c0 = blockIdx.y * NCT * 16;
    const int tl = lane & (KB - 1);
    int arow[RT];
    double rmask[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int r = 16 * rt + ar;
        arow[rt] = tk.row0 + min(r, nrow - 1);
        rmask[rt] = r < nrow ? 1.0 : 0.0;
    }
    int xcol[NCT];
    double cmask[NCT];
#pragma unroll
    for (int g = 0; g < NCT; ++g) {
        const int col = c0 + 16 * g + ar;
        xcol[g] = min(col, R - 1);
        cmask[g] = col < R ? 1.0 : 0.0;
    }
    // raw operands of the batch in flight and their column masks; the masks are applied when the
    // batch is used, so that nothing waits for the loads before the next batch's MFMAs
    double a[KU][RT], b[KU][NCT], km[KU];
    auto load_ops = [&](int k0, const NearCol &ev) {
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int src = (4 * u + ak) << 2;  // byte address of the lane holding this column
            const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)ev.aoff);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(ev.aoff >> 32));
            const int xr = __builtin_amdgcn_ds_bpermute(src, ev.xrow);
            const int64_t rs = TRANS ? (int64_t)__builtin_amdgcn_ds_bpermute(src, ev.rstride) : 1;
            km[u] = (k0 + 4 * u + ak < ncol) ? 1.0 : 0.0;
            const V *__restrict__ ap = vals + (int64_t)(((uint64_t)hi << 32) | lo);
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) a[u][rt] = (double)ap[arow[rt] * rs];
            const double *__restrict__ xp = xs + (int64_t)xr * R;
#pragma unroll
            for (int g = 0; g < NCT; ++g) b[u][g] = xp[xcol[g]];
        }
    };
    double4_t acc[RT][NCT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int g = 0; g < NCT; ++g) acc[rt][g] = double4_t{0.0, 0.0, 0.0, 0.0};
    const int nb = (ncol + KB - 1) / KB;
    NearCol tb0 = cl[min(tl, ncol - 1)], tb1 = cl[min(KB + tl, ncol - 1)];
    load_ops(0, tb0);
    for (int ib = 0; ib < nb; ++ib) {
        double ac[KU][RT], bc[KU][NCT];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) ac[u][rt] = a[u][rt] * (rmask[rt] * km[u]);
#pragma unroll
            for (int g = 0; g < NCT; ++g) bc[u][g] = b[u][g] * (cmask[g] * km[u]);
        }
        const NearCol tb2 = cl[min((ib + 2) * KB + tl, ncol - 1)];
        load_ops((ib + 1) * KB, tb1);
        // keep the next batch's loads ahead of this batch's MFMAs
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < KU; ++u)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int g = 0; g < NCT; ++g)
                    acc[rt][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[u][rt], bc[u][g], acc[rt][g], 0, 0, 0);
        tb1 = tb2;
    }
    double *__restrict__ yo = ys + (tk.out0 + tk.row0) * R;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int g = 0; g < NCT; ++g) {
            const int col = c0 + 16 * g + ar;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * rt + mf_d_row(lane, v);
                if (row < nrow && col < R) yo[(int64_t)row * R + col] = acc[rt][g][v];
            }
        }
}

// The mixed near field of real maps on FP64 MFMA from FP32 operands in load order.  As above, one
// wave per task (RT row tiles) and NCT 16-column tiles per wave, but the k-steps of a group of 16
// columns are assigned so that each lane's operands are contiguous: at k-step u of group g, lane
// (m, q) = (lane % 16, lane / 16) contributes column 16 g + 4 q + u.  Its matrix values of 4
// k-steps are then one 16-byte load per row tile from the task's row-major FP32 copy, its 4
// source points one 16-byte load from the padded point table, and its NCT source columns of a
// point one load from the interleaved FP32 copy of x.  Rows past nrow read row nrow - 1 (their
// results are not stored), padding columns have zero values and repeat a valid point, so the loop
// needs no masks.  The point table runs two groups ahead of its use, the operands one group.
// M32: the products on v_mfma_f32_16x16x4f32 (twice the FP64 MFMA rate on gfx950) with FP32
// partial sums of FLUSH groups (16 FLUSH products per output) added into FP64 accumulators; else
// on v_mfma_f64_16x16x4f64 with FP64 accumulation throughout.
template <int RT, int NCT, bool M32>
__global__ void __launch_bounds__(TPB, 2) near_frag_kernel(const NearTask *__restrict__ tasks, int64_t t0, int64_t t1,
                                                        const int32_t *__restrict__ xrow,
                                                        const float *__restrict__ frag,
                                                        const float *__restrict__ x32, int64_t pitch,
                                                        double *__restrict__ ys, int R) {
    typedef float float4_t __attribute__((ext_vector_type(4)));
    typedef int int4_t __attribute__((ext_vector_type(4)));
    typedef float floatn_t __attribute__((ext_vector_type(NCT)));
    int wg = blockIdx.x;
    const int n8 = gridDim.x & ~7;
    if (wg < n8) wg = (wg & 7) * (n8 >> 3) + (wg >> 3);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t ti = t0 + (int64_t)wg * (TPB / 64) + wave;
    if (ti >= t1) return;
    const NearTask tk = tasks[ti];
    const int nrow = tk.nrow;
    const int ncp = (int)near_ncolp(tk.ncol);
    const int ng = ncp >> 4;
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, q = lane >> 4;
    const int c0 = blockIdx.y * NCT * 16;
    const float *__restrict__ ap[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) ap[rt] = frag + tk.foff + (int64_t)min(16 * rt + m, nrow - 1) * ncp + 4 * q;
    const int32_t *__restrict__ xq = xrow + tk.xoff + 4 * q;
    const float *__restrict__ xb = x32 + c0 + NCT * m;
    auto load_pt = [&](int g) { return *reinterpret_cast<const int4_t *>(xq + 16 * min(g, ng - 1)); };
    float4_t a[RT];
    floatn_t bv[4];
    auto load_ops = [&](int g, const int4_t &pt) {
        const int gc = min(g, ng - 1);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) a[rt] = *reinterpret_cast<const float4_t *>(ap[rt] + 16 * gc);
#pragma unroll
        for (int u = 0; u < 4; ++u) bv[u] = *reinterpret_cast<const floatn_t *>(xb + (int64_t)pt[u] * pitch);
    };
    constexpr int FLUSH = 2;
    double4_t acc[RT][NCT];
    float4_t acc32[RT][NCT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int t = 0; t < NCT; ++t) {
            acc[rt][t] = double4_t{0.0, 0.0, 0.0, 0.0};
            acc32[rt][t] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
    auto flush = [&]() {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int t = 0; t < NCT; ++t) {
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[rt][t][v] += (double)acc32[rt][t][v];
                acc32[rt][t] = float4_t{0.f, 0.f, 0.f, 0.f};
            }
    };
    int4_t pt1 = load_pt(1);
    load_ops(0, load_pt(0));
    for (int g = 0; g < ng; ++g) {
        float4_t ac[RT];
        floatn_t bc[4];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) ac[rt] = a[rt];
#pragma unroll
        for (int u = 0; u < 4; ++u) bc[u] = bv[u];
        const int4_t pt2 = load_pt(g + 2);
        load_ops(g + 1, pt1);
        // keep the next group's loads ahead of this group's MFMAs
        __asm__ volatile("" ::: "memory");
        if constexpr (M32) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                    for (int t = 0; t < NCT; ++t)
                        acc32[rt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[rt][u], bc[u][t], acc32[rt][t], 0, 0, 0);
            if (g % FLUSH == FLUSH - 1) flush();
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                    for (int t = 0; t < NCT; ++t)
                        acc[rt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ac[rt][u], (double)bc[u][t],
                                                                          acc[rt][t], 0, 0, 0);
        }
        pt1 = pt2;
    }
    if constexpr (M32) flush();
    double *__restrict__ yo = ys + (tk.out0 + tk.row0) * R;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int t = 0; t < NCT; ++t) {
            const int col = c0 + 16 * t + m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // D[m][n] in lane n + 16 (m / 4), element m % 4 (FP32); lane n + 16 (m % 4), element m / 4 (FP64)
                const int row = 16 * rt + (M32 ? 4 * (lane >> 4) + v : mf_d_row(lane, v));
                if (row < nrow && col < R) yo[(int64_t)row * R + col] = acc[rt][t][v];
            }
        }
}

// Near field of complex maps on FP64 MFMA, one row-block product per task as near_task_kernel,
// in real form: a complex product Y = A X is the real product of the matrix with entries
// [Re A, -Im A; Im A, Re A] (row 2i + c, column 2j + b) and [Re X; Im X] (row 2j + b), so a
// 16-row tile holds 8 complex rows and a k-step 2 complex columns.  A lane's A operand is one FP64
// part of its entry (Re when c == b, else Im, negated for c = 0 -- c = 1 for the adjoint, which
// conjugates A) and its B operand one part of its source value; D row 2i + c is part c of row i.
// Tasks have at most 4 tiles (32 complex rows).
template <bool TRANS, bool CONJ, int RT, int KU, int NCT>
__global__ void __launch_bounds__(TPB, 2)
    near_task_cplx_kernel(const NearTask *__restrict__ tasks, int64_t t0, int64_t t1, const NearCol *__restrict__ cols,
                          const double *__restrict__ vals, const double *__restrict__ xs, double *__restrict__ ys,
                          int R) {
    constexpr int KBC = 2 * KU;  // complex columns per batch
    int wg = blockIdx.x;
    const int n8 = gridDim.x & ~7;
    if (wg < n8) wg = (wg & 7) * (n8 >> 3) + (wg >> 3);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t ti = t0 + (int64_t)wg * (TPB / 64) + wave;
    if (ti >= t1) return;
    const NearTask tk = tasks[ti];
    const int nrow = tk.nrow, ncol = tk.ncol;
    const NearCol *__restrict__ cl = cols + tk.col0;
    const int lane = threadIdx.x & 63;
    const int ar = mf_a_row(lane), ak = mf_a_k(lane);
    // this lane's A part (Re or +-Im) and B part (Re or Im); its column within a k-step
    const int cpart = ar & 1, bpart = ak & 1, jk = ak >> 1;
    const int apart = cpart == bpart ? 0 : 1;
    const double asgn = apart == 0 ? 1.0 : ((cpart == 0) != CONJ ? -1.0 : 1.0);
    const int c0 = blockIdx.y * NCT * 16;
    const int tl = lane & (KBC - 1);
    int64_t arow[RT];
    double rmask[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int r = 8 * rt + (ar >> 1);
        arow[rt] = tk.row0 + min(r, nrow - 1);
        rmask[rt] = r < nrow ? asgn : 0.0;
    }
    int xcol[NCT];
    double cmask[NCT];
#pragma unroll
    for (int g = 0; g < NCT; ++g) {
        const int col = c0 + 16 * g + ar;
        xcol[g] = min(col, R - 1);
        cmask[g] = col < R ? 1.0 : 0.0;
    }
    double a[KU][RT], b[KU][NCT], km[KU];
    auto load_ops = [&](int k0, const NearCol &ev) {
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int src = (2 * u + jk) << 2;  // byte address of the lane holding this column's entry
            const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)ev.aoff);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(ev.aoff >> 32));
            const int xr = __builtin_amdgcn_ds_bpermute(src, ev.xrow);
            const int64_t rs = TRANS ? (int64_t)__builtin_amdgcn_ds_bpermute(src, ev.rstride) : 1;
            km[u] = (k0 + 2 * u + jk < ncol) ? 1.0 : 0.0;
            const double *__restrict__ ap = vals + 2 * (int64_t)(((uint64_t)hi << 32) | lo) + apart;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) a[u][rt] = ap[2 * arow[rt] * rs];
            const double *__restrict__ xp = xs + 2 * (int64_t)xr * R + bpart;
#pragma unroll
            for (int g = 0; g < NCT; ++g) b[u][g] = xp[2 * xcol[g]];
        }
    };
    double4_t acc[RT][NCT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int g = 0; g < NCT; ++g) acc[rt][g] = double4_t{0.0, 0.0, 0.0, 0.0};
    const int nb = (ncol + KBC - 1) / KBC;
    NearCol tb0 = cl[min(tl, ncol - 1)], tb1 = cl[min(KBC + tl, ncol - 1)];
    load_ops(0, tb0);
    for (int ib = 0; ib < nb; ++ib) {
        double ac[KU][RT], bc[KU][NCT];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) ac[u][rt] = a[u][rt] * (rmask[rt] * km[u]);
#pragma unroll
            for (int g = 0; g < NCT; ++g) bc[u][g] = b[u][g] * (cmask[g] * km[u]);
        }
        const NearCol tb2 = cl[min((ib + 2) * KBC + tl, ncol - 1)];
        load_ops((ib + 1) * KBC, tb1);
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < KU; ++u)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int g = 0; g < NCT; ++g)
                    acc[rt][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[u][rt], bc[u][g], acc[rt][g], 0, 0, 0);
        tb1 = tb2;
    }
    double *__restrict__ yo = ys + 2 * (tk.out0 + tk.row0) * R + (ak & 1);
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int g = 0; g < NCT; ++g) {
            const int col = c0 + 16 * g + ar;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // D row q + 4v (q = lane / 16) is part q % 2 of complex row q / 2 + 2v of the tile
                const int row = 8 * rt + (ak >> 1) + 2 * v;
                if (row < nrow && col < R) yo[2 * ((int64_t)row * R + col)] = acc[rt][g][v];
            }
        }
}

template <class T> struct vec4_of;
template <> struct vec4_of<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct vec4_of<double> { typedef double type __attribute__((ext_vector_type(4))); };

// p[k0 .. k0 + 3] as doubles, zero where !valid or beyond K: one vector load when K is a multiple of
// 4 (k0 is then one too, and rows of K values keep 16-byte alignment), else clamped scalar loads
template <class T>
__device__ inline void load4_k(const T *__restrict__ p, int k0, int K, bool valid, double (&o)[4]) {
    if ((K & 3) == 0) {
        using V4 = typename vec4_of<T>::type;
        const V4 v = *reinterpret_cast<const V4 *>(p + min(k0, K - 4));
#pragma unroll
        for (int u = 0; u < 4; ++u) o[u] = (valid && k0 + u < K) ? (double)v[u] : 0.0;
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) o[u] = (double)p[min(k0 + u, K - 1)] * ((valid && k0 + u < K) ? 1.0 : 0.0);
    }
}

// Leaf expansions on FP64 MFMA (real maps, R >= MFMA_MIN_R), FP64 accumulation.
//   L2P: Y_leaf (n x R) += alpha C^T (n x K) L (K x R)      rows = points, contraction over k
//   P2M: M (K x R)      = C (K x n) X (n x R)                rows = k, contraction over points
// A k-step's 4 contraction indices are assigned so that each lane's operands over 4 consecutive
// k-steps are 4 consecutive values (index 16 j + 4 k' + u for lane row k' and k-step 4 j + u):
// the contraction order does not change the product, and every lane reads its operand rows
// contiguously.  P2M: one workgroup per (leaf, group of NCT 16-column RHS tiles), its waves over
// the row tiles of k.  L2P of large leaves (STAGE): the same, its waves over the leaf's 16-point
// row tiles; L2P of small leaves: one wave per leaf (8 leaves per workgroup), since a leaf of at
// most 16 points has one row tile and a workgroup per leaf would leave 7 of its 8 waves idle.
template <class T, int NCT, bool STAGE>
__global__ void __launch_bounds__(L2P_NT)
    l2p_mfma_kernel(const int32_t *__restrict__ leaves, int n_leaves, const int64_t *__restrict__ begin,
                    const int64_t *__restrict__ end, const T *__restrict__ coll, int K, const T *__restrict__ loc,
                    int R, double *__restrict__ ys, double alpha) {
    // STAGE (leaves of many points): the leaf's expansions of this workgroup's columns, [col][k]
    // with a row pitch of K + 1, read from the [box][col][k] layout with coalesced loads (the
    // MFMA operand order reads 16 columns K values apart); small leaves read them directly, where
    // the staging would cost more than the operand loads it saves
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lt = reinterpret_cast<T *>(smem_raw);
    const int KP = K + 1;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int li = STAGE ? (int)blockIdx.x : (int)blockIdx.x * (L2P_NT / 64) + wave;
    if (!STAGE && li >= n_leaves) return;  // no barrier below without STAGE
    const int b = leaves[li];
    const int64_t j0 = begin[b];
    const int n = (int)(end[b] - j0);
    const int c0 = blockIdx.y * NCT * 16;
    const int ncol = min(NCT * 16, R - c0);
    const T *lb = loc + ((int64_t)b * R + c0) * K;
    if (STAGE) {
        if ((K & 3) == 0) {
            // one 16-byte (FP32) or 32-byte (FP64) load per thread: 4 values of one column
            using V4 = typename vec4_of<T>::type;
            for (int e = 4 * (int)threadIdx.x; e < ncol * K; e += 4 * (int)blockDim.x) {
                const int c = e / K, k = e - c * K;
                const V4 v = *reinterpret_cast<const V4 *>(lb + e);
#pragma unroll
                for (int u = 0; u < 4; ++u) lt[c * KP + k + u] = v[u];
            }
        } else {
            for (int e = threadIdx.x; e < ncol * K; e += (int)blockDim.x) {
                const int c = e / K, k = e - c * K;
                lt[c * KP + k] = lb[e];
            }
        }
        __syncthreads();
    }
    const T *lsrc = STAGE ? lt : lb;
    const int pitch = STAGE ? KP : K;
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, kq = lane >> 4;
    for (int rt = STAGE ? wave : 0; rt * 16 < n; rt += STAGE ? (int)(blockDim.x >> 6) : 1) {
        const int row = rt * 16 + m;
        const T *crow = coll + (j0 + min(row, n - 1)) * K;
        double4_t acc[NCT];
#pragma unroll
        for (int g = 0; g < NCT; ++g) acc[g] = double4_t{0.0, 0.0, 0.0, 0.0};
        for (int jb = 0; jb < K; jb += 16) {  // uniform trip count
            const int kb = jb + 4 * kq;
            double a[4], bv[NCT][4];
            // loads from clamped (valid) addresses, out-of-range operands zero (no divergent
            // branches in the loop, so the accumulators stay in the MFMA registers); the global
            // rows go as one vector load per lane and k-step group
            load4_k(crow, kb, K, row < n, a);
#pragma unroll
            for (int g = 0; g < NCT; ++g) {
                const int cl = min(16 * g + m, ncol - 1);
                const T *lc = lsrc + cl * pitch;
                if (STAGE) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        bv[g][u] = (double)lc[min(kb + u, K - 1)] * ((16 * g + m < ncol && kb + u < K) ? 1.0 : 0.0);
                } else {
                    load4_k(lc, kb, K, 16 * g + m < ncol, bv[g]);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int g = 0; g < NCT; ++g) acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bv[g][u], acc[g], 0, 0, 0);
        }
#pragma unroll
        for (int g = 0; g < NCT; ++g) {
            const int col = c0 + 16 * g + m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ro = rt * 16 + mf_d_row(lane, v);
                if (ro < n && col < R) {
                    double *yo = ys + (j0 + ro) * R + col;
                    *yo = *yo + alpha * acc[g][v];
                }
            }
        }
    }
}

template <class T, int NCT>
__global__ void __launch_bounds__(P2M_NT)
    p2m_mfma_kernel(const int32_t *__restrict__ leaves, const int64_t *__restrict__ begin,
                    const int64_t *__restrict__ end, const T *__restrict__ coll, int K, const double *__restrict__ xs,
                    int R, T *__restrict__ mom) {
    // results go through LDS ([col][k], row pitch K + 1) and leave with coalesced stores in the
    // [box][col][k] layout (the MFMA result order would write 16 columns K values apart)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *mt = reinterpret_cast<T *>(smem_raw);
    const int KP = K + 1;
    const int b = leaves[blockIdx.x];
    const int64_t j0 = begin[b];
    const int n = (int)(end[b] - j0);
    const int c0 = blockIdx.y * NCT * 16;
    const int ncol = min(NCT * 16, R - c0);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, kq = lane >> 4;
    for (int rt = wave; rt * 16 < K; rt += (int)(blockDim.x >> 6)) {
        const int krow = rt * 16 + m;
        double4_t acc[NCT];
#pragma unroll
        for (int g = 0; g < NCT; ++g) acc[g] = double4_t{0.0, 0.0, 0.0, 0.0};
        for (int jb = 0; jb < n; jb += 16) {  // uniform trip count
            const int pb = jb + 4 * kq;
            double a[4], bv[NCT][4];
            // unconditional loads from clamped (valid) addresses, out-of-range operands multiplied
            // by zero (no branches in the loop, so the accumulators stay in the MFMA registers)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                a[u] = (double)coll[(j0 + min(pb + u, n - 1)) * K + min(krow, K - 1)] *
                       ((pb + u < n && krow < K) ? 1.0 : 0.0);
#pragma unroll
            for (int g = 0; g < NCT; ++g) {
                const int col = c0 + 16 * g + m;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    bv[g][u] = xs[(j0 + min(pb + u, n - 1)) * R + min(col, R - 1)] * ((pb + u < n && col < R) ? 1.0 : 0.0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int g = 0; g < NCT; ++g) acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bv[g][u], acc[g], 0, 0, 0);
        }
#pragma unroll
        for (int g = 0; g < NCT; ++g) {
            const int cl = 16 * g + m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ko = rt * 16 + mf_d_row(lane, v);
                if (ko < K && cl < ncol) mt[cl * KP + ko] = (T)acc[g][v];
            }
        }
    }
    __syncthreads();
    T *mb = mom + ((int64_t)b * R + c0) * K;
    if ((K & 3) == 0) {
        // one 16-byte (FP32) or 32-byte (FP64) store per thread: 4 values of one column
        using V4 = typename vec4_of<T>::type;
        for (int e = 4 * (int)threadIdx.x; e < ncol * K; e += 4 * (int)blockDim.x) {
            const int c = e / K, k = e - c * K;
            V4 v;
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = mt[c * KP + k + u];
            *reinterpret_cast<V4 *>(mb + e) = v;
        }
    } else {
        for (int e = threadIdx.x; e < ncol * K; e += (int)blockDim.x) {
            const int c = e / K, k = e - c * K;
            mb[e] = mt[c * KP + k];
        }
    }
}

// The FP32 row-major copy of one task per workgroup: row r, column j at foff + r * ncolp + j, the
// value of row row0 + r of column j's block (zero for the padding columns j >= ncol).
__global__ void near_frag_fill_kernel(const NearTask *__restrict__ tasks, const NearCol *__restrict__ cols,
                                      const double *__restrict__ vals, float *__restrict__ frag) {
    const NearTask tk = tasks[blockIdx.x];
    const int64_t ncp = near_ncolp(tk.ncol);
    const int64_t n = (int64_t)tk.nrow * ncp;
    float *__restrict__ dst = frag + tk.foff;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        const int64_t r = e / ncp, j = e - r * ncp;
        float v = 0.f;
        if (j < tk.ncol) {
            const NearCol cc = cols[tk.col0 + j];
            v = (float)vals[cc.aoff + (int64_t)(tk.row0 + r) * cc.rstride];
        }
        dst[e] = v;
    }
}

// FP32 source columns of the mixed near field: x32 [point][pitch] from xs [point][R].  Within each
// chunk of W = 16 NCT columns, column chunk + 16 t + m sits at chunk + NCT m + t, so the lane of
// MFMA column m reads its NCT columns (m, m + 16, ...) of a point with one load; columns >= R are
// zero.
template <int NCT>
__global__ void near_x32_kernel(const double *__restrict__ xs, int R, int64_t n, int64_t pitch,
                                float *__restrict__ x32) {
    constexpr int W = 16 * NCT;
    const int64_t total = n * pitch;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pt = i / pitch;
        const int q = (int)(i - pt * pitch);
        const int rem = q % W;
        const int c = q - rem + (rem % NCT) * 16 + rem / NCT;
        x32[i] = c < R ? (float)xs[pt * R + c] : 0.f;
    }
}

// probe of the operand / result lane layout used above: D = A (16x4) * B (4x16)
__global__ void mfma_f64_probe_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                      double *__restrict__ D) {
    const int lane = threadIdx.x & 63;
    const double a = A[mf_a_row(lane) * 4 + mf_a_k(lane)];
    const double b = B[(lane >> 4) * 16 + (lane & 15)];
    double4_t acc = double4_t{0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) D[mf_d_row(lane, v) * 16 + (lane & 15)] = acc[v];
}
// probe of the FP32 16x16x4 layout the list M2L kernel uses: operands as for FP64, result
// D[m][n] in lane n + 16 (m / 4), element m % 4
__global__ void mfma_f32_probe_kernel(const float *__restrict__ A, const float *__restrict__ B, float *__restrict__ D) {
    typedef float float4_t __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const float a = A[mf_a_row(lane) * 4 + mf_a_k(lane)];
    const float b = B[(lane >> 4) * 16 + (lane & 15)];
    float4_t acc = float4_t{0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) D[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = acc[v];
}
// probe of the BF16 16x16x32 layout the split-BF16 dense M2L kernel uses: lane l holds
// A[l % 16][8 (l / 16) + u] and B[8 (l / 16) + u][l % 16], u = 0..7; D[m][n] in lane
// n + 16 (m / 4), element m % 4
__global__ void mfma_bf16_probe_kernel(const float *__restrict__ A, const float *__restrict__ B, float *__restrict__ D) {
    typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
    typedef float float4_t __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    bf16x8 a, b;
    for (int u = 0; u < 8; ++u) {
        a[u] = (__bf16)A[(lane & 15) * 32 + 8 * (lane >> 4) + u];
        b[u] = (__bf16)B[(8 * (lane >> 4) + u) * 16 + (lane & 15)];
    }
    float4_t acc = float4_t{0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) D[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = acc[v];
}
}  // namespace

bool launch_l2p_mfma(int cdtype, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                     const void *coll, int K, const void *loc, int R, void *ys, double alpha, bool stage,
                     hipStream_t s) {
    if (R < MFMA_MIN_R || (cdtype != H2G_F64 && cdtype != H2G_F32)) return false;
    // column tiles per workgroup: up to 4, as many as the LDS block of 64 KB allows
    const size_t es = cdtype == H2G_F64 ? 8 : 4;
    const int nct = std::min({4, (R + 15) / 16, stage ? (int)(65536 / (es * 16 * (size_t)(K + 1))) : 4});
    if (nct < 1) return false;
    const int per_wg = stage ? 1 : L2P_NT / 64;  // leaves per workgroup
    const dim3 grid((unsigned)((n + per_wg - 1) / per_wg), (unsigned)((R + 16 * nct - 1) / (16 * nct)));
    auto go = [&](auto kern, auto tv) {
        using T = decltype(tv);
        const size_t sh = stage ? sizeof(T) * (size_t)nct * 16 * (K + 1) : 0;
        kern<<<grid, L2P_NT, sh, s>>>(leaves, n, begin, end, (const T *)coll, K, (const T *)loc, R, (double *)ys,
                                      alpha);
    };
    const bool f64 = cdtype == H2G_F64;
    switch (nct) {
    case 1: f64 ? go((stage ? l2p_mfma_kernel<double, 1, true> : l2p_mfma_kernel<double, 1, false>), 0.0) : go((stage ? l2p_mfma_kernel<float, 1, true> : l2p_mfma_kernel<float, 1, false>), 0.0f); break;
    case 2: f64 ? go((stage ? l2p_mfma_kernel<double, 2, true> : l2p_mfma_kernel<double, 2, false>), 0.0) : go((stage ? l2p_mfma_kernel<float, 2, true> : l2p_mfma_kernel<float, 2, false>), 0.0f); break;
    case 3: f64 ? go((stage ? l2p_mfma_kernel<double, 3, true> : l2p_mfma_kernel<double, 3, false>), 0.0) : go((stage ? l2p_mfma_kernel<float, 3, true> : l2p_mfma_kernel<float, 3, false>), 0.0f); break;
    default: f64 ? go((stage ? l2p_mfma_kernel<double, 4, true> : l2p_mfma_kernel<double, 4, false>), 0.0) : go((stage ? l2p_mfma_kernel<float, 4, true> : l2p_mfma_kernel<float, 4, false>), 0.0f); break;
    }
    H2G_CHECK_LAUNCH();
    return true;
}

bool launch_p2m_mfma(int cdtype, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                     const void *coll, int K, const void *xs, int R, void *mom, hipStream_t s) {
    if (R < MFMA_MIN_R || (cdtype != H2G_F64 && cdtype != H2G_F32)) return false;
    // column tiles per workgroup: up to 4, as many as the LDS block of 64 KB allows
    const size_t es = cdtype == H2G_F64 ? 8 : 4;
    const int nct = std::min({4, (R + 15) / 16, (int)(65536 / (es * 16 * (size_t)(K + 1)))});
    if (nct < 1) return false;
    const dim3 grid((unsigned)n, (unsigned)((R + 16 * nct - 1) / (16 * nct)));
    auto go = [&](auto kern, auto tv) {
        using T = decltype(tv);
        const size_t sh = sizeof(T) * (size_t)nct * 16 * (K + 1);
        kern<<<grid, P2M_NT, sh, s>>>(leaves, begin, end, (const T *)coll, K, (const double *)xs, R, (T *)mom);
    };
    const bool f64 = cdtype == H2G_F64;
    switch (nct) {
    case 1: f64 ? go(p2m_mfma_kernel<double, 1>, 0.0) : go(p2m_mfma_kernel<float, 1>, 0.0f); break;
    case 2: f64 ? go(p2m_mfma_kernel<double, 2>, 0.0) : go(p2m_mfma_kernel<float, 2>, 0.0f); break;
    case 3: f64 ? go(p2m_mfma_kernel<double, 3>, 0.0) : go(p2m_mfma_kernel<float, 3>, 0.0f); break;
    default: f64 ? go(p2m_mfma_kernel<double, 4>, 0.0) : go(p2m_mfma_kernel<float, 4>, 0.0f); break;
    }
    H2G_CHECK_LAUNCH();
    return true;
}

bool launch_near_mfma(int op, const Core &c, const void *xs, void *ys, int R, float *x32, int near_cols,
                      hipStream_t s) {
    if ((c.dtype != H2G_F64 && c.dtype != H2G_C64) || R < MFMA_MIN_R) return false;
    const bool trans = op != H2G_OP_N;
    const NearPlan &P = trans ? c.near_rev_plan : c.near_fwd_plan;
    if (!P.ok) return false;
    using std::integral_constant;
    if (c.dtype == H2G_C64) {
        // complex maps (both forms: the near field computes in ComplexF64), up to 32 RHS per
        // workgroup, tasks of 4, 3, 2, 1 tiles of 8 complex rows.  Measured on Helmholtz maps of
        // 1.2M (Toeplitz) and 1M (dense) points on the sphere: the block kernels win at 1 and 2
        // columns (3.0 / 6.1 ms against 9.3 / 9.6 ms forward), this one from 4 columns on (9.8
        // against 11.8 ms at 4 columns, 14 against 27 ms forward and 16 against 82 ms adjoint at
        // 32); batches of 16 complex columns for tasks of 1-2 tiles and 8 for 3-4 tiles (8 for
        // all: up to 10 % slower; 8 and 4: up to 35 % slower)
        if (R < 3) return false;
        const int nct = R > 16 ? 2 : 1;
        const unsigned gy = (unsigned)((R + 16 * nct - 1) / (16 * nct));
        for (int i = 0; i < 4; ++i) {
            const int64_t t0 = P.cls[i], t1 = P.cls[i + 1];
            if (t1 == t0) continue;
            const dim3 grid((unsigned)((t1 - t0 + TPB / 64 - 1) / (TPB / 64)), gy);
            auto pick = [&](auto rt_c, auto ku_c, auto nct_c) {
                constexpr int RT = decltype(rt_c)::value, KU = decltype(ku_c)::value, NC = decltype(nct_c)::value;
                auto kern = op == H2G_OP_N   ? near_task_cplx_kernel<false, false, RT, KU, NC>
                            : op == H2G_OP_T ? near_task_cplx_kernel<true, false, RT, KU, NC>
                                             : near_task_cplx_kernel<true, true, RT, KU, NC>;
                hipLaunchKernelGGL(kern, grid, TPB, 0, s, P.task.as<NearTask>(), t0, t1, P.col.as<NearCol>(),
                                   (const double *)c.near_vals.ptr, (const double *)xs, (double *)ys, R);
            };
            auto by_nct = [&](auto rt_c, auto ku_c) {
                if (nct == 1)
                    pick(rt_c, ku_c, integral_constant<int, 1>{});
                else
                    pick(rt_c, ku_c, integral_constant<int, 2>{});
            };
            switch (4 - i) {
            case 1: by_nct(integral_constant<int, 1>{}, integral_constant<int, 8>{}); break;
            case 2: by_nct(integral_constant<int, 2>{}, integral_constant<int, 8>{}); break;
            case 3: by_nct(integral_constant<int, 3>{}, integral_constant<int, 4>{}); break;
            default: by_nct(integral_constant<int, 4>{}, integral_constant<int, 4>{}); break;
            }
            H2G_CHECK_LAUNCH();
        }
        return true;
    }
    if (x32 && P.frag.ptr) {
        // mixed form: FP32 source columns, then the tasks on their row-major FP32 values
        const int nct = near_x32_nct(R, near_cols);
        const int64_t pitch = near_x32_pitch(R, near_cols);
        const int64_t n_src = (trans ? c.rev : c.fwd).src->n_dofs;
        const unsigned gx = (unsigned)std::min<int64_t>((n_src * pitch + TPB - 1) / TPB, 1 << 20);
        if (gx > 0) {
            switch (nct) {
            case 1: hipLaunchKernelGGL((near_x32_kernel<1>), gx, TPB, 0, s, (const double *)xs, R, n_src, pitch, x32); break;
            case 2: hipLaunchKernelGGL((near_x32_kernel<2>), gx, TPB, 0, s, (const double *)xs, R, n_src, pitch, x32); break;
            default: hipLaunchKernelGGL((near_x32_kernel<4>), gx, TPB, 0, s, (const double *)xs, R, n_src, pitch, x32); break;
            }
            H2G_CHECK_LAUNCH();
        }
        const unsigned gy = (unsigned)(pitch / (16 * nct));
        for (int i = 0; i < 4; ++i) {  // tasks of 4, 3, 2, 1 row tiles
            const int64_t t0 = P.cls[i], t1 = P.cls[i + 1];
            if (t1 == t0) continue;
            const dim3 grid((unsigned)((t1 - t0 + TPB / 64 - 1) / (TPB / 64)), gy);
            auto go = [&](auto rt_c, auto nct_c) {
                constexpr int RT = decltype(rt_c)::value, NC = decltype(nct_c)::value;
                hipLaunchKernelGGL((near_frag_kernel<RT, NC, true>), grid, TPB, 0, s, P.task.as<NearTask>(), t0, t1,
                                   P.xrow.as<int32_t>(), P.frag.as<float>(), (const float *)x32, pitch, (double *)ys, R);
            };
            auto by_nct = [&](auto rt_c) {
                switch (nct) {
                case 1: go(rt_c, integral_constant<int, 1>{}); break;
                case 2: go(rt_c, integral_constant<int, 2>{}); break;
                default: go(rt_c, integral_constant<int, 4>{}); break;
                }
            };
            switch (4 - i) {
            case 1: by_nct(integral_constant<int, 1>{}); break;
            case 2: by_nct(integral_constant<int, 2>{}); break;
            case 3: by_nct(integral_constant<int, 3>{}); break;
            default: by_nct(integral_constant<int, 4>{}); break;
            }
            H2G_CHECK_LAUNCH();
        }
        return true;
    }
    // NCT column tiles per workgroup (up to 64 RHS); wider batches take several workgroups
    const int nct = near_cols ? near_cols / 16 : std::min(4, (R + 15) / 16);
    const unsigned gy = (unsigned)((R + 16 * nct - 1) / (16 * nct));
    for (int i = 0; i < 4; ++i) {  // tasks of 4, 3, 2, 1 row tiles
        const int64_t t0 = P.cls[i], t1 = P.cls[i + 1];
        if (t1 == t0) continue;
        const dim3 grid((unsigned)((t1 - t0 + TPB / 64 - 1) / (TPB / 64)), gy);
        // the orientation selects the instantiation
        auto pick = [&](auto rt_c, auto ku_c, auto nct_c) {
            constexpr int RT = decltype(rt_c)::value, KU = decltype(ku_c)::value, NC = decltype(nct_c)::value;
            auto kern = trans ? near_task_kernel<true, RT, KU, NC, double> : near_task_kernel<false, RT, KU, NC, double>;
            hipLaunchKernelGGL(kern, grid, TPB, 0, s, P.task.as<NearTask>(), t0, t1, P.col.as<NearCol>(),
                               c.near_vals.as<double>(), (const double *)xs, (double *)ys, R);
        };
        auto by_nct = [&](auto rt_c, auto ku_c) {
            switch (nct) {
            case 1: pick(rt_c, ku_c, integral_constant<int, 1>{}); break;
            case 2: pick(rt_c, ku_c, integral_constant<int, 2>{}); break;
            case 3: pick(rt_c, ku_c, integral_constant<int, 3>{}); break;
            default: pick(rt_c, ku_c, integral_constant<int, 4>{}); break;
            }
        };
        switch (4 - i) {
        case 1: by_nct(integral_constant<int, 1>{}, integral_constant<int, 4>{}); break;
        case 2: by_nct(integral_constant<int, 2>{}, integral_constant<int, 2>{}); break;
        case 3: by_nct(integral_constant<int, 3>{}, integral_constant<int, 2>{}); break;
        default: by_nct(integral_constant<int, 4>{}, integral_constant<int, 2>{}); break;
        }
        H2G_CHECK_LAUNCH();
    }
    return true;
}

void build_near_frags(Core &c) {
    if (c.dtype != H2G_F64 || !c.near_vals.ptr) return;
    for (NearPlan *P : {&c.near_fwd_plan, &c.near_rev_plan}) {
        if (!P->ok || P->n_frag == 0) continue;
        if (!P->frag.ptr) P->frag.alloc((size_t)P->n_frag * sizeof(float));
        const int64_t nt = P->cls[4];
        if (nt > 0)
            hipLaunchKernelGGL(near_frag_fill_kernel, (unsigned)nt, TPB, 0, 0, P->task.as<NearTask>(), P->col.as<NearCol>(),
                               c.near_vals.as<double>(), P->frag.as<float>());
        H2G_CHECK_LAUNCH();
    }
    H2G_HIP(hipDeviceSynchronize());
}

}  // namespace h2g

// Test hook: runs the MFMA layout probe on host arrays (A 16x4, B 4x16 row-major; D 16x16).
// Test hook: the FP32 layout probe (A 16x4, B 4x16 row-major; D 16x16).
extern "C" int h2g_probe_mfma_f32(const float *A, const float *B, float *D) {
    using namespace h2g;
    try {
        DevBuf a, b, d;
        a.upload(A, 64 * sizeof(float));
        b.upload(B, 64 * sizeof(float));
        d.alloc(256 * sizeof(float));
        mfma_f32_probe_kernel<<<1, 64>>>(a.as<float>(), b.as<float>(), d.as<float>());
        H2G_HIP(hipGetLastError());
        H2G_HIP(hipMemcpy(D, d.ptr, 256 * sizeof(float), hipMemcpyDeviceToHost));
        return H2G_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
}

// Test hook: the BF16 layout probe (A 16x32, B 32x16 row-major; D 16x16).
extern "C" int h2g_probe_mfma_bf16(const float *A, const float *B, float *D) {
    using namespace h2g;
    try {
        DevBuf a, b, d;
        a.upload(A, 512 * sizeof(float));
        b.upload(B, 512 * sizeof(float));
        d.alloc(256 * sizeof(float));
        hipLaunchKernelGGL((mfma_bf16_probe_kernel), 1, 64, 0, 0, a.as<float>(), b.as<float>(), d.as<float>());
        H2G_HIP(hipGetLastError());
        H2G_HIP(hipMemcpy(D, d.ptr, 256 * sizeof(float), hipMemcpyDeviceToHost));
        return H2G_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
}

extern "C" int h2g_probe_mfma_f64(const double *A, const double *B, double *D) {
    using namespace h2g;
    try {
        DevBuf a, b, d;
        a.upload(A, 64 * sizeof(double));
        b.upload(B, 64 * sizeof(double));
        d.alloc(256 * sizeof(double));
        mfma_f64_probe_kernel<<<1, 64>>>(a.as<double>(), b.as<double>(), d.as<double>());
        H2G_HIP(hipGetLastError());
        H2G_HIP(hipMemcpy(D, d.ptr, 256 * sizeof(double), hipMemcpyDeviceToHost));
        return H2G_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
}

// Lattice M2L coupling stage on FP64 MFMA (v_mfma_f64_16x16x4f64).
//
// For one level, one frequency f and one block (receiving parent P with child targets t = 0..7)
// the coupling stage is
//     O[t][r] = sum_q E_f[q][t] * S[src(P, q)][r],   q = (neighbour d of P, sector s), 216 slots,
// a complex (8 x 216) x (216 x R) product whose left factor E_f is the same for every block of
// the level.  In real form (re/im split) it is one GEMM per (f, column tile):
//     D (16 x 16) += A (16 x 4) B (4 x 16),   m = 2t + c (c: re / im of the output),
//                                             k = 2q + b (b: re / im of the source),
// A[(t,c)][(q,b)] = Re E (c == b), -Im E (c = 0, b = 1), Im E (c = 1, b = 0), i.e. the 2 x 2
// real matrix of multiplication by E, and B[(q,b)][n] = re / im part of the source of column n.
// A is stored expanded per frequency ("EA": [f][q][m][b], 32 FP64 values per slot q), so each
// lane loads its A element with one LDS read; the adjoint negates the Im parts (c != b).
//
// Two kernels share this:
//  * cluster kernel (any precision, narrow batches): tiles of up to 16 receiving parents
//    (consecutive in Morton order) with the union of their present sources staged in LDS,
//    columns = (block, r), 16 per MFMA column tile; every tile runs the 208 slots q whose
//    neighbour is not P itself (those 8 couple only adjacent boxes), absent sources read zeros.
//  * list kernel: up to 12 receiving parents per tile (consecutive in Morton order) whose present
//    sources, at most 768, are staged in LDS; 16 columns = the 16 right-hand sides of one block,
//    and each block runs only its own present slots (a packed list built on the host), which
//    skips the absent sources of sparse levels.  FP32 spectra, R >= 16.
// The cluster kernel runs on FP64 MFMA (v_mfma_f64_16x16x4f64; the FP64 MFMA rate equals the
// FP64 / unpacked FP32 FMA rate on CDNA4).  The list kernel, whose operands are FP32 (spectra as
// stored, FP32 coupling table), runs on FP32 MFMA (v_mfma_f32_16x16x4f32, twice the FP64 rate)
// and adds each partial sum of 4 slot pairs (16 products) into an FP64 accumulator, which keeps
// the mixed form's error at the FP32 rounding of its inputs.
// A workgroup runs a group of consecutive tiles (Morton order, so spatially compact) for one
// frequency: it loads the frequency's coupling table once and stages the region of tile t + 1
// in registers while tile t runs on the MFMAs.  Workgroup i takes group (i % 8) * (gridDim.x / 8)
// + i / 8: consecutive workgroups go to different XCDs, so each XCD gets a contiguous range of
// groups and its L2 holds only its part of the frequency slice.
#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "h2g_internal.h"

namespace h2g {

namespace {

typedef double double4_m __attribute__((ext_vector_type(4)));
typedef float float4_m __attribute__((ext_vector_type(4)));

constexpr int NT = 256;  // threads per workgroup (4 waves)

template <class T> struct real_t { using type = T; };
template <> struct real_t<cplx<float>> { using type = float; };
template <> struct real_t<cplx<double>> { using type = double; };

__device__ inline int group_of_workgroup() {
    const int i = blockIdx.x;
    return (i & 7) * (int)(gridDim.x >> 3) + (i >> 3);
}

template <class EA, int NTH>
__device__ inline void fill_ea(EA *ea, const EA *__restrict__ EAf) {
    constexpr int N = 216 * 32, PT = (N + NTH - 1) / NTH;
    EA v[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
        const int e = u * NTH + (int)threadIdx.x;
        v[u] = e < N ? EAf[e] : EA(0);
    }
#pragma unroll
    for (int u = 0; u < PT; ++u) {
        const int e = u * NTH + (int)threadIdx.x;
        if (e < N) ea[e] = v[u];
    }
}

// Region staging, software-pipelined over the tiles of a workgroup: the slot indices of tile
// t + 2 and the spectra of tile t + 1 are loaded while tile t computes, and stored to LDS after
// the next barrier.  The region moves in chunks of CE values (16 bytes, or one value when a
// slot row is shorter): flat chunk e = thread + u * NTH is chunk e % CH of slot e / CH, so
// consecutive lanes load consecutive chunks of a slot row (coalesced) and store consecutive
// 16-byte LDS words (no bank conflicts).
template <class C, int RC, int NSLOT, int NTH>
struct RegionPipe {
    static constexpr int CE = (16 / (int)sizeof(C)) < RC ? 16 / (int)sizeof(C) : RC;  // values per chunk
    static constexpr int CH = RC / CE;                                                // chunks per slot
    static constexpr int CPT = (NSLOT * CH + NTH - 1) / NTH;                          // chunks per thread
    int idx[CPT];
    C val[CPT][CE];
    __device__ void load_idx(const int32_t *__restrict__ ts, int lane) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = lane + u * NTH;
            idx[u] = e < NSLOT * CH ? ts[e / CH] : -1;
        }
    }
    // vec: 16-byte chunks are aligned and inside R (checked by the caller for this chunk of columns)
    __device__ void load_val(const C *__restrict__ Sf, int R, int r0, bool vec, int lane) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = (lane + u * NTH) % CH;
            if (idx[u] >= 0) {
                const C *src = Sf + (int64_t)idx[u] * R + r0 + c * CE;
                if (CE * sizeof(C) == 16 && vec) {
                    const float4 t = *reinterpret_cast<const float4 *>(src);
                    __builtin_memcpy(&val[u][0], &t, 16);
                } else {
#pragma unroll
                    for (int i = 0; i < CE; ++i) val[u][i] = r0 + c * CE + i < R ? src[i] : zero_v<C>();
                }
            } else {
#pragma unroll
                for (int i = 0; i < CE; ++i) val[u][i] = zero_v<C>();
            }
        }
    }
    __device__ void store(C *reg, int lane) const {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = lane + u * NTH;
            if (e < NSLOT * CH)
#pragma unroll
                for (int i = 0; i < CE; ++i) reg[e * CE + i] = val[u][i];
        }
    }
};

// ---------------------------------------------------------------------------------------------
// Cluster kernel (any precision, narrow batches): tiles of up to 16 receiving parents,
// consecutive in Morton order, with the union of their present sources (at most 1024 slots) and,
// per block, the union index of each of its 216 slots (1024: absent, a zero row).  Columns =
// (block, r), 16 per MFMA column tile, all running the same 208 slot pairs.  Each wave takes its
// own tiles of the group with its own LDS region, so one wave's staging overlaps the MFMAs of
// the others; the coupling table of the frequency is shared.  Workgroup = (group of tiles, f,
// chunk of RC columns).
// ---------------------------------------------------------------------------------------------
constexpr int CL_NB = 16, CL_U = 1024;

template <class C, class EA, int RC, bool CONJ>
__global__ void __launch_bounds__(NT)
    m2l_cluster_kernel(int n_tiles, int tg, const int32_t *__restrict__ tile_slot, const int32_t *__restrict__ tile_qt,
                       const int32_t *__restrict__ tile_npb, const int32_t *__restrict__ tile_tgt,
                       const EA *__restrict__ EAg, int n_src, int n_recv, const C *__restrict__ S, C *__restrict__ O,
                       int R, int vec_ok) {
    using Rl = typename real_t<C>::type;
    using Pipe = RegionPipe<C, RC, CL_U, 64>;
    constexpr int NB = CL_NB, U = CL_U, G = RC;  // 16 * RC columns = RC column tiles
    constexpr int QTW = NB * 216 / 2;            // int32 words of the int16 slot table
    constexpr int QPL = (QTW + 63) / 64, TPL = NB * 8 / 64;
    constexpr size_t WAVE_BYTES = sizeof(C) * (size_t)(U + 1) * RC + sizeof(int32_t) * (QTW + NB * 8);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    EA *ea = reinterpret_cast<EA *>(smem_raw);  // [216][32], shared
    const int grp = group_of_workgroup();
    const int t0 = grp * tg, t1 = min(n_tiles, t0 + tg);
    if (t0 >= t1) return;
    const int f = blockIdx.y;
    const int r0 = blockIdx.z * RC;
    const bool vec = vec_ok && r0 + RC <= R;
    const C *Sf = S + (int64_t)f * n_src * R;
    fill_ea<EA, NT>(ea, EAg + (int64_t)f * 216 * 32);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;  // wave index is uniform
    unsigned char *wb = smem_raw + sizeof(EA) * 216 * 32 + (size_t)w * WAVE_BYTES;
    C *reg = reinterpret_cast<C *>(wb);                                                  // [U + 1][RC]
    int32_t *qt32 = reinterpret_cast<int32_t *>(wb + sizeof(C) * (size_t)(U + 1) * RC);  // [NB][216] int16
    int32_t *s_tgt = qt32 + QTW;                                                         // [NB][8]
    const int16_t *qt = reinterpret_cast<const int16_t *>(qt32);
    if (l < RC) reg[U * RC + l] = zero_v<C>();
    const int m = l & 15, b = (l >> 4) & 1, h = l >> 5;
    const int lane_a = h * 32 + m * 2 + b;
    const bool neg = CONJ && ((m & 1) != b);
    const Rl *regR = reinterpret_cast<const Rl *>(reg);
    Rl *OfR = reinterpret_cast<Rl *>(O + (int64_t)f * n_recv * R);
    // per-wave pipeline: the slot indices of tile t + 2 and everything else of tile t + 1 are
    // loaded while tile t computes
    constexpr int WS = NT / 64;  // waves; wave w takes tiles t0 + w, t0 + w + WS, ...
    Pipe pipe;
    int qpre[QPL], tpre[TPL], npre = 0, idx_n[Pipe::CPT];
    auto load_meta = [&](int t) {
#pragma unroll
        for (int u = 0; u < QPL; ++u) qpre[u] = u * 64 + l < QTW ? tile_qt[(int64_t)t * QTW + u * 64 + l] : 0;
#pragma unroll
        for (int u = 0; u < TPL; ++u) tpre[u] = tile_tgt[(int64_t)t * NB * 8 + u * 64 + l];
        npre = tile_npb[t];
    };
    int t = t0 + w;
    if (t < t1) {
        pipe.load_idx(tile_slot + (int64_t)t * U, l);
        pipe.load_val(Sf, R, r0, vec, l);
        load_meta(t);
        if (t + WS < t1) {
            Pipe tmp;
            tmp.load_idx(tile_slot + (int64_t)(t + WS) * U, l);
#pragma unroll
            for (int u = 0; u < Pipe::CPT; ++u) idx_n[u] = tmp.idx[u];
        }
    }
    for (; t < t1; t += WS) {
        // this wave's previous tile is done with its LDS region (LDS ops of a wave run in order)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        pipe.store(reg, l);
#pragma unroll
        for (int u = 0; u < QPL; ++u)
            if (u * 64 + l < QTW) qt32[u * 64 + l] = qpre[u];
#pragma unroll
        for (int u = 0; u < TPL; ++u) s_tgt[u * 64 + l] = tpre[u];
        const int npb = __builtin_amdgcn_readfirstlane(npre);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (t + WS < t1) {
#pragma unroll
            for (int u = 0; u < Pipe::CPT; ++u) pipe.idx[u] = idx_n[u];
            pipe.load_val(Sf, R, r0, vec, l);
            load_meta(t + WS);
            if (t + 2 * WS < t1) {
                Pipe tmp;
                tmp.load_idx(tile_slot + (int64_t)(t + 2 * WS) * U, l);
#pragma unroll
                for (int u = 0; u < Pipe::CPT; ++u) idx_n[u] = tmp.idx[u];
            }
        }
        // column (block, r) of this lane in column tile g; columns past the tile's blocks read
        // block 0's slots and are not written
        int cblk[G], crr[G], qbase[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = 16 * g + m;
            cblk[g] = c / RC < npb ? c / RC : -1;
            crr[g] = c % RC;
            qbase[g] = (cblk[g] >= 0 ? cblk[g] : 0) * 216 + h;
        }
        double4_m acc[G][2];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g][0] = acc[g][1] = double4_m{0.0, 0.0, 0.0, 0.0};
        for (int ks0 = 0; ks0 < 104; ks0 += 8) {
            double a[8], bv[G][8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int ks = ks0 + i;
                const int q0 = ks < 52 ? 2 * ks : 2 * ks + 8;  // slot pairs, skipping q = 104 .. 111
                a[i] = (double)ea[q0 * 32 + lane_a];
                if (neg) a[i] = -a[i];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const int ri = qt[qbase[g] + q0];
                    bv[g][i] = (double)regR[(ri * RC + crr[g]) * 2 + b];
                }
            }
            // raised wave priority while the MFMAs issue (1-column mixed products 11.7 -> 11.2 ms)
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int g = 0; g < G; ++g)
                    acc[g][i & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[g][i], acc[g][i & 1], 0, 0, 0);
            __builtin_amdgcn_s_setprio(0);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (cblk[g] < 0 || r0 + crr[g] >= R) continue;
            const int32_t *tt = s_tgt + cblk[g] * 8;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int mo = (l >> 4) + 4 * v;
                const int tgi = tt[mo >> 1];
                if (tgi >= 0) OfR[((int64_t)tgi * R + r0 + crr[g]) * 2 + (mo & 1)] = (Rl)(acc[g][0][v] + acc[g][1][v]);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// List kernel: up to LIST_NB parents per tile, 16 right-hand sides, FP32 spectra; one wave per block,
// w-th present block of the tile over its own slot list (LDS reads of 8 slot pairs ahead of
// their MFMAs, four accumulators); workgroup = (group of TG tiles, f, chunk of 16 columns).
// ---------------------------------------------------------------------------------------------
constexpr int NT_LIST = 64 * LIST_NB;  // one wave per block

template <bool CONJ>
__global__ void __launch_bounds__(NT_LIST)
    m2l_list_kernel(int n_tiles, int tg, const int32_t *__restrict__ tile_slot, const int32_t *__restrict__ tile_npb,
                    const int32_t *__restrict__ tile_tgt, const int32_t *__restrict__ tile_list,
                    const int32_t *__restrict__ tile_nlist, const float *__restrict__ EAg, int n_src, int n_recv,
                    const cplx<float> *__restrict__ S, cplx<float> *__restrict__ O, int R, int vec_ok) {
    using C = cplx<float>;
    constexpr int NTL = NT_LIST, RC = 16, NB = LIST_NB, QL = LIST_QL, NSLOT = LIST_NSLOT, LW = NB * QL * 2,
                  LPT = (LW + NTL - 1) / NTL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *reg = reinterpret_cast<C *>(smem_raw);                      // [NSLOT + 1][RC], last slot zero
    float *ea = reinterpret_cast<float *>(reg + (NSLOT + 1) * RC);  // [216][32]
    int32_t *lst = reinterpret_cast<int32_t *>(ea + 216 * 32);     // [NB][QL] (region slot, q) pairs
    int32_t *s_tgt = lst + LW;                                     // [NB][8]
    int32_t *s_nl = s_tgt + NB * 8;                                // [NB]
    int32_t *s_npb = s_nl + NB;
    const int grp = group_of_workgroup();
    const int t0 = grp * tg, t1 = min(n_tiles, t0 + tg);
    if (t0 >= t1) return;
    const int f = blockIdx.y;
    const int r0 = blockIdx.z * RC;
    const bool vec = vec_ok && r0 + RC <= R;
    const C *Sf = S + (int64_t)f * n_src * R;
    if (threadIdx.x < RC) reg[NSLOT * RC + threadIdx.x] = C{0.f, 0.f};
    fill_ea<float, NTL>(ea, EAg + (int64_t)f * 216 * 32);
    RegionPipe<C, RC, NSLOT, NTL> pipe;
    pipe.load_idx(tile_slot + (int64_t)t0 * NSLOT, threadIdx.x);
    pipe.load_val(Sf, R, r0, vec, threadIdx.x);
    int idx_n[RegionPipe<C, RC, NSLOT, NTL>::CPT];
    int lpre[LPT], m_tg = -1, m_nl = 0, m_np = 0;
    auto load_meta = [&](int t) {
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
            const int e = (int)threadIdx.x + u * NTL;
            lpre[u] = e < LW ? tile_list[(int64_t)t * LW + e] : 0;
        }
        if (threadIdx.x < NB * 8) m_tg = tile_tgt[(int64_t)t * NB * 8 + threadIdx.x];
        if (threadIdx.x < NB) m_nl = tile_nlist[(int64_t)t * NB + threadIdx.x];
        m_np = tile_npb[t];
    };
    load_meta(t0);
    if (t0 + 1 < t1) {
        RegionPipe<C, RC, NSLOT, NTL> tmp;
        tmp.load_idx(tile_slot + (int64_t)(t0 + 1) * NSLOT, threadIdx.x);
#pragma unroll
        for (int u = 0; u < RegionPipe<C, RC, NSLOT, NTL>::CPT; ++u) idx_n[u] = tmp.idx[u];
    }
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;  // wave index is uniform
    const int m = l & 15, b = (l >> 4) & 1, h = l >> 5;
    const bool neg = CONJ && ((m & 1) != b);
    const float *regR = reinterpret_cast<const float *>(reg);
    float *OfR = reinterpret_cast<float *>(O + (int64_t)f * n_recv * R);
    for (int t = t0; t < t1; ++t) {
        __syncthreads();  // the previous tile's compute is done with the LDS
        pipe.store(reg, threadIdx.x);
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
            const int e = (int)threadIdx.x + u * NTL;
            if (e < LW) lst[e] = lpre[u];
        }
        if (threadIdx.x < NB * 8) s_tgt[threadIdx.x] = m_tg;
        if (threadIdx.x < NB) s_nl[threadIdx.x] = m_nl;
        if (threadIdx.x == 0) *s_npb = m_np;
        __syncthreads();
        if (t + 1 < t1) {
#pragma unroll
            for (int u = 0; u < RegionPipe<C, RC, NSLOT, NTL>::CPT; ++u) pipe.idx[u] = idx_n[u];
            pipe.load_val(Sf, R, r0, vec, threadIdx.x);
            load_meta(t + 1);
            if (t + 2 < t1) {
                RegionPipe<C, RC, NSLOT, NTL> tmp;
                tmp.load_idx(tile_slot + (int64_t)(t + 2) * NSLOT, threadIdx.x);
#pragma unroll
                for (int u = 0; u < RegionPipe<C, RC, NSLOT, NTL>::CPT; ++u) idx_n[u] = tmp.idx[u];
            }
        }
        const int npb = __builtin_amdgcn_readfirstlane(*s_npb);
        const int j = w;
        if (j < npb) {
            // the list holds (region row, coupling row) byte offsets, padded with zero entries to
            // whole batches of 8 slot pairs; lane half h takes the second slot of each pair
            const int nk = __builtin_amdgcn_readfirstlane(s_nl[j]) / 2;  // slot pairs
            const int2 *lj = reinterpret_cast<const int2 *>(lst) + j * QL + h;
            const char *regb = reinterpret_cast<const char *>(reg) + (m * 2 + b) * (int)sizeof(float);
            const char *eab = reinterpret_cast<const char *>(ea) + (m * 2 + b) * (int)sizeof(float);
            // FP32 MFMA (FP32 operands, as stored) over 4 slot pairs per partial sum, partial sums
            // added in FP64: 16 rounded products per FP32 accumulation
            double4_m acc = double4_m{0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < nk; k0 += 8) {
                int2 e[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) e[i] = lj[2 * (k0 + i)];
                float a[8], bv[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    a[i] = *reinterpret_cast<const float *>(eab + e[i].y);
                    if (neg) a[i] = -a[i];
                    bv[i] = *reinterpret_cast<const float *>(regb + e[i].x);
                }
                float4_m f0 = float4_m{0.f, 0.f, 0.f, 0.f}, f1 = f0;
#pragma unroll
                for (int i = 0; i < 8; i += 2) {
                    f0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bv[i], f0, 0, 0, 0);
                    f1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i + 1], bv[i + 1], f1, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] += (double)f0[v] + (double)f1[v];
            }
            const double4_m sum = acc;
            if (r0 + m < R) {
                const int32_t *tt = s_tgt + j * 8;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int mo = 4 * (l >> 4) + v;  // FP32 16x16x4 result layout: rows 4 (l / 16) + v
                    const int tgi = tt[mo >> 1];
                    if (tgi >= 0) OfR[((int64_t)tgi * R + r0 + m) * 2 + (mo & 1)] = (float)sum[v];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Direct kernel (FP32 spectra, R >= DR_MIN_R): no staging of source spectra.  Workgroup = (group of
// consecutive blocks in Morton order, f, chunk of 32 columns); the frequency's coupling table is
// loaded into LDS once, then every wave takes its own blocks of the group and, per slot pair of
// the block's list, reads the two source rows straight from global memory (L2: the group's blocks
// and their neighbours on one XCD share them) and the two coupling rows from LDS.  A wave covers
// NCH 16-column tiles of the chunk, so every coupling value it reads feeds NCH MFMAs.  A block's
// list is copied into the wave's part of LDS (one coalesced load per lane, issued while the
// previous block computes), so the per-pair entries, which all lanes of a half share, are read
// from LDS and not through the vector memory return path.  Source values load one batch (of 8
// pairs) ahead into the other of two operand sets, list entries two batches ahead; FP32 MFMA over
// 4 slot pairs per partial sum, each batch's partial sums added in FP64.  116-120 registers and
// 35 KB of LDS per workgroup leave room for 4 workgroups (16 waves) per CU.
// ---------------------------------------------------------------------------------------------
constexpr int DR_NT = 512, DR_MAXB = 16;  // threads; batches of a block's list held in LDS
// workgroups per frequency (groups of consecutive blocks); measured on the N = 2M, p = 6 map with
// 4 workgroups per CU: 128 groups 33.7 ms, 256 32.5 ms, 512 to 1024 32.0 to 32.3 ms (levels 4-6)
constexpr int DR_GROUPS = 512;
// operand sets in flight: batch k's MFMAs read set k % DR_SETS while the loads of batch k + 1 fill
// the other set (three sets, two batches ahead, need more registers than 4 waves per SIMD allow
// and measured slower)
constexpr int DR_SETS = 2;
// narrowest batch for the direct kernel: it computes whole chunks of 32 columns (padding columns
// are zero), which costs less than the cluster kernel's per-column work from about 4 columns on
// (N = 2M, p = 6: 4 columns 44 ms either way, 8 columns 74 ms on the cluster kernel)
constexpr int DR_MIN_R = 4;
// batches of at most 16 columns run one 16-column tile per chunk (rows of 16 columns)
constexpr bool DR_NCH1 = true;

// One task of the direct kernel: block group `grp` of a level, frequency f, columns r0 .. r0 +
// 16 NCH - 1.  `ea` (the frequency's coupling table) and `lsd` (per wave: the list of its current
// block; the next block's list waits in registers and is written after the wave has finished
// reading the current one) are the workgroup's LDS.  Every thread of the workgroup calls it with
// the same arguments (the early return and the barrier are workgroup-uniform).
template <bool CONJ, int NCH>
__device__ __forceinline__ void direct_task(int n_blocks, int gsz, const int32_t *__restrict__ blk_ptr,
                                            const int4 *__restrict__ list, int n_list,
                                            const int32_t *__restrict__ blk_tgt, const float *__restrict__ EAg,
                                            int n_src, int n_recv, const cplx<float> *S, int Rrow,
                                            cplx<float> *__restrict__ O, int R, int grp, int f, int r0, float *ea,
                                            int4 (*lsd)[DR_MAXB * 8]) {
    const int bb0 = grp * gsz, bb1 = min(n_blocks, bb0 + gsz);
    if (bb0 >= bb1) return;
    // NCH 16-column tiles per chunk of columns: 2 (rows padded to 32 columns in the permuted
    // layout) or 1 (rows of 16 columns in natural order, batches of at most 16 columns)
    using BV = typename std::conditional<NCH == 2, float4, float2>::type;
    fill_ea<float, DR_NT>(ea, EAg + (int64_t)f * 216 * 32);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
    const int m = l & 15, bq = (l >> 4) & 1, h = l >> 5;
    // adjoint: the Im parts of the couplings (rows c != b) change sign
    const float sgn = (CONJ && ((m & 1) != bq)) ? -1.f : 1.f;
    // source rows of frequency f.  NCH = 2: permuted column layout (columns c and c + 16 of a
    // chunk of 32 adjacent), 16-byte unit m of chunk r0 / 32 holds columns r0 + m and r0 + 16 + m;
    // NCH = 1: the 8 bytes of column r0 + m
    const char *Sfb = reinterpret_cast<const char *>(S + (int64_t)f * n_src * Rrow);
    const uint32_t rowb = (uint32_t)Rrow * (uint32_t)sizeof(cplx<float>);
    const uint32_t lo4 = (uint32_t)(r0 * (int)sizeof(cplx<float>) + m * (int)sizeof(BV));
    const char *eab = reinterpret_cast<const char *>(ea) + (m * 2 + bq) * (int)sizeof(float);
    float *__restrict__ Of = reinterpret_cast<float *>(O + (int64_t)f * n_recv * R);
    // list copies: lane l moves entries l and 64 + l (int4, two pairs each) of a block's list
    // (two named registers: an indexed array would be placed in LDS by the compiler)
    int4 nx0, nx1;
    auto fetch = [&](int blk) {
        const int base = blk_ptr[blk] * 8;
        nx0 = list[min(base + l, n_list - 1)];
        nx1 = list[min(base + 64 + l, n_list - 1)];
    };
    fetch(bb0 + w);
    int4 *lw = lsd[w];
    for (int blk = bb0 + w; blk < bb1; blk += DR_NT / 64) {
        const int k0 = blk_ptr[blk], k1 = blk_ptr[blk + 1];
        lw[l] = nx0;
        lw[64 + l] = nx1;
        if (blk + DR_NT / 64 < bb1) fetch(blk + DR_NT / 64);
        // batch k, half h: 8 entries (slot, coupling row byte offset) = 4 int4 at (k - k0) * 8 + 4 h
        const int4 *lh = lw + h * 4;
        auto ldl = [&](int k, int4(&L)[4]) {
            const int4 *p = lh + (min(k, k1 - 1) - k0) * 8;
#pragma unroll
            for (int j = 0; j < 4; ++j) L[j] = p[j];
        };
        // per entry pair j (slot pairs 2j, 2j + 1): the lane of 16-lane row bq loads its half's
        // source row of pair 2j + bq (one 16-byte unit); the swap at use time sorts the values
        auto issue = [&](const int4(&L)[4], float(&an)[8], BV(&bn)[4]) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                an[2 * j] = *reinterpret_cast<const float *>(eab + L[j].y);
                an[2 * j + 1] = *reinterpret_cast<const float *>(eab + L[j].w);
                const uint32_t slot = (uint32_t)(bq ? L[j].z : L[j].x);
                bn[j] = *reinterpret_cast<const BV *>(Sfb + (slot * rowb + lo4));
            }
        };
        // DR_SETS operand sets rotate: batch k's MFMA operands come from set k % DR_SETS while
        // batch k + DR_SETS - 1 loads into set (k - 1) % DR_SETS, whose operands have been taken;
        // Lx is the list of the next batch to issue
        int4 Lx[4];
        float an[DR_SETS][8];
        BV bn[DR_SETS][4];
#pragma unroll
        for (int u = 0; u < DR_SETS - 1; ++u) {
            ldl(k0 + u, Lx);
            issue(Lx, an[u], bn[u]);
        }
        ldl(k0 + DR_SETS - 1, Lx);
        double4_m acc[NCH];
#pragma unroll
        for (int g = 0; g < NCH; ++g) acc[g] = double4_m{0.0, 0.0, 0.0, 0.0};
        // a batch past the block's last (the trip's tail) loads clamped (valid) operands and
        // skips its MFMAs.  The FP32 partial sums f0 / f1 run over the DR_SETS batches of one
        // trip (32 products each) and are added into the FP64 accumulators once per trip.
        float4_m f0[NCH], f1[NCH];
        auto step = [&](int k, auto u_, const float(&anc)[8], const BV(&bnc)[4], float(&ann)[8], BV(&bnn)[4]) {
            constexpr int u = decltype(u_)::value;
            float ac[8], bc[8][NCH];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ac[2 * j] = anc[2 * j] * sgn;
                ac[2 * j + 1] = anc[2 * j + 1] * sgn;
                // the lane holds (re, im) of columns m and 16 + m of pair 2j + bq; swapping the
                // odd 16-lane rows of the re parts with the even rows of the im parts leaves it
                // with part bq of those columns for pairs 2j (x, z) and 2j + 1 (y, w).  (Elements
                // are copied out of the returned pairs before reinterpreting them:
                // __builtin_bit_cast of a vector element reads element 0 with this compiler.)
                const auto s0 = __builtin_amdgcn_permlane16_swap(__float_as_uint(bnc[j].x), __float_as_uint(bnc[j].y),
                                                                 false, false);
                const uint32_t x = s0[0], y = s0[1];
                bc[2 * j][0] = __uint_as_float(x);
                bc[2 * j + 1][0] = __uint_as_float(y);
                if constexpr (NCH == 2) {
                    const auto s1 = __builtin_amdgcn_permlane16_swap(__float_as_uint(bnc[j].z),
                                                                     __float_as_uint(bnc[j].w), false, false);
                    const uint32_t z = s1[0], ww = s1[1];
                    bc[2 * j][NCH - 1] = __uint_as_float(z);
                    bc[2 * j + 1][NCH - 1] = __uint_as_float(ww);
                }
            }
            issue(Lx, ann, bnn);
            ldl(k + DR_SETS, Lx);
            // the next batches' loads stay ahead of this batch's MFMAs: not sunk to their uses
            // (memory clobber) and not scheduled after the MFMAs (scheduling barrier)
            __asm__ volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (u == 0) {
#pragma unroll
                for (int g = 0; g < NCH; ++g) {
                    f0[g] = float4_m{0.f, 0.f, 0.f, 0.f};
                    f1[g] = f0[g];
                }
            }
            if (k < k1) {
                // raised wave priority while this batch's MFMAs issue, so that they go out back to
                // back rather than interleaved with other waves' loads (coupling stage 28.4 ->
                // 27.4 ms on the N = 2M, p = 6 map)
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int i = 0; i < 8; i += 2)
#pragma unroll
                    for (int g = 0; g < NCH; ++g) {
                        f0[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[i], bc[i][g], f0[g], 0, 0, 0);
                        f1[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ac[i + 1], bc[i + 1][g], f1[g], 0, 0, 0);
                    }
                __builtin_amdgcn_s_setprio(0);
            }
            if constexpr (u == DR_SETS - 1) {
#pragma unroll
                for (int g = 0; g < NCH; ++g)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[g][v] += (double)f0[g][v] + (double)f1[g][v];
            }
        };
        for (int k = k0; k < k1; k += DR_SETS) {
            step(k, std::integral_constant<int, 0>{}, an[0], bn[0], an[DR_SETS - 1], bn[DR_SETS - 1]);
            step(k + 1, std::integral_constant<int, 1>{}, an[1], bn[1], an[0], bn[0]);
        }
        static_assert(DR_SETS == 2, "the trip above unrolls two batches");
#pragma unroll
        for (int g = 0; g < NCH; ++g) {
            // FP32 16x16x4 result layout: rows 4 (l / 16) + v = (target 2 (l / 16) + v / 2, re / im v % 2)
            const int col = r0 + 16 * g + m;
            if (col < R) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int tgi = blk_tgt[blk * 8 + 2 * (l >> 4) + u];
                    if (tgi >= 0) {
                        float2 o;
                        o.x = (float)acc[g][2 * u];
                        o.y = (float)acc[g][2 * u + 1];
                        // streaming store: the receiving spectra are read once, by modulate-out
                        uint64_t bits;
                        __builtin_memcpy(&bits, &o, 8);
                        __builtin_nontemporal_store(bits, reinterpret_cast<uint64_t *>(Of + ((int64_t)tgi * R + col) * 2));
                    }
                }
            }
        }
    }
}

template <bool CONJ, int NCH>
__global__ void __launch_bounds__(DR_NT, 1024 / DR_NT)
    m2l_direct_kernel(int n_blocks, int gsz, const int32_t *__restrict__ blk_ptr, const int4 *__restrict__ list,
                      int n_list, const int32_t *__restrict__ blk_tgt, const float *__restrict__ EAg, int n_src,
                      int n_recv, const cplx<float> *S, int Rrow, cplx<float> *__restrict__ O, int R) {
    __shared__ __align__(16) float ea[216 * 32];
    __shared__ int4 lsd[DR_NT / 64][DR_MAXB * 8];
    direct_task<CONJ, NCH>(n_blocks, gsz, blk_ptr, list, n_list, blk_tgt, EAg, n_src, n_recv, S, Rrow, O, R,
                           group_of_workgroup(), blockIdx.y, blockIdx.z * 16 * NCH, ea, lsd);
}

// levels of one persistent launch
constexpr int DP_MAXL = 8;
struct DirectLevel {
    int n_blocks, gsz, ng, n_list, n_src, n_recv, F, pad;
    const int32_t *blk_ptr;
    const int4 *list;
    const int32_t *blk_tgt;
    const float *EA;
    const cplx<float> *S;
    cplx<float> *O;
};
struct DirectPersist {
    int n_levels, Rrow, R, pad;
    DirectLevel lv[DP_MAXL];
    int count[DP_MAXL][8];  // tasks of level i on XCD x
    int total[8];
};

// The coupling stage of every lattice level in one persistent launch.  Each XCD x takes its own
// task list: for each level (largest first), each 32-column chunk and each frequency, the block
// groups [x ng / 8, (x + 1) ng / 8) of the level -- a contiguous range of blocks of one frequency
// at a time per XCD, as in the per-level launches (workgroup b runs on XCD b mod 8).  A workgroup
// takes the XCD's next task from counters[x] until the list is exhausted; tasks are independent,
// so no workgroup waits for another and every workgroup exits once its counter passes the end.
// The counters are zeroed on the stream before the launch.
template <bool CONJ, int NCH>
__global__ void __launch_bounds__(DR_NT, 1024 / DR_NT)
    m2l_direct_persistent_kernel(DirectPersist P, int *__restrict__ counters) {
    __shared__ __align__(16) float ea[216 * 32];
    __shared__ int4 lsd[DR_NT / 64][DR_MAXB * 8];
    __shared__ int task;
    const int x = blockIdx.x & 7;
    for (;;) {
        __syncthreads();  // the previous task's readers of ea, lsd and task are done
        if (threadIdx.x == 0) task = atomicAdd(&counters[x], 1);
        __syncthreads();
        int t = task;
        if (t >= P.total[x]) break;
        int i = 0;
        while (i + 1 < P.n_levels && t >= P.count[i][x]) t -= P.count[i][x], ++i;
        const DirectLevel &L = P.lv[i];
        const int g0 = (int)((int64_t)x * L.ng / 8), g1 = (int)((int64_t)(x + 1) * L.ng / 8);
        const int ngx = g1 - g0;
        const int g = g0 + t % ngx;
        const int fc = t / ngx;  // chunk-major, then frequency
        const int f = fc % L.F, chunk = fc / L.F;
        direct_task<CONJ, NCH>(L.n_blocks, L.gsz, L.blk_ptr, L.list, L.n_list, L.blk_tgt, L.EA, L.n_src, L.n_recv,
                               L.S, P.Rrow, L.O, P.R, g, f, chunk * 16 * NCH, ea, lsd);
    }
}

// EA[f][q][m][b] from the level's coupling table: D[op(t, q)] at frequency f, through the
// symmetry permutation of (t, q) for the symmetry-reduced kinds; zero where (t, q) does not couple.
template <class EA>
__global__ void build_ea_kernel(const int32_t *__restrict__ tq_op, const int32_t *__restrict__ tq_sym,
                                const int32_t *__restrict__ perms, const cplx<double> *__restrict__ ops, int L, int FX,
                                int F, EA *__restrict__ out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)F * 216 * 32) return;
    const int b = (int)(idx & 1), m = (int)((idx >> 1) & 15);
    const int q = (int)((idx >> 5) % 216);
    const int f = (int)(idx / (216 * 32));
    const int t = m >> 1, c = m & 1;
    const int L3 = L * L * L;
    int fx, fy, fz;
    if (FX == L) {
        fx = f % L;
        fy = (f / L) % L;
        fz = f / (L * L);
    } else {
        half_freq_of(L, f, fx, fy, fz);
    }
    const int ffull = fx + L * (fy + L * fz);
    const int op = tq_op[t * 216 + q];
    double v = 0.0;
    if (op >= 0) {
        const int sym = tq_sym ? tq_sym[t * 216 + q] : -1;
        const int ff = sym >= 0 ? perms[(int64_t)sym * L3 + ffull] : ffull;
        const cplx<double> e = ops[(int64_t)op * L3 + ff];
        v = c == b ? e.re : (c == 0 ? -e.im : e.im);
    }
    out[idx] = (EA)v;
}

template <class K>
bool allow_lds_m(K kernel, size_t bytes) {
    int dev = 0, lim = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0)
        lim = 65536;
    if (bytes > (size_t)lim) return false;
    if (bytes > 65536)
        return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
               hipSuccess;
    return true;
}

template <class C, class EA, int RC>
void cluster_go(bool conj, const LatticeLevel &LL, const EA *EAg, int n_src, int n_recv, const void *S, void *O,
                int R, hipStream_t s) {
    const LatticeTiles &T = LL.clus;
    const size_t sh = sizeof(EA) * 216 * 32 +
                      (NT / 64) * (sizeof(C) * (size_t)(CL_U + 1) * RC + sizeof(int32_t) * (CL_NB * 216 / 2 + CL_NB * 8));
    auto k = conj ? m2l_cluster_kernel<C, EA, RC, true> : m2l_cluster_kernel<C, EA, RC, false>;
    if (!allow_lds_m(k, sh)) throw Error{H2G_EUNSUPPORTED, "lattice M2L tile does not fit the LDS of this device"};
    // each wave takes every 4th tile of its workgroup's group
    const int tg = std::max(4, std::min(32, T.n_tiles / 32 / 4 * 4));
    const int ng = (T.n_tiles + tg - 1) / tg;
    const dim3 grid((unsigned)((ng + 7) / 8 * 8), (unsigned)LL.F, (unsigned)((R + RC - 1) / RC));
    const int vec_ok = ((size_t)R * sizeof(C)) % 16 == 0;
    k<<<grid, NT, sh, s>>>(T.n_tiles, tg, T.slot.as<int32_t>(), T.qt.as<int32_t>(), T.npb.as<int32_t>(),
                           T.tgt.as<int32_t>(), EAg, n_src, n_recv, (const C *)S, (C *)O, R, vec_ok);
}

}  // namespace

int m2l_direct_rows(int cdtype, const LatticeLevel &LL, int n_src, int R) {
    const bool f64 = cdtype == H2G_F64 || cdtype == H2G_C64;
    // batches of at most 16 columns: one 16-column tile per chunk, rows of 16 columns; wider
    // ones: chunks of two tiles, rows padded to 32 columns
    const int64_t Rrow = R <= 16 && DR_NCH1 ? 16 : (R + 31) / 32 * 32;
    const bool ok = !f64 && R >= DR_MIN_R && LL.ok && LL.direct.n_blocks > 0 && LL.direct.max_batches <= DR_MAXB &&
                    LL.EA32.ptr && (int64_t)n_src * Rrow * (int64_t)sizeof(cplx<float>) < INT32_MAX;
    return ok ? (int)Rrow : 0;
}

void launch_m2l_lattice(int cdtype, bool conj, const LatticeLevel &LL, int n_src, int n_recv, const void *S, void *O,
                        int R, int groups, hipStream_t s) {
    const bool f64 = cdtype == H2G_F64 || cdtype == H2G_C64;
    if (!LL.EA64.ptr) throw Error{H2G_EINVAL, "lattice coupling table missing"};
    if (const int Rrow = m2l_direct_rows(cdtype, LL, n_src, R)) {
        const LatticeDirect &D = LL.direct;
        // groups of consecutive blocks: at least one block per wave (the coupling table is
        // loaded once per group), and about `groups` groups per frequency, so that the
        // workgroups resident on an XCD work on few frequencies at a time (L2 reuse of source
        // rows)
        const int gsz = std::max(DR_NT / 64, D.n_blocks / (groups > 0 ? groups : DR_GROUPS));
        const int ng = (D.n_blocks + gsz - 1) / gsz;
        const int ch = Rrow == 16 ? 16 : 32;  // columns per chunk (NCH 16-column tiles)
        const dim3 grid((unsigned)((ng + 7) / 8 * 8), (unsigned)LL.F, (unsigned)(Rrow / ch));
        auto go = [&](auto kern) {
            kern<<<grid, DR_NT, 0, s>>>(D.n_blocks, gsz, D.ptr.as<int32_t>(), D.list.as<int4>(),
                                        (int)(D.list.bytes / sizeof(int4)), D.tgt.as<int32_t>(), LL.EA32.as<float>(),
                                        n_src, n_recv, (const cplx<float> *)S, Rrow, (cplx<float> *)O, R);
        };
        if (ch == 16)
            conj ? go(m2l_direct_kernel<true, 1>) : go(m2l_direct_kernel<false, 1>);
        else
            conj ? go(m2l_direct_kernel<true, 2>) : go(m2l_direct_kernel<false, 2>);
    } else if (!f64 && R >= 16 && LL.list.n_tiles > 0 && LL.EA32.ptr) {
        const LatticeTiles &T = LL.list;
        const size_t sh = sizeof(cplx<float>) * (size_t)(LIST_NSLOT + 1) * 16 + sizeof(float) * 216 * 32 +
                          sizeof(int32_t) * (LIST_NB * LIST_QL * 2 + LIST_NB * 8 + LIST_NB + 1);
        auto k = conj ? m2l_list_kernel<true> : m2l_list_kernel<false>;
        if (!allow_lds_m(k, sh)) throw Error{H2G_EUNSUPPORTED, "lattice M2L tile does not fit the LDS of this device"};
        const int tg = std::max(1, std::min(16, T.n_tiles / 32));
        const int ng = (T.n_tiles + tg - 1) / tg;
        const int vec_ok = (R % 2) == 0;  // 16-byte rows of 16 complex FP32 values
        const dim3 grid((unsigned)((ng + 7) / 8 * 8), (unsigned)LL.F, (unsigned)((R + 15) / 16));
        k<<<grid, NT_LIST, sh, s>>>(T.n_tiles, tg, T.slot.as<int32_t>(), T.npb.as<int32_t>(), T.tgt.as<int32_t>(),
                               T.list.as<int32_t>(), T.nlist.as<int32_t>(), LL.EA32.as<float>(), n_src, n_recv,
                               (const cplx<float> *)S, (cplx<float> *)O, R, vec_ok);
    } else if (LL.clus.n_tiles > 0) {
        const double *E = LL.EA64.as<double>();
        if (f64) cluster_go<cplx<double>, double, 1>(conj, LL, E, n_src, n_recv, S, O, R, s);
        else if (R == 1) cluster_go<cplx<float>, double, 1>(conj, LL, E, n_src, n_recv, S, O, R, s);
        else cluster_go<cplx<float>, double, 2>(conj, LL, E, n_src, n_recv, S, O, R, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{H2G_EHIP, std::string("launch: ") + hipGetErrorString(e)};
}

void launch_m2l_direct_persistent(bool conj, const std::vector<DirectLaunchLevel> &levels, int R, int groups,
                                  int *counters, hipStream_t s) {
    if (levels.empty()) return;
    if ((int)levels.size() > DP_MAXL) throw Error{H2G_EINVAL, "persistent coupling stage: too many lattice levels"};
    DirectPersist P{};
    P.n_levels = (int)levels.size();
    P.R = R;
    P.Rrow = 0;
    for (int i = 0; i < P.n_levels; ++i) {
        const DirectLaunchLevel &d = levels[i];
        const LatticeLevel &LL = *d.LL;
        const int Rrow = m2l_direct_rows(H2G_F32, LL, d.n_src, R);
        if (!Rrow || (P.Rrow && Rrow != P.Rrow)) throw Error{H2G_EINVAL, "persistent coupling stage: level not on the direct path"};
        P.Rrow = Rrow;
        const LatticeDirect &D = LL.direct;
        DirectLevel &L = P.lv[i];
        L.n_blocks = D.n_blocks;
        L.gsz = std::max(DR_NT / 64, D.n_blocks / (groups > 0 ? groups : DR_GROUPS));
        L.ng = (D.n_blocks + L.gsz - 1) / L.gsz;
        L.n_list = (int)(D.list.bytes / sizeof(int4));
        L.n_src = d.n_src;
        L.n_recv = d.n_recv;
        L.F = LL.F;
        L.blk_ptr = D.ptr.as<int32_t>();
        L.list = D.list.as<int4>();
        L.blk_tgt = D.tgt.as<int32_t>();
        L.EA = LL.EA32.as<float>();
        L.S = (const cplx<float> *)d.S;
        L.O = (cplx<float> *)d.O;
    }
    const int ch = P.Rrow == 16 ? 16 : 32;
    const int chunks = P.Rrow / ch;
    for (int x = 0; x < 8; ++x) {
        int64_t tot = 0;
        for (int i = 0; i < P.n_levels; ++i) {
            const DirectLevel &L = P.lv[i];
            const int64_t ngx = (int64_t)(x + 1) * L.ng / 8 - (int64_t)x * L.ng / 8;
            const int64_t c = ngx * L.F * chunks;
            P.count[i][x] = (int)c;
            tot += c;
        }
        if (tot > INT32_MAX / 2) throw Error{H2G_EINVAL, "persistent coupling stage: too many tasks"};
        P.total[x] = (int)tot;
    }
    auto kern = ch == 16 ? (conj ? m2l_direct_persistent_kernel<true, 1> : m2l_direct_persistent_kernel<false, 1>)
                         : (conj ? m2l_direct_persistent_kernel<true, 2> : m2l_direct_persistent_kernel<false, 2>);
    // as many workgroups as are resident at once, a multiple of the 8 XCDs (every instantiation
    // has the same resources)
    static const int grid = [&] {
        int dev = 0, ncu = 0, per_cu = 0;
        H2G_HIP(hipGetDevice(&dev));
        H2G_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        H2G_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, DR_NT, 0));
        return std::max(8, (ncu * std::max(1, per_cu)) / 8 * 8);
    }();
    launch_fill_zero(counters, 8 * sizeof(int), s);
    kern<<<grid, DR_NT, 0, s>>>(P, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{H2G_EHIP, std::string("launch: ") + hipGetErrorString(e)};
}

void build_lattice_ea(const int32_t *tq_op, const int32_t *tq_sym, const int32_t *perms, const void *ops, int L,
                      bool half, bool want32, LatticeLevel &LL) {
    const int FX = half ? (L + 1) / 2 : L;
    LL.F = half ? half_freq_count(L) : L * L * L;
    const int64_t n = (int64_t)LL.F * 216 * 32;
    const unsigned nb = (unsigned)((n + 255) / 256);
    LL.EA64.alloc((size_t)n * sizeof(double));
    build_ea_kernel<double><<<nb, 256>>>(tq_op, tq_sym, perms, (const cplx<double> *)ops, L, FX, LL.F,
                                         LL.EA64.as<double>());
    if (want32) {
        LL.EA32.alloc((size_t)n * sizeof(float));
        build_ea_kernel<float><<<nb, 256>>>(tq_op, tq_sym, perms, (const cplx<double> *)ops, L, FX, LL.F,
                                            LL.EA32.as<float>());
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{H2G_EHIP, std::string("launch: ") + hipGetErrorString(e)};
    H2G_HIP(hipDeviceSynchronize());
}

}  // namespace h2g

// Internal declarations of libh2g: scalar types, error plumbing, device-side core and
// workspace structures, kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "h2g.h"

namespace h2g {

// ------------------------------------------------------------------------------------------
// Scalars: real (double / float) and interleaved complex (cplx<double> / cplx<float>).
// ------------------------------------------------------------------------------------------
template <class R>
struct cplx {
    R re, im;
};

template <class T> struct real_of { using type = T; };
template <class R> struct real_of<cplx<R>> { using type = R; };
template <class T> struct is_complex { static constexpr bool value = false; };
template <class R> struct is_complex<cplx<R>> { static constexpr bool value = true; };
template <class T> struct complex_of { using type = cplx<T>; };
template <class R> struct complex_of<cplx<R>> { using type = cplx<R>; };

__host__ __device__ inline double conjv(double a) { return a; }
__host__ __device__ inline float conjv(float a) { return a; }
template <class R>
__host__ __device__ inline cplx<R> conjv(cplx<R> a) { return {a.re, -a.im}; }

template <class R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__host__ __device__ inline cplx<R> operator*(R a, cplx<R> b) { return {a * b.re, a * b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, R b) { return {a.re * b, a.im * b}; }
template <class R>
__host__ __device__ inline cplx<R> &operator+=(cplx<R> &a, cplx<R> b) { a.re += b.re; a.im += b.im; return a; }

// acc += a * b (real or complex, same type)
__device__ inline void fmacc(double &acc, double a, double b) { acc = fma(a, b, acc); }
__device__ inline void fmacc(float &acc, float a, float b) { acc = fmaf(a, b, acc); }
template <class R>
__device__ inline void fmacc(cplx<R> &acc, cplx<R> a, cplx<R> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
// acc(complex) += a(real or complex) * b(complex)
template <class R>
__device__ inline void fmacc(cplx<R> &acc, R a, cplx<R> b) {
    acc.re = fma(a, b.re, acc.re);
    acc.im = fma(a, b.im, acc.im);
}

template <class T> __host__ __device__ inline T zero_v() { return T(0); }
template <> __host__ __device__ inline cplx<double> zero_v<cplx<double>>() { return {0.0, 0.0}; }
template <> __host__ __device__ inline cplx<float> zero_v<cplx<float>>() { return {0.0f, 0.0f}; }

// conversions between element types (complex -> real keeps the real part)
template <class To> struct cvt;
template <> struct cvt<double> {
    __host__ __device__ static double from(double a) { return a; }
    __host__ __device__ static double from(float a) { return a; }
    template <class R> __host__ __device__ static double from(cplx<R> a) { return (double)a.re; }
};
template <> struct cvt<float> {
    __host__ __device__ static float from(double a) { return (float)a; }
    __host__ __device__ static float from(float a) { return a; }
    template <class R> __host__ __device__ static float from(cplx<R> a) { return (float)a.re; }
};
template <class R> struct cvt<cplx<R>> {
    __host__ __device__ static cplx<R> from(double a) { return {(R)a, (R)0}; }
    __host__ __device__ static cplx<R> from(float a) { return {(R)a, (R)0}; }
    template <class S> __host__ __device__ static cplx<R> from(cplx<S> a) { return {(R)a.re, (R)a.im}; }
};

// scalar * value for the alpha/beta of the entry points (alpha carried as c64)
template <class T> __device__ inline T scale_by(cplx<double> s, T v);
template <> __device__ inline double scale_by<double>(cplx<double> s, double v) { return s.re * v; }
template <> __device__ inline float scale_by<float>(cplx<double> s, float v) { return (float)(s.re * v); }
template <> __device__ inline cplx<double> scale_by<cplx<double>>(cplx<double> s, cplx<double> v) { return s * v; }
template <> __device__ inline cplx<float> scale_by<cplx<float>>(cplx<double> s, cplx<float> v) {
    cplx<double> r = s * cplx<double>{(double)v.re, (double)v.im};
    return {(float)r.re, (float)r.im};
}

// ------------------------------------------------------------------------------------------
// Errors
// ------------------------------------------------------------------------------------------
void set_error(const std::string &msg);
struct Error {
    int code;
    std::string msg;
};
#define H2G_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::h2g::Error{H2G_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)};  \
    } while (0)

// after a kernel launch: turns a launch error into an Error
#define H2G_CHECK_LAUNCH()                                                                   \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess) throw Error{H2G_EHIP, std::string("launch: ") + hipGetErrorString(e_)}; \
    } while (0)

inline size_t elem_size(int dtype) {
    switch (dtype) {
    case H2G_F64: return 8;
    case H2G_C64: return 16;
    case H2G_F32: return 4;
    case H2G_C32: return 8;
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// Device buffers
// ------------------------------------------------------------------------------------------
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;  // false for a view (borrow) of memory the buffer does not free
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            throw Error{H2G_ENOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e)};
        }
        bytes = n;
    }
    void upload(const void *host, size_t n) {
        alloc(n);
        if (n) H2G_HIP(hipMemcpy(ptr, host, n, hipMemcpyHostToDevice));
    }
    template <class T> void upload_vec(const std::vector<T> &v) { upload(v.data(), v.size() * sizeof(T)); }
    // a view of n bytes at p (not freed by this buffer)
    void borrow(void *p, size_t n) {
        release();
        ptr = n ? p : nullptr;
        bytes = n;
        owned = false;
    }
    void release() {
        if (ptr && owned) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        owned = true;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Per-level list: [begin, begin+count) into a concatenated int32 device array.
struct Span {
    int64_t begin = 0;
    int32_t count = 0;
};

// A family of K x K operators in MFMA fragment order for the dense kernels (dense_m2l.hip): the
// real form (Kr = K rows, or 2K for complex operators) in KrT tiles of 16, per orientation
// (0: A, 1: A^T, 2: A^H, 3: conj(A)) and representation (FP64, FP32, split-BF16 planes).
constexpr int DENSE_MAX_RTW = 7;  // at most 4 * 7 row tiles of 16 (Kr <= 448)
enum { DENSE_REP_F64 = 1, DENSE_REP_F32 = 2, DENSE_REP_BF16 = 4, DENSE_REP_BF16X2 = 8, DENSE_REP_BF16X3 = 16 };
struct DenseOps {
    bool ok = false;
    bool cf = false;  // complex maps: f64 / f32 hold complex fragments (half the real form's values)
    int Kr = 0, KrT = 0;
    int64_t n = 0;
    DevBuf f64[4], f32[4], bf[4];
    DenseOps() = default;
    DenseOps &operator=(DenseOps &&o) noexcept {
        ok = o.ok;
        cf = o.cf;
        Kr = o.Kr;
        KrT = o.KrT;
        n = o.n;
        for (int i = 0; i < 4; ++i) {
            std::swap(f64[i].ptr, o.f64[i].ptr), std::swap(f64[i].bytes, o.f64[i].bytes);
            std::swap(f32[i].ptr, o.f32[i].ptr), std::swap(f32[i].bytes, o.f32[i].bytes);
            std::swap(bf[i].ptr, o.bf[i].ptr), std::swap(bf[i].bytes, o.bf[i].bytes);
        }
        return *this;
    }
};

// One side of the map.
struct Side {
    int64_t n_dofs = 0;
    int n_levels = 0, n_boxes = 0;
    std::vector<int32_t> level_offsets, parent, sector;
    std::vector<int64_t> begin, end;
    std::vector<uint8_t> is_leaf;
    std::vector<int64_t> perm;
    std::vector<int32_t> coord;  // [n_boxes][3] or empty
    DevBuf d_perm, d_iperm, d_begin, d_end;  // d_iperm: input index -> tree index
    DevBuf coll, coll32;            // K x n_dofs core type / reduced type
    // complex map with real transfers (sep_cplx), a real collection and a real alpha: the
    // collection kept real (FP64 / FP32, `coll` and `coll32` released), so P2M and L2P run the real
    // leaf kernels over x, y viewed as 2R real columns and planar expansion rows
    bool coll_real = false;
    DevBuf coll_re, coll32_re;
    DevBuf transfers, transfers32;  // [n_levels][8] K x K
    bool separable = false;         // every transfer is Az (x) Ay (x) Ax
    // complex map whose transfers are real and separable: `factors` hold their real factors, and
    // the transfers run on the real separable kernels over expansion rows made planar (all real
    // parts, then all imaginary parts), i.e. a real pool of twice the columns
    bool sep_cplx = false;
    bool sep_shared = false;        // ... with one factor per axis and child bit (stored per sector)
    int sep_p = 0;
    DevBuf factors, factors32;      // [n_levels*8][3][p][p]
    DenseOps tr_dense;              // the transfers on the dense MFMA kernel (when not separable)
    int level_of(int b) const {
        int l = 0;
        while (l + 1 < (int)level_offsets.size() && b >= level_offsets[l + 1]) ++l;
        return l;
    }
};

// Parent-block tables of one level for the lattice M2L kernels.
// Tiles of receiving parents for the lattice M2L kernels: up to 16 (cluster tiles) or LIST_NB
// (list tiles) parents, consecutive in Morton order, and the union of their present sources.
// List tiles: blocks per tile (one wave each) and the cap on the union of their present sources.
constexpr int LIST_NB = 12, LIST_NSLOT = 768, LIST_QL = 208;

struct LatticeTiles {
    int n_tiles = 0;
    DevBuf slot;   // int32 [tile][union size 1024 / LIST_NSLOT]: source slot (with couplings) or -1
    DevBuf npb;    // int32 [tile]: blocks in the tile
    DevBuf tgt;    // int32 [tile][blocks][8 children]: receiving index of the k-th block's children or -1
    DevBuf qt;     // int16 [tile][16][216]: union index of each slot q of the k-th block, 1024 if absent
                   // (cluster tiles)
    DevBuf list;   // int32 [tile][LIST_NB][208][2]: byte offsets (union index * 128, q * 128) of the
                   // k-th block's present slots in the list kernel's LDS region and coupling table,
                   // padded to a multiple of 16 entries with (LIST_NSLOT, 104) (list tiles)
    DevBuf nlist;  // int32 [tile][LIST_NB]: list lengths (multiples of 16)
};

// Direct lists (batches of >= 16 columns, FP32 spectra): blocks in Morton order, each with its
// present, coupled slots as slot pairs in batches of 8 pairs.  Batch k holds 16 entries (slot,
// q * 128) laid out [pair half h][pair i]: pair i couples slots of the two halves (one even and
// one odd q where possible); padding entries are (0, 104 * 128), a zero coupling row.
struct LatticeDirect {
    int n_blocks = 0, max_batches = 0;
    DevBuf ptr;   // int32 [n_blocks + 1]: first batch of each block
    DevBuf list;  // int32 [batches][2][8][2]
    DevBuf tgt;   // int32 [n_blocks][8]: receiving index of the block's children or -1
};

// Stored frequencies of the lattice levels' frequency-major spectra.  Full spectra: f = fx + L (fy
// + L fz).  Half spectra (real input, Hermitian couplings): a value at (fx, fy, fz) is the
// conjugate of the one at (-fx, -fy, -fz) (indices mod L), so only fx < (L + 1) / 2 is kept and,
// in the self-conjugate plane fx = 0, only one of each pair: fz = 0 with fy <= (L - 1) / 2, and
// 1 <= fz <= (L - 1) / 2 with any fy; (L^2 + 1) / 2 values, stored first, then the planes fx >= 1.
__host__ __device__ inline int half_freq_count(int L) { return (L * L + 1) / 2 + (L - 1) / 2 * L * L; }
__host__ __device__ inline bool half_freq_kept(int L, int fx, int fy, int fz) {
    return fx > 0 || (fz == 0 ? fy <= (L - 1) / 2 : fz <= (L - 1) / 2);
}
// index of a kept frequency
__host__ __device__ inline int half_freq_index(int L, int fx, int fy, int fz) {
    if (fx == 0) return fz == 0 ? fy : (L + 1) / 2 + (fz - 1) * L + fy;
    return (L * L + 1) / 2 + (fx - 1) + (L - 1) / 2 * (fy + L * fz);
}
__host__ __device__ inline void half_freq_of(int L, int f, int &fx, int &fy, int &fz) {
    const int h = (L * L + 1) / 2;
    if (f < h) {
        fx = 0;
        if (f < (L + 1) / 2) {
            fz = 0;
            fy = f;
        } else {
            fz = 1 + (f - (L + 1) / 2) / L;
            fy = (f - (L + 1) / 2) % L;
        }
    } else {
        const int FX1 = (L - 1) / 2, t = f - h;
        fx = 1 + t % FX1;
        fy = (t / FX1) % L;
        fz = t / FX1 / L;
    }
}

struct LatticeLevel {
    bool ok = false;
    int F = 0;                // stored frequencies (L^3, or half_freq_count(L) for half spectra)
    LatticeTiles clus;        // cluster tiles (any batch width)
    LatticeTiles list;        // list tiles (batches of >= 16 columns, FP32 spectra)
    LatticeDirect direct;     // direct lists (batches of >= 16 columns, FP32 spectra)
    DevBuf EA64;              // double [F][216 slots q][16 rows m = 2t + c][2 b]: real form of the couplings
    DevBuf EA32;              // the same in FP32 (cores with the mixed precision)
};

// Near field of one output leaf (or of its rows row0 .. row0 + nrow - 1, at most 64) as one
// product Y (nrow x R) = A (nrow x ncol) X (ncol x R) whose ncol columns are the source points of
// all the leaf's near blocks, concatenated.  Column j of A starts at value aoff (row 0) with row
// stride rstride (1 for a block applied as stored, the block's row count for its transpose), and
// multiplies source point xrow.
//
// For the mixed form of real maps the values are also kept per task in FP32, row-major with the
// columns padded to a multiple of 16 (zeros): row r of the task at foff + r * ncolp, so that a lane
// reads 4 consecutive columns with one 16-byte load; xoff is the task's first entry in the padded
// table of source points (ncolp int32 per leaf, padding entries repeat a valid point).
struct NearTask {
    int64_t out0;  // first output point of the leaf
    int64_t col0;  // first entry in the column table
    int32_t nrow, ncol, row0, pad;
    int64_t foff;  // first FP32 value of the task's row-major copy
    int64_t xoff;  // first entry of the leaf's padded source-point table
};
__host__ __device__ inline int64_t near_ncolp(int64_t ncol) { return (ncol + 15) & ~int64_t(15); }
struct NearCol {
    int64_t aoff;
    int32_t xrow, rstride;
};
struct NearPlan {
    DevBuf task, col;
    DevBuf frag;      // float: the tasks' row-major FP32 copies (mixed cores of real maps)
    DevBuf xrow;      // int32: the leaves' padded source-point tables
    int64_t n_frag = 0;
    int64_t cls[5] = {0, 0, 0, 0, 0};  // tasks of 4, 3, 2, 1 row tiles: [cls[i], cls[i+1])
    bool ok = false;
};

// Work lists of one orientation (forward: trial -> test; reverse: test -> trial).
struct OrientPlan {
    std::vector<LatticeLevel> lattice;  // per dst level
    const Side *src = nullptr, *dst = nullptr;
    DevBuf lists;                       // concatenated int32 box lists
    std::vector<Span> p2m, m2m, l2l, l2p, m2l_recv;  // per level
    std::vector<int32_t> l2p_avg;                    // per level: mean points per L2P leaf
    // M2M items: CSR over m2m targets (all levels concatenated); item = (src child box, matrix index)
    DevBuf m2m_ptr, m2m_src, m2m_mat;
    std::vector<int64_t> m2m_ptr_base;  // per level: offset of that level's targets into m2m_ptr
    // L2L items: one per target: (parent box, matrix index, accumulate flag)
    DevBuf l2l_ptr, l2l_src, l2l_mat, l2l_acc;
    std::vector<int64_t> l2l_base;
    // L2L grouped by parent (separable transfers): per level the parents, and per parent its
    // children (box, matrix index, accumulate flag) as a CSR
    std::vector<Span> l2l_par;            // per level: span of the parent list in l2lp_box
    std::vector<int64_t> l2lp_base;       // per level: offset of the level's parents into l2lp_ptr
    DevBuf l2lp_box, l2lp_ptr, l2lp_child, l2lp_mat, l2lp_acc;
    // M2L items: CSR over receiving boxes (concatenated by level); m2l_fop: the item's expanded
    // (translation id, symmetry) operator of the dense MFMA kernel
    DevBuf m2l_ptr, m2l_src, m2l_op, m2l_sym, m2l_fop;
    std::vector<int64_t> m2l_base;
    // dense M2L on receiver pairs (operators read once for both receivers of a pair): per level the
    // receivers paired with a lattice neighbour where one is free, else with the next free one; per
    // pair its items merged by operator (operator, source of each receiver or -1)
    DevBuf m2l2_recv, m2l2_ptr, m2l2_src, m2l2_fop;
    std::vector<int64_t> m2l2_base;   // per level: offset of the level's pairs into m2l2_ptr
    std::vector<int64_t> m2l2_rbase;  // per level: offset of the level's pairs into m2l2_recv (2 entries each)
    std::vector<int32_t> m2l2_count;  // per level: receiver pairs
    int min_level = 0, max_level = -1;
    int64_t max_src_level_boxes = 0, max_recv_level_boxes = 0;
    DevBuf out_mask;                    // uint8 per tree-ordered output dof: leaf initialised
};

struct Core {
    std::atomic<int> refs{1};
    int device = 0;
    int dtype = H2G_F64;
    int K = 0, p = 0, L = 0, L3 = 0;
    bool galerkin = true;
    Side test, trial_own;
    const Side *trial = nullptr;
    int m2l_kind = H2G_M2L_DENSE;
    int64_t n_pairs = 0, n_ops = 0;
    int n_perms = 0;
    DevBuf ops, ops32;        // dense: core type / reduced; toeplitz: c64 spectra / c32
    DevBuf perms, perms_inv;  // int32
    // near: blocks
    int64_t n_near_blocks = 0, n_near_values = 0;
    DevBuf near_vals;
    DevBuf near_fwd_ptr, near_fwd_blk;   // CSR over test leaves (by list index)
    DevBuf near_rev_ptr, near_rev_blk;   // CSR over trial leaves
    DevBuf near_fwd_tgt, near_rev_tgt;   // target leaf lists
    int32_t n_near_fwd_tgt = 0, n_near_rev_tgt = 0;
    DevBuf near_blk_rb, near_blk_cb, near_blk_nr, near_blk_nc, near_blk_off;  // per block
    // near field of real maps as one row-block product per output leaf (wide batches): per
    // orientation, tasks sorted by row-tile count and the concatenated source-column table
    NearPlan near_fwd_plan, near_rev_plan;
    cplx<double> alpha{1.0, 0.0};
    int precisions = 1;
    // counts for the diagnostics (h2g_core_summary / h2g_core_sanitycheck)
    int64_t n_unique_translations = 0, n_self_pairs = 0, n_bad_ops = 0;
    // a workspace precision other than PLAIN is supported (FP32 copies of the operators)
    bool reduced_precisions() const { return (precisions & ~(1 << H2G_PREC_PLAIN)) != 0; }
    OrientPlan fwd, rev;
    DevBuf twiddle;  // c64 [L]: exp(-2 pi i n / L)
    // dense / symmetry-reduced dense M2L on MFMA: every distinct (translation id, symmetry) of the
    // pairs expanded into MFMA fragment order
    DenseOps m2l_dense;
    bool half_spectrum = false;  // real map with Hermitian spectra: real-to-complex modulation
    int64_t device_bytes() const;
};

struct GraphKey {
    int op, which, nrhs;
    bool operator<(const GraphKey &o) const {
        return std::tie(op, which, nrhs) < std::tie(o.op, o.which, o.nrhs);
    }
};

struct Workspace {
    Core *core = nullptr;
    h2g_ws_opts opts{};
    std::atomic<bool> busy{false};
    hipStream_t stream = nullptr;
    // the near field runs on `side` (default priority) while the far field's upward pass and
    // M2L run on `stream`; the first far-field write of the result waits for it (ev_near)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_near = nullptr;
    hipEvent_t ev_done = nullptr;  // recorded after each product (h2g_ws_query / h2g_ws_sync)
    bool own_streams = true;       // false: stream (== side) belongs to the caller
    int R_cap = 1;
    DevBuf xs, ys;        // tree-ordered FP64/C64, [n][R]
    DevBuf x32;           // FP32 source columns of the mixed near field of real maps
    DevBuf mom, loc;      // [box][R][K] compute type
    DevBuf spec_in, spec_out;  // toeplitz spectra [slot][R][L3] complex compute type
    DevBuf counters;           // 8 int32: task counters of the persistent coupling stage
    DevBuf xhost, yhost;  // device staging for host pointers
    DevBuf scal;          // unused placeholder
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::vector<h2g_timing_record> timings;
    int cur_op = 0;  // orientation of the product being recorded
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t n_events_used = 0;
    // records [start, stop) of one phase when opts.timing is set (eager passes only)
    void mark_begin(int phase, int level);
    void mark_end();
    int64_t device_bytes() const;
};

// ------------------------------------------------------------------------------------------
// Kernel launchers (kernels.hip), all on `s`.
// T = compute scalar, S = storage scalar of xs/ys (double or cplx<double>).
// ------------------------------------------------------------------------------------------
// perm: tree index -> input index; iperm: its inverse (for the tiled batch forms)
void launch_gather(int dtype_core, int dtype_x, const void *x, int64_t ldx, const int64_t *perm,
                   const int64_t *iperm, int64_t n, int R, void *xs, hipStream_t s);
void launch_scatter(int dtype_core, int dtype_y, const void *ys, const int64_t *perm, const int64_t *iperm, int64_t n,
                    int R, void *y, int64_t ldy, cplx<double> a, cplx<double> b, const uint8_t *mask, hipStream_t s);
// vals32: read the FP32 copy of the near values where the core has one (mixed precision)
// x32: for the mixed form of real maps, a workspace buffer of near_x32_elems(n, R) floats that the
// near field fills with its FP32 source columns (null: the FP64 path)
// near_cols: columns per workgroup of the MFMA near kernels (16, 32, 64; 0: the default)
void launch_near(int dtype_core, int op, const Core &c, const void *xs, void *ys, int R, float *x32, int near_cols,
                 hipStream_t s);
// FP32 row-major near values of both orientations' tasks, from the FP64 values (after they are set)
void build_near_frags(Core &c);
// column tiles per workgroup of the mixed near kernel: the requested width, else the narrowest
// of 1, 2, 4 tiles that holds the batch
inline int near_x32_nct(int R, int near_cols = 0) {
    if (near_cols == 16 || near_cols == 32 || near_cols == 64) return near_cols / 16;
    return R <= 16 ? 1 : R <= 32 ? 2 : 4;
}
inline int64_t near_x32_pitch(int R, int near_cols = 0) {
    const int w = 16 * near_x32_nct(R, near_cols);
    return (int64_t)(R + w - 1) / w * w;
}
void launch_p2m(int cdtype, bool conj, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                const void *coll, int K, const void *xs, int R, void *mom, hipStream_t s);
void launch_l2p(int cdtype, bool conj, const int32_t *leaves, int n, const int64_t *begin, const int64_t *end,
                const void *coll, int K, const void *loc, int R, void *ys, cplx<double> alpha, int avg_points, hipStream_t s);
// out[t] = (acc ? out[t] : 0) + sum_items opA(mats[item_mat]) * in[item_src]  (optionally permuted)
// acc_mode: 0 overwrite, 1 accumulate, 2 accumulate when acc_flags[item_ptr[t]] (first item of t)
// opA: 0 = A, 1 = A^T, 2 = A^H, 3 = conj(A)
void launch_boxgemm(int cdtype, int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                    const int32_t *item_src, const int32_t *item_mat, const int32_t *item_sym,
                    const int32_t *perms, const int32_t *perms_inv, const void *mats, int K, const void *in,
                    void *out, int R, int acc_mode, const uint8_t *acc_flags, hipStream_t s);
void launch_fft_in(int cdtype, bool inverse, const int32_t *boxes, int n, int slot_base, const void *mom, int K,
                   int p, int L, int R, void *spec, const void *twiddle, hipStream_t s);
void launch_hadamard(int cdtype, bool conj, int n_recv, const int64_t *item_ptr, const int32_t *item_src,
                     const int32_t *item_op, const int32_t *item_sym, const int32_t *perms, const void *ops,
                     int L3, int src_slot_base, const void *spec_in, void *spec_out, int R, hipStream_t s);
void launch_fill_zero(void *p, size_t bytes, hipStream_t s);
// Diagonal couplings: out[t][r][k] = sum over the items of t of c(op)[k] in[src][r][k], c conjugated
// for the adjoint (overwrites out[t])
void launch_diag_m2l(int cdtype, bool conj, int n_recv, const int32_t *recv, const int64_t *item_ptr,
                     const int32_t *item_src, const int32_t *item_op, const void *ops, int K, const void *in, void *out,
                     int R, hipStream_t s);
// M2M of all children of each target (shared factors: Side::sep_shared), overwriting the target
// Real collection of a complex map (Side::coll_real) when `S` has real separable transfers, the core
// a real alpha and every collection entry a zero imaginary part; otherwise nothing changes.
void make_real_collection(Side &S, const Core &c);
// In place, for the boxes of a list: complex expansion rows (R per box, K values each) between the
// interleaved layout (K complex values) and the planar one (K real parts, then K imaginary parts).
// cdtype: H2G_C64 or H2G_C32.
void launch_cplx_planar(int cdtype, const int32_t *boxes, int n, void *pool, int R, int K, bool to_planar,
                        hipStream_t s);
void launch_sep_m2m(int cdtype, int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                    const int32_t *item_src, const int32_t *item_mat, const void *factors, int mat_base, int p,
                    void *mom, int R, hipStream_t s);
void launch_sep_l2l(int cdtype, int opA, int n_parents, const int32_t *parents, const int64_t *child_ptr,
                    const int32_t *child_box, const int32_t *child_mat, const uint8_t *child_acc, const void *factors,
                    int mat_base, int p, bool shared, void *loc, int R, hipStream_t s);
void launch_sep_transfer(int cdtype, int opA, int n_targets, const int32_t *targets, const int64_t *item_ptr,
                         const int32_t *item_src, const int32_t *item_mat, const void *factors, int mat_base, int p,
                         const void *in, void *out, int R, int acc_mode, const uint8_t *acc_flags, hipStream_t s);
// Rrow >= R: stored columns per row (padding columns zero); perm: the 32-column permuted layout
void launch_fft_in_fm(int cdtype, bool half, bool inverse, int box0, int nbox, const void *mom, int K, int p, int L,
                      int R, int Rrow, bool perm, void *S, const void *twiddle, hipStream_t s);
void launch_fft_out_fm(int cdtype, bool half, bool inverse, const int32_t *recv, int nrecv, const void *O, int K,
                       int p, int L, int R, void *loc, const void *twiddle, hipStream_t s);
// The lattice M2L of a level runs the direct kernel for FP32 spectra and R >= 4 columns; its
// source spectra are then stored in rows of this many columns (16: natural order, batches of at
// most 16 columns; a multiple of 32: the permuted layout of launch_fft_in_fm, perm).  0 when the
// level does not run the direct kernel.
int m2l_direct_rows(int cdtype, const LatticeLevel &LL, int n_src, int R);
// The direct coupling kernel of several lattice levels (FP32 batches on the direct path) in one
// persistent launch; counters: 8 int32 of device memory (zeroed on `s` by the launcher).
struct DirectLaunchLevel {
    const LatticeLevel *LL;
    int n_src, n_recv;
    const void *S;
    void *O;
};
void launch_m2l_direct_persistent(bool conj, const std::vector<DirectLaunchLevel> &levels, int R, int groups,
                                  int *counters, hipStream_t s);
// groups: workgroups per frequency of the direct kernel (0: the default)
void launch_m2l_lattice(int cdtype, bool conj, const LatticeLevel &LL, int n_src, int n_recv, const void *S, void *O,
                        int R, int groups, hipStream_t s);
void build_lattice_ea(const int32_t *tq_op, const int32_t *tq_sym, const int32_t *perms, const void *ops, int L,
                      bool half, bool want32, LatticeLevel &LL);
void launch_fft_out(int cdtype, bool inverse, const int32_t *boxes, int n, const void *spec, int K, int p, int L,
                    int R, void *loc, const void *twiddle, hipStream_t s);
// Builds D from n operators on the device (`ops`: core type, column-major K x K; operator e is
// ops[fop_op[e]] with symmetry fop_sym[e] applied, A_e[k][j] = C[perms_inv(k)][perms_inv(j)]),
// for the orientations of orient_mask and the representations of reps (DENSE_REP_F64 / F32 / BF16).
void build_dense_ops(DenseOps &D, int dtype, int K, const void *ops, const int32_t *fop_op, const int32_t *fop_sym,
                     const int32_t *perms_inv, int64_t n, int orient_mask, int reps);
// out[t] = (acc_flags && acc_flags[first item of t] ? out[t] : 0) + sum over the items of t of
// A_{orient}(fop(item)) in[src(item)], per receiving box t of `recv`; rep: DENSE_REP_F64 (FP64
// data), DENSE_REP_F32 (FP32 data) or DENSE_REP_BF16X2 / X3 (FP32 data, split-BF16 products).
void launch_dense(const DenseOps &D, int orient, int rep, const int32_t *recv, int n_recv, const int64_t *item_ptr,
                  const int32_t *item_src, const int32_t *item_fop, const void *in, void *out, int R,
                  const uint8_t *acc_flags, hipStream_t s);
// M2L on receiver pairs: per pair of receiving boxes recv2[2 i], recv2[2 i + 1] (-1: none), items
// item_ptr[i] .. item_ptr[i + 1] of (operator item_fop, sources item_src2[2 item], [2 item + 1], -1
// where that receiver has no such partner).  want: 1 pairs wherever the pair kernel applies, -1
// never, 0 for batches of at most 16 columns when the operators of the representation exceed
// DENSE_PAIR_MIN_BYTES.  False when pairs are not wanted or the map has no pair kernel (complex
// operators, split BF16); the caller then runs launch_dense on the unpaired items.
constexpr int64_t DENSE_PAIR_MIN_BYTES = 192ll << 20;
bool launch_dense_pair(const DenseOps &D, int orient, int rep, int want, const int32_t *recv2, int n_pairs,
                       const int64_t *item_ptr, const int32_t *item_src2, const int32_t *item_fop, const void *in,
                       void *out, int R, hipStream_t s);

}  // namespace h2g

struct h2g_core : h2g::Core {};
struct h2g_ws : h2g::Workspace {};

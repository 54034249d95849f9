// Octree and traversal plans built on the device: Morton codes and their sort, the level-by-level
// split of the boxes, the M2L pairs with their translation ids (and lattice symmetries), the near
// pairs with their value offsets.
//
// In the reference these come from H2Trees (not vendored): the tree a method passes to
// H2Factory, the plans of `assembleplans` (src/maps/interactionstrategy.jl:50-60, `isnear` and
// the translating-nodes iterator) and the translation ids of `translationmetadata`
// (src/farinteractions/i2otranslations/i2otranslationcollection.jl:19-84).  The definitions
// reproduced here are the ones of the host builder h2g/tree.py, bit for bit:
//   * points get 21-bit integer coordinates floor((x - origin) / size * 2^21), clipped, and are
//     ordered by Morton code (x in the lowest bit of each triple), stably;
//   * a box is split while it holds more than leaf_size points and its level is below
//     max_level; only non-empty children exist; boxes are stored level by level in Morton order;
//   * M2L pairs: same-level boxes whose parents are adjacent and that are not adjacent themselves,
//     grouped by test box with trial boxes ascending; the translation id of a pair is the rank of
//     its (level, offset) -- or (level, canonical offset) with a symmetry id for the symmetric
//     kinds -- among the present ones;
//   * near pairs: a test leaf and a trial leaf whose ancestors at the coarser of the two leaf
//     levels are adjacent (Chebyshev distance <= 1), grouped by test leaf, trial ascending; block
//     q holds nr x nc values at the exclusive prefix sum of the block sizes.
//
// Every kernel here is a plain one-thread-per-item kernel (no wave intrinsics), so the host replay
// of the library runs them under AddressSanitizer; sorting and scans go through prims.hip.
#include <algorithm>
#include <deque>
#include <memory>

#include "h2g_internal.h"

namespace h2g {

// device-wide primitives (prims.hip)
void prim_sort_keys_u64(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t s);
void prim_sort_pairs_u64(const uint64_t *kin, uint64_t *kout, const int64_t *vin, int64_t *vout, int64_t n,
                         int end_bit, hipStream_t s);
void prim_inclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);
void prim_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);

namespace {

constexpr int DEPTH = 21;
constexpr int TPB = 256;

inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

__device__ inline uint64_t interleave3(int64_t x, int64_t y, int64_t z, int bits) {
    uint64_t k = 0;
    for (int b = 0; b < bits; ++b)
        k |= ((uint64_t)((x >> b) & 1) << (3 * b)) | ((uint64_t)((y >> b) & 1) << (3 * b + 1)) |
             ((uint64_t)((z >> b) & 1) << (3 * b + 2));
    return k;
}

// Morton code of each point (tree order is the stable sort of these)
__global__ void morton_kernel(const double *__restrict__ pts, int64_t n, double o0, double o1, double o2,
                              double size, uint64_t *__restrict__ code, int64_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double scale = (double)(1 << DEPTH);
    const double o[3] = {o0, o1, o2};
    int64_t q[3];
    for (int a = 0; a < 3; ++a) {
        double t = floor((pts[3 * i + a] - o[a]) / size * scale);
        t = t < 0.0 ? 0.0 : (t > scale - 1.0 ? scale - 1.0 : t);
        q[a] = (int64_t)t;
    }
    code[i] = interleave3(q[0], q[1], q[2], DEPTH);
    idx[i] = i;
}

__global__ void split_kernel(const int64_t *__restrict__ beg, const int64_t *__restrict__ end, int64_t nb,
                             int64_t leaf, int can_split, int64_t *__restrict__ split) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    split[j] = (can_split && end[j] - beg[j] > leaf) ? 1 : 0;
}

// head[i] = 1 where point i opens a box of the next level inside a split box of this level
__global__ void heads_kernel(const uint64_t *__restrict__ code, int64_t n, const int32_t *__restrict__ pbox,
                             int64_t off, const int64_t *__restrict__ beg, const int64_t *__restrict__ split,
                             int shift, int64_t *__restrict__ head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t h = 0;
    const int32_t b = pbox[i];
    if (b >= 0) {
        const int64_t j = b - off;
        if (split[j]) h = (i == beg[j]) || ((code[i] >> shift) != (code[i - 1] >> shift));
    }
    head[i] = h;
}

__global__ void newbox_kernel(const uint64_t *__restrict__ code, int64_t n, const int64_t *__restrict__ head,
                              const int64_t *__restrict__ cnt, const int32_t *__restrict__ pbox, int shift,
                              int64_t *__restrict__ nbeg, uint64_t *__restrict__ nkey, int32_t *__restrict__ npar) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    const int64_t j = cnt[i] - 1;
    nbeg[j] = i;
    nkey[j] = code[i] >> shift;
    npar[j] = pbox[i];
}

__global__ void newend_kernel(int64_t nn, const int64_t *__restrict__ nbeg, const int32_t *__restrict__ npar,
                              const int64_t *__restrict__ end, int64_t off, int64_t *__restrict__ nend) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    nend[j] = (j + 1 < nn && npar[j + 1] == npar[j]) ? nbeg[j + 1] : end[npar[j] - off];
}

// each point moves to its box of the next level (or drops out once its box is a leaf)
__global__ void pbox_kernel(int64_t n, int64_t off, const int64_t *__restrict__ split, int64_t noff,
                            const int64_t *__restrict__ cnt, int32_t *__restrict__ pbox) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = pbox[i];
    pbox[i] = (b >= 0 && split[b - off]) ? (int32_t)(noff + cnt[i] - 1) : -1;
}

// per box: level, lattice coordinate, sector; the parent learns it has a child
__global__ void finish_kernel(int64_t nb, const int64_t *__restrict__ loff, int nlev, const uint64_t *__restrict__ key,
                              const int32_t *__restrict__ parent, int32_t *__restrict__ level,
                              int32_t *__restrict__ coord, int32_t *__restrict__ sector, uint8_t *__restrict__ has_child) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    int l = 0;
    while (l + 1 < nlev && loff[l + 1] <= j) ++l;
    level[j] = l;
    const uint64_t k = key[j];
    int32_t c[3] = {0, 0, 0};
    for (int b = 0; b < l; ++b)
        for (int a = 0; a < 3; ++a) c[a] |= (int32_t)((k >> (3 * b + a)) & 1) << b;
    coord[3 * j] = c[0];
    coord[3 * j + 1] = c[1];
    coord[3 * j + 2] = c[2];
    sector[j] = j == 0 ? -1 : (int32_t)(k & 7);
    if (j > 0) has_child[parent[j]] = 1;
}

__global__ void leaf_flag_kernel(int64_t nb, const uint8_t *__restrict__ has_child, uint8_t *__restrict__ is_leaf,
                                 const int64_t *__restrict__ beg, uint64_t *__restrict__ lkey,
                                 int64_t *__restrict__ lval) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const uint8_t lf = has_child[j] ? 0 : 1;
    is_leaf[j] = lf;
    // leaves first, in dof order (begin < 2^40); other boxes after them
    lkey[j] = lf ? (uint64_t)beg[j] : ((uint64_t)1 << 40);
    lval[j] = j;
}

// box id of the level-l box with integer coordinate c in a tree, or -1
__device__ inline int64_t find_box(const uint64_t *__restrict__ key, const int64_t *__restrict__ loff, int nlev,
                                   int l, int64_t cx, int64_t cy, int64_t cz) {
    if (l >= nlev) return -1;
    const int64_t n = (int64_t)1 << l;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= n || cy >= n || cz >= n) return -1;
    const uint64_t k = interleave3(cx, cy, cz, l);
    int64_t lo = loff[l], hi = loff[l + 1];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (key[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return (lo < loff[l + 1] && key[lo] == k) ? lo : -1;
}

__device__ inline int64_t lower_bound_i64(const int64_t *__restrict__ a, int64_t n, int64_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

struct TreeView {
    const uint64_t *key;
    const int64_t *loff;  // device copy of the level offsets
    int nlev;
    const int32_t *level, *coord, *sector;
    const int64_t *beg, *end;
    const uint8_t *is_leaf;
    const int64_t *leaf_beg;  // begins of the leaves in dof order
    const int64_t *leaf_id;   // the leaves in dof order
    int64_t n_leaves;
};

// M2L candidates of test box a: children of the parent's neighbours that are not adjacent.
// The offsets per axis run over [-2, 3] for a child bit 0 and [-3, 2] for bit 1.
template <bool FILL>
__global__ void m2l_kernel(int64_t nb, TreeView A, TreeView B, const int64_t *__restrict__ pos,
                           int64_t *__restrict__ cnt, uint64_t *__restrict__ out) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    const int l = A.level[a];
    int64_t c = 0;
    if (l >= 1 && l < B.nlev) {
        const int s = A.sector[a];
        const int64_t x = A.coord[3 * a], y = A.coord[3 * a + 1], z = A.coord[3 * a + 2];
        const int lx = (s & 1) ? -3 : -2, ly = (s & 2) ? -3 : -2, lz = (s & 4) ? -3 : -2;
        int64_t w = FILL ? pos[a] : 0;
        for (int oz = lz; oz < lz + 6; ++oz)
            for (int oy = ly; oy < ly + 6; ++oy)
                for (int ox = lx; ox < lx + 6; ++ox) {
                    if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
                    const int64_t b = find_box(B.key, B.loff, B.nlev, l, x + ox, y + oy, z + oz);
                    if (b < 0) continue;
                    if (FILL) out[w++] = ((uint64_t)a << 32) | (uint64_t)b;
                    ++c;
                }
    }
    if (!FILL) cnt[a] = c;
}

// Near pairs seen from the leaves of side A against the boxes of side B at the leaf's level.
// FIRST: the test side -- same-level leaves pair directly, non-leaf boxes give their leaf
// descendants.  Otherwise the trial side -- only non-leaf test boxes, their leaf descendants
// paired with the trial leaf (pairs stored as (test, trial)).
template <bool FILL, bool FIRST>
__global__ void near_kernel(int64_t nb, TreeView A, TreeView B, const int64_t *__restrict__ pos,
                            int64_t *__restrict__ cnt, uint64_t *__restrict__ out) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    int64_t c = 0;
    int64_t w = FILL ? pos[a] : 0;
    if (A.is_leaf[a]) {
        const int l = A.level[a];
        const int64_t x = A.coord[3 * a], y = A.coord[3 * a + 1], z = A.coord[3 * a + 2];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int64_t b = find_box(B.key, B.loff, B.nlev, l, x + dx, y + dy, z + dz);
                    if (b < 0) continue;
                    if (B.is_leaf[b]) {
                        if (FIRST) {
                            if (FILL) out[w++] = ((uint64_t)a << 32) | (uint64_t)b;
                            ++c;
                        }
                        continue;
                    }
                    const int64_t lo = lower_bound_i64(B.leaf_beg, B.n_leaves, B.beg[b]);
                    const int64_t hi = lower_bound_i64(B.leaf_beg, B.n_leaves, B.end[b]);
                    if (FILL)
                        for (int64_t k = lo; k < hi; ++k) {
                            const uint64_t d = (uint64_t)B.leaf_id[k];
                            out[w++] = FIRST ? (((uint64_t)a << 32) | d) : ((d << 32) | (uint64_t)a);
                        }
                    c += hi - lo;
                }
    }
    if (!FILL) cnt[a] = c;
}

// index of a signed permutation in the 48-element group, enumerated as permutations of (0,1,2)
// in lexicographic order times sign patterns (+,+,+), (+,+,-), ..., (-,-,-)
__device__ inline int group_index(const int perm[3], const int neg[3]) {
    return (perm[0] * 2 + (perm[1] > perm[2] ? 1 : 0)) * 8 + neg[0] * 4 + neg[1] * 2 + neg[2];
}

// per M2L pair: test/trial ids, translation code and (symmetric plans) symmetry id
__global__ void pair_code_kernel(int64_t np, const uint64_t *__restrict__ keys, TreeView A, TreeView B, int symmetric,
                                 int32_t *__restrict__ pt, int32_t *__restrict__ ps, int32_t *__restrict__ code,
                                 int32_t *__restrict__ sym, int32_t *__restrict__ present) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int64_t a = (int64_t)(keys[i] >> 32), b = (int64_t)(keys[i] & 0xffffffffu);
    pt[i] = (int32_t)a;
    ps[i] = (int32_t)b;
    int o[3];
    for (int ax = 0; ax < 3; ++ax) o[ax] = B.coord[3 * b + ax] - A.coord[3 * a + ax];
    int s = 0;
    if (symmetric) {
        // canonical offset: |o| sorted ascending (stable), S with S * canon == o
        int ab[3] = {o[0] < 0 ? -o[0] : o[0], o[1] < 0 ? -o[1] : o[1], o[2] < 0 ? -o[2] : o[2]};
        int order[3] = {0, 1, 2};
        for (int u = 1; u < 3; ++u)
            for (int v = u; v > 0 && ab[order[v - 1]] > ab[order[v]]; --v) {
                const int t = order[v];
                order[v] = order[v - 1];
                order[v - 1] = t;
            }
        int perm[3], neg[3];
        for (int u = 0; u < 3; ++u) perm[order[u]] = u;
        for (int u = 0; u < 3; ++u) neg[u] = o[u] < 0 ? 1 : 0;
        s = group_index(perm, neg);
        const int c0 = ab[order[0]], c1 = ab[order[1]], c2 = ab[order[2]];
        o[0] = c0, o[1] = c1, o[2] = c2;
    }
    const int c = ((A.level[a] * 8 + o[0] + 3) * 8 + o[1] + 3) * 8 + o[2] + 3;
    code[i] = c;
    if (sym) sym[i] = s;
    present[c] = 1;
}

__global__ void pair_op_kernel(int64_t np, const int32_t *__restrict__ code, const int32_t *__restrict__ rank,
                               int32_t *__restrict__ op) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    op[i] = rank[code[i]];
}

__global__ void near_size_kernel(int64_t np, const uint64_t *__restrict__ keys, TreeView A, TreeView B,
                                 int32_t *__restrict__ nt, int32_t *__restrict__ ns, int64_t *__restrict__ size) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int64_t a = (int64_t)(keys[i] >> 32), b = (int64_t)(keys[i] & 0xffffffffu);
    nt[i] = (int32_t)a;
    ns[i] = (int32_t)b;
    size[i] = (A.end[a] - A.beg[a]) * (B.end[b] - B.beg[b]);
}

__global__ void fill_i64_kernel(int64_t *p, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_u8_kernel(uint8_t *p, int64_t n, uint8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int bits_for(int64_t n) {
    int b = 1;
    while (b < 62 && ((int64_t)1 << b) < n) ++b;
    return b;
}

template <class T> T last_value(const DevBuf &b, int64_t n) {
    T v{};
    if (n > 0) H2G_HIP(hipMemcpy(&v, b.as<T>() + (n - 1), sizeof(T), hipMemcpyDeviceToHost));
    return v;
}

}  // namespace

// One side's tree on the device.
struct DevTree {
    int64_t n = 0;
    int nlev = 0;
    int64_t nb = 0, n_leaves = 0;
    std::vector<int64_t> loff;
    DevBuf perm, key, beg, end, parent, level, coord, sector, is_leaf, d_loff, leaf_beg, leaf_id;

    TreeView view() const {
        return TreeView{key.as<uint64_t>(), d_loff.as<int64_t>(), nlev, level.as<int32_t>(), coord.as<int32_t>(),
                        sector.as<int32_t>(), beg.as<int64_t>(), end.as<int64_t>(), is_leaf.as<uint8_t>(),
                        leaf_beg.as<int64_t>(), leaf_id.as<int64_t>(), n_leaves};
    }

    void build(const double *pts_host, int64_t npts, const h2g_plan_opts &o) {
        n = npts;
        hipStream_t s = 0;
        DevBuf pts, code, idx, scode;
        pts.upload(pts_host, (size_t)n * 3 * sizeof(double));
        code.alloc(n * 8), idx.alloc(n * 8), scode.alloc(n * 8), perm.alloc(n * 8);
        hipLaunchKernelGGL(morton_kernel, nblk(n), TPB, 0, s, pts.as<double>(), n, o.origin[0], o.origin[1],
                           o.origin[2], o.size, code.as<uint64_t>(), idx.as<int64_t>());
        H2G_CHECK_LAUNCH();
        prim_sort_pairs_u64(code.as<uint64_t>(), scode.as<uint64_t>(), idx.as<int64_t>(), perm.as<int64_t>(), n,
                            3 * DEPTH, s);
        const uint64_t *sc = scode.as<uint64_t>();
        const int max_level = o.max_level > 0 ? o.max_level : DEPTH;

        // level 0: the root
        std::deque<DevBuf> lb, le, lk, lp;  // per level: begin, end, key, parent
        auto push_level = [&](int64_t cnt) {
            lb.emplace_back(), le.emplace_back(), lk.emplace_back(), lp.emplace_back();
            lb.back().alloc(cnt * 8), le.back().alloc(cnt * 8), lk.back().alloc(cnt * 8), lp.back().alloc(cnt * 4);
        };
        push_level(1);
        {
            const int64_t b0 = 0, e0 = n;
            const uint64_t k0 = 0;
            const int32_t p0 = -1;
            H2G_HIP(hipMemcpy(lb[0].ptr, &b0, 8, hipMemcpyHostToDevice));
            H2G_HIP(hipMemcpy(le[0].ptr, &e0, 8, hipMemcpyHostToDevice));
            H2G_HIP(hipMemcpy(lk[0].ptr, &k0, 8, hipMemcpyHostToDevice));
            H2G_HIP(hipMemcpy(lp[0].ptr, &p0, 4, hipMemcpyHostToDevice));
        }
        loff = {0, 1};
        DevBuf pbox, head, cnt, split;
        pbox.alloc(n * 4), head.alloc(n * 8), cnt.alloc(n * 8);
        hipLaunchKernelGGL(fill_i32_kernel, nblk(n), TPB, 0, s, pbox.as<int32_t>(), n, 0);
        H2G_CHECK_LAUNCH();
        for (int l = 0;; ++l) {
            const int64_t nbl = loff[l + 1] - loff[l];
            split.alloc(nbl * 8);
            hipLaunchKernelGGL(split_kernel, nblk(nbl), TPB, 0, s, lb[l].as<int64_t>(), le[l].as<int64_t>(), nbl,
                               (int64_t)o.leaf_size, l < max_level ? 1 : 0, split.as<int64_t>());
            H2G_CHECK_LAUNCH();
            const int shift = 3 * (DEPTH - (l + 1));
            if (l >= max_level) break;
            hipLaunchKernelGGL(heads_kernel, nblk(n), TPB, 0, s, sc, n, pbox.as<int32_t>(), loff[l],
                               lb[l].as<int64_t>(), split.as<int64_t>(), shift, head.as<int64_t>());
            H2G_CHECK_LAUNCH();
            prim_inclusive_scan_i64(head.as<int64_t>(), cnt.as<int64_t>(), n, s);
            const int64_t nn = last_value<int64_t>(cnt, n);
            if (nn == 0) break;
            push_level(nn);
            hipLaunchKernelGGL(newbox_kernel, nblk(n), TPB, 0, s, sc, n, head.as<int64_t>(), cnt.as<int64_t>(),
                               pbox.as<int32_t>(), shift, lb[l + 1].as<int64_t>(), lk[l + 1].as<uint64_t>(),
                               lp[l + 1].as<int32_t>());
            H2G_CHECK_LAUNCH();
            hipLaunchKernelGGL(newend_kernel, nblk(nn), TPB, 0, s, nn, lb[l + 1].as<int64_t>(),
                               lp[l + 1].as<int32_t>(), le[l].as<int64_t>(), loff[l], le[l + 1].as<int64_t>());
            H2G_CHECK_LAUNCH();
            hipLaunchKernelGGL(pbox_kernel, nblk(n), TPB, 0, s, n, loff[l], split.as<int64_t>(), loff[l + 1],
                               cnt.as<int64_t>(), pbox.as<int32_t>());
            H2G_CHECK_LAUNCH();
            if (loff.back() + nn > INT32_MAX) throw Error{H2G_EINVAL, "plan: more than 2^31 boxes"};
            loff.push_back(loff.back() + nn);
        }
        nlev = (int)loff.size() - 1;
        nb = loff.back();
        beg.alloc(nb * 8), end.alloc(nb * 8), key.alloc(nb * 8), parent.alloc(nb * 4);
        for (int l = 0; l < nlev; ++l) {
            const int64_t c = loff[l + 1] - loff[l];
            H2G_HIP(hipMemcpy(beg.as<int64_t>() + loff[l], lb[l].ptr, c * 8, hipMemcpyDeviceToDevice));
            H2G_HIP(hipMemcpy(end.as<int64_t>() + loff[l], le[l].ptr, c * 8, hipMemcpyDeviceToDevice));
            H2G_HIP(hipMemcpy(key.as<uint64_t>() + loff[l], lk[l].ptr, c * 8, hipMemcpyDeviceToDevice));
            H2G_HIP(hipMemcpy(parent.as<int32_t>() + loff[l], lp[l].ptr, c * 4, hipMemcpyDeviceToDevice));
        }
        d_loff.upload(loff.data(), loff.size() * 8);
        level.alloc(nb * 4), coord.alloc(nb * 12), sector.alloc(nb * 4), is_leaf.alloc(nb);
        DevBuf has_child, lkey, lval, skey, sval;
        has_child.alloc(nb);
        hipLaunchKernelGGL(fill_u8_kernel, nblk(nb), TPB, 0, s, has_child.as<uint8_t>(), nb, (uint8_t)0);
        H2G_CHECK_LAUNCH();
        hipLaunchKernelGGL(finish_kernel, nblk(nb), TPB, 0, s, nb, d_loff.as<int64_t>(), nlev, key.as<uint64_t>(),
                           parent.as<int32_t>(), level.as<int32_t>(), coord.as<int32_t>(), sector.as<int32_t>(),
                           has_child.as<uint8_t>());
        H2G_CHECK_LAUNCH();
        lkey.alloc(nb * 8), lval.alloc(nb * 8), skey.alloc(nb * 8), sval.alloc(nb * 8);
        hipLaunchKernelGGL(leaf_flag_kernel, nblk(nb), TPB, 0, s, nb, has_child.as<uint8_t>(), is_leaf.as<uint8_t>(),
                           beg.as<int64_t>(), lkey.as<uint64_t>(), lval.as<int64_t>());
        H2G_CHECK_LAUNCH();
        prim_sort_pairs_u64(lkey.as<uint64_t>(), skey.as<uint64_t>(), lval.as<int64_t>(), sval.as<int64_t>(), nb, 41, s);
        std::vector<uint8_t> lf(nb);
        H2G_HIP(hipMemcpy(lf.data(), is_leaf.ptr, nb, hipMemcpyDeviceToHost));
        n_leaves = 0;
        for (uint8_t v : lf) n_leaves += v;
        // the sorted keys of the leaves are their begins
        leaf_beg.alloc(n_leaves * 8), leaf_id.alloc(n_leaves * 8);
        H2G_HIP(hipMemcpy(leaf_beg.ptr, skey.ptr, n_leaves * 8, hipMemcpyDeviceToDevice));
        H2G_HIP(hipMemcpy(leaf_id.ptr, sval.ptr, n_leaves * 8, hipMemcpyDeviceToDevice));
        H2G_HIP(hipDeviceSynchronize());
    }

    void copy_out(const h2g_plan_tree_out &t) const {
        auto cp = [](void *dst, const DevBuf &b, size_t bytes) {
            if (dst && bytes) H2G_HIP(hipMemcpy(dst, b.ptr, bytes, hipMemcpyDeviceToHost));
        };
        cp(t.perm, perm, n * 8);
        if (t.level_offsets) std::copy(loff.begin(), loff.end(), t.level_offsets);
        cp(t.level, level, nb * 4);
        cp(t.coord, coord, nb * 12);
        cp(t.key, key, nb * 8);
        cp(t.parent, parent, nb * 4);
        cp(t.sector, sector, nb * 4);
        cp(t.begin, beg, nb * 8);
        cp(t.end, end, nb * 8);
        cp(t.is_leaf, is_leaf, nb);
    }
};

}  // namespace h2g

struct h2g_plan {
    int device = 0;
    bool galerkin = true;
    h2g::DevTree test, trial;
    int64_t n_pairs = 0, n_near = 0, n_near_values = 0;
    bool symmetric = false;
    std::vector<int32_t> keys;  // [n_keys][4]
    h2g::DevBuf pair_test, pair_trial, pair_op, pair_sym, near_test, near_trial, near_offset;
};

using namespace h2g;

namespace {

void build_pairs(h2g_plan &P) {
    hipStream_t s = 0;
    const DevTree &A = P.test;
    const DevTree &B = P.galerkin ? P.test : P.trial;
    const TreeView va = A.view(), vb = B.view();
    const int end_bit = 32 + bits_for(A.nb);
    // M2L pairs
    {
        DevBuf cnt, pos, raw, sorted;
        cnt.alloc(A.nb * 8), pos.alloc(A.nb * 8);
        hipLaunchKernelGGL((m2l_kernel<false>), nblk(A.nb), TPB, 0, s, A.nb, va, vb, nullptr, cnt.as<int64_t>(),
                           nullptr);
        H2G_CHECK_LAUNCH();
        prim_exclusive_scan_i64(cnt.as<int64_t>(), pos.as<int64_t>(), A.nb, s);
        const int64_t np = last_value<int64_t>(pos, A.nb) + last_value<int64_t>(cnt, A.nb);
        P.n_pairs = np;
        if (np > INT32_MAX) throw Error{H2G_EINVAL, "plan: more than 2^31 M2L pairs"};
        raw.alloc(np * 8), sorted.alloc(np * 8);
        if (np > 0) {
            hipLaunchKernelGGL((m2l_kernel<true>), nblk(A.nb), TPB, 0, s, A.nb, va, vb, pos.as<int64_t>(), nullptr,
                               raw.as<uint64_t>());
            H2G_CHECK_LAUNCH();
            prim_sort_keys_u64(raw.as<uint64_t>(), sorted.as<uint64_t>(), np, end_bit, s);
        }
        const int ncode = (std::max(A.nlev, B.nlev) + 1) * 512;
        DevBuf code, present, rank;
        present.alloc((size_t)ncode * 4);
        hipLaunchKernelGGL(fill_i32_kernel, nblk(ncode), TPB, 0, s, present.as<int32_t>(), (int64_t)ncode, 0);
        H2G_CHECK_LAUNCH();
        P.pair_test.alloc(np * 4), P.pair_trial.alloc(np * 4), P.pair_op.alloc(np * 4), code.alloc(np * 4);
        if (P.symmetric) P.pair_sym.alloc(np * 4);
        if (np > 0) {
            hipLaunchKernelGGL(pair_code_kernel, nblk(np), TPB, 0, s, np, sorted.as<uint64_t>(), va, vb,
                               P.symmetric ? 1 : 0, P.pair_test.as<int32_t>(), P.pair_trial.as<int32_t>(),
                               code.as<int32_t>(), P.symmetric ? P.pair_sym.as<int32_t>() : nullptr,
                               present.as<int32_t>());
            H2G_CHECK_LAUNCH();
        }
        std::vector<int32_t> pres(ncode), rk(ncode, -1);
        H2G_HIP(hipMemcpy(pres.data(), present.ptr, (size_t)ncode * 4, hipMemcpyDeviceToHost));
        P.keys.clear();
        int32_t r = 0;
        for (int c = 0; c < ncode; ++c)
            if (pres[c]) {
                rk[c] = r++;
                P.keys.insert(P.keys.end(), {c / 512, (c / 64) % 8 - 3, (c / 8) % 8 - 3, c % 8 - 3});
            }
        rank.upload_vec(rk);
        if (np > 0) {
            hipLaunchKernelGGL(pair_op_kernel, nblk(np), TPB, 0, s, np, code.as<int32_t>(), rank.as<int32_t>(),
                               P.pair_op.as<int32_t>());
            H2G_CHECK_LAUNCH();
        }
        H2G_HIP(hipDeviceSynchronize());
    }
    // near pairs: from the test leaves, then from the trial leaves
    {
        DevBuf c1, p1, c2, p2, raw, sorted, size;
        c1.alloc(A.nb * 8), p1.alloc(A.nb * 8), c2.alloc(B.nb * 8), p2.alloc(B.nb * 8);
        hipLaunchKernelGGL((near_kernel<false, true>), nblk(A.nb), TPB, 0, s, A.nb, va, vb, nullptr,
                           c1.as<int64_t>(), nullptr);
        H2G_CHECK_LAUNCH();
        hipLaunchKernelGGL((near_kernel<false, false>), nblk(B.nb), TPB, 0, s, B.nb, vb, va, nullptr,
                           c2.as<int64_t>(), nullptr);
        H2G_CHECK_LAUNCH();
        prim_exclusive_scan_i64(c1.as<int64_t>(), p1.as<int64_t>(), A.nb, s);
        prim_exclusive_scan_i64(c2.as<int64_t>(), p2.as<int64_t>(), B.nb, s);
        const int64_t n1 = last_value<int64_t>(p1, A.nb) + last_value<int64_t>(c1, A.nb);
        const int64_t n2 = last_value<int64_t>(p2, B.nb) + last_value<int64_t>(c2, B.nb);
        const int64_t nn = n1 + n2;
        if (nn > INT32_MAX) throw Error{H2G_EINVAL, "plan: more than 2^31 near blocks"};
        P.n_near = nn;
        raw.alloc(nn * 8), sorted.alloc(nn * 8);
        if (n1 > 0) {
            hipLaunchKernelGGL((near_kernel<true, true>), nblk(A.nb), TPB, 0, s, A.nb, va, vb, p1.as<int64_t>(),
                               nullptr, raw.as<uint64_t>());
            H2G_CHECK_LAUNCH();
        }
        if (n2 > 0) {
            hipLaunchKernelGGL((near_kernel<true, false>), nblk(B.nb), TPB, 0, s, B.nb, vb, va, p2.as<int64_t>(),
                               nullptr, raw.as<uint64_t>() + n1);
            H2G_CHECK_LAUNCH();
        }
        if (nn > 0) prim_sort_keys_u64(raw.as<uint64_t>(), sorted.as<uint64_t>(), nn, end_bit, s);
        P.near_test.alloc(nn * 4), P.near_trial.alloc(nn * 4), P.near_offset.alloc(nn * 8), size.alloc(nn * 8);
        if (nn > 0) {
            hipLaunchKernelGGL(near_size_kernel, nblk(nn), TPB, 0, s, nn, sorted.as<uint64_t>(), va, vb,
                               P.near_test.as<int32_t>(), P.near_trial.as<int32_t>(), size.as<int64_t>());
            H2G_CHECK_LAUNCH();
            prim_exclusive_scan_i64(size.as<int64_t>(), P.near_offset.as<int64_t>(), nn, s);
        }
        P.n_near_values = last_value<int64_t>(P.near_offset, nn) + last_value<int64_t>(size, nn);
        H2G_HIP(hipDeviceSynchronize());
    }
}

template <class F>
int plan_guard(F &&f) {
    try {
        f();
        return H2G_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        set_error(e.what());
        return H2G_EINVAL;
    }
}

}  // namespace

extern "C" {

int h2g_plan_build(const double *test_points, int64_t n_test, const double *trial_points, int64_t n_trial,
                   const h2g_plan_opts *opts, h2g_plan **out) {
    return plan_guard([&] {
        if (!test_points || !opts || !out) throw Error{H2G_EINVAL, "h2g_plan_build: null argument"};
        *out = nullptr;
        const h2g_plan_opts &o = *opts;
        if (n_test < 1 || (trial_points && n_trial < 1)) throw Error{H2G_EINVAL, "h2g_plan_build: no points"};
        if (n_test > INT32_MAX || n_trial > INT32_MAX) throw Error{H2G_EINVAL, "h2g_plan_build: more than 2^31 points"};
        if (o.leaf_size < 1) throw Error{H2G_EINVAL, "h2g_plan_build: leaf_size must be >= 1"};
        if (o.max_level < 0 || o.max_level > 21) throw Error{H2G_EINVAL, "h2g_plan_build: max_level must be in 0..21"};
        if (!(o.size > 0.0)) throw Error{H2G_EINVAL, "h2g_plan_build: size must be positive"};
        int ndev = 0;
        H2G_HIP(hipGetDeviceCount(&ndev));
        if (o.device < 0 || o.device >= ndev) throw Error{H2G_EINVAL, "h2g_plan_build: device ordinal out of range"};
        H2G_HIP(hipSetDevice(o.device));
        auto P = std::make_unique<h2g_plan>();
        P->device = o.device;
        P->galerkin = trial_points == nullptr;
        P->symmetric = o.symmetric != 0;
        P->test.build(test_points, n_test, o);
        if (!P->galerkin) P->trial.build(trial_points, n_trial, o);
        build_pairs(*P);
        *out = P.release();
    });
}

int h2g_plan_info(const h2g_plan *plan, h2g_plan_counts *out) {
    return plan_guard([&] {
        if (!plan || !out) throw Error{H2G_EINVAL, "h2g_plan_info: null argument"};
        const DevTree &B = plan->galerkin ? plan->test : plan->trial;
        out->n_dofs[0] = plan->test.n, out->n_dofs[1] = B.n;
        out->n_levels[0] = plan->test.nlev, out->n_levels[1] = B.nlev;
        out->n_boxes[0] = (int32_t)plan->test.nb, out->n_boxes[1] = (int32_t)B.nb;
        out->n_pairs = plan->n_pairs;
        out->n_keys = (int64_t)plan->keys.size() / 4;
        out->n_near_pairs = plan->n_near;
        out->n_near_values = plan->n_near_values;
        out->galerkin = plan->galerkin ? 1 : 0;
    });
}

int h2g_plan_copy(const h2g_plan *plan, const h2g_plan_out *out) {
    return plan_guard([&] {
        if (!plan || !out) throw Error{H2G_EINVAL, "h2g_plan_copy: null argument"};
        H2G_HIP(hipSetDevice(plan->device));
        plan->test.copy_out(out->test);
        if (!plan->galerkin) plan->trial.copy_out(out->trial);
        auto cp = [](void *dst, const DevBuf &b, size_t bytes) {
            if (dst && bytes) H2G_HIP(hipMemcpy(dst, b.ptr, bytes, hipMemcpyDeviceToHost));
        };
        const size_t np = (size_t)plan->n_pairs, nn = (size_t)plan->n_near;
        cp(out->pair_test, plan->pair_test, np * 4);
        cp(out->pair_trial, plan->pair_trial, np * 4);
        cp(out->pair_op, plan->pair_op, np * 4);
        if (out->pair_sym) {
            if (!plan->symmetric) throw Error{H2G_EINVAL, "h2g_plan_copy: pair symmetries of a plan built without them"};
            cp(out->pair_sym, plan->pair_sym, np * 4);
        }
        if (out->keys) std::copy(plan->keys.begin(), plan->keys.end(), out->keys);
        cp(out->near_test, plan->near_test, nn * 4);
        cp(out->near_trial, plan->near_trial, nn * 4);
        cp(out->near_offset, plan->near_offset, nn * 8);
    });
}

int h2g_plan_destroy(h2g_plan *plan) {
    return plan_guard([&] {
        if (!plan) return;
        (void)hipSetDevice(plan->device);
        delete plan;
    });
}

}  // extern "C"

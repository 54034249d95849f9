// Host replay of libh2g -- TEST INFRASTRUCTURE ONLY.
//
// The MFMA near-field and leaf kernels (leaf_mfma.hip) and the lattice Toeplitz kernels
// (m2l_lattice.hip, m2l_mfma.hip) use gfx950 instructions and cannot run on the host.  These
// stand-ins let the replay link: the MFMA launchers decline, as they do on the GPU for complex
// element types, so a real map takes the same generic kernels here that a complex map takes
// there; the dense coupling kernel declines too, so dense maps take the generic box GEMM here; and
// the lattice entry points refuse -- the replay covers maps whose couplings run on the generic
// kernels (dense, symmetry-reduced dense, and Toeplitz maps built without lattice coordinates).
#include "h2g_internal.h"

namespace h2g {

bool launch_near_mfma(int, const Core &, const void *, void *, int, float *, int, hipStream_t) { return false; }
// the FP32 row-major near values feed only the MFMA near kernel, which declines here
void build_near_frags(Core &) {}
bool launch_p2m_mfma(int, const int32_t *, int, const int64_t *, const int64_t *, const void *, int, const void *,
                     int, void *, hipStream_t) {
    return false;
}
bool launch_l2p_mfma(int, const int32_t *, int, const int64_t *, const int64_t *, const void *, int, const void *,
                     int, void *, double, bool, hipStream_t) {
    return false;
}

// the dense MFMA kernels are not built: dense maps and general transfers take the generic box GEMM
void build_dense_ops(DenseOps &D, int, int, const void *, const int32_t *, const int32_t *, const int32_t *, int64_t,
                     int, int) {
    D.ok = false;
}
void launch_dense(const DenseOps &, int, int, const int32_t *, int, const int64_t *, const int32_t *, const int32_t *,
                  const void *, void *, int, const uint8_t *, hipStream_t) {
    throw Error{H2G_EUNSUPPORTED, "host replay: the dense MFMA kernels run on the GPU only"};
}
bool launch_dense_pair(const DenseOps &, int, int, int, const int32_t *, int, const int64_t *, const int32_t *,
                       const int32_t *, const void *, void *, int, hipStream_t) {
    throw Error{H2G_EUNSUPPORTED, "host replay: the dense MFMA kernels run on the GPU only"};
}

[[noreturn]] static void lattice_refused() {
    throw Error{H2G_EUNSUPPORTED, "host replay: the lattice Toeplitz kernels run on the GPU only"};
}
int m2l_direct_rows(int, const LatticeLevel &, int, int) { lattice_refused(); }
void launch_m2l_lattice(int, bool, const LatticeLevel &, int, int, const void *, void *, int, int, hipStream_t) {
    lattice_refused();
}
void build_lattice_ea(const int32_t *, const int32_t *, const int32_t *, const void *, int, bool, bool, LatticeLevel &) {
    lattice_refused();
}
void launch_fft_in_fm(int, bool, bool, int, int, const void *, int, int, int, int, int, bool, void *, const void *,
                      hipStream_t) {
    lattice_refused();
}
void launch_fft_out_fm(int, bool, bool, const int32_t *, int, const void *, int, int, int, int, void *, const void *,
                       hipStream_t) {
    lattice_refused();
}

}  // namespace h2g

// Device-wide sort and scan of the plan builder (prims.hip runs them on rocPRIM): host stand-ins.
// The radix sorts order keys by their low end_bit bits, stably; the callers' keys have no bits above.
#include <numeric>

namespace h2g {

void prim_sort_keys_u64(const uint64_t *in, uint64_t *out, int64_t n, int, hipStream_t) {
    std::vector<uint64_t> v(in, in + n);
    std::stable_sort(v.begin(), v.end());
    std::copy(v.begin(), v.end(), out);
}

void prim_sort_pairs_u64(const uint64_t *kin, uint64_t *kout, const int64_t *vin, int64_t *vout, int64_t n, int,
                         hipStream_t) {
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return kin[a] < kin[b]; });
    std::vector<uint64_t> k(n);
    std::vector<int64_t> v(n);
    for (int64_t i = 0; i < n; ++i) k[i] = kin[idx[i]], v[i] = vin[idx[i]];
    std::copy(k.begin(), k.end(), kout);
    std::copy(v.begin(), v.end(), vout);
}

void prim_inclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t) {
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) out[i] = (s += in[i]);
}

void prim_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t) {
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t v = in[i];
        out[i] = s;
        s += v;
    }
}

}  // namespace h2g

namespace h2g {
// the persistent coupling stage is a lattice kernel too
void launch_m2l_direct_persistent(bool, const std::vector<DirectLaunchLevel> &, int, int, int *, hipStream_t) {
    throw Error{H2G_EUNSUPPORTED, "host replay: the lattice Toeplitz kernels run on the GPU only"};
}
}  // namespace h2g

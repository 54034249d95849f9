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

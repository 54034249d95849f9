// Device assembly of the interpolation method's data: near-field blocks of a point kernel and
// tensor-Lagrange leaf collections.  In the reference this is the downstream method's work
// (the near assemblers of src/nearinteractions/strategies.jl:67-81 and the collection
// builders the method hands to the core); here it runs on the GPU so that a two-million-point
// map is built in seconds without staging its 22 GB near field through host memory.
#include <cmath>

#include "h2g_internal.h"

namespace h2g {

void launch_fill_zero(void *p, size_t bytes, hipStream_t s);

namespace {

constexpr double FOUR_PI = 12.566370614359172;

template <class S>
__device__ inline S kernel_value(int kind, double k, double r);

template <>
__device__ inline double kernel_value<double>(int, double, double r) {
    return r > 0.0 ? 1.0 / (FOUR_PI * r) : 0.0;
}

template <>
__device__ inline cplx<double> kernel_value<cplx<double>>(int kind, double k, double r) {
    if (!(r > 0.0)) return {0.0, 0.0};
    const double g = 1.0 / (FOUR_PI * r);
    if (kind == 0) return {g, 0.0};
    double s, c;
    sincos(k * r, &s, &c);
    return {g * c, g * s};
}

// one workgroup per near block; column-major nr x nc values
template <class S>
__global__ void near_fill_kernel(const int64_t *__restrict__ rb, const int64_t *__restrict__ cb,
                                 const int32_t *__restrict__ nr, const int32_t *__restrict__ nc,
                                 const int64_t *__restrict__ off, const int64_t *__restrict__ perm_t,
                                 const int64_t *__restrict__ perm_s, const double *__restrict__ pt,
                                 const double *__restrict__ ps, int kind, double wavenumber, S *__restrict__ vals) {
    const int q = blockIdx.x;
    const int m = nr[q], n = nc[q];
    S *B = vals + off[q];
    for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
        const int i = e % m, j = e / m;
        const int64_t a = perm_t[rb[q] + i], b = perm_s[cb[q] + j];
        const double dx = pt[3 * a] - ps[3 * b];
        const double dy = pt[3 * a + 1] - ps[3 * b + 1];
        const double dz = pt[3 * a + 2] - ps[3 * b + 2];
        B[e] = kernel_value<S>(kind, wavenumber, sqrt(dx * dx + dy * dy + dz * dz));
    }
}

// one workgroup per leaf; C[k + K*j] = prod_axes l_{k_axis}((y_j - c)/h)
template <class S>
__global__ void coll_fill_kernel(const int32_t *__restrict__ leaves, const int64_t *__restrict__ begin,
                                 const int64_t *__restrict__ end, const int64_t *__restrict__ perm,
                                 const double *__restrict__ pts, const double *__restrict__ center,
                                 const double *__restrict__ halfw, const double *__restrict__ nodes, int p,
                                 S *__restrict__ coll) {
    const int b = leaves[blockIdx.x];
    const int64_t j0 = begin[b], n = end[b] - j0;
    const int K = p * p * p;
    const double h = halfw[b];
    for (int64_t e = threadIdx.x; e < n * K; e += blockDim.x) {
        const int64_t j = e / K;
        const int k = (int)(e - j * K);
        const int64_t g = perm[j0 + j];
        int kk[3] = {k % p, (k / p) % p, k / (p * p)};
        double v = 1.0;
        for (int ax = 0; ax < 3; ++ax) {
            const double t = (pts[3 * g + ax] - center[3 * b + ax]) / h;
            const int ka = kk[ax];
            double l = 1.0;
            for (int mm = 0; mm < p; ++mm)
                if (mm != ka) l *= (t - nodes[mm]) / (nodes[ka] - nodes[mm]);
            v *= l;
        }
        coll[(j0 + j) * K + k] = cvt<S>::from(v);
    }
}

__global__ void narrow_kernel(const double *__restrict__ src, float *__restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

void fill_side_collection(const Side &S, const h2g_method_side_desc &m, int dtype, int p, const double *d_nodes) {
    DevBuf pts, cen, hw, leaves;
    pts.upload(m.points, (size_t)S.n_dofs * 3 * sizeof(double));
    cen.upload(m.box_center, (size_t)S.n_boxes * 3 * sizeof(double));
    hw.upload(m.box_halfwidth, (size_t)S.n_boxes * sizeof(double));
    std::vector<int32_t> lv;
    for (int b = 0; b < S.n_boxes; ++b)
        if (S.is_leaf[b]) lv.push_back(b);
    leaves.upload_vec(lv);
    if (lv.empty()) return;
    if (dtype == H2G_F64)
        hipLaunchKernelGGL((coll_fill_kernel<double>), (unsigned)lv.size(), 256, 0, 0, leaves.as<int32_t>(), S.d_begin.as<int64_t>(),
                                                               S.d_end.as<int64_t>(), S.d_perm.as<int64_t>(),
                                                               pts.as<double>(), cen.as<double>(), hw.as<double>(),
                                                               d_nodes, p, S.coll.as<double>());
    else
        hipLaunchKernelGGL((coll_fill_kernel<cplx<double>>), (unsigned)lv.size(), 256, 0, 0,
            leaves.as<int32_t>(), S.d_begin.as<int64_t>(), S.d_end.as<int64_t>(), S.d_perm.as<int64_t>(),
            pts.as<double>(), cen.as<double>(), hw.as<double>(), d_nodes, p, S.coll.as<cplx<double>>());
    H2G_HIP(hipGetLastError());
    H2G_HIP(hipDeviceSynchronize());
    if (S.coll32.ptr) {
        const int64_t n = (int64_t)(S.coll.bytes / sizeof(double));
        hipLaunchKernelGGL((narrow_kernel), (unsigned)((n + 255) / 256), 256, 0, 0, S.coll.as<double>(), S.coll32.as<float>(), n);
        H2G_HIP(hipGetLastError());
        H2G_HIP(hipDeviceSynchronize());
    }
}

}  // namespace
}  // namespace h2g

using namespace h2g;

extern "C" int h2g_method_fill(h2g_core *core, const h2g_method_desc *desc) {
    try {
        if (!core || !desc) throw Error{H2G_EINVAL, "null argument"};
        const h2g_method_desc &d = *desc;
        Core &c = *core;
        if (d.kernel != 0 && d.kernel != 1) throw Error{H2G_EINVAL, "kernel must be 0 (Laplace) or 1 (Helmholtz)"};
        if (d.kernel == 1 && c.dtype != H2G_C64) throw Error{H2G_EINVAL, "the Helmholtz kernel needs a ComplexF64 core"};
        H2G_HIP(hipSetDevice(c.device));
        const Side &tr = *c.trial;
        if (d.fill_near && c.n_near_blocks > 0) {
            if (!d.test.points || !(c.galerkin || d.trial.points)) throw Error{H2G_EINVAL, "missing points"};
            DevBuf pt, ps;
            pt.upload(d.test.points, (size_t)c.test.n_dofs * 3 * sizeof(double));
            const double *dps = pt.as<double>();
            if (!c.galerkin) {
                ps.upload(d.trial.points, (size_t)tr.n_dofs * 3 * sizeof(double));
                dps = ps.as<double>();
            }
            const unsigned nb = (unsigned)c.n_near_blocks;
            if (c.dtype == H2G_F64)
                hipLaunchKernelGGL((near_fill_kernel<double>), nb, 256, 0, 0, c.near_blk_rb.as<int64_t>(), c.near_blk_cb.as<int64_t>(),
                                                      c.near_blk_nr.as<int32_t>(), c.near_blk_nc.as<int32_t>(),
                                                      c.near_blk_off.as<int64_t>(), c.test.d_perm.as<int64_t>(),
                                                      tr.d_perm.as<int64_t>(), pt.as<double>(), dps, d.kernel,
                                                      d.wavenumber, c.near_vals.as<double>());
            else
                hipLaunchKernelGGL((near_fill_kernel<cplx<double>>), nb, 256, 0, 0,
                    c.near_blk_rb.as<int64_t>(), c.near_blk_cb.as<int64_t>(), c.near_blk_nr.as<int32_t>(),
                    c.near_blk_nc.as<int32_t>(), c.near_blk_off.as<int64_t>(), c.test.d_perm.as<int64_t>(),
                    tr.d_perm.as<int64_t>(), pt.as<double>(), dps, d.kernel, d.wavenumber,
                    c.near_vals.as<cplx<double>>());
            H2G_HIP(hipGetLastError());
            H2G_HIP(hipDeviceSynchronize());
            if (c.dtype == H2G_F64) build_near_frags(c);
            H2G_HIP(hipDeviceSynchronize());
        }
        if (d.fill_collections) {
            if (d.p < 1 || d.p * d.p * d.p != c.K || !d.nodes) throw Error{H2G_EINVAL, "collections need K == p^3 and nodes"};
            if (c.test.coll_real || c.trial->coll_real)
                throw Error{H2G_EINVAL, "the collections were filled already (and are kept real)"};
            DevBuf nodes;
            nodes.upload(d.nodes, (size_t)d.p * sizeof(double));
            fill_side_collection(c.test, d.test, c.dtype, d.p, nodes.as<double>());
            if (!c.galerkin) fill_side_collection(*c.trial, d.trial, c.dtype, d.p, nodes.as<double>());
            // interpolation collections are real for every kernel
            make_real_collection(c.test, c);
            if (!c.galerkin) make_real_collection(c.trial_own, c);
        }
        return H2G_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
}

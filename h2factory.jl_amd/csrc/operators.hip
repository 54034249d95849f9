// Operators of the point-kernel interpolation method built on the device: dense couplings
// (kernel values between two boxes' interpolation grids), Toeplitz kernel spectra (the FP64
// 3-D DFT of the kernel sampled on the difference grid, one workgroup per translation id,
// computed in LDS), the level transfers (Lagrange bases of the parent at the children's nodes)
// and the lattice-symmetry permutation tables of the symmetry-reduced kinds.
//
// The reference takes these from the method through its translation builders
// (src/farinteractions/i2otranslations/i2otranslationcollection.jl:52-70: one operator per
// translation id from a representative pair; ext/H2Toeplitz/toeplitzi2ocollection.jl:14-31: the
// spectra D = fft(kappa)), the level translator (src/maps/leveltranslator.jl:100-158) and the
// permutation builders (src/farinteractions/i2otranslations/symmetryreducedcollection.jl:162-232:
// toeplitzpermutation applies S^-1 to wrap-order differences, latticegridpermutation reverses
// the grid per negative sign).  The conventions are the ones of the core description (h2g.h).
#include <cmath>

#include "h2g_internal.h"

namespace h2g {
namespace {

constexpr double FOUR_PI = 12.566370614359172;
constexpr int TPB = 256;

// G(r) of the kernel: Laplace 1/(4 pi r), Helmholtz exp(i k r)/(4 pi r); 0 at r = 0
__device__ inline void kernel_g(int kind, double k, double r, double &re, double &im) {
    if (!(r > 0.0)) {
        re = 0.0, im = 0.0;
        return;
    }
    const double g = 1.0 / (FOUR_PI * r);
    if (kind == 0) {
        re = g, im = 0.0;
        return;
    }
    double s, c;
    sincos(k * r, &s, &c);
    re = c * g, im = s * g;
}

// one thread per entry (key o, row k, column j): M_o[k, j] = G(x_k - y_j - width * offset),
// x_k = grid point k of the receiving box at the origin, column-major K x K per key
template <bool CPLX>
__global__ void dense_ops_kernel(int64_t n_keys, const int32_t *__restrict__ keys, const double *__restrict__ nodes,
                                 int p, double root, int kind, double wavenumber, double *__restrict__ out) {
    const int K = p * p * p;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_keys * K * K) return;
    const int64_t o = e / ((int64_t)K * K);
    const int rem = (int)(e - o * K * K);
    const int j = rem / K, k = rem % K;
    const int lvl = keys[4 * o];
    const double width = root / (double)(1ll << lvl), h = width / 2.0;
    const int ka[3] = {k % p, (k / p) % p, k / (p * p)};
    const int ja[3] = {j % p, (j / p) % p, j / (p * p)};
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double d = (nodes[ka[a]] * h - nodes[ja[a]] * h) - width * (double)keys[4 * o + 1 + a];
        r2 += d * d;
    }
    double re, im;
    kernel_g(kind, wavenumber, sqrt(r2), re, im);
    if (CPLX) {
        out[2 * e] = re;
        out[2 * e + 1] = im;
    } else {
        out[e] = re;
    }
}

// one workgroup per key: kappa[z][y][x] = G(delta (d - (p-1) o)) on the wrap-order difference
// grid d in {0..p-1, -(p-1)..-1}, delta = width / (p-1), then D = 3-D DFT (exp(-2 pi i f n / L))
// along x, y and z in LDS.  LDS: two L^3 complex buffers + the twiddles.
__global__ void toeplitz_ops_kernel(const int32_t *__restrict__ keys, int p, double root, int kind, double wavenumber,
                                    double *__restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *lds = reinterpret_cast<double *>(smem_raw);
    const int L = 2 * p - 1, L3 = L * L * L;
    double *a = lds;             // [L3][2]
    double *b = lds + 2 * L3;    // [L3][2]
    double *tw = lds + 4 * L3;   // [L][2]: exp(-2 pi i m / L)
    const int o = blockIdx.x;
    const int lvl = keys[4 * o];
    const double width = root / (double)(1ll << lvl);
    const double delta = width / (double)(p - 1);
    const double ox = keys[4 * o + 1], oy = keys[4 * o + 2], oz = keys[4 * o + 3];
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)m / (double)L, &s, &c);
        tw[2 * m] = c;
        tw[2 * m + 1] = -s;
    }
    for (int e = threadIdx.x; e < L3; e += blockDim.x) {
        const int x = e % L, y = (e / L) % L, z = e / (L * L);
        const double dx = x <= L / 2 ? x : x - L, dy = y <= L / 2 ? y : y - L, dz = z <= L / 2 ? z : z - L;
        const double vx = delta * (dx - (double)(p - 1) * ox);
        const double vy = delta * (dy - (double)(p - 1) * oy);
        const double vz = delta * (dz - (double)(p - 1) * oz);
        double re, im;
        kernel_g(kind, wavenumber, sqrt(vx * vx + vy * vy + vz * vz), re, im);
        a[2 * e] = re;
        a[2 * e + 1] = im;
    }
    __syncthreads();
    double *src = a, *dst = b;
    for (int axis = 0; axis < 3; ++axis) {
        const int stride = axis == 0 ? 1 : (axis == 1 ? L : L * L);
        for (int e = threadIdx.x; e < L3; e += blockDim.x) {
            const int f = (e / stride) % L;       // output frequency along the axis
            const int base = e - f * stride;      // the line's first element
            double sr = 0.0, si = 0.0;
            for (int n = 0; n < L; ++n) {
                const int m = (f * n) % L;
                const double vr = src[2 * (base + n * stride)], vi = src[2 * (base + n * stride) + 1];
                sr += vr * tw[2 * m] - vi * tw[2 * m + 1];
                si += vr * tw[2 * m + 1] + vi * tw[2 * m];
            }
            dst[2 * e] = sr;
            dst[2 * e + 1] = si;
        }
        __syncthreads();
        double *t = src;
        src = dst;
        dst = t;
    }
    double *D = out + (int64_t)o * 2 * L3;
    for (int e = threadIdx.x; e < L3; e += blockDim.x) {
        D[2 * e] = src[2 * e];
        D[2 * e + 1] = src[2 * e + 1];
    }
}

__device__ inline double lagrange1(const double *__restrict__ nodes, int p, int k, double t) {
    double v = 1.0;
    for (int m = 0; m < p; ++m)
        if (m != k) v *= (t - nodes[m]) / (nodes[k] - nodes[m]);
    return v;
}

// T_s[k, i] = prod_axes l_{k_a}(((bit_a(s) * 2 - 1) + nodes[i_a]) / 2): the parent's basis k at
// the child's node i, child of sector s; column-major K x K per sector
__global__ void transfer_ops_kernel(const double *__restrict__ nodes, int p, double *__restrict__ out) {
    const int K = p * p * p;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 8ll * K * K) return;
    const int s = (int)(e / ((int64_t)K * K));
    const int rem = (int)(e - (int64_t)s * K * K);
    const int i = rem / K, k = rem % K;
    const int ka[3] = {k % p, (k / p) % p, k / (p * p)};
    const int ia[3] = {i % p, (i / p) % p, i / (p * p)};
    double v = 1.0;
    for (int a = 0; a < 3; ++a) {
        const double t = ((double)(((s >> a) & 1) * 2 - 1) + nodes[ia[a]]) / 2.0;
        v *= lagrange1(nodes, p, ka[a], t);
    }
    out[e] = v;
}

// the signed permutation of group index g: perm[3] and sign[3] (+1 / -1); permutations of
// (0, 1, 2) in lexicographic order times sign patterns (+++), (++-), ..., (---)
__device__ inline void group_element(int g, int perm[3], int sign[3]) {
    const int pi = g / 8, si = g % 8;
    const int first = pi / 2;
    int rest[2], r = 0;
    for (int v = 0; v < 3; ++v)
        if (v != first) rest[r++] = v;
    perm[0] = first;
    perm[1] = (pi & 1) ? rest[1] : rest[0];
    perm[2] = (pi & 1) ? rest[0] : rest[1];
    sign[0] = (si & 4) ? -1 : 1;
    sign[1] = (si & 2) ? -1 : 1;
    sign[2] = (si & 1) ? -1 : 1;
}

// SYMDENSE: perm_g[i] = index of the image of grid point i (component a of the image is
// u[perm[a]], reversed u -> p-1-u where sign[a] < 0).
// SYMTOEPLITZ: perm_g[j] = wrap index of S^-1 d_j, (S^-1 v)[a] = sign_inv[a] v[perm_inv[a]].
__global__ void perm_ops_kernel(int toeplitz, int n, int *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int N3 = n * n * n;
    if (e >= 48ll * N3) return;
    const int g = (int)(e / N3), j = (int)(e % N3);
    int perm[3], sign[3];
    group_element(g, perm, sign);
    const int u[3] = {j % n, (j / n) % n, j / (n * n)};
    int img[3];
    if (!toeplitz) {
        for (int a = 0; a < 3; ++a) img[a] = sign[a] > 0 ? u[perm[a]] : n - 1 - u[perm[a]];
    } else {
        int inv[3], isg[3];
        for (int a = 0; a < 3; ++a) inv[perm[a]] = a;
        for (int a = 0; a < 3; ++a) isg[a] = sign[inv[a]];
        int d[3];
        for (int a = 0; a < 3; ++a) d[a] = u[a] <= n / 2 ? u[a] : u[a] - n;
        for (int a = 0; a < 3; ++a) img[a] = ((isg[a] * d[inv[a]]) % n + n) % n;
    }
    out[e] = img[0] + n * (img[1] + n * img[2]);
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

}  // namespace
}  // namespace h2g

using namespace h2g;

extern "C" int h2g_method_operators(const h2g_ops_desc *desc, const h2g_ops_out *out) {
    try {
        if (!desc || !out) throw Error{H2G_EINVAL, "h2g_method_operators: null argument"};
        const h2g_ops_desc &d = *desc;
        if (d.kernel != 0 && d.kernel != 1) throw Error{H2G_EINVAL, "kernel must be 0 (Laplace) or 1 (Helmholtz)"};
        if (d.p < 1 || d.p > 16 || !d.nodes) throw Error{H2G_EINVAL, "h2g_method_operators: need 1 <= p <= 16 and nodes"};
        if (!(d.root_size > 0.0)) throw Error{H2G_EINVAL, "h2g_method_operators: root_size must be positive"};
        if (d.n_keys < 0 || (d.n_keys > 0 && !d.keys)) throw Error{H2G_EINVAL, "h2g_method_operators: missing keys"};
        const bool toep = d.m2l_kind == H2G_M2L_TOEPLITZ || d.m2l_kind == H2G_M2L_SYMTOEPLITZ;
        const bool dense = d.m2l_kind == H2G_M2L_DENSE || d.m2l_kind == H2G_M2L_SYMDENSE;
        if (!toep && !dense) throw Error{H2G_EINVAL, "h2g_method_operators: m2l_kind must be a dense or Toeplitz kind"};
        if (toep && d.p < 2) throw Error{H2G_EINVAL, "h2g_method_operators: Toeplitz couplings need p >= 2"};
        for (int64_t o = 0; o < d.n_keys; ++o) {
            const int32_t *k = d.keys + 4 * o;
            if (k[0] < 0 || k[0] > 40) throw Error{H2G_EINVAL, "h2g_method_operators: key level out of range"};
        }
        int ndev = 0;
        H2G_HIP(hipGetDeviceCount(&ndev));
        if (d.device < 0 || d.device >= ndev) throw Error{H2G_EINVAL, "h2g_method_operators: device ordinal out of range"};
        H2G_HIP(hipSetDevice(d.device));
        const int p = d.p, K = p * p * p, L = 2 * p - 1, L3 = L * L * L;
        DevBuf nodes, keys;
        nodes.upload(d.nodes, (size_t)p * 8);
        if (d.n_keys > 0) keys.upload(d.keys, (size_t)d.n_keys * 16);
        if (out->ops && d.n_keys > 0) {
            DevBuf ops;
            if (dense) {
                const bool cp = d.kernel == 1;
                const int64_t n = d.n_keys * K * K;
                ops.alloc((size_t)n * (cp ? 16 : 8));
                if (cp)
                    hipLaunchKernelGGL((dense_ops_kernel<true>), nblk(n), TPB, 0, 0, d.n_keys, keys.as<int32_t>(),
                                       nodes.as<double>(), p, d.root_size, d.kernel, d.wavenumber, ops.as<double>());
                else
                    hipLaunchKernelGGL((dense_ops_kernel<false>), nblk(n), TPB, 0, 0, d.n_keys, keys.as<int32_t>(),
                                       nodes.as<double>(), p, d.root_size, d.kernel, d.wavenumber, ops.as<double>());
            } else {
                ops.alloc((size_t)d.n_keys * L3 * 16);
                const size_t lds = (size_t)(4 * L3 + 2 * L) * sizeof(double);
                if (lds > 160 * 1024) throw Error{H2G_EINVAL, "h2g_method_operators: spectrum grid too large for LDS"};
                hipLaunchKernelGGL(toeplitz_ops_kernel, (unsigned)d.n_keys, TPB, lds, 0, keys.as<int32_t>(), p,
                                   d.root_size, d.kernel, d.wavenumber, ops.as<double>());
            }
            H2G_CHECK_LAUNCH();
            H2G_HIP(hipMemcpy(out->ops, ops.ptr, ops.bytes, hipMemcpyDeviceToHost));
        }
        if (out->transfers) {
            DevBuf tr;
            const int64_t n = 8ll * K * K;
            tr.alloc((size_t)n * 8);
            hipLaunchKernelGGL(transfer_ops_kernel, nblk(n), TPB, 0, 0, nodes.as<double>(), p, tr.as<double>());
            H2G_CHECK_LAUNCH();
            H2G_HIP(hipMemcpy(out->transfers, tr.ptr, tr.bytes, hipMemcpyDeviceToHost));
        }
        if (out->perms && (d.m2l_kind == H2G_M2L_SYMDENSE || d.m2l_kind == H2G_M2L_SYMTOEPLITZ)) {
            DevBuf pm;
            const int n = d.m2l_kind == H2G_M2L_SYMTOEPLITZ ? L : p;
            const int64_t cnt = 48ll * n * n * n;
            pm.alloc((size_t)cnt * 4);
            hipLaunchKernelGGL(perm_ops_kernel, nblk(cnt), TPB, 0, 0, d.m2l_kind == H2G_M2L_SYMTOEPLITZ ? 1 : 0, n,
                               pm.as<int>());
            H2G_CHECK_LAUNCH();
            H2G_HIP(hipMemcpy(out->perms, pm.ptr, pm.bytes, hipMemcpyDeviceToHost));
        }
        return H2G_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    }
}

// C ABI of libh2g: core / workspace lifetime, validation, traversal plans and the matvec.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <array>
#include <complex>
#include <set>
#include <unordered_map>
#include <map>
#include <mutex>

#include "h2g_internal.h"

namespace h2g {

void launch_fill_zero(void *p, size_t bytes, hipStream_t s);

namespace {
thread_local std::string g_last_error;

constexpr int VERSION = 100;  // 0.1.0

template <class F>
int guarded(F &&f) {
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

[[noreturn]] void fail(int code, const std::string &msg) { throw Error{code, msg}; }

bool is_complex_dtype(int d) { return d == H2G_C64 || d == H2G_C32; }

// ------------------------------------------------------------------------------------------
// Validation (host only, before any device call)
// ------------------------------------------------------------------------------------------
void validate_side(const h2g_side_desc &s, const char *name, int K) {
    std::string n(name);
    if (s.n_dofs < 0 || s.n_levels < 1 || s.n_boxes < 1) fail(H2G_EINVAL, n + ": empty tree");
    if (!s.perm || !s.level_offsets || !s.parent || !s.sector || !s.dof_begin || !s.dof_end || !s.is_leaf ||
        !s.transfers)
        fail(H2G_EINVAL, n + ": missing array");
    if (s.level_offsets[0] != 0 || s.level_offsets[s.n_levels] != s.n_boxes)
        fail(H2G_EDIM, n + ": level_offsets must run from 0 to n_boxes");
    for (int l = 0; l < s.n_levels; ++l)
        if (s.level_offsets[l + 1] < s.level_offsets[l]) fail(H2G_EINVAL, n + ": level_offsets not monotone");
    std::vector<uint8_t> seen(s.n_dofs, 0);
    for (int64_t i = 0; i < s.n_dofs; ++i) {
        if (s.perm[i] < 0 || s.perm[i] >= s.n_dofs || seen[s.perm[i]]) fail(H2G_EINVAL, n + ": perm is not a permutation");
        seen[s.perm[i]] = 1;
    }
    for (int l = 0; l < s.n_levels; ++l)
        for (int b = s.level_offsets[l]; b < s.level_offsets[l + 1]; ++b) {
            if (s.dof_begin[b] < 0 || s.dof_end[b] < s.dof_begin[b] || s.dof_end[b] > s.n_dofs)
                fail(H2G_EDIM, n + ": box dof range out of bounds");
            if (l == 0) {
                if (s.parent[b] != -1) fail(H2G_EINVAL, n + ": level-0 box with a parent");
                continue;
            }
            const int p = s.parent[b];
            if (p < s.level_offsets[l - 1] || p >= s.level_offsets[l]) fail(H2G_EINVAL, n + ": parent not on the level above");
            if (s.sector[b] < 0 || s.sector[b] > 7) fail(H2G_EINVAL, n + ": sector out of 0..7");
            if (s.is_leaf[p]) fail(H2G_EINVAL, n + ": a leaf has children");
            if (s.dof_begin[b] < s.dof_begin[p] || s.dof_end[b] > s.dof_end[p]) fail(H2G_EINVAL, n + ": child range outside parent");
            if (b > s.level_offsets[l] && s.parent[b] < s.parent[b - 1]) fail(H2G_EINVAL, n + ": children not grouped by parent");
        }
    if (s.level_offsets[1] != 1) fail(H2G_EINVAL, n + ": level 0 must hold exactly one root box");
    if (s.coord)
        for (int b = 1; b < s.n_boxes; ++b)
            for (int a = 0; a < 3; ++a)
                if (s.coord[3 * b + a] != 2 * s.coord[3 * s.parent[b] + a] + ((s.sector[b] >> a) & 1))
                    fail(H2G_EINVAL, n + ": box coordinate is not 2*parent + sector bits");
    // leaves tile the dofs
    std::vector<uint8_t> cover(s.n_dofs, 0);
    for (int b = 0; b < s.n_boxes; ++b)
        if (s.is_leaf[b])
            for (int64_t j = s.dof_begin[b]; j < s.dof_end[b]; ++j) {
                if (cover[j]) fail(H2G_EINVAL, n + ": leaves overlap");
                cover[j] = 1;
            }
    for (int64_t j = 0; j < s.n_dofs; ++j)
        if (!cover[j]) fail(H2G_EINVAL, n + ": leaves do not cover every dof");
    (void)K;
}

void validate_desc(const h2g_core_desc &d) {
    if (d.dtype != H2G_F64 && d.dtype != H2G_C64) fail(H2G_EINVAL, "core dtype must be H2G_F64 or H2G_C64");
    if (d.K < 1) fail(H2G_EINVAL, "K must be positive");
    if (d.m2l_kind < 0 || d.m2l_kind > H2G_M2L_DIAGONAL) fail(H2G_EINVAL, "unknown m2l_kind");
    const bool toep = d.m2l_kind == H2G_M2L_TOEPLITZ || d.m2l_kind == H2G_M2L_SYMTOEPLITZ;
    const bool sym = d.m2l_kind == H2G_M2L_SYMDENSE || d.m2l_kind == H2G_M2L_SYMTOEPLITZ;
    if (toep) {
        if (d.p < 2 || d.p * d.p * d.p != d.K)
            fail(H2G_EINVAL, "Toeplitz couplings need K == p^3 with p >= 2 (ToeplitzModulation)");
        if (d.p > 6) fail(H2G_EUNSUPPORTED, "Toeplitz couplings support p <= 6");
    }
    validate_side(d.test, "test", d.K);
    const h2g_side_desc &tr = d.galerkin ? d.test : d.trial;
    if (!d.galerkin) validate_side(d.trial, "trial", d.K);
    if (d.n_pairs < 0 || (d.n_pairs > 0 && (!d.pair_test || !d.pair_trial || !d.pair_op)))
        fail(H2G_EINVAL, "missing coupling pair arrays");
    if (d.n_pairs > 0 && d.n_ops < 1) fail(H2G_EINVAL, "couplings without operators");
    if (d.n_ops > 0 && !d.ops) fail(H2G_EINVAL, "missing coupling operators");
    const int L = 2 * d.p - 1;
    const int64_t plen = d.m2l_kind == H2G_M2L_SYMDENSE ? d.K : (int64_t)L * L * L;
    if (sym) {
        if (d.n_pairs > 0 && (!d.pair_sym || d.n_perms < 1 || !d.perms))
            fail(H2G_EINVAL, "symmetry-reduced couplings need pair_sym and perms");
        for (int64_t q = 0; q < (int64_t)d.n_perms * plen; ++q)
            if (d.perms[q] < 0 || d.perms[q] >= plen)
                fail(H2G_EDIM, "permutation entry out of range: the permutation was built for a different grid");
        for (int q = 0; q < d.n_perms; ++q) {
            std::vector<uint8_t> seen(plen, 0);
            for (int64_t i = 0; i < plen; ++i) {
                int v = d.perms[q * plen + i];
                if (seen[v]) fail(H2G_EINVAL, "perms entry is not a permutation");
                seen[v] = 1;
            }
        }
    }
    for (int64_t k = 0; k < d.n_pairs; ++k) {
        const int b = d.pair_test[k], c = d.pair_trial[k];
        if (b < 0 || b >= d.test.n_boxes || c < 0 || c >= tr.n_boxes) fail(H2G_EINVAL, "pair box out of range");
        if (d.pair_op[k] < 0 || d.pair_op[k] >= d.n_ops) fail(H2G_EINVAL, "pair operator id out of range");
        if (sym && (d.pair_sym[k] < 0 || d.pair_sym[k] >= d.n_perms)) fail(H2G_EINVAL, "pair permutation id out of range");
        if (toep) {
            int lb = 0, lc = 0;
            while (b >= d.test.level_offsets[lb + 1]) ++lb;
            while (c >= tr.level_offsets[lc + 1]) ++lc;
            if (lb != lc) fail(H2G_EUNSUPPORTED, "Toeplitz couplings must join boxes of one level");
        }
    }
    if (d.n_near_blocks < 0 || (d.n_near_blocks > 0 && (!d.near_test || !d.near_trial || !d.near_offset)))
        fail(H2G_EINVAL, "missing near-field arrays");
    for (int64_t q = 0; q < d.n_near_blocks; ++q) {
        const int b = d.near_test[q], c = d.near_trial[q];
        if (b < 0 || b >= d.test.n_boxes || c < 0 || c >= tr.n_boxes || !d.test.is_leaf[b] || !tr.is_leaf[c])
            fail(H2G_EINVAL, "near block must join a test leaf and a trial leaf");
        const int64_t nr = d.test.dof_end[b] - d.test.dof_begin[b], nc = tr.dof_end[c] - tr.dof_begin[c];
        if (d.near_offset[q] < 0 || d.near_offset[q] + nr * nc > d.n_near_values)
            fail(H2G_EDIM, "near block values out of range");
    }
}

// ------------------------------------------------------------------------------------------
// Upload helpers
// ------------------------------------------------------------------------------------------
// Kronecker factors of every transfer (T = Az (x) Ay (x) Ax, p x p each), if they exist.
template <class V>
bool extract_factors(const V *tr, int n_levels, int K, int p, std::vector<V> &fac) {
    auto mag = [](const V &v) { return std::abs(v); };
    fac.assign((size_t)n_levels * 8 * 3 * p * p, V(0));
    for (int l = 1; l < n_levels; ++l)
        for (int s = 0; s < 8; ++s) {
            const V *T = tr + ((size_t)l * 8 + s) * K * K;
            size_t best = 0;
            double mx = 0.0;
            for (size_t i = 0; i < (size_t)K * K; ++i)
                if (mag(T[i]) > mx) mx = mag(T[i]), best = i;
            if (mx == 0.0) return false;
            const int ks = (int)(best % K), js = (int)(best / K);
            const int kx = ks % p, ky = (ks / p) % p, kz = ks / (p * p);
            const int jx = js % p, jy = (js / p) % p, jz = js / (p * p);
            auto at = [&](int a, int b, int c, int d, int e, int f) {
                return T[(a + p * (b + p * c)) + (size_t)K * (d + p * (e + p * f))];
            };
            const V piv = at(kx, ky, kz, jx, jy, jz);
            V *F = fac.data() + ((size_t)l * 8 + s) * 3 * p * p;
            for (int a = 0; a < p; ++a)
                for (int b = 0; b < p; ++b) {
                    F[a + p * b] = at(a, ky, kz, b, jy, jz);
                    F[p * p + a + p * b] = at(kx, a, kz, jx, b, jz) / piv;
                    F[2 * p * p + a + p * b] = at(kx, ky, a, jx, jy, b) / piv;
                }
            for (int c = 0; c < p; ++c)
                for (int f = 0; f < p; ++f)
                    for (int b = 0; b < p; ++b)
                        for (int e = 0; e < p; ++e)
                            for (int a = 0; a < p; ++a)
                                for (int d = 0; d < p; ++d) {
                                    const V v = F[a + p * d] * F[p * p + b + p * e] * F[2 * p * p + c + p * f];
                                    if (mag(v - at(a, b, c, d, e, f)) > 1e-12 * mx) return false;
                                }
        }
    return true;
}

// Shared factors: when each sector's transfer is C_z[bit z] (x) C_y[bit y] (x) C_x[bit x] with one
// p x p factor per axis and child bit (interpolation transfers), rewrite the per-sector factors of
// every level as those shared ones (same values in every sector with the same bit), so that
// kernels may share line stages between sectors.  The per-sector factors fix each sector up to
// a scale per axis; C_a[1] takes the scales of the other two axes relative to sector 0.  Returns
// false (factors unchanged) unless every sector is reproduced to 1e-12.
template <class V>
bool share_factors(const V *tr, int n_levels, int K, int p, std::vector<V> &fac) {
    auto mag = [](const V &v) { return std::abs(v); };
    const int pp = p * p;
    std::vector<V> out(fac);
    for (int l = 1; l < n_levels; ++l) {
        auto Fs = [&](int sec, int ax) { return fac.data() + (((size_t)l * 8 + sec) * 3 + ax) * pp; };
        auto ratio = [&](const V *X, const V *Y) {  // X = ratio * Y
            int best = 0;
            for (int i = 1; i < pp; ++i)
                if (mag(Y[i]) > mag(Y[best])) best = i;
            return X[best] / Y[best];
        };
        std::vector<V> C(3 * 2 * pp);  // [axis][bit][p x p]
        for (int ax = 0; ax < 3; ++ax) {
            std::copy(Fs(0, ax), Fs(0, ax) + pp, C.begin() + (ax * 2) * pp);
            V sc = V(1);
            for (int o = 0; o < 3; ++o)
                if (o != ax) sc *= ratio(Fs(1 << ax, o), Fs(0, o));
            for (int i = 0; i < pp; ++i) C[(ax * 2 + 1) * pp + i] = Fs(1 << ax, ax)[i] * sc;
        }
        for (int sec = 0; sec < 8; ++sec) {
            const V *T = tr + ((size_t)l * 8 + sec) * K * K;
            double mx = 0.0;
            for (size_t i = 0; i < (size_t)K * K; ++i) mx = std::max(mx, (double)mag(T[i]));
            const V *Cx = C.data() + (0 * 2 + (sec & 1)) * pp, *Cy = C.data() + (1 * 2 + ((sec >> 1) & 1)) * pp,
                    *Cz = C.data() + (2 * 2 + ((sec >> 2) & 1)) * pp;
            for (int c = 0; c < p; ++c)
                for (int f = 0; f < p; ++f)
                    for (int b = 0; b < p; ++b)
                        for (int e = 0; e < p; ++e)
                            for (int a = 0; a < p; ++a)
                                for (int d = 0; d < p; ++d) {
                                    const V v = Cx[a + p * d] * Cy[b + p * e] * Cz[c + p * f];
                                    const V t = T[(a + p * (b + p * c)) + (size_t)K * (d + p * (e + p * f))];
                                    if (mag(v - t) > 1e-12 * mx) return false;
                                }
            std::copy(Cx, Cx + pp, out.begin() + (((size_t)l * 8 + sec) * 3 + 0) * pp);
            std::copy(Cy, Cy + pp, out.begin() + (((size_t)l * 8 + sec) * 3 + 1) * pp);
            std::copy(Cz, Cz + pp, out.begin() + (((size_t)l * 8 + sec) * 3 + 2) * pp);
        }
    }
    fac.swap(out);
    return true;
}

void copy_side(Side &S, const h2g_side_desc &d, int dtype, int K, bool want32) {
    S.n_dofs = d.n_dofs;
    S.n_levels = d.n_levels;
    S.n_boxes = d.n_boxes;
    S.level_offsets.assign(d.level_offsets, d.level_offsets + d.n_levels + 1);
    S.parent.assign(d.parent, d.parent + d.n_boxes);
    S.sector.assign(d.sector, d.sector + d.n_boxes);
    S.begin.assign(d.dof_begin, d.dof_begin + d.n_boxes);
    S.end.assign(d.dof_end, d.dof_end + d.n_boxes);
    S.is_leaf.assign(d.is_leaf, d.is_leaf + d.n_boxes);
    S.perm.assign(d.perm, d.perm + d.n_dofs);
    if (d.coord) S.coord.assign(d.coord, d.coord + 3 * (size_t)d.n_boxes);
    S.d_perm.upload_vec(S.perm);
    {
        // inverse permutation when perm is one (every input index exactly once)
        std::vector<int64_t> inv(S.n_dofs, -1);
        bool ok = true;
        for (int64_t i = 0; i < S.n_dofs && ok; ++i) {
            const int64_t j = S.perm[i];
            if (j < 0 || j >= S.n_dofs || inv[j] >= 0) ok = false;
            else inv[j] = i;
        }
        if (ok) S.d_iperm.upload_vec(inv);
    }
    S.d_begin.upload_vec(S.begin);
    S.d_end.upload_vec(S.end);
    const size_t es = elem_size(dtype);
    const size_t ncoll = (size_t)K * d.n_dofs, ntr = (size_t)d.n_levels * 8 * K * K;
    if (d.collection)
        S.coll.upload(d.collection, ncoll * es);
    else
        S.coll.alloc(ncoll * es);
    S.transfers.upload(d.transfers, ntr * es);
    {
        int p = 1;
        while (p * p * p < K) ++p;
        // Separable transfers run only for real maps: the register-line kernel faulted the GPU
        // for ComplexF64 (p = 3) with the cause not yet found, so complex maps use the dense
        // transfer kernel.
        if (dtype == H2G_F64 && p * p * p == K && p > 1 && p <= 6 && d.n_levels > 1) {
            bool ok;
            size_t nf;
            if (dtype == H2G_F64) {
                std::vector<double> fac;
                ok = extract_factors((const double *)d.transfers, d.n_levels, K, p, fac);
                if (ok) {
                    S.sep_shared = share_factors((const double *)d.transfers, d.n_levels, K, p, fac);
                    S.factors.upload_vec(fac);
                }
                nf = fac.size();
            } else {
                std::vector<std::complex<double>> fac;
                ok = extract_factors((const std::complex<double> *)d.transfers, d.n_levels, K, p, fac);
                if (ok) S.factors.upload_vec(fac);
                nf = fac.size() * 2;
            }
            if (ok) {
                S.separable = true;
                S.sep_p = p;
                if (want32) {
                    std::vector<double> host(nf);
                    H2G_HIP(hipMemcpy(host.data(), S.factors.ptr, nf * sizeof(double), hipMemcpyDeviceToHost));
                    std::vector<float> f32(host.begin(), host.end());
                    S.factors32.upload_vec(f32);
                }
            }
        }
        // complex maps whose transfers are real (interpolation transfers are, whatever the kernel):
        // their real factors, for the real separable kernels on planar expansion rows
        if (dtype == H2G_C64 && p * p * p == K && p > 1 && p <= 6 && d.n_levels > 1) {
            const double *tc = (const double *)d.transfers;  // (re, im) pairs
            bool real = true;
            for (size_t i = 0; i < ntr && real; ++i) real = tc[2 * i + 1] == 0.0;
            if (real) {
                std::vector<double> tr(ntr), fac;
                for (size_t i = 0; i < ntr; ++i) tr[i] = tc[2 * i];
                if (extract_factors(tr.data(), d.n_levels, K, p, fac)) {
                    S.sep_shared = share_factors(tr.data(), d.n_levels, K, p, fac);
                    S.factors.upload_vec(fac);
                    if (want32) {
                        std::vector<float> f32(fac.begin(), fac.end());
                        S.factors32.upload_vec(f32);
                    }
                    S.sep_cplx = true;
                    S.sep_p = p;
                }
            }
        }
    }
    if (want32) {
        auto narrow = [&](const void *src, size_t n, DevBuf &dst) {
            if (dtype == H2G_F64) {
                std::vector<float> v(n);
                const double *s = (const double *)src;
                for (size_t i = 0; i < n; ++i) v[i] = (float)s[i];
                dst.upload_vec(v);
            } else {
                std::vector<float> v(2 * n);
                const double *s = (const double *)src;
                for (size_t i = 0; i < 2 * n; ++i) v[i] = (float)s[i];
                dst.upload_vec(v);
            }
        };
        if (d.collection)
            narrow(d.collection, ncoll, S.coll32);
        else
            S.coll32.alloc(ncoll * es / 2);
        narrow(d.transfers, ntr, S.transfers32);
    }
}

// Spectrum elements per column one orientation needs: the largest level's source boxes (in) and
// receiving boxes (out) times the frequencies the level stores (lattice levels: LL.F, half spectra
// for real maps; the per-pair Hadamard levels: L^3).
// Spectrum elements per column: the largest level's (per-level launches) or, with `all_levels`
// (the persistent coupling stage), the sum over the levels.
void spectrum_sizes(const Core &c, const OrientPlan &P, const Side &src, int64_t &fin, int64_t &fout,
                    bool all_levels = false) {
    int64_t sin = 0, sout = 0;
    for (int l = 0; l < (int)P.m2l_recv.size(); ++l) {
        if (P.m2l_recv[l].count == 0) continue;
        const int64_t F = l < (int)P.lattice.size() && P.lattice[l].ok ? P.lattice[l].F : c.L3;
        const int64_t a = l < src.n_levels ? (int64_t)(src.level_offsets[l + 1] - src.level_offsets[l]) * F : 0;
        const int64_t b = (int64_t)P.m2l_recv[l].count * F;
        fin = std::max<int64_t>(fin, a);
        fout = std::max<int64_t>(fout, b);
        sin += a;
        sout += b;
    }
    if (all_levels) {
        fin = std::max(fin, sin);
        fout = std::max(fout, sout);
    }
}

int level_of(const std::vector<int32_t> &off, int b) {
    return (int)(std::upper_bound(off.begin(), off.end(), b) - off.begin()) - 1;
}

void build_lattice_level(const Core &c, OrientPlan &P, int l, const std::vector<int32_t> &lists,
                         const std::vector<int64_t> &iptr_level, const std::vector<int32_t> &isrc,
                         const std::vector<int32_t> &iop, const std::vector<int32_t> &isym, bool sym,
                         const std::vector<uint8_t> &partner);

void build_orient(Core &c, OrientPlan &P, const Side &src, const Side &dst, const int32_t *recv_of_pair,
                  const int32_t *part_of_pair, const int32_t *pair_op, const int32_t *pair_sym,
                  const int32_t *pair_fop, int64_t n_pairs) {
    P.src = &src;
    P.dst = &dst;
    // M2L CSR over receiving (dst) boxes, partners sorted for a deterministic order
    std::vector<int64_t> cnt(dst.n_boxes + 1, 0);
    for (int64_t k = 0; k < n_pairs; ++k) cnt[recv_of_pair[k] + 1]++;
    std::partial_sum(cnt.begin(), cnt.end(), cnt.begin());
    std::vector<int64_t> order(n_pairs), fill(cnt.begin(), cnt.end() - 1);
    for (int64_t k = 0; k < n_pairs; ++k) order[fill[recv_of_pair[k]]++] = k;
    for (int b = 0; b < dst.n_boxes; ++b)
        std::sort(order.begin() + cnt[b], order.begin() + cnt[b + 1],
                  [&](int64_t a, int64_t q) { return part_of_pair[a] < part_of_pair[q]; });
    std::vector<uint8_t> has_m2l(dst.n_boxes, 0), partner(src.n_boxes, 0);
    for (int64_t k = 0; k < n_pairs; ++k) {
        has_m2l[recv_of_pair[k]] = 1;
        partner[part_of_pair[k]] = 1;
    }
    std::vector<uint8_t> need(src.n_boxes, 0), init(dst.n_boxes, 0);
    for (int b = 0; b < src.n_boxes; ++b) need[b] = partner[b] || (src.parent[b] >= 0 && need[src.parent[b]]);
    for (int b = 0; b < dst.n_boxes; ++b) init[b] = has_m2l[b] || (dst.parent[b] >= 0 && init[dst.parent[b]]);

    std::vector<int32_t> lists;
    auto span_of = [&](std::vector<Span> &v, int nl, auto pred, const Side &S) {
        v.assign(nl, Span{});
        for (int l = 0; l < nl; ++l) {
            v[l].begin = (int64_t)lists.size();
            for (int b = S.level_offsets[l]; b < S.level_offsets[l + 1]; ++b)
                if (pred(b)) lists.push_back(b);
            v[l].count = (int32_t)((int64_t)lists.size() - v[l].begin);
        }
    };
    span_of(P.p2m, src.n_levels, [&](int b) { return need[b] && src.is_leaf[b]; }, src);
    span_of(P.m2m, src.n_levels, [&](int b) { return need[b] && !src.is_leaf[b]; }, src);
    span_of(P.m2l_recv, dst.n_levels, [&](int b) { return (bool)has_m2l[b]; }, dst);
    span_of(P.l2l, dst.n_levels, [&](int b) { return dst.parent[b] >= 0 && init[dst.parent[b]]; }, dst);
    span_of(P.l2p, dst.n_levels, [&](int b) { return init[b] && dst.is_leaf[b]; }, dst);
    P.l2p_avg.assign(dst.n_levels, 0);
    for (int l = 0; l < dst.n_levels; ++l) {
        int64_t pts = 0;
        for (int i = 0; i < P.l2p[l].count; ++i) {
            const int b = lists[P.l2p[l].begin + i];
            pts += dst.end[b] - dst.begin[b];
        }
        if (P.l2p[l].count) P.l2p_avg[l] = (int32_t)(pts / P.l2p[l].count);
    }
    P.lists.upload_vec(lists);

    // M2M items: children of a box are contiguous on the next level
    std::vector<std::vector<int32_t>> kids(src.n_boxes);
    for (int ch = 1; ch < src.n_boxes; ++ch)
        if (src.parent[ch] >= 0) kids[src.parent[ch]].push_back(ch);
    std::vector<int64_t> mptr;
    std::vector<int32_t> msrc, mmat;
    P.m2m_ptr_base.assign(src.n_levels, 0);
    for (int l = 0; l < src.n_levels; ++l) {
        P.m2m_ptr_base[l] = (int64_t)mptr.size();
        for (int i = 0; i < P.m2m[l].count; ++i) {
            const int b = lists[P.m2m[l].begin + i];
            mptr.push_back((int64_t)msrc.size());
            for (int ch : kids[b]) {
                msrc.push_back(ch);
                mmat.push_back((l + 1) * 8 + src.sector[ch]);
            }
        }
        mptr.push_back((int64_t)msrc.size());
    }
    P.m2m_ptr.upload_vec(mptr);
    P.m2m_src.upload_vec(msrc);
    P.m2m_mat.upload_vec(mmat);

    // L2L items (one per target)
    std::vector<int32_t> lsrc, lmat;
    std::vector<uint8_t> lacc;
    std::vector<int64_t> lptr;
    P.l2l_base.assign(dst.n_levels, 0);
    for (int l = 0; l < dst.n_levels; ++l) {
        P.l2l_base[l] = (int64_t)lptr.size();
        for (int i = 0; i < P.l2l[l].count; ++i) {
            const int b = lists[P.l2l[l].begin + i];
            lptr.push_back((int64_t)lsrc.size());
            lsrc.push_back(dst.parent[b]);
            lmat.push_back(l * 8 + dst.sector[b]);
            lacc.push_back(has_m2l[b]);
        }
        lptr.push_back((int64_t)lsrc.size());
    }
    P.l2l_src.upload_vec(lsrc);
    P.l2l_mat.upload_vec(lmat);
    P.l2l_acc.upload_vec(lacc);  // indexed by item: one item per L2L target
    P.l2l_ptr.upload_vec(lptr);
    // the same items grouped by parent (children of one parent are consecutive in level order)
    {
        std::vector<int32_t> pbox, cbox, cmat;
        std::vector<uint8_t> cacc;
        std::vector<int64_t> pptr;
        P.l2l_par.assign(dst.n_levels, Span{});
        P.l2lp_base.assign(dst.n_levels, 0);
        for (int l = 0; l < dst.n_levels; ++l) {
            P.l2l_par[l].begin = (int64_t)pbox.size();
            P.l2lp_base[l] = (int64_t)pptr.size();
            int last = -1;
            for (int i = 0; i < P.l2l[l].count; ++i) {
                const int b = lists[P.l2l[l].begin + i];
                if (dst.parent[b] != last) {
                    last = dst.parent[b];
                    pbox.push_back(last);
                    pptr.push_back((int64_t)cbox.size());
                }
                cbox.push_back(b);
                cmat.push_back(l * 8 + dst.sector[b]);
                cacc.push_back(has_m2l[b]);
            }
            pptr.push_back((int64_t)cbox.size());
            P.l2l_par[l].count = (int32_t)((int64_t)pbox.size() - P.l2l_par[l].begin);
        }
        P.l2lp_box.upload_vec(pbox);
        P.l2lp_ptr.upload_vec(pptr);
        P.l2lp_child.upload_vec(cbox);
        P.l2lp_mat.upload_vec(cmat);
        P.l2lp_acc.upload_vec(cacc);
    }

    // M2L items
    std::vector<int64_t> iptr;
    std::vector<int32_t> isrc, iop, isym, ifop;
    P.m2l_base.assign(dst.n_levels, 0);
    P.min_level = dst.n_levels;
    P.max_level = -1;
    P.max_recv_level_boxes = 0;
    P.max_src_level_boxes = 0;
    for (int l = 0; l < dst.n_levels; ++l) {
        P.m2l_base[l] = (int64_t)iptr.size();
        for (int i = 0; i < P.m2l_recv[l].count; ++i) {
            const int b = lists[P.m2l_recv[l].begin + i];
            iptr.push_back((int64_t)isrc.size());
            for (int64_t q = cnt[b]; q < cnt[b + 1]; ++q) {
                const int64_t k = order[q];
                isrc.push_back(part_of_pair[k]);
                iop.push_back(pair_op[k]);
                if (pair_sym) isym.push_back(pair_sym[k]);
                if (pair_fop) ifop.push_back(pair_fop[k]);
            }
        }
        iptr.push_back((int64_t)isrc.size());
        if (P.m2l_recv[l].count > 0) {
            P.min_level = std::min(P.min_level, l);
            P.max_level = std::max(P.max_level, l);
            P.max_recv_level_boxes = std::max<int64_t>(P.max_recv_level_boxes, P.m2l_recv[l].count);
            if (l < src.n_levels)
                P.max_src_level_boxes =
                    std::max<int64_t>(P.max_src_level_boxes, src.level_offsets[l + 1] - src.level_offsets[l]);
        }
    }
    P.m2l_ptr.upload_vec(iptr);
    P.m2l_src.upload_vec(isrc);
    P.m2l_op.upload_vec(iop);
    if (pair_sym) P.m2l_sym.upload_vec(isym);
    if (pair_fop) P.m2l_fop.upload_vec(ifop);
    if (pair_fop && c.dtype != H2G_C64) {
        // receiver pairs for the dense pair kernel (real operators only)
        std::vector<int32_t> r2, s2, f2;
        std::vector<int64_t> p2;
        P.m2l2_base.assign(dst.n_levels, 0);
        P.m2l2_count.assign(dst.n_levels, 0);
        P.m2l2_rbase.assign(dst.n_levels, 0);
        std::vector<std::pair<int32_t, int32_t>> ea, eb;
        auto items_of = [&](int l, int i, std::vector<std::pair<int32_t, int32_t>> &e) {
            e.clear();
            const int64_t q0 = iptr[P.m2l_base[l] + i], q1 = iptr[P.m2l_base[l] + i + 1];
            for (int64_t q = q0; q < q1; ++q) e.emplace_back(ifop[q], isrc[q]);
            std::sort(e.begin(), e.end());
        };
        for (int l = 0; l < dst.n_levels; ++l) {
            P.m2l2_base[l] = (int64_t)p2.size();
            P.m2l2_rbase[l] = (int64_t)r2.size() / 2;
            const int n = P.m2l_recv[l].count;
            const int32_t *rl = lists.data() + P.m2l_recv[l].begin;
            // partner of each receiver: a free lattice neighbour (x, y, z, then the opposite sides)
            // when the tree has coordinates, else the next free receiver
            std::vector<int> mate(n, -1);
            std::vector<uint8_t> used(n, 0);
            std::unordered_map<uint64_t, int> at;
            auto ckey = [](int64_t x, int64_t y, int64_t z) {
                return ((uint64_t)(x & 0x1FFFFF) << 42) | ((uint64_t)(y & 0x1FFFFF) << 21) | (uint64_t)(z & 0x1FFFFF);
            };
            const bool geo = !dst.coord.empty();
            if (geo)
                for (int i = 0; i < n; ++i) {
                    const int b = rl[i];
                    at.emplace(ckey(dst.coord[3 * (size_t)b], dst.coord[3 * (size_t)b + 1], dst.coord[3 * (size_t)b + 2]), i);
                }
            static const int nb[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
            int nxt = 0;  // the first free receiver after i (free ones only become used: monotone)
            for (int i = 0; i < n; ++i) {
                if (used[i]) continue;
                used[i] = 1;
                if (geo) {
                    const int b = rl[i];
                    for (const auto &d : nb) {
                        auto it = at.find(ckey(dst.coord[3 * (size_t)b] + d[0], dst.coord[3 * (size_t)b + 1] + d[1],
                                               dst.coord[3 * (size_t)b + 2] + d[2]));
                        if (it != at.end() && !used[it->second]) {
                            mate[i] = it->second;
                            break;
                        }
                    }
                }
                if (mate[i] < 0) {
                    nxt = std::max(nxt, i + 1);
                    while (nxt < n && used[nxt]) ++nxt;
                    if (nxt < n) mate[i] = nxt;
                }
                if (mate[i] >= 0) used[mate[i]] = 1;
                items_of(l, i, ea);
                if (mate[i] >= 0) items_of(l, mate[i], eb);
                else eb.clear();
                r2.push_back(rl[i]);
                r2.push_back(mate[i] >= 0 ? rl[mate[i]] : -1);
                p2.push_back((int64_t)f2.size());
                size_t x = 0, y = 0;
                while (x < ea.size() || y < eb.size()) {
                    if (y == eb.size() || (x < ea.size() && ea[x].first < eb[y].first)) {
                        f2.push_back(ea[x].first), s2.push_back(ea[x].second), s2.push_back(-1), ++x;
                    } else if (x == ea.size() || eb[y].first < ea[x].first) {
                        f2.push_back(eb[y].first), s2.push_back(-1), s2.push_back(eb[y].second), ++y;
                    } else {
                        f2.push_back(ea[x].first), s2.push_back(ea[x].second), s2.push_back(eb[y].second), ++x, ++y;
                    }
                }
                ++P.m2l2_count[l];
            }
            p2.push_back((int64_t)f2.size());
        }
        P.m2l2_recv.upload_vec(r2);
        P.m2l2_ptr.upload_vec(p2);
        P.m2l2_src.upload_vec(s2);
        P.m2l2_fop.upload_vec(f2);
    }
    P.lattice = std::vector<LatticeLevel>(dst.n_levels);
    if (c.m2l_kind == H2G_M2L_TOEPLITZ || c.m2l_kind == H2G_M2L_SYMTOEPLITZ)
        for (int l = 0; l < dst.n_levels; ++l) {
            if (P.m2l_recv[l].count == 0) continue;
            std::vector<int64_t> lvl(iptr.begin() + P.m2l_base[l], iptr.begin() + P.m2l_base[l] + P.m2l_recv[l].count + 1);
            build_lattice_level(c, P, l, lists, lvl, isrc, iop, isym, pair_sym != nullptr, partner);
        }

    // output mask: dofs of initialised dst leaves
    std::vector<uint8_t> mask(dst.n_dofs, 0);
    for (int b = 0; b < dst.n_boxes; ++b)
        if (init[b] && dst.is_leaf[b])
            for (int64_t j = dst.begin[b]; j < dst.end[b]; ++j) mask[j] = 1;
    P.out_mask.upload_vec(mask);
}

// Parent-block tables for the lattice M2L kernels of one level; ok = false when the level's
// couplings do not follow the parent-neighbour lattice with one operator per offset.
void build_lattice_level(const Core &c, OrientPlan &P, int l, const std::vector<int32_t> &lists,
                         const std::vector<int64_t> &iptr_level, const std::vector<int32_t> &isrc,
                         const std::vector<int32_t> &iop, const std::vector<int32_t> &isym, bool sym,
                         const std::vector<uint8_t> &partner) {
    LatticeLevel &LL = P.lattice[l];
    LL.ok = false;
    const Side &src = *P.src, &dst = *P.dst;
    if (src.coord.empty() || dst.coord.empty() || l >= src.n_levels || l < 1) return;
    const int n = P.m2l_recv[l].count;
    if (n == 0) return;
    auto cd = [&](const Side &S, int b, int a) { return S.coord[3 * (size_t)b + a]; };
    std::vector<int32_t> op_tab(343, -1), sym_tab(343, -1);
    // blocks = receiving parents
    std::map<int, int> block_of;
    std::vector<int> parents;
    for (int i = 0; i < n; ++i) {
        const int b = lists[P.m2l_recv[l].begin + i];
        const int pa = dst.parent[b];
        if (pa < 0) return;
        if (!block_of.count(pa)) {
            block_of[pa] = (int)parents.size();
            parents.push_back(pa);
        }
    }
    const int nb = (int)parents.size();
    std::vector<int32_t> tgt((size_t)nb * 8, -1);
    int64_t n_items = 0;
    for (int i = 0; i < n; ++i) {
        const int b = lists[P.m2l_recv[l].begin + i];
        tgt[(size_t)block_of[dst.parent[b]] * 8 + dst.sector[b]] = i;
        for (int64_t it = iptr_level[i]; it < iptr_level[i + 1]; ++it) {
            const int cc = isrc[it];
            int o[3], dq[3];
            for (int a = 0; a < 3; ++a) {
                o[a] = cd(src, cc, a) - cd(dst, b, a);
                dq[a] = cd(src, src.parent[cc], a) - cd(dst, dst.parent[b], a);
                if (o[a] < -3 || o[a] > 3 || dq[a] < -1 || dq[a] > 1) return;
            }
            if (std::max({std::abs(o[0]), std::abs(o[1]), std::abs(o[2])}) <= 1) return;
            const int oi = (o[0] + 3) + 7 * (o[1] + 3) + 49 * (o[2] + 3);
            const int sv = sym ? isym[it] : -1;
            if (op_tab[oi] < 0) {
                op_tab[oi] = iop[it];
                sym_tab[oi] = sv;
            } else if (op_tab[oi] != iop[it] || sym_tab[oi] != sv) {
                return;
            }
            ++n_items;
        }
    }
    // source slots by coordinate lookup
    std::map<std::tuple<int, int, int>, int> src_at;
    for (int b = src.level_offsets[l]; b < src.level_offsets[l + 1]; ++b)
        src_at[{cd(src, b, 0), cd(src, b, 1), cd(src, b, 2)}] = b - src.level_offsets[l];
    std::vector<int32_t> bsrc((size_t)216 * nb, -1);
    std::vector<int32_t> tq_op(8 * 216, -1), tq_sym(8 * 216, -1);
    for (int q = 0; q < 216; ++q) {
        const int dn = q / 8, s = q % 8;
        const int d[3] = {dn % 3 - 1, (dn / 3) % 3 - 1, dn / 9 - 1};
        for (int t = 0; t < 8; ++t) {
            int o[3];
            for (int a = 0; a < 3; ++a) o[a] = 2 * d[a] + ((s >> a) & 1) - ((t >> a) & 1);
            if (std::max({std::abs(o[0]), std::abs(o[1]), std::abs(o[2])}) <= 1) continue;
            const int oi = (o[0] + 3) + 7 * (o[1] + 3) + 49 * (o[2] + 3);
            tq_op[t * 216 + q] = op_tab[oi];
            tq_sym[t * 216 + q] = sym_tab[oi];
        }
    }
    int64_t implied = 0;
    for (int k = 0; k < nb; ++k) {
        const int pa = parents[k];
        for (int q = 0; q < 216; ++q) {
            const int dn = q / 8, s = q % 8;
            const int cx = 2 * (cd(dst, pa, 0) + dn % 3 - 1) + (s & 1);
            const int cy = 2 * (cd(dst, pa, 1) + (dn / 3) % 3 - 1) + ((s >> 1) & 1);
            const int cz = 2 * (cd(dst, pa, 2) + dn / 9 - 1) + ((s >> 2) & 1);
            auto it = src_at.find({cx, cy, cz});
            if (it == src_at.end()) continue;
            bsrc[(size_t)q * nb + k] = it->second;
            for (int t = 0; t < 8; ++t) {
                if (tgt[(size_t)k * 8 + t] < 0) continue;
                int o[3] = {2 * (dn % 3 - 1) + (s & 1) - (t & 1), 2 * ((dn / 3) % 3 - 1) + ((s >> 1) & 1) - ((t >> 1) & 1),
                            2 * (dn / 9 - 1) + ((s >> 2) & 1) - ((t >> 2) & 1)};
                if (std::max({std::abs(o[0]), std::abs(o[1]), std::abs(o[2])}) <= 1) continue;
                if (tq_op[t * 216 + q] < 0) return;  // a lattice partner with no coupling
                ++implied;
            }
        }
    }
    if (implied != n_items) return;
    // Tiles of parent blocks.  A region child is staged only when it is a source of some
    // coupling (its moment is computed); the others read as zeros.
    auto key3 = [](int64_t x, int64_t y, int64_t z) { return ((x + 1) << 42) | ((y + 1) << 21) | (z + 1); };
    std::unordered_map<int64_t, int> src_slot, block_at;
    src_slot.reserve(src_at.size() * 2);
    for (const auto &kv : src_at)
        if (partner[src.level_offsets[l] + kv.second])
            src_slot[key3(std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first))] = kv.second;
    for (int k = 0; k < nb; ++k)
        block_at[key3(cd(dst, parents[k], 0), cd(dst, parents[k], 1), cd(dst, parents[k], 2))] = k;
    auto morton = [](uint32_t x, uint32_t y, uint32_t z) {
        uint64_t m = 0;
        for (int bit = 0; bit < 21; ++bit)
            m |= ((uint64_t)((x >> bit) & 1) << (3 * bit)) | ((uint64_t)((y >> bit) & 1) << (3 * bit + 1)) |
                 ((uint64_t)((z >> bit) & 1) << (3 * bit + 2));
        return m;
    };
    constexpr int QL = 208;
    // blocks in Morton order of their parent coordinate; present sources of each block
    std::vector<std::pair<uint64_t, int>> ord(nb);
    for (int k = 0; k < nb; ++k)
        ord[k] = {morton(cd(dst, parents[k], 0), cd(dst, parents[k], 1), cd(dst, parents[k], 2)), k};
    std::sort(ord.begin(), ord.end());
    auto sources_of = [&](int k) {  // (source slot, q) of block k's present, coupled slots
        const int px = cd(dst, parents[k], 0), py = cd(dst, parents[k], 1), pz = cd(dst, parents[k], 2);
        const int32_t *bt = tgt.data() + (size_t)k * 8;
        std::vector<std::pair<int, int>> mine;
        for (int q = 0; q < 216; ++q) {
            const int dn = q / 8, sc = q % 8;
            const int64_t cx = 2 * (int64_t)(px + dn % 3 - 1) + (sc & 1);
            const int64_t cy = 2 * (int64_t)(py + (dn / 3) % 3 - 1) + ((sc >> 1) & 1);
            const int64_t cz = 2 * (int64_t)(pz + dn / 9 - 1) + ((sc >> 2) & 1);
            if (cx < 0 || cy < 0 || cz < 0) continue;
            auto it = src_slot.find(key3(cx, cy, cz));
            if (it == src_slot.end()) continue;
            bool any = false;
            for (int t = 0; t < 8; ++t) any = any || (bt[t] >= 0 && tq_op[t * 216 + q] >= 0);
            if (any) mine.push_back({it->second, q});
        }
        return mine;
    };
    // Cluster tiles: up to 16 blocks, union of present sources at most 1024.
    {
        LatticeTiles &T = LL.clus;
        constexpr int NB = 16, U = 1024;
        std::vector<int32_t> slot, npb, ttgt;
        std::vector<int16_t> qt;
        std::unordered_map<int, int> uni;
        std::vector<std::vector<std::pair<int, int>>> bsrc;
        std::vector<int> bid;
        auto close_tile = [&]() {
            if (bid.empty()) return;
            const size_t ti = npb.size();
            slot.resize((ti + 1) * U, -1);
            for (const auto &kv : uni) slot[ti * U + kv.second] = kv.first;
            ttgt.resize((ti + 1) * NB * 8, -1);
            qt.resize((ti + 1) * NB * 216, (int16_t)U);
            for (size_t j = 0; j < bid.size(); ++j) {
                for (int t = 0; t < 8; ++t) ttgt[(ti * NB + j) * 8 + t] = tgt[(size_t)bid[j] * 8 + t];
                for (const auto &e : bsrc[j]) qt[(ti * NB + j) * 216 + e.second] = (int16_t)uni[e.first];
            }
            npb.push_back((int)bid.size());
            uni.clear();
            bsrc.clear();
            bid.clear();
        };
        for (const auto &ok : ord) {
            auto mine = sources_of(ok.second);
            int add = 0;
            for (const auto &e : mine) add += uni.count(e.first) ? 0 : 1;
            if ((int)bid.size() == NB || (int)uni.size() + add > U) close_tile();
            for (const auto &e : mine)
                if (!uni.count(e.first)) {
                    const int u = (int)uni.size();
                    uni[e.first] = u;
                }
            bsrc.push_back(std::move(mine));
            bid.push_back(ok.second);
        }
        close_tile();
        T.n_tiles = (int)npb.size();
        T.slot.upload_vec(slot);
        T.npb.upload_vec(npb);
        T.tgt.upload_vec(ttgt);
        T.qt.upload_vec(qt);
    }
    // List tiles: up to LIST_NB receiving parents, consecutive in Morton order, whose present
    // sources (the children with couplings around them) number at most LIST_NSLOT together; the
    // tile's slot table is that union, and each block lists (union index, q) of its own sources.
    // On sparse levels a tile gathers scattered parents, so the region staged per block stays small.
    {
        LatticeTiles &T = LL.list;
        constexpr int NB = LIST_NB, NSLOT = LIST_NSLOT;
        std::vector<int32_t> slot, npb, ttgt, tlist, tnl;
        std::unordered_map<int, int> uni;  // source slot -> union index of the open tile
        std::vector<std::vector<std::pair<int, int>>> blk_src;  // per block of the open tile: (slot, q)
        std::vector<int> blk_id;
        auto close_tile = [&]() {
            if (blk_id.empty()) return;
            const size_t ti = npb.size();
            slot.resize((ti + 1) * NSLOT, -1);
            for (const auto &kv : uni) slot[ti * NSLOT + kv.second] = kv.first;
            ttgt.resize((ti + 1) * NB * 8, -1);
            tlist.resize((ti + 1) * NB * QL * 2, 0);
            tnl.resize((ti + 1) * NB, 0);
            for (size_t j = 0; j < blk_id.size(); ++j) {
                for (int t = 0; t < 8; ++t) ttgt[(ti * NB + j) * 8 + t] = tgt[(size_t)blk_id[j] * 8 + t];
                // pair entries across the LDS bank halves of their region row and coupling row
                std::vector<int32_t> cls[4];
                for (const auto &e : blk_src[j]) {
                    const int ri = uni[e.first], q = e.second;
                    cls[((q & 1) << 1) | (ri & 1)].push_back(ri | (q << 16));
                }
                std::vector<int32_t> o;
                for (int a = 0; a < 2; ++a) {
                    auto &x = cls[a], &y = cls[3 - a];
                    const size_t k = std::min(x.size(), y.size());
                    for (size_t i = 0; i < k; ++i) {
                        o.push_back(x[i]);
                        o.push_back(y[i]);
                    }
                    x.erase(x.begin(), x.begin() + k);
                    y.erase(y.begin(), y.begin() + k);
                }
                for (auto &x : cls) o.insert(o.end(), x.begin(), x.end());
                // whole batches of 8 slot pairs: pad with (zero slot, zero coupling row q = 104)
                while (o.size() % 16) o.push_back(NSLOT | (104 << 16));
                // byte offsets of the region row (16 complex FP32 values) and of the coupling row
                // (32 FP32 values) in the list kernel's LDS arrays
                for (size_t i = 0; i < o.size(); ++i) {
                    tlist[((ti * NB + j) * QL + i) * 2] = (o[i] & 0xffff) * 16 * 8;
                    tlist[((ti * NB + j) * QL + i) * 2 + 1] = (o[i] >> 16) * 32 * 4;
                }
                tnl[ti * NB + j] = (int)o.size();
            }
            npb.push_back((int)blk_id.size());
            uni.clear();
            blk_src.clear();
            blk_id.clear();
        };
        for (const auto &ok : ord) {
            const int k = ok.second;
            auto mine = sources_of(k);
            int add = 0;
            for (const auto &e : mine) add += uni.count(e.first) ? 0 : 1;
            if ((int)blk_id.size() == NB || (int)uni.size() + add > NSLOT) close_tile();
            for (const auto &e : mine)
                if (!uni.count(e.first)) {
                    const int u = (int)uni.size();
                    uni[e.first] = u;
                }
            blk_src.push_back(std::move(mine));
            blk_id.push_back(k);
        }
        close_tile();
        T.n_tiles = (int)npb.size();
        T.slot.upload_vec(slot);
        T.npb.upload_vec(npb);
        T.tgt.upload_vec(ttgt);
        T.list.upload_vec(tlist);
        T.nlist.upload_vec(tnl);
    }
    // Direct lists: per block (Morton order) its present, coupled slots as pairs of one even and one
    // odd q where possible (the two halves of a pair read coupling rows in opposite LDS bank
    // halves), in batches of 8 pairs padded with (slot 0, zero coupling row 104).
    {
        LatticeDirect &D = LL.direct;
        D.max_batches = 0;
        std::vector<int32_t> dptr, dlist, dtgt;
        dptr.push_back(0);
        for (const auto &ok : ord) {
            const int k = ok.second;
            const auto mine = sources_of(k);
            std::vector<std::pair<int, int>> cls[2];
            for (const auto &e : mine) cls[e.second & 1].push_back(e);
            // each class in the source boxes' Morton order: the waves of a workgroup (Morton-
            // consecutive blocks) then read the source rows they share at about the same time
            for (auto &cl : cls) std::sort(cl.begin(), cl.end());
            std::vector<std::pair<int, int>> h0, h1;
            const size_t both = std::min(cls[0].size(), cls[1].size());
            for (size_t i = 0; i < both; ++i) {
                h0.push_back(cls[0][i]);
                h1.push_back(cls[1][i]);
            }
            std::vector<std::pair<int, int>> rest;
            for (int a = 0; a < 2; ++a) rest.insert(rest.end(), cls[a].begin() + both, cls[a].end());
            for (size_t i = 0; i < rest.size(); i += 2) {
                h0.push_back(rest[i]);
                h1.push_back(i + 1 < rest.size() ? rest[i + 1] : std::make_pair(0, 104));
            }
            // at least one batch, so that every block writes its children (zeros without couplings)
            while (h0.empty() || h0.size() % 8) {
                h0.push_back({0, 104});
                h1.push_back({0, 104});
            }
            for (size_t b0 = 0; b0 < h0.size(); b0 += 8)
                for (const auto *h : {&h0, &h1})
                    for (size_t i = b0; i < b0 + 8; ++i) {
                        dlist.push_back((*h)[i].first);
                        dlist.push_back((*h)[i].second * 32 * 4);
                    }
            dptr.push_back(dptr.back() + (int32_t)(h0.size() / 8));
            D.max_batches = std::max(D.max_batches, (int)(h0.size() / 8));
            for (int t = 0; t < 8; ++t) dtgt.push_back(tgt[(size_t)k * 8 + t]);
        }
        D.n_blocks = nb;
        D.ptr.upload_vec(dptr);
        D.list.upload_vec(dlist);
        D.tgt.upload_vec(dtgt);
    }
    DevBuf d_op, d_sym;
    d_op.upload_vec(tq_op);
    if (sym) d_sym.upload_vec(tq_sym);
    build_lattice_ea(d_op.as<int32_t>(), sym ? d_sym.as<int32_t>() : nullptr, c.perms.as<int32_t>(), c.ops.ptr, c.L,
                     c.half_spectrum, c.reduced_precisions(), LL);
    LL.ok = true;
}

// Row-block tasks of one orientation: for each output leaf (CSR row t = blocks order[p[t]] ..
// order[p[t+1]-1], all with the same output points), the source columns of all its blocks in
// order, and one task per 4 tiles of `tile_rows` output rows (16 for real maps; 8 for complex
// maps, whose MFMA tiles hold the real form).  Tasks are grouped by tile count, largest first.
// Source indices must fit int32; otherwise the plan stays empty and the block kernels run.
// With `frag`, also the padded source-point tables and each task's place in the FP32 row-major
// copy (filled by build_near_frags once the values are set).
void build_near_plan(NearPlan &plan, bool trans, const std::vector<int64_t> &p, const std::vector<int32_t> &order,
                     const std::vector<int64_t> &rb, const std::vector<int64_t> &cb, const std::vector<int32_t> &nr,
                     const std::vector<int32_t> &nc, const std::vector<int64_t> &off, bool frag, int tile_rows) {
    plan.task.release();
    plan.col.release();
    plan.frag.release();
    plan.xrow.release();
    plan.n_frag = 0;
    plan.ok = false;
    std::vector<NearCol> cols;
    std::vector<int32_t> xrow;
    std::vector<NearTask> byrt[4];
    const int64_t ntgt = (int64_t)p.size() - 1;
    for (int64_t t = 0; t < ntgt; ++t) {
        const int first = order[p[t]];
        const int n_out = trans ? nc[first] : nr[first];
        const int64_t out0 = trans ? cb[first] : rb[first];
        const int64_t col0 = (int64_t)cols.size();
        for (int64_t q = p[t]; q < p[t + 1]; ++q) {
            const int bl = order[q];
            const int kdim = trans ? nr[bl] : nc[bl];
            const int64_t src0 = trans ? rb[bl] : cb[bl];
            if (src0 + kdim > INT32_MAX) return;
            for (int kc = 0; kc < kdim; ++kc) {
                NearCol e;
                e.aoff = trans ? off[bl] + kc : off[bl] + (int64_t)nr[bl] * kc;
                e.xrow = (int32_t)(src0 + kc);
                e.rstride = trans ? nr[bl] : 1;
                cols.push_back(e);
            }
        }
        const int64_t ncol = (int64_t)cols.size() - col0;
        if (ncol == 0) continue;  // nothing to add: the output rows stay zero
        if (ncol > INT32_MAX) return;
        const int64_t xoff = (int64_t)xrow.size();
        if (frag) {
            for (int64_t j = 0; j < near_ncolp(ncol); ++j) xrow.push_back(cols[col0 + std::min(j, ncol - 1)].xrow);
        }
        for (int r0 = 0; r0 < n_out; r0 += 4 * tile_rows) {
            NearTask k;
            k.out0 = out0;
            k.col0 = col0;
            k.nrow = std::min(4 * tile_rows, n_out - r0);
            k.ncol = (int32_t)ncol;
            k.row0 = r0;
            k.pad = 0;
            k.foff = 0;
            k.xoff = xoff;
            byrt[(k.nrow + tile_rows - 1) / tile_rows - 1].push_back(k);
        }
    }
    std::vector<NearTask> tasks;
    for (int i = 0; i < 4; ++i) {
        plan.cls[i] = (int64_t)tasks.size();
        tasks.insert(tasks.end(), byrt[3 - i].begin(), byrt[3 - i].end());
    }
    plan.cls[4] = (int64_t)tasks.size();
    if (frag) {
        int64_t f = 0;
        for (NearTask &k : tasks) {
            k.foff = f;
            f += (int64_t)k.nrow * near_ncolp(k.ncol);
        }
        plan.n_frag = f;
        plan.xrow.upload_vec(xrow);
    }
    plan.task.upload_vec(tasks);
    plan.col.upload_vec(cols);
    plan.ok = true;
}

void build_near(Core &c, const h2g_core_desc &d) {
    const h2g_side_desc &tr = d.galerkin ? d.test : d.trial;
    const int64_t nb = d.n_near_blocks;
    c.n_near_blocks = nb;
    c.n_near_values = d.n_near_values;
    if (d.near_values)
        c.near_vals.upload(d.near_values, (size_t)d.n_near_values * elem_size(d.dtype));
    else
        c.near_vals.alloc((size_t)d.n_near_values * elem_size(d.dtype));
    std::vector<int64_t> rb(nb), cb(nb), off(nb);
    std::vector<int32_t> nr(nb), nc(nb);
    for (int64_t q = 0; q < nb; ++q) {
        const int b = d.near_test[q], cc = d.near_trial[q];
        rb[q] = d.test.dof_begin[b];
        cb[q] = tr.dof_begin[cc];
        nr[q] = (int32_t)(d.test.dof_end[b] - d.test.dof_begin[b]);
        nc[q] = (int32_t)(tr.dof_end[cc] - tr.dof_begin[cc]);
        off[q] = d.near_offset[q];
    }
    c.near_blk_rb.upload_vec(rb);
    c.near_blk_cb.upload_vec(cb);
    c.near_blk_nr.upload_vec(nr);
    c.near_blk_nc.upload_vec(nc);
    c.near_blk_off.upload_vec(off);
    auto csr = [&](const int32_t *key, int nkeys, DevBuf &tgt, DevBuf &ptr, DevBuf &blk, int32_t &ntgt,
                   NearPlan &plan, bool trans) {
        std::vector<int64_t> cnt(nkeys + 1, 0);
        for (int64_t q = 0; q < nb; ++q) cnt[key[q] + 1]++;
        std::partial_sum(cnt.begin(), cnt.end(), cnt.begin());
        std::vector<int32_t> order(nb);
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t q = 0; q < nb; ++q) order[fill[key[q]]++] = (int32_t)q;
        std::vector<int32_t> targets;
        std::vector<int64_t> p;
        for (int k = 0; k < nkeys; ++k)
            if (cnt[k + 1] > cnt[k]) {
                targets.push_back(k);
                p.push_back(cnt[k]);
            }
        p.push_back(nb);
        ntgt = (int32_t)targets.size();
        tgt.upload_vec(targets);
        ptr.upload_vec(p);
        blk.upload_vec(order);
        // the mixed form of real maps reads the near values rounded to FP32, per task in row-major
        // order (half the HBM traffic of the near field; products and sums stay FP64)
        if (c.dtype == H2G_F64)
            build_near_plan(plan, trans, p, order, rb, cb, nr, nc, off, c.reduced_precisions() && nb > 0, 16);
        else if (c.dtype == H2G_C64)  // tiles of 8 complex rows (16 rows of the real form)
            build_near_plan(plan, trans, p, order, rb, cb, nr, nc, off, false, 8);
    };
    csr(d.near_test, d.test.n_boxes, c.near_fwd_tgt, c.near_fwd_ptr, c.near_fwd_blk, c.n_near_fwd_tgt,
        c.near_fwd_plan, false);
    csr(d.near_trial, tr.n_boxes, c.near_rev_tgt, c.near_rev_ptr, c.near_rev_blk, c.n_near_rev_tgt,
        c.near_rev_plan, true);
    if (d.near_values) build_near_frags(c);
}

// ------------------------------------------------------------------------------------------
// Far pass
// ------------------------------------------------------------------------------------------
bool is_toeplitz(const Core &c) { return c.m2l_kind == H2G_M2L_TOEPLITZ || c.m2l_kind == H2G_M2L_SYMTOEPLITZ; }

// representations of dense operators a core needs for its precisions (split BF16 only for couplings)
int dense_reps(int precisions, bool couplings) {
    int r = 0;
    if (precisions & (1 << H2G_PREC_PLAIN)) r |= DENSE_REP_F64;
    if (precisions & ~(1 << H2G_PREC_PLAIN)) r |= DENSE_REP_F32;
    if (couplings && (precisions & ((1 << H2G_PREC_BF16X2) | (1 << H2G_PREC_BF16X3)))) r |= DENSE_REP_BF16;
    return r;
}

// representation of a workspace precision for the dense kernel
int dense_rep(int precision, bool couplings) {
    if (precision == H2G_PREC_PLAIN) return DENSE_REP_F64;
    if (couplings && precision == H2G_PREC_BF16X2) return DENSE_REP_BF16X2;
    if (couplings && precision == H2G_PREC_BF16X3) return DENSE_REP_BF16X3;
    return DENSE_REP_F32;
}

// join: an event the first write of the result (L2P) waits for (the near field on the side
// stream), or null
// Whether the coupling stage runs as one persistent launch: a workspace asking for it, Toeplitz
// couplings, and every level with coupling pairs a lattice level on the direct kernel's path.
bool persistent_stage(const Workspace &w, const OrientPlan &P, const Side &src, const Side &dst, int cd, int R) {
    if (!w.opts.persistent || !w.counters.ptr || !is_toeplitz(*w.core) || (cd != H2G_F32 && cd != H2G_C32)) return false;
    bool any = false;
    for (int l = 0; l < dst.n_levels; ++l) {
        if (P.m2l_recv[l].count == 0) continue;
        if (l >= src.n_levels || l >= (int)P.lattice.size() || !P.lattice[l].ok) return false;
        const int nsrc = src.level_offsets[l + 1] - src.level_offsets[l];
        if (!m2l_direct_rows(cd, P.lattice[l], nsrc, R)) return false;
        any = true;
    }
    return any;
}

void far_pass(Workspace &w, int op, int R, hipEvent_t join) {
    Core &c = *w.core;
    const OrientPlan &P = op == H2G_OP_N ? c.fwd : c.rev;
    const Side &src = *P.src, &dst = *P.dst;
    const bool mixed = w.opts.precision != H2G_PREC_PLAIN;
    const int cd = mixed ? (c.dtype == H2G_F64 ? H2G_F32 : H2G_C32) : c.dtype;
    const int rd = mixed ? H2G_F32 : H2G_F64;  // the real type of the compute type
    const bool conj = op == H2G_OP_C;
    const bool cmap = c.dtype == H2G_C64;
    const int32_t *L = P.lists.as<int32_t>();
    hipStream_t s = w.stream;
    const void *scoll = mixed ? src.coll32.ptr : src.coll.ptr;
    const void *dcoll = mixed ? dst.coll32.ptr : dst.coll.ptr;
    const void *str = mixed ? src.transfers32.ptr : src.transfers.ptr;
    const void *dtr = mixed ? dst.transfers32.ptr : dst.transfers.ptr;
    const void *ops = mixed ? c.ops32.ptr : c.ops.ptr;
    const int K = c.K;

    // upward pass
    for (int l = src.n_levels - 1; l >= 0; --l) {
        if (P.p2m[l].count) {
            w.mark_begin(2, l);
            if (src.sep_cplx && src.coll_real) {
                // real collection: the real leaf GEMM over x as 2 R real columns writes planar rows
                launch_p2m(rd, false, L + P.p2m[l].begin, P.p2m[l].count, src.d_begin.as<int64_t>(),
                           src.d_end.as<int64_t>(), mixed ? src.coll32_re.ptr : src.coll_re.ptr, K, w.xs.ptr,
                           2 * R, w.mom.ptr, s);
            } else {
                launch_p2m(cd, conj, L + P.p2m[l].begin, P.p2m[l].count, src.d_begin.as<int64_t>(),
                           src.d_end.as<int64_t>(), scoll, K, w.xs.ptr, R, w.mom.ptr, s);
                // real transfers of a complex map: the leaves' moments become planar rows for M2M
                if (src.sep_cplx)
                    launch_cplx_planar(cd, L + P.p2m[l].begin, P.p2m[l].count, w.mom.ptr, R, K, true, s);
            }
            w.mark_end();
        }
        if (P.m2m[l].count) {
            w.mark_begin(3, l);
            if (src.sep_cplx) {
                // real factors on planar rows (2 R real columns); conjugation leaves them unchanged
                const void *fac = mixed ? src.factors32.ptr : src.factors.ptr;
                if (src.sep_shared)
                    launch_sep_m2m(rd, 0, P.m2m[l].count, L + P.m2m[l].begin, P.m2m_ptr.as<int64_t>() + P.m2m_ptr_base[l],
                                   P.m2m_src.as<int32_t>(), P.m2m_mat.as<int32_t>(), fac, (l + 1) * 8, src.sep_p,
                                   w.mom.ptr, 2 * R, s);
                else
                    launch_sep_transfer(rd, 0, P.m2m[l].count, L + P.m2m[l].begin,
                                        P.m2m_ptr.as<int64_t>() + P.m2m_ptr_base[l], P.m2m_src.as<int32_t>(),
                                        P.m2m_mat.as<int32_t>(), fac, (l + 1) * 8, src.sep_p, w.mom.ptr, w.mom.ptr,
                                        2 * R, 0, nullptr, s);
            } else if (src.separable && src.sep_shared)
                launch_sep_m2m(cd, conj ? 3 : 0, P.m2m[l].count, L + P.m2m[l].begin,
                               P.m2m_ptr.as<int64_t>() + P.m2m_ptr_base[l], P.m2m_src.as<int32_t>(),
                               P.m2m_mat.as<int32_t>(), mixed ? src.factors32.ptr : src.factors.ptr, (l + 1) * 8,
                               src.sep_p, w.mom.ptr, R, s);
            else if (src.separable)
                launch_sep_transfer(cd, conj ? 3 : 0, P.m2m[l].count, L + P.m2m[l].begin,
                                    P.m2m_ptr.as<int64_t>() + P.m2m_ptr_base[l], P.m2m_src.as<int32_t>(),
                                    P.m2m_mat.as<int32_t>(), mixed ? src.factors32.ptr : src.factors.ptr, (l + 1) * 8,
                                    src.sep_p,
                                    w.mom.ptr, w.mom.ptr, R, 0, nullptr, s);
            else if (src.tr_dense.ok)
                launch_dense(src.tr_dense, conj && cmap ? 3 : 0, dense_rep(w.opts.precision, false), L + P.m2m[l].begin,
                             P.m2m[l].count, P.m2m_ptr.as<int64_t>() + P.m2m_ptr_base[l], P.m2m_src.as<int32_t>(),
                             P.m2m_mat.as<int32_t>(), w.mom.ptr, w.mom.ptr, R, nullptr, s);
            else
                launch_boxgemm(cd, conj ? 3 : 0, P.m2m[l].count, L + P.m2m[l].begin,
                               P.m2m_ptr.as<int64_t>() + P.m2m_ptr_base[l], P.m2m_src.as<int32_t>(),
                               P.m2m_mat.as<int32_t>(), nullptr, nullptr, nullptr, str, K, w.mom.ptr, w.mom.ptr, R, 0,
                               nullptr, s);
            w.mark_end();
        }
    }
    // the couplings read interleaved moments
    if (src.sep_cplx) {
        w.mark_begin(3, -1);
        for (int l = 0; l < src.n_levels; ++l) {
            launch_cplx_planar(cd, L + P.p2m[l].begin, P.p2m[l].count, w.mom.ptr, R, K, false, s);
            launch_cplx_planar(cd, L + P.m2m[l].begin, P.m2m[l].count, w.mom.ptr, R, K, false, s);
        }
        w.mark_end();
    }
    // couplings
    const int opA = op == H2G_OP_N ? 0 : (op == H2G_OP_T ? 1 : 2);
    const bool sym = c.m2l_kind == H2G_M2L_SYMDENSE || c.m2l_kind == H2G_M2L_SYMTOEPLITZ;
    const bool persist = persistent_stage(w, P, src, dst, cd, R);
    if (persist) {
        // every coupling level is a lattice level on the direct path: modulate-in of all levels
        // into their own parts of the spectrum buffers, one persistent coupling launch over the
        // levels (largest first), modulate-out of all levels
        const bool inv_in = op == H2G_OP_T;
        std::vector<DirectLaunchLevel> lv;
        std::vector<std::pair<int64_t, int>> order;
        int64_t oin = 0, oout = 0;
        for (int l = 0; l < dst.n_levels; ++l) {
            const int n = P.m2l_recv[l].count;
            if (n == 0) continue;
            const int box0 = src.level_offsets[l];
            const int nsrc = src.level_offsets[l + 1] - box0;
            const LatticeLevel &LL = P.lattice[l];
            const int drows = m2l_direct_rows(cd, LL, nsrc, R);
            cplx<float> *S = w.spec_in.as<cplx<float>>() + oin;
            cplx<float> *O = w.spec_out.as<cplx<float>>() + oout;
            w.mark_begin(4, l);
            launch_fft_in_fm(cd, c.half_spectrum, inv_in, box0, nsrc, w.mom.ptr, K, c.p, c.L, R, drows, drows > 16,
                             S, c.twiddle.ptr, s);
            w.mark_end();
            order.push_back({(int64_t)LL.direct.n_blocks * LL.F, (int)lv.size()});
            lv.push_back(DirectLaunchLevel{&LL, nsrc, n, S, O});
            oin += (int64_t)LL.F * nsrc * drows;
            oout += (int64_t)LL.F * n * R;
        }
        std::sort(order.begin(), order.end(), [](auto &a, auto &b) { return a.first > b.first; });
        std::vector<DirectLaunchLevel> sorted;
        for (auto &o : order) sorted.push_back(lv[o.second]);
        w.mark_begin(5, -1);
        launch_m2l_direct_persistent(conj, sorted, R, w.opts.m2l_groups, w.counters.as<int>(), s);
        w.mark_end();
        int li = 0;
        for (int l = 0; l < dst.n_levels; ++l) {
            const int n = P.m2l_recv[l].count;
            if (n == 0) continue;
            w.mark_begin(6, l);
            launch_fft_out_fm(cd, c.half_spectrum, !inv_in, L + P.m2l_recv[l].begin, n, lv[li].O, K, c.p, c.L, R,
                              w.loc.ptr, c.twiddle.ptr, s);
            w.mark_end();
            ++li;
        }
    }
    for (int l = 0; l < dst.n_levels && !persist; ++l) {
        const int n = P.m2l_recv[l].count;
        if (n == 0) continue;
        const int32_t *recv = L + P.m2l_recv[l].begin;
        const int64_t *iptr = P.m2l_ptr.as<int64_t>() + P.m2l_base[l];
        if (c.m2l_kind == H2G_M2L_DIAGONAL) {
            w.mark_begin(5, l);
            launch_diag_m2l(cd, conj, n, recv, iptr, P.m2l_src.as<int32_t>(), P.m2l_op.as<int32_t>(), ops, K, w.mom.ptr,
                            w.loc.ptr, R, s);
            w.mark_end();
        } else if (!is_toeplitz(c)) {
            w.mark_begin(5, l);
            const int orient = op == H2G_OP_N ? 0 : (op == H2G_OP_T || !cmap ? 1 : 2);
            const int rep = dense_rep(w.opts.precision, true);
            if (c.m2l_dense.ok && P.m2l2_ptr.ptr &&
                launch_dense_pair(c.m2l_dense, orient, rep, w.opts.dense_pairs, P.m2l2_recv.as<int32_t>() + 2 * P.m2l2_rbase[l],
                                  P.m2l2_count[l], P.m2l2_ptr.as<int64_t>() + P.m2l2_base[l],
                                  P.m2l2_src.as<int32_t>(), P.m2l2_fop.as<int32_t>(), w.mom.ptr, w.loc.ptr, R, s)) {
            } else if (c.m2l_dense.ok)
                launch_dense(c.m2l_dense, orient, rep, recv, n, iptr, P.m2l_src.as<int32_t>(),
                             P.m2l_fop.as<int32_t>(), w.mom.ptr, w.loc.ptr, R, nullptr, s);
            else
                launch_boxgemm(cd, opA, n, recv, iptr, P.m2l_src.as<int32_t>(), P.m2l_op.as<int32_t>(),
                               sym ? P.m2l_sym.as<int32_t>() : nullptr, c.perms.as<int32_t>(),
                               c.perms_inv.as<int32_t>(), ops, K, w.mom.ptr, w.loc.ptr, R, 0, nullptr, s);
            w.mark_end();
        } else {
            // forward / adjoint: fft in, ifft out; transpose swaps the two plans
            const bool inv_in = op == H2G_OP_T;
            if (l >= src.n_levels) continue;
            const int box0 = src.level_offsets[l];
            const int nsrc = src.level_offsets[l + 1] - box0;
            const LatticeLevel &LL = P.lattice[l];
            if (LL.ok) {
                const bool half = c.half_spectrum;
                w.mark_begin(4, l);
                // the direct kernel's row layout: 16 columns, or a multiple of 32 permuted
                const int drows = m2l_direct_rows(cd, LL, nsrc, R);
                launch_fft_in_fm(cd, half, inv_in, box0, nsrc, w.mom.ptr, K, c.p, c.L, R, drows ? drows : R, drows > 16,
                                 w.spec_in.ptr,
                                 c.twiddle.ptr, s);
                w.mark_end();
                w.mark_begin(5, l);
                launch_m2l_lattice(cd, conj, LL, nsrc, n, w.spec_in.ptr, w.spec_out.ptr, R, w.opts.m2l_groups, s);
                w.mark_end();
                w.mark_begin(6, l);
                launch_fft_out_fm(cd, half, !inv_in, recv, n, w.spec_out.ptr, K, c.p, c.L, R, w.loc.ptr,
                                  c.twiddle.ptr, s);
                w.mark_end();
                continue;
            }
            w.mark_begin(4, l);
            launch_fft_in(cd, inv_in, nullptr, nsrc, box0, w.mom.ptr, K, c.p, c.L, R, w.spec_in.ptr, c.twiddle.ptr, s);
            w.mark_end();
            w.mark_begin(5, l);
            launch_hadamard(cd, conj, n, iptr, P.m2l_src.as<int32_t>(), P.m2l_op.as<int32_t>(),
                            sym ? P.m2l_sym.as<int32_t>() : nullptr, c.perms.as<int32_t>(), ops, c.L3, box0,
                            w.spec_in.ptr, w.spec_out.ptr, R, s);
            w.mark_end();
            w.mark_begin(6, l);
            launch_fft_out(cd, !inv_in, recv, n, w.spec_out.ptr, K, c.p, c.L, R, w.loc.ptr, c.twiddle.ptr, s);
            w.mark_end();
        }
    }
    // downward pass
    const cplx<double> a = conj ? conjv(c.alpha) : c.alpha;
    // real transfers of a complex map: the couplings' results become planar rows for L2L
    if (dst.sep_cplx) {
        w.mark_begin(7, -1);
        for (int l = 0; l < dst.n_levels; ++l)
            launch_cplx_planar(cd, L + P.m2l_recv[l].begin, P.m2l_recv[l].count, w.loc.ptr, R, K, true, s);
        w.mark_end();
    }
    for (int l = 0; l < dst.n_levels; ++l) {
        if (P.l2l[l].count) {
            w.mark_begin(7, l);
            if (dst.sep_cplx)
                launch_sep_l2l(rd, 1, P.l2l_par[l].count, P.l2lp_box.as<int32_t>() + P.l2l_par[l].begin,
                               P.l2lp_ptr.as<int64_t>() + P.l2lp_base[l], P.l2lp_child.as<int32_t>(),
                               P.l2lp_mat.as<int32_t>(), P.l2lp_acc.as<uint8_t>(),
                               mixed ? dst.factors32.ptr : dst.factors.ptr, l * 8, dst.sep_p, dst.sep_shared, w.loc.ptr,
                               2 * R, s);
            else if (dst.separable)
                launch_sep_l2l(cd, conj ? 2 : 1, P.l2l_par[l].count, P.l2lp_box.as<int32_t>() + P.l2l_par[l].begin,
                               P.l2lp_ptr.as<int64_t>() + P.l2lp_base[l], P.l2lp_child.as<int32_t>(),
                               P.l2lp_mat.as<int32_t>(), P.l2lp_acc.as<uint8_t>(),
                               mixed ? dst.factors32.ptr : dst.factors.ptr, l * 8, dst.sep_p, dst.sep_shared, w.loc.ptr, R, s);
            else if (dst.tr_dense.ok)
                launch_dense(dst.tr_dense, conj && cmap ? 2 : 1, dense_rep(w.opts.precision, false), L + P.l2l[l].begin,
                             P.l2l[l].count, P.l2l_ptr.as<int64_t>() + P.l2l_base[l], P.l2l_src.as<int32_t>(),
                             P.l2l_mat.as<int32_t>(), w.loc.ptr, w.loc.ptr, R, P.l2l_acc.as<uint8_t>(), s);
            else
                launch_boxgemm(cd, conj ? 2 : 1, P.l2l[l].count, L + P.l2l[l].begin,
                               P.l2l_ptr.as<int64_t>() + P.l2l_base[l], P.l2l_src.as<int32_t>(),
                               P.l2l_mat.as<int32_t>(), nullptr, nullptr, nullptr, dtr, K, w.loc.ptr, w.loc.ptr, R, 2,
                               P.l2l_acc.as<uint8_t>(), s);
            w.mark_end();
        }
        if (P.l2p[l].count) {
            if (join) {
                H2G_HIP(hipStreamWaitEvent(s, join, 0));
                join = nullptr;
            }
            w.mark_begin(8, l);
            if (dst.sep_cplx && dst.coll_real) {
                // real collection (and alpha): the real leaf GEMM from planar rows into y as 2 R real columns
                launch_l2p(rd, false, L + P.l2p[l].begin, P.l2p[l].count, dst.d_begin.as<int64_t>(),
                           dst.d_end.as<int64_t>(), mixed ? dst.coll32_re.ptr : dst.coll_re.ptr, K, w.loc.ptr,
                           2 * R, w.ys.ptr, a, P.l2p_avg[l], s);
            } else {
                if (dst.sep_cplx)
                    launch_cplx_planar(cd, L + P.l2p[l].begin, P.l2p[l].count, w.loc.ptr, R, K, false, s);
                launch_l2p(cd, conj, L + P.l2p[l].begin, P.l2p[l].count, dst.d_begin.as<int64_t>(),
                           dst.d_end.as<int64_t>(), dcoll, K, w.loc.ptr, R, w.ys.ptr, a, P.l2p_avg[l], s);
            }
            w.mark_end();
        }
    }
    if (join) H2G_HIP(hipStreamWaitEvent(s, join, 0));
}

// Device buffers of one workspace of `R` columns (bytes): tree-ordered x and y in the core type,
// the moment pools in the compute type and, for Toeplitz couplings, the spectra.
struct WsLayout {
    size_t xs = 0, ys = 0, mom = 0, loc = 0, spec_in = 0, spec_out = 0, x32 = 0;
    size_t total(size_t align) const {
        size_t s = 0;
        for (size_t b : {xs, ys, mom, loc, spec_in, spec_out, x32}) s += (b + align - 1) / align * align;
        return s;
    }
};

WsLayout ws_layout(const Core &c, bool mixed, int R, int near_cols, bool persistent) {
    WsLayout L;
    const int64_t nmax = std::max(c.test.n_dofs, c.trial->n_dofs);
    const int64_t bmax = std::max(c.test.n_boxes, c.trial->n_boxes);
    const size_t se = elem_size(c.dtype);
    const size_t ce = mixed ? se / 2 : se;
    L.xs = L.ys = (size_t)nmax * R * se;
    L.mom = L.loc = (size_t)bmax * c.K * R * ce;
    if (mixed && (c.near_fwd_plan.n_frag > 0 || c.near_rev_plan.n_frag > 0))
        L.x32 = (size_t)nmax * near_x32_pitch(R, near_cols) * sizeof(float);
    if (is_toeplitz(c)) {
        const size_t cc = mixed ? 8 : 16;
        // per column, the largest level's (boxes x stored frequencies) of either orientation
        int64_t fin = 0, fout = 0;
        const bool all = persistent && mixed && R >= 4;
        spectrum_sizes(c, c.fwd, *c.trial, fin, fout, all);
        spectrum_sizes(c, c.rev, c.test, fin, fout, all);
        // source spectra of FP32 batches of >= 4 columns may be stored with rows padded to 32
        // columns (the direct M2L kernel's layout)
        const int64_t rin = mixed && R >= 4 ? (R + 31) / 32 * 32 : R;
        L.spec_in = (size_t)fin * rin * cc;
        L.spec_out = (size_t)fout * R * cc;
    }
    return L;
}

// y = a op(A) x + b y (five) or y = op(A) x over column blocks of the workspace's capacity, all on
// the workspace stream.  x and y are device pointers.  overlap_near: the near field runs on the
// workspace's side stream, concurrent with the far field up to its first write of the result.
void run_passes(Workspace &w, int op, int which, int dtype, int64_t nrhs, const void *dx, int64_t lx, void *dy,
                int64_t ly, cplx<double> a, cplx<double> b, bool five, bool overlap_near, bool use_graph) {
    Core &c = *w.core;
    hipStream_t s = w.stream;
    const size_t xe = elem_size(dtype), se = elem_size(c.dtype);
    const OrientPlan &P = op == H2G_OP_N ? c.fwd : c.rev;
    const Side &src = *P.src, &dst = *P.dst;
    const bool mixed_ws = w.opts.precision != H2G_PREC_PLAIN;
    float *x32 = mixed_ws && w.x32.ptr ? w.x32.as<float>() : nullptr;
    for (int64_t c0 = 0; c0 < nrhs; c0 += w.R_cap) {
        const int R = (int)std::min<int64_t>(w.R_cap, nrhs - c0);
        w.mark_begin(0, -1);
        launch_gather(c.dtype, dtype, (const char *)dx + (size_t)c0 * lx * xe, lx, src.d_perm.as<int64_t>(),
                      src.d_iperm.as<int64_t>(), src.n_dofs, R, w.xs.ptr, s);
        w.mark_end();
        auto body = [&]() {
            launch_fill_zero(w.ys.ptr, (size_t)dst.n_dofs * R * se, s);
            if ((which & H2G_NEARMUL) && (which & H2G_FARMUL) && overlap_near && !w.opts.timing) {
                // the near field (which sets its rows of the result) on the side stream,
                // concurrent with the far field up to its first write of the result
                H2G_HIP(hipEventRecord(w.ev_fork, s));
                H2G_HIP(hipStreamWaitEvent(w.side, w.ev_fork, 0));
                launch_near(c.dtype, op, c, w.xs.ptr, w.ys.ptr, R, x32, w.opts.near_cols, w.side);
                H2G_HIP(hipEventRecord(w.ev_near, w.side));
                far_pass(w, op, R, w.ev_near);
                return;
            }
            if (which & H2G_NEARMUL) {
                w.mark_begin(1, -1);
                launch_near(c.dtype, op, c, w.xs.ptr, w.ys.ptr, R, x32, w.opts.near_cols, s);
                w.mark_end();
            }
            if (which & H2G_FARMUL) far_pass(w, op, R, nullptr);
        };
        if (use_graph) {
            GraphKey key{op, which, R};
            auto it = w.graphs.find(key);
            if (it == w.graphs.end()) {
                hipGraph_t g;
                H2G_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    body();
                } catch (...) {
                    (void)hipStreamEndCapture(s, &g);
                    throw;
                }
                H2G_HIP(hipStreamEndCapture(s, &g));
                hipGraphExec_t ge;
                H2G_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                it = w.graphs.emplace(key, ge).first;
            }
            H2G_HIP(hipGraphLaunch(it->second, s));
        } else {
            body();
        }
        const uint8_t *mask = (which == H2G_FARMUL && five) ? P.out_mask.as<uint8_t>() : nullptr;
        w.mark_begin(9, -1);
        launch_scatter(c.dtype, dtype, w.ys.ptr, dst.d_perm.as<int64_t>(), dst.d_iperm.as<int64_t>(), dst.n_dofs, R,
                       (char *)dy + (size_t)c0 * ly * xe, ly, a, five ? b : cplx<double>{0.0, 0.0}, mask, s);
        w.mark_end();
    }
}

// Argument checks shared by the product entry points; returns the scalars (1, 0 without them).
void check_call(const Core &c, int op, int which, int dtype, int64_t nrhs, const void *x, int64_t ldx, const void *y,
                int64_t ldy, const void *alpha, const void *beta, cplx<double> &a, cplx<double> &b) {
    if (op < 0 || op > 2) fail(H2G_EINVAL, "op must be H2G_OP_N, H2G_OP_T or H2G_OP_C");
    if (which < 1 || which > 3) fail(H2G_EINVAL, "which must be H2G_MUL, H2G_FARMUL or H2G_NEARMUL");
    if (dtype < 0 || dtype > 3) fail(H2G_EINVAL, "unknown vector dtype");
    if (c.dtype == H2G_F64 && is_complex_dtype(dtype)) fail(H2G_EINVAL, "a real map takes real vectors");
    if ((alpha == nullptr) != (beta == nullptr)) fail(H2G_EINVAL, "give both alpha and beta or neither");
    if (nrhs < 0) fail(H2G_EINVAL, "nrhs must be >= 0");
    const int64_t n_in = op == H2G_OP_N ? c.trial->n_dofs : c.test.n_dofs;
    const int64_t n_out = op == H2G_OP_N ? c.test.n_dofs : c.trial->n_dofs;
    if (nrhs > 0 && (ldx < n_in || ldy < n_out)) fail(H2G_EDIM, "leading dimension smaller than the vector length");
    if (nrhs > 0 && (!x || !y)) fail(H2G_EINVAL, "null x or y");
    a = {1.0, 0.0};
    b = {0.0, 0.0};
    if (alpha) {
        // scalars are host values of the vectors' element type
        auto rd = [&](const void *p) -> cplx<double> {
            switch (dtype) {
            case H2G_F64: return {*(const double *)p, 0.0};
            case H2G_F32: return {(double)*(const float *)p, 0.0};
            case H2G_C64: return {((const double *)p)[0], ((const double *)p)[1]};
            default: return {(double)((const float *)p)[0], (double)((const float *)p)[1]};
            }
        };
        a = rd(alpha);
        b = rd(beta);
    }
}

}  // namespace

}  // namespace h2g

// Scratch of h2g_matvec_noalloc: the workspace buffers (x, y, moments, spectra) of its products
// are views into this array.  Each GPU holds its own copy, at its own address, until the process
// ends; host code reaches it only through the address hipGetSymbolAddress returns per device.
static __device__ __align__(256) unsigned char scratch[size_t(128) << 20];

namespace h2g {
namespace {
constexpr size_t NOALLOC_BYTES = size_t(128) << 20, NOALLOC_ALIGN = 256;
constexpr int NOALLOC_MAX_DEVICES = 64;
std::array<std::atomic<bool>, NOALLOC_MAX_DEVICES> g_noalloc_busy{};

// Device address of `scratch` on the current device `device`, looked up once per device.
void *noalloc_scratch(int device) {
    static std::mutex mu;
    static std::array<void *, NOALLOC_MAX_DEVICES> ptrs{};
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[device]) {
        void *p = nullptr;
        H2G_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(H2G_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        ptrs[device] = p;
    }
    return ptrs[device];
}
}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }

void Workspace::mark_begin(int phase, int level) {
    if (!opts.timing) return;
    if (n_events_used == events.size()) {
        hipEvent_t a, b;
        H2G_HIP(hipEventCreate(&a));
        H2G_HIP(hipEventCreate(&b));
        events.emplace_back(a, b);
    }
    H2G_HIP(hipEventRecord(events[n_events_used].first, stream));
    timings.push_back(h2g_timing_record{phase, level, 0.0f, cur_op, -1});
}

void Workspace::mark_end() {
    if (!opts.timing) return;
    H2G_HIP(hipEventRecord(events[n_events_used].second, stream));
    ++n_events_used;
}

int64_t Core::device_bytes() const {
    int64_t b = 0;
    auto dense = [&](const DenseOps &D) {
        for (int o = 0; o < 4; ++o) b += D.f64[o].bytes + D.f32[o].bytes + D.bf[o].bytes;
    };
    auto side = [&](const Side &S) {
        b += S.d_perm.bytes + S.d_iperm.bytes + S.d_begin.bytes + S.d_end.bytes + S.coll.bytes + S.coll32.bytes +
             S.coll_re.bytes + S.coll32_re.bytes + S.transfers.bytes +
             S.transfers32.bytes + S.factors.bytes + S.factors32.bytes;
        dense(S.tr_dense);
    };
    dense(m2l_dense);
    side(test);
    if (!galerkin) side(trial_own);
    b += ops.bytes + ops32.bytes + perms.bytes + perms_inv.bytes + near_vals.bytes;
    b += near_fwd_plan.frag.bytes + near_fwd_plan.xrow.bytes + near_rev_plan.frag.bytes + near_rev_plan.xrow.bytes;
    b += near_fwd_plan.task.bytes + near_fwd_plan.col.bytes + near_rev_plan.task.bytes + near_rev_plan.col.bytes;
    return b;
}

int64_t Workspace::device_bytes() const {
    return xs.bytes + ys.bytes + mom.bytes + loc.bytes + spec_in.bytes + spec_out.bytes + x32.bytes + xhost.bytes +
           yhost.bytes;
}

}  // namespace h2g

using namespace h2g;

extern "C" {

int h2g_version(void) { return VERSION; }

int h2g_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return H2G_EINVAL;
    std::snprintf(buf, len, "%s", g_last_error.c_str());
    return H2G_OK;
}

int h2g_core_create(const h2g_core_desc *desc, h2g_core **out) {
    return guarded([&] {
        if (!desc || !out) fail(H2G_EINVAL, "null argument");
        *out = nullptr;
        validate_desc(*desc);
        const h2g_core_desc &d = *desc;
        int ndev = 0;
        H2G_HIP(hipGetDeviceCount(&ndev));
        if (d.device < 0 || d.device >= ndev) fail(H2G_EINVAL, "device ordinal out of range");
        H2G_HIP(hipSetDevice(d.device));
        auto *c = new h2g_core();
        try {
            c->device = d.device;
            c->dtype = d.dtype;
            c->K = d.K;
            c->p = d.p;
            c->L = d.p > 0 ? 2 * d.p - 1 : 0;
            c->L3 = c->L * c->L * c->L;
            c->galerkin = d.galerkin != 0;
            c->m2l_kind = d.m2l_kind;
            c->n_pairs = d.n_pairs;
            c->n_ops = d.n_ops;
            c->n_perms = d.n_perms;
            c->alpha = {d.alpha_re, d.alpha_im};
            c->precisions = d.precisions ? d.precisions : (1 << H2G_PREC_PLAIN);
            const bool want32 = c->reduced_precisions();
            copy_side(c->test, d.test, d.dtype, d.K, want32);
            if (c->galerkin) {
                c->trial = &c->test;
            } else {
                copy_side(c->trial_own, d.trial, d.dtype, d.K, want32);
                c->trial = &c->trial_own;
            }
            // given collections that are real (a collection to be filled by h2g_method_fill is
            // checked after the fill)
            if (d.test.collection) make_real_collection(c->test, *c);
            if (!c->galerkin && d.trial.collection) make_real_collection(c->trial_own, *c);
            const bool toep = is_toeplitz(*c);
            const bool diag = d.m2l_kind == H2G_M2L_DIAGONAL;
            const size_t opelems = toep ? (size_t)d.n_ops * c->L3 : (size_t)d.n_ops * d.K * (diag ? 1 : d.K);
            const int opdtype = toep ? H2G_C64 : d.dtype;
            c->ops.upload(d.ops, opelems * elem_size(opdtype));
            if (want32) {
                const size_t nreal = opelems * (opdtype == H2G_C64 ? 2 : 1);
                std::vector<float> v(nreal);
                const double *s = (const double *)d.ops;
                for (size_t i = 0; i < nreal; ++i) v[i] = (float)s[i];
                c->ops32.upload_vec(v);
            }
            if (toep && d.dtype == H2G_F64) {
                // real map: real-to-complex modulation when every spectrum is Hermitian
                const cplx<double> *D = (const cplx<double> *)d.ops;
                const int L = c->L;
                double mx = 0.0, dev = 0.0;
                for (int64_t o = 0; o < d.n_ops; ++o)
                    for (int z = 0; z < L; ++z)
                        for (int y = 0; y < L; ++y)
                            for (int x = 0; x < L; ++x) {
                                const cplx<double> a = D[o * c->L3 + x + L * (y + L * z)];
                                const cplx<double> b =
                                    D[o * c->L3 + (L - x) % L + L * ((L - y) % L + L * ((L - z) % L))];
                                mx = std::max(mx, std::hypot(a.re, a.im));
                                dev = std::max(dev, std::hypot(a.re - b.re, a.im + b.im));
                            }
                c->half_spectrum = dev <= 1e-12 * mx;
            }
            if (d.n_perms > 0) {
                const int64_t plen = d.m2l_kind == H2G_M2L_SYMDENSE ? d.K : c->L3;
                std::vector<int32_t> pv(d.perms, d.perms + (size_t)d.n_perms * plen), pinv(pv.size());
                for (int q = 0; q < d.n_perms; ++q)
                    for (int64_t i = 0; i < plen; ++i) pinv[q * plen + pv[q * plen + i]] = (int32_t)i;
                c->perms.upload_vec(pv);
                c->perms_inv.upload_vec(pinv);
            }
            if (c->L > 0) {
                std::vector<cplx<double>> tw(c->L);
                for (int i = 0; i < c->L; ++i) {
                    const double ang = -2.0 * M_PI * (double)i / (double)c->L;
                    tw[i] = {std::cos(ang), std::sin(ang)};
                }
                c->twiddle.upload_vec(tw);
            }
            build_near(*c, d);
            {
                std::set<std::pair<int32_t, int32_t>> uniq;
                for (int64_t k = 0; k < d.n_pairs; ++k) {
                    uniq.emplace(d.pair_op[k], d.pair_sym ? d.pair_sym[k] : -1);
                    if (d.galerkin && d.pair_test[k] == d.pair_trial[k]) ++c->n_self_pairs;
                }
                c->n_unique_translations = (int64_t)uniq.size();
            }
            // dense kinds: the distinct (translation id, symmetry) operators of the pairs
            std::vector<int32_t> pair_fop, fop_op, fop_sym;
            if (!toep && !diag && d.n_pairs > 0) {
                std::map<std::pair<int32_t, int32_t>, int32_t> fops;
                pair_fop.resize(d.n_pairs);
                for (int64_t k = 0; k < d.n_pairs; ++k) {
                    const auto key = std::make_pair(d.pair_op[k], d.pair_sym ? d.pair_sym[k] : -1);
                    auto it = fops.find(key);
                    if (it == fops.end()) {
                        it = fops.emplace(key, (int32_t)fop_op.size()).first;
                        fop_op.push_back(key.first);
                        fop_sym.push_back(key.second);
                    }
                    pair_fop[k] = it->second;
                }
                build_dense_ops(c->m2l_dense, d.dtype, d.K, c->ops.ptr, fop_op.data(),
                                d.pair_sym ? fop_sym.data() : nullptr, c->perms_inv.as<int32_t>(),
                                (int64_t)fop_op.size(), d.dtype == H2G_C64 ? 0x7 : 0x3, dense_reps(c->precisions, true));
            }
            const int32_t *pf = c->m2l_dense.ok ? pair_fop.data() : nullptr;
            // transfers that are not separable (complex maps, general operators) on the dense kernel
            for (Side *S : {&c->test, c->galerkin ? nullptr : &c->trial_own}) {
                if (!S || S->separable || S->sep_cplx || S->n_levels < 2) continue;
                std::vector<int32_t> mats(S->n_levels * 8);
                for (int i = 0; i < (int)mats.size(); ++i) mats[i] = i;
                build_dense_ops(S->tr_dense, d.dtype, d.K, S->transfers.ptr, mats.data(), nullptr, nullptr,
                                (int64_t)mats.size(), d.dtype == H2G_C64 ? 0xF : 0x3, dense_reps(c->precisions, false));
            }
            build_orient(*c, c->fwd, *c->trial, c->test, d.pair_test, d.pair_trial, d.pair_op, d.pair_sym, pf, d.n_pairs);
            build_orient(*c, c->rev, c->test, *c->trial, d.pair_trial, d.pair_test, d.pair_op, d.pair_sym, pf, d.n_pairs);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int h2g_core_summary(const h2g_core *core, h2g_summary *out) {
    if (!core || !out) return H2G_EINVAL;
    const Core &c = *core;
    h2g_summary s{};
    s.rows = c.test.n_dofs;
    s.cols = c.trial->n_dofs;
    s.dtype = c.dtype;
    s.m2l_kind = c.m2l_kind;
    s.galerkin = c.galerkin ? 1 : 0;
    s.n_levels = c.test.n_levels;
    s.n_nodes = c.test.n_boxes;
    int64_t mn = INT64_MAX, mx = 0, tot = 0;
    for (int b = 0; b < c.test.n_boxes; ++b)
        if (c.test.is_leaf[b]) {
            const int64_t n = c.test.end[b] - c.test.begin[b];
            ++s.n_leaves;
            mn = std::min(mn, n);
            mx = std::max(mx, n);
            tot += n;
        }
    s.min_leaf_dofs = s.n_leaves ? mn : 0;
    s.max_leaf_dofs = mx;
    s.mean_leaf_dofs = s.n_leaves ? (double)tot / s.n_leaves : 0.0;
    s.far_pairs = c.n_pairs;
    s.unique_translations = c.n_unique_translations;
    s.reuse = c.n_unique_translations ? (double)c.n_pairs / (double)c.n_unique_translations : NAN;
    s.near_blocks = c.n_near_blocks;
    s.near_values = c.n_near_values;
    for (const LatticeLevel &LL : c.fwd.lattice) s.lattice_levels += LL.ok ? 1 : 0;
    s.dense_mfma = c.m2l_dense.ok ? 1 : 0;
    s.device_bytes = c.device_bytes();
    *out = s;
    return H2G_OK;
}

int h2g_core_sanitycheck(const h2g_core *core, h2g_issue *out, int32_t max, int32_t *n) {
    if (!core || !n) return H2G_EINVAL;
    return guarded([&] {
    const Core &c = *core;
    H2G_HIP(hipSetDevice(c.device));
    std::vector<h2g_issue> is;
    if (c.n_pairs == 0) is.push_back({1, H2G_ISSUE_NO_FAR_PAIRS, 0, -1});
    if (c.n_self_pairs) is.push_back({2, H2G_ISSUE_SELF_FAR_PAIR, c.n_self_pairs, -1});
    // operator ids and moments: a core is only created when every pair resolves (validate_desc)
    // and every box on a far pair's side gets a moment of K coefficients; the plan is re-checked
    // here from the orientation's work lists: each receiving box and each partner must be a box
    // whose moment the passes produce
    for (const OrientPlan *P : {&c.fwd, &c.rev}) {
        std::vector<int64_t> ptr(P->m2l_ptr.bytes / sizeof(int64_t));
        std::vector<int32_t> src(P->m2l_src.bytes / sizeof(int32_t));
        if (!ptr.empty()) H2G_HIP(hipMemcpy(ptr.data(), P->m2l_ptr.ptr, P->m2l_ptr.bytes, hipMemcpyDeviceToHost));
        if (!src.empty()) H2G_HIP(hipMemcpy(src.data(), P->m2l_src.ptr, P->m2l_src.bytes, hipMemcpyDeviceToHost));
        int64_t missing = 0;
        for (int32_t s : src)
            if (s < 0 || s >= P->src->n_boxes) ++missing;
        if (missing) is.push_back({2, H2G_ISSUE_MISSING_MOMENT, missing, -1});
        if (P == &c.fwd && is_toeplitz(c))
            for (int l = 0; l < (int)P->m2l_recv.size(); ++l)
                if (P->m2l_recv[l].count && !(l < (int)P->lattice.size() && P->lattice[l].ok))
                    is.push_back({0, H2G_ISSUE_GENERIC_COUPLING_LEVEL, P->m2l_recv[l].count, l});
    }
    *n = (int32_t)is.size();
    if (out)
        for (int32_t i = 0; i < std::min(max, *n); ++i) out[i] = is[i];
    });
}

int h2g_core_retain(h2g_core *core) {
    if (!core) return H2G_EINVAL;
    core->refs.fetch_add(1);
    return H2G_OK;
}

int h2g_core_release(h2g_core *core) {
    if (!core) return H2G_EINVAL;
    if (core->refs.fetch_sub(1) == 1) {
        (void)hipSetDevice(core->device);
        delete core;
    }
    return H2G_OK;
}

int h2g_core_size(const h2g_core *core, int64_t *rows, int64_t *cols) {
    if (!core || !rows || !cols) return H2G_EINVAL;
    *rows = core->test.n_dofs;
    *cols = core->trial->n_dofs;
    return H2G_OK;
}

int h2g_core_device_bytes(const h2g_core *core, int64_t *bytes) {
    if (!core || !bytes) return H2G_EINVAL;
    *bytes = core->device_bytes();
    return H2G_OK;
}

int h2g_ws_create(h2g_core *core, const h2g_ws_opts *opts, h2g_ws **out) {
    return guarded([&] {
        if (!core || !out) fail(H2G_EINVAL, "null argument");
        *out = nullptr;
        h2g_ws_opts o{};
        o.precision = H2G_PREC_PLAIN;
        o.max_rhs = 1;
        if (opts) o = *opts;
        if (o.max_rhs < 1) o.max_rhs = 1;
        if (o.precision < H2G_PREC_PLAIN || o.precision > H2G_PREC_BF16X3) fail(H2G_EINVAL, "unknown precision");
        if (o.m2l_groups < 0) fail(H2G_EINVAL, "m2l_groups must be >= 0");
        if (o.near_cols != 0 && o.near_cols != 16 && o.near_cols != 32 && o.near_cols != 64)
            fail(H2G_EINVAL, "near_cols must be 0, 16, 32 or 64");
        if (!(core->precisions & (1 << o.precision)))
            fail(H2G_EINVAL, "the core was built without this precision (h2g_core_desc.precisions)");
        H2G_HIP(hipSetDevice(core->device));
        auto *w = new h2g_ws();
        try {
            w->core = core;
            w->opts = o;
            w->R_cap = o.max_rhs;
            if (o.stream) {
                w->stream = w->side = static_cast<hipStream_t>(o.stream);
                w->own_streams = false;
            } else {
                H2G_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
                H2G_HIP(hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking));
            }
            H2G_HIP(hipEventCreateWithFlags(&w->ev_fork, hipEventDisableTiming));
            H2G_HIP(hipEventCreateWithFlags(&w->ev_near, hipEventDisableTiming));
            H2G_HIP(hipEventCreateWithFlags(&w->ev_done, hipEventDisableTiming));
            if (o.persistent != 0 && o.persistent != 1) fail(H2G_EINVAL, "persistent must be 0 or 1");
            if (o.dense_pairs < -1 || o.dense_pairs > 1) fail(H2G_EINVAL, "dense_pairs must be -1, 0 or 1");
            const WsLayout L = ws_layout(*core, o.precision != H2G_PREC_PLAIN, w->R_cap, o.near_cols, o.persistent);
            w->xs.alloc(L.xs);
            w->ys.alloc(L.ys);
            w->x32.alloc(L.x32);
            w->mom.alloc(L.mom);
            w->loc.alloc(L.loc);
            w->spec_in.alloc(L.spec_in);
            w->spec_out.alloc(L.spec_out);
            if (o.persistent) w->counters.alloc(8 * sizeof(int));
        } catch (...) {
            if (w->ev_done) (void)hipEventDestroy(w->ev_done);
            if (w->ev_near) (void)hipEventDestroy(w->ev_near);
            if (w->ev_fork) (void)hipEventDestroy(w->ev_fork);
            if (w->own_streams && w->side) (void)hipStreamDestroy(w->side);
            if (w->own_streams && w->stream) (void)hipStreamDestroy(w->stream);
            delete w;
            throw;
        }
        h2g_core_retain(core);
        *out = w;
    });
}

int h2g_ws_destroy(h2g_ws *ws) {
    if (!ws) return H2G_EINVAL;
    (void)hipSetDevice(ws->core->device);
    if (ws->ev_done) (void)hipEventSynchronize(ws->ev_done);
    if (ws->own_streams && ws->stream) (void)hipStreamSynchronize(ws->stream);
    if (ws->own_streams && ws->side) (void)hipStreamSynchronize(ws->side);
    for (auto &kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &e : ws->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (ws->ev_done) (void)hipEventDestroy(ws->ev_done);
    if (ws->ev_near) (void)hipEventDestroy(ws->ev_near);
    if (ws->ev_fork) (void)hipEventDestroy(ws->ev_fork);
    if (ws->own_streams && ws->side) (void)hipStreamDestroy(ws->side);
    if (ws->own_streams && ws->stream) (void)hipStreamDestroy(ws->stream);
    h2g_core *c = static_cast<h2g_core *>(ws->core);
    delete ws;
    return h2g_core_release(c);
}

int h2g_ws_device_bytes(const h2g_ws *ws, int64_t *bytes) {
    if (!ws || !bytes) return H2G_EINVAL;
    *bytes = ws->device_bytes();
    return H2G_OK;
}

void *h2g_ws_stream(h2g_ws *ws) { return ws ? (void *)ws->stream : nullptr; }

int h2g_ws_set_stream(h2g_ws *ws, void *stream) {
    if (!ws || !stream) return H2G_EINVAL;
    if (ws->own_streams) {
        set_error("h2g_ws_set_stream: the workspace runs on streams of its own");
        return H2G_EINVAL;
    }
    if (ws->busy.load()) {
        set_error("ConcurrentWorkspaceError: the workspace is already in use");
        return H2G_EBUSY;
    }
    return guarded([&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (s != ws->stream) {
            H2G_HIP(hipSetDevice(ws->core->device));
            H2G_HIP(hipStreamWaitEvent(s, ws->ev_done, 0));
            ws->stream = ws->side = s;
        }
    });
}

int h2g_ws_acquire(h2g_ws *ws) {
    if (!ws) return H2G_EINVAL;
    bool expected = false;
    if (!ws->busy.compare_exchange_strong(expected, true)) {
        set_error("ConcurrentWorkspaceError: the workspace is already in use");
        return H2G_EBUSY;
    }
    return H2G_OK;
}

int h2g_ws_release(h2g_ws *ws) {
    if (!ws) return H2G_EINVAL;
    ws->busy.store(false);
    return H2G_OK;
}

int h2g_ws_sync(h2g_ws *ws) {
    if (!ws) return H2G_EINVAL;
    return guarded([&] { H2G_HIP(hipEventSynchronize(ws->ev_done)); });
}

int h2g_ws_query(h2g_ws *ws) {
    if (!ws) return H2G_EINVAL;
    const hipError_t e = hipEventQuery(ws->ev_done);
    if (e == hipSuccess) return H2G_OK;
    if (e == hipErrorNotReady) return H2G_EAGAIN;
    set_error(std::string("hipStreamQuery: ") + hipGetErrorString(e));
    return H2G_EHIP;
}

int h2g_ws_timing(h2g_ws *ws, h2g_timing_record *out, int32_t max, int32_t *n) {
    if (!ws || !n) return H2G_EINVAL;
    *n = (int32_t)ws->timings.size();
    if (out)
        for (int32_t i = 0; i < std::min(max, *n); ++i) out[i] = ws->timings[i];
    return H2G_OK;
}

int h2g_matvec(h2g_ws *ws, int op, int which, int dtype, int64_t nrhs, const void *x, int64_t ldx, void *y,
               int64_t ldy, const void *alpha, const void *beta, int flags) {
    if (!ws) {
        set_error("null workspace");
        return H2G_EINVAL;
    }
    int rc = h2g_ws_acquire(ws);
    if (rc != H2G_OK) return rc;
    rc = guarded([&] {
        Core &c = *ws->core;
        cplx<double> a, b;
        check_call(c, op, which, dtype, nrhs, x, ldx, y, ldy, alpha, beta, a, b);
        const bool five = alpha != nullptr;
        if (nrhs == 0) return;
        H2G_HIP(hipSetDevice(c.device));
        hipStream_t s = ws->stream;
        const size_t xe = elem_size(dtype);
        const int64_t n_in = op == H2G_OP_N ? c.trial->n_dofs : c.test.n_dofs;
        const int64_t n_out = op == H2G_OP_N ? c.test.n_dofs : c.trial->n_dofs;
        const void *dx = x;
        void *dy = y;
        int64_t lx = ldx, ly = ldy;
        const bool host = flags & H2G_HOST_PTRS;
        const bool reads_y = five && (b.re != 0.0 || b.im != 0.0 || which == H2G_FARMUL);
        if (host) {
            // staging sized for the workspace's column capacity, so later calls do not reallocate
            // (a reallocation would wait for in-flight asynchronous products)
            const int64_t cols = std::max<int64_t>(nrhs, ws->R_cap);
            if (ws->xhost.bytes < (size_t)(n_in * nrhs) * xe) ws->xhost.alloc((size_t)(n_in * cols) * xe);
            if (ws->yhost.bytes < (size_t)(n_out * nrhs) * xe) ws->yhost.alloc((size_t)(n_out * cols) * xe);
            H2G_HIP(hipMemcpy2DAsync(ws->xhost.ptr, n_in * xe, x, ldx * xe, n_in * xe, nrhs, hipMemcpyHostToDevice, s));
            if (reads_y)
                H2G_HIP(hipMemcpy2DAsync(ws->yhost.ptr, n_out * xe, y, ldy * xe, n_out * xe, nrhs, hipMemcpyHostToDevice, s));
            dx = ws->xhost.ptr;
            dy = ws->yhost.ptr;
            lx = n_in;
            ly = n_out;
        }
        ws->timings.clear();
        ws->n_events_used = 0;
        ws->cur_op = op;
        run_passes(*ws, op, which, dtype, nrhs, dx, lx, dy, ly, a, b, five, true, ws->opts.use_graph && !ws->opts.timing);
        if (host)
            H2G_HIP(hipMemcpy2DAsync(y, ldy * xe, ws->yhost.ptr, n_out * xe, n_out * xe, nrhs, hipMemcpyDeviceToHost, s));
        H2G_HIP(hipEventRecord(ws->ev_done, s));
        if (!(flags & H2G_ASYNC) || ws->opts.timing) H2G_HIP(hipEventSynchronize(ws->ev_done));
        for (size_t i = 0; i < ws->n_events_used && i < ws->timings.size(); ++i)
            H2G_HIP(hipEventElapsedTime(&ws->timings[i].ms, ws->events[i].first, ws->events[i].second));
    });
    h2g_ws_release(ws);
    return rc;
}

int h2g_matvec_noalloc(h2g_core *core, int precision, int op, int which, int dtype, int64_t nrhs, const void *x,
                       int64_t ldx, void *y, int64_t ldy, const void *alpha, const void *beta, void *stream,
                       int flags) {
    if (!core) {
        set_error("null core");
        return H2G_EINVAL;
    }
    if (core->device < 0 || core->device >= NOALLOC_MAX_DEVICES) {
        set_error("h2g_matvec_noalloc: device ordinal out of range");
        return H2G_EINVAL;
    }
    std::atomic<bool> &busy = g_noalloc_busy[core->device];
    bool expected = false;
    if (!busy.compare_exchange_strong(expected, true)) {
        set_error("h2g_matvec_noalloc: another call is using this GPU's scratch");
        return H2G_EBUSY;
    }
    const int rc = guarded([&] {
        Core &c = *core;
        if (precision < H2G_PREC_PLAIN || precision > H2G_PREC_BF16X3) fail(H2G_EINVAL, "unknown precision");
        if (!(c.precisions & (1 << precision)))
            fail(H2G_EINVAL, "the core was built without this precision (h2g_core_desc.precisions)");
        if (flags & ~H2G_ASYNC) fail(H2G_EINVAL, "h2g_matvec_noalloc takes device pointers only (flags: 0 or H2G_ASYNC)");
        cplx<double> a, b;
        check_call(c, op, which, dtype, nrhs, x, ldx, y, ldy, alpha, beta, a, b);
        if (nrhs == 0) return;
        H2G_HIP(hipSetDevice(c.device));
        const bool mixed = precision != H2G_PREC_PLAIN;
        // the widest column block whose buffers fit the scratch
        int R = (int)std::min<int64_t>(nrhs, 1 << 16);
        while (R > 1 && ws_layout(c, mixed, R, 0, false).total(NOALLOC_ALIGN) > NOALLOC_BYTES) R = R > 8 ? R * 3 / 4 : R - 1;
        const WsLayout L = ws_layout(c, mixed, R, 0, false);
        if (L.total(NOALLOC_ALIGN) > NOALLOC_BYTES)
            fail(H2G_ENOMEM, "h2g_matvec_noalloc: one column of this map needs " + std::to_string(L.total(NOALLOC_ALIGN)) +
                                 " bytes of scratch, more than the " + std::to_string(NOALLOC_BYTES) +
                                 " bytes it holds; use h2g_matvec with a workspace");
        unsigned char *base = static_cast<unsigned char *>(noalloc_scratch(c.device));
        Workspace w;
        w.core = &c;
        w.opts.precision = precision;
        w.opts.max_rhs = R;
        w.R_cap = R;
        w.stream = static_cast<hipStream_t>(stream);
        w.side = w.stream;
        size_t off = 0;
        auto carve = [&](DevBuf &buf, size_t n) {
            buf.borrow(base + off, n);
            off += (n + NOALLOC_ALIGN - 1) / NOALLOC_ALIGN * NOALLOC_ALIGN;
        };
        carve(w.xs, L.xs);
        carve(w.ys, L.ys);
        carve(w.x32, L.x32);
        carve(w.mom, L.mom);
        carve(w.loc, L.loc);
        carve(w.spec_in, L.spec_in);
        carve(w.spec_out, L.spec_out);
        run_passes(w, op, which, dtype, nrhs, x, ldx, y, ldy, a, b, alpha != nullptr, false, false);
        if (!(flags & H2G_ASYNC)) H2G_HIP(hipStreamSynchronize(w.stream));
    });
    busy.store(false);
    return rc;
}


}  // extern "C"

// Lattice-blocked Toeplitz M2L.
//
// When the couplings of a level follow the parent-neighbour lattice (every receiving box b
// translates exactly the children of the 27 neighbours of its parent that are not adjacent to
// b), the level's M2L is a stencil over parent blocks: block = one receiving parent P, its 8
// child targets t and 216 source slots q = (neighbour d of P, child sector s).  The coupling
// of (t, q) depends only on the level and the offset 2d + bits(s) - bits(t).
//
// This file holds the modulation stages around the coupling stage (which runs on MFMA, in
// m2l_mfma.hip).  Spectra are frequency-major: S[f][slot][r] (slot = source box index within
// the level) and O[f][recv][r].  For a real map with Hermitian spectra only the half spectrum
// fx < (L+1)/2 is stored (real-to-complex modulation); complex maps keep the full L^3 spectrum.
#include <algorithm>
#include <type_traits>

#include "h2g_internal.h"

namespace h2g {

namespace {

constexpr int TPB = 256;
// FP32 half spectra (the mixed form of real maps): jobs and threads per workgroup of the two
// modulation kernels, measured on the N = 2M, p = 6 map (levels 5 + 6, ms; 256 threads and 16 jobs
// each: 6.83 in, 6.15 out): modulate-in 32 jobs (157 KB of LDS, one workgroup per CU) on 1024
// threads 5.81, on 512 threads 5.92; modulate-out 16 jobs on 1024 threads 5.12, on 512 threads
// 5.38, 32 jobs on 1024 threads 5.37.  More threads per job shorten each stage's rounds over the
// workgroup's lines; more jobs per workgroup lengthen modulate-in's contiguous stores per frequency
// (32 jobs x 8 bytes).
constexpr int FFT_IN_JB = 32, FFT_IN_TPB = 1024, FFT_OUT_JB = 16, FFT_OUT_TPB = 1024;

// streaming (nontemporal) access to the spectra, which are written once and read once from HBM
// (measured on the N = 2M, p = 6 map: modulate-in, modulate-out and the coupling stage together
// 0.8 ms faster than with ordinary stores and loads)
template <class C>
__device__ inline void store_stream(C *p, const C &v) {
    if constexpr (sizeof(C) == 8) {
        uint64_t u;
        __builtin_memcpy(&u, &v, 8);
        __builtin_nontemporal_store(u, reinterpret_cast<uint64_t *>(p));
    } else {
        *p = v;
    }
}
template <class C>
__device__ inline C load_stream(const C *p) {
    if constexpr (sizeof(C) == 8) {
        const uint64_t u = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(p));
        C v;
        __builtin_memcpy(&v, &u, 8);
        return v;
    } else {
        return *p;
    }
}

// --------------------------------------------------------------------------------------------
// Modulate-in: pad + 3-D DFT of JB (box, rhs) jobs per workgroup, frequency-major output.
//   S[f][slot][r], f = fx + FX*(fy + L*fz), FX = HALF ? (L+1)/2 : L.
// --------------------------------------------------------------------------------------------
// Each stage is a batch of 1-D DFTs along one axis: a thread loads one input line (p or L
// values) into registers and writes the whole output line, with the L twiddles in registers.
template <class C, int L>
__device__ inline void load_twiddles(C (&tw)[L], const cplx<double> *__restrict__ twiddle, bool inverse) {
    using Rl = typename real_of<C>::type;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const cplx<double> w = twiddle[i];
        tw[i] = C{(Rl)w.re, (Rl)(inverse ? -w.im : w.im)};
    }
}

// Rrow: stored columns per source row (R; 16 for the direct kernel's narrow batches; R rounded up
// to 32 when perm); perm: within each
// chunk of 32 columns, column c is stored at 2 (c % 16) + c / 16, so that columns c and c + 16
// are adjacent (one 16-byte load holds both); padding columns (>= R) are zero.
template <class T, bool HALF, int JB, int P>
__global__ void fft_in_fm_kernel(int box0, int nbox, const T *__restrict__ mom, int K, int R, int Rrow, bool perm,
                                 typename complex_of<T>::type *__restrict__ S,
                                 const cplx<double> *__restrict__ twiddle, bool inverse) {
    using C = typename complex_of<T>::type;
    using Rl = typename real_of<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int p = P, L = 2 * P - 1;
    constexpr int FX = HALF ? (L + 1) / 2 : L;
    constexpr int n1 = FX * p * p, n2 = FX * L * p;
    C *b1 = reinterpret_cast<C *>(smem_raw);  // [JB][FX*p*p]
    C *b2 = b1 + JB * n1;                     // [JB][FX*L*p]
    C tw[L];
    load_twiddles(tw, twiddle, inverse);
    const int64_t njobs = (int64_t)nbox * Rrow;
    const int64_t job0 = (int64_t)blockIdx.x * JB;
    // x: line (iy, iz) of a job -> b1[fx + FX*(iy + p*iz)], fx < FX
    for (int i = threadIdx.x; i < JB * p * p; i += blockDim.x) {
        const int j = i / (p * p), rest = i - j * (p * p);
        const int64_t job = job0 + j;
        C in[p];
        const int64_t slot = job / Rrow;
        const int q = (int)(job - slot * Rrow);
        const int r = perm ? (q & ~31) + ((q & 31) >> 1) + 16 * (q & 1) : q;
        if (job < njobs && r < R) {
            const T *mj = mom + ((int64_t)(box0 + slot) * R + r) * K + p * rest;
#pragma unroll
            for (int ix = 0; ix < p; ++ix) in[ix] = cvt<C>::from(mj[ix]);
        } else {
#pragma unroll
            for (int ix = 0; ix < p; ++ix) in[ix] = zero_v<C>();
        }
#pragma unroll
        for (int fx = 0; fx < FX; ++fx) {
            C acc = zero_v<C>();
            if constexpr (is_complex<T>::value) {
#pragma unroll
                for (int ix = 0; ix < p; ++ix) fmacc(acc, tw[(fx * ix) % L], in[ix]);
            } else {
                // real moments: real x complex twiddle (2 FMA per term instead of 4)
#pragma unroll
                for (int ix = 0; ix < p; ++ix) fmacc(acc, in[ix].re, tw[(fx * ix) % L]);
            }
            b1[j * n1 + fx + FX * rest] = acc;
        }
    }
    __syncthreads();
    // y: line (fx, iz) -> b2[fx + FX*(fy + L*iz)], fy < L
    for (int i = threadIdx.x; i < JB * FX * p; i += blockDim.x) {
        const int j = i / (FX * p), e = i - j * (FX * p);
        const int fx = e % FX, iz = e / FX;
        C in[p];
#pragma unroll
        for (int iy = 0; iy < p; ++iy) in[iy] = b1[j * n1 + fx + FX * (iy + p * iz)];
#pragma unroll
        for (int fy = 0; fy < L; ++fy) {
            C acc = zero_v<C>();
#pragma unroll
            for (int iy = 0; iy < p; ++iy) fmacc(acc, tw[(fy * iy) % L], in[iy]);
            b2[j * n2 + fx + FX * (fy + L * iz)] = acc;
        }
    }
    __syncthreads();
    // z: line (fx, fy) -> S[f][job], f = fx + FX*(fy + L*fz); jobs fastest for contiguous writes
    const Rl scale = inverse ? (Rl)(1.0 / (double)(L * L * L)) : (Rl)1;
    for (int i = threadIdx.x; i < JB * FX * L; i += blockDim.x) {
        const int j = i % JB, fxy = i / JB;
        const int64_t job = job0 + j;
        if (job >= njobs) continue;
        C in[p];
#pragma unroll
        for (int iz = 0; iz < p; ++iz) in[iz] = b2[j * n2 + fxy + FX * L * iz];
#pragma unroll
        for (int fz = 0; fz < L; ++fz) {
            // half spectra: the conjugate half of the plane fx = 0 is not stored
            if (HALF && !half_freq_kept(L, fxy % FX, fxy / FX, fz)) continue;
            C acc = zero_v<C>();
#pragma unroll
            for (int iz = 0; iz < p; ++iz) fmacc(acc, tw[(fz * iz) % L], in[iz]);
            const int f = HALF ? half_freq_index(L, fxy % FX, fxy / FX, fz) : fxy + FX * L * fz;
            store_stream(&S[(int64_t)f * njobs + job], acc * scale);
        }
    }
}

// --------------------------------------------------------------------------------------------
// Modulate-out: 3-D inverse DFT + crop of JB (recv, rhs) jobs, from O[f][recv][r]; writes
// (overwrites) loc[box][r][k].  HALF: real output from the half spectrum.
// --------------------------------------------------------------------------------------------
template <class T, bool HALF, int JB, int P>
__global__ void fft_out_fm_kernel(const int32_t *__restrict__ recv_boxes, int nrecv,
                                  const typename complex_of<T>::type *__restrict__ O, int K, int R,
                                  T *__restrict__ loc, const cplx<double> *__restrict__ twiddle, bool inverse) {
    using C = typename complex_of<T>::type;
    using Rl = typename real_of<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int p = P, L = 2 * P - 1;
    constexpr int FX = HALF ? (L + 1) / 2 : L;
    constexpr int n1 = FX * L * p, n2 = FX * p * p;
    C *a1 = reinterpret_cast<C *>(smem_raw);  // [JB][FX*L*p]
    C *a2 = a1 + JB * n1;                     // [JB][FX*p*p]
    C tw[L];
    load_twiddles(tw, twiddle, inverse);
    const int64_t njobs = (int64_t)nrecv * R;
    const int64_t job0 = (int64_t)blockIdx.x * JB;
    // z: line (fx, fy) of the spectrum -> a1[fxy + FX*L*iz], iz < p
    for (int i = threadIdx.x; i < JB * FX * L; i += blockDim.x) {
        const int j = i % JB, fxy = i / JB;
        const int64_t job = job0 + j;
        C in[L];
        // every load is issued (clamped job, valid address) and selected afterwards, so the L loads
        // of a line go out together; half spectra: a value of the plane fx = 0 that is not stored
        // is the conjugate of the stored one at (0, -fy, -fz)
        const int64_t jc = job < njobs ? job : njobs - 1;
        const int fx = fxy % FX, fy = fxy / FX;
#pragma unroll
        for (int fz = 0; fz < L; ++fz) {
            int f = fxy + FX * L * fz;
            bool conj = false;
            if (HALF) {
                conj = !half_freq_kept(L, fx, fy, fz);
                f = conj ? half_freq_index(L, 0, (L - fy) % L, (L - fz) % L) : half_freq_index(L, fx, fy, fz);
            }
            const C v = load_stream(&O[(int64_t)f * njobs + jc]);
            const Rl z = job < njobs ? (Rl)1 : (Rl)0;
            in[fz] = C{v.re * z, (conj ? -v.im : v.im) * z};
        }
#pragma unroll
        for (int iz = 0; iz < p; ++iz) {
            C acc = zero_v<C>();
#pragma unroll
            for (int fz = 0; fz < L; ++fz) fmacc(acc, tw[(fz * iz) % L], in[fz]);
            a1[j * n1 + fxy + FX * L * iz] = acc;
        }
    }
    __syncthreads();
    // y: line (fx, iz) -> a2[fx + FX*(iy + p*iz)], iy < p
    for (int i = threadIdx.x; i < JB * FX * p; i += blockDim.x) {
        const int j = i / (FX * p), e = i - j * (FX * p);
        const int fx = e % FX, iz = e / FX;
        C in[L];
#pragma unroll
        for (int fy = 0; fy < L; ++fy) in[fy] = a1[j * n1 + fx + FX * (fy + L * iz)];
#pragma unroll
        for (int iy = 0; iy < p; ++iy) {
            C acc = zero_v<C>();
#pragma unroll
            for (int fy = 0; fy < L; ++fy) fmacc(acc, tw[(fy * iy) % L], in[fy]);
            a2[j * n2 + fx + FX * (iy + p * iz)] = acc;
        }
    }
    __syncthreads();
    // x: line (iy, iz) -> loc[... + ix + p*(iy + p*iz)], ix < p
    const Rl scale = inverse ? (Rl)(1.0 / (double)(L * L * L)) : (Rl)1;
    for (int i = threadIdx.x; i < JB * p * p; i += blockDim.x) {
        const int j = i / (p * p), rest = i - j * (p * p);
        const int64_t job = job0 + j;
        if (job >= njobs) continue;
        C in[FX];
#pragma unroll
        for (int fx = 0; fx < FX; ++fx) in[fx] = a2[j * n2 + fx + FX * rest];
        const int64_t slot = job / R;
        const int r = (int)(job - slot * R);
        T *dst = loc + ((int64_t)recv_boxes[slot] * R + r) * K + p * rest;
#pragma unroll
        for (int ix = 0; ix < p; ++ix) {
            C acc = zero_v<C>();
            if (HALF) {
                // Hermitian spectrum: y = a2[0] + 2 Re sum_{fx>=1} a2[fx] w^(fx*ix)
                C tail = zero_v<C>();
#pragma unroll
                for (int fx = 1; fx < FX; ++fx) fmacc(tail, tw[(fx * ix) % L], in[fx]);
                acc = C{in[0].re + (Rl)2 * tail.re, (Rl)0};
            } else {
#pragma unroll
                for (int fx = 0; fx < FX; ++fx) fmacc(acc, tw[(fx * ix) % L], in[fx]);
            }
            dst[ix] = cvt<T>::from(acc * scale);
        }
    }
}

template <int JB, class T, bool HALF>
size_t fft_smem(int p, int L) {
    using C = typename complex_of<T>::type;
    const int FX = HALF ? (L + 1) / 2 : L;
    return sizeof(C) * (size_t)(JB * (FX * p * p + FX * L * p));
}

// jobs per workgroup (other than FP32 half spectra): 16 FP32, 8 FP64 (78 KB of LDS at p = 6)
template <class T, bool HALF>
constexpr int fft_jobs() {
    return (sizeof(typename real_of<T>::type) == 4 ? 16 : 8) / (HALF ? 1 : 2) / (is_complex<T>::value ? 1 : 1);
}

// LDS a workgroup may allocate on the current device (queried once per device)
size_t device_lds_limit() {
    static size_t cache[64] = {0};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) return 65536;
    if (!cache[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        cache[dev] = (size_t)v;
    }
    return cache[dev];
}

template <class K>
bool allow_lds(K kernel, size_t bytes) {
    if (bytes > device_lds_limit()) return false;
    if (bytes > 65536)
        return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
               hipSuccess;
    return true;
}

}  // namespace

#define H2G_CHECK_LAUNCH2()                                                                   \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) throw Error{H2G_EHIP, std::string("launch: ") + hipGetErrorString(e_)}; \
    } while (0)

// FP32 half spectra (the mixed form of real maps): jobs and threads per workgroup of each
// modulation kernel
template <class T, bool HALF>
constexpr bool fft_f32_half() {
    return HALF && !is_complex<T>::value && sizeof(T) == 4;
}

template <class T, bool HALF, int P>
static void fft_in_fm_p(int box0, int nbox, const void *mom, int K, int R, int Rrow, bool perm, void *S, const void *tw,
                        bool inverse, hipStream_t s) {
    constexpr int JB = fft_f32_half<T, HALF>() ? FFT_IN_JB : fft_jobs<T, HALF>();
    // ComplexF32 full spectra (complex maps, mixed form) on 512 threads: on a Helmholtz map of 1.2M
    // points at 32 columns 11.5 ms on 256 threads, 9.7 on 512, 9.5 on 1024 (modulate-out unchanged)
    constexpr int NT = fft_f32_half<T, HALF>() ? FFT_IN_TPB : (std::is_same<T, cplx<float>>::value ? 512 : TPB);
    constexpr int JH = JB / 2;
    constexpr int L = 2 * P - 1;
    const int64_t njobs = (int64_t)nbox * Rrow;
    const size_t sh = fft_smem<JB, T, HALF>(P, L);
    using C = typename complex_of<T>::type;
    if (allow_lds(fft_in_fm_kernel<T, HALF, JB, P>, sh)) {
        fft_in_fm_kernel<T, HALF, JB, P><<<(unsigned)((njobs + JB - 1) / JB), NT, sh, s>>>(
            box0, nbox, (const T *)mom, K, R, Rrow, perm, (C *)S, (const cplx<double> *)tw, inverse);
    } else {
        fft_in_fm_kernel<T, HALF, JH, P><<<(unsigned)((njobs + JH - 1) / JH), NT, fft_smem<JH, T, HALF>(P, L), s>>>(
            box0, nbox, (const T *)mom, K, R, Rrow, perm, (C *)S, (const cplx<double> *)tw, inverse);
    }
    H2G_CHECK_LAUNCH2();
}

template <class T, bool HALF, int P>
static void fft_out_fm_p(const int32_t *recv, int nrecv, const void *O, int K, int R, void *loc, const void *tw,
                         bool inverse, hipStream_t s) {
    constexpr int JB = fft_f32_half<T, HALF>() ? FFT_OUT_JB : fft_jobs<T, HALF>();
    constexpr int NT = fft_f32_half<T, HALF>() ? FFT_OUT_TPB : TPB;
    constexpr int JH = JB / 2;
    constexpr int L = 2 * P - 1;
    const int64_t njobs = (int64_t)nrecv * R;
    const size_t sh = fft_smem<JB, T, HALF>(P, L);
    using C = typename complex_of<T>::type;
    if (allow_lds(fft_out_fm_kernel<T, HALF, JB, P>, sh)) {
        fft_out_fm_kernel<T, HALF, JB, P><<<(unsigned)((njobs + JB - 1) / JB), NT, sh, s>>>(
            recv, nrecv, (const C *)O, K, R, (T *)loc, (const cplx<double> *)tw, inverse);
    } else {
        fft_out_fm_kernel<T, HALF, JH, P><<<(unsigned)((njobs + JH - 1) / JH), NT, fft_smem<JH, T, HALF>(P, L), s>>>(
            recv, nrecv, (const C *)O, K, R, (T *)loc, (const cplx<double> *)tw, inverse);
    }
    H2G_CHECK_LAUNCH2();
}

template <class T, bool HALF>
static void fft_in_fm_t(int box0, int nbox, const void *mom, int K, int p, int L, int R, int Rrow, bool perm, void *S,
                        const void *tw, bool inverse, hipStream_t s) {
    switch (p) {
    case 2: fft_in_fm_p<T, HALF, 2>(box0, nbox, mom, K, R, Rrow, perm, S, tw, inverse, s); break;
    case 3: fft_in_fm_p<T, HALF, 3>(box0, nbox, mom, K, R, Rrow, perm, S, tw, inverse, s); break;
    case 4: fft_in_fm_p<T, HALF, 4>(box0, nbox, mom, K, R, Rrow, perm, S, tw, inverse, s); break;
    case 5: fft_in_fm_p<T, HALF, 5>(box0, nbox, mom, K, R, Rrow, perm, S, tw, inverse, s); break;
    case 6: fft_in_fm_p<T, HALF, 6>(box0, nbox, mom, K, R, Rrow, perm, S, tw, inverse, s); break;
    default: throw Error{H2G_EUNSUPPORTED, "Toeplitz modulation supports 2 <= p <= 6"};
    }
    (void)L;
}

template <class T, bool HALF>
static void fft_out_fm_t(const int32_t *recv, int nrecv, const void *O, int K, int p, int L, int R, void *loc,
                         const void *tw, bool inverse, hipStream_t s) {
    switch (p) {
    case 2: fft_out_fm_p<T, HALF, 2>(recv, nrecv, O, K, R, loc, tw, inverse, s); break;
    case 3: fft_out_fm_p<T, HALF, 3>(recv, nrecv, O, K, R, loc, tw, inverse, s); break;
    case 4: fft_out_fm_p<T, HALF, 4>(recv, nrecv, O, K, R, loc, tw, inverse, s); break;
    case 5: fft_out_fm_p<T, HALF, 5>(recv, nrecv, O, K, R, loc, tw, inverse, s); break;
    case 6: fft_out_fm_p<T, HALF, 6>(recv, nrecv, O, K, R, loc, tw, inverse, s); break;
    default: throw Error{H2G_EUNSUPPORTED, "Toeplitz modulation supports 2 <= p <= 6"};
    }
    (void)L;
}

void launch_fft_in_fm(int cdtype, bool half, bool inverse, int box0, int nbox, const void *mom, int K, int p, int L,
                      int R, int Rrow, bool perm, void *S, const void *twiddle, hipStream_t s) {
    if (nbox == 0) return;
    switch (cdtype) {
    case H2G_F64:
        if (half) fft_in_fm_t<double, true>(box0, nbox, mom, K, p, L, R, Rrow, perm, S, twiddle, inverse, s);
        else fft_in_fm_t<double, false>(box0, nbox, mom, K, p, L, R, Rrow, perm, S, twiddle, inverse, s);
        break;
    case H2G_F32:
        if (half) fft_in_fm_t<float, true>(box0, nbox, mom, K, p, L, R, Rrow, perm, S, twiddle, inverse, s);
        else fft_in_fm_t<float, false>(box0, nbox, mom, K, p, L, R, Rrow, perm, S, twiddle, inverse, s);
        break;
    case H2G_C64: fft_in_fm_t<cplx<double>, false>(box0, nbox, mom, K, p, L, R, Rrow, perm, S, twiddle, inverse, s); break;
    default: fft_in_fm_t<cplx<float>, false>(box0, nbox, mom, K, p, L, R, Rrow, perm, S, twiddle, inverse, s);
    }
}

void launch_fft_out_fm(int cdtype, bool half, bool inverse, const int32_t *recv, int nrecv, const void *O, int K,
                       int p, int L, int R, void *loc, const void *twiddle, hipStream_t s) {
    if (nrecv == 0) return;
    switch (cdtype) {
    case H2G_F64:
        if (half) fft_out_fm_t<double, true>(recv, nrecv, O, K, p, L, R, loc, twiddle, inverse, s);
        else fft_out_fm_t<double, false>(recv, nrecv, O, K, p, L, R, loc, twiddle, inverse, s);
        break;
    case H2G_F32:
        if (half) fft_out_fm_t<float, true>(recv, nrecv, O, K, p, L, R, loc, twiddle, inverse, s);
        else fft_out_fm_t<float, false>(recv, nrecv, O, K, p, L, R, loc, twiddle, inverse, s);
        break;
    case H2G_C64: fft_out_fm_t<cplx<double>, false>(recv, nrecv, O, K, p, L, R, loc, twiddle, inverse, s); break;
    default: fft_out_fm_t<cplx<float>, false>(recv, nrecv, O, K, p, L, R, loc, twiddle, inverse, s);
    }
}

}  // namespace h2g

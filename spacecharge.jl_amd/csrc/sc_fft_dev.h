// sc_fft_dev.h -- workgroup-cooperative mixed-radix FFT in LDS (gfx950).
//
// A workgroup of NT lanes (kFftThreads, or kFftThreadsXL for the longest
// pencils) transforms a batch of nb pencils of one length n held in LDS,
// each lane holding up to E complex values of a stage (nb*n <= E*NT; E =
// fft_elems<T>() by default, twice that for the long pencils of axes beyond
// fft_cap<T>()).  Element e of pencil b lives
// at buf[e*es + b*ps]:
//   es = nb, ps = 1  ("element-major") for pencils that are strided in HBM
//                    and batched over the contiguous axis, so a global
//                    load of one element row is one coalesced segment;
//   es = 1,  ps = n  ("pencil-major") for pencils that are contiguous in HBM.
// The algorithm is Stockham autosort, decimation in time: stage s with radix
// R reads x[j + q*n/R] (q < R), twiddles by w_{Ns*R}^{q*(j mod Ns)}, does an
// R-point DFT and writes to (j/Ns)*Ns*R + (j mod Ns) + q*Ns.  Each stage
// reads everything a lane needs into registers, barriers, then writes in
// place, so one LDS buffer suffices.  Twiddles come from an LDS copy of
// tw[m] = exp(-2 pi i m/n); sign = -1 is the forward transform, +1 the
// unnormalised backward transform.
#pragma once

#include <hip/hip_runtime.h>

#include "sc_internal.h"

namespace sc {

template <typename T>
struct cpx {
    T x, y;
};

// Loads and stores through the global address space.  Kernels that may take
// their buffers from a descriptor table in device memory (the grouped
// launches) cannot show the compiler that those pointers are global, and
// plain accesses become flat instructions, which also count against the LDS
// wait counter: every LDS exchange of a transform then waits for the memory
// operations still in flight.  Complex values go as their two components
// (the compiler merges them into one 8- or 16-byte access).
template <typename T>
using gbl = __attribute__((address_space(1))) T;

template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
    return *(const gbl<T> *)p;
}
template <typename T>
__device__ __forceinline__ cpx<T> gld(const cpx<T> *p)
{
    const gbl<T> *q = (const gbl<T> *)p;
    return cpx<T>{q[0], q[1]};
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v)
{
    *(gbl<T> *)p = v;
}
template <typename T>
__device__ __forceinline__ void gst(cpx<T> *p, cpx<T> v)
{
    gbl<T> *q = (gbl<T> *)p;
    q[0] = v.x;
    q[1] = v.y;
}

template <typename T>
__device__ __forceinline__ cpx<T> c_add(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> c_sub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> c_mul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i*sign... i.e. by -i for the forward (sign<0), +i for backward
template <typename T>
__device__ __forceinline__ cpx<T> c_rot(cpx<T> a, int sign)
{
    return sign < 0 ? cpx<T>{a.y, -a.x} : cpx<T>{-a.y, a.x};
}

// R-point DFT of v in registers.  w_R^m = tw[m * (n/R)] (conjugated for
// the backward transform).
template <typename T, int R>
__device__ __forceinline__ void dft_r(cpx<T> *v, const cpx<T> *tw, int n, int sign)
{
    if constexpr (R == 2) {
        cpx<T> a = v[0], b = v[1];
        v[0] = c_add(a, b);
        v[1] = c_sub(a, b);
    } else if constexpr (R == 4) {
        cpx<T> t0 = c_add(v[0], v[2]), t1 = c_sub(v[0], v[2]);
        cpx<T> t2 = c_add(v[1], v[3]), t3 = c_rot(c_sub(v[1], v[3]), sign);
        v[0] = c_add(t0, t2);
        v[1] = c_add(t1, t3);
        v[2] = c_sub(t0, t2);
        v[3] = c_sub(t1, t3);
    } else if constexpr (R == 8) {
        // two radix-4 halves (even / odd inputs), then w_8^k combination
        cpx<T> e0 = c_add(v[0], v[4]), e1 = c_sub(v[0], v[4]);
        cpx<T> e2 = c_add(v[2], v[6]), e3 = c_rot(c_sub(v[2], v[6]), sign);
        cpx<T> E0 = c_add(e0, e2), E1 = c_add(e1, e3), E2 = c_sub(e0, e2), E3 = c_sub(e1, e3);
        cpx<T> o0 = c_add(v[1], v[5]), o1 = c_sub(v[1], v[5]);
        cpx<T> o2 = c_add(v[3], v[7]), o3 = c_rot(c_sub(v[3], v[7]), sign);
        cpx<T> O0 = c_add(o0, o2), O1 = c_add(o1, o3), O2 = c_sub(o0, o2), O3 = c_sub(o1, o3);
        const T h = (T)0.70710678118654752440084436210484903928;
        // w8^1 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 backward
        cpx<T> W1O1 = sign < 0 ? cpx<T>{(O1.x + O1.y) * h, (O1.y - O1.x) * h}
                               : cpx<T>{(O1.x - O1.y) * h, (O1.y + O1.x) * h};
        cpx<T> W2O2 = c_rot(O2, sign);
        // w8^3 = (-1 - i)/sqrt2 forward, (-1 + i)/sqrt2 backward
        cpx<T> W3O3 = sign < 0 ? cpx<T>{(-O3.x + O3.y) * h, (-O3.y - O3.x) * h}
                               : cpx<T>{(-O3.x - O3.y) * h, (-O3.y + O3.x) * h};
        v[0] = c_add(E0, O0);
        v[4] = c_sub(E0, O0);
        v[1] = c_add(E1, W1O1);
        v[5] = c_sub(E1, W1O1);
        v[2] = c_add(E2, W2O2);
        v[6] = c_sub(E2, W2O2);
        v[3] = c_add(E3, W3O3);
        v[7] = c_sub(E3, W3O3);
    }
}

// Odd-prime R: output p of the R-point DFT, straight from the definition.
template <typename T, int R, int P>
__device__ __forceinline__ cpx<T> dft_prime_out(const cpx<T> *v, const cpx<T> *tw, int step,
                                                int sign)
{
    cpx<T> s = v[0];
#pragma unroll
    for (int q = 1; q < R; ++q) {
        cpx<T> w = tw[((P * q) % R) * step];
        if (sign > 0) w.y = -w.y;
        s = c_add(s, c_mul(v[q], w));
    }
    return s;
}

template <typename T, int R, int P>
__device__ __forceinline__ void dft_prime_store(const cpx<T> *v, const cpx<T> *tw, int step,
                                                int sign, cpx<T> *buf, int base, int Ns, int es,
                                                int bps)
{
    if constexpr (P < R) {
        buf[(base + P * Ns) * es + bps] = dft_prime_out<T, R, P>(v, tw, step, sign);
        dft_prime_store<T, R, P + 1>(v, tw, step, sign, buf, base, Ns, es, bps);
    }
}

template <typename T, int R, int E, int NT>
__device__ __forceinline__ void fft_stage(cpx<T> *buf, const cpx<T> *tw, int n, int Ns, int nb,
                                          int es, int ps, bool bfast, int sign)
{
    constexpr int MAXB = (E + R - 1) / R;
    const int nR = n / R;
    const int TB = nb * nR;
    const int tstep = n / (Ns * R);
    cpx<T> v[MAXB][R];
#pragma unroll
    for (int s = 0; s < MAXB; ++s) {
        const int bf = (int)threadIdx.x + s * NT;
        if (bf < TB) {
            int b, j;
            if (bfast) { b = bf % nb; j = bf / nb; }
            else { j = bf % nR; b = bf / nR; }
#pragma unroll
            for (int q = 0; q < R; ++q) v[s][q] = buf[(j + q * nR) * es + b * ps];
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXB; ++s) {
        const int bf = (int)threadIdx.x + s * NT;
        if (bf < TB) {
            int b, j;
            if (bfast) { b = bf % nb; j = bf / nb; }
            else { j = bf % nR; b = bf / nR; }
            const int k = j % Ns;
            if (k != 0) {
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    cpx<T> w = tw[q * k * tstep];
                    if (sign > 0) w.y = -w.y;
                    v[s][q] = c_mul(v[s][q], w);
                }
            }
            const int base = (j / Ns) * Ns * R + k;
            if constexpr (R == 2 || R == 4 || R == 8) {
                dft_r<T, R>(v[s], tw, n, sign);
#pragma unroll
                for (int q = 0; q < R; ++q) buf[(base + q * Ns) * es + b * ps] = v[s][q];
            } else {
                // every lane read its inputs before the barrier above, so each
                // output can go to LDS as soon as it is formed
                dft_prime_store<T, R, 0>(v[s], tw, n / R, sign, buf, base, Ns, es, b * ps);
            }
        }
    }
    __syncthreads();
}

// Transform nb pencils of length pl.n in LDS.  All NT lanes of the
// workgroup must call it (it contains barriers).  Requires nb*n <= E*NT,
// which the host guarantees.  tw may be in LDS or in global memory.
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__device__ void lds_fft(cpx<T> *buf, const cpx<T> *tw, const FftPlan &pl, int nb, int es, int ps,
                        bool bfast, int sign)
{
    const int n = pl.n;
    int Ns = 1;
    for (int st = 0; st < pl.nst; ++st) {
        const int R = pl.rad[st];
        switch (R) {
        case 2: fft_stage<T, 2, E, NT>(buf, tw, n, Ns, nb, es, ps, bfast, sign); break;
        case 3: fft_stage<T, 3, E, NT>(buf, tw, n, Ns, nb, es, ps, bfast, sign); break;
        case 4: fft_stage<T, 4, E, NT>(buf, tw, n, Ns, nb, es, ps, bfast, sign); break;
        case 5: fft_stage<T, 5, E, NT>(buf, tw, n, Ns, nb, es, ps, bfast, sign); break;
        case 7: fft_stage<T, 7, E, NT>(buf, tw, n, Ns, nb, es, ps, bfast, sign); break;
        case 8: fft_stage<T, 8, E, NT>(buf, tw, n, Ns, nb, es, ps, bfast, sign); break;
        default: break;  // unreachable: make_plan rejects other radices
        }
        Ns *= R;
    }
}

// Copy a twiddle table of n entries into LDS (all lanes call; ends with a barrier).
template <typename T>
__device__ __forceinline__ void load_twiddles(cpx<T> *dst, const cpx<T> *src, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

}  // namespace sc

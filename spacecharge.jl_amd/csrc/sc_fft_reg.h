// sc_fft_reg.h -- register-resident pencil FFT for power-of-two lengths.
//
// A pencil of length L is held by TPP = L/E lanes, E elements each (E = 16
// for Float32, 8 for Float64).  Lane j starts with x[j + r*TPP], r < E (the
// "input layout"), runs Stockham passes of radix E (plus one smaller final
// radix when L is not a power of E) entirely in registers, and exchanges
// data through LDS only between passes.  After the last pass lane j holds
// X[j + r*TPP] again, with a compile-time register permutation, so a
// forward transform can feed a backward one without touching memory.
//
// Stockham pass with radix R, NS = product of the earlier radices: butterfly
// b (b < L/R) reads x[b + q*L/R], multiplies element q by w_L^(q*(b mod NS)*
// L/(NS*R)), does an R-point DFT and writes output q to
// (b/NS)*NS*R + (b mod NS) + q*NS.  Lane j owns butterflies j + s*TPP,
// s < E/R.
#pragma once

#include <hip/hip_runtime.h>

#include "sc_fft_dev.h"

namespace sc {
namespace reg {

template <typename T>
constexpr int elems() { return sizeof(T) == 4 ? 16 : 8; }

constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// multiply by -i (DIR < 0) or +i (DIR > 0)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rot(cpx<T> a)
{
    if constexpr (DIR < 0) return cpx<T>{a.y, -a.x};
    else return cpx<T>{-a.y, a.x};
}

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cpx<T> &a, cpx<T> &b)
{
    const cpx<T> t = a;
    a = c_add(t, b);
    b = c_sub(t, b);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cpx<T> &a0, cpx<T> &a1, cpx<T> &a2, cpx<T> &a3)
{
    const cpx<T> t0 = c_add(a0, a2), t1 = c_sub(a0, a2);
    const cpx<T> t2 = c_add(a1, a3), t3 = rot<DIR>(c_sub(a1, a3));
    a0 = c_add(t0, t2);
    a2 = c_sub(t0, t2);
    a1 = c_add(t1, t3);
    a3 = c_sub(t1, t3);
}

// w_16^m for m in [0, 16): cos and sin(2*pi*m/16) with the direction's sign
template <int DIR, typename T, int M>
__device__ __forceinline__ cpx<T> w16(cpx<T> a)
{
    constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                              0.38268343236508977173, 0.0, -0.38268343236508977173,
                              -0.70710678118654752440, -0.92387953251128675613, -1.0,
                              -0.92387953251128675613, -0.70710678118654752440,
                              -0.38268343236508977173, 0.0, 0.38268343236508977173,
                              0.70710678118654752440, 0.92387953251128675613};
    constexpr int m = M & 15;
    constexpr double cr = c[m];
    constexpr double si = c[(m + 12) & 15];  // sin(2 pi m/16) = cos(2 pi (m-4)/16)
    constexpr double wi = DIR < 0 ? -si : si;
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) return rot<DIR>(a);
    else if constexpr (m == 8) return cpx<T>{-a.x, -a.y};
    else if constexpr (m == 12) return rot<-DIR>(a);
    else return cpx<T>{a.x * (T)cr - a.y * (T)wi, a.x * (T)wi + a.y * (T)cr};
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8(cpx<T> *v)
{
    // even / odd radix-4 halves, then w_8^k = w_16^(2k)
    cpx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = w16<DIR, T, 2>(o1);
    o2 = w16<DIR, T, 4>(o2);
    o3 = w16<DIR, T, 6>(o3);
    v[0] = c_add(e0, o0);
    v[4] = c_sub(e0, o0);
    v[1] = c_add(e1, o1);
    v[5] = c_sub(e1, o1);
    v[2] = c_add(e2, o2);
    v[6] = c_sub(e2, o2);
    v[3] = c_add(e3, o3);
    v[7] = c_sub(e3, o3);
}

// 16-point DFT as 4 x 4: X[k1 + 4 k2] = sum_n2 w16^(n2 k1) w4^(n2 k2) DFT4_n1(x[4 n1 + n2])[k1]
template <int DIR, typename T>
__device__ __forceinline__ void dft16(cpx<T> *v)
{
    cpx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        y[n2][0] = v[n2];
        y[n2][1] = v[4 + n2];
        y[n2][2] = v[8 + n2];
        y[n2][3] = v[12 + n2];
        dft4<DIR>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    y[1][1] = w16<DIR, T, 1>(y[1][1]);
    y[1][2] = w16<DIR, T, 2>(y[1][2]);
    y[1][3] = w16<DIR, T, 3>(y[1][3]);
    y[2][1] = w16<DIR, T, 2>(y[2][1]);
    y[2][2] = w16<DIR, T, 4>(y[2][2]);
    y[2][3] = w16<DIR, T, 6>(y[2][3]);
    y[3][1] = w16<DIR, T, 3>(y[3][1]);
    y[3][2] = w16<DIR, T, 6>(y[3][2]);
    y[3][3] = w16<DIR, T, 9>(y[3][3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cpx<T> a0 = y[0][k1], a1 = y[1][k1], a2 = y[2][k1], a3 = y[3][k1];
        dft4<DIR>(a0, a1, a2, a3);
        v[k1] = a0;
        v[k1 + 4] = a1;
        v[k1 + 8] = a2;
        v[k1 + 12] = a3;
    }
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cpx<T> *v)
{
    if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 16) dft16<DIR>(v);
    else static_assert(R == 2, "radix must be 2, 4, 8 or 16");
}

// LDS slice of a pencil stored contiguously (es == 1): element e lives at
// e + e/E, one pad slot per E elements.  After the first pass lane j stores
// its E outputs to consecutive elements E j .. E j + E - 1, so without the
// pad the lanes of one store sit E elements apart and land on a few banks
// (and every pencil's slice starts on the same bank); with it they are E + 1
// apart.  Pencils stored interleaved (es = pencils per workgroup) need no pad.
template <int L, int E>
constexpr int padded() { return L + L / E; }

template <int E>
__device__ __forceinline__ int slot(int e, int es) { return es == 1 ? e + e / E : e * es; }

// Radix of pass p for length L with E elements per lane.
template <int L, int E>
constexpr int npasses()
{
    int n = 0, m = L;
    while (m > 1) {
        m = (m % E == 0) ? m / E : 1;
        ++n;
    }
    return n;
}
template <int L, int E>
constexpr int radix(int p)
{
    int m = L;
    for (int i = 0; i < p; ++i) m = (m % E == 0) ? m / E : 1;
    return (m % E == 0) ? E : m;
}
template <int L, int E>
constexpr int ns_before(int p)
{
    int ns = 1;
    for (int i = 0; i < p; ++i) ns *= radix<L, E>(i);
    return ns;
}

// Work of one pass in registers: twiddles and butterflies.
template <typename T, int L, int E, int P, int DIR>
__device__ __forceinline__ void pass_compute(cpx<T> *v, int j, const cpx<T> *__restrict__ tw)
{
    constexpr int R = radix<L, E>(P), NS = ns_before<L, E>(P);
    constexpr int TPP = L / E, NBF = E / R;
#pragma unroll
    for (int s = 0; s < NBF; ++s) {
        if constexpr (NS > 1) {
            const int k = (j + s * TPP) & (NS - 1);
#pragma unroll
            for (int q = 1; q < R; ++q) {
                cpx<T> w = tw[q * k * (L / (NS * R))];
                if constexpr (DIR > 0) w.y = -w.y;
                v[s * R + q] = c_mul(v[s * R + q], w);
            }
        }
        dft<R, DIR>(v + s * R);
    }
}

// LDS exchange after pass P (all lanes of the workgroup participate).
// `lds` is this pencil's slice: element e at lds[slot<E>(e, es)] (a
// contiguous slice, es == 1, holds padded<L, E>() entries).
template <typename T, int L, int E, int P>
__device__ __forceinline__ void pass_exchange(cpx<T> *v, int j, cpx<T> *lds, int es)
{
    constexpr int R = radix<L, E>(P), NS = ns_before<L, E>(P);
    constexpr int R2 = radix<L, E>(P + 1);
    constexpr int TPP = L / E, NBF = E / R, NBF2 = E / R2;
#pragma unroll
    for (int s = 0; s < NBF; ++s) {
        const int b = j + s * TPP;
        const int base = (b / NS) * NS * R + (b & (NS - 1));
#pragma unroll
        for (int q = 0; q < R; ++q) lds[slot<E>(base + q * NS, es)] = v[s * R + q];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NBF2; ++s) {
        const int b = j + s * TPP;
#pragma unroll
        for (int q = 0; q < R2; ++q) v[s * R2 + q] = lds[slot<E>(b + q * (L / R2), es)];
    }
}

// Full transform.  v holds the input layout on entry and the output in the
// input layout on exit (x[j + r*TPP] <-> v[r]).  Every lane of the
// workgroup must call it; it contains (npasses - 1) exchanges, each with a
// barrier, plus one barrier at entry so that `lds` may be reused.
template <typename T, int L, int E, int DIR>
__device__ __forceinline__ void fft(cpx<T> *v, int j, cpx<T> *lds, int es,
                                    const cpx<T> *__restrict__ tw)
{
    constexpr int NP = npasses<L, E>();
    static_assert(is_pow2(L) && is_pow2(E) && L >= E, "power-of-two lengths only");
    __syncthreads();
    pass_compute<T, L, E, 0, DIR>(v, j, tw);
    if constexpr (NP > 1) {
        pass_exchange<T, L, E, 0>(v, j, lds, es);
        pass_compute<T, L, E, 1, DIR>(v, j, tw);
    }
    if constexpr (NP > 2) {
        __syncthreads();
        pass_exchange<T, L, E, 1>(v, j, lds, es);
        pass_compute<T, L, E, 2, DIR>(v, j, tw);
    }
    if constexpr (NP > 3) {
        __syncthreads();
        pass_exchange<T, L, E, 2>(v, j, lds, es);
        pass_compute<T, L, E, 3, DIR>(v, j, tw);
    }
    static_assert(NP <= 4, "at most four passes");
    // last pass: butterfly b = j + s*TPP holds X[b + q*L/R] in v[s*R + q];
    // put X[j + r*TPP] into v[r] (r = s + q*NBF)
    constexpr int RL = radix<L, E>(NP - 1);
    constexpr int NBF = E / RL;
    if constexpr (NBF > 1) {
        cpx<T> t[E];
#pragma unroll
        for (int s = 0; s < NBF; ++s)
#pragma unroll
            for (int q = 0; q < RL; ++q) t[s + q * NBF] = v[s * RL + q];
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = t[r];
    }
}

}  // namespace reg
}  // namespace sc

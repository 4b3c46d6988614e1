// sc_green.hip -- transformed integrated Green functions on gfx950.
//
// Reference: get_green_function! / get_green_kernel! /
// apply_8point_differencing! (src/green_functions.jl:41-112) build, per
// field component, the pointwise antiderivative F (:35-38) on the whole
// (2nx, 2ny, 2nz) node lattice, difference it in 8 points and FFT the
// result (src/solvers/free_space.jl:79-89).
//
// Here:
//   green_f  F_c in Float64 at the nodes of the non-negative octant (free
//            space; the cathode image needs the full z range) for all
//            components in one launch.
//   green_1  8-point difference in Float64 (the only place where the large
//            F values cancel), rounded to the mesh type, then the x
//            transform of the cell rows' symmetric extension (odd for G_x,
//            even otherwise), two real rows per complex transform.
//   green_2  y transform, two kx columns per complex transform.
//   green_3  z transform; free space keeps kz in [0, nz] of the symmetric
//            extension, the image takes the full transform of its 2nz-1
//            cells in circular order.  FPEI, the cell-volume factor of
//            :76-80 and the 1/(8 nx ny nz) of the backward transform are
//            applied here.
// G_x is odd in x and even in y, z (x/r^3), and cyclically for G_y, G_z, so
// every axis transform of the free-space function is a sine or cosine
// transform: the spectrum is i*R_c with R_c real on (nx+1)(ny+1)(nz+1)
// points.  The image function (offset along z) keeps the x, y symmetry and
// has a complex z spectrum.
//
// The scalar potential (GreenArgs::pot) runs the same passes on one
// component, the antiderivative of 1/r (src/green_functions.jl:13-22), whose
// cell integrals are even in every axis: its spectrum is real.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

#include "sc_fft_reg.h"
#include "sc_solve_common.h"

namespace sc {

namespace {

// Float64 log and atan of a few ulp (at most 6.9 and 2.4 ulp against long
// double over 2e7 arguments each, checked on the CPU), about half the
// instructions of the library's correctly rounded forms; used for the Green
// function of Float32 meshes, whose cell integrals need F to ~1e-14 only.
// log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
// s = (m - 1)/(m + 1), |s| <= 0.1716, the atanh series to s^17; ln 2 split
// so that e * hi is exact.
__device__ __forceinline__ double fast_log(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m += m;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 17.0;
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double t = s + s;
    const double de = (double)e;
    return fma(de, 0.6931471526622772, fma(t * z, p, fma(de, 2.7897668087737545e-08, t)));
}

// atan(num / den), one quotient for the range reduction:
// u = min(|num|, |den|) / max(|num|, |den|) (|num| > |den| ->
// pi/2 - atan(u)); u > tan(pi/12) -> pi/6 + atan((sqrt(3) u - 1)/(u + sqrt(3))),
// then |u| <= tan(pi/12) and the odd Taylor series to u^25.  The sign is that
// of num / den (a zero den gives +-pi/2); 0/0 gives NaN, as atan(0/0).
__device__ __forceinline__ double fast_atan_q(double num, double den)
{
    const double an = fabs(num), ad = fabs(den);
    const bool inv = an > ad;
    double u = (inv ? ad : an) / (inv ? an : ad);
    const bool shift = u > 0.2679491924311227;
    if (shift) u = fma(u, 1.7320508075688772, -1.0) / (u + 1.7320508075688772);
    const double z = u * u;
    double p = 1.0 / 25.0;
    p = fma(p, z, -1.0 / 23.0);
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, -1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, -1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, -1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, -1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, -1.0 / 3.0);
    double r = fma(u * z, p, u);
    if (shift) r += 0.5235987755982989;
    if (inv) r = 1.5707963267948966 - r;
    return (signbit(num) != signbit(den)) ? -r : r;
}

template <bool FAST>
__device__ __forceinline__ double log_(double x)
{
    if constexpr (FAST) return fast_log(x);
    else return log(x);
}

// atan(num / den)
template <bool FAST>
__device__ __forceinline__ double atan_q(double num, double den)
{
    if constexpr (FAST) return fast_atan_q(num, den);
    else return atan(num / den);
}

// src/green_functions.jl:35-38, in Float64.
template <bool FAST>
__device__ __forceinline__ double field_green(double x, double y, double z)
{
    const double r = sqrt((x * x + y * y) + z * z);
    const double t1 = x * atan_q<FAST>(y * z, r * x);
    const double t2 = z * log_<FAST>(r + y);
    const double t3 = (y * log_<FAST>((r - z) / (r + z))) * 0.5;
    return (t1 - t2) + t3;
}

// src/green_functions.jl:13-22, in Float64 (zero at the origin, which is
// never a lattice node: node coordinates are odd multiples of half a cell).
template <bool FAST>
__device__ __forceinline__ double potential_green(double x, double y, double z)
{
    const double r = sqrt((x * x + y * y) + z * z);
    if (r == 0.0) return 0.0;
    double t = -0.5 * (z * z) * atan_q<FAST>(x * y, z * r);
    t = t - 0.5 * (y * y) * atan_q<FAST>(x * z, y * r);
    t = t - 0.5 * (x * x) * atan_q<FAST>(y * z, x * r);
    t = t + (y * z) * log_<FAST>(x + r);
    t = t + (x * z) * log_<FAST>(y + r);
    t = t + (x * y) * log_<FAST>(z + r);
    return t;
}

// Component permutation of get_green_kernel! (src/green_functions.jl:90-98);
// c == 3 is the scalar potential.
template <bool FAST>
__device__ __forceinline__ double field_green_comp(int c, double u, double v, double w)
{
    if (c == 0) return field_green<FAST>(u, v, w);
    if (c == 1) return field_green<FAST>(v, w, u);
    if (c == 2) return field_green<FAST>(w, u, v);
    return potential_green<FAST>(u, v, w);
}

struct GreenArgs {
    Dims d;
    double dx, dy, dz;  // dz = gamma * delta_z (green_functions.jl:73)
    double offzg;       // image z offset times gamma (:84)
    double scale[4];    // FPEI * factor_c / (8 nx ny nz); [3] for the potential
    int pot;            // 1: one component, the scalar potential (c = 3)
    // device-resident mesh geometry (lo at 0, hi at 3, delta at 6): when set,
    // the spacings, the image offset and the scales above are formed from it
    // in the kernels, with the same operations as build_green
    const double *geom;
    double gamma;
    int64_t f_free, f_img;  // F elements per component
    int64_t r_free, r_img;  // R elements per component
};

// z node planes / z cell rows of component set `image`
__device__ __forceinline__ int z_nodes(const GreenArgs &a, int image) { return image ? a.d.Lz : a.d.nz + 1; }
__device__ __forceinline__ int z_cells(const GreenArgs &a, int image) { return image ? a.d.Lz - 1 : a.d.nz; }
__device__ __forceinline__ int64_t f_off(const GreenArgs &a, int q)
{
    return q < 3 ? q * a.f_free : 3 * a.f_free + (q - 3) * a.f_img;
}
__device__ __forceinline__ int64_t r_off(const GreenArgs &a, int q)
{
    return q < 3 ? q * a.r_free : 3 * a.r_free + (q - 3) * a.r_img;
}

// Unpack two real sequences transformed together as A + iB whose extensions
// share one parity: even -> both transforms real; odd -> both i*real, and
// the stored value is the coefficient of i.
template <typename T>
__device__ __forceinline__ void unpack_pair(cpx<T> X, bool odd, T &ra, T &rb)
{
    if (odd) { ra = X.y; rb = -X.x; }
    else { ra = X.x; rb = X.y; }
}

// Spacings (dz stretched by gamma) and the image offset times gamma.
__device__ __forceinline__ void green_geometry(const GreenArgs &a, double &dx, double &dy, double &dz,
                                               double &offzg)
{
    if (a.geom) {
        dx = a.geom[6];
        dy = a.geom[7];
        dz = a.geom[8] * a.gamma;
        offzg = (a.geom[2] + a.geom[5]) * a.gamma;
    } else {
        dx = a.dx;
        dy = a.dy;
        dz = a.dz;
        offzg = a.offzg;
    }
}

// FPEI * factor_c / (8 nx ny nz) of component c (3: the potential).
__device__ __forceinline__ double green_scale(const GreenArgs &a, int c)
{
    // constant indices: a runtime index into the argument copy (which the
    // grouped launches overwrite) would place it in scratch memory
    if (!a.geom) return c == 0 ? a.scale[0] : c == 1 ? a.scale[1] : c == 2 ? a.scale[2] : a.scale[3];
    double dx, dy, dz, offzg;
    green_geometry(a, dx, dy, dz, offzg);
    const double vol = (dx * dy) * dz;
    const double P = (double)a.d.Lx * (double)a.d.Ly * (double)a.d.Lz;
    return kFPEI * ((c != 2 ? a.gamma : 1.0) / vol) / P;
}

// A grouped launch's bunch: its spacings, image offset and scales.
__device__ __forceinline__ void green_from_table(GreenArgs &a, const SBunch *sb)
{
    a.dx = sb->dx;
    a.dy = sb->dy;
    a.dz = sb->dz;
    a.offzg = sb->offzg;
#pragma unroll
    for (int c = 0; c < 4; ++c) a.scale[c] = sb->scale[c];
}

// The function computed by component set q: field component q % 3 (image
// set q / 3), or the potential (c = 3, odd in no axis).
__device__ __forceinline__ int comp_of(const GreenArgs &a, int q) { return a.pot ? 3 : q % 3; }

// grid (x-y plane blocks, z node planes, components); in-plane index math
// in 32 bits (this kernel is bound by its Float64 instruction count).
// FAST: the few-ulp log and atan (Green functions of Float32 meshes).
template <bool FAST>
__global__ void __launch_bounds__(256)
k_green_f(double *__restrict__ F, GreenArgs a, const SBunch *__restrict__ tab, int per)
{
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) {
        F = sb->F;
        green_from_table(a, sb);
    }
    const int q = blockIdx.z, c = comp_of(a, q), image = q / 3;
    const int nx1 = a.d.nx + 1, ny1 = a.d.ny + 1;
    const int kn = blockIdx.y;
    if (kn >= z_nodes(a, image)) return;
    const int pidx = bx * blockDim.x + threadIdx.x;
    if (pidx >= nx1 * ny1) return;
    const int in = pidx % nx1;
    const int jn = pidx / nx1;
    const int64_t idx = (int64_t)kn * nx1 * ny1 + pidx;
    double dx, dy, dz, offzg;
    green_geometry(a, dx, dy, dz, offzg);
    const double u = ((double)in - 0.5) * dx;
    const double v = ((double)jn - 0.5) * dy;
    const double w = image ? ((double)(kn - a.d.nz) + 0.5) * dz + offzg : ((double)kn - 0.5) * dz;
    gst(F + f_off(a, q) + idx, field_green_comp<FAST>(c, u, v, w));
}

// The three field components of a node from shared terms.
// With r = |(u, v, w)| and F_x = f(u, v, w), F_y = f(v, w, u),
// F_z = f(w, u, v) (field_green), each component's quotient log follows
// from the logs of r + u, r + v, r + w and of sums of two squares, since
// (r - w)(r + w) = u^2 + v^2:
//   log((r - w) / (r + w)) = log(u^2 + v^2) - 2 log(r + w)   (and cyclically),
// which also avoids the cancellation in r - w.  The three atan terms satisfy
//   atan(vw/(ru)) + atan(wu/(rv)) + atan(uv/(rw)) = sign(uvw) pi/2
// (the tangent of the first two summed is rw/(uv), and their product
// w^2/r^2 < 1), so the third is formed from the other two.  A lane takes a
// node (x, y) of kF3Planes consecutive z planes, so log(u^2 + v^2) is formed
// once per lane; log(v^2 + w^2) depends on the node's y row and z plane
// only, and log(w^2 + u^2) on its x column and z plane, so the workgroup
// forms them once per row and column of its 256 nodes (tables in LDS; the
// column table for planes of at most 256 columns).  Per node this is
// 1 sqrt, 2 atan and about 3.8 logs on a 129^2 plane against 3 sqrt, 3 atan,
// 6 logs and 3 quotients for the three components taken one by one.  FAST:
// the few-ulp log and atan (Float32 meshes); otherwise the library's
// (Float64 meshes).
// Grid (x-y plane blocks, z plane groups, image sets).
constexpr int kF3Planes = 4;
constexpr int kF3Rows = 128;  // y rows of one workgroup's 256 nodes (rows of >= 3 nodes)
constexpr int kF3Cols = 256;  // x columns of the column table

template <bool FAST>
__global__ void __launch_bounds__(256)
k_green_f3(double *__restrict__ F, GreenArgs a, const SBunch *__restrict__ tab, int per)
{
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) {
        F = sb->F;
        green_from_table(a, sb);
    }
    const int image = blockIdx.z;
    const int nx1 = a.d.nx + 1, ny1 = a.d.ny + 1;
    const int pidx = bx * blockDim.x + threadIdx.x;
    const int np = nx1 * ny1;
    double dx, dy, dz, offzg;
    green_geometry(a, dx, dy, dz, offzg);
    const int zn = z_nodes(a, image);
    const int k0 = blockIdx.y * kF3Planes;
    auto coord_u = [&](int i) { return ((double)i - 0.5) * dx; };
    auto coord_v = [&](int j) { return ((double)j - 0.5) * dy; };
    auto coord_w = [&](int k) {
        return image ? ((double)(k - a.d.nz) + 0.5) * dz + offzg : ((double)k - 0.5) * dz;
    };
    // the row and column logs of this workgroup's nodes (every lane takes
    // part, live or not: the tables end with a barrier)
    __shared__ double qvw_t[kF3Planes][kF3Rows];
    __shared__ double qwu_t[kF3Planes][kF3Cols];
    const int p0 = bx * blockDim.x;
    const int p1 = (p0 + (int)blockDim.x < np ? p0 + (int)blockDim.x : np) - 1;
    const int jn0 = p0 / nx1;
    const int nrows = p0 < np ? p1 / nx1 - jn0 + 1 : 0;  // <= 255 / nx1 + 2 <= kF3Rows
    const bool coltab = nx1 <= kF3Cols;
    for (int t = threadIdx.x; t < kF3Planes * nrows; t += blockDim.x) {
        const int kk = t / nrows, jr = t - kk * nrows;
        const double v = coord_v(jn0 + jr), w = coord_w(k0 + kk);
        const double vv = v * v, ww = w * w;
        qvw_t[kk][jr] = log_<FAST>(vv + ww);
    }
    if (coltab) {
        for (int t = threadIdx.x; t < kF3Planes * nx1; t += blockDim.x) {
            const int kk = t / nx1, i = t - kk * nx1;
            const double u = coord_u(i), w = coord_w(k0 + kk);
            const double uu = u * u, ww = w * w;
            qwu_t[kk][i] = log_<FAST>(ww + uu);
        }
    }
    __syncthreads();
    if (pidx >= np) return;
    const int in = pidx % nx1;
    const int jn = pidx / nx1;
    const double u = coord_u(in);
    const double v = coord_v(jn);
    const double uu = u * u, vv = v * v;
    const double Quv = log_<FAST>(uu + vv);
    const int64_t plane = (int64_t)nx1 * ny1;
    gbl<double> *Fx = (gbl<double> *)(F + f_off(a, 3 * image));
    gbl<double> *Fy = (gbl<double> *)(F + f_off(a, 3 * image + 1));
    gbl<double> *Fz = (gbl<double> *)(F + f_off(a, 3 * image + 2));
#pragma unroll
    for (int kk = 0; kk < kF3Planes; ++kk) {
        const int kn = k0 + kk;
        if (kn >= zn) break;
        const double w = coord_w(kn);
        const double ww = w * w;
        const double r = sqrt((uu + vv) + ww);
        const double Lu = log_<FAST>(r + u), Lv = log_<FAST>(r + v), Lw = log_<FAST>(r + w);
        const double Qvw = qvw_t[kk][jn - jn0];
        const double Qwu = coltab ? qwu_t[kk][in] : log_<FAST>(ww + uu);
        const double ax = atan_q<FAST>(v * w, r * u);
        const double ay = atan_q<FAST>(w * u, r * v);
        const double hp = w == 0.0 ? 0.0
                                   : ((signbit(u) != signbit(v)) != signbit(w) ? -1.5707963267948966
                                                                               : 1.5707963267948966);
        const double az = (hp - ax) - ay;
        const int64_t idx = (int64_t)kn * plane + pidx;
        Fx[idx] = (u * ax - w * Lv) + (v * (Quv - 2.0 * Lw)) * 0.5;
        Fy[idx] = (v * ay - u * Lw) + (w * (Qvw - 2.0 * Lu)) * 0.5;
        Fz[idx] = (w * az - v * Lu) + (u * (Qwu - 2.0 * Lv)) * 0.5;
    }
}

// The 8-point difference of apply_8point_differencing!
// (src/green_functions.jl:108-111), the mixed third difference of F over the
// cell, on values f<y><z><x> (x, y, z in {0, 1} relative to the cell's lower
// corner).  It is formed as differences of differences, x neighbours first:
// neighbouring F values agree in their leading digits, so each first
// difference is exact or nearly so (Sterbenz) and the later ones subtract
// small numbers.  Summed in the reference's order instead, the partial sums
// swing between |F| and the small result and round at ulp(F) several times;
// this is the dominant Float64 error of the reference's field at large grids.
__device__ __forceinline__ double diff8v(double f000, double f001, double f010, double f011,
                                         double f100, double f101, double f110, double f111)
{
    const double g00 = f001 - f000, g01 = f011 - f010;  // along x at (y, z) = (0, 0), (0, 1)
    const double g10 = f101 - f100, g11 = f111 - f110;  // (1, 0), (1, 1)
    return (g11 - g10) - (g01 - g00);
}

// The same difference for cell (i, j, k) read from the F lattice.
__device__ __forceinline__ double diff8(const double *__restrict__ F, int i, int j, int k, int nx1,
                                        int ny1)
{
    const int64_t sy = nx1, sz = (int64_t)nx1 * ny1;
    const double *p = F + i + sy * j + sz * k;
    return diff8v(p[0], p[1], p[sz], p[1 + sz], p[sy], p[1 + sy], p[sy + sz], p[1 + sy + sz]);
}

template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_green_1(const double *__restrict__ F, T *__restrict__ R, const cpx<T> *__restrict__ twg,
          FftPlan pl, GreenArgs a, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + a.d.Lx;
    const int q = blockIdx.y, c = comp_of(a, q), image = q / 3;
    const int nx = a.d.nx, ny = a.d.ny, Lx = a.d.Lx, Kx = a.d.Kx, Kp = a.d.Kp;
    const int bpairs = (ny + 1) / 2;
    const int npencils = z_cells(a, image) * bpairs;
    const double *Fq = F + f_off(a, q);
    T *Rq = R + r_off(a, q);
    load_twiddles(tw, twg, Lx);
    const int p0 = blockIdx.x * nb;
    for (int idx = threadIdx.x; idx < nb * nx; idx += blockDim.x) {
        const int p = idx / nx, i = idx - p * nx;
        const int P = p0 + p;
        double ga = 0.0, gb = 0.0;
        if (P < npencils) {
            const int k = P / bpairs, j = 2 * (P - k * bpairs);
            ga = diff8(Fq, i, j, k, nx + 1, ny + 1);
            if (j + 1 < ny) gb = diff8(Fq, i, j + 1, k, nx + 1, ny + 1);
        }
        buf[p * Lx + i] = cpx<T>{(T)ga, (T)gb};
    }
    __syncthreads();
    const T sgn = (c == 0) ? (T)-1 : (T)1;
    for (int idx = threadIdx.x; idx < nb * nx; idx += blockDim.x) {
        const int p = idx / nx, e = nx + (idx - p * nx);
        cpx<T> v{(T)0, (T)0};
        if (e > nx) {
            const cpx<T> s = buf[p * Lx + Lx - e];
            v = cpx<T>{sgn * s.x, sgn * s.y};
        }
        buf[p * Lx + e] = v;
    }
    __syncthreads();
    lds_fft<T, E, NT>(buf, tw, pl, nb, 1, Lx, false, -1);
    for (int idx = threadIdx.x; idx < nb * Kx; idx += blockDim.x) {
        const int p = idx / Kx, kx = idx - p * Kx;
        const int P = p0 + p;
        if (P < npencils) {
            const int k = P / bpairs, j = 2 * (P - k * bpairs);
            T ra, rb;
            unpack_pair(buf[p * Lx + kx], c == 0, ra, rb);
            Rq[((size_t)k * (ny + 1) + j) * Kp + kx] = ra;
            if (j + 1 < ny) Rq[((size_t)k * (ny + 1) + j + 1) * Kp + kx] = rb;
        }
    }
}

template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_green_2(T *__restrict__ R, const cpx<T> *__restrict__ twg, FftPlan pl, GreenArgs a, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + a.d.Ly;
    const int q = blockIdx.z, c = comp_of(a, q), image = q / 3;
    const int k = blockIdx.y;
    if (k >= z_cells(a, image)) return;  // uniform over the workgroup
    const int ny = a.d.ny, Ly = a.d.Ly, Kx = a.d.Kx, Kp = a.d.Kp;
    T *Rrow = R + r_off(a, q) + (size_t)k * (ny + 1) * Kp;
    load_twiddles(tw, twg, Ly);
    const int kp0 = blockIdx.x * nb;
    for (int idx = threadIdx.x; idx < nb * ny; idx += blockDim.x) {
        const int e = idx / nb, p = idx - e * nb;
        const int kxa = 2 * (kp0 + p), kxb = kxa + 1;
        const T va = kxa < Kx ? Rrow[(size_t)e * Kp + kxa] : (T)0;
        const T vb = kxb < Kx ? Rrow[(size_t)e * Kp + kxb] : (T)0;
        buf[e * nb + p] = cpx<T>{va, vb};
    }
    __syncthreads();
    const T sgn = (c == 1) ? (T)-1 : (T)1;
    for (int idx = threadIdx.x; idx < nb * ny; idx += blockDim.x) {
        const int e = ny + idx / nb, p = idx % nb;
        cpx<T> v{(T)0, (T)0};
        if (e > ny) {
            const cpx<T> s = buf[(Ly - e) * nb + p];
            v = cpx<T>{sgn * s.x, sgn * s.y};
        }
        buf[e * nb + p] = v;
    }
    __syncthreads();
    lds_fft<T, E, NT>(buf, tw, pl, nb, nb, 1, true, -1);
    for (int idx = threadIdx.x; idx < nb * (ny + 1); idx += blockDim.x) {
        const int ky = idx / nb, p = idx - ky * nb;
        const int kxa = 2 * (kp0 + p), kxb = kxa + 1;
        T ra, rb;
        unpack_pair(buf[ky * nb + p], c == 1, ra, rb);
        if (kxa < Kx) Rrow[(size_t)ky * Kp + kxa] = ra;
        if (kxb < Kx) Rrow[(size_t)ky * Kp + kxb] = rb;
    }
}

template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_green_3(const T *__restrict__ R, T *__restrict__ Gf, cpx<T> *__restrict__ Gi,
          const cpx<T> *__restrict__ twg, FftPlan pl, GreenArgs a, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + a.d.Lz;
    const int q = blockIdx.z, c = comp_of(a, q), image = q / 3;
    const int ky = blockIdx.y;
    const int nz = a.d.nz, Lz = a.d.Lz, Kx = a.d.Kx, Kp = a.d.Kp, ny1 = a.d.ny + 1;
    const T *Rq = R + r_off(a, q);
    load_twiddles(tw, twg, Lz);
    const int kp0 = blockIdx.x * nb;
    for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) {
        const int e = idx / nb, p = idx - e * nb;
        const int kxa = 2 * (kp0 + p), kxb = kxa + 1;
        int cell = -1;  // which z cell row feeds slot e
        if (!image) {
            if (e < nz) cell = e;
        } else if (e != nz) {
            // slot e holds displacement e (e < nz) or e - Lz (e > nz);
            // cell row = displacement + nz - 1
            cell = (e < nz) ? e + nz - 1 : e - nz - 1;
        }
        cpx<T> v{(T)0, (T)0};
        if (cell >= 0) {
            const T *row = Rq + ((size_t)cell * ny1 + ky) * Kp;
            v.x = kxa < Kx ? row[kxa] : (T)0;
            v.y = kxb < Kx ? row[kxb] : (T)0;
        }
        buf[idx] = v;
    }
    __syncthreads();
    if (!image) {
        // symmetric extension along z (odd for G_z)
        const T sgn = (c == 2) ? (T)-1 : (T)1;
        for (int idx = threadIdx.x; idx < nb * nz; idx += blockDim.x) {
            const int e = nz + idx / nb, p = idx % nb;
            if (e > nz) {
                const cpx<T> s = buf[(Lz - e) * nb + p];
                buf[e * nb + p] = cpx<T>{sgn * s.x, sgn * s.y};
            }
        }
        __syncthreads();
    }
    lds_fft<T, E, NT>(buf, tw, pl, nb, nb, 1, true, -1);
    const T scale = (T)green_scale(a, c);
    if (!image) {
        T *G = Gf + (size_t)(q % 3) * (nz + 1) * ny1 * Kp;
        for (int idx = threadIdx.x; idx < nb * (nz + 1); idx += blockDim.x) {
            const int kz = idx / nb, p = idx - kz * nb;
            const int kxa = 2 * (kp0 + p), kxb = kxa + 1;
            T ra, rb;
            unpack_pair(buf[kz * nb + p], c == 2, ra, rb);
            T *row = G + ((size_t)kz * ny1 + ky) * Kp;
            if (kxa < Kx) row[kxa] = ra * scale;
            if (kxb < Kx) row[kxb] = rb * scale;
        }
    } else {
        // separate the two real sequences: A_k = (X_k + conj X_{-k})/2,
        // B_k = (X_k - conj X_{-k})/(2i)
        cpx<T> *G = Gi + (size_t)(q % 3) * Lz * ny1 * Kp;
        for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) {
            const int kz = idx / nb, p = idx - kz * nb;
            const int kxa = 2 * (kp0 + p), kxb = kxa + 1;
            const cpx<T> X = buf[kz * nb + p];
            const cpx<T> Y = buf[((Lz - kz) % Lz) * nb + p];
            const T h = (T)0.5 * scale;
            cpx<T> *row = G + ((size_t)kz * ny1 + ky) * Kp;
            if (kxa < Kx) row[kxa] = cpx<T>{(X.x + Y.x) * h, (X.y - Y.y) * h};
            if (kxb < Kx) row[kxb] = cpx<T>{(X.y + Y.y) * h, (Y.x - X.x) * h};
        }
    }
}

// ---- register-resident versions for power-of-two lengths (sc_fft_reg.h) ----
constexpr int kRThreads = 256;

template <typename T, int L>
struct RShape {
    static constexpr int E = reg::elems<T>() < L ? reg::elems<T>() : L;
    static constexpr int TPP = L / E;
    static constexpr int NB = kRThreads / TPP;
};

template <typename T, int L>
__global__ void __launch_bounds__(kRThreads)
k_green_1r(const double *__restrict__ F, T *__restrict__ R, const cpx<T> *__restrict__ twg,
           GreenArgs a, const SBunch *__restrict__ tab, int per)
{
    using S = RShape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * reg::padded<L, E>()];
    // one bunch: workgroups in XCD order over the whole grid (pencils of
    // plane k + 1 read the F rows of plane k's pencils a few workgroups
    // earlier, which then sit on the same XCD and its L2)
    const SBunch *sb = nullptr;
    int bx, q;
    if (tab) {
        bx = group_block(tab, per, sb);
        q = (int)blockIdx.y;
    } else {
        const int task = xcd_task((int)(blockIdx.x + blockIdx.y * gridDim.x),
                                  (int)(gridDim.x * gridDim.y));
        bx = task % (int)gridDim.x;
        q = task / (int)gridDim.x;
    }
    if (sb) {
        F = sb->F;
        R = static_cast<T *>(sb->R);
    }
    for (int i = threadIdx.x; i < L; i += kRThreads) tw[i] = gld(twg + i);
    const int c = comp_of(a, q), image = q / 3;
    const int nx = a.d.nx, ny = a.d.ny, Kx = a.d.Kx, Kp = a.d.Kp;
    const int bpairs = (ny + 1) / 2;
    const int npencils = z_cells(a, image) * bpairs;
    const int b = threadIdx.x / TPP, j = threadIdx.x % TPP;
    const int P = bx * NB + b;
    const bool live = P < npencils;
    const int k = live ? P / bpairs : 0, ya = live ? 2 * (P - k * bpairs) : 0;
    const bool hasb = ya + 1 < ny;
    const gbl<double> *Fq = (const gbl<double> *)(F + f_off(a, q));
    const T sgn = (c == 0) ? (T)-1 : (T)1;
    // The extension's upper half (e > nx) mirrors cells nx - 1 .. 1, so each
    // cell row is differenced once (e < nx, the first E/2 registers when
    // L = 2 nx) and the mirror is read back through this pencil's LDS slice.
    cpx<T> *row = lds + b * reg::padded<L, E>();
    cpx<T> v[E];
    // F rows ya, ya+1, ya+2 (clamped: without a second row the third is
    // unused) of planes k, k+1, loaded without branches two cells at a time
    // so that their 24 loads are in flight together
    const int64_t sy = nx + 1, sz = (int64_t)(nx + 1) * (ny + 1);
    const int yr[3] = {ya, ya + 1, ya + 2 <= ny ? ya + 2 : ny};
    constexpr int kCells = 2;
    static_assert((E / 2) % kCells == 0, "cells per batch");
#pragma unroll
    for (int r0 = 0; r0 < E / 2; r0 += kCells) {
        double f[kCells][3][2][2];  // [cell][y][z][x]
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const int e = j + (r0 + u) * TPP;
#pragma unroll
            for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                for (int zz = 0; zz < 2; ++zz) {
                    const gbl<double> *p = Fq + e + sy * yr[yy] + sz * (k + zz);
                    f[u][yy][zz][0] = p[0];
                    f[u][yy][zz][1] = p[1];
                }
        }
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const int r = r0 + u;
            const double ga = diff8v(f[u][0][0][0], f[u][0][0][1], f[u][0][1][0], f[u][0][1][1],
                                     f[u][1][0][0], f[u][1][0][1], f[u][1][1][0], f[u][1][1][1]);
            const double gb = diff8v(f[u][1][0][0], f[u][1][0][1], f[u][1][1][0], f[u][1][1][1],
                                     f[u][2][0][0], f[u][2][0][1], f[u][2][1][0], f[u][2][1][1]);
            v[r].x = live ? (T)ga : (T)0;
            v[r].y = (live && hasb) ? (T)gb : (T)0;
            row[j + r * TPP] = v[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = E / 2; r < E; ++r) {
        const int e = j + r * TPP;
        const cpx<T> m = row[(L - e) & (L - 1)];  // e == nx reads row[nx], unused
        v[r] = e == nx ? cpx<T>{(T)0, (T)0} : cpx<T>{sgn * m.x, sgn * m.y};
    }
    reg::fft<T, L, E, -1>(v, j, row, 1, tw);
    if (live) {
        gbl<T> *Rq = (gbl<T> *)(R + r_off(a, q));
#pragma unroll
        for (int r = 0; r <= E / 2; ++r) {
            const int kx = j + r * TPP;
            if (r < E / 2 || kx == nx) {
                T ra, rb;
                unpack_pair(v[r], c == 0, ra, rb);
                Rq[((size_t)k * (ny + 1) + ya) * Kp + kx] = ra;
                if (hasb) Rq[((size_t)k * (ny + 1) + ya + 1) * Kp + kx] = rb;
            }
        }
    }
}

template <typename T, int L>
__global__ void __launch_bounds__(kRThreads)
k_green_2r(T *__restrict__ R, const cpx<T> *__restrict__ twg, GreenArgs a,
           const SBunch *__restrict__ tab, int per)
{
    using S = RShape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * L];
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) R = static_cast<T *>(sb->R);
    const int q = blockIdx.z, c = comp_of(a, q), image = q / 3;
    const int k = blockIdx.y;
    if (k >= z_cells(a, image)) return;  // uniform over the workgroup
    for (int i = threadIdx.x; i < L; i += kRThreads) tw[i] = gld(twg + i);
    const int ny = a.d.ny, Kx = a.d.Kx, Kp = a.d.Kp;
    const int b = threadIdx.x % NB, j = threadIdx.x / NB;
    const int kxa = 2 * (bx * NB + b), kxb = kxa + 1;
    gbl<T> *Rrow = (gbl<T> *)(R + r_off(a, q) + (size_t)k * (ny + 1) * Kp);
    const T sgn = (c == 1) ? (T)-1 : (T)1;
    // every load is in range (clamped row and column) and issued without a
    // branch, so a lane's loads are all in flight together; the selects
    // below apply the extension and the bounds
    const int kxa_c = kxa < Kx ? kxa : Kx - 1, kxb_c = kxb < Kx ? kxb : Kx - 1;
    T ra[E], rb[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int e = j + r * TPP;
        const int y = e < ny ? e : (e == ny ? 0 : L - e);
        ra[r] = Rrow[(size_t)y * Kp + kxa_c];
        rb[r] = Rrow[(size_t)y * Kp + kxb_c];
    }
    cpx<T> v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int e = j + r * TPP;
        const T s = e < ny ? (T)1 : sgn;
        v[r].x = (kxa < Kx && e != ny) ? s * ra[r] : (T)0;
        v[r].y = (kxb < Kx && e != ny) ? s * rb[r] : (T)0;
    }
    reg::fft<T, L, E, -1>(v, j, lds + b, NB, tw);
#pragma unroll
    for (int r = 0; r <= E / 2; ++r) {
        const int ky = j + r * TPP;
        if (r < E / 2 || ky == ny) {
            T ra, rb;
            unpack_pair(v[r], c == 1, ra, rb);
            if (kxa < Kx) Rrow[(size_t)ky * Kp + kxa] = ra;
            if (kxb < Kx) Rrow[(size_t)ky * Kp + kxb] = rb;
        }
    }
}

template <typename T, int L>
__global__ void __launch_bounds__(kRThreads)
k_green_3r(const T *__restrict__ R, T *__restrict__ Gf, cpx<T> *__restrict__ Gi,
           const cpx<T> *__restrict__ twg, GreenArgs a, const SBunch *__restrict__ tab, int per)
{
    using S = RShape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * L];
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) {
        R = static_cast<const T *>(sb->R);
        Gf = static_cast<T *>(sb->Gf);
        Gi = static_cast<cpx<T> *>(sb->Gi);
        green_from_table(a, sb);
    }
    for (int i = threadIdx.x; i < L; i += kRThreads) tw[i] = gld(twg + i);
    const int q = blockIdx.z, c = comp_of(a, q), image = q / 3;
    const int ky = blockIdx.y;
    const int nz = a.d.nz, Kx = a.d.Kx, Kp = a.d.Kp, ny1 = a.d.ny + 1;
    const int b = threadIdx.x % NB, j = threadIdx.x / NB;
    const int kxa = 2 * (bx * NB + b), kxb = kxa + 1;
    const gbl<T> *Rq = (const gbl<T> *)(R + r_off(a, q));
    const T sgn = (c == 2) ? (T)-1 : (T)1;
    // loads clamped in range and issued without branches (see k_green_2r)
    const int kxa_c = kxa < Kx ? kxa : Kx - 1, kxb_c = kxb < Kx ? kxb : Kx - 1;
    T ra[E], rb[E], sc[E];
    bool on[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int e = j + r * TPP;
        int cell = 0;
        T s = (T)1;
        on[r] = e != nz;  // slot e == nz is zero
        if (!image) {
            if (e != nz) {
                cell = e < nz ? e : L - e;
                s = e < nz ? (T)1 : sgn;
            }
        } else if (e != nz) {
            // slot e holds displacement e (e < nz) or e - L (e > nz); row = displacement + nz - 1
            cell = e < nz ? e + nz - 1 : e - nz - 1;
        }
        const gbl<T> *row = Rq + ((size_t)cell * ny1 + ky) * Kp;
        ra[r] = row[kxa_c];
        rb[r] = row[kxb_c];
        sc[r] = s;
    }
    cpx<T> v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        v[r].x = (kxa < Kx && on[r]) ? sc[r] * ra[r] : (T)0;
        v[r].y = (kxb < Kx && on[r]) ? sc[r] * rb[r] : (T)0;
    }
    reg::fft<T, L, E, -1>(v, j, lds + b, NB, tw);
    const T scale = (T)green_scale(a, c);
    if (!image) {
        gbl<T> *G = (gbl<T> *)(Gf + (size_t)(q % 3) * (nz + 1) * ny1 * Kp);
#pragma unroll
        for (int r = 0; r <= E / 2; ++r) {
            const int kz = j + r * TPP;
            if (r < E / 2 || kz == nz) {
                T ra, rb;
                unpack_pair(v[r], c == 2, ra, rb);
                gbl<T> *row = G + ((size_t)kz * ny1 + ky) * Kp;
                if (kxa < Kx) row[kxa] = ra * scale;
                if (kxb < Kx) row[kxb] = rb * scale;
            }
        }
    } else {
        // the separation needs X[-kz], held by another lane: go through LDS
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) lds[b + (j + r * TPP) * NB] = v[r];
        __syncthreads();
        cpx<T> *G = Gi + (size_t)(q % 3) * L * ny1 * Kp;
        const T h = (T)0.5 * scale;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const int kz = j + r * TPP;
            const cpx<T> X = v[r];
            const cpx<T> Y = lds[b + ((L - kz) & (L - 1)) * NB];
            cpx<T> *row = G + ((size_t)kz * ny1 + ky) * Kp;
            if (kxa < Kx) gst(row + kxa, cpx<T>{(X.x + Y.x) * h, (X.y - Y.y) * h});
            if (kxb < Kx) gst(row + kxb, cpx<T>{(X.y + Y.y) * h, (Y.x - X.x) * h});
        }
    }
}

template <typename T, int L>
void green_pass_r(int pass, dim3 grid, hipStream_t st, const Workspace &ws, const GreenArgs &a,
                  const cpx<T> *tw)
{
    T *R = static_cast<T *>(ws.R);
    if (pass == 1)
        hipLaunchKernelGGL((k_green_1r<T, L>), grid, dim3(kRThreads), 0, st, (const double *)ws.F, R,
                           tw, a, (const SBunch *)nullptr, 0);
    else if (pass == 2)
        hipLaunchKernelGGL((k_green_2r<T, L>), grid, dim3(kRThreads), 0, st, R, tw, a,
                           (const SBunch *)nullptr, 0);
    else
        hipLaunchKernelGGL((k_green_3r<T, L>), grid, dim3(kRThreads), 0, st, (const T *)R,
                           static_cast<T *>(ws.Gf), static_cast<cpx<T> *>(ws.Gi), tw, a,
                           (const SBunch *)nullptr, 0);
}

// One register-engine Green pass if its length is supported; false otherwise.
template <typename T>
bool green_pass_fast(int pass, int L, int ncomp, const Dims &d, int cath, hipStream_t st,
                     const Workspace &ws, const GreenArgs &a, const cpx<T> *tw)
{
    auto grid_for = [&](int nb) -> dim3 {
        const int zc = cath ? d.Lz - 1 : d.nz;
        const int kp = (d.Kx + 1) / 2;
        if (pass == 1) return dim3(div_up(zc * ((d.ny + 1) / 2), nb), ncomp);
        if (pass == 2) return dim3(div_up(kp, nb), zc, ncomp);
        return dim3(div_up(kp, nb), d.ny + 1, ncomp);
    };
    switch (L) {
#define SC_GREEN_CASE(LL)                                                                   \
    case LL:                                                                                \
        green_pass_r<T, LL>(pass, grid_for(RShape<T, LL>::NB), st, ws, a, tw);             \
        return true;
        SC_GREEN_CASE(32)
        SC_GREEN_CASE(64)
        SC_GREEN_CASE(128)
        SC_GREEN_CASE(256)
        SC_GREEN_CASE(512)
#undef SC_GREEN_CASE
    default: return false;
    }
}

}  // namespace

template <typename T>
int build_green(Mesh *m, const double dl[3], double gamma, double offz, int mode)
{
    const int cath = mode == kFieldCathode;
    // without host spacings the kernels read the mesh's device-resident geometry
    static const double unit[3] = {1.0, 1.0, 1.0};
    const double *dlh = dl ? dl : unit;
    Ctx *ctx = m->ctx;
    Workspace &ws = m->ws;
    hipStream_t st = ctx->stream;
    GreenArgs a;
    a.d = make_dims(m);
    const Dims &d = a.d;
    a.dx = dlh[0];
    a.dy = dlh[1];
    a.dz = dlh[2] * gamma;
    a.offzg = offz * gamma;
    a.geom = dl ? nullptr : m->geom;
    a.gamma = gamma;
    const double vol = (a.dx * a.dy) * a.dz;
    const double P = (double)d.Lx * (double)d.Ly * (double)d.Lz;
    for (int c = 0; c < 3; ++c) a.scale[c] = kFPEI * ((c < 2 ? gamma : 1.0) / vol) / P;
    // the lab-frame potential gamma * phi_rest takes the transverse factor
    a.scale[3] = kFPEI * (gamma / vol) / P;
    a.pot = mode == kPotential;
    a.f_free = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    a.f_img = (int64_t)(d.nx + 1) * (d.ny + 1) * d.Lz;
    a.r_free = (int64_t)d.nz * (d.ny + 1) * d.Kp;
    a.r_img = (int64_t)(d.Lz - 1) * (d.ny + 1) * d.Kp;
    const int ncomp = a.pot ? 1 : (cath ? 6 : 3);
    const cpx<T> *twx = static_cast<const cpx<T> *>(ws.twT[0]);
    const cpx<T> *twy = static_cast<const cpx<T> *>(ws.twT[1]);
    const cpx<T> *twz = static_cast<const cpx<T> *>(ws.twT[2]);
    T *R = static_cast<T *>(ws.R);
    {
        const int plane = (d.nx + 1) * (d.ny + 1);
        const int znodes = cath ? d.Lz : d.nz + 1;
        launch_begin(ctx, K_GREEN_F);
        if (!a.pot)
            hipLaunchKernelGGL(k_green_f3<sizeof(T) == 4>,
                               dim3((unsigned)div_up(plane, 256), div_up(znodes, kF3Planes), cath ? 2 : 1),
                               dim3(256), 0, st, ws.F, a, (const SBunch *)nullptr, 0);
        else
            hipLaunchKernelGGL(k_green_f<sizeof(T) == 4>,
                               dim3((unsigned)div_up(plane, 256), znodes, ncomp), dim3(256), 0, st,
                               ws.F, a, (const SBunch *)nullptr, 0);
        launch_end(ctx, K_GREEN_F);
    }
    const bool fast_ok = !m->generic_transforms;
    launch_begin(ctx, K_GREEN_X);
    if (!(fast_ok && green_pass_fast<T>(1, d.Lx, ncomp, d, cath, st, ws, a, twx))) {
        with_fft_elems<T>(d.Lx, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Lx, E * NT, 16);
            const int zc = cath ? d.Lz - 1 : d.nz;
            const int np = zc * ((d.ny + 1) / 2);
            const size_t sh = sizeof(cpx<T>) * (d.Lx + (size_t)nb * d.Lx);
            hipLaunchKernelGGL((k_green_1<T, E, NT>), dim3(div_up(np, nb), ncomp), dim3(NT), sh,
                               st, (const double *)ws.F, R, twx, ws.plan[0], a, nb);
        });
    }
    launch_end(ctx, K_GREEN_X);
    const int kp = (d.Kx + 1) / 2;
    launch_begin(ctx, K_GREEN_Y);
    if (!(fast_ok && green_pass_fast<T>(2, d.Ly, ncomp, d, cath, st, ws, a, twy))) {
        with_fft_elems<T>(d.Ly, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Ly, E * NT, 16);
            const int zc = cath ? d.Lz - 1 : d.nz;
            const size_t sh = sizeof(cpx<T>) * (d.Ly + (size_t)nb * d.Ly);
            hipLaunchKernelGGL((k_green_2<T, E, NT>), dim3(div_up(kp, nb), zc, ncomp), dim3(NT),
                               sh, st, R, twy, ws.plan[1], a, nb);
        });
    }
    launch_end(ctx, K_GREEN_Y);
    launch_begin(ctx, K_GREEN_Z);
    if (!(fast_ok && green_pass_fast<T>(3, d.Lz, ncomp, d, cath, st, ws, a, twz))) {
        with_fft_elems<T>(d.Lz, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Lz, E * NT, 16);
            const size_t sh = sizeof(cpx<T>) * (d.Lz + (size_t)nb * d.Lz);
            hipLaunchKernelGGL((k_green_3<T, E, NT>), dim3(div_up(kp, nb), d.ny + 1, ncomp),
                               dim3(NT), sh, st, (const T *)R, static_cast<T *>(ws.Gf),
                               static_cast<cpx<T> *>(ws.Gi), twz, ws.plan[2], a, nb);
        });
    }
    launch_end(ctx, K_GREEN_Z);
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

// The Green build of nb same-shape bunches, each kernel once for all of
// them (x stacked); field modes only, register engine lengths only.
template <typename T>
void build_green_grouped(const Mesh *m0, int nb, const SBunch *tab, double gamma, int mode,
                         hipStream_t st)
{
    const int cath = mode == kFieldCathode;
    GreenArgs a;
    std::memset(&a, 0, sizeof(a));
    a.d = make_dims(m0);
    const Dims &d = a.d;
    a.pot = 0;
    a.geom = nullptr;
    a.gamma = gamma;
    a.f_free = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    a.f_img = (int64_t)(d.nx + 1) * (d.ny + 1) * d.Lz;
    a.r_free = (int64_t)d.nz * (d.ny + 1) * d.Kp;
    a.r_img = (int64_t)(d.Lz - 1) * (d.ny + 1) * d.Kp;
    const int ncomp = cath ? 6 : 3;
    const Workspace &ws = m0->ws;
    const cpx<T> *tw[3] = {static_cast<const cpx<T> *>(ws.twT[0]),
                           static_cast<const cpx<T> *>(ws.twT[1]),
                           static_cast<const cpx<T> *>(ws.twT[2])};
    {
        const int per = div_up((d.nx + 1) * (d.ny + 1), 256);
        const int znodes = cath ? d.Lz : d.nz + 1;
        hipLaunchKernelGGL(k_green_f3<sizeof(T) == 4>,
                           dim3((unsigned)(per * nb), div_up(znodes, kF3Planes), cath ? 2 : 1),
                           dim3(256), 0, st, (double *)nullptr, a, tab, per);
    }
    const int zc = cath ? d.Lz - 1 : d.nz;
    const int kp = (d.Kx + 1) / 2;
    for (int pass = 1; pass <= 3; ++pass) {
        const int L = pass == 1 ? d.Lx : (pass == 2 ? d.Ly : d.Lz);
        auto go = [&](auto lc) {
            constexpr int LL = decltype(lc)::value;
            constexpr int NB = RShape<T, LL>::NB;
            if (pass == 1) {
                const int per = div_up(zc * ((d.ny + 1) / 2), NB);
                hipLaunchKernelGGL((k_green_1r<T, LL>), dim3(per * nb, ncomp), dim3(kRThreads), 0, st,
                                   (const double *)nullptr, (T *)nullptr, tw[0], a, tab, per);
            } else if (pass == 2) {
                const int per = div_up(kp, NB);
                hipLaunchKernelGGL((k_green_2r<T, LL>), dim3(per * nb, zc, ncomp), dim3(kRThreads), 0,
                                   st, (T *)nullptr, tw[1], a, tab, per);
            } else {
                const int per = div_up(kp, NB);
                hipLaunchKernelGGL((k_green_3r<T, LL>), dim3(per * nb, d.ny + 1, ncomp),
                                   dim3(kRThreads), 0, st, (const T *)nullptr, (T *)nullptr,
                                   (cpx<T> *)nullptr, tw[2], a, tab, per);
            }
        };
        switch (L) {
        case 32: go(std::integral_constant<int, 32>{}); break;
        case 64: go(std::integral_constant<int, 64>{}); break;
        case 128: go(std::integral_constant<int, 128>{}); break;
        case 256: go(std::integral_constant<int, 256>{}); break;
        case 512: go(std::integral_constant<int, 512>{}); break;
        default: break;  // the caller groups only supported lengths
        }
    }
}

template void build_green_grouped<float>(const Mesh *, int, const SBunch *, double, int, hipStream_t);
template void build_green_grouped<double>(const Mesh *, int, const SBunch *, double, int,
                                          hipStream_t);

template int build_green<float>(Mesh *, const double[3], double, double, int);
template int build_green<double>(Mesh *, const double[3], double, double, int);

}  // namespace sc

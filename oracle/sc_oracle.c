/*
 * sc_oracle.c -- CPU oracle for the space-charge step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * SpaceCharge.jl v1.2.0 CPU path (mesh bounds, CIC deposit, integrated Green
 * function, zero-padded 2N FFT convolution, cathode image solve, trilinear
 * interpolation).  It is the checker the tests compare the HIP library
 * against, and the CPU baseline the bench times.  Nothing in the product
 * (the HIP library or its Python host layer) links, loads or calls it.
 *
 * Parity pins (each checked by tests/test_oracle.py):
 *   - FFT: numpy.fft on random data, incl. radix-3/5/7 and prime lengths.
 *   - field_green_function: 40-digit mpmath values and the identity
 *     d3F/dxdydz = x/r^3.
 *   - IGF cell values: mpmath/Gauss-Legendre cubature of x/r^3 over cells.
 *   - solve: brute-force direct sum with independently integrated cells,
 *     the analytical isotropic Gaussian (test/analytical_test.jl), and the
 *     cathode-as-mirrored-free-space identity.
 *   - deposit/interpolate: hand-computed CIC weights, charge conservation,
 *     constant/linear-field exactness and the deposit/interp adjoint identity.
 *   - bounds: the closed form of src/mesh.jl:95-174 and test/test_mesh.jl.
 *
 * Deviation from the reference, shared with the HIP path: a particle whose
 * lower cell index floor(norm) falls outside [0, n-2] is clamped to that
 * range and keeps its (now extrapolating) weights, so no node outside the
 * mesh is ever written or read.  The reference leaves this undefined
 * (@inbounds, src/deposition.jl:183); for in-mesh particles both agree.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SCO_F32 4
#define SCO_F64 8

/* src/utils.jl:7-8 */
#define SCO_CLIGHT 299792458.0
#define SCO_FPEI (SCO_CLIGHT * SCO_CLIGHT * 1.0e-7)

/* cos and sin of -2*pi*k/n in x87 extended precision (rounded by the caller
 * to the transform's precision). */
static void sco_twiddle(long long k, long long n, long double *c, long double *s)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    long long kk = k % n;
    long double a = (2.0L * pi * (long double)kk) / (long double)n;
    *c = cosl(a);
    *s = -sinl(a);
}

static int sco_threads(int nthreads)
{
    return nthreads > 0 ? nthreads : omp_get_max_threads();
}

/* ---- double instantiation ---- */
#define R double
#define SFX(name) name##_f64
#define sqrt__f64 sqrt
#define atan__f64 atan
#define log__f64 log
#include "sc_oracle_body.h"
#undef R
#undef SFX

/* ---- float instantiation ---- */
#define R float
#define SFX(name) name##_f32
#define sqrt__f32 sqrtf
#define atan__f32 atanf
#define log__f32 logf
#include "sc_oracle_body.h"
#undef R
#undef SFX

/* ---- x87 extended instantiation: the same algorithm with a 64-bit
 * mantissa, used by the tests as the "truth" against which both the Float64
 * oracle and the GPU's Float64 path are measured ---- */
#define R long double
#define SFX(name) name##_f80
#define sqrt__f80 sqrtl
#define atan__f80 atanl
#define log__f80 logl
#include "sc_oracle_body.h"
#undef R
#undef SFX

/* ------------------------------------------------------------------------ */
/* Deposit and interpolation, for every (particle type, mesh type) pair.     */
/* Arithmetic is done in the wider of the two types, as Julia promotes it.   */
/* ------------------------------------------------------------------------ */

/* Lower cell index with the clamp policy described in the file header. */
static inline long long sco_cell(double fl, long long n)
{
    if (!(fl >= 0.0)) return 0; /* also catches NaN */
    if (fl > (double)(n - 2)) return n - 2;
    return (long long)fl;
}

/* src/deposition.jl:28-86 (deposit_particle!, use_atomic=false) driven by
 * the serial loop of deposit_vectorized! (:167-197): one thread, particle
 * order.  Accumulates into rho (the caller clears, as deposit! does at
 * :230-232). */
#define DEF_DEPOSIT(PT, MT, CT, NAME)                                                  \
    static void NAME(const PT *px, const PT *py, const PT *pz, const PT *pq,          \
                     long long n, const long long g[3], const double lo[3],            \
                     const double dl[3], MT *rho)                                      \
    {                                                                                  \
        const CT lx = (CT)(MT)lo[0], ly = (CT)(MT)lo[1], lz = (CT)(MT)lo[2];           \
        const CT ddx = (CT)(MT)dl[0], ddy = (CT)(MT)dl[1], ddz = (CT)(MT)dl[2];        \
        for (long long p = 0; p < n; ++p) {                                            \
            CT nx_ = ((CT)px[p] - lx) / ddx;                                           \
            CT ny_ = ((CT)py[p] - ly) / ddy;                                           \
            CT nz_ = ((CT)pz[p] - lz) / ddz;                                           \
            long long ix = sco_cell((double)floor(nx_), g[0]);                         \
            long long iy = sco_cell((double)floor(ny_), g[1]);                         \
            long long iz = sco_cell((double)floor(nz_), g[2]);                         \
            CT dx = nx_ - (CT)ix, dy = ny_ - (CT)iy, dz = nz_ - (CT)iz;                \
            CT wx0 = (CT)1 - dx, wx1 = dx, wy0 = (CT)1 - dy, wy1 = dy;                 \
            CT wz0 = (CT)1 - dz, wz1 = dz;                                             \
            CT c = (CT)pq[p];                                                          \
            long long b = ix + g[0] * (iy + g[1] * iz);                                \
            long long sy = g[0], sz = g[0] * g[1];                                     \
            rho[b] = (MT)((CT)rho[b] + ((c * wx0) * wy0) * wz0);                       \
            rho[b + 1] = (MT)((CT)rho[b + 1] + ((c * wx1) * wy0) * wz0);               \
            rho[b + sy] = (MT)((CT)rho[b + sy] + ((c * wx0) * wy1) * wz0);             \
            rho[b + sy + 1] = (MT)((CT)rho[b + sy + 1] + ((c * wx1) * wy1) * wz0);     \
            rho[b + sz] = (MT)((CT)rho[b + sz] + ((c * wx0) * wy0) * wz1);             \
            rho[b + sz + 1] = (MT)((CT)rho[b + sz + 1] + ((c * wx1) * wy0) * wz1);     \
            rho[b + sz + sy] = (MT)((CT)rho[b + sz + sy] + ((c * wx0) * wy1) * wz1);   \
            rho[b + sz + sy + 1] =                                                     \
                (MT)((CT)rho[b + sz + sy + 1] + ((c * wx1) * wy1) * wz1);              \
        }                                                                              \
    }

DEF_DEPOSIT(float, float, float, dep_f32_f32)
DEF_DEPOSIT(double, double, double, dep_f64_f64)
DEF_DEPOSIT(double, float, double, dep_f64_f32)
DEF_DEPOSIT(float, double, double, dep_f32_f64)

/* src/interpolation.jl:17-86 (interpolate_kernel!), outputs in the particle
 * type as similar(particles_x) makes them (:110-112). */
#define DEF_INTERP(PT, MT, CT, NAME)                                                   \
    static void NAME(const MT *ef, const PT *px, const PT *py, const PT *pz,          \
                     long long n, const long long g[3], const double lo[3],            \
                     const double dl[3], PT *Ex, PT *Ey, PT *Ez, int nth)              \
    {                                                                                  \
        const CT lx = (CT)(MT)lo[0], ly = (CT)(MT)lo[1], lz = (CT)(MT)lo[2];           \
        const CT ddx = (CT)(MT)dl[0], ddy = (CT)(MT)dl[1], ddz = (CT)(MT)dl[2];        \
        const long long M = g[0] * g[1] * g[2], sy = g[0], sz = g[0] * g[1];           \
        _Pragma("omp parallel for schedule(static) num_threads(nth)")                  \
        for (long long p = 0; p < n; ++p) {                                            \
            CT nx_ = ((CT)px[p] - lx) / ddx;                                           \
            CT ny_ = ((CT)py[p] - ly) / ddy;                                           \
            CT nz_ = ((CT)pz[p] - lz) / ddz;                                           \
            long long ix = sco_cell((double)floor(nx_), g[0]);                         \
            long long iy = sco_cell((double)floor(ny_), g[1]);                         \
            long long iz = sco_cell((double)floor(nz_), g[2]);                         \
            CT dx = nx_ - (CT)ix, dy = ny_ - (CT)iy, dz = nz_ - (CT)iz;                \
            CT w000 = (((CT)1 - dx) * ((CT)1 - dy)) * ((CT)1 - dz);                    \
            CT w100 = (dx * ((CT)1 - dy)) * ((CT)1 - dz);                              \
            CT w010 = (((CT)1 - dx) * dy) * ((CT)1 - dz);                              \
            CT w110 = (dx * dy) * ((CT)1 - dz);                                        \
            CT w001 = (((CT)1 - dx) * ((CT)1 - dy)) * dz;                              \
            CT w101 = (dx * ((CT)1 - dy)) * dz;                                        \
            CT w011 = (((CT)1 - dx) * dy) * dz;                                        \
            CT w111 = (dx * dy) * dz;                                                  \
            long long b = ix + sy * iy + sz * iz;                                      \
            PT *out[3] = {Ex, Ey, Ez};                                                 \
            for (int c = 0; c < 3; ++c) {                                              \
                const MT *e = ef + (long long)c * M + b;                               \
                CT s = (CT)e[0] * w000 + (CT)e[1] * w100;                              \
                s = s + (CT)e[sy] * w010;                                              \
                s = s + (CT)e[sy + 1] * w110;                                          \
                s = s + (CT)e[sz] * w001;                                              \
                s = s + (CT)e[sz + 1] * w101;                                          \
                s = s + (CT)e[sz + sy] * w011;                                         \
                s = s + (CT)e[sz + sy + 1] * w111;                                     \
                out[c][p] = (PT)s;                                                     \
            }                                                                          \
        }                                                                              \
    }

DEF_INTERP(float, float, float, itp_f32_f32)
DEF_INTERP(double, double, double, itp_f64_f64)
DEF_INTERP(double, float, double, itp_f64_f32)
DEF_INTERP(float, double, double, itp_f32_f64)

/* ------------------------------------------------------------------------ */
/* Exported entry points (ctypes).  dtype codes: 4 = Float32, 8 = Float64.   */
/* Bounds and spacings cross as doubles holding the mesh-type values.        */
/* ------------------------------------------------------------------------ */

int sco_version(void) { return 1; }

/* src/mesh.jl:95-174 -- bounds of the particle-extrema constructor.
 * Returns 0, or -1 on a bad grid, -2 on empty input. */
int sco_bounds_auto(int pdt, const void *x, const void *y, const void *z, long long n,
                    const long long grid[3], int mdt, double lo[3], double hi[3],
                    double delta[3])
{
    for (int a = 0; a < 3; ++a)
        if (grid[a] <= 1) return -1;                      /* :106-108 */
    if (n <= 0) return -2;                                /* :110-112 */
    const void *P[3] = {x, y, z};
    double mn[3], mx[3], d0[3];
    for (int a = 0; a < 3; ++a) {
        if (pdt == SCO_F32) {
            const float *v = (const float *)P[a];
            float lo_ = v[0], hi_ = v[0];
            for (long long i = 1; i < n; ++i) {
                if (v[i] < lo_) lo_ = v[i];
                if (v[i] > hi_) hi_ = v[i];
            }
            /* delta in the particle type: Float32 / Int -> Float32 (:127-129) */
            float d = (hi_ - lo_) / (float)(grid[a] - 1);
            mn[a] = lo_;
            mx[a] = hi_;
            d0[a] = d;
        } else {
            const double *v = (const double *)P[a];
            double lo_ = v[0], hi_ = v[0];
            for (long long i = 1; i < n; ++i) {
                if (v[i] < lo_) lo_ = v[i];
                if (v[i] > hi_) hi_ = v[i];
            }
            mn[a] = lo_;
            mx[a] = hi_;
            d0[a] = (hi_ - lo_) / (double)(grid[a] - 1);
        }
    }
    for (int a = 0; a < 3; ++a) {
        /* padding by the Float64 literal 1e-6 promotes to Float64 (:132-137) */
        double l = mn[a] - 1e-6 * d0[a];
        double h = mx[a] + 1e-6 * d0[a];
        double d = (h - l) / (double)(grid[a] - 1);       /* :140-142 */
        if (d == 0.0) d = 1e-6;                           /* :145-149 */
        if (mdt == SCO_F32) {                             /* :150-156 */
            lo[a] = (float)l;
            hi[a] = (float)h;
            delta[a] = (float)d;
        } else {
            lo[a] = l;
            hi[a] = h;
            delta[a] = d;
        }
    }
    return 0;
}

/* src/mesh.jl:196-238 -- manual-bounds constructor.  -1 bad grid, -3 bounds. */
int sco_bounds_manual(const double mn[3], const double mx[3], const long long grid[3],
                      int mdt, double lo[3], double hi[3], double delta[3])
{
    for (int a = 0; a < 3; ++a)
        if (grid[a] <= 1) return -1;                      /* :206-208 */
    for (int a = 0; a < 3; ++a)
        if (!(mx[a] > mn[a])) return -3;                  /* :209-211 */
    for (int a = 0; a < 3; ++a) {
        if (mdt == SCO_F32) {
            float l = (float)mn[a], h = (float)mx[a];
            lo[a] = l;
            hi[a] = h;
            delta[a] = (h - l) / (float)(grid[a] - 1);    /* :219 */
        } else {
            lo[a] = mn[a];
            hi[a] = mx[a];
            delta[a] = (mx[a] - mn[a]) / (double)(grid[a] - 1);
        }
    }
    return 0;
}

int sco_deposit(int pdt, int mdt, const void *x, const void *y, const void *z,
                const void *q, long long n, const long long grid[3], const double lo[3],
                const double delta[3], void *rho)
{
    if (pdt == SCO_F32 && mdt == SCO_F32)
        dep_f32_f32(x, y, z, q, n, grid, lo, delta, rho);
    else if (pdt == SCO_F64 && mdt == SCO_F64)
        dep_f64_f64(x, y, z, q, n, grid, lo, delta, rho);
    else if (pdt == SCO_F64 && mdt == SCO_F32)
        dep_f64_f32(x, y, z, q, n, grid, lo, delta, rho);
    else if (pdt == SCO_F32 && mdt == SCO_F64)
        dep_f32_f64(x, y, z, q, n, grid, lo, delta, rho);
    else
        return -4;
    return 0;
}

int sco_interpolate(int pdt, int mdt, const void *efield, const void *x, const void *y,
                    const void *z, long long n, const long long grid[3], const double lo[3],
                    const double delta[3], void *Ex, void *Ey, void *Ez, int nthreads)
{
    int nth = sco_threads(nthreads);
    if (pdt == SCO_F32 && mdt == SCO_F32)
        itp_f32_f32(efield, x, y, z, n, grid, lo, delta, Ex, Ey, Ez, nth);
    else if (pdt == SCO_F64 && mdt == SCO_F64)
        itp_f64_f64(efield, x, y, z, n, grid, lo, delta, Ex, Ey, Ez, nth);
    else if (pdt == SCO_F64 && mdt == SCO_F32)
        itp_f64_f32(efield, x, y, z, n, grid, lo, delta, Ex, Ey, Ez, nth);
    else if (pdt == SCO_F32 && mdt == SCO_F64)
        itp_f32_f64(efield, x, y, z, n, grid, lo, delta, Ex, Ey, Ez, nth);
    else
        return -4;
    return 0;
}

int sco_solve(int mdt, const long long grid[3], const double lo[3], const double hi[3],
              const double delta[3], double gamma, int at_cathode, const void *rho,
              void *efield, int nthreads)
{
    const int g[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    int nth = sco_threads(nthreads);
    if (mdt == SCO_F64) {
        return solve_f64(rho, efield, g, lo, hi, delta, gamma, at_cathode, nth);
    } else if (mdt == SCO_F32) {
        const float l[3] = {(float)lo[0], (float)lo[1], (float)lo[2]};
        const float h[3] = {(float)hi[0], (float)hi[1], (float)hi[2]};
        const float d[3] = {(float)delta[0], (float)delta[1], (float)delta[2]};
        return solve_f32(rho, efield, g, l, h, d, (float)gamma, at_cathode, nth);
    }
    return -4;
}

/* A persistent solve workspace (src/mesh.jl:49-72) of Float64 or Float32
 * type for a grid; NULL on allocation failure. */
void *sco_workspace_create(int mdt, const long long grid[3])
{
    const int g[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    if (mdt == SCO_F64) return ws_create_f64(g);
    if (mdt == SCO_F32) return ws_create_f32(g);
    return NULL;
}

void sco_workspace_free(int mdt, void *ws)
{
    if (mdt == SCO_F64) ws_free_f64((workspace_f64 *)ws);
    else if (mdt == SCO_F32) ws_free_f32((workspace_f32 *)ws);
}

/* sco_solve on a workspace from sco_workspace_create (same grid and type). */
int sco_solve_ws(void *ws, int mdt, const double lo[3], const double hi[3], const double delta[3],
                 double gamma, int at_cathode, const void *rho, void *efield, int nthreads)
{
    if (!ws) return -1;
    int nth = sco_threads(nthreads);
    if (mdt == SCO_F64)
        return solve_ws_f64((workspace_f64 *)ws, rho, efield, lo, hi, delta, gamma, at_cathode, nth);
    if (mdt == SCO_F32) {
        const float l[3] = {(float)lo[0], (float)lo[1], (float)lo[2]};
        const float h[3] = {(float)hi[0], (float)hi[1], (float)hi[2]};
        const float d[3] = {(float)delta[0], (float)delta[1], (float)delta[2]};
        return solve_ws_f32((workspace_f32 *)ws, rho, efield, l, h, d, (float)gamma, at_cathode,
                            nth);
    }
    return -4;
}

/* The solve in x87 extended precision: rho and efield cross as Float64
 * arrays, bounds/spacings/gamma as the Float64 mesh values. */
int sco_solve_extended(const long long grid[3], const double lo[3], const double hi[3],
                       const double delta[3], double gamma, int at_cathode, const double *rho,
                       double *efield, int nthreads)
{
    const int g[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    const long long M = grid[0] * grid[1] * grid[2];
    long double *r = (long double *)malloc(sizeof(long double) * (size_t)M);
    long double *e = (long double *)malloc(sizeof(long double) * 3 * (size_t)M);
    if (!r || !e) { free(r); free(e); return -1; }
    for (long long i = 0; i < M; ++i) r[i] = rho[i];
    const long double l[3] = {lo[0], lo[1], lo[2]};
    const long double h[3] = {hi[0], hi[1], hi[2]};
    const long double d[3] = {delta[0], delta[1], delta[2]};
    int rc = solve_f80(r, e, g, l, h, d, (long double)gamma, at_cathode, sco_threads(nthreads));
    for (long long i = 0; i < 3 * M; ++i) efield[i] = (double)e[i];
    free(r);
    free(e);
    return rc;
}

/* Scalar potential of rho in free space (no reference solver; the Green
 * function is src/green_functions.jl:13-22).  phi has M elements. */
int sco_solve_potential(int mdt, const long long grid[3], const double delta[3], double gamma,
                        const void *rho, void *phi, int nthreads)
{
    const int g[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    int nth = sco_threads(nthreads);
    if (mdt == SCO_F64) return solve_potential_f64(rho, phi, g, delta, gamma, nth);
    if (mdt == SCO_F32) {
        const float d[3] = {(float)delta[0], (float)delta[1], (float)delta[2]};
        return solve_potential_f32(rho, phi, g, d, (float)gamma, nth);
    }
    return -4;
}

int sco_solve_potential_extended(const long long grid[3], const double delta[3], double gamma,
                                 const double *rho, double *phi, int nthreads)
{
    const int g[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    const long long M = grid[0] * grid[1] * grid[2];
    long double *r = (long double *)malloc(sizeof(long double) * (size_t)M);
    long double *p = (long double *)malloc(sizeof(long double) * (size_t)M);
    if (!r || !p) { free(r); free(p); return -1; }
    for (long long i = 0; i < M; ++i) r[i] = rho[i];
    const long double d[3] = {delta[0], delta[1], delta[2]};
    int rc = solve_potential_f80(r, p, g, d, (long double)gamma, sco_threads(nthreads));
    for (long long i = 0; i < M; ++i) phi[i] = (double)p[i];
    free(r);
    free(p);
    return rc;
}

/* Transpose of the field solve (no reference counterpart): rho_g = S^T g,
 * g of 3M elements, rho_g of M. */
int sco_solve_transpose(int mdt, const long long grid[3], const double lo[3], const double hi[3],
                        const double delta[3], double gamma, int at_cathode, const void *g,
                        void *rho_g, int nthreads)
{
    const int gr[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    int nth = sco_threads(nthreads);
    if (mdt == SCO_F64) return solve_transpose_f64(g, rho_g, gr, lo, hi, delta, gamma, at_cathode, nth);
    if (mdt == SCO_F32) {
        const float l[3] = {(float)lo[0], (float)lo[1], (float)lo[2]};
        const float h[3] = {(float)hi[0], (float)hi[1], (float)hi[2]};
        const float d[3] = {(float)delta[0], (float)delta[1], (float)delta[2]};
        return solve_transpose_f32(g, rho_g, gr, l, h, d, (float)gamma, at_cathode, nth);
    }
    return -4;
}

int sco_solve_transpose_extended(const long long grid[3], const double lo[3], const double hi[3],
                                 const double delta[3], double gamma, int at_cathode,
                                 const double *g, double *rho_g, int nthreads)
{
    const int gr[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    const long long M = grid[0] * grid[1] * grid[2];
    long double *a = (long double *)malloc(sizeof(long double) * 3 * (size_t)M);
    long double *r = (long double *)malloc(sizeof(long double) * (size_t)M);
    if (!a || !r) { free(a); free(r); return -1; }
    for (long long i = 0; i < 3 * M; ++i) a[i] = g[i];
    const long double l[3] = {lo[0], lo[1], lo[2]};
    const long double h[3] = {hi[0], hi[1], hi[2]};
    const long double d[3] = {delta[0], delta[1], delta[2]};
    int rc = solve_transpose_f80(a, r, gr, l, h, d, (long double)gamma, at_cathode,
                                 sco_threads(nthreads));
    for (long long i = 0; i < M; ++i) rho_g[i] = (double)r[i];
    free(a);
    free(r);
    return rc;
}

/* ---- internals exposed for the oracle's own tests ---- */

double sco_potential_green_function_f64(double x, double y, double z)
{
    return potential_green_function_f64(x, y, z);
}

double sco_field_green_function_f64(double x, double y, double z)
{
    return field_green_function_f64(x, y, z);
}

float sco_field_green_function_f32(float x, float y, float z)
{
    return field_green_function_f32(x, y, z);
}

/* The (2nx,2ny,2nz) complex Green lattice after differencing, before its FFT. */
int sco_green_lattice(int mdt, const long long grid[3], const double delta[3], double gamma,
                      int icomp, const double offset[3], void *cgrn, int nthreads)
{
    const int g[3] = {(int)grid[0], (int)grid[1], (int)grid[2]};
    const long long P = 8 * grid[0] * grid[1] * grid[2];
    int nth = sco_threads(nthreads);
    if (mdt == SCO_F64) {
        double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)P);
        get_green_function_f64(cgrn, tmp, g, delta, gamma, icomp, offset, nth);
        free(tmp);
    } else if (mdt == SCO_F32) {
        float *tmp = (float *)malloc(sizeof(float) * 2 * (size_t)P);
        const float d[3] = {(float)delta[0], (float)delta[1], (float)delta[2]};
        const float o[3] = {(float)offset[0], (float)offset[1], (float)offset[2]};
        get_green_function_f32(cgrn, tmp, g, d, (float)gamma, icomp, o, nth);
        free(tmp);
    } else {
        return -4;
    }
    return 0;
}

/* In-place unnormalised 3-D transform, sign -1 forward / +1 backward. */
int sco_fft3d(int mdt, const long long dims[3], void *data, int sign, int nthreads)
{
    const int d[3] = {(int)dims[0], (int)dims[1], (int)dims[2]};
    int nth = sco_threads(nthreads);
    if (mdt == SCO_F64) fft3d_f64(data, d, sign, nth);
    else if (mdt == SCO_F32) fft3d_f32(data, d, sign, nth);
    else return -4;
    return 0;
}

double sco_fpei(void) { return SCO_FPEI; }

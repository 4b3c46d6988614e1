/*
 * sc_oracle_body.h -- precision-generic body of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: this file is included twice by sc_oracle.c, once
 * with R = double and once with R = float.  Only tests/, the smoke check and
 * the bench's cpu_baseline leg may use the oracle; the product never does.
 *
 * Every function restates one piece of the SpaceCharge.jl v1.2.0 CPU path in
 * the reference's own operation order (citations are file:line in the
 * reference repository).  Julia evaluates `a - b + c` as `(a - b) + c`, `x^2`
 * as `x*x` and never contracts a*b+c into an FMA, so this file is compiled
 * with -ffp-contract=off and keeps every parenthesisation explicit.
 *
 * Requires (from the includer): R, SFX(name), and the shared helpers
 * sco_cpx_* / sco_omp_* defined in sc_oracle.c.
 */

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, any length.  The reference calls      */
/* FFTW 3.3.11 through AbstractFFTs (src/mesh.jl:60-61, plan_fft!/plan_ifft!).*/
/* FFTW's forward transform is X[k] = sum_j x[j] exp(-2 pi i jk/N); the       */
/* backward transform has +i and no scaling (ifft's 1/N is applied by         */
/* AbstractFFTs' ScaledPlan afterwards, see sco_solve_freespace below).       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;
    int nst;
    int rad[64];
    R *tw; /* tw[2m], tw[2m+1] = cos, sin of -2 pi m / n, m in [0, n) */
} SFX(fftplan);

static void SFX(fftplan_init)(SFX(fftplan) *p, int n)
{
    p->n = n;
    p->nst = 0;
    int m = n;
    while (m % 4 == 0) { p->rad[p->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p->rad[p->nst++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->rad[p->nst++] = f; m /= f; }
    }
    p->tw = (R *)malloc(sizeof(R) * 2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        long double c, s;
        sco_twiddle(k, n, &c, &s);
        p->tw[2 * k] = (R)c;
        p->tw[2 * k + 1] = (R)s;
    }
}

static void SFX(fftplan_free)(SFX(fftplan) *p)
{
    free(p->tw);
    p->tw = NULL;
}

/* One length-n transform of x (interleaved complex), using work (2n reals).
 * sign = -1 forward, +1 backward (unnormalised). */
static void SFX(fft1d)(const SFX(fftplan) *p, R *x, R *work, int sign)
{
    const int n = p->n;
    if (n == 1) return;
    R *in = x, *out = work;
    int Ns = 1;
    R va[2 * 64], vb[2 * 64];
    for (int st = 0; st < p->nst; ++st) {
        const int Rd = p->rad[st];
        const int nR = n / Rd;
        const int tstep = n / (Ns * Rd);
        for (int j = 0; j < nR; ++j) {
            const int k = j % Ns;
            for (int q = 0; q < Rd; ++q) {
                R ar = in[2 * (j + q * nR)], ai = in[2 * (j + q * nR) + 1];
                if (q > 0 && k > 0) {
                    int m = (int)(((long long)q * k * tstep) % n);
                    R wr = p->tw[2 * m], wi = p->tw[2 * m + 1];
                    if (sign > 0) wi = -wi;
                    R tr = ar * wr - ai * wi;
                    R ti = ar * wi + ai * wr;
                    ar = tr;
                    ai = ti;
                }
                va[2 * q] = ar;
                va[2 * q + 1] = ai;
            }
            if (Rd == 2) {
                vb[0] = va[0] + va[2];
                vb[1] = va[1] + va[3];
                vb[2] = va[0] - va[2];
                vb[3] = va[1] - va[3];
            } else if (Rd == 4) {
                R t0r = va[0] + va[4], t0i = va[1] + va[5];
                R t1r = va[0] - va[4], t1i = va[1] - va[5];
                R t2r = va[2] + va[6], t2i = va[3] + va[7];
                R t3r = va[2] - va[6], t3i = va[3] - va[7];
                /* forward: multiply t3 by -i; backward: by +i */
                R ur, ui;
                if (sign < 0) { ur = t3i; ui = -t3r; }
                else { ur = -t3i; ui = t3r; }
                vb[0] = t0r + t2r; vb[1] = t0i + t2i;
                vb[2] = t1r + ur;  vb[3] = t1i + ui;
                vb[4] = t0r - t2r; vb[5] = t0i - t2i;
                vb[6] = t1r - ur;  vb[7] = t1i - ui;
            } else {
                /* generic odd-prime DFT from the definition */
                for (int pp = 0; pp < Rd; ++pp) {
                    R sr = 0, si = 0;
                    for (int q = 0; q < Rd; ++q) {
                        int m = (int)(((long long)pp * q % Rd) * (n / Rd));
                        R wr = p->tw[2 * m], wi = p->tw[2 * m + 1];
                        if (sign > 0) wi = -wi;
                        sr += va[2 * q] * wr - va[2 * q + 1] * wi;
                        si += va[2 * q] * wi + va[2 * q + 1] * wr;
                    }
                    vb[2 * pp] = sr;
                    vb[2 * pp + 1] = si;
                }
            }
            const int base = (j / Ns) * Ns * Rd + k;
            for (int q = 0; q < Rd; ++q) {
                out[2 * (base + q * Ns)] = vb[2 * q];
                out[2 * (base + q * Ns) + 1] = vb[2 * q + 1];
            }
        }
        R *t = in; in = out; out = t;
        Ns *= Rd;
    }
    if (in != x) memcpy(x, in, sizeof(R) * 2 * (size_t)n);
}

/* In-place 3-D complex transform of a column-major (d0, d1, d2) array,
 * x fastest (Julia's layout), with the plans pl[0..2] of the three lengths.
 * Rows of every axis are transformed independently; OpenMP splits rows over
 * threads (deterministic). */
static void SFX(fft3d_pl)(R *a, const SFX(fftplan) pl[3], int sign, int nthreads)
{
    const long long d0 = pl[0].n, d1 = pl[1].n, d2 = pl[2].n;
    const int B = 16; /* pencils gathered together for the strided axes */
    /* axis 0: contiguous rows */
    #pragma omp parallel num_threads(nthreads)
    {
        R *work = (R *)malloc(sizeof(R) * 2 * (size_t)d0);
        #pragma omp for schedule(static)
        for (long long r = 0; r < d1 * d2; ++r)
            SFX(fft1d)(&pl[0], a + 2 * r * d0, work, sign);
        free(work);
    }
    /* axis 1: stride d0 */
    #pragma omp parallel num_threads(nthreads)
    {
        R *buf = (R *)malloc(sizeof(R) * 2 * (size_t)d1 * B);
        R *work = (R *)malloc(sizeof(R) * 2 * (size_t)d1);
        long long nblk = (d0 + B - 1) / B;
        #pragma omp for schedule(static)
        for (long long t = 0; t < nblk * d2; ++t) {
            long long k = t / nblk, i0 = (t % nblk) * B;
            int nb = (int)((d0 - i0) < B ? (d0 - i0) : B);
            for (long long j = 0; j < d1; ++j)
                for (int b = 0; b < nb; ++b) {
                    long long src = 2 * ((i0 + b) + d0 * (j + d1 * k));
                    buf[2 * (b * d1 + j)] = a[src];
                    buf[2 * (b * d1 + j) + 1] = a[src + 1];
                }
            for (int b = 0; b < nb; ++b) SFX(fft1d)(&pl[1], buf + 2 * (size_t)b * d1, work, sign);
            for (long long j = 0; j < d1; ++j)
                for (int b = 0; b < nb; ++b) {
                    long long dst = 2 * ((i0 + b) + d0 * (j + d1 * k));
                    a[dst] = buf[2 * (b * d1 + j)];
                    a[dst + 1] = buf[2 * (b * d1 + j) + 1];
                }
        }
        free(buf);
        free(work);
    }
    /* axis 2: stride d0*d1 */
    #pragma omp parallel num_threads(nthreads)
    {
        R *buf = (R *)malloc(sizeof(R) * 2 * (size_t)d2 * B);
        R *work = (R *)malloc(sizeof(R) * 2 * (size_t)d2);
        long long nblk = (d0 + B - 1) / B;
        #pragma omp for schedule(static)
        for (long long t = 0; t < nblk * d1; ++t) {
            long long j = t / nblk, i0 = (t % nblk) * B;
            int nb = (int)((d0 - i0) < B ? (d0 - i0) : B);
            for (long long k = 0; k < d2; ++k)
                for (int b = 0; b < nb; ++b) {
                    long long src = 2 * ((i0 + b) + d0 * (j + d1 * k));
                    buf[2 * (b * d2 + k)] = a[src];
                    buf[2 * (b * d2 + k) + 1] = a[src + 1];
                }
            for (int b = 0; b < nb; ++b) SFX(fft1d)(&pl[2], buf + 2 * (size_t)b * d2, work, sign);
            for (long long k = 0; k < d2; ++k)
                for (int b = 0; b < nb; ++b) {
                    long long dst = 2 * ((i0 + b) + d0 * (j + d1 * k));
                    a[dst] = buf[2 * (b * d2 + k)];
                    a[dst + 1] = buf[2 * (b * d2 + k) + 1];
                }
        }
        free(buf);
        free(work);
    }
}

/* fft3d_pl with plans made for this one call. */
static void SFX(fft3d)(R *a, const int d[3], int sign, int nthreads)
{
    SFX(fftplan) pl[3];
    for (int ax = 0; ax < 3; ++ax) SFX(fftplan_init)(&pl[ax], d[ax]);
    SFX(fft3d_pl)(a, pl, sign, nthreads);
    for (int ax = 0; ax < 3; ++ax) SFX(fftplan_free)(&pl[ax]);
}

/* One padded transform: with the cached plans of a workspace when pl is
 * given (src/mesh.jl:60-61 plans once per mesh), else planned per call. */
static void SFX(fft3d_any)(R *a, const int d[3], const SFX(fftplan) *pl, int sign, int nthreads)
{
    if (pl) SFX(fft3d_pl)(a, pl, sign, nthreads);
    else SFX(fft3d)(a, d, sign, nthreads);
}

/* ------------------------------------------------------------------------ */
/* Green functions                                                          */
/* ------------------------------------------------------------------------ */

/* src/green_functions.jl:35-38 -- antiderivative of x/r^3:
 *   F = x*atan((y*z)/(r*x)) - z*log(r+y) + y*log((r-z)/(r+z))/2 */
static R SFX(field_green_function)(R x, R y, R z)
{
    R r = SFX(sqrt_)((x * x + y * y) + z * z);
    R t1 = x * SFX(atan_)((y * z) / (r * x));
    R t2 = z * SFX(log_)(r + y);
    R t3 = (y * SFX(log_)((r - z) / (r + z))) / (R)2;
    return (t1 - t2) + t3;
}

/* src/green_functions.jl:13-22 -- antiderivative of 1/r (the scalar
 * potential), zero at r = 0:
 *   -z^2 atan(xy/(zr))/2 - y^2 atan(xz/(yr))/2 - x^2 atan(yz/(xr))/2
 *   + yz log(x+r) + xz log(y+r) + xy log(z+r) */
static R SFX(potential_green_function)(R x, R y, R z)
{
    R r = SFX(sqrt_)((x * x + y * y) + z * z);
    if (r == (R)0) return (R)0;
    const R half = (R)1 / (R)2;
    R t = -half * (z * z) * SFX(atan_)((x * y) / (z * r));
    t = t - half * (y * y) * SFX(atan_)((x * z) / (y * r));
    t = t - half * (x * x) * SFX(atan_)((y * z) / (x * r));
    t = t + (y * z) * SFX(log_)(x + r);
    t = t + (x * z) * SFX(log_)(y + r);
    t = t + (x * y) * SFX(log_)(z + r);
    return t;
}

/* src/green_functions.jl:41-112 -- get_green_function!: pointwise F*factor on
 * the (2nx,2ny,2nz) node lattice (get_green_kernel!, :69-101), then the
 * 8-point difference (apply_8point_differencing!, :103-112) written to a
 * scratch array and copied back over [1:end-1, 1:end-1, 1:end-1] (:63-66).
 * The last i, j and k planes keep the raw F*factor values.
 * cgrn and temp are interleaved complex arrays, column-major (2nx,2ny,2nz). */
static void SFX(get_green_function)(R *cgrn, R *temp, const int grid[3],
                                    const R delta[3], R gamma, int icomp,
                                    const R offset[3], int nthreads)
{
    const long long isz = 2LL * grid[0], jsz = 2LL * grid[1], ksz = 2LL * grid[2];
    /* get_green_kernel! (:69-101) */
    const R dx = delta[0], dy = delta[1], dz = delta[2] * gamma;
    R factor;
    /* icomp 4: the scalar potential (no solver in the reference uses it);
     * the lab-frame potential gamma * phi_rest takes the transverse factor */
    if (icomp == 1 || icomp == 2 || icomp == 4) factor = gamma / ((dx * dy) * dz);
    else factor = (R)1 / ((dx * dy) * dz);
    const R umin = ((R)(1 - isz) * dx) / (R)2 + offset[0];
    const R vmin = ((R)(1 - jsz) * dy) / (R)2 + offset[1];
    const R wmin = ((R)(1 - ksz) * dz) / (R)2 + offset[2] * gamma;
    #pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
    for (long long k = 0; k < ksz; ++k)
        for (long long j = 0; j < jsz; ++j) {
            const R w = (R)k * dz + wmin;
            const R v = (R)j * dy + vmin;
            for (long long i = 0; i < isz; ++i) {
                const R u = (R)i * dx + umin;
                R g;
                if (icomp == 1) g = SFX(field_green_function)(u, v, w) * factor;
                else if (icomp == 2) g = SFX(field_green_function)(v, w, u) * factor;
                else if (icomp == 3) g = SFX(field_green_function)(w, u, v) * factor;
                else if (icomp == 4) g = SFX(potential_green_function)(u, v, w) * factor;
                else g = (R)0;
                long long idx = i + isz * (j + jsz * k);
                cgrn[2 * idx] = g;
                cgrn[2 * idx + 1] = (R)0;
            }
        }
    /* apply_8point_differencing! (:103-112) into temp, ndrange (isz-1, jsz-1, ksz-1) */
#define G_(a, b, c) (cgrn + 2 * ((a) + isz * ((b) + jsz * (c))))
    #pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
    for (long long k = 0; k < ksz - 1; ++k)
        for (long long j = 0; j < jsz - 1; ++j)
            for (long long i = 0; i < isz - 1; ++i) {
                for (int part = 0; part < 2; ++part) {
                    R s = G_(i + 1, j + 1, k + 1)[part] - G_(i, j + 1, k + 1)[part];
                    s = s - G_(i + 1, j, k + 1)[part];
                    s = s - G_(i + 1, j + 1, k)[part];
                    s = s - G_(i, j, k)[part];
                    s = s + G_(i, j, k + 1)[part];
                    s = s + G_(i, j + 1, k)[part];
                    s = s + G_(i + 1, j, k)[part];
                    temp[2 * (i + isz * (j + jsz * k)) + part] = s;
                }
            }
    /* copy back (:63-66) */
    #pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
    for (long long k = 0; k < ksz - 1; ++k)
        for (long long j = 0; j < jsz - 1; ++j)
            for (long long i = 0; i < isz - 1; ++i) {
                long long idx = i + isz * (j + jsz * k);
                cgrn[2 * idx] = temp[2 * idx];
                cgrn[2 * idx + 1] = temp[2 * idx + 1];
            }
#undef G_
}

/* ------------------------------------------------------------------------ */
/* Solvers                                                                   */
/* ------------------------------------------------------------------------ */

/* src/solvers/free_space.jl:56-101 -- solve_freespace!.
 * ws holds three (2nx,2ny,2nz) complex arrays (crho, cgrn, temp), like the
 * lazily created workspace of src/mesh.jl:49-72. */
static void SFX(solve_freespace)(const R *rho, R *efield, const int grid[3],
                                 const R delta[3], R gamma, const R offset[3],
                                 R *crho, R *cgrn, R *temp, const SFX(fftplan) *pl,
                                 int nthreads, int c0, int c1)
{
    const long long nx = grid[0], ny = grid[1], nz = grid[2];
    const long long isz = 2 * nx, jsz = 2 * ny, ksz = 2 * nz;
    const long long P = isz * jsz * ksz, M = nx * ny * nz;
    const int dims[3] = {(int)isz, (int)jsz, (int)ksz};
    /* fill!(crho, 0); crho[1:nx,1:ny,1:nz] .= rho  (:68-69) */
    memset(crho, 0, sizeof(R) * 2 * (size_t)P);
    for (long long k = 0; k < nz; ++k)
        for (long long j = 0; j < ny; ++j)
            for (long long i = 0; i < nx; ++i)
                crho[2 * (i + isz * (j + jsz * k))] = rho[i + nx * (j + ny * k)];
    /* fft_plan_inplace * crho  (:72) */
    SFX(fft3d_any)(crho, dims, pl, -1, nthreads);
    /* factr = T(FPEI)  (:75) */
    const R factr = (R)SCO_FPEI;
    /* AbstractFFTs normalisation of plan_ifft!: one(T)/prod(size) */
    const R scale = (R)1 / (R)P;
    for (int icomp = c0; icomp <= c1; ++icomp) {
        SFX(get_green_function)(cgrn, temp, grid, delta, gamma, icomp, offset, nthreads);
        SFX(fft3d_any)(cgrn, dims, pl, -1, nthreads);               /* :89 */
        #pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long long e = 0; e < P; ++e) {                          /* :92 */
            R ar = crho[2 * e], ai = crho[2 * e + 1];
            R br = cgrn[2 * e], bi = cgrn[2 * e + 1];
            temp[2 * e] = ar * br - ai * bi;
            temp[2 * e + 1] = ar * bi + ai * br;
        }
        SFX(fft3d_any)(temp, dims, pl, +1, nthreads);               /* :95 */
        #pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long long e = 0; e < P; ++e) {
            temp[2 * e] = temp[2 * e] * scale;
            temp[2 * e + 1] = temp[2 * e + 1] * scale;
        }
        /* efield[:,:,:,icomp] = factr*real(temp[nx:2nx-1, ...])  (:98-99) */
        R *E = efield + (size_t)(icomp - c0) * M;
        for (long long k = 0; k < nz; ++k)
            for (long long j = 0; j < ny; ++j)
                for (long long i = 0; i < nx; ++i) {
                    long long src = (i + nx - 1) + isz * ((j + ny - 1) + jsz * (k + nz - 1));
                    E[i + nx * (j + ny * k)] = factr * temp[2 * src];
                }
    }
}

/* src/solvers/free_space.jl:14-47 -- solve!(mesh; at_cathode).  The image
 * charge is -rho flipped in z (:34), solved with z-offset
 * 2*zmin + (zmax - zmin) (:39) and added to the free-space field (:45). */
static int SFX(solve_in)(const R *rho, R *efield, const int grid[3], const R lo[3],
                         const R hi[3], const R delta[3], R gamma, int at_cathode,
                         R *crho, R *cgrn, R *temp, const SFX(fftplan) *pl, int nthreads)
{
    const long long nx = grid[0], ny = grid[1], nz = grid[2];
    const long long M = nx * ny * nz;
    const R zero3[3] = {0, 0, 0};
    SFX(solve_freespace)(rho, efield, grid, delta, gamma, zero3, crho, cgrn, temp, pl, nthreads,
                         1, 3);
    if (at_cathode) {
        R *irho = (R *)malloc(sizeof(R) * (size_t)M);
        R *ief = (R *)malloc(sizeof(R) * 3 * (size_t)M);
        for (long long k = 0; k < nz; ++k)
            for (long long j = 0; j < ny; ++j)
                for (long long i = 0; i < nx; ++i)
                    irho[i + nx * (j + ny * k)] = -rho[i + nx * (j + ny * (nz - 1 - k))];
        const R offz = (R)2 * lo[2] + (hi[2] - lo[2]);
        const R off[3] = {0, 0, offz};
        SFX(solve_freespace)(irho, ief, grid, delta, gamma, off, crho, cgrn, temp, pl, nthreads, 1,
                             3);
        for (long long e = 0; e < 3 * M; ++e) efield[e] = efield[e] + ief[e];
        free(irho);
        free(ief);
    }
    return 0;
}

/* solve! with a workspace allocated (and plans made) for this one call. */
static int SFX(solve)(const R *rho, R *efield, const int grid[3], const R lo[3],
                      const R hi[3], const R delta[3], R gamma, int at_cathode,
                      int nthreads)
{
    const long long P = 8 * (long long)grid[0] * grid[1] * grid[2];
    R *crho = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *cgrn = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *temp = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    int rc = -1;
    if (crho && cgrn && temp)
        rc = SFX(solve_in)(rho, efield, grid, lo, hi, delta, gamma, at_cathode, crho, cgrn, temp,
                           NULL, nthreads);
    free(crho);
    free(cgrn);
    free(temp);
    return rc;
}

/* The mesh's lazily created workspace, src/mesh.jl:49-72: the three padded
 * complex arrays (crho, cgrn, temp_result) and the transform plans, made once
 * and reused by every solve of the mesh. */
typedef struct {
    int grid[3];
    R *crho, *cgrn, *temp;
    SFX(fftplan) pl[3];
} SFX(workspace);

static inline void SFX(ws_free)(SFX(workspace) *w)
{
    if (!w) return;
    free(w->crho);
    free(w->cgrn);
    free(w->temp);
    for (int a = 0; a < 3; ++a)
        if (w->pl[a].tw) SFX(fftplan_free)(&w->pl[a]);
    free(w);
}

static inline SFX(workspace) *SFX(ws_create)(const int grid[3])
{
    SFX(workspace) *w = (SFX(workspace) *)calloc(1, sizeof(SFX(workspace)));
    if (!w) return NULL;
    const long long P = 8 * (long long)grid[0] * grid[1] * grid[2];
    for (int a = 0; a < 3; ++a) w->grid[a] = grid[a];
    w->crho = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    w->cgrn = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    w->temp = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    if (!w->crho || !w->cgrn || !w->temp) { SFX(ws_free)(w); return NULL; }
    /* touch every page once, as the first solve! does */
    memset(w->crho, 0, sizeof(R) * 2 * (size_t)P);
    memset(w->cgrn, 0, sizeof(R) * 2 * (size_t)P);
    memset(w->temp, 0, sizeof(R) * 2 * (size_t)P);
    for (int a = 0; a < 3; ++a) SFX(fftplan_init)(&w->pl[a], 2 * grid[a]);
    return w;
}

static inline int SFX(solve_ws)(SFX(workspace) *w, const R *rho, R *efield, const R lo[3],
                                const R hi[3], const R delta[3], R gamma, int at_cathode,
                                int nthreads)
{
    return SFX(solve_in)(rho, efield, w->grid, lo, hi, delta, gamma, at_cathode, w->crho, w->cgrn,
                         w->temp, w->pl, nthreads);
}

/* Scalar potential, free space: phi = FPEI * gamma/(dx dy dz_rest) *
 * sum_n rho[n] * (cell integral of 1/r), the same linear convolution as the
 * field solve (src/solvers/free_space.jl:56-101) with the antiderivative of
 * src/green_functions.jl:13-22.  The reference defines that function but no
 * solver calls it. */
static int SFX(solve_potential)(const R *rho, R *phi, const int grid[3], const R delta[3],
                                R gamma, int nthreads)
{
    const long long P = 8LL * grid[0] * grid[1] * grid[2];
    R *crho = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *cgrn = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *temp = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    if (!crho || !cgrn || !temp) { free(crho); free(cgrn); free(temp); return -1; }
    const R zero3[3] = {0, 0, 0};
    SFX(solve_freespace)(rho, phi, grid, delta, gamma, zero3, crho, cgrn, temp, NULL, nthreads, 4,
                         4);
    free(crho);
    free(cgrn);
    free(temp);
    return 0;
}

/* Transpose of SFX(solve), rho_g = S^T g for g = (g_x, g_y, g_z) (no solver
 * in the reference; this is the vector-Jacobian product of solve!).
 * S_c convolves rho with the cell integral G_c of src/green_functions.jl,
 * which is odd along axis c and even along the other two, so the transpose
 * of the free-space part is S_c^T g_c = -(S_c g_c).  The cathode image term
 * convolves -flip_z(rho) with G_c(r + off), off = (0, 0, zmin + zmax)
 * (src/solvers/free_space.jl:34-45); G_c(r + off) = -G_c(-r - off), so its
 * transpose is minus the convolution with the Green function at offset -off,
 * and -flip_z is its own transpose:
 *   rho_g = -sum_c S_c g_c + flip_z(sum_c S_c^(-off) g_c). */
static int SFX(solve_transpose)(const R *g, R *out, const int grid[3], const R lo[3],
                                const R hi[3], const R delta[3], R gamma, int at_cathode,
                                int nthreads)
{
    const long long nx = grid[0], ny = grid[1], nz = grid[2];
    const long long M = nx * ny * nz, P = 8 * M;
    R *crho = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *cgrn = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *temp = (R *)malloc(sizeof(R) * 2 * (size_t)P);
    R *ef = (R *)malloc(sizeof(R) * (size_t)M);
    R *acc = (R *)malloc(sizeof(R) * (size_t)M);
    if (!crho || !cgrn || !temp || !ef || !acc) {
        free(crho); free(cgrn); free(temp); free(ef); free(acc);
        return -1;
    }
    const R zero3[3] = {0, 0, 0};
    for (long long e = 0; e < M; ++e) out[e] = (R)0;
    for (int c = 0; c < 3; ++c) {
        SFX(solve_freespace)(g + (size_t)c * M, ef, grid, delta, gamma, zero3, crho, cgrn, temp,
                             NULL, nthreads, c + 1, c + 1);
        for (long long e = 0; e < M; ++e) out[e] = out[e] - ef[e];
    }
    if (at_cathode) {
        const R off[3] = {0, 0, -((R)2 * lo[2] + (hi[2] - lo[2]))};
        for (long long e = 0; e < M; ++e) acc[e] = (R)0;
        for (int c = 0; c < 3; ++c) {
            SFX(solve_freespace)(g + (size_t)c * M, ef, grid, delta, gamma, off, crho, cgrn, temp,
                                 NULL, nthreads, c + 1, c + 1);
            for (long long e = 0; e < M; ++e) acc[e] = acc[e] + ef[e];
        }
        for (long long k = 0; k < nz; ++k)
            for (long long j = 0; j < ny; ++j)
                for (long long i = 0; i < nx; ++i)
                    out[i + nx * (j + ny * k)] += acc[i + nx * (j + ny * (nz - 1 - k))];
    }
    free(crho);
    free(cgrn);
    free(temp);
    free(ef);
    free(acc);
    return 0;
}

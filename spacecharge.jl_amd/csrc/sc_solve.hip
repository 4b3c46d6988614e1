// sc_solve.hip -- the free-space / cathode field solve on gfx950.
//
// Reference: solve! and solve_freespace! (src/solvers/free_space.jl:14-101).
// The reference zero-pads rho to (2nx,2ny,2nz), builds each G_c on the full
// padded lattice, runs 7 dense 3-D complex FFTs (14 with the cathode) and
// extracts the physical octant.  The same linear convolution is computed
// here with far less HBM traffic:
//
//  * Green spectra: sc_green.hip (symmetric transforms of the non-negative
//    octant, Float64 antiderivative and differencing).
//  * rho.  x transform of the real rows (kx in [0, nx] kept), then the y
//    transform; both skip the zero half of the padded input.
//  * z convolution, fused: per (kx, ky) pencil, forward z transform of rho
//    (and of the flipped, negated image charge), product with all three
//    Green spectra, backward z transform, only z < nz written.  The cathode
//    image adds its product before the same backward transform, so the
//    cathode solve costs one extra forward pencil transform, not a second
//    solve.
//  * backward y transform keeping y < ny, then the Hermitian (complex to
//    real) backward x transform keeping x < nx; FPEI, the cell-volume factor
//    and 1/(8 nx ny nz) are folded into the Green spectra.
// Only rows that can be non-zero are read, only outputs that are used are
// written.  Workspace: 4 spectra of nz*2ny*(nx+1) complex T, plus the Green
// spectra and their build scratch.
//
// The same kernels run the scalar potential (one Green spectrum, real) and
// the transpose of the field solve, g_rho = S^T g_E (the vector-Jacobian
// product of solve!): S_c is a real convolution, so S_c^T is the correlation
// with G_c, i.e. the product with conj(G_c^): the three components of g_E
// go through the forward x and y passes, the z convolution sums their
// products into one spectrum, and one inverse y, x pass gives g_rho.  The
// image term rho_img = -flip_z(rho) transposes to -flip_z of its correlation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sc_solve_common.h"

namespace sc {

namespace {

// ---------------------------------------------------------------------------
// x pass of the real input: rows (y, z) of component blockIdx.y, zero-padded
// to 2nx, kx in [0, nx] to that component's spectrum.
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_rho_fwd_x(const T *__restrict__ in, cpx<T> *__restrict__ out, const cpx<T> *__restrict__ twg,
            FftPlan pl, Dims d, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + d.Lx;
    const int Lx = d.Lx, nx = d.mx, nrows = d.my * d.mz;
    const T *rho = in + (size_t)blockIdx.y * nx * nrows;
    cpx<T> *A = out + (size_t)blockIdx.y * d.mz * d.Ly * d.Kp;
    load_twiddles(tw, twg, Lx);
    const int r0 = blockIdx.x * nb;
    for (int idx = threadIdx.x; idx < nb * Lx; idx += blockDim.x) {
        const int b = idx / Lx, e = idx - b * Lx;
        const int r = r0 + b;
        T v = (T)0;
        if (r < nrows && e < nx) v = rho[(size_t)r * nx + e];
        buf[idx] = cpx<T>{v, (T)0};
    }
    __syncthreads();
    lds_fft<T, E, NT>(buf, tw, pl, nb, 1, Lx, false, -1);
    for (int idx = threadIdx.x; idx < nb * d.Kx; idx += blockDim.x) {
        const int b = idx / d.Kx, kx = idx - b * d.Kx;
        const int r = r0 + b;
        if (r < nrows) {
            const int y = r % d.my, z = r / d.my;
            A[((size_t)z * d.Ly + y) * d.Kp + kx] = buf[b * Lx + kx];
        }
    }
}

// y pass: pencils (kx, z < nz) of component blockIdx.z along y, ny inputs
// zero-padded to 2ny.
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_rho_fwd_y(cpx<T> *__restrict__ A, const cpx<T> *__restrict__ twg, FftPlan pl, Dims d, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + d.Ly;
    const int Ly = d.Ly, Kx = d.Kx, Kp = d.Kp;
    load_twiddles(tw, twg, Ly);
    const int kx0 = blockIdx.x * nb, z = blockIdx.y;
    cpx<T> *Ap = A + ((size_t)blockIdx.z * d.mz + z) * Ly * Kp;
    for (int idx = threadIdx.x; idx < nb * Ly; idx += blockDim.x) {
        const int e = idx / nb, b = idx - e * nb;
        const int kx = kx0 + b;
        cpx<T> v{(T)0, (T)0};
        if (kx < Kx && e < d.my) v = Ap[(size_t)e * Kp + kx];
        buf[idx] = v;
    }
    __syncthreads();
    lds_fft<T, E, NT>(buf, tw, pl, nb, nb, 1, true, -1);
    for (int idx = threadIdx.x; idx < nb * Ly; idx += blockDim.x) {
        const int e = idx / nb, b = idx - e * nb;
        const int kx = kx0 + b;
        if (kx < Kx) Ap[(size_t)e * Kp + kx] = buf[idx];
    }
}

// Fused z convolution.  Per (kx, ky): forward z transform of rho,
// products with the three Green spectra (with the cathode also the image
// charge -rho[.., mz-1-z], whose spectrum follows from rho's), backward z
// transforms, z < nz of each field spectrum to B.  The potential (mode
// kPotential) has one Green spectrum, real and even in every axis.
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_zconv(const cpx<T> *__restrict__ A, cpx<T> *__restrict__ B, const T *__restrict__ Gf,
        const cpx<T> *__restrict__ Gi, const cpx<T> *__restrict__ twg, FftPlan pl, Dims d, int nb,
        int mode)
{
    const bool cathode = mode == kFieldCathode, pot = mode == kPotential;
    extern __shared__ __align__(16) unsigned char smem[];
    // twiddles in LDS, except on the widest workgroups (the longest
    // pencils), whose two pencil buffers fill the LDS: they read them from
    // global memory
    constexpr bool kLdsTw = NT == kFftThreads;
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem);
    const cpx<T> *tw = kLdsTw ? lds : twg;
    cpx<T> *P0 = kLdsTw ? lds + d.Lz : lds;
    cpx<T> *W = P0 + nb * d.Lz;
    // nz, ny: transform half-lengths (spectral octant); mz: mesh z nodes
    const int Lz = d.Lz, nz = d.nz, mz = d.mz, Kx = d.Kx, Kp = d.Kp, Ly = d.Ly, ny = d.ny;
    const int ny1 = ny + 1;
    if constexpr (kLdsTw) load_twiddles(lds, twg, Lz);
    const int kx0 = blockIdx.x * nb, ky = blockIdx.y;
    const size_t zs = (size_t)Ly * Kp;  // stride between z planes of A and B
    for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) {
        const int e = idx / nb, b = idx - e * nb;
        const int kx = kx0 + b;
        cpx<T> v{(T)0, (T)0};
        if (kx < Kx && e < mz) v = A[(size_t)e * zs + (size_t)ky * Kp + kx];
        P0[idx] = v;
    }
    __syncthreads();
    lds_fft<T, E, NT>(P0, tw, pl, nb, nb, 1, true, -1);
    const bool kyneg = ky > ny;
    const int kyp = kyneg ? Ly - ky : ky;
    const size_t M1 = (size_t)(nz + 1) * ny1 * Kp;  // one free-space Green spectrum
    const size_t MI = (size_t)Lz * ny1 * Kp;        // one image Green spectrum
    const int ncomp = pot ? 1 : 3;
    for (int c = 0; c < ncomp; ++c) {
        const T sy = (!pot && c == 1 && kyneg) ? (T)-1 : (T)1;
        for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) {
            const int kz = idx / nb, b = idx - kz * nb;
            const int kx = kx0 + b;
            cpx<T> out{(T)0, (T)0};
            if (kx < Kx) {
                const bool kzneg = kz > nz;
                const int kzp = kzneg ? Lz - kz : kz;
                const T sz = (!pot && c == 2 && kzneg) ? (T)-1 : (T)1;
                const T g = sy * sz * Gf[c * M1 + ((size_t)kzp * ny1 + kyp) * Kp + kx];
                const cpx<T> p = P0[idx];
                // rho_hat * (i g), or rho_hat * g for the potential
                out = pot ? cpx<T>{p.x * g, p.y * g} : cpx<T>{-p.y * g, p.x * g};
                if (cathode) {
                    cpx<T> q = Gi[c * MI + ((size_t)kz * ny1 + kyp) * Kp + kx];
                    q.x *= sy;
                    q.y *= sy;
                    if (c < 2) q = cpx<T>{-q.y, q.x};  // times i (odd in x or y)
                    // the image -rho(mz - 1 - z) (z < mz, zero-padded) has the
                    // spectrum -w_Lz^((mz-1) kz) P[-kz]: no transform of its own
                    const int km = kz == 0 ? 0 : Lz - kz;
                    const cpx<T> r = c_mul(tw[(int)(((int64_t)(mz - 1) * kz) % Lz)], P0[km * nb + b]);
                    out = c_add(out, c_mul(cpx<T>{-r.x, -r.y}, q));
                }
            }
            W[idx] = out;
        }
        __syncthreads();
        lds_fft<T, E, NT>(W, tw, pl, nb, nb, 1, true, +1);
        cpx<T> *Bc = B + (size_t)c * mz * zs;
        for (int idx = threadIdx.x; idx < nb * mz; idx += blockDim.x) {
            const int z = idx / nb, b = idx - z * nb;
            const int kx = kx0 + b;
            if (kx < Kx) Bc[(size_t)z * zs + (size_t)ky * Kp + kx] = W[idx];
        }
        __syncthreads();
    }
}

// Transposed z convolution.  Per (kx, ky): forward z transform of each of
// the three g_E spectra in Gs, product with conj(G_c^) summed over c, and
// with the cathode minus the image sum Wi = sum_c p conj(Gi_c^) reversed in
// z (out[z] = W[z] - Wi[mz-1-z] after the backward transform, formed in the
// spectrum), one backward z transform, z < mz to out.
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_zconv_adj(const cpx<T> *__restrict__ Gs, cpx<T> *__restrict__ out, const T *__restrict__ Gf,
            const cpx<T> *__restrict__ Gi, const cpx<T> *__restrict__ twg, FftPlan pl, Dims d,
            int nb, int cathode)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // twiddles as in k_zconv (global memory on the widest workgroups)
    constexpr bool kLdsTw = NT == kFftThreads;
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem);
    const cpx<T> *tw = kLdsTw ? lds : twg;
    cpx<T> *Q = kLdsTw ? lds + d.Lz : lds;
    cpx<T> *W = Q + nb * d.Lz;
    // nz, ny: transform half-lengths (spectral octant); mz: mesh z nodes
    const int Lz = d.Lz, nz = d.nz, mz = d.mz, Kx = d.Kx, Kp = d.Kp, Ly = d.Ly, ny = d.ny;
    const int ny1 = ny + 1;
    if constexpr (kLdsTw) load_twiddles(lds, twg, Lz);
    const int kx0 = blockIdx.x * nb, ky = blockIdx.y;
    const size_t zs = (size_t)Ly * Kp;
    const bool kyneg = ky > ny;
    const int kyp = kyneg ? Ly - ky : ky;
    const size_t M1 = (size_t)(nz + 1) * ny1 * Kp;
    const size_t MI = (size_t)Lz * ny1 * Kp;
    for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) W[idx] = cpx<T>{(T)0, (T)0};
    for (int c = 0; c < 3; ++c) {
        const cpx<T> *Gc = Gs + (size_t)c * mz * zs;
        for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) {
            const int e = idx / nb, b = idx - e * nb;
            const int kx = kx0 + b;
            cpx<T> v{(T)0, (T)0};
            if (kx < Kx && e < mz) v = Gc[(size_t)e * zs + (size_t)ky * Kp + kx];
            Q[idx] = v;
        }
        __syncthreads();
        lds_fft<T, E, NT>(Q, tw, pl, nb, nb, 1, true, -1);
        const T sy = (c == 1 && kyneg) ? (T)-1 : (T)1;
        for (int idx = threadIdx.x; idx < nb * Lz; idx += blockDim.x) {
            const int kz = idx / nb, b = idx - kz * nb;
            const int kx = kx0 + b;
            if (kx < Kx) {
                const bool kzneg = kz > nz;
                const int kzp = kzneg ? Lz - kz : kz;
                const T sz = (c == 2 && kzneg) ? (T)-1 : (T)1;
                const T g = sy * sz * Gf[c * M1 + ((size_t)kzp * ny1 + kyp) * Kp + kx];
                const cpx<T> p = Q[idx];
                // p * conj(i g) = p * (-i g)
                cpx<T> acc = c_add(W[idx], cpx<T>{p.y * g, -p.x * g});
                if (cathode) {
                    // the image sum W1 = sum_c p conj(q), reversed in z after its
                    // backward transform, is the backward transform of
                    // U[k] = W1[-k] w_Lz^((mz-1) k): its terms, taken at the
                    // mirror bin, go into W (one backward transform for both)
                    const int km = kz == 0 ? 0 : Lz - kz;
                    cpx<T> q = Gi[c * MI + ((size_t)km * ny1 + kyp) * Kp + kx];
                    q.x *= sy;
                    q.y *= sy;
                    if (c < 2) q = cpx<T>{-q.y, q.x};
                    const cpx<T> t = c_mul(Q[km * nb + b], cpx<T>{q.x, -q.y});
                    acc = c_sub(acc, c_mul(tw[(int)(((int64_t)(mz - 1) * kz) % Lz)], t));
                }
                W[idx] = acc;
            }
        }
        __syncthreads();
    }
    lds_fft<T, E, NT>(W, tw, pl, nb, nb, 1, true, +1);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nb * mz; idx += blockDim.x) {
        const int z = idx / nb, b = idx - z * nb;
        const int kx = kx0 + b;
        if (kx < Kx) out[(size_t)z * zs + (size_t)ky * Kp + kx] = W[idx];
    }
}

// Backward y pass: pencils (kx, z < nz) of each field spectrum, y < ny kept.
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_inv_y(cpx<T> *__restrict__ B, const cpx<T> *__restrict__ twg, FftPlan pl, Dims d, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + d.Ly;
    const int Ly = d.Ly, Kx = d.Kx, Kp = d.Kp;
    load_twiddles(tw, twg, Ly);
    const int kx0 = blockIdx.x * nb, z = blockIdx.y, c = blockIdx.z;
    cpx<T> *Bp = B + ((size_t)c * d.mz + z) * Ly * Kp;
    for (int idx = threadIdx.x; idx < nb * Ly; idx += blockDim.x) {
        const int e = idx / nb, b = idx - e * nb;
        const int kx = kx0 + b;
        cpx<T> v{(T)0, (T)0};
        if (kx < Kx) v = Bp[(size_t)e * Kp + kx];
        buf[idx] = v;
    }
    __syncthreads();
    lds_fft<T, E, NT>(buf, tw, pl, nb, nb, 1, true, +1);
    for (int idx = threadIdx.x; idx < nb * d.my; idx += blockDim.x) {
        const int e = idx / nb, b = idx - e * nb;
        const int kx = kx0 + b;
        if (kx < Kx) Bp[(size_t)e * Kp + kx] = buf[idx];
    }
}

// Backward x pass (complex to real): rows (y < ny, z < nz) of each field
// spectrum, Hermitian completion, x < nx of the real part to efield[c].
template <typename T, int E = fft_elems<T>(), int NT = kFftThreads>
__global__ void __launch_bounds__(NT)
k_inv_x(const cpx<T> *__restrict__ B, T *__restrict__ efield, const cpx<T> *__restrict__ twg, FftPlan pl,
        Dims d, int nb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem);
    cpx<T> *buf = tw + d.Lx;
    // hx: transform half-length (Hermitian completion); nx: mesh x nodes
    const int Lx = d.Lx, hx = d.nx, nx = d.mx, Kx = d.Kx, Kp = d.Kp, nrows = d.my * d.mz;
    load_twiddles(tw, twg, Lx);
    const int r0 = blockIdx.x * nb, c = blockIdx.y;
    const cpx<T> *Bc = B + (size_t)c * d.mz * d.Ly * Kp;
    for (int idx = threadIdx.x; idx < nb * Kx; idx += blockDim.x) {
        const int b = idx / Kx, kx = idx - b * Kx;
        const int r = r0 + b;
        cpx<T> v{(T)0, (T)0};
        if (r < nrows) {
            const int y = r % d.my, z = r / d.my;
            v = Bc[((size_t)z * d.Ly + y) * Kp + kx];
        }
        buf[b * Lx + kx] = v;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nb * (hx - 1); idx += blockDim.x) {
        const int b = idx / (hx - 1), k = idx - b * (hx - 1) + 1;
        const cpx<T> v = buf[b * Lx + k];
        buf[b * Lx + Lx - k] = cpx<T>{v.x, -v.y};
    }
    __syncthreads();
    lds_fft<T, E, NT>(buf, tw, pl, nb, 1, Lx, false, +1);
    T *Ec = efield + (size_t)c * nx * nrows;
    for (int idx = threadIdx.x; idx < nb * nx; idx += blockDim.x) {
        const int b = idx / nx, x = idx - b * nx;
        const int r = r0 + b;
        if (r < nrows) Ec[(size_t)r * nx + x] = buf[b * Lx + x].x;
    }
}

template <typename T>
int upload_twiddles(Ctx *ctx, int n, void **dst)
{
    std::vector<cpx<T>> h(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; ++m) {
        const long double a = (2.0L * pi * (long double)m) / (long double)n;
        h[m].x = (T)cosl(a);
        h[m].y = (T)(-sinl(a));
    }
    SC_HIP(ctx, hipMalloc(dst, sizeof(cpx<T>) * n));
    SC_HIP(ctx, hipMemcpy(*dst, h.data(), sizeof(cpx<T>) * n, hipMemcpyHostToDevice));
    return SC_OK;
}

template <typename T>
int run_solve(Mesh *m, const double lo[3], const double hi[3], const double dl[3], double gamma,
              const SolveIO &io)
{
    const int mode = io.mode;
    const int gmode = green_mode(mode);
    const int cath = mode_cathode(mode);
    Ctx *ctx = m->ctx;
    Workspace &ws = m->ws;
    hipStream_t st = ctx->stream;
    const Dims d = make_dims(m);
    const cpx<T> *twx = static_cast<const cpx<T> *>(ws.twT[0]);
    const cpx<T> *twy = static_cast<const cpx<T> *>(ws.twT[1]);
    const cpx<T> *twz = static_cast<const cpx<T> *>(ws.twT[2]);

    // ---- Green spectra ----
    // dl == NULL: the geometry lives on the device (sc_mesh_auto_bounds) and
    // the Green kernels read it there; the host cannot key a cache on it
    const bool devgeom = dl == nullptr;
    const double offz = devgeom ? 0.0 : lo[2] + hi[2];  // 2*zmin + (zmax - zmin), free_space.jl:39
    double key[6] = {0.0, 0.0, 0.0, gamma, 0.0, (double)gmode};
    if (!devgeom) {
        for (int a = 0; a < 3; ++a) key[a] = dl[a];
        key[4] = cath ? offz : 0.0;
    }
    const bool reuse = !devgeom && m->green_reuse && ws.g_valid && ws.g_cath == gmode &&
                       std::memcmp(key, ws.g_key, sizeof(key)) == 0;
    // The Green spectra depend on the spacing, gamma and the cathode offset
    // only, not on rho: built on the context's side stream they overlap the
    // deposit kernels still running on `st`.  Under stream capture the side
    // stream joins the capture through the fork event and rejoins `st`
    // before the z convolution, so the graph keeps the overlap; it must
    // exist already (no stream is created while capturing).
    unsigned long long cid = 0;
    const bool cap = stream_capturing(st, &cid);
    bool overlap = false;
    if (!reuse) {
        overlap = !ctx->in_batch &&
                  (cap ? ctx->side && ws.green_done && ws.deposit_start : ensure_side_stream(m) == SC_OK);
        if (overlap) {
            // the build forks where this step's deposit started (recorded by
            // a deposit of this capture, or by an eager deposit; an older
            // eager record only if no capture has used this mesh), else at
            // the current end of `st`, which may still read this mesh's
            // Green spectra
            const bool from_deposit =
                cap ? ws.fork_capture == cid : (ws.fork_ready || (ws.start_recorded && !ws.captured));
            if (!from_deposit) {
                SC_HIP(ctx, hipEventRecord(ws.deposit_start, st));
                if (!cap) ws.start_recorded = true;
            }
            if (cap) ws.captured = true;
            ws.fork_ready = false;
            ws.fork_capture = 0;
            // an eager build also waits for the last eager z convolution,
            // which may have run on another stream (a graph is ordered by its
            // launch stream instead)
            if (!cap && ws.gf_free_recorded)
                SC_HIP(ctx, hipStreamWaitEvent(ctx->side, ws.gf_free, 0));
            SC_HIP(ctx, hipStreamWaitEvent(ctx->side, ws.deposit_start, 0));
            ctx->stream = ctx->side;  // the build launches (and times) on ctx->stream
            const int rc = build_green<T>(m, dl, gamma, offz, gmode);
            ctx->stream = st;
            if (rc) return rc;
            SC_HIP(ctx, hipEventRecord(ws.green_done, ctx->side));
        } else {
            if (int rc = build_green<T>(m, dl, gamma, offz, gmode)) return rc;
        }
        ws.g_valid = !devgeom;
        ws.g_cath = gmode;
        std::memcpy(ws.g_key, key, sizeof(key));
    }

    // Each transform kernel runs on the register-resident engine when its
    // length is a supported power of two, else on the generic LDS engine.
    const bool fast_ok = !m->generic_transforms && !dims_padded(d);
    auto fast = [&](int which) { return fast_ok && launch_fast<T>(m, which, io); };

    // ---- forward x, y ----
    cpx<T> *F = static_cast<cpx<T> *>(io.fwd);
    cpx<T> *B = static_cast<cpx<T> *>(io.inv);
    if (!fast(0)) {
        with_fft_elems<T>(d.Lx, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Lx, E * NT, 16);
            const size_t sh = sizeof(cpx<T>) * (d.Lx + (size_t)nb * d.Lx);
            launch_begin(ctx, K_RHO_FWD_X);
            hipLaunchKernelGGL((k_rho_fwd_x<T, E, NT>), dim3(div_up(d.my * d.mz, nb), io.nin),
                               dim3(NT), sh, st, static_cast<const T *>(io.in), F, twx,
                               ws.plan[0], d, nb);
            launch_end(ctx, K_RHO_FWD_X);
        });
    }
    if (!fast(1)) {
        with_fft_elems<T>(d.Ly, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Ly, E * NT, 16);
            const size_t sh = sizeof(cpx<T>) * (d.Ly + (size_t)nb * d.Ly);
            launch_begin(ctx, K_RHO_FWD_Y);
            hipLaunchKernelGGL((k_rho_fwd_y<T, E, NT>), dim3(div_up(d.Kx, nb), d.mz, io.nin),
                               dim3(NT), sh, st, F, twy, ws.plan[1], d, nb);
            launch_end(ctx, K_RHO_FWD_Y);
        });
    }
    // ---- fused z convolution ----
    if (overlap) SC_HIP(ctx, hipStreamWaitEvent(st, ws.green_done, 0));
    if (!fast(2)) {
        with_fft_elems<T>(d.Lz, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            // LDS buffers of nb pencils: the spectrum and the product, after
            // the twiddles (read from global memory on the widest workgroups)
            const int nbuf = 2;
            const size_t ntw = NT == kFftThreads ? (size_t)d.Lz : 0;
            // at most 64 KiB of LDS where more pencils fit (occupancy); one
            // long pencil may take more (up to the device's 160 KiB)
            int nb = pick_nb(d.Lz, E * NT, 16);
            while (nb > 1 && sizeof(cpx<T>) * (ntw + (size_t)nbuf * nb * d.Lz) > 65536) nb /= 2;
            const size_t sh = sizeof(cpx<T>) * (ntw + (size_t)nbuf * nb * d.Lz);
            launch_begin(ctx, K_ZCONV);
            if (mode_adjoint(mode))
                hipLaunchKernelGGL((k_zconv_adj<T, E, NT>), dim3(div_up(d.Kx, nb), d.Ly), dim3(NT),
                                   sh, st, (const cpx<T> *)F, B, static_cast<const T *>(ws.Gf),
                                   static_cast<const cpx<T> *>(ws.Gi), twz, ws.plan[2], d, nb, cath);
            else
                hipLaunchKernelGGL((k_zconv<T, E, NT>), dim3(div_up(d.Kx, nb), d.Ly), dim3(NT), sh,
                                   st, (const cpx<T> *)F, B, static_cast<const T *>(ws.Gf),
                                   static_cast<const cpx<T> *>(ws.Gi), twz, ws.plan[2], d, nb, mode);
            launch_end(ctx, K_ZCONV);
        });
    }
    if (ws.green_done && !cap) {  // the side stream exists: the next build waits for this point
        SC_HIP(ctx, hipEventRecord(ws.gf_free, st));
        ws.gf_free_recorded = true;
    }
    // ---- backward y, x ----
    if (!fast(3)) {
        with_fft_elems<T>(d.Ly, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Ly, E * NT, 16);
            const size_t sh = sizeof(cpx<T>) * (d.Ly + (size_t)nb * d.Ly);
            launch_begin(ctx, K_INV_Y);
            hipLaunchKernelGGL((k_inv_y<T, E, NT>), dim3(div_up(d.Kx, nb), d.mz, io.nout), dim3(NT),
                               sh, st, B, twy, ws.plan[1], d, nb);
            launch_end(ctx, K_INV_Y);
        });
    }
    if (!fast(4)) {
        with_fft_elems<T>(d.Lx, [&](auto e, auto t) {
            constexpr int E = decltype(e)::value, NT = decltype(t)::value;
            const int nb = pick_nb(d.Lx, E * NT, 16);
            const size_t sh = sizeof(cpx<T>) * (d.Lx + (size_t)nb * d.Lx);
            launch_begin(ctx, K_INV_X);
            hipLaunchKernelGGL((k_inv_x<T, E, NT>), dim3(div_up(d.my * d.mz, nb), io.nout),
                               dim3(NT), sh, st, (const cpx<T> *)B, static_cast<T *>(io.out),
                               twx, ws.plan[0], d, nb);
            launch_end(ctx, K_INV_X);
        });
    }
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

// Runs the solve described by `io` (mode, components, real input and
// output); the spectra buffers come from the mesh's workspace: A holds the
// single-component spectrum and B the three-component ones.
int run_solve_io(Mesh *m, const double lo[3], const double hi[3], const double dl[3], double gamma,
                 SolveIO io)
{
    int rc = ensure_workspace(m);
    if (rc != SC_OK) return rc;
    const bool adj = mode_adjoint(io.mode);
    io.fwd = adj ? m->ws.B : m->ws.A;
    io.inv = adj ? m->ws.A : m->ws.B;
    if (m->dtype == SC_F32) return run_solve<float>(m, lo, hi, dl, gamma, io);
    return run_solve<double>(m, lo, hi, dl, gamma, io);
}

}  // namespace

bool make_plan(int n, FftPlan *p)
{
    p->n = n;
    p->nst = 0;
    int m = n;
    while (m % 8 == 0) { p->rad[p->nst++] = 8; m /= 8; }
    while (m % 4 == 0) { p->rad[p->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p->rad[p->nst++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) {
            if (f > kFftMaxRadix || p->nst >= kMaxStages) return false;
            p->rad[p->nst++] = f;
            m /= f;
        }
    }
    return true;
}

bool stream_capturing(hipStream_t st, unsigned long long *id)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    unsigned long long cid = 0;
    if (hipStreamGetCaptureInfo(st, &cs, &cid) != hipSuccess) cs = hipStreamCaptureStatusNone;
    const bool active = cs == hipStreamCaptureStatusActive;
    *id = active ? cid : 0ull;
    return active;
}

int mark_deposit_start(Mesh *m)
{
    Workspace &ws = m->ws;
    if (!ws.deposit_start) return SC_OK;
    hipStream_t st = m->ctx->stream;
    unsigned long long cid = 0;
    const bool cap = stream_capturing(st, &cid);
    SC_HIP(m->ctx, hipEventRecord(ws.deposit_start, st));
    ws.fork_ready = !cap;
    ws.fork_capture = cid;
    if (cap) ws.captured = true;
    else ws.start_recorded = true;
    return SC_OK;
}

int ensure_side_stream(Mesh *m)
{
    Ctx *ctx = m->ctx;
    Workspace &ws = m->ws;
    if (!ctx->side && hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
        ctx->side = nullptr;
        return SC_ERR_DEVICE;
    }
    if (!ws.green_done && hipEventCreateWithFlags(&ws.green_done, hipEventDisableTiming) != hipSuccess) {
        ws.green_done = nullptr;
        return SC_ERR_DEVICE;
    }
    if (!ws.gf_free && hipEventCreateWithFlags(&ws.gf_free, hipEventDisableTiming) != hipSuccess) {
        ws.gf_free = nullptr;
        return SC_ERR_DEVICE;
    }
    if (!ws.deposit_start &&
        hipEventCreateWithFlags(&ws.deposit_start, hipEventDisableTiming) != hipSuccess) {
        ws.deposit_start = nullptr;
        return SC_ERR_DEVICE;
    }
    return SC_OK;
}

void free_workspace(Mesh *m)
{
    Workspace &ws = m->ws;
    // a Green build may still be running on the side stream
    if (m->ctx && m->ctx->side) (void)hipStreamSynchronize(m->ctx->side);
    if (ws.green_done) (void)hipEventDestroy(ws.green_done);
    if (ws.gf_free) (void)hipEventDestroy(ws.gf_free);
    if (ws.deposit_start) (void)hipEventDestroy(ws.deposit_start);
    void *ptrs[] = {ws.A, ws.B, ws.Gf, ws.Gi, ws.F, ws.R, ws.twT[0], ws.twT[1], ws.twT[2]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ws = Workspace();
}

int ensure_workspace(Mesh *m)
{
    Workspace &ws = m->ws;
    if (ws.ready) return SC_OK;
    Ctx *ctx = m->ctx;
    const Dims d = make_dims(m);
    // transform half-lengths (padded beyond the mesh's nodes where 2n is not 7-smooth)
    const int64_t nx = d.nx, ny = d.ny, nz = d.nz;
    const int L[3] = {d.Lx, d.Ly, d.Lz};
    const int64_t mn[3] = {m->n[0], m->n[1], m->n[2]};
    const int cap = (m->dtype == SC_F32) ? fft_cap_xl<float>() : fft_cap_xl<double>();
    for (int a = 0; a < 3; ++a) {
        // one pencil must fit one workgroup (registers and LDS)
        if (L[a] > cap)
            return set_err(ctx, SC_ERR_UNSUPPORTED,
                           "grid size " + std::to_string(mn[a]) + " needs a transform of length " +
                               std::to_string(L[a]) + ", beyond the maximum of " +
                               std::to_string(cap) + " for this mesh type (at most " +
                               std::to_string(cap / 2) + " nodes per axis)");
        if (!make_plan(L[a], &ws.plan[a]))  // unreachable: transform_half gives 7-smooth lengths
            return set_err(ctx, SC_ERR_UNSUPPORTED,
                           "no transform plan for length " + std::to_string(L[a]));
    }
    const size_t s = (m->dtype == SC_F32) ? 4 : 8;
    const size_t Kp = d.Kp;  // row stride of every kx-indexed array
    // complex elements of one spectrum: [mesh z][Ly][Kp]
    const size_t spec = (size_t)d.mz * d.Ly * Kp;
    const size_t gf = (size_t)(nz + 1) * (ny + 1) * Kp;
    const size_t gi = (size_t)2 * nz * (ny + 1) * Kp;
    // Green build scratch: F at the nodes (Float64), R rows (T), free + image
    const size_t fn = (size_t)(nx + 1) * (ny + 1) * ((nz + 1) + 2 * nz);
    const size_t rn = (size_t)(ny + 1) * Kp * (nz + (2 * nz - 1));
    auto fail = [&](void) {
        free_workspace(m);
        return set_err(ctx, SC_ERR_NOMEM, "device allocation of the solver workspace failed");
    };
    if (hipMalloc(&ws.A, 2 * s * spec) != hipSuccess) return fail();
    if (hipMalloc(&ws.B, 3 * 2 * s * spec) != hipSuccess) return fail();
    if (hipMalloc(&ws.Gf, 3 * s * gf) != hipSuccess) return fail();
    if (hipMalloc(&ws.Gi, 3 * 2 * s * gi) != hipSuccess) return fail();
    if (hipMalloc((void **)&ws.F, sizeof(double) * 3 * fn) != hipSuccess) return fail();
    if (hipMalloc(&ws.R, s * 3 * rn) != hipSuccess) return fail();
    ws.bytes = (int64_t)(8 * s * spec + 3 * s * gf + 6 * s * gi + 24 * fn + 3 * s * rn);
    for (int a = 0; a < 3; ++a) {
        int rc = (m->dtype == SC_F32) ? upload_twiddles<float>(ctx, L[a], &ws.twT[a])
                                      : upload_twiddles<double>(ctx, L[a], &ws.twT[a]);
        if (rc != SC_OK) { free_workspace(m); return rc; }
    }
    ws.dtype = m->dtype;
    ws.ready = true;
    return SC_OK;
}

int solve_mesh(Mesh *m, const double lo[3], const double hi[3], const double dl[3], double gamma,
               int at_cathode)
{
    SolveIO io{at_cathode ? kFieldCathode : kFieldFree, 1, 3, m->rho, m->efield, nullptr, nullptr};
    return run_solve_io(m, lo, hi, dl, gamma, io);
}

int solve_potential_mesh(Mesh *m, const double lo[3], const double hi[3], const double dl[3],
                         double gamma, void *phi)
{
    SolveIO io{kPotential, 1, 1, m->rho, phi, nullptr, nullptr};
    return run_solve_io(m, lo, hi, dl, gamma, io);
}

int solve_adjoint_mesh(Mesh *m, const double lo[3], const double hi[3], const double dl[3],
                       double gamma, int at_cathode, const void *g_efield, void *g_rho)
{
    SolveIO io{at_cathode ? kAdjointCathode : kAdjointFree, 3, 1, g_efield, g_rho, nullptr, nullptr};
    return run_solve_io(m, lo, hi, dl, gamma, io);
}


// ---- grouped solves: same-shape bunches of a batched solve ----
// The bunches of one class (mesh shape, type, cathode flag; register-engine
// lengths) run each Green and transform kernel once for all of them, with
// their buffers and Green geometry from a table in the context's device
// memory, written by store launches that carry the entries as arguments
// (so a captured graph rewrites them on every replay).  Each class owns a
// disjoint range of the table, so classes may run on different streams.

namespace {

struct SStore {
    int first, count;
    SBunch b[kGroupStore];
};
static_assert(sizeof(SStore) <= 2048, "store-launch arguments stay small");

// one lane per 8-byte word (vector stores)
__global__ void __launch_bounds__(256) k_sgroup_store(SStore s, SBunch *__restrict__ tab)
{
    constexpr int kWords = (int)(sizeof(SBunch) / 8);
    static_assert(sizeof(SBunch) % 8 == 0, "entry of whole words");
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(s.b);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(tab + s.first);
    for (int i = threadIdx.x; i < s.count * kWords; i += blockDim.x) dst[i] = src[i];
}

bool reg_length(int L) { return L == 32 || L == 64 || L == 128 || L == 256 || L == 512; }

template <typename T>
int solve_class_t(Ctx *ctx, Mesh *const *ms, int k, const double *const *lo, const double *const *hi,
                  const double *const *dl, const double *gamma, int cath, SBunch *tab)
{
    hipStream_t st = ctx->stream;
    unsigned long long cid = 0;
    const int mode = cath ? kFieldCathode : kFieldFree;
    for (int i = 0; i < k; ++i)
        if (int rc = ensure_workspace(ms[i])) return rc;
    const Dims d = make_dims(ms[0]);
    const double P = (double)d.Lx * (double)d.Ly * (double)d.Lz;
    SStore s;
    std::memset(&s, 0, sizeof(s));
    for (int b0 = 0; b0 < k; b0 += kGroupStore) {
        s.first = b0;
        s.count = k - b0 < kGroupStore ? k - b0 : kGroupStore;
        for (int j = 0; j < s.count; ++j) {
            const int i = b0 + j;
            Mesh *m = ms[i];
            Workspace &ws = m->ws;
            SBunch &e = s.b[j];
            e.in = m->rho;
            e.out = m->efield;
            e.fwd = ws.A;
            e.inv = ws.B;
            e.Gf = ws.Gf;
            e.Gi = ws.Gi;
            e.R = ws.R;
            e.F = ws.F;
            // the expressions of build_green and run_solve
            e.dx = dl[i][0];
            e.dy = dl[i][1];
            e.dz = dl[i][2] * gamma[i];
            e.offzg = (lo[i][2] + hi[i][2]) * gamma[i];
            const double vol = (e.dx * e.dy) * e.dz;
            for (int c = 0; c < 3; ++c) e.scale[c] = kFPEI * ((c < 2 ? gamma[i] : 1.0) / vol) / P;
            e.scale[3] = kFPEI * (gamma[i] / vol) / P;
        }
        hipLaunchKernelGGL(k_sgroup_store, dim3(1), dim3(256), 0, st, s, tab);
    }
    launch_begin(ctx, K_GREEN_F);
    build_green_grouped<T>(ms[0], k, tab, 1.0, mode, st);
    launch_end(ctx, K_GREEN_F);
    static const int kid[5] = {K_RHO_FWD_X, K_RHO_FWD_Y, K_ZCONV, K_INV_Y, K_INV_X};
    for (int which = 0; which < 5; ++which) {
        launch_begin(ctx, kid[which]);
        launch_fast_grouped<T>(ms[0], k, tab, which, mode, 1, 3, st);
        launch_end(ctx, kid[which]);
        if (which == 2 && !stream_capturing(st, &cid)) {
            // a later single solve's side-stream Green build waits for this
            for (int i = 0; i < k; ++i) {
                Workspace &ws = ms[i]->ws;
                if (ws.gf_free) {
                    SC_HIP(ctx, hipEventRecord(ws.gf_free, st));
                    ws.gf_free_recorded = true;
                }
            }
        }
    }
    for (int i = 0; i < k; ++i) ms[i]->ws.g_valid = false;  // the reuse cache was not kept
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

}  // namespace

bool solve_groupable(const Mesh *m)
{
    if (m->generic_transforms || m->green_reuse) return false;
    const Dims d = make_dims(m);
    return !dims_padded(d) && reg_length(d.Lx) && reg_length(d.Ly) && reg_length(d.Lz);
}

int64_t solve_group_bytes(const Mesh *m)
{
    const Dims d = make_dims(m);
    const int64_t s = m->dtype == SC_F32 ? 4 : 8;
    // four spectra (rho and three fields) of [mesh z][Ly][Kp] complex
    return 4 * 2 * s * (int64_t)d.mz * d.Ly * d.Kp;
}

int ensure_solve_table(Ctx *ctx)
{
    if (ctx->stable) return SC_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SC_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return SC_ERR_UNSUPPORTED;
    if (hipMalloc(&ctx->stable, sizeof(SBunch) * kGroupMax) != hipSuccess) {
        ctx->stable = nullptr;
        return set_err(ctx, SC_ERR_NOMEM, "device allocation of the batch descriptor table failed");
    }
    return SC_OK;
}

int solve_class(Ctx *ctx, Mesh *const *ms, int k, const double *const *lo, const double *const *hi,
                const double *const *dl, const double *gamma, int cath, int slot)
{
    SBunch *tab = static_cast<SBunch *>(ctx->stable) + slot;
    if (ms[0]->dtype == SC_F32) return solve_class_t<float>(ctx, ms, k, lo, hi, dl, gamma, cath, tab);
    return solve_class_t<double>(ctx, ms, k, lo, hi, dl, gamma, cath, tab);
}

}  // namespace sc

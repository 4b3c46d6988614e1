// sc_solve_fast.hip -- register-resident transform kernels of the field solve
// for power-of-two padded lengths (32..1024).
//
// Same data flow and layouts as the generic kernels in sc_solve.hip (which
// remain the path for other lengths), with every pencil held in registers:
// a lane loads its E elements straight from HBM, the pencil is transformed
// with one LDS exchange per pass boundary, and the lane stores straight to
// HBM.  Which inputs are zero (the padded half) and which outputs are kept
// (the physical octant, the Hermitian half) depends only on the register
// index, so the pruning is resolved at compile time.
#include <hip/hip_runtime.h>

#include "sc_fft_reg.h"
#include "sc_solve_common.h"

namespace sc {

namespace {

constexpr int kThreads = 256;

template <typename T, int L, int THREADS = kThreads>
struct Shape {
    static constexpr int E = reg::elems<T>() < L ? reg::elems<T>() : L;
    static constexpr int TPP = L / E;         // lanes per pencil
    static constexpr int NB = THREADS / TPP;  // pencils per workgroup
};

// Workgroup size of the z convolution.  (512 threads, which make each
// workgroup's Green row segment a whole 128-byte line, cut its fetched bytes
// by a quarter on MI355X but ran slower: the kernel is bound by its
// transform arithmetic, not by memory.)
template <typename T, int L>
constexpr int zconv_threads()
{
    return kThreads;
}

// x pass of the real input: rows (y, z) of component blockIdx.y, zero-padded
// to 2nx, kx in [0, nx] to that component's spectrum.
template <typename T, int L>
__global__ void __launch_bounds__(kThreads)
k_rho_fwd_x_r(const T *__restrict__ in, cpx<T> *__restrict__ spec, const cpx<T> *__restrict__ twg,
              Dims d, const SBunch *__restrict__ tab, int per)
{
    using S = Shape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * reg::padded<L, E>()];
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) {
        in = static_cast<const T *>(sb->in);
        spec = static_cast<cpx<T> *>(sb->fwd);
    }
    for (int i = threadIdx.x; i < L; i += kThreads) tw[i] = gld(twg + i);
    const int b = threadIdx.x / TPP, j = threadIdx.x % TPP;
    const int r = bx * NB + b;
    const int nrows = d.ny * d.nz, nx = d.nx;
    const T *rho = in + (size_t)blockIdx.y * nx * nrows;
    cpx<T> *A = spec + (size_t)blockIdx.y * d.nz * d.Ly * d.Kp;
    const bool live = r < nrows;
    cpx<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        v[q] = cpx<T>{(T)0, (T)0};
        if (q < E / 2 && live) v[q].x = gld(rho + (size_t)r * nx + j + q * TPP);
    }
    reg::fft<T, L, E, -1>(v, j, lds + b * reg::padded<L, E>(), 1, tw);
    if (live) {
        const int y = r % d.ny, z = r / d.ny;
        cpx<T> *out = A + ((size_t)z * d.Ly + y) * d.Kp;
#pragma unroll
        for (int q = 0; q <= E / 2; ++q) {
            const int k = j + q * TPP;
            if (q < E / 2 || k == nx) gst(out + k, v[q]);
        }
    }
}

// y pass: pencils (kx, z < nz) of component blockIdx.z along y, ny inputs
// zero-padded to 2ny.
template <typename T, int L>
__global__ void __launch_bounds__(kThreads)
k_rho_fwd_y_r(cpx<T> *__restrict__ A, const cpx<T> *__restrict__ twg, Dims d,
              const SBunch *__restrict__ tab, int per)
{
    using S = Shape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * L];
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) A = static_cast<cpx<T> *>(sb->fwd);
    for (int i = threadIdx.x; i < L; i += kThreads) tw[i] = gld(twg + i);
    const int b = threadIdx.x % NB, j = threadIdx.x / NB;
    const int kx = bx * NB + b, z = blockIdx.y;
    const int Kx = d.Kx, Kp = d.Kp;
    const bool live = kx < Kx;
    cpx<T> *Ap = A + ((size_t)blockIdx.z * d.nz + z) * d.Ly * Kp + kx;
    cpx<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        v[q] = cpx<T>{(T)0, (T)0};
        if (q < E / 2 && live) v[q] = gld(Ap + (size_t)(j + q * TPP) * Kp);
    }
    reg::fft<T, L, E, -1>(v, j, lds + b, NB, tw);
    if (live) {
#pragma unroll
        for (int q = 0; q < E; ++q) gst(Ap + (size_t)(j + q * TPP) * Kp, v[q]);
    }
}

// Fused z convolution (see k_zconv in sc_solve.hip for the algebra), MODE
// a SolveMode.  Free space: three waves per SIMD; the forward pencil stays
// in registers across the three inverse transforms, and unconstrained the
// kernel takes 169 VGPRs (two waves); 168 fit without spilling.  With the
// cathode image the z-reversed pencil's spectrum is held as well (formed
// from the forward spectrum, not transformed), which does not fit in 168
// VGPRs (held to it, the compiler spilled 25-62 VGPRs to scratch), so that
// mode runs two waves per SIMD.
template <typename T, int L, int MODE, int THREADS>
__global__ void __launch_bounds__(THREADS)
__attribute__((amdgpu_waves_per_eu(MODE == kFieldCathode ? 2 : 3)))
k_zconv_r(const cpx<T> *__restrict__ A, cpx<T> *__restrict__ B, const T *__restrict__ Gf,
          const cpx<T> *__restrict__ Gi, const cpx<T> *__restrict__ twg, Dims d,
          const SBunch *__restrict__ tab, int per)
{
    // grouped launches and the cathode mode: grid (x blocks [* bunches], Ly);
    // one bunch in free space or for the potential: a 1-D grid of per * Ly
    // workgroups in XCD order, per x blocks per row (the cathode mode, at two
    // waves per SIMD, measured 2 us slower in that order in Float64)
    const SBunch *sb = nullptr;
    int bx, ky;
    if (tab) {
        bx = group_block(tab, per, sb);
        ky = (int)blockIdx.y;
    } else if (MODE == kFieldCathode) {
        bx = (int)blockIdx.x;
        ky = (int)blockIdx.y;
    } else {
        zconv_pencils(xcd_task((int)blockIdx.x, (int)gridDim.x), per, d.Ly, bx, ky);
    }
    if (sb) {
        A = static_cast<const cpx<T> *>(sb->fwd);
        B = static_cast<cpx<T> *>(sb->inv);
        Gf = static_cast<const T *>(sb->Gf);
        Gi = static_cast<const cpx<T> *>(sb->Gi);
    }
    constexpr bool CATH = MODE == kFieldCathode, POT = MODE == kPotential;
    constexpr int NCOMP = POT ? 1 : 3;
    using S = Shape<T, L, THREADS>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * L];
    const int Kp = d.Kp, Ly = d.Ly, ny = d.ny, nz = d.nz, ny1 = ny + 1;
    for (int i = threadIdx.x; i < L; i += THREADS) tw[i] = gld(twg + i);
    const int b = threadIdx.x % NB, j = threadIdx.x / NB;
    const int kx = bx * NB + b;
    const bool live = kx < d.Kx;
    const size_t zs = (size_t)Ly * Kp;
    const cpx<T> *Ap = A + (size_t)ky * Kp + kx;
    cpx<T> p[E], pi[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        p[q] = cpx<T>{(T)0, (T)0};
        if (q < E / 2 && live) p[q] = gld(Ap + (size_t)(j + q * TPP) * zs);
    }
    reg::fft<T, L, E, -1>(p, j, lds + b, NB, tw);
    if constexpr (CATH) {
        // The image -rho(nz - 1 - z) (z < nz, zero-padded) has the spectrum
        // -w_L^((nz-1) k) P[-k]; P[-k] is held by another lane, so the
        // spectrum goes through LDS once instead of a second transform.
        __syncthreads();  // every lane past the transform's last exchange
#pragma unroll
        for (int q = 0; q < E; ++q) lds[b + (j + q * TPP) * NB] = p[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int k = j + q * TPP;
            const cpx<T> m = lds[b + ((L - k) & (L - 1)) * NB];
            const cpx<T> r = c_mul(tw[((nz - 1) * k) & (L - 1)], m);
            pi[q] = cpx<T>{-r.x, -r.y};
        }
    }
    const bool kyneg = ky > ny;
    const int kyp = kyneg ? Ly - ky : ky;
    const size_t M1 = (size_t)(nz + 1) * ny1 * Kp;
    const size_t MI = (size_t)L * ny1 * Kp;
    const int kxc = live ? kx : 0;  // keep dead lanes' loads in bounds
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) {
        const T sy = (!POT && c == 1 && kyneg) ? (T)-1 : (T)1;
        cpx<T> w[E];
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int kz = j + q * TPP;
            const bool kzneg = kz > nz;
            const int kzp = kzneg ? L - kz : kz;
            const T sz = (!POT && c == 2 && kzneg) ? (T)-1 : (T)1;
            const T g = sy * sz * gld(Gf + c * M1 + ((size_t)kzp * ny1 + kyp) * Kp + kxc);
            if constexpr (POT)
                w[q] = cpx<T>{p[q].x * g, p[q].y * g};  // rho_hat * g
            else
                w[q] = cpx<T>{-p[q].y * g, p[q].x * g};  // rho_hat * (i g)
            if constexpr (CATH) {
                cpx<T> gi = gld(Gi + c * MI + ((size_t)kz * ny1 + kyp) * Kp + kxc);
                gi.x *= sy;
                gi.y *= sy;
                if (c < 2) gi = cpx<T>{-gi.y, gi.x};
                w[q] = c_add(w[q], c_mul(pi[q], gi));
            }
        }
        reg::fft<T, L, E, +1>(w, j, lds + b, NB, tw);
        if (live) {
            cpx<T> *Bc = B + (size_t)c * nz * zs + (size_t)ky * Kp + kx;
#pragma unroll
            for (int q = 0; q < E / 2; ++q) gst(Bc + (size_t)(j + q * TPP) * zs, w[q]);
        }
    }
}

// Transposed z convolution (see k_zconv_adj in sc_solve.hip): the three g_E
// spectra in Gs are transformed one at a time and their products with
// conj(G_c^) summed in registers; the image sum, reversed in z, is
// subtracted in the spectrum (a phase factor and the mirror bin, through
// LDS), so one backward transform serves both.
template <typename T, int L, bool CATH, int THREADS>
__global__ void __launch_bounds__(THREADS)
k_zconv_adj_r(const cpx<T> *__restrict__ Gs, cpx<T> *__restrict__ out, const T *__restrict__ Gf,
              const cpx<T> *__restrict__ Gi, const cpx<T> *__restrict__ twg, Dims d)
{
    using S = Shape<T, L, THREADS>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * L];
    const int Kp = d.Kp, Ly = d.Ly, ny = d.ny, nz = d.nz, ny1 = ny + 1;
    for (int i = threadIdx.x; i < L; i += THREADS) tw[i] = gld(twg + i);
    const int b = threadIdx.x % NB, j = threadIdx.x / NB;
    // free space: a 1-D grid of div_up(Kx, NB) * Ly workgroups in XCD order;
    // with the cathode image a grid (div_up(Kx, NB), Ly), as in k_zconv_r
    int bx, ky;
    if (CATH) {
        bx = (int)blockIdx.x;
        ky = (int)blockIdx.y;
    } else {
        zconv_pencils(xcd_task((int)blockIdx.x, (int)gridDim.x), (d.Kx + NB - 1) / NB, Ly, bx, ky);
    }
    const int kx = bx * NB + b;
    const bool live = kx < d.Kx;
    const size_t zs = (size_t)Ly * Kp;
    const bool kyneg = ky > ny;
    const int kyp = kyneg ? Ly - ky : ky;
    const size_t M1 = (size_t)(nz + 1) * ny1 * Kp;
    const size_t MI = (size_t)L * ny1 * Kp;
    const int kxc = live ? kx : 0;  // keep dead lanes' loads in bounds
    cpx<T> w[E], wi[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        w[q] = cpx<T>{(T)0, (T)0};
        wi[q] = cpx<T>{(T)0, (T)0};
    }
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        const cpx<T> *Gp = Gs + (size_t)c * nz * zs + (size_t)ky * Kp + kxc;
        cpx<T> p[E];
#pragma unroll
        for (int q = 0; q < E; ++q) {
            p[q] = cpx<T>{(T)0, (T)0};
            if (q < E / 2 && live) p[q] = gld(Gp + (size_t)(j + q * TPP) * zs);
        }
        reg::fft<T, L, E, -1>(p, j, lds + b, NB, tw);
        const T sy = (c == 1 && kyneg) ? (T)-1 : (T)1;
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int kz = j + q * TPP;
            const bool kzneg = kz > nz;
            const int kzp = kzneg ? L - kz : kz;
            const T sz = (c == 2 && kzneg) ? (T)-1 : (T)1;
            const T g = sy * sz * gld(Gf + c * M1 + ((size_t)kzp * ny1 + kyp) * Kp + kxc);
            w[q] = c_add(w[q], cpx<T>{p[q].y * g, -p[q].x * g});  // p * conj(i g)
            if constexpr (CATH) {
                cpx<T> gi = gld(Gi + c * MI + ((size_t)kz * ny1 + kyp) * Kp + kxc);
                gi.x *= sy;
                gi.y *= sy;
                if (c < 2) gi = cpx<T>{-gi.y, gi.x};
                wi[q] = c_add(wi[q], c_mul(p[q], cpx<T>{gi.x, -gi.y}));
            }
        }
    }
    if constexpr (CATH) {
        // out[z] = IFFT(W)[z] - IFFT(WI)[nz-1-z].  The reversed transform is
        // IFFT(U) with U[k] = WI[-k] w_L^((nz-1) k), so one backward
        // transform of W - U gives both; WI[-k] is held by another lane.
        __syncthreads();  // every lane past the last forward transform's exchange
#pragma unroll
        for (int q = 0; q < E; ++q) lds[b + (j + q * TPP) * NB] = wi[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int k = j + q * TPP;
            const cpx<T> m = lds[b + ((L - k) & (L - 1)) * NB];
            w[q] = c_sub(w[q], c_mul(tw[((nz - 1) * k) & (L - 1)], m));
        }
    }
    reg::fft<T, L, E, +1>(w, j, lds + b, NB, tw);
    if (live) {
        cpx<T> *o = out + (size_t)ky * Kp + kx;
#pragma unroll
        for (int q = 0; q < E / 2; ++q) gst(o + (size_t)(j + q * TPP) * zs, w[q]);
    }
}

// Backward y pass: pencils (kx, z < nz) of each field spectrum, y < ny kept.
template <typename T, int L>
__global__ void __launch_bounds__(kThreads)
k_inv_y_r(cpx<T> *__restrict__ B, const cpx<T> *__restrict__ twg, Dims d,
          const SBunch *__restrict__ tab, int per)
{
    using S = Shape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * L];
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) B = static_cast<cpx<T> *>(sb->inv);
    for (int i = threadIdx.x; i < L; i += kThreads) tw[i] = gld(twg + i);
    const int b = threadIdx.x % NB, j = threadIdx.x / NB;
    const int kx = bx * NB + b, z = blockIdx.y, c = blockIdx.z;
    const int Kx = d.Kx, Kp = d.Kp;
    const bool live = kx < Kx;
    cpx<T> *Bp = B + ((size_t)c * d.nz + z) * d.Ly * Kp + (live ? kx : 0);
    cpx<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = gld(Bp + (size_t)(j + q * TPP) * Kp);
    reg::fft<T, L, E, +1>(v, j, lds + b, NB, tw);
    if (live) {
#pragma unroll
        for (int q = 0; q < E / 2; ++q) gst(Bp + (size_t)(j + q * TPP) * Kp, v[q]);
    }
}

// Backward x pass (complex to real) with Hermitian completion.
template <typename T, int L>
__global__ void __launch_bounds__(kThreads)
k_inv_x_r(const cpx<T> *__restrict__ B, T *__restrict__ E_, const cpx<T> *__restrict__ twg, Dims d,
          const SBunch *__restrict__ tab, int per)
{
    using S = Shape<T, L>;
    constexpr int E = S::E, TPP = S::TPP, NB = S::NB;
    __shared__ cpx<T> tw[L];
    __shared__ cpx<T> lds[NB * reg::padded<L, E>()];
    const SBunch *sb;
    const int bx = group_block(tab, per, sb);
    if (sb) {
        B = static_cast<const cpx<T> *>(sb->inv);
        E_ = static_cast<T *>(sb->out);
    }
    for (int i = threadIdx.x; i < L; i += kThreads) tw[i] = gld(twg + i);
    const int b = threadIdx.x / TPP, j = threadIdx.x % TPP;
    const int r = bx * NB + b, c = blockIdx.y;
    const int nrows = d.ny * d.nz, nx = d.nx;
    const bool live = r < nrows;
    const int rr = live ? r : 0;
    const int y = rr % d.ny, z = rr / d.ny;
    const cpx<T> *row = B + (size_t)c * d.nz * d.Ly * d.Kp + ((size_t)z * d.Ly + y) * d.Kp;
    // one unconditional load per element (the Hermitian half reads the
    // mirror), so all of a lane's loads are in flight together
    cpx<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int k = j + q * TPP;
        const bool direct = q < E / 2 || k == nx;
        v[q] = gld(row + (direct ? k : L - k));
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int k = j + q * TPP;
        if (!(q < E / 2 || k == nx)) v[q].y = -v[q].y;
    }
    reg::fft<T, L, E, +1>(v, j, lds + b * reg::padded<L, E>(), 1, tw);
    if (live) {
        T *out = E_ + (size_t)c * nx * nrows + (size_t)r * nx;
#pragma unroll
        for (int q = 0; q < E / 2; ++q) gst(out + j + q * TPP, v[q].x);
    }
}

template <typename T, int L>
void launch_all(Mesh *m, const Dims &d, int axis_kernel, const SolveIO &io)
{
    Workspace &ws = m->ws;
    Ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    using S = Shape<T, L>;
    const cpx<T> *tw = static_cast<const cpx<T> *>(ws.twT[axis_kernel == 0 || axis_kernel == 4 ? 0
                                                           : axis_kernel == 2         ? 2
                                                                                      : 1]);
    cpx<T> *F = static_cast<cpx<T> *>(io.fwd);
    cpx<T> *B = static_cast<cpx<T> *>(io.inv);
    switch (axis_kernel) {
    case 0:
        launch_begin(ctx, K_RHO_FWD_X);
        hipLaunchKernelGGL((k_rho_fwd_x_r<T, L>), dim3(div_up(d.ny * d.nz, S::NB), io.nin),
                           dim3(kThreads), 0, st, static_cast<const T *>(io.in), F, tw, d,
                           (const SBunch *)nullptr, 0);
        launch_end(ctx, K_RHO_FWD_X);
        break;
    case 1:
        launch_begin(ctx, K_RHO_FWD_Y);
        hipLaunchKernelGGL((k_rho_fwd_y_r<T, L>), dim3(div_up(d.Kx, S::NB), d.nz, io.nin),
                           dim3(kThreads), 0, st, F, tw, d, (const SBunch *)nullptr, 0);
        launch_end(ctx, K_RHO_FWD_Y);
        break;
    case 2: {
        constexpr int ZT = zconv_threads<T, L>();
        using SZ = Shape<T, L, ZT>;
        launch_begin(ctx, K_ZCONV);
        const int nbx = div_up(d.Kx, SZ::NB);
        const bool cath_mode = io.mode == kFieldCathode || io.mode == kAdjointCathode;
        const dim3 zgrid = cath_mode ? dim3(nbx, d.Ly) : dim3(nbx * d.Ly);
        const cpx<T> *Fc = F;
        const T *Gf = static_cast<const T *>(ws.Gf);
        const cpx<T> *Gi = static_cast<const cpx<T> *>(ws.Gi);
        switch (io.mode) {
        case kFieldCathode:
            hipLaunchKernelGGL((k_zconv_r<T, L, kFieldCathode, ZT>), zgrid, dim3(ZT), 0, st, Fc, B, Gf,
                               Gi, tw, d, (const SBunch *)nullptr, nbx);
            break;
        case kPotential:
            hipLaunchKernelGGL((k_zconv_r<T, L, kPotential, ZT>), zgrid, dim3(ZT), 0, st, Fc, B, Gf, Gi,
                               tw, d, (const SBunch *)nullptr, nbx);
            break;
        case kAdjointFree:
            hipLaunchKernelGGL((k_zconv_adj_r<T, L, false, ZT>), zgrid, dim3(ZT), 0, st, Fc, B, Gf, Gi,
                               tw, d);
            break;
        case kAdjointCathode:
            hipLaunchKernelGGL((k_zconv_adj_r<T, L, true, ZT>), zgrid, dim3(ZT), 0, st, Fc, B, Gf, Gi,
                               tw, d);
            break;
        default:
            hipLaunchKernelGGL((k_zconv_r<T, L, kFieldFree, ZT>), zgrid, dim3(ZT), 0, st, Fc, B, Gf, Gi,
                               tw, d, (const SBunch *)nullptr, nbx);
            break;
        }
        launch_end(ctx, K_ZCONV);
        break;
    }
    case 3:
        launch_begin(ctx, K_INV_Y);
        hipLaunchKernelGGL((k_inv_y_r<T, L>), dim3(div_up(d.Kx, S::NB), d.nz, io.nout),
                           dim3(kThreads), 0, st, B, tw, d, (const SBunch *)nullptr, 0);
        launch_end(ctx, K_INV_Y);
        break;
    case 4:
        launch_begin(ctx, K_INV_X);
        hipLaunchKernelGGL((k_inv_x_r<T, L>), dim3(div_up(d.ny * d.nz, S::NB), io.nout),
                           dim3(kThreads), 0, st, (const cpx<T> *)B, static_cast<T *>(io.out), tw, d,
                           (const SBunch *)nullptr, 0);
        launch_end(ctx, K_INV_X);
        break;
    }
}

// Solve kernel `which` of nb same-shape bunches in one launch (x stacked).
template <typename T, int L>
void launch_all_grouped(const Dims &d, const cpx<T> *tw, int nb, const SBunch *tab, int which,
                        int mode, int nin, int nout, hipStream_t st)
{
    using S = Shape<T, L>;
    switch (which) {
    case 0: {
        const int per = div_up(d.ny * d.nz, S::NB);
        hipLaunchKernelGGL((k_rho_fwd_x_r<T, L>), dim3(per * nb, nin), dim3(kThreads), 0, st,
                           (const T *)nullptr, (cpx<T> *)nullptr, tw, d, tab, per);
        break;
    }
    case 1: {
        const int per = div_up(d.Kx, S::NB);
        hipLaunchKernelGGL((k_rho_fwd_y_r<T, L>), dim3(per * nb, d.nz, nin), dim3(kThreads), 0, st,
                           (cpx<T> *)nullptr, tw, d, tab, per);
        break;
    }
    case 2: {
        constexpr int ZT = zconv_threads<T, L>();
        using SZ = Shape<T, L, ZT>;
        const int per = div_up(d.Kx, SZ::NB);
        const dim3 zgrid(per * nb, d.Ly);
        if (mode == kFieldCathode)
            hipLaunchKernelGGL((k_zconv_r<T, L, kFieldCathode, ZT>), zgrid, dim3(ZT), 0, st,
                               (const cpx<T> *)nullptr, (cpx<T> *)nullptr, (const T *)nullptr,
                               (const cpx<T> *)nullptr, tw, d, tab, per);
        else
            hipLaunchKernelGGL((k_zconv_r<T, L, kFieldFree, ZT>), zgrid, dim3(ZT), 0, st,
                               (const cpx<T> *)nullptr, (cpx<T> *)nullptr, (const T *)nullptr,
                               (const cpx<T> *)nullptr, tw, d, tab, per);
        break;
    }
    case 3: {
        const int per = div_up(d.Kx, S::NB);
        hipLaunchKernelGGL((k_inv_y_r<T, L>), dim3(per * nb, d.nz, nout), dim3(kThreads), 0, st,
                           (cpx<T> *)nullptr, tw, d, tab, per);
        break;
    }
    case 4: {
        const int per = div_up(d.ny * d.nz, S::NB);
        hipLaunchKernelGGL((k_inv_x_r<T, L>), dim3(per * nb, nout), dim3(kThreads), 0, st,
                           (const cpx<T> *)nullptr, (T *)nullptr, tw, d, tab, per);
        break;
    }
    }
}

}  // namespace

template <typename T>
void launch_fast_grouped(const Mesh *m0, int nb, const SBunch *tab, int which, int mode, int nin,
                         int nout, hipStream_t st)
{
    const Dims d = make_dims(m0);
    const int axis = (which == 0 || which == 4) ? 0 : (which == 2 ? 2 : 1);
    const int L = axis == 0 ? d.Lx : (axis == 2 ? d.Lz : d.Ly);
    const cpx<T> *tw = static_cast<const cpx<T> *>(m0->ws.twT[axis]);
    switch (L) {
    case 32: launch_all_grouped<T, 32>(d, tw, nb, tab, which, mode, nin, nout, st); break;
    case 64: launch_all_grouped<T, 64>(d, tw, nb, tab, which, mode, nin, nout, st); break;
    case 128: launch_all_grouped<T, 128>(d, tw, nb, tab, which, mode, nin, nout, st); break;
    case 256: launch_all_grouped<T, 256>(d, tw, nb, tab, which, mode, nin, nout, st); break;
    case 512: launch_all_grouped<T, 512>(d, tw, nb, tab, which, mode, nin, nout, st); break;
    default: break;  // the caller groups only supported lengths
    }
}

template void launch_fast_grouped<float>(const Mesh *, int, const SBunch *, int, int, int, int,
                                         hipStream_t);
template void launch_fast_grouped<double>(const Mesh *, int, const SBunch *, int, int, int, int,
                                          hipStream_t);

// Runs one solve kernel (0 rho_fwd_x, 1 rho_fwd_y, 2 zconv, 3 inv_y,
// 4 inv_x) with the register engine if its transform length is a supported
// power of two; returns false (nothing launched) otherwise.
template <typename T>
bool launch_fast(Mesh *m, int which, const SolveIO &io)
{
    const Dims d = make_dims(m);
    // the register engine's passes take the mesh's nodes as the half-length
    if (dims_padded(d)) return false;
    const int L = (which == 0 || which == 4) ? d.Lx : (which == 2 ? d.Lz : d.Ly);
    switch (L) {
    case 32: launch_all<T, 32>(m, d, which, io); return true;
    case 64: launch_all<T, 64>(m, d, which, io); return true;
    case 128: launch_all<T, 128>(m, d, which, io); return true;
    case 256: launch_all<T, 256>(m, d, which, io); return true;
    case 512: launch_all<T, 512>(m, d, which, io); return true;
    default: return false;
    }
}

template bool launch_fast<float>(Mesh *, int, const SolveIO &);
template bool launch_fast<double>(Mesh *, int, const SolveIO &);

}  // namespace sc

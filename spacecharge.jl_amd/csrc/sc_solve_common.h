// sc_solve_common.h -- pieces shared by the Green-function build and the
// field transforms.
#pragma once

#include <hip/hip_runtime.h>

#include "sc_fft_dev.h"
#include "sc_internal.h"

namespace sc {

constexpr double kFPEI = 299792458.0 * 299792458.0 * 1.0e-7;  // src/utils.jl:7-8

// true when L has no prime factor above 7 (the radices of the transforms)
inline bool smooth7(int L)
{
    if (L < 1) return false;
    for (int f : {2, 3, 5, 7})
        while (L % f == 0) L /= f;
    return L == 1;
}

// Transform half-length of an axis of n nodes: the smallest t >= n whose 2t
// has no prime factor above 7.  The reference zero-pads each axis to exactly
// 2n (src/mesh.jl:51-61) and FFTW transforms any length.  The convolution it
// computes is linear: output node m < n sums rho[j] G(m - j) over j < n, lags
// |m - j| <= n - 1, so any cyclic length of at least 2n - 1 gives the same
// outputs.  Axes whose 2n has a larger prime factor are therefore padded
// further, to 2t: rho stays zero beyond its n nodes, the Green function is
// built on t lags (its first n are the reference's cell integrals; it depends
// on the spacing, gamma and the image offset only, not on n), and only the n
// physical outputs are written.  t == n for every axis whose 2n is 7-smooth.
inline int transform_half(int n)
{
    int t = n;
    while (!smooth7(2 * t)) ++t;
    return t;
}

// Shape of a solve.  nx, ny, nz are the transform half-lengths (the Green
// function's lags and spectral octant); mx, my, mz the mesh's nodes (the
// extent of rho, efield and the real rows the passes read and write).  They
// differ only on axes padded beyond 2n (transform_half).
struct Dims {
    int nx, ny, nz;  // transform half-lengths: L = 2n, spectral octant n + 1
    int mx, my, mz;  // mesh nodes (data extent)
    int Kx;          // nx + 1 kept x frequencies
    int Kp;          // row stride of kx-indexed arrays: Kx rounded up to 32, so that the
                     // 16-lane row segments of every transform start on a 128-byte line
    int Lx, Ly, Lz;  // padded lengths
};

inline Dims make_dims(const Mesh *m)
{
    Dims d;
    d.mx = (int)m->n[0];
    d.my = (int)m->n[1];
    d.mz = (int)m->n[2];
    d.nx = transform_half(d.mx);
    d.ny = transform_half(d.my);
    d.nz = transform_half(d.mz);
    d.Kx = d.nx + 1;
    d.Kp = (d.Kx + 31) / 32 * 32;
    d.Lx = 2 * d.nx;
    d.Ly = 2 * d.ny;
    d.Lz = 2 * d.nz;
    return d;
}

// true when some axis is padded beyond twice its nodes (generic engine only)
inline bool dims_padded(const Dims &d) { return d.nx != d.mx || d.ny != d.my || d.nz != d.mz; }

inline int div_up(int a, int b) { return (a + b - 1) / b; }

// Largest batch with nb*n <= cap elements, at most nbmax pencils.
inline int pick_nb(int n, int cap, int nbmax)
{
    int nb = cap / n;
    if (nb > nbmax) nb = nbmax;
    if (nb < 1) nb = 1;
    return nb;
}

// What one solve computes: the three field components in free space or with
// the cathode image charge, the scalar potential in free space, or the
// transpose of the field solve (free space or cathode): g_rho = S^T g_E.
enum SolveMode {
    kFieldFree = 0,
    kFieldCathode = 1,
    kPotential = 2,
    kAdjointFree = 3,
    kAdjointCathode = 4
};

inline bool mode_cathode(int mode) { return mode == kFieldCathode || mode == kAdjointCathode; }
inline bool mode_adjoint(int mode) { return mode == kAdjointFree || mode == kAdjointCathode; }
// The Green spectra a mode needs (the transpose uses the field's).
inline int green_mode(int mode)
{
    return mode == kPotential ? kPotential : (mode_cathode(mode) ? kFieldCathode : kFieldFree);
}

// Buffers of one solve: the real input `in` (nin components of the mesh's
// mx*my*mz nodes) is transformed along x and y into the spectra `fwd` (nin of
// them), the z convolution writes nout spectra to `inv`, and the inverse y
// and x passes write the real output `out` (nout components).  Spectra are
// [mz][Ly][Kp] complex each, consecutive.
struct SolveIO {
    int mode;
    int nin, nout;
    const void *in;
    void *out;
    void *fwd;
    void *inv;
};

// ---- grouped solves: bunches of one mesh shape, one launch per kernel ----
// Per-bunch buffers and Green geometry of a grouped solve (a table in device
// memory).  A grouped launch stacks the bunches along x: workgroup x belongs
// to bunch x / per and acts as workgroup x % per of that bunch's own launch.
// Kernels given no table (the single-bunch launch) use their arguments.
struct SBunch {
    const void *in;
    void *out, *fwd, *inv, *Gf, *Gi, *R;
    double *F;
    double dx, dy, dz, offzg;  // Green spacings (dz times gamma) and image offset times gamma
    double scale[4];
};

__device__ __forceinline__ int group_block(const SBunch *tab, int per, const SBunch *&sb)
{
    if (!tab) {
        sb = nullptr;
        return (int)blockIdx.x;
    }
    const int k = (int)blockIdx.x / per;
    sb = tab + k;
    return (int)blockIdx.x - k * per;
}

// Pencil (x block, ky) of z-convolution task `task`, with nbx x blocks per ky
// row.  The rows go in the order 0, 1, Ly-1, 2, Ly-2, ..., ny, because rows
// ky and Ly - ky read the same Green row (the spectrum is even in ky), and
// neighbouring x blocks share the 128-byte lines of that row.
__device__ __forceinline__ void zconv_pencils(int task, int nbx, int Ly, int &bx, int &ky)
{
    bx = task % nbx;
    const int u = task / nbx;
    const int kyp = (u + 1) >> 1;
    ky = u == 0 ? 0 : ((u & 1) ? kyp : Ly - kyp);
}

// Builds the transformed Green functions for green_mode `mode` into ws.Gf
// (and ws.Gi for kFieldCathode): three field components, or the
// potential's one (stored in the first component's slot).
template <typename T>
int build_green(Mesh *m, const double dl[3], double gamma, double offz, int mode);

// Solve kernel `which` (0 forward x, 1 forward y, 2 z convolution, 3 inverse
// y, 4 inverse x) of `io` on the register-resident engine; false when its
// transform length is not one of the supported powers of two (nothing is
// launched then).
template <typename T>
bool launch_fast(Mesh *m, int which, const SolveIO &io);

// Grouped forms, on `st`, for nb bunches of mesh shape d whose transform
// lengths are supported powers of two: the Green build (field modes) of
// every bunch, from m0's twiddles and the table's spacings, and solve
// kernel `which` of every bunch with the table's buffers.
template <typename T>
void build_green_grouped(const Mesh *m0, int nb, const SBunch *tab, double gamma, int mode,
                         hipStream_t st);
template <typename T>
void launch_fast_grouped(const Mesh *m0, int nb, const SBunch *tab, int which, int mode, int nin,
                         int nout, hipStream_t st);

}  // namespace sc

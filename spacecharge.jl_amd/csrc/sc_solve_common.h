// sc_solve_common.h -- pieces shared by the Green-function build and the
// field transforms.
#pragma once

#include <hip/hip_runtime.h>

#include "sc_fft_dev.h"
#include "sc_internal.h"

namespace sc {

constexpr double kFPEI = 299792458.0 * 299792458.0 * 1.0e-7;  // src/utils.jl:7-8

struct Dims {
    int nx, ny, nz;
    int Kx;          // nx + 1 kept x frequencies
    int Kp;          // row stride of kx-indexed arrays: Kx rounded up to 32, so that the
                     // 16-lane row segments of every transform start on a 128-byte line
    int Lx, Ly, Lz;  // padded lengths
};

inline Dims make_dims(const Mesh *m)
{
    Dims d;
    d.nx = (int)m->n[0];
    d.ny = (int)m->n[1];
    d.nz = (int)m->n[2];
    d.Kx = d.nx + 1;
    d.Kp = (d.Kx + 31) / 32 * 32;
    d.Lx = 2 * d.nx;
    d.Ly = 2 * d.ny;
    d.Lz = 2 * d.nz;
    return d;
}

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
// the cathode image charge, or the scalar potential in free space.
enum SolveMode { kFieldFree = 0, kFieldCathode = 1, kPotential = 2 };

// Builds the transformed Green functions of one solve into ws.Gf (and
// ws.Gi for kFieldCathode): three field components, or the potential's one
// (stored in the first component's slot).
template <typename T>
int build_green(Mesh *m, const double dl[3], double gamma, double offz, int mode);

// Solve kernel `which` (0 rho_fwd_x, 1 rho_fwd_y, 2 zconv, 3 inv_y,
// 4 inv_x) on the register-resident engine; false when its transform length
// is not one of the supported powers of two (nothing is launched then).
// inv_x writes the solve's components to `out` (ncomp * nx*ny*nz elements).
template <typename T>
bool launch_fast(Mesh *m, int which, int mode, void *out);

}  // namespace sc

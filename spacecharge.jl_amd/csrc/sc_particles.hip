// sc_particles.hip -- particle <-> mesh kernels on gfx950.
//
//  * deposit: cloud-in-cell (src/deposition.jl:28-86, GPU kernel :106-129).
//  * interpolate: trilinear gather (src/interpolation.jl:17-86).
//  * bounds: min/max reduction for the particle-extrema mesh constructor
//    (src/mesh.jl:120-122).
//  * vector-Jacobian products of deposit and interpolation (no reference
//    counterpart): gathers of the node cotangent / field with the CIC
//    weights and their position derivatives.
//
// Per-particle arithmetic keeps the reference's evaluation order without
// FMA contraction, so a particle's eight deposit contributions and its
// interpolated fields are bit-identical to the CPU path; only the order in
// which the atomics add up differs.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

#include "sc_internal.h"

namespace sc {

namespace {

template <typename A, typename B>
using wider_t = typename std::conditional<(sizeof(A) >= sizeof(B)), A, B>::type;

// Lower cell index of a normalised coordinate, clamped to [0, n-2] (NaN -> 0).
template <typename C>
__device__ __forceinline__ int cell_index(C norm, int n)
{
    const C f = floor(norm);
    if (!(f >= (C)0)) return 0;
    if (f > (C)(n - 2)) return n - 2;
    return (int)f;
}

// Lower bounds and spacing from the mesh's device-resident geometry (lo at
// 0, delta at 6) when the call came without host bounds.
__device__ __forceinline__ void device_geometry(const double *geom, double &lx, double &ly,
                                                double &lz, double &dx, double &dy, double &dz)
{
    if (geom) {
        lx = geom[0];
        ly = geom[1];
        lz = geom[2];
        dx = geom[6];
        dy = geom[7];
        dz = geom[8];
    }
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(256)
k_deposit(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
          const PT *__restrict__ pq, int64_t n, MT *__restrict__ rho, int nx, int ny, int nz,
          double lx_, double ly_, double lz_, double dx_, double dy_, double dz_,
          const double *__restrict__ geom)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    device_geometry(geom, lx_, ly_, lz_, dx_, dy_, dz_);
    const C lx = (C)(MT)lx_, ly = (C)(MT)ly_, lz = (C)(MT)lz_;
    const C ddx = (C)(MT)dx_, ddy = (C)(MT)dy_, ddz = (C)(MT)dz_;
    const C nxn = ((C)px[p] - lx) / ddx;
    const C nyn = ((C)py[p] - ly) / ddy;
    const C nzn = ((C)pz[p] - lz) / ddz;
    const int ix = cell_index(nxn, nx), iy = cell_index(nyn, ny), iz = cell_index(nzn, nz);
    const C dx = nxn - (C)ix, dy = nyn - (C)iy, dz = nzn - (C)iz;
    const C wx0 = (C)1 - dx, wx1 = dx, wy0 = (C)1 - dy, wy1 = dy, wz0 = (C)1 - dz, wz1 = dz;
    const C c = (C)pq[p];
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    MT *b = rho + ix + sy * iy + sz * iz;
    atomicAdd(b, (MT)(((c * wx0) * wy0) * wz0));
    atomicAdd(b + 1, (MT)(((c * wx1) * wy0) * wz0));
    atomicAdd(b + sy, (MT)(((c * wx0) * wy1) * wz0));
    atomicAdd(b + sy + 1, (MT)(((c * wx1) * wy1) * wz0));
    atomicAdd(b + sz, (MT)(((c * wx0) * wy0) * wz1));
    atomicAdd(b + sz + 1, (MT)(((c * wx1) * wy0) * wz1));
    atomicAdd(b + sz + sy, (MT)(((c * wx0) * wy1) * wz1));
    atomicAdd(b + sz + sy + 1, (MT)(((c * wx1) * wy1) * wz1));
}

// Node-interleaved copy of the field, (Ex, Ey, Ez, 0) per node, so that the
// gather below needs one vector load per node instead of three scalar ones.
template <typename T>
struct V4;
template <>
struct V4<float> {
    using type = float4;
};
template <>
struct V4<double> {
    using type = double4;
};

// Large Float64 meshes (P = 2) store node pairs: entry i holds nodes i and
// i + 1 (the next node along x; zero past the row's end, which no cell
// reads), 64 bytes, so that a cell's two x neighbours never straddle a
// 128-byte line.  Measured on 4M particles, 128^3: the gather 307 -> 245 us,
// the copy 26 -> 47 us, the step 0.834 -> 0.792 ms.  Float32 meshes keep one
// node per entry (16-byte nodes straddle a line in one pair of eight; on 4M
// particles, 128^3, pairs took the gather from 99 to 134 us).
// The six values of a pair are packed (Ex, Ey, Ez of node i, then of node
// i + 1, then 16 bytes of zeros), so a gather of both nodes is three
// 16-byte loads (pair_nodes) instead of four: on 4M particles, 128^3, the
// gather 215 -> 194 us, the copy 45 -> 48 us.  Packed, pairs also pay on
// smaller Float64 meshes (1M particles, 64^3: gather 46 -> 36 us, step
// 0.167 -> 0.161 ms; the 2M cathode bunch on 64x64x128: 103 -> 92 us), so
// every Float64 mesh takes them.  With P = 2 a lane writes one 16-byte
// quarter of an entry, so the stores are contiguous.
template <typename MT>
__device__ __forceinline__ void interleave_body(int64_t i, const MT *__restrict__ ef,
                                                typename V4<MT>::type *__restrict__ efi, int64_t M,
                                                int nx, int P)
{
    if constexpr (sizeof(MT) == 8) {
        if (P == 2) {
            if (i >= 4 * M) return;
            const int64_t e = i >> 2;
            const int part = (int)(i & 3);
            const bool hi = e % nx != nx - 1;  // node e + 1 is in the row
            // values 2 part and 2 part + 1 of the entry's six: value s is
            // component s % 3 of node e + s / 3 (no branch per quarter)
            const int s0 = 2 * part, s1 = s0 + 1;
            const bool ok0 = s0 < 3 || (s0 < 6 && hi), ok1 = s1 < 3 || (s1 < 6 && hi);
            const int64_t i0 = ok0 ? (s0 % 3) * M + e + s0 / 3 : 0;
            const int64_t i1 = ok1 ? (s1 % 3) * M + e + s1 / 3 : 0;
            const double a = ef[i0], b = ef[i1];
            // streamed past the L2: a pair copy (64 bytes per node) outgrows
            // it anyway, and the gather then finds it in the last-level cache
            // (4M particles, 128^3: the copy 47.5 -> 34.1 us, the gather
            // 194.4 -> 196.8 us).  The padding quarter is written too: left
            // out, the partial lines made the copy slower (55 us).
            double *dst = reinterpret_cast<double *>(efi) + 2 * i;
            __builtin_nontemporal_store(ok0 ? a : 0.0, dst);
            __builtin_nontemporal_store(ok1 ? b : 0.0, dst + 1);
            return;
        }
    }
    if (i >= M) return;
    typename V4<MT>::type v;
    v.x = ef[i];
    v.y = ef[M + i];
    v.z = ef[2 * M + i];
    v.w = (MT)0;
    efi[i] = v;
}

template <typename MT>
__global__ void __launch_bounds__(256)
k_interleave(const MT *__restrict__ ef, typename V4<MT>::type *__restrict__ efi, int64_t M, int nx,
             int P)
{
    interleave_body<MT>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, ef, efi, M, nx, P);
}

// Nodes i and i + 1 of a packed node-pair entry (Float64, P = 2): the
// first 32 bytes and the next 16, as two one-node entries.
__device__ __forceinline__ void pair_nodes(const double4 *__restrict__ ent, double4 &lo, double4 &hi)
{
    const double4 a = ent[0];
    const double2 b = *reinterpret_cast<const double2 *>(ent + 1);
    lo = make_double4(a.x, a.y, a.z, 0.0);
    hi = make_double4(a.w, b.x, b.y, 0.0);
}

// PP: the entry layout when known at compile time (1 or 2), else 0 and P
// gives it (the grouped launches, where it varies by bunch).
template <typename PT, typename MT, int PP = 0>
__device__ __forceinline__ void interp_body(
    int64_t p, const typename V4<MT>::type *__restrict__ efi, const PT *__restrict__ px,
    const PT *__restrict__ py, const PT *__restrict__ pz, int64_t n, int nx, int ny, int nz,
    double lx_, double ly_, double lz_, double dx_, double dy_, double dz_,
    const double *__restrict__ geom, PT *__restrict__ Ex, PT *__restrict__ Ey, PT *__restrict__ Ez,
    int P)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    if constexpr (PP == 0) {
        // the two layouts as separate code, so neither pays for the other's loads
        if (P == 2)
            interp_body<PT, MT, 2>(p, efi, px, py, pz, n, nx, ny, nz, lx_, ly_, lz_, dx_, dy_, dz_,
                                   geom, Ex, Ey, Ez, P);
        else
            interp_body<PT, MT, 1>(p, efi, px, py, pz, n, nx, ny, nz, lx_, ly_, lz_, dx_, dy_, dz_,
                                   geom, Ex, Ey, Ez, P);
        return;
    }
    if (p >= n) return;
    device_geometry(geom, lx_, ly_, lz_, dx_, dy_, dz_);
    const C lx = (C)(MT)lx_, ly = (C)(MT)ly_, lz = (C)(MT)lz_;
    const C ddx = (C)(MT)dx_, ddy = (C)(MT)dy_, ddz = (C)(MT)dz_;
    const C nxn = ((C)px[p] - lx) / ddx;
    const C nyn = ((C)py[p] - ly) / ddy;
    const C nzn = ((C)pz[p] - lz) / ddz;
    const int ix = cell_index(nxn, nx), iy = cell_index(nyn, ny), iz = cell_index(nzn, nz);
    const C dx = nxn - (C)ix, dy = nyn - (C)iy, dz = nzn - (C)iz;
    const C one = (C)1;
    const C w000 = ((one - dx) * (one - dy)) * (one - dz);
    const C w100 = (dx * (one - dy)) * (one - dz);
    const C w010 = ((one - dx) * dy) * (one - dz);
    const C w110 = (dx * dy) * (one - dz);
    const C w001 = ((one - dx) * (one - dy)) * dz;
    const C w101 = (dx * (one - dy)) * dz;
    const C w011 = ((one - dx) * dy) * dz;
    const C w111 = (dx * dy) * dz;
    const int64_t sy = (int64_t)P * nx, sz = (int64_t)P * nx * ny;
    const typename V4<MT>::type *e = efi + P * (ix + (int64_t)nx * iy + (int64_t)nx * ny * iz);
    typename V4<MT>::type n000, n100, n010, n110, n001, n101, n011, n111;
    if constexpr (sizeof(MT) == 8) {
        if (PP == 2) {
            pair_nodes(e, n000, n100);
            pair_nodes(e + sy, n010, n110);
            pair_nodes(e + sz, n001, n101);
            pair_nodes(e + sz + sy, n011, n111);
        } else {
            n000 = e[0], n100 = e[1], n010 = e[sy], n110 = e[sy + 1];
            n001 = e[sz], n101 = e[sz + 1], n011 = e[sz + sy], n111 = e[sz + sy + 1];
        }
    } else {
        n000 = e[0], n100 = e[1], n010 = e[sy], n110 = e[sy + 1];
        n001 = e[sz], n101 = e[sz + 1], n011 = e[sz + sy], n111 = e[sz + sy + 1];
    }
    // the reference's sum order, per component (src/interpolation.jl:56-85)
#define SC_ITP(f)                                              \
    ((((((((C)n000.f * w000 + (C)n100.f * w100) + (C)n010.f * w010) + (C)n110.f * w110) + \
         (C)n001.f * w001) + (C)n101.f * w101) + (C)n011.f * w011) + (C)n111.f * w111)
    Ex[p] = (PT)SC_ITP(x);
    Ey[p] = (PT)SC_ITP(y);
    Ez[p] = (PT)SC_ITP(z);
#undef SC_ITP
}

// Float64 fields: at most 4 waves per SIMD, so that the compiler may keep
// every gather of a particle in flight at once (70 VGPRs).  Left to its own
// occupancy target it fitted 62 VGPRs by issuing the z components' loads
// only after the first loads had returned, a second memory round trip: on
// 4M particles, 128^3, the gather went from 240 to 210 us (node pairs) and
// from 303 to 264 us (one node per entry).
template <typename PT, typename MT, int PP>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(1, sizeof(MT) == 8 ? 4 : 8)))
k_interp(const typename V4<MT>::type *__restrict__ efi, const PT *__restrict__ px,
         const PT *__restrict__ py, const PT *__restrict__ pz, int64_t n, int nx, int ny, int nz,
         double lx_, double ly_, double lz_, double dx_, double dy_, double dz_,
         const double *__restrict__ geom, PT *__restrict__ Ex, PT *__restrict__ Ey,
         PT *__restrict__ Ez, int P)
{
    interp_body<PT, MT, PP>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, efi, px, py, pz, n, nx,
                            ny, nz, lx_, ly_, lz_, dx_, dy_, dz_, geom, Ex, Ey, Ez, P);
}

// ---- vector-Jacobian products of the CIC weights ----
// At a particle with cell (i, j, k) and fractions (dx, dy, dz) the weight of
// node (i+a, j+b, k+c) is Wa(dx) Wb(dy) Wc(dz), W0(t) = 1 - t, W1(t) = t,
// and dt/dx = 1/delta_x with the cell index held fixed: the weights are
// linear in the position inside a cell, and a particle on a cell face takes
// the derivative of the cell it is assigned to (as its weights do).  For
// node values f[a + 2b + 4c] this returns the interpolant (in the order of
// k_interp) and its x, y, z derivatives times the spacings.
template <typename C>
__device__ __forceinline__ void trilinear_grad(const C f[8], C dx, C dy, C dz, C &val, C &gx, C &gy,
                                               C &gz)
{
#pragma clang fp contract(off)
    const C one = (C)1;
    const C ex = one - dx, ey = one - dy, ez = one - dz;
    val = ((((((f[0] * ((ex * ey) * ez) + f[1] * ((dx * ey) * ez)) + f[2] * ((ex * dy) * ez)) +
               f[3] * ((dx * dy) * ez)) +
              f[4] * ((ex * ey) * dz)) +
             f[5] * ((dx * ey) * dz)) +
            f[6] * ((ex * dy) * dz)) +
          f[7] * ((dx * dy) * dz);
    gx = (((f[1] - f[0]) * (ey * ez) + (f[3] - f[2]) * (dy * ez)) + (f[5] - f[4]) * (ey * dz)) +
         (f[7] - f[6]) * (dy * dz);
    gy = (((f[2] - f[0]) * (ex * ez) + (f[3] - f[1]) * (dx * ez)) + (f[6] - f[4]) * (ex * dz)) +
         (f[7] - f[5]) * (dx * dz);
    gz = (((f[4] - f[0]) * (ex * ey) + (f[5] - f[1]) * (dx * ey)) + (f[6] - f[2]) * (ex * dy)) +
         (f[7] - f[3]) * (dx * dy);
}

// VJP of deposit at cotangent g (one value per node): gq = sum_8 w g and
// g_pos = q * d/dpos (sum_8 w g).  Null outputs are skipped.
template <typename PT, typename MT>
__global__ void __launch_bounds__(256)
k_deposit_vjp(const MT *__restrict__ g, const PT *__restrict__ px, const PT *__restrict__ py,
              const PT *__restrict__ pz, const PT *__restrict__ pq, int64_t n, int nx, int ny, int nz,
              double lx_, double ly_, double lz_, double dx_, double dy_, double dz_,
              const double *__restrict__ geom, PT *__restrict__ gq, PT *__restrict__ gx,
              PT *__restrict__ gy, PT *__restrict__ gz, int accumulate)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    device_geometry(geom, lx_, ly_, lz_, dx_, dy_, dz_);
    const C ddx = (C)(MT)dx_, ddy = (C)(MT)dy_, ddz = (C)(MT)dz_;
    const C nxn = ((C)px[p] - (C)(MT)lx_) / ddx;
    const C nyn = ((C)py[p] - (C)(MT)ly_) / ddy;
    const C nzn = ((C)pz[p] - (C)(MT)lz_) / ddz;
    const int ix = cell_index(nxn, nx), iy = cell_index(nyn, ny), iz = cell_index(nzn, nz);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const MT *b = g + (ix + sy * iy + sz * iz);
    const C f[8] = {(C)b[0],  (C)b[1],      (C)b[sy],      (C)b[sy + 1],
                    (C)b[sz], (C)b[sz + 1], (C)b[sz + sy], (C)b[sz + sy + 1]};
    C val, Gx, Gy, Gz;
    trilinear_grad(f, nxn - (C)ix, nyn - (C)iy, nzn - (C)iz, val, Gx, Gy, Gz);
    if (gq) gq[p] = (PT)val;
    if (gx) {
        const C q = (C)pq[p];
        const C tx = q * (Gx / ddx), ty = q * (Gy / ddy), tz = q * (Gz / ddz);
        if (accumulate) {
            gx[p] = (PT)((C)gx[p] + tx);
            gy[p] = (PT)((C)gy[p] + ty);
            gz[p] = (PT)((C)gz[p] + tz);
        } else {
            gx[p] = (PT)tx;
            gy[p] = (PT)ty;
            gz[p] = (PT)tz;
        }
    }
}

// Position part of the VJP of interpolation at cotangents (gEx, gEy, gEz):
// g_pos = sum_c gE_c * d/dpos (sum_8 w E_c), from the interleaved field
// (entry layout PP, as P).
template <typename PT, typename MT, int PP>
__global__ void __launch_bounds__(256)
k_interp_vjp(const typename V4<MT>::type *__restrict__ efi, const PT *__restrict__ px,
             const PT *__restrict__ py, const PT *__restrict__ pz, const PT *__restrict__ gEx,
             const PT *__restrict__ gEy, const PT *__restrict__ gEz, int64_t n, int nx, int ny,
             int nz, double lx_, double ly_, double lz_, double dx_, double dy_, double dz_,
             const double *__restrict__ geom, PT *__restrict__ gx, PT *__restrict__ gy,
             PT *__restrict__ gz, int accumulate, int P)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    device_geometry(geom, lx_, ly_, lz_, dx_, dy_, dz_);
    const C ddx = (C)(MT)dx_, ddy = (C)(MT)dy_, ddz = (C)(MT)dz_;
    const C nxn = ((C)px[p] - (C)(MT)lx_) / ddx;
    const C nyn = ((C)py[p] - (C)(MT)ly_) / ddy;
    const C nzn = ((C)pz[p] - (C)(MT)lz_) / ddz;
    const int ix = cell_index(nxn, nx), iy = cell_index(nyn, ny), iz = cell_index(nzn, nz);
    const C dx = nxn - (C)ix, dy = nyn - (C)iy, dz = nzn - (C)iz;
    const int64_t sy = (int64_t)P * nx, sz = (int64_t)P * nx * ny;
    const typename V4<MT>::type *e = efi + P * (ix + (int64_t)nx * iy + (int64_t)nx * ny * iz);
    typename V4<MT>::type nd[8];
    bool packed = false;
    if constexpr (sizeof(MT) == 8) {
        if (PP == 2) {
            pair_nodes(e, nd[0], nd[1]);
            pair_nodes(e + sy, nd[2], nd[3]);
            pair_nodes(e + sz, nd[4], nd[5]);
            pair_nodes(e + sz + sy, nd[6], nd[7]);
            packed = true;
        }
    }
    if (!packed) {
        nd[0] = e[0];
        nd[1] = e[1];
        nd[2] = e[sy];
        nd[3] = e[sy + 1];
        nd[4] = e[sz];
        nd[5] = e[sz + 1];
        nd[6] = e[sz + sy];
        nd[7] = e[sz + sy + 1];
    }
    C f[8], val, Gx, Gy, Gz;
    C sx = (C)0, sy_ = (C)0, sz_ = (C)0;
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = (C)nd[k].x;
    trilinear_grad(f, dx, dy, dz, val, Gx, Gy, Gz);
    const C wx = (C)gEx[p];
    sx = wx * Gx;
    sy_ = wx * Gy;
    sz_ = wx * Gz;
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = (C)nd[k].y;
    trilinear_grad(f, dx, dy, dz, val, Gx, Gy, Gz);
    const C wy = (C)gEy[p];
    sx = sx + wy * Gx;
    sy_ = sy_ + wy * Gy;
    sz_ = sz_ + wy * Gz;
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = (C)nd[k].z;
    trilinear_grad(f, dx, dy, dz, val, Gx, Gy, Gz);
    const C wz = (C)gEz[p];
    sx = (sx + wz * Gx) / ddx;
    sy_ = (sy_ + wz * Gy) / ddy;
    sz_ = (sz_ + wz * Gz) / ddz;
    if (accumulate) {
        gx[p] = (PT)((C)gx[p] + sx);
        gy[p] = (PT)((C)gy[p] + sy_);
        gz[p] = (PT)((C)gz[p] + sz_);
    } else {
        gx[p] = (PT)sx;
        gy[p] = (PT)sy_;
        gz[p] = (PT)sz_;
    }
}

// ---- bounds: grid-stride min/max per block, then one block folds them ----
constexpr int kRedBlocks = 1024;

template <typename PT>
__global__ void __launch_bounds__(256)
k_bounds_partial(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
                 int64_t n, double *__restrict__ part)
{
    __shared__ double sm[6][256];
    PT mn[3], mx[3];
    const PT *P[3] = {px, py, pz};
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    const int64_t s0 = first < n ? first : 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) mn[a] = mx[a] = P[a][s0];
    for (int64_t i = first; i < n; i += step) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const PT v = P[a][i];
            mn[a] = v < mn[a] ? v : mn[a];
            mx[a] = v > mx[a] ? v : mx[a];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sm[a][threadIdx.x] = (double)mn[a];
        sm[3 + a][threadIdx.x] = (double)mx[a];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double o = sm[a][threadIdx.x + w];
                sm[a][threadIdx.x] = o < sm[a][threadIdx.x] ? o : sm[a][threadIdx.x];
                const double O = sm[3 + a][threadIdx.x + w];
                sm[3 + a][threadIdx.x] = O > sm[3 + a][threadIdx.x] ? O : sm[3 + a][threadIdx.x];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(size_t)blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_bounds_final(double *__restrict__ part, int nblk)
{
    __shared__ double sm[6][256];
    for (int a = 0; a < 6; ++a) {
        double v = part[a];
        for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
            const double o = part[(size_t)i * 6 + a];
            v = (a < 3) ? (o < v ? o : v) : (o > v ? o : v);
        }
        sm[a][threadIdx.x] = v;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int a = 0; a < 6; ++a) {
                const double o = sm[a][threadIdx.x + w];
                double &t = sm[a][threadIdx.x];
                t = (a < 3) ? (o < t ? o : t) : (o > t ? o : t);
            }
        }
        __syncthreads();
    }
    // all reads of part[] happened before the first barrier; the final
    // values go to a separate slot past the partials
    if (threadIdx.x < 6) part[(size_t)nblk * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

// Padded bounds from the reduced extrema (fin: min x, y, z, max x, y, z):
// the Mesh3D particle-extrema rules of src/mesh.jl:120-156 -- first spacing
// in the particle type, Float64 padding by 1e-6 of it, spacing again, the
// zero patch, conversion to the mesh type -- with the same operations in the
// same order as the host path (no contraction), so both give the same bits.
// One lane per axis.
__global__ void __launch_bounds__(64)
k_bounds_geom(const double *__restrict__ fin, double *__restrict__ geom, int pdt, int mdt,
              long long g0, long long g1, long long g2)
{
#pragma clang fp contract(off)
    const int a = threadIdx.x;
    if (a >= 3) return;
    const double mn = fin[a], mx = fin[3 + a];
    const double nm1 = (double)((a == 0 ? g0 : a == 1 ? g1 : g2) - 1);
    double d0;
    if (pdt == SC_F32) d0 = (double)(((float)mx - (float)mn) / (float)nm1);
    else d0 = (mx - mn) / nm1;
    const double l = mn - 1e-6 * d0;
    const double h = mx + 1e-6 * d0;
    double d = (h - l) / nm1;
    if (d == 0.0) d = 1e-6;
    if (mdt == SC_F32) {
        geom[a] = (double)(float)l;
        geom[3 + a] = (double)(float)h;
        geom[6 + a] = (double)(float)d;
    } else {
        geom[a] = l;
        geom[3 + a] = h;
        geom[6 + a] = d;
    }
}

__global__ void __launch_bounds__(256) k_clear(uint32_t *__restrict__ p, int64_t nwords)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < nwords; k += step) p[k] = 0u;
}

int64_t blocks_for(int64_t n, int bs) { return (n + bs - 1) / bs; }

// Bounds for a launch: the host values, or (lo == NULL) the mesh's
// device-resident geometry, which the kernel reads itself.
struct LaunchGeo {
    double l[3], d[3];
    const double *dev;
};

LaunchGeo launch_geo(const Mesh *m, const double *lo, const double *dl)
{
    LaunchGeo h;
    h.dev = lo ? nullptr : m->geom;
    for (int a = 0; a < 3; ++a) {
        h.l[a] = lo ? lo[a] : 0.0;
        h.d[a] = lo ? dl[a] : 1.0;
    }
    return h;
}

template <typename PT, typename MT>
void dep(Mesh *m, const void *x, const void *y, const void *z, const void *q, int64_t n,
         const double lo[3], const double dl[3], void *out)
{
    const LaunchGeo h = launch_geo(m, lo, dl);
    hipLaunchKernelGGL((k_deposit<PT, MT>), dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                       m->ctx->stream, (const PT *)x, (const PT *)y, (const PT *)z, (const PT *)q, n,
                       (MT *)out, (int)m->n[0], (int)m->n[1], (int)m->n[2], h.l[0], h.l[1], h.l[2],
                       h.d[0], h.d[1], h.d[2], h.dev);
}

// The node-interleaved copy of m->efield in `efi` (P entries per node).
template <typename MT>
void interleave(Mesh *m, void *efi, int P)
{
    const int64_t M = m->n[0] * m->n[1] * m->n[2];
    hipLaunchKernelGGL((k_interleave<MT>), dim3((unsigned)blocks_for(P == 2 ? 4 * M : M, 256)),
                       dim3(256), 0,
                       m->ctx->stream, (const MT *)m->efield,
                       static_cast<typename V4<MT>::type *>(efi), M, (int)m->n[0], P);
}

template <typename PT, typename MT>
void itp(Mesh *m, const void *efi, int P, const void *x, const void *y, const void *z, int64_t n,
         const double lo[3], const double dl[3], void *Ex, void *Ey, void *Ez)
{
    const LaunchGeo h = launch_geo(m, lo, dl);
    auto kern = P == 2 ? k_interp<PT, MT, 2> : k_interp<PT, MT, 1>;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                       m->ctx->stream, static_cast<const typename V4<MT>::type *>(efi),
                       (const PT *)x, (const PT *)y, (const PT *)z, n, (int)m->n[0], (int)m->n[1],
                       (int)m->n[2], h.l[0], h.l[1], h.l[2], h.d[0], h.d[1], h.d[2], h.dev,
                       (PT *)Ex, (PT *)Ey, (PT *)Ez, P);
}

}  // namespace

namespace {

template <typename PT, typename MT>
void dep_vjp(Mesh *m, const void *x, const void *y, const void *z, const void *q, int64_t n,
             const double lo[3], const double dl[3], const void *g, void *gq, void *gx, void *gy,
             void *gz, int accumulate)
{
    const LaunchGeo h = launch_geo(m, lo, dl);
    hipLaunchKernelGGL((k_deposit_vjp<PT, MT>), dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                       m->ctx->stream, (const MT *)g, (const PT *)x, (const PT *)y, (const PT *)z,
                       (const PT *)q, n, (int)m->n[0], (int)m->n[1], (int)m->n[2], h.l[0], h.l[1],
                       h.l[2], h.d[0], h.d[1], h.d[2], h.dev, (PT *)gq, (PT *)gx, (PT *)gy, (PT *)gz,
                       accumulate);
}

template <typename PT, typename MT>
void itp_vjp(Mesh *m, const void *x, const void *y, const void *z, int64_t n, const double lo[3],
             const double dl[3], const void *gEx, const void *gEy, const void *gEz, void *gx,
             void *gy, void *gz, int accumulate)
{
    const LaunchGeo h = launch_geo(m, lo, dl);
    auto kern = m->efi_pair == 2 ? k_interp_vjp<PT, MT, 2> : k_interp_vjp<PT, MT, 1>;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                       m->ctx->stream, static_cast<const typename V4<MT>::type *>(m->efi),
                       (const PT *)x, (const PT *)y, (const PT *)z, (const PT *)gEx, (const PT *)gEy,
                       (const PT *)gEz, n, (int)m->n[0], (int)m->n[1], (int)m->n[2], h.l[0], h.l[1],
                       h.l[2], h.d[0], h.d[1], h.d[2], h.dev, (PT *)gx, (PT *)gy, (PT *)gz,
                       accumulate, m->efi_pair);
}

int ensure_efi(Mesh *m)
{
    if (!m->efi) {
        const int64_t M = m->n[0] * m->n[1] * m->n[2];
        // node pairs for Float64 fields (see interleave_body), unless the
        // mesh asks for a layout
        if (m->efi_layout) m->efi_pair = m->efi_layout;
        else m->efi_pair = m->dtype == SC_F64 ? 2 : 1;
        if (hipMalloc(&m->efi, (size_t)M * 4 * m->dtype * m->efi_pair) != hipSuccess)
            return set_err(m->ctx, SC_ERR_NOMEM, "device allocation of the interleaved field failed");
    }
    return SC_OK;
}

}  // namespace

int launch_clear(Ctx *ctx, void *p, int64_t bytes)
{
    if (bytes <= 0) return SC_OK;
    // all buffers cleared here are whole 4-byte words (float/double arrays)
    const int64_t words = bytes / 4;
    int64_t nb = blocks_for(words, 256);
    if (nb > 4096) nb = 4096;
    launch_begin(ctx, K_CLEAR);
    hipLaunchKernelGGL(k_clear, dim3((unsigned)nb), dim3(256), 0, ctx->stream, (uint32_t *)p, words);
    launch_end(ctx, K_CLEAR);
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

int launch_deposit_to(Mesh *m, int pdt, const void *x, const void *y, const void *z, const void *q,
                      int64_t n, const double lo[3], const double dl[3], int clear, void *out)
{
    const int64_t rho_bytes = m->n[0] * m->n[1] * m->n[2] * m->dtype;
    if (n <= 0) return clear ? launch_clear(m->ctx, out, rho_bytes) : SC_OK;
    const int rc = launch_deposit_binned(m, pdt, x, y, z, q, n, lo, dl, clear, out);
    if (rc != SC_ERR_UNSUPPORTED) return rc;
    // more tiles than the binned path holds: one atomic per node contribution
    if (clear) {
        if (int rc2 = launch_clear(m->ctx, out, rho_bytes)) return rc2;
    }
    if ((n + 255) / 256 > 0x7fffffffLL)
        return set_err(m->ctx, SC_ERR_INVALID, "too many particles for one launch");
    launch_begin(m->ctx, K_DEPOSIT);
    if (pdt == SC_F32 && m->dtype == SC_F32) dep<float, float>(m, x, y, z, q, n, lo, dl, out);
    else if (pdt == SC_F64 && m->dtype == SC_F64) dep<double, double>(m, x, y, z, q, n, lo, dl, out);
    else if (pdt == SC_F64 && m->dtype == SC_F32) dep<double, float>(m, x, y, z, q, n, lo, dl, out);
    else dep<float, double>(m, x, y, z, q, n, lo, dl, out);
    launch_end(m->ctx, K_DEPOSIT);
    SC_HIP(m->ctx, hipGetLastError());
    return SC_OK;
}

int launch_deposit(Mesh *m, int pdt, const void *x, const void *y, const void *z, const void *q,
                   int64_t n, const double lo[3], const double dl[3], int clear)
{
    // marks where this deposit starts on the stream; a following solve's
    // Green build (side stream) waits for it, see run_solve
    if (int rc = mark_deposit_start(m)) return rc;
    return launch_deposit_to(m, pdt, x, y, z, q, n, lo, dl, clear, m->rho);
}

int launch_interp_into(Mesh *m, void *efi, int P, int pdt, const void *x, const void *y,
                       const void *z, int64_t n, const double lo[3], const double dl[3], void *Ex,
                       void *Ey, void *Ez)
{
    if (n <= 0) return SC_OK;
    if ((n + 255) / 256 > 0x7fffffffLL)
        return set_err(m->ctx, SC_ERR_INVALID, "too many particles for one launch");
    // efield may have been written by the caller since the last solve, so the
    // interleaved copy is refreshed on every call
    launch_begin(m->ctx, K_INTERLEAVE);
    if (m->dtype == SC_F32) interleave<float>(m, efi, P);
    else interleave<double>(m, efi, P);
    launch_end(m->ctx, K_INTERLEAVE);
    launch_begin(m->ctx, K_INTERP);
    if (pdt == SC_F32 && m->dtype == SC_F32)
        itp<float, float>(m, efi, P, x, y, z, n, lo, dl, Ex, Ey, Ez);
    else if (pdt == SC_F64 && m->dtype == SC_F64)
        itp<double, double>(m, efi, P, x, y, z, n, lo, dl, Ex, Ey, Ez);
    else if (pdt == SC_F64 && m->dtype == SC_F32)
        itp<double, float>(m, efi, P, x, y, z, n, lo, dl, Ex, Ey, Ez);
    else
        itp<float, double>(m, efi, P, x, y, z, n, lo, dl, Ex, Ey, Ez);
    launch_end(m->ctx, K_INTERP);
    SC_HIP(m->ctx, hipGetLastError());
    return SC_OK;
}

int launch_interp(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n,
                  const double lo[3], const double dl[3], void *Ex, void *Ey, void *Ez)
{
    if (n <= 0) return SC_OK;
    if (int rc = ensure_efi(m)) return rc;
    return launch_interp_into(m, m->efi, m->efi_pair, pdt, x, y, z, n, lo, dl, Ex, Ey, Ez);
}


int launch_deposit_vjp(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                       const void *q, int64_t n, const double lo[3], const double dl[3],
                       const void *g_rho, void *gq, void *gx, void *gy, void *gz, int accumulate)
{
    if (n <= 0 || (!gq && !gx)) return SC_OK;
    if ((n + 255) / 256 > 0x7fffffffLL)
        return set_err(m->ctx, SC_ERR_INVALID, "too many particles for one launch");
    launch_begin(m->ctx, K_DEPOSIT_VJP);
    if (pdt == SC_F32 && m->dtype == SC_F32)
        dep_vjp<float, float>(m, x, y, z, q, n, lo, dl, g_rho, gq, gx, gy, gz, accumulate);
    else if (pdt == SC_F64 && m->dtype == SC_F64)
        dep_vjp<double, double>(m, x, y, z, q, n, lo, dl, g_rho, gq, gx, gy, gz, accumulate);
    else if (pdt == SC_F64 && m->dtype == SC_F32)
        dep_vjp<double, float>(m, x, y, z, q, n, lo, dl, g_rho, gq, gx, gy, gz, accumulate);
    else
        dep_vjp<float, double>(m, x, y, z, q, n, lo, dl, g_rho, gq, gx, gy, gz, accumulate);
    launch_end(m->ctx, K_DEPOSIT_VJP);
    SC_HIP(m->ctx, hipGetLastError());
    return SC_OK;
}

int launch_interp_vjp(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n,
                      const double lo[3], const double dl[3], const void *gEx, const void *gEy,
                      const void *gEz, void *g_efield, void *gx, void *gy, void *gz, int accumulate)
{
    const int64_t M = m->n[0] * m->n[1] * m->n[2];
    if (g_efield) {
        // the transpose of the gather is the deposit of the cotangents: the
        // three channels share one binning pass where the binned path applies.
        // Its start is a deposit start: a following (adjoint) solve's Green
        // build forks there and runs beside it
        if (int rc = mark_deposit_start(m)) return rc;
        const void *gE[3] = {gEx, gEy, gEz};
        void *outs[3];
        for (int c = 0; c < 3; ++c) outs[c] = static_cast<char *>(g_efield) + (size_t)c * M * m->dtype;
        int rc = n > 0 ? launch_deposit_binned_multi(m, pdt, x, y, z, gE, 3, n, lo, dl, 1, outs)
                       : SC_ERR_UNSUPPORTED;
        if (rc == SC_ERR_UNSUPPORTED) {
            for (int c = 0; c < 3; ++c)
                if ((rc = launch_deposit_to(m, pdt, x, y, z, gE[c], n, lo, dl, 1, outs[c]))) return rc;
        } else if (rc) {
            return rc;
        }
    }
    if (n <= 0 || !gx) return SC_OK;
    if ((n + 255) / 256 > 0x7fffffffLL)
        return set_err(m->ctx, SC_ERR_INVALID, "too many particles for one launch");
    if (int rc = ensure_efi(m)) return rc;
    launch_begin(m->ctx, K_INTERLEAVE);
    if (m->dtype == SC_F32) interleave<float>(m, m->efi, m->efi_pair);
    else interleave<double>(m, m->efi, m->efi_pair);
    launch_end(m->ctx, K_INTERLEAVE);
    launch_begin(m->ctx, K_INTERP_VJP);
    if (pdt == SC_F32 && m->dtype == SC_F32)
        itp_vjp<float, float>(m, x, y, z, n, lo, dl, gEx, gEy, gEz, gx, gy, gz, accumulate);
    else if (pdt == SC_F64 && m->dtype == SC_F64)
        itp_vjp<double, double>(m, x, y, z, n, lo, dl, gEx, gEy, gEz, gx, gy, gz, accumulate);
    else if (pdt == SC_F64 && m->dtype == SC_F32)
        itp_vjp<double, float>(m, x, y, z, n, lo, dl, gEx, gEy, gEz, gx, gy, gz, accumulate);
    else
        itp_vjp<float, double>(m, x, y, z, n, lo, dl, gEx, gEy, gEz, gx, gy, gz, accumulate);
    launch_end(m->ctx, K_INTERP_VJP);
    SC_HIP(m->ctx, hipGetLastError());
    return SC_OK;
}

namespace {

// Extrema of the particle coordinates into ctx->red[6*nblk .. 6*nblk+5]
// (min x, y, z, max x, y, z), stream-ordered; returns nblk through *nb.
int reduce_extrema(Ctx *ctx, int pdt, const void *x, const void *y, const void *z, int64_t n,
                   int64_t *nb)
{
    int64_t nblk = blocks_for(n, 256);
    if (nblk > kRedBlocks) nblk = kRedBlocks;
    if (!ctx->red) {
        if (hipMalloc((void **)&ctx->red, sizeof(double) * 6 * (kRedBlocks + 1)) != hipSuccess)
            return set_err(ctx, SC_ERR_NOMEM, "device allocation for the bounds reduction failed");
    }
    double *part = ctx->red;
    launch_begin(ctx, K_BOUNDS_PARTIAL);
    if (pdt == SC_F32)
        hipLaunchKernelGGL(k_bounds_partial<float>, dim3((unsigned)nblk), dim3(256), 0, ctx->stream,
                           (const float *)x, (const float *)y, (const float *)z, n, part);
    else
        hipLaunchKernelGGL(k_bounds_partial<double>, dim3((unsigned)nblk), dim3(256), 0, ctx->stream,
                           (const double *)x, (const double *)y, (const double *)z, n, part);
    launch_end(ctx, K_BOUNDS_PARTIAL);
    launch_begin(ctx, K_BOUNDS_FINAL);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, ctx->stream, part, (int)nblk);
    launch_end(ctx, K_BOUNDS_FINAL);
    *nb = nblk;
    return SC_OK;
}

}  // namespace

int launch_bounds_geom(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n)
{
    Ctx *ctx = m->ctx;
    if (!m->geom && hipMalloc((void **)&m->geom, sizeof(double) * 12) != hipSuccess) {
        m->geom = nullptr;
        return set_err(ctx, SC_ERR_NOMEM, "device allocation of the mesh geometry failed");
    }
    int64_t nblk = 0;
    if (int rc = reduce_extrema(ctx, pdt, x, y, z, n, &nblk)) return rc;
    launch_begin(ctx, K_BOUNDS_FINAL);
    hipLaunchKernelGGL(k_bounds_geom, dim3(1), dim3(64), 0, ctx->stream,
                       (const double *)(ctx->red + 6 * nblk), m->geom, pdt, m->dtype,
                       (long long)m->n[0], (long long)m->n[1], (long long)m->n[2]);
    launch_end(ctx, K_BOUNDS_FINAL);
    SC_HIP(ctx, hipGetLastError());
    m->geom_set = true;
    return SC_OK;
}

int launch_bounds(Ctx *ctx, int pdt, const void *x, const void *y, const void *z, int64_t n,
                  double mn[3], double mx[3])
{
    int64_t nblk = 0;
    if (int rc = reduce_extrema(ctx, pdt, x, y, z, n, &nblk)) return rc;
    double *part = ctx->red;
    double h[6];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, part + 6 * nblk, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_err(ctx, e, "bounds reduction");
    for (int a = 0; a < 3; ++a) {
        mn[a] = h[a];
        mx[a] = h[3 + a];
    }
    return SC_OK;
}


// ---- grouped launches: the batched interpolation, one launch per phase ----
// (the scheme of launch_deposit_grouped: a descriptor table in the
// context's device memory, written by store launches whose arguments carry
// the descriptors, then one interleave and one gather launch for all
// bunches, workgroup -> bunch by the per-phase block starts)

namespace {

enum { kIInterleave, kIGather, kIPhases };

struct IBunch {
    const void *ef;
    void *efi;
    const void *x, *y, *z;
    void *Ex, *Ey, *Ez;
    int64_t n, M;
    double l[3], d[3];
    int nx, ny, nz, P;  // P: entries per node of efi
};

struct ITable {
    unsigned start[kIPhases][kGroupMax + 1];
    IBunch b[kGroupMax];
    int nb;
};

struct IStore {
    int first, count, nb, pad;
    unsigned start[kIPhases][kGroupStore + 1];
    IBunch b[kGroupStore];
};
static_assert(sizeof(IStore) <= 2048, "store-launch arguments stay small");

// one lane per 8-byte word (vector stores)
__global__ void __launch_bounds__(256) k_igroup_store(IStore s, ITable *__restrict__ T)
{
    constexpr int kWords = (int)(sizeof(IBunch) / 8);
    static_assert(sizeof(IBunch) % 8 == 0, "descriptor of whole words");
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(s.b);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(T->b + s.first);
    for (int i = threadIdx.x; i < s.count * kWords; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kIPhases * (s.count + 1); i += blockDim.x) {
        const int ph = i / (s.count + 1), k = i - ph * (s.count + 1);
        T->start[ph][s.first + k] = s.start[ph][k];
    }
    if (threadIdx.x == 0) T->nb = s.nb;
}

__device__ __forceinline__ int igroup_of(const unsigned *__restrict__ start, int nb, unsigned blk)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename MT>
__global__ void __launch_bounds__(256) k_interleave_g(const ITable *__restrict__ T)
{
    const int b = igroup_of(T->start[kIInterleave], T->nb, blockIdx.x);
    const IBunch &d = T->b[b];
    const int64_t lb = blockIdx.x - T->start[kIInterleave][b];
    interleave_body<MT>(lb * blockDim.x + threadIdx.x, static_cast<const MT *>(d.ef),
                        static_cast<typename V4<MT>::type *>(d.efi), d.M, d.nx, d.P);
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(1, sizeof(MT) == 8 ? 4 : 8)))
k_interp_g(const ITable *__restrict__ T)
{
    const int b = igroup_of(T->start[kIGather], T->nb, blockIdx.x);
    const IBunch &d = T->b[b];
    const int64_t lb = blockIdx.x - T->start[kIGather][b];
    interp_body<PT, MT>(lb * blockDim.x + threadIdx.x,
                        static_cast<const typename V4<MT>::type *>(d.efi), (const PT *)d.x,
                        (const PT *)d.y, (const PT *)d.z, d.n, d.nx, d.ny, d.nz, d.l[0], d.l[1],
                        d.l[2], d.d[0], d.d[1], d.d[2], nullptr, (PT *)d.Ex, (PT *)d.Ey, (PT *)d.Ez,
                        d.P);
}

template <typename PT, typename MT>
void launch_igroups(Ctx *ctx, ITable *T, const unsigned *total)
{
    launch_begin(ctx, K_INTERLEAVE);
    hipLaunchKernelGGL(k_interleave_g<MT>, dim3(total[kIInterleave]), dim3(256), 0, ctx->stream,
                       (const ITable *)T);
    launch_end(ctx, K_INTERLEAVE);
    launch_begin(ctx, K_INTERP);
    hipLaunchKernelGGL((k_interp_g<PT, MT>), dim3(total[kIGather]), dim3(256), 0, ctx->stream,
                       (const ITable *)T);
    launch_end(ctx, K_INTERP);
}

}  // namespace

// Qualifies: at most kGroupMax bunches, one mesh type, every bunch with
// particles; the interleaved field of each mesh is allocated outside.
int launch_interp_grouped(Ctx *ctx, Mesh *const *ms, int nb, int pdt, const void *const *x,
                          const void *const *y, const void *const *z, const int64_t *n,
                          const double *lo, const double *dl, void *const *Ex, void *const *Ey,
                          void *const *Ez)
{
    if (!(ctx->grouped & SC_GROUP_PARTICLES) || nb < 1 || nb > kGroupMax || sizeof(ITable) > group_table_bytes())
        return SC_ERR_UNSUPPORTED;
    const int mdt = ms[0]->dtype;
    int64_t blocks = 0;  // the gather's grid over all bunches
    for (int b = 0; b < nb; ++b) {
        if (ms[b]->dtype != mdt || n[b] < 1) return SC_ERR_UNSUPPORTED;
        blocks += (n[b] + 255) / 256;
        const int64_t M = ms[b]->n[0] * ms[b]->n[1] * ms[b]->n[2];
        if (blocks > 0x7fffffffLL || (M + 255) / 256 > 0x7fffffffLL) return SC_ERR_UNSUPPORTED;
    }
    if (int rc = ensure_group_table(ctx, group_table_bytes())) return rc;
    for (int b = 0; b < nb; ++b)
        if (int rc = ensure_efi(ms[b])) return rc;
    ITable *T = static_cast<ITable *>(ctx->gtable);
    unsigned total[kIPhases] = {};
    IStore s;
    std::memset(&s, 0, sizeof(s));
    s.nb = nb;
    for (int b0 = 0; b0 < nb; b0 += kGroupStore) {
        s.first = b0;
        s.count = nb - b0 < kGroupStore ? nb - b0 : kGroupStore;
        for (int k = 0; k < s.count; ++k) {
            const int b = b0 + k;
            const Mesh *m = ms[b];
            IBunch &d = s.b[k];
            d.ef = m->efield;
            d.efi = m->efi;
            d.x = x[b];
            d.y = y[b];
            d.z = z[b];
            d.Ex = Ex[b];
            d.Ey = Ey[b];
            d.Ez = Ez[b];
            d.n = n[b];
            d.M = m->n[0] * m->n[1] * m->n[2];
            for (int a = 0; a < 3; ++a) {
                d.l[a] = lo[3 * b + a];
                d.d[a] = dl[3 * b + a];
            }
            d.nx = (int)m->n[0];
            d.ny = (int)m->n[1];
            d.nz = (int)m->n[2];
            d.P = m->efi_pair;
            const unsigned blocks[kIPhases] = {(unsigned)blocks_for(d.P == 2 ? 4 * d.M : d.M, 256),
                                               (unsigned)blocks_for(d.n, 256)};
            for (int ph = 0; ph < kIPhases; ++ph) {
                s.start[ph][k] = total[ph];
                total[ph] += blocks[ph];
            }
        }
        for (int ph = 0; ph < kIPhases; ++ph) s.start[ph][s.count] = total[ph];
        hipLaunchKernelGGL(k_igroup_store, dim3(1), dim3(256), 0, ctx->stream, s, T);
    }
    if (pdt == SC_F32 && mdt == SC_F32) launch_igroups<float, float>(ctx, T, total);
    else if (pdt == SC_F64 && mdt == SC_F64) launch_igroups<double, double>(ctx, T, total);
    else if (pdt == SC_F64 && mdt == SC_F32) launch_igroups<double, float>(ctx, T, total);
    else launch_igroups<float, double>(ctx, T, total);
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

}  // namespace sc

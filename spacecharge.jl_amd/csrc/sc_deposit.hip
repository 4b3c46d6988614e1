// sc_deposit.hip -- tile-binned cloud-in-cell deposit on gfx950.
//
// Reference: deposit! / deposit_kernel! (src/deposition.jl:106-158,
// 218-247) issue eight global atomics per particle straight into rho; under
// a Gaussian bunch most of them land on the few hundred nodes of the core
// and serialise in the atomic units (device-scope float atomics reach only
// ~18 G/s on MI355X under that contention).  Here the particles are first
// binned by the 16x16x16-cell tile that holds their cell, then each
// workgroup accumulates one chunk of a tile's particles into a private LDS
// copy of the tile's (17x17x17-node) sub-grid, and a node-parallel pass sums
// the sub-grids into rho.  No global atomic touches rho.
//
//   bin_count   per-particle tile, workgroup histogram in LDS, one global
//               atomic per (workgroup, non-empty tile)
//   bin_scan    one workgroup: tile offsets and the (tile, start, length)
//               chunk table, chunks of at most kChunk particles
//   bin_scatter recount with LDS ranks, reserve each tile's run with one
//               global atomic, write (x, y, z, q) records in tile order
//   deposit_tiles  one workgroup per chunk: Float64 LDS sub-grid, CIC with
//               the reference's arithmetic, sub-grid stored to `partials`
//   gather_tiles   one lane per node: sum over the chunks of the (up to 8)
//               tiles whose sub-grid holds the node, in a fixed order
//
// A particle's eight contributions are computed exactly as in the reference
// (same operation order, no FMA contraction); only the order of the sums
// differs, and the sums run in Float64.  Grids whose tile count exceeds
// kMaxTiles use the one-pass atomic kernel instead.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sc_internal.h"

namespace sc {

namespace {

constexpr int kTile = 16;            // cells per tile edge
constexpr int kSub = kTile + 1;      // nodes per sub-grid edge
constexpr int kSubVol = kSub * kSub * kSub;
constexpr int kMaxTiles = 4096;      // LDS histogram size
constexpr int kBinThreads = 256;
constexpr int kBinPerThread = 16;    // particles per lane in the binning passes
constexpr int kBinChunk = kBinThreads * kBinPerThread;
constexpr int kChunk = 8192;         // particles per deposit workgroup

template <typename A, typename B>
using wider_t = typename std::conditional<(sizeof(A) >= sizeof(B)), A, B>::type;

template <typename P>
struct alignas(4 * sizeof(P)) Rec {
    P x, y, z, q;
};

struct Geo {
    int nx, ny, nz;
    int tx, ty, tz;  // tiles per axis
    double lo[3], dl[3];
};

template <typename C>
__device__ __forceinline__ int cell_of(C norm, int n)
{
    const C f = floor(norm);
    if (!(f >= (C)0)) return 0;
    if (f > (C)(n - 2)) return n - 2;
    return (int)f;
}

// Tile of a particle; the arithmetic matches the deposit's normalisation.
template <typename PT, typename MT>
__device__ __forceinline__ int tile_of(PT px, PT py, PT pz, const Geo &g)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    const C nxn = ((C)px - (C)(MT)g.lo[0]) / (C)(MT)g.dl[0];
    const C nyn = ((C)py - (C)(MT)g.lo[1]) / (C)(MT)g.dl[1];
    const C nzn = ((C)pz - (C)(MT)g.lo[2]) / (C)(MT)g.dl[2];
    const int ix = cell_of(nxn, g.nx), iy = cell_of(nyn, g.ny), iz = cell_of(nzn, g.nz);
    return ix / kTile + g.tx * (iy / kTile + g.ty * (iz / kTile));
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(kBinThreads)
k_bin_count(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
            int64_t n, Geo g, unsigned *__restrict__ counts)
{
    __shared__ unsigned hist[kMaxTiles];
    const int ntiles = g.tx * g.ty * g.tz;
    for (int t = threadIdx.x; t < ntiles; t += kBinThreads) hist[t] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kBinChunk;
#pragma unroll 4
    for (int s = 0; s < kBinPerThread; ++s) {
        const int64_t p = base + s * kBinThreads + threadIdx.x;
        if (p < n) atomicAdd(&hist[tile_of<PT, MT>(px[p], py[p], pz[p], g)], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += kBinThreads) {
        const unsigned c = hist[t];
        if (c) atomicAdd(&counts[t], c);
    }
}

// One workgroup.  counts -> cursor (exclusive offsets), each tile's first
// chunk, and the chunk table.
__global__ void __launch_bounds__(1024)
k_bin_scan(const unsigned *__restrict__ counts, unsigned *__restrict__ cursor,
           unsigned *__restrict__ chunk_base, int4 *__restrict__ chunks,
           unsigned *__restrict__ nchunks, int ntiles)
{
    __shared__ unsigned s_off[kMaxTiles];
    __shared__ unsigned s_chk[kMaxTiles];
    __shared__ unsigned s_sum[2][1024];
    const int T = blockDim.x;
    const int per = (ntiles + T - 1) / T;
    const int t0 = threadIdx.x * per;
    unsigned a = 0, b = 0;
    for (int i = 0; i < per; ++i) {
        const int t = t0 + i;
        if (t < ntiles) {
            const unsigned c = counts[t];
            a += c;
            b += (c + kChunk - 1) / kChunk;
        }
    }
    s_sum[0][threadIdx.x] = a;
    s_sum[1][threadIdx.x] = b;
    __syncthreads();
    // inclusive Hillis-Steele scan of the per-lane sums
    for (int d = 1; d < T; d <<= 1) {
        unsigned va = 0, vb = 0;
        if ((int)threadIdx.x >= d) {
            va = s_sum[0][threadIdx.x - d];
            vb = s_sum[1][threadIdx.x - d];
        }
        __syncthreads();
        s_sum[0][threadIdx.x] += va;
        s_sum[1][threadIdx.x] += vb;
        __syncthreads();
    }
    unsigned oa = threadIdx.x ? s_sum[0][threadIdx.x - 1] : 0u;
    unsigned ob = threadIdx.x ? s_sum[1][threadIdx.x - 1] : 0u;
    for (int i = 0; i < per; ++i) {
        const int t = t0 + i;
        if (t < ntiles) {
            const unsigned c = counts[t];
            s_off[t] = oa;
            s_chk[t] = ob;
            cursor[t] = oa;
            chunk_base[t] = ob;
            oa += c;
            ob += (c + kChunk - 1) / kChunk;
        }
    }
    if (threadIdx.x == T - 1) *nchunks = s_sum[1][T - 1];
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += T) {
        const unsigned c = counts[t];
        const unsigned k0 = s_chk[t];
        const unsigned nk = (c + kChunk - 1) / kChunk;
        for (unsigned k = 0; k < nk; ++k) {
            const unsigned start = s_off[t] + k * kChunk;
            const unsigned len = (c - k * kChunk) < (unsigned)kChunk ? (c - k * kChunk) : kChunk;
            chunks[k0 + k] = make_int4(t, (int)start, (int)len, 0);
        }
    }
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(kBinThreads)
k_bin_scatter(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
              const PT *__restrict__ pq, int64_t n, Geo g, unsigned *__restrict__ cursor,
              Rec<PT> *__restrict__ sorted)
{
    __shared__ unsigned hist[kMaxTiles];
    const int ntiles = g.tx * g.ty * g.tz;
    for (int t = threadIdx.x; t < ntiles; t += kBinThreads) hist[t] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kBinChunk;
    int tl[kBinPerThread];
    unsigned rk[kBinPerThread];
#pragma unroll
    for (int s = 0; s < kBinPerThread; ++s) {
        const int64_t p = base + s * kBinThreads + threadIdx.x;
        tl[s] = -1;
        rk[s] = 0;
        if (p < n) {
            tl[s] = tile_of<PT, MT>(px[p], py[p], pz[p], g);
            rk[s] = atomicAdd(&hist[tl[s]], 1u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += kBinThreads) {
        const unsigned c = hist[t];
        if (c) hist[t] = atomicAdd(&cursor[t], c);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kBinPerThread; ++s) {
        const int64_t p = base + s * kBinThreads + threadIdx.x;
        if (tl[s] >= 0) {
            Rec<PT> r;
            r.x = px[p];
            r.y = py[p];
            r.z = pz[p];
            r.q = pq[p];
            sorted[hist[tl[s]] + rk[s]] = r;
        }
    }
}

// CIC contributions of one record, in the reference's arithmetic
// (src/deposition.jl:38-84), added to the Float64 LDS sub-grid.  The LDS
// accumulator is Float64 for every mesh type: gfx950 executes LDS atomic
// adds on doubles several times faster than on floats, and each node's
// chunk sum is then exact to double rounding before it meets the mesh type.
template <typename PT, typename MT>
__device__ __forceinline__ void deposit_one(const Rec<PT> &r, const Geo &g, int ox, int oy, int oz,
                                            double *__restrict__ sub)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    const C nxn = ((C)r.x - (C)(MT)g.lo[0]) / (C)(MT)g.dl[0];
    const C nyn = ((C)r.y - (C)(MT)g.lo[1]) / (C)(MT)g.dl[1];
    const C nzn = ((C)r.z - (C)(MT)g.lo[2]) / (C)(MT)g.dl[2];
    const int ix = cell_of(nxn, g.nx), iy = cell_of(nyn, g.ny), iz = cell_of(nzn, g.nz);
    const C dx = nxn - (C)ix, dy = nyn - (C)iy, dz = nzn - (C)iz;
    const C wx0 = (C)1 - dx, wx1 = dx, wy0 = (C)1 - dy, wy1 = dy, wz0 = (C)1 - dz, wz1 = dz;
    const C c = (C)r.q;
    double *b = sub + (ix - ox) + kSub * ((iy - oy) + kSub * (iz - oz));
    constexpr int sy = kSub, sz = kSub * kSub;
    atomicAdd(b, (double)(((c * wx0) * wy0) * wz0));
    atomicAdd(b + 1, (double)(((c * wx1) * wy0) * wz0));
    atomicAdd(b + sy, (double)(((c * wx0) * wy1) * wz0));
    atomicAdd(b + sy + 1, (double)(((c * wx1) * wy1) * wz0));
    atomicAdd(b + sz, (double)(((c * wx0) * wy0) * wz1));
    atomicAdd(b + sz + 1, (double)(((c * wx1) * wy0) * wz1));
    atomicAdd(b + sz + sy, (double)(((c * wx0) * wy1) * wz1));
    atomicAdd(b + sz + sy + 1, (double)(((c * wx1) * wy1) * wz1));
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(256)
k_deposit_tiles(const Rec<PT> *__restrict__ sorted, const int4 *__restrict__ chunks,
                const unsigned *__restrict__ nchunks, Geo g, double *__restrict__ partials)
{
    __shared__ double sub[kSubVol];
    if (blockIdx.x >= *nchunks) return;
    const int4 ch = chunks[blockIdx.x];
    const int t = ch.x;
    const int ox = (t % g.tx) * kTile, oy = ((t / g.tx) % g.ty) * kTile, oz = (t / (g.tx * g.ty)) * kTile;
    for (int i = threadIdx.x; i < kSubVol; i += blockDim.x) sub[i] = 0.0;
    __syncthreads();
    // records are fetched kBatch at a time so their loads overlap
    constexpr int kBatch = 8;
    const Rec<PT> *src = sorted + ch.y;
    for (int i0 = 0; i0 < ch.z; i0 += kBatch * 256) {
        Rec<PT> rb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = i0 + u * 256 + (int)threadIdx.x;
            if (i < ch.z) rb[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (i0 + u * 256 + (int)threadIdx.x < ch.z) deposit_one<PT, MT>(rb[u], g, ox, oy, oz, sub);
        }
    }
    __syncthreads();
    double *out = partials + (size_t)blockIdx.x * kSubVol;
    for (int i = threadIdx.x; i < kSubVol; i += blockDim.x) out[i] = sub[i];
}

// Node-parallel sum of the chunk sub-grids: node g lies in at most two tiles
// per axis (two when it is on a tile face), each tile contributes the sum
// over its chunks.  Plain stores, fixed summation order; with accumulate
// the sum is added to the existing rho (deposit!(...; clear=false)).
template <typename MT>
__global__ void __launch_bounds__(256)
k_gather_tiles(const double *__restrict__ partials, const unsigned *__restrict__ counts,
               const unsigned *__restrict__ chunk_base, Geo g, MT *__restrict__ rho, int accumulate)
{
    const int64_t M = (int64_t)g.nx * g.ny * g.nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int gx = (int)(i % g.nx);
    const int gy = (int)((i / g.nx) % g.ny);
    const int gz = (int)(i / ((int64_t)g.nx * g.ny));
    int cx[2], cy[2], cz[2], nxc = 0, nyc = 0, nzc = 0;
    if (gx / kTile < g.tx) cx[nxc++] = gx / kTile;
    if (gx % kTile == 0 && gx >= kTile) cx[nxc++] = gx / kTile - 1;
    if (gy / kTile < g.ty) cy[nyc++] = gy / kTile;
    if (gy % kTile == 0 && gy >= kTile) cy[nyc++] = gy / kTile - 1;
    if (gz / kTile < g.tz) cz[nzc++] = gz / kTile;
    if (gz % kTile == 0 && gz >= kTile) cz[nzc++] = gz / kTile - 1;
    double s = 0.0;
    for (int c = 0; c < nzc; ++c)
        for (int b = 0; b < nyc; ++b)
            for (int a = 0; a < nxc; ++a) {
                const int t = cx[a] + g.tx * (cy[b] + g.ty * cz[c]);
                const unsigned nk = (counts[t] + kChunk - 1) / kChunk;
                if (nk == 0) continue;
                const int local = (gx - kTile * cx[a]) + kSub * ((gy - kTile * cy[b]) + kSub * (gz - kTile * cz[c]));
                const double *p = partials + (size_t)chunk_base[t] * kSubVol + local;
                // four independent partial sums keep several loads in flight
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                unsigned k = 0;
                for (; k + 4 <= nk; k += 4) {
                    s0 += p[(size_t)k * kSubVol];
                    s1 += p[(size_t)(k + 1) * kSubVol];
                    s2 += p[(size_t)(k + 2) * kSubVol];
                    s3 += p[(size_t)(k + 3) * kSubVol];
                }
                for (; k < nk; ++k) s0 += p[(size_t)k * kSubVol];
                s += (s0 + s1) + (s2 + s3);
            }
    rho[i] = accumulate ? (MT)((double)rho[i] + s) : (MT)s;
}

int ensure_bins(Mesh *m, int64_t n, int pdt, int ntiles)
{
    DepositBins &B = m->bins;
    const size_t rec = 4 * (size_t)pdt;
    const size_t need = (size_t)n * rec;
    auto fail = [&]() {
        return set_err(m->ctx, SC_ERR_NOMEM, "device allocation of the deposit bins failed");
    };
    if (B.sorted_bytes < need) {
        if (B.sorted) (void)hipFree(B.sorted);
        B.sorted = nullptr;
        B.sorted_bytes = 0;
        if (hipMalloc(&B.sorted, need) != hipSuccess) return fail();
        B.sorted_bytes = need;
    }
    const int64_t maxchunks = ntiles + n / kChunk + 1;
    if (B.chunk_cap < maxchunks) {
        if (B.chunks) (void)hipFree(B.chunks);
        if (B.partials) (void)hipFree(B.partials);
        B.chunks = B.partials = nullptr;
        B.chunk_cap = 0;
        if (hipMalloc(&B.chunks, sizeof(int4) * maxchunks) != hipSuccess) return fail();
        // one sub-grid per chunk, in the wider mesh type (double covers both)
        if (hipMalloc(&B.partials, sizeof(double) * kSubVol * (size_t)maxchunks) != hipSuccess)
            return fail();
        B.chunk_cap = maxchunks;
    }
    if (!B.counts) {
        // counts | cursor | chunk_base | nchunks
        if (hipMalloc((void **)&B.counts, sizeof(unsigned) * (3 * kMaxTiles + 1)) != hipSuccess)
            return fail();
    }
    return SC_OK;
}

template <typename PT, typename MT>
int binned(Mesh *m, const void *x, const void *y, const void *z, const void *q, int64_t n,
           const Geo &g, int clear)
{
    Ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    DepositBins &B = m->bins;
    const int ntiles = g.tx * g.ty * g.tz;
    unsigned *counts = B.counts, *cursor = B.counts + kMaxTiles;
    unsigned *cbase = B.counts + 2 * kMaxTiles, *nch = B.counts + 3 * kMaxTiles;
    int4 *chunks = static_cast<int4 *>(B.chunks);
    const unsigned nbin = (unsigned)((n + kBinChunk - 1) / kBinChunk);
    if (int rc = launch_clear(ctx, counts, sizeof(unsigned) * ntiles)) return rc;
    launch_begin(ctx, K_BIN_COUNT);
    hipLaunchKernelGGL((k_bin_count<PT, MT>), dim3(nbin), dim3(kBinThreads), 0, st, (const PT *)x,
                       (const PT *)y, (const PT *)z, n, g, counts);
    launch_end(ctx, K_BIN_COUNT);
    launch_begin(ctx, K_BIN_SCAN);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, (const unsigned *)counts, cursor,
                       cbase, chunks, nch, ntiles);
    launch_end(ctx, K_BIN_SCAN);
    launch_begin(ctx, K_BIN_SCATTER);
    hipLaunchKernelGGL((k_bin_scatter<PT, MT>), dim3(nbin), dim3(kBinThreads), 0, st,
                       (const PT *)x, (const PT *)y, (const PT *)z, (const PT *)q, n, g, cursor,
                       static_cast<Rec<PT> *>(B.sorted));
    launch_end(ctx, K_BIN_SCATTER);
    const unsigned maxchunks = (unsigned)(ntiles + n / kChunk + 1);
    launch_begin(ctx, K_DEPOSIT_TILES);
    hipLaunchKernelGGL((k_deposit_tiles<PT, MT>), dim3(maxchunks), dim3(256), 0, st,
                       (const Rec<PT> *)B.sorted, (const int4 *)chunks, (const unsigned *)nch, g,
                       static_cast<double *>(B.partials));
    launch_end(ctx, K_DEPOSIT_TILES);
    const int64_t M = (int64_t)g.nx * g.ny * g.nz;
    launch_begin(ctx, K_GATHER_TILES);
    hipLaunchKernelGGL((k_gather_tiles<MT>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st,
                       (const double *)B.partials, (const unsigned *)counts, (const unsigned *)cbase, g,
                       static_cast<MT *>(m->rho), clear ? 0 : 1);
    launch_end(ctx, K_GATHER_TILES);
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

}  // namespace

void free_bins(Mesh *m)
{
    DepositBins &B = m->bins;
    if (B.sorted) (void)hipFree(B.sorted);
    if (B.chunks) (void)hipFree(B.chunks);
    if (B.partials) (void)hipFree(B.partials);
    if (B.counts) (void)hipFree(B.counts);
    B = DepositBins();
}

// Binned deposit; returns SC_ERR_UNSUPPORTED (without launching anything)
// when the grid has more tiles than the LDS histogram holds.  Writes every
// node of rho (clear != 0) or adds to it (clear == 0).
int launch_deposit_binned(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                          const void *q, int64_t n, const double lo[3], const double dl[3],
                          int clear)
{
    Geo g;
    g.nx = (int)m->n[0];
    g.ny = (int)m->n[1];
    g.nz = (int)m->n[2];
    g.tx = (g.nx - 1 + kTile - 1) / kTile;
    g.ty = (g.ny - 1 + kTile - 1) / kTile;
    g.tz = (g.nz - 1 + kTile - 1) / kTile;
    if ((int64_t)g.tx * g.ty * g.tz > kMaxTiles) return SC_ERR_UNSUPPORTED;
    if (n > 0x7fffffffLL) return SC_ERR_UNSUPPORTED;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo[a];
        g.dl[a] = dl[a];
    }
    if (int rc = ensure_bins(m, n, pdt, g.tx * g.ty * g.tz)) return rc;
    if (pdt == SC_F32 && m->dtype == SC_F32) return binned<float, float>(m, x, y, z, q, n, g, clear);
    if (pdt == SC_F64 && m->dtype == SC_F64) return binned<double, double>(m, x, y, z, q, n, g, clear);
    if (pdt == SC_F64 && m->dtype == SC_F32) return binned<double, float>(m, x, y, z, q, n, g, clear);
    return binned<float, double>(m, x, y, z, q, n, g, clear);
}

}  // namespace sc

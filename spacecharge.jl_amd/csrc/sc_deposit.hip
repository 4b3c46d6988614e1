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
//   bin_count   per-particle tile; each workgroup histograms one contiguous
//               slice of the particles in LDS and stores it as one column of
//               a [tile][slice] count matrix (no global atomics)
//   bin_rowscan one wave per tile: prefix of its row over the slices, and
//               the tile's total; on meshes of one binning range also the
//               tile's run and chunk table entries (one 64-bit atomic per
//               tile reserves them), and bin_scan is not launched
//   bin_scan    one workgroup: tile offsets and the (tile, start, length)
//               chunk table, chunks of at most 2^chunk_log2 particles
//               (meshes of several binning ranges, the sort permutation,
//               grouped launches)
//   bin_scatter each slice writes (x, y, z, q) records into its reserved run
//               of each tile, slots from an LDS cursor per tile
//   deposit_tiles  one workgroup per chunk: Float64 LDS sub-grid, CIC with
//               the reference's arithmetic, sub-grid stored to `partials`
//   tile_reduce    adds the further chunks of each multi-chunk tile to its
//               first chunk's sub-grid (slot = tile index), zeroes empty
//               tiles' slots; one lane per sub-grid node
//   gather_tiles   one lane per node: sum over the (up to 8) tiles whose
//               sub-grid holds the node, in a fixed order
//
// A particle's eight contributions are computed exactly as in the reference
// (same operation order, no FMA contraction); only the order of the sums
// differs, and the sums run in Float64.  Meshes with more tiles than one
// LDS histogram holds (kRange) bin in ranges of kRange tiles, one per
// workgroup row of the count and scatter launches; grids whose tile count
// exceeds kMaxTiles use the one-pass atomic kernel instead.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

#include "sc_internal.h"

namespace sc {

namespace {

constexpr int kTile = 16;            // cells per tile edge
constexpr int kSub = kTile + 1;      // nodes per sub-grid edge
constexpr int kSubVol = kSub * kSub * kSub;
constexpr int kRange = 4096;         // tiles per LDS histogram (one binning range)
constexpr int kMaxTiles = 65535;     // tiles of the binned path (a grid axis of the reduction)
constexpr int64_t kCmatBudget = (int64_t)1 << 22;  // count-matrix entries (slices x tiles)
constexpr int kBinThreads = 256;
constexpr int kBatch = 8;            // particles a lane loads before computing any
// particles per deposit workgroup: 2^chunk_log2, a per-mesh setting
// (Mesh::chunk_log2, kChunkLog2Min..kChunkLog2Max)
constexpr int kDepThreads = 512;     // deposit workgroup size
constexpr int kBinMaxSlices = 2048;  // binning workgroups (columns of the count matrix)

inline int div_up4(int a) { return (a + 3) / 4; }

// The deposit chain runs beside the Green-function build of the solve (on
// the library's second stream) and is the step's critical path while the
// build has slack.  Its small latency-bound kernels (row scan, tile scan,
// reduction, gather) take instruction-issue priority over the build's waves
// where the two share a SIMD; the binning and tile-deposit kernels do not
// (measured: slower with it).
constexpr int kChainPrio = 2;

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
    const double *dev;  // device-resident geometry (lo at 0, delta at 6) when set
};

// Takes the lower bounds and spacing from device memory when the mesh
// geometry lives there.
__device__ __forceinline__ void resolve_geo(Geo &g)
{
    if (g.dev) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = g.dev[a];
            g.dl[a] = g.dev[6 + a];
        }
    }
}

template <typename C>
__device__ __forceinline__ int cell_of(C norm, int n)
{
    const C f = floor(norm);
    if (!(f >= (C)0)) return 0;
    if (f > (C)(n - 2)) return n - 2;
    return (int)f;
}

// Tile of a particle; the arithmetic matches the deposit's normalisation.
// `ext` is set when a weight leaves [0, 1] (a coordinate outside the mesh
// extent: the clamped cell extrapolates) or is NaN.
template <typename PT, typename MT>
__device__ __forceinline__ int tile_of(PT px, PT py, PT pz, const Geo &g, bool &ext)
{
#pragma clang fp contract(off)
    using C = wider_t<PT, MT>;
    const C nxn = ((C)px - (C)(MT)g.lo[0]) / (C)(MT)g.dl[0];
    const C nyn = ((C)py - (C)(MT)g.lo[1]) / (C)(MT)g.dl[1];
    const C nzn = ((C)pz - (C)(MT)g.lo[2]) / (C)(MT)g.dl[2];
    const int ix = cell_of(nxn, g.nx), iy = cell_of(nyn, g.ny), iz = cell_of(nzn, g.nz);
    ext = !(nxn >= (C)0 && nxn <= (C)(g.nx - 1) && nyn >= (C)0 && nyn <= (C)(g.ny - 1) &&
            nzn >= (C)0 && nzn <= (C)(g.nz - 1));
    return ix / kTile + g.tx * (iy / kTile + g.ty * (iz / kTile));
}

template <typename PT, typename MT>
__device__ __forceinline__ int tile_of(PT px, PT py, PT pz, const Geo &g)
{
    bool ext;
    return tile_of<PT, MT>(px, py, pz, g, ext);
}

// The binning workgroups split the particles into G contiguous slices of
// `slice` particles; all three binning kernels use the same split.
struct Split {
    int64_t n, slice;
    int G;
};

// Per-slice tile histogram, written as column w of the [tile][G] matrix.
// Each lane loads kBatch particles before it computes any, so the loads of
// a wave overlap.
// Tiles [r0, r0 + rn) only (rn <= kRange): the workgroup row of this range.
template <typename PT, typename MT>
__device__ __forceinline__ void bin_count_body(unsigned w, const PT *__restrict__ px,
                                               const PT *__restrict__ py, const PT *__restrict__ pz,
                                               Split sp, Geo g, unsigned *__restrict__ cmat, int r0,
                                               int rn)
{
    resolve_geo(g);
    __shared__ unsigned hist[kRange];
    for (int t = threadIdx.x; t < rn; t += kBinThreads) hist[t] = 0u;
    __syncthreads();
    const int64_t first = (int64_t)w * sp.slice;
    const int64_t end = first + sp.slice < sp.n ? first + sp.slice : sp.n;
    for (int64_t base = first; base < end; base += kBatch * kBinThreads) {
        PT x[kBatch], y[kBatch], z[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            int64_t p = base + s * kBinThreads + threadIdx.x;
            p = p < end ? p : end - 1;  // unconditional loads stay in flight together
            x[s] = px[p];
            y[s] = py[p];
            z[s] = pz[p];
        }
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            const unsigned lt = (unsigned)(tile_of<PT, MT>(x[s], y[s], z[s], g) - r0);
            const bool ok = base + s * kBinThreads + threadIdx.x < end && lt < (unsigned)rn;
            atomicAdd(&hist[ok ? lt : 0u], ok ? 1u : 0u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < rn; t += kBinThreads)
        cmat[(size_t)(r0 + t) * sp.G + w] = hist[t];
}

// Grid (slices, tile ranges); the slices go to workgroups in XCD order
// (xcd_task), so that the workgroups of one XCD own neighbouring columns of
// the count matrix and share its lines in their L2.  A non-null `resv` (the
// row scan's reservation counter, k_bin_rowscan) is zeroed here, one launch
// ahead of its first use.
template <typename PT, typename MT>
__global__ void __launch_bounds__(kBinThreads)
k_bin_count(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
            Split sp, Geo g, unsigned *__restrict__ cmat, unsigned long long *__restrict__ resv)
{
    const int ntiles = g.tx * g.ty * g.tz;
    const int r0 = (int)blockIdx.y * kRange;
    if (resv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *resv = 0ull;
    bin_count_body<PT, MT>(xcd_task((int)blockIdx.x, (int)gridDim.x), px, py, pz, sp, g, cmat, r0,
                           ntiles - r0 < kRange ? ntiles - r0 : kRange);
}

// What the row scan also does for a mesh of one binning range (then
// k_bin_scan is not launched): each tile's wave reserves the tile's run of
// records and its chunk table entries with one 64-bit atomic on `resv`
// (records in the low half, chunks in the high half; zeroed by k_bin_count)
// and writes the tile's start (`cursor`), first extra sub-grid slot, chunk
// entries and cleared extrapolation flag.  The runs then lie in the order
// the tiles reserved them, but chunk k of a tile still holds the records at
// positions k 2^chunk_log2 .. of the tile's run, and the reduction adds a
// tile's chunks in the same order, so the sums are those of the tile-ordered
// layout of k_bin_scan.  Sub-grid slots:
// a tile's first chunk uses slot t, its further chunks ntiles + c + k for
// chunk entries c .. c + nk - 1 (slots below 2 ntiles + chunks).
struct RowscanChunks {
    unsigned long long *resv;  // null: k_bin_scan does this
    unsigned *cursor, *extra_base, *ext_flags;
    int4 *chunks;
    int chunk_log2;
};

// One wave per tile: exclusive prefix of the tile's row over the slices (in
// place), the tile's total, and with rc.resv the reservation above.
__device__ __forceinline__ void bin_rowscan_body(unsigned bx, unsigned *__restrict__ cmat, int G,
                                                 int ntiles, unsigned *__restrict__ tot,
                                                 RowscanChunks rc = RowscanChunks{})
{
    constexpr int kPer = kBinMaxSlices / 64;
    const int lane = threadIdx.x & 63;
    const int t = (int)bx * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    unsigned *row = cmat + (size_t)t * G;
    // lane l owns the V consecutive slices [l V, l V + V); all loads first
    const int V = (G + 63) / 64;
    const int w0 = lane * V;
    unsigned v[kPer];
    unsigned s = 0u;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        v[i] = (i < V && w0 + i < G) ? row[w0 + i] : 0u;
        s += v[i];
    }
    unsigned incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    unsigned run = incl - s;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        if (i < V && w0 + i < G) row[w0 + i] = run;
        run += v[i];
    }
    if (lane == 63) tot[t] = incl;
    if (!rc.resv) return;
    // the tile's total is lane 63's inclusive prefix
    const unsigned total = __shfl(incl, 63, 64);
    const unsigned chunk = 1u << rc.chunk_log2;
    const unsigned nk = (total + chunk - 1) >> rc.chunk_log2;
    unsigned long long base = 0ull;
    if (lane == 0 && total)
        base = atomicAdd(rc.resv, ((unsigned long long)nk << 32) | (unsigned long long)total);
    base = __shfl(base, 0, 64);
    const unsigned r = (unsigned)base, c = (unsigned)(base >> 32);
    if (lane == 0) {
        rc.cursor[t] = r;
        rc.extra_base[t] = (unsigned)ntiles + c + 1u;
        rc.ext_flags[t] = 0u;
    }
    for (unsigned k = (unsigned)lane; k < nk; k += 64u) {
        const unsigned len = total - k * chunk < chunk ? total - k * chunk : chunk;
        const unsigned slot = k == 0 ? (unsigned)t : (unsigned)ntiles + c + k;
        rc.chunks[c + k] = make_int4(t, (int)(r + k * chunk), (int)len, (int)slot);
    }
}

__global__ void __launch_bounds__(256)
k_bin_rowscan(unsigned *__restrict__ cmat, int G, int ntiles, unsigned *__restrict__ tot,
              RowscanChunks rc)
{
    __builtin_amdgcn_s_setprio(kChainPrio);
    bin_rowscan_body(blockIdx.x, cmat, G, ntiles, tot, rc);
}

// One workgroup.  Per-tile totals -> tile start (exclusive offsets), each
// tile's first chunk, the chunk table; clears the extrapolation flags.
// Tiles go in blocks of 256 * 16, each block's prefix offset by the totals
// of the blocks before it.
// (Folding this into k_bin_rowscan as "last workgroup done" costs more on
// MI355X than the launch it saves: the device-scope fences write back L2.)
__device__ __forceinline__ void bin_scan_body(const unsigned *__restrict__ counts,
                                              unsigned *__restrict__ cursor,
                                              unsigned *__restrict__ extra_base,
                                              int4 *__restrict__ chunks,
                                              unsigned *__restrict__ nchunks,
                                              unsigned *__restrict__ ext_flags, int ntiles,
                                              int chunk_log2)
{
    const unsigned chunk = 1u << chunk_log2;
    constexpr int kPer = 16;  // at most this many tiles per lane per block
    __shared__ unsigned s_w[3][4];
    __shared__ unsigned s_carry[3];  // particles, chunks, further chunks before this block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x < 3) s_carry[threadIdx.x] = 0u;
    // one block of 256 per tiles when ntiles <= 256 kPer, spread over all lanes
    const int per = (ntiles + 255) / 256 < kPer ? (ntiles + 255) / 256 : kPer;
    for (int blk = 0; blk < ntiles; blk += 256 * per) {
        // thread i owns the per consecutive tiles [blk + i per, blk + i per + per)
        const int t0 = blk + (int)threadIdx.x * per;
        unsigned c[kPer];
        unsigned a = 0u, b = 0u, x = 0u;  // particles, chunks, chunks after each tile's first
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            c[i] = (i < per && t0 + i < ntiles) ? counts[t0 + i] : 0u;
            const unsigned nk = (c[i] + chunk - 1) >> chunk_log2;
            a += c[i];
            b += nk;
            x += nk ? nk - 1 : 0u;
        }
        unsigned ia = a, ib = b, ix = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned oa = __shfl_up(ia, d, 64), ob = __shfl_up(ib, d, 64),
                           ox = __shfl_up(ix, d, 64);
            if (lane >= d) {
                ia += oa;
                ib += ob;
                ix += ox;
            }
        }
        if (lane == 63) {
            s_w[0][wv] = ia;
            s_w[1][wv] = ib;
            s_w[2][wv] = ix;
        }
        __syncthreads();
        unsigned oa = s_carry[0] + ia - a, ob = s_carry[1] + ib - b, ox = s_carry[2] + ix - x;
        for (int k = 0; k < wv; ++k) {
            oa += s_w[0][k];
            ob += s_w[1][k];
            ox += s_w[2][k];
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int t = t0 + i;
            if (i < per && t < ntiles) {
                ext_flags[t] = 0u;
                cursor[t] = oa;
                // sub-grid slots: the tile's first chunk uses slot t (which then
                // holds the tile's sum), its further chunks ntiles + ox, ...
                extra_base[t] = (unsigned)ntiles + ox;
                const unsigned nk = (c[i] + chunk - 1) >> chunk_log2;
                for (unsigned k = 0; k < nk; ++k) {
                    const unsigned len = (c[i] - k * chunk) < chunk ? (c[i] - k * chunk) : chunk;
                    const unsigned slot = k == 0 ? (unsigned)t : (unsigned)ntiles + ox + k - 1;
                    chunks[ob + k] = make_int4(t, (int)(oa + k * chunk), (int)len, (int)slot);
                }
                oa += c[i];
                ob += nk;
                ox += nk ? nk - 1 : 0u;
            }
        }
        __syncthreads();  // every lane has read s_w and the carry
        if (threadIdx.x == 0) {
            for (int k = 0; k < 4; ++k) {
                s_carry[0] += s_w[0][k];
                s_carry[1] += s_w[1][k];
                s_carry[2] += s_w[2][k];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *nchunks = s_carry[1];
}

__global__ void __launch_bounds__(256)
k_bin_scan(const unsigned *__restrict__ counts, unsigned *__restrict__ cursor,
           unsigned *__restrict__ extra_base, int4 *__restrict__ chunks,
           unsigned *__restrict__ nchunks, unsigned *__restrict__ ext_flags, int ntiles,
           int chunk_log2)
{
    __builtin_amdgcn_s_setprio(kChainPrio);
    bin_scan_body(counts, cursor, extra_base, chunks, nchunks, ext_flags, ntiles, chunk_log2);
}

// Each slice's particles go to its reserved run inside each tile: the LDS
// cursor of tile t starts at start[t] + (particles of t in earlier slices),
// and an LDS atomic hands out the slots.  The same split and the same
// tile_of as k_bin_count make every run exactly as long as counted.  Also
// flags the tiles that hold a particle with weights outside [0, 1] and
// takes max|q| (both steer the deposit's accumulator).
// With NCH = 3 the charges q1, q2 of two more channels are written in the
// same order to q1s, q2s and their per-slice max|q| after channel 0's
// (kBinMaxSlices apart).
template <typename PT, typename MT, int NCH>
__device__ __forceinline__ void bin_scatter_body(
    unsigned w, const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
    const PT *__restrict__ pq, Split sp, Geo g, const unsigned *__restrict__ cmat,
    const unsigned *__restrict__ start, Rec<PT> *__restrict__ sorted,
    unsigned *__restrict__ ext_flags, unsigned long long *__restrict__ slice_qmax,
    const PT *__restrict__ q1, const PT *__restrict__ q2, PT *__restrict__ q1s, PT *__restrict__ q2s,
    int r0, int rn)
{
    resolve_geo(g);
    __shared__ unsigned cur[kRange];
    __shared__ unsigned char ext_lds[kRange];
    for (int t = threadIdx.x; t < rn; t += kBinThreads) {
        cur[t] = start[r0 + t] + cmat[(size_t)(r0 + t) * sp.G + w];
        ext_lds[t] = 0;
    }
    __syncthreads();
    double qm = 0.0, qm1 = 0.0, qm2 = 0.0;
    const int64_t first = (int64_t)w * sp.slice;
    const int64_t end = first + sp.slice < sp.n ? first + sp.slice : sp.n;
    for (int64_t base = first; base < end; base += kBatch * kBinThreads) {
        Rec<PT> r[kBatch];
        PT c1[kBatch], c2[kBatch];
        unsigned slot[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            int64_t p = base + s * kBinThreads + threadIdx.x;
            p = p < end ? p : end - 1;  // unconditional loads stay in flight together
            r[s].x = px[p];
            r[s].y = py[p];
            r[s].z = pz[p];
            r[s].q = pq[p];
            if constexpr (NCH == 3) {
                c1[s] = q1[p];
                c2[s] = q2[p];
            }
        }
        // tiles first, then the slot atomics back to back, then the flags
        // and stores (an LDS store between the atomics would order them)
        int tl[kBatch];
        bool ex[kBatch], ok[kBatch], in[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            ok[s] = base + s * kBinThreads + threadIdx.x < end;
            const unsigned lt = (unsigned)(tile_of<PT, MT>(r[s].x, r[s].y, r[s].z, g, ex[s]) - r0);
            in[s] = ok[s] && lt < (unsigned)rn;  // this workgroup row's tile range
            tl[s] = in[s] ? (int)lt : 0;
        }
        // no branch around the atomics (a lane past the slice end or outside
        // the range adds 0)
#pragma unroll
        for (int s = 0; s < kBatch; ++s) slot[s] = atomicAdd(&cur[tl[s]], in[s] ? 1u : 0u);
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            if (ok[s]) {
                // every range's workgroup takes max|q| over the whole slice,
                // so their plain stores below agree
                const double aq = fabs((double)r[s].q);
                qm = (aq > qm || aq != aq) ? aq : qm;  // NaN sticks
                if constexpr (NCH == 3) {
                    const double a1 = fabs((double)c1[s]), a2 = fabs((double)c2[s]);
                    qm1 = (a1 > qm1 || a1 != a1) ? a1 : qm1;
                    qm2 = (a2 > qm2 || a2 != a2) ? a2 : qm2;
                }
            }
            if (in[s]) {
                if (ex[s]) ext_lds[tl[s]] = 1;  // same value from every writer
                sorted[slot[s]] = r[s];
                if constexpr (NCH == 3) {
                    q1s[slot[s]] = c1[s];
                    q2s[slot[s]] = c2[s];
                }
            }
        }
    }
    // |q| >= 0, so its bit pattern orders like its value (NaN above +inf);
    // one plain store per slice (the deposit reduces them)
    __shared__ unsigned long long wmax[NCH][kBinThreads / 64];
    const double qms[3] = {qm, qm1, qm2};
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        unsigned long long bits = (unsigned long long)__double_as_longlong(qms[c]);
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            const unsigned long long o = __shfl_xor(bits, d, 64);
            bits = o > bits ? o : bits;
        }
        if ((threadIdx.x & 63) == 0) wmax[c][threadIdx.x >> 6] = bits;
    }
    __syncthreads();
    if (threadIdx.x < NCH) {
        const int c = threadIdx.x;
        unsigned long long m = wmax[c][0];
        for (int k = 1; k < kBinThreads / 64; ++k) m = wmax[c][k] > m ? wmax[c][k] : m;
        slice_qmax[(size_t)c * kBinMaxSlices + w] = m;
    }
    for (int t = threadIdx.x; t < rn; t += kBinThreads)
        if (ext_lds[t]) atomicOr(&ext_flags[r0 + t], 1u);  // rare: particles outside the mesh
}

template <typename PT, typename MT, int NCH>
__global__ void __launch_bounds__(kBinThreads)
k_bin_scatter(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz,
              const PT *__restrict__ pq, Split sp, Geo g, const unsigned *__restrict__ cmat,
              const unsigned *__restrict__ start, Rec<PT> *__restrict__ sorted,
              unsigned *__restrict__ ext_flags, unsigned long long *__restrict__ slice_qmax,
              const PT *__restrict__ q1, const PT *__restrict__ q2, PT *__restrict__ q1s,
              PT *__restrict__ q2s)
{
    // grid (slices, tile ranges), slices in XCD order as in k_bin_count
    const int ntiles = g.tx * g.ty * g.tz;
    const int r0 = (int)blockIdx.y * kRange;
    bin_scatter_body<PT, MT, NCH>(xcd_task((int)blockIdx.x, (int)gridDim.x), px, py, pz, pq, sp,
                                  g, cmat, start, sorted,
                                  ext_flags, slice_qmax, q1, q2, q1s, q2s, r0,
                                  ntiles - r0 < kRange ? ntiles - r0 : kRange);
}

// The particle order of the binning as a permutation: perm[slot] = particle
// index, slots in tile order (the same cursors as k_bin_scatter).
template <typename PT, typename MT>
__global__ void __launch_bounds__(kBinThreads)
k_bin_perm(const PT *__restrict__ px, const PT *__restrict__ py, const PT *__restrict__ pz, Split sp,
           Geo g, const unsigned *__restrict__ cmat, const unsigned *__restrict__ start,
           int *__restrict__ perm)
{
    resolve_geo(g);
    __shared__ unsigned cur[kRange];
    // grid (slices, tile ranges), as the scatter
    const int ntiles = g.tx * g.ty * g.tz;
    const int r0 = (int)blockIdx.y * kRange;
    const int rn = ntiles - r0 < kRange ? ntiles - r0 : kRange;
    for (int t = threadIdx.x; t < rn; t += kBinThreads)
        cur[t] = start[r0 + t] + cmat[(size_t)(r0 + t) * sp.G + blockIdx.x];
    __syncthreads();
    const int64_t first = (int64_t)blockIdx.x * sp.slice;
    const int64_t end = first + sp.slice < sp.n ? first + sp.slice : sp.n;
    for (int64_t base = first; base < end; base += kBatch * kBinThreads) {
        int tl[kBatch];
        bool ok[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) {
            int64_t p = base + s * kBinThreads + threadIdx.x;
            ok[s] = p < end;
            p = ok[s] ? p : end - 1;
            const unsigned lt = (unsigned)(tile_of<PT, MT>(px[p], py[p], pz[p], g) - r0);
            ok[s] = ok[s] && lt < (unsigned)rn;
            tl[s] = ok[s] ? (int)lt : 0;
        }
        unsigned slot[kBatch];
#pragma unroll
        for (int s = 0; s < kBatch; ++s) slot[s] = atomicAdd(&cur[tl[s]], ok[s] ? 1u : 0u);
#pragma unroll
        for (int s = 0; s < kBatch; ++s)
            if (ok[s]) perm[slot[s]] = (int)(base + s * kBinThreads + threadIdx.x);
    }
}

// CIC contributions of one record, in the reference's arithmetic
// (src/deposition.jl:38-84), added to the chunk's LDS sub-grid.  Two
// accumulators, chosen per chunk:
//  * Float64 LDS atomics (FIXED = false): any mesh type, any weights;
//  * 64-bit fixed point (FIXED = true, Float32 meshes): the contribution is
//    rounded to a multiple of 1/S (S a power of two with max|q| * S < 2^49
//    and 2^chunk_log2 * max|q| * S < 2^62, so no sum can overflow) and added with an
//    integer LDS atomic, which gfx950 executes about twice as fast as the
//    Float64 one.  The sum is exact on that grid; the rounding is at most
//    2^-49 max|q| per contribution, far below Float32 resolution.  Only
//    chunks whose weights all lie in [0, 1] take it.
template <typename PT, typename MT, bool FIXED>
__device__ __forceinline__ void deposit_one(const Rec<PT> &r, const Geo &g, int ox, int oy, int oz,
                                            double *__restrict__ sub, double S)
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
    const int o = (ix - ox) + kSub * ((iy - oy) + kSub * (iz - oz));
    constexpr int sy = kSub, sz = kSub * kSub;
    const C v[8] = {((c * wx0) * wy0) * wz0, ((c * wx1) * wy0) * wz0,
                    ((c * wx0) * wy1) * wz0, ((c * wx1) * wy1) * wz0,
                    ((c * wx0) * wy0) * wz1, ((c * wx1) * wy0) * wz1,
                    ((c * wx0) * wy1) * wz1, ((c * wx1) * wy1) * wz1};
    constexpr int off[8] = {0, 1, sy, sy + 1, sz, sz + 1, sz + sy, sz + sy + 1};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if constexpr (FIXED) {
            // round-to-nearest integer of v*S (|v S| < 2^49): adding 1.5 * 2^52
            // leaves it in the low mantissa bits of the sum
            unsigned long long *u = reinterpret_cast<unsigned long long *>(sub + o + off[k]);
            const double t = (double)v[k] * S + 6755399441055744.0;
            atomicAdd(u, (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000LL));
        } else {
            atomicAdd(sub + o + off[k], (double)v[k]);
        }
    }
}

// qsrc: the chunk's charges of another channel (same order as the
// records), or null for the records' own charges.
template <typename PT, typename MT, bool FIXED>
__device__ __forceinline__ void deposit_chunk(const Rec<PT> *__restrict__ src,
                                              const PT *__restrict__ qsrc, int len, const Geo &g,
                                              int ox, int oy, int oz, double *__restrict__ sub,
                                              double S)
{
    // records are fetched kBatch at a time so their loads overlap
    constexpr int kBatch = 8;
    for (int i0 = 0; i0 < len; i0 += kBatch * kDepThreads) {
        // unconditional loads (index clamped into the chunk) so that all
        // kBatch are in flight together; len >= 1 for every chunk
        Rec<PT> rb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = i0 + u * kDepThreads + (int)threadIdx.x;
            rb[u] = src[i < len ? i : len - 1];
        }
        if (qsrc) {  // uniform over the workgroup
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int i = i0 + u * kDepThreads + (int)threadIdx.x;
                rb[u].q = qsrc[i < len ? i : len - 1];
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (i0 + u * kDepThreads + (int)threadIdx.x < len)
                deposit_one<PT, MT, FIXED>(rb[u], g, ox, oy, oz, sub, S);
        }
    }
}

// Grid (chunks, channels): channel c takes its charges from the records
// (c = 0) or from qs[c - 1] in the records' order, its max|q| from the
// slice maxima kBinMaxSlices * c in, and writes sub-grid set c (pstride
// doubles apart).
template <typename PT, typename MT>
__device__ __forceinline__ void deposit_tiles_body(
    unsigned bx, int chan, const Rec<PT> *__restrict__ sorted, const PT *__restrict__ q1s,
    const PT *__restrict__ q2s, const int4 *__restrict__ chunks, const unsigned *__restrict__ nchunks,
    Geo g, double *__restrict__ partials, size_t pstride, const unsigned *__restrict__ ext_flags,
    const unsigned long long *__restrict__ slice_qmax, int nslices, int chunk_log2)
{
    resolve_geo(g);
    __shared__ double sub[kSubVol];
    __shared__ unsigned long long wmax[kDepThreads / 64];
    if (bx >= *nchunks) return;
    slice_qmax += (size_t)chan * kBinMaxSlices;
    partials += chan * pstride;
    const int4 ch = chunks[bx];
    const int t = ch.x;
    const int ox = (t % g.tx) * kTile, oy = ((t / g.tx) % g.ty) * kTile, oz = (t / (g.tx * g.ty)) * kTile;
    // fixed point for Float32 meshes when every weight of the tile is in
    // [0, 1] and max|q| is finite; the branch is uniform over the workgroup
    bool fixed = false;
    double S = 1.0, invS = 1.0;
    if constexpr (sizeof(MT) == 4) {
        // max|q| over all particles: the maximum of the per-slice maxima
        unsigned long long m = 0ull;
        for (int i = threadIdx.x; i < nslices; i += blockDim.x) m = slice_qmax[i] > m ? slice_qmax[i] : m;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            const unsigned long long o = __shfl_xor(m, d, 64);
            m = o > m ? o : m;
        }
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
        __syncthreads();
        m = wmax[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = wmax[k] > m ? wmax[k] : m;
        const double qm = __longlong_as_double((long long)m);
        // finite max|q| not so small that the scale 2^(48 - ilogb) overflows
        if (ext_flags[t] == 0u && qm <= 1.7976931348623157e308 && (qm == 0.0 || qm >= 1e-270)) {
            fixed = true;
            if (qm > 0.0) {
                // |q 2^e| < 2^49 (the rounding below needs < 2^51) and
                // 2^chunk_log2 * qm * 2^e < 2^62 (no sum overflows)
                const int e = (chunk_log2 <= 13 ? 48 : 61 - chunk_log2) - ilogb(qm);
                S = ldexp(1.0, e);
                invS = ldexp(1.0, -e);
            }
        }
    }
    for (int i = threadIdx.x; i < kSubVol; i += blockDim.x) sub[i] = 0.0;  // also integer zero
    __syncthreads();
    const Rec<PT> *src = sorted + ch.y;
    const PT *qsrc = chan == 0 ? nullptr : (chan == 1 ? q1s : q2s) + ch.y;
    if (fixed) deposit_chunk<PT, MT, true>(src, qsrc, ch.z, g, ox, oy, oz, sub, S);
    else deposit_chunk<PT, MT, false>(src, qsrc, ch.z, g, ox, oy, oz, sub, S);
    __syncthreads();
    double *out = partials + (size_t)ch.w * kSubVol;
    if (fixed) {
        const long long *si = reinterpret_cast<const long long *>(sub);
        for (int i = threadIdx.x; i < kSubVol; i += blockDim.x) out[i] = (double)si[i] * invS;
    } else {
        for (int i = threadIdx.x; i < kSubVol; i += blockDim.x) out[i] = sub[i];
    }
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(kDepThreads)
k_deposit_tiles(const Rec<PT> *__restrict__ sorted, const PT *__restrict__ q1s,
                const PT *__restrict__ q2s, const int4 *__restrict__ chunks,
                const unsigned *__restrict__ nchunks, Geo g, double *__restrict__ partials,
                size_t pstride, const unsigned *__restrict__ ext_flags,
                const unsigned long long *__restrict__ slice_qmax, int nslices, int chunk_log2)
{
    deposit_tiles_body<PT, MT>(blockIdx.x, blockIdx.y, sorted, q1s, q2s, chunks, nchunks, g,
                               partials, pstride, ext_flags, slice_qmax, nslices, chunk_log2);
}

// Completes the per-tile sums in sub-grid slots [0, ntiles): adds a tile's
// further chunks (slots extra_base[t], ...) to slot t.  An empty tile's
// slot is zeroed, unless skip_empty (meshes of more than one binning range,
// where most tiles are empty): the gather then skips empty tiles.  Grid (ceil(kSubVol / 256), tiles, channels), one lane per
// node; channel c's sub-grids start c * pstride doubles in.
__device__ __forceinline__ void tile_reduce_body(unsigned bx, int t, int chan,
                                                 double *__restrict__ partials, size_t pstride,
                                                 const unsigned *__restrict__ counts,
                                                 const unsigned *__restrict__ extra_base,
                                                 int chunk_log2, int skip_empty)
{
    partials += chan * pstride;
    const unsigned nk = (counts[t] + (1u << chunk_log2) - 1) >> chunk_log2;
    if (nk == 1 || (nk == 0 && skip_empty)) return;  // slot t holds the sum, or stays unset
    const int local = (int)bx * blockDim.x + threadIdx.x;
    if (local >= kSubVol) return;
    double *dst = partials + (size_t)t * kSubVol + local;
    if (nk == 0) {
        *dst = 0.0;
        return;
    }
    const double *p = partials + (size_t)extra_base[t] * kSubVol + local;
    // eight independent partial sums keep eight loads in flight (the
    // densest tiles hold a few dozen chunks)
    double acc[8] = {*dst, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const unsigned nx = nk - 1;  // further chunks
    unsigned k = 0;
    for (; k + 8 <= nx; k += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += p[(size_t)(k + u) * kSubVol];
    }
#pragma unroll
    for (int u = 0; u < 7; ++u)
        if (k + u < nx) acc[u] += p[(size_t)(k + u) * kSubVol];
    *dst = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

__global__ void __launch_bounds__(256)
k_tile_reduce(double *__restrict__ partials, size_t pstride, const unsigned *__restrict__ counts,
              const unsigned *__restrict__ extra_base, int chunk_log2, int skip_empty)
{
    __builtin_amdgcn_s_setprio(kChainPrio);
    tile_reduce_body(blockIdx.x, blockIdx.y, blockIdx.z, partials, pstride, counts, extra_base,
                     chunk_log2, skip_empty);
}

// Node-parallel sum over the (up to 8) tiles whose sub-grid holds the node:
// a node lies in two tiles along an axis when it is on a tile face.  Tile
// t's sum is in sub-grid slot t (k_tile_reduce), so every address follows
// from the node index alone.  Plain stores, fixed summation order; with
// accumulate the sum is added to the existing rho (deposit!(...; clear=false)).
// Grid (ceil(M / 256), channels): channel c reads sub-grid set c (pstride
// doubles apart) and writes rho_c.
// With skip_empty, empty tiles' slots are unset (tile_reduce_body) and are
// not added.
template <typename MT>
__device__ __forceinline__ void gather_tiles_body(unsigned bx, const double *__restrict__ partials,
                                                  const unsigned *__restrict__ counts, const Geo &g,
                                                  MT *__restrict__ rho, int accumulate,
                                                  int skip_empty)
{
    // at most kMaxTiles tiles of 16^3 cells: node indices fit in 32 bits
    const int M = g.nx * g.ny * g.nz;
    const int i = (int)(bx * blockDim.x + threadIdx.x);
    if (i >= M) return;
    const int gx = i % g.nx;
    const int gy = (i / g.nx) % g.ny;
    const int gz = i / (g.nx * g.ny);
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
                const int local = (gx - kTile * cx[a]) + kSub * ((gy - kTile * cy[b]) + kSub * (gz - kTile * cz[c]));
                const double v = partials[(size_t)t * kSubVol + local];
                if (skip_empty) {
                    // read beside the count (no dependent load), not added when empty
                    s += counts[t] != 0u ? v : 0.0;
                } else {
                    s += v;
                }
            }
    rho[i] = accumulate ? (MT)((double)rho[i] + s) : (MT)s;
}

template <typename MT>
__global__ void __launch_bounds__(256)
k_gather_tiles(const double *__restrict__ partials, size_t pstride,
               const unsigned *__restrict__ counts, Geo g, MT *__restrict__ rho0,
               MT *__restrict__ rho1, MT *__restrict__ rho2, int accumulate, int skip_empty)
{
    __builtin_amdgcn_s_setprio(kChainPrio);
    const int chan = blockIdx.y;
    gather_tiles_body<MT>(blockIdx.x, partials + chan * pstride, counts, g,
                          chan == 0 ? rho0 : (chan == 1 ? rho1 : rho2), accumulate, skip_empty);
}

// At most kBinMaxSlices slices of about 2^slice_log2 particles, and at
// most kCmatBudget / ntiles (the count matrix has a row per tile).
Split split_of(int64_t n, int slice_log2, int ntiles)
{
    Split sp;
    sp.n = n;
    const int64_t per = (int64_t)1 << slice_log2;
    int64_t G = (n + per - 1) / per;
    if (G > kBinMaxSlices) G = kBinMaxSlices;
    const int64_t gmax = kCmatBudget / (ntiles > 0 ? ntiles : 1);
    if (G > gmax) G = gmax;
    if (G < 1) G = 1;
    sp.G = (int)G;
    sp.slice = (n + G - 1) / G;
    return sp;
}

// Entries of each per-tile block of the counts array (totals, tile start,
// first extra slot, extrapolation flags): kRange for meshes of at most one
// binning range (the layout the grouped launches assume), else the tiles.
int counts_stride(int ntiles) { return ntiles > kRange ? ntiles : kRange; }

int ensure_bins(Mesh *m, int64_t n, int pdt, int ntiles, int nchan)
{
    const int64_t chunk = (int64_t)1 << m->chunk_log2;
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
    // chunks: at most ntiles + n / chunk; sub-grid slots: below 2 ntiles + chunks
    // (slots of the row scan's reservation order, k_bin_rowscan)
    const int64_t maxchunks = 2 * (int64_t)ntiles + n / chunk + 1;
    if (B.chunk_cap < maxchunks || B.partial_sets < nchan) {
        const int64_t cap = B.chunk_cap > maxchunks ? B.chunk_cap : maxchunks;
        const int sets = B.partial_sets > nchan ? B.partial_sets : nchan;
        if (B.chunks) (void)hipFree(B.chunks);
        if (B.partials) (void)hipFree(B.partials);
        B.chunks = B.partials = nullptr;
        B.chunk_cap = 0;
        B.partial_sets = 0;
        if (hipMalloc(&B.chunks, sizeof(int4) * cap) != hipSuccess) return fail();
        // one sub-grid per chunk and channel, in the wider mesh type (double
        // covers both); channel c's set starts c * cap sub-grids in
        if (hipMalloc(&B.partials, sizeof(double) * kSubVol * (size_t)cap * sets) != hipSuccess)
            return fail();
        B.chunk_cap = cap;
        B.partial_sets = sets;
    }
    if (!B.counts) {
        // totals | tile start | first extra sub-grid slot | extrapolation flags | nchunks, pad |
        // per-slice max|q| of up to three channels (u64 bit patterns)
        const size_t bytes =
            sizeof(unsigned) * (4 * (size_t)counts_stride(ntiles) + 4 + 3 * 2 * kBinMaxSlices);
        if (hipMalloc((void **)&B.counts, bytes) != hipSuccess) return fail();
    }
    const int64_t need_cmat = (int64_t)ntiles * split_of(n, m->slice_log2, ntiles).G;
    if (B.cmat_cap < need_cmat) {
        if (B.cmat) (void)hipFree(B.cmat);
        B.cmat = nullptr;
        B.cmat_cap = 0;
        if (hipMalloc((void **)&B.cmat, sizeof(unsigned) * (size_t)need_cmat) != hipSuccess)
            return fail();
        B.cmat_cap = need_cmat;
    }
    return SC_OK;
}

template <typename PT, typename MT>
int binned(Mesh *m, const void *x, const void *y, const void *z, const void *const *q, int nchan,
           int64_t n, const Geo &g, int clear, void *const *out)
{
    Ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    DepositBins &B = m->bins;
    const int ntiles = g.tx * g.ty * g.tz;
    const int cs = counts_stride(ntiles);
    const unsigned ranges = (unsigned)((ntiles + kRange - 1) / kRange);
    unsigned *counts = B.counts, *cursor = B.counts + cs;
    unsigned *cbase = B.counts + 2 * cs, *ext = B.counts + 3 * cs;
    unsigned *nch = B.counts + 4 * cs;
    auto *slice_qmax = reinterpret_cast<unsigned long long *>(B.counts + 4 * cs + 4);
    int4 *chunks = static_cast<int4 *>(B.chunks);
    const Split sp = split_of(n, m->slice_log2, ntiles);
    // one binning range: the row scan reserves the tiles' runs and chunks
    // (records and chunks counted in the two halves of one 64-bit word after
    // nchunks; the chunk count is its high half) and the tile scan is not
    // launched
    const bool by_rowscan = ranges == 1 && n < ((int64_t)1 << 31);
    auto *resv = reinterpret_cast<unsigned long long *>(B.counts + 4 * cs + 2);
    if (by_rowscan) nch = B.counts + 4 * cs + 3;
    launch_begin(ctx, K_BIN_COUNT);
    hipLaunchKernelGGL((k_bin_count<PT, MT>), dim3(sp.G, ranges), dim3(kBinThreads), 0, st,
                       (const PT *)x, (const PT *)y, (const PT *)z, sp, g, B.cmat,
                       by_rowscan ? resv : nullptr);
    launch_end(ctx, K_BIN_COUNT);
    launch_begin(ctx, K_BIN_ROWSCAN);
    const RowscanChunks rc = by_rowscan
                                 ? RowscanChunks{resv, cursor, cbase, ext, chunks, m->chunk_log2}
                                 : RowscanChunks{};
    hipLaunchKernelGGL(k_bin_rowscan, dim3((unsigned)div_up4(ntiles)), dim3(256), 0, st, B.cmat,
                       sp.G, ntiles, counts, rc);
    launch_end(ctx, K_BIN_ROWSCAN);
    if (!by_rowscan) {
        launch_begin(ctx, K_BIN_SCAN);
        hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(256), 0, st, (const unsigned *)counts, cursor,
                           cbase, chunks, nch, ext, ntiles, m->chunk_log2);
        launch_end(ctx, K_BIN_SCAN);
    }
    Rec<PT> *sorted = static_cast<Rec<PT> *>(B.sorted);
    PT *q1s = static_cast<PT *>(B.qs), *q2s = q1s ? q1s + n : nullptr;
    launch_begin(ctx, K_BIN_SCATTER);
    if (nchan == 3)
        hipLaunchKernelGGL((k_bin_scatter<PT, MT, 3>), dim3(sp.G, ranges), dim3(kBinThreads), 0, st,
                           (const PT *)x, (const PT *)y, (const PT *)z, (const PT *)q[0], sp, g,
                           (const unsigned *)B.cmat, (const unsigned *)cursor, sorted, ext,
                           slice_qmax, (const PT *)q[1], (const PT *)q[2], q1s, q2s);
    else
        hipLaunchKernelGGL((k_bin_scatter<PT, MT, 1>), dim3(sp.G, ranges), dim3(kBinThreads), 0, st,
                           (const PT *)x, (const PT *)y, (const PT *)z, (const PT *)q[0], sp, g,
                           (const unsigned *)B.cmat, (const unsigned *)cursor, sorted, ext,
                           slice_qmax, (const PT *)nullptr, (const PT *)nullptr, (PT *)nullptr,
                           (PT *)nullptr);
    launch_end(ctx, K_BIN_SCATTER);
    // every channel in each launch (grid axis), on the same chunks
    const unsigned maxchunks = (unsigned)(ntiles + (n >> m->chunk_log2) + 1);
    const int64_t M = (int64_t)g.nx * g.ny * g.nz;
    const size_t pstride = (size_t)kSubVol * (size_t)B.chunk_cap;
    launch_begin(ctx, K_DEPOSIT_TILES);
    hipLaunchKernelGGL((k_deposit_tiles<PT, MT>), dim3(maxchunks, nchan), dim3(kDepThreads), 0, st,
                       (const Rec<PT> *)sorted, (const PT *)q1s, (const PT *)q2s,
                       (const int4 *)chunks, (const unsigned *)nch, g,
                       static_cast<double *>(B.partials), pstride, (const unsigned *)ext,
                       (const unsigned long long *)slice_qmax, sp.G, m->chunk_log2);
    launch_end(ctx, K_DEPOSIT_TILES);
    launch_begin(ctx, K_TILE_REDUCE);
    hipLaunchKernelGGL(k_tile_reduce, dim3((unsigned)((kSubVol + 255) / 256), (unsigned)ntiles, nchan),
                       dim3(256), 0, st, static_cast<double *>(B.partials), pstride,
                       (const unsigned *)counts, (const unsigned *)cbase, m->chunk_log2,
                       ranges > 1 ? 1 : 0);
    launch_end(ctx, K_TILE_REDUCE);
    launch_begin(ctx, K_GATHER_TILES);
    MT *o0 = static_cast<MT *>(out[0]);
    MT *o1 = nchan > 1 ? static_cast<MT *>(out[1]) : o0;
    MT *o2 = nchan > 2 ? static_cast<MT *>(out[2]) : o0;
    hipLaunchKernelGGL((k_gather_tiles<MT>), dim3((unsigned)((M + 255) / 256), nchan), dim3(256), 0,
                       st, (const double *)B.partials, pstride, (const unsigned *)counts, g, o0, o1,
                       o2, clear ? 0 : 1, ranges > 1 ? 1 : 0);
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
    if (B.cmat) (void)hipFree(B.cmat);
    if (B.qs) (void)hipFree(B.qs);
    B = DepositBins();
}

namespace {

template <typename PT, typename MT>
int perm_binned(Mesh *m, const void *x, const void *y, const void *z, int64_t n, const Geo &g,
                int *perm)
{
    Ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    DepositBins &B = m->bins;
    const int ntiles = g.tx * g.ty * g.tz;
    const int cs = counts_stride(ntiles);
    const unsigned ranges = (unsigned)((ntiles + kRange - 1) / kRange);
    unsigned *counts = B.counts, *cursor = B.counts + cs;
    unsigned *cbase = B.counts + 2 * cs, *ext = B.counts + 3 * cs;
    unsigned *nch = B.counts + 4 * cs;
    int4 *chunks = static_cast<int4 *>(B.chunks);
    const Split sp = split_of(n, m->slice_log2, ntiles);
    launch_begin(ctx, K_BIN_COUNT);
    hipLaunchKernelGGL((k_bin_count<PT, MT>), dim3(sp.G, ranges), dim3(kBinThreads), 0, st,
                       (const PT *)x, (const PT *)y, (const PT *)z, sp, g, B.cmat,
                       (unsigned long long *)nullptr);
    launch_end(ctx, K_BIN_COUNT);
    launch_begin(ctx, K_BIN_ROWSCAN);
    hipLaunchKernelGGL(k_bin_rowscan, dim3((unsigned)div_up4(ntiles)), dim3(256), 0, st, B.cmat,
                       sp.G, ntiles, counts, RowscanChunks{});
    launch_end(ctx, K_BIN_ROWSCAN);
    launch_begin(ctx, K_BIN_SCAN);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(256), 0, st, (const unsigned *)counts, cursor,
                       cbase, chunks, nch, ext, ntiles, m->chunk_log2);
    launch_end(ctx, K_BIN_SCAN);
    launch_begin(ctx, K_BIN_SCATTER);
    hipLaunchKernelGGL((k_bin_perm<PT, MT>), dim3(sp.G, ranges), dim3(kBinThreads), 0, st, (const PT *)x,
                       (const PT *)y, (const PT *)z, sp, g, (const unsigned *)B.cmat,
                       (const unsigned *)cursor, perm);
    launch_end(ctx, K_BIN_SCATTER);
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

}  // namespace

int launch_sort_perm(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n,
                     const double lo[3], const double dl[3], int *perm)
{
    Geo g;
    g.nx = (int)m->n[0];
    g.ny = (int)m->n[1];
    g.nz = (int)m->n[2];
    g.tx = (g.nx - 1 + kTile - 1) / kTile;
    g.ty = (g.ny - 1 + kTile - 1) / kTile;
    g.tz = (g.nz - 1 + kTile - 1) / kTile;
    if ((int64_t)g.tx * g.ty * g.tz > kMaxTiles)
        return set_err(m->ctx, SC_ERR_UNSUPPORTED, "the mesh has too many tiles to sort by");
    if (n > 0x7fffffffLL) return set_err(m->ctx, SC_ERR_INVALID, "too many particles to sort");
    g.dev = lo ? nullptr : m->geom;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo ? lo[a] : 0.0;
        g.dl[a] = lo ? dl[a] : 1.0;
    }
    if (int rc = ensure_bins(m, n, pdt, g.tx * g.ty * g.tz, 1)) return rc;
    if (pdt == SC_F32 && m->dtype == SC_F32) return perm_binned<float, float>(m, x, y, z, n, g, perm);
    if (pdt == SC_F64 && m->dtype == SC_F64)
        return perm_binned<double, double>(m, x, y, z, n, g, perm);
    if (pdt == SC_F64 && m->dtype == SC_F32) return perm_binned<double, float>(m, x, y, z, n, g, perm);
    return perm_binned<float, double>(m, x, y, z, n, g, perm);
}

int launch_deposit_binned_multi(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                                const void *const *q, int nchan, int64_t n, const double lo[3],
                                const double dl[3], int clear, void *const *out)
{
    if (nchan != 1 && nchan != 3) return SC_ERR_UNSUPPORTED;
    Geo g;
    g.nx = (int)m->n[0];
    g.ny = (int)m->n[1];
    g.nz = (int)m->n[2];
    g.tx = (g.nx - 1 + kTile - 1) / kTile;
    g.ty = (g.ny - 1 + kTile - 1) / kTile;
    g.tz = (g.nz - 1 + kTile - 1) / kTile;
    if ((int64_t)g.tx * g.ty * g.tz > kMaxTiles) return SC_ERR_UNSUPPORTED;
    if (n > 0x7fffffffLL) return SC_ERR_UNSUPPORTED;
    // without host bounds the kernels read the mesh's device-resident ones
    g.dev = lo ? nullptr : m->geom;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo ? lo[a] : 0.0;
        g.dl[a] = lo ? dl[a] : 1.0;
    }
    if (int rc = ensure_bins(m, n, pdt, g.tx * g.ty * g.tz, nchan)) return rc;
    if (nchan > 1) {
        DepositBins &B = m->bins;
        const size_t need = (size_t)2 * n * pdt;
        if (B.qs_bytes < need) {
            if (B.qs) (void)hipFree(B.qs);
            B.qs = nullptr;
            B.qs_bytes = 0;
            if (hipMalloc(&B.qs, need) != hipSuccess)
                return set_err(m->ctx, SC_ERR_NOMEM, "device allocation of the deposit bins failed");
            B.qs_bytes = need;
        }
    }
    if (pdt == SC_F32 && m->dtype == SC_F32)
        return binned<float, float>(m, x, y, z, q, nchan, n, g, clear, out);
    if (pdt == SC_F64 && m->dtype == SC_F64)
        return binned<double, double>(m, x, y, z, q, nchan, n, g, clear, out);
    if (pdt == SC_F64 && m->dtype == SC_F32)
        return binned<double, float>(m, x, y, z, q, nchan, n, g, clear, out);
    return binned<float, double>(m, x, y, z, q, nchan, n, g, clear, out);
}

// Binned deposit into `out` (nx*ny*nz of the mesh type); returns
// SC_ERR_UNSUPPORTED (without launching anything) when the grid has more
// tiles than the LDS histogram holds.  Writes every node of out
// (clear != 0) or adds to it (clear == 0).
int launch_deposit_binned(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                          const void *q, int64_t n, const double lo[3], const double dl[3],
                          int clear, void *out)
{
    return launch_deposit_binned_multi(m, pdt, x, y, z, &q, 1, n, lo, dl, clear, &out);
}


// ---- grouped launches: the batched deposit with one launch per phase ----
//
// Each phase of the binned deposit is one launch over the workgroups of all
// bunches: workgroup `blk` of a phase belongs to the last bunch b with
// start[phase][b] <= blk, and runs that bunch's body with local index
// blk - start[phase][b].  The per-bunch descriptors and the block starts
// live in a table in device memory written by small store launches whose
// arguments carry the descriptors by value, so a captured graph rewrites
// the table on every replay.  All launches are on one stream: a later
// grouped call's store runs after this call's phases have read the table.

namespace {

enum { kGCount, kGRowscan, kGScan, kGScatter, kGDeposit, kGReduce, kGGather, kGPhases };
constexpr int kReduceBlocks = (kSubVol + 255) / 256;  // k_tile_reduce workgroups per tile

struct GBunch {
    const void *x, *y, *z, *q;
    Split sp;
    Geo g;
    unsigned *counts;  // the bins' counts block (totals | cursor | chunk base | flags | nchunks | qmax)
    unsigned *cmat;
    int4 *chunks;
    void *sorted;
    double *partials;
    void *rho;
    int chunk_log2;
};

struct GTable {
    unsigned start[kGPhases][kGroupMax + 1];
    GBunch b[kGroupMax];
    int nb, accumulate;
};

// by-value argument of one store launch: descriptors first .. first+count-1
// and their block starts (entry count = the next bunch's start)
struct GStore {
    int first, count, nb, accumulate;
    unsigned start[kGPhases][kGroupStore + 1];
    GBunch b[kGroupStore];
};
static_assert(sizeof(GStore) <= 2048, "store-launch arguments stay small");

// one lane per 8-byte word (vector stores)
__global__ void __launch_bounds__(256) k_group_store(GStore s, GTable *__restrict__ T)
{
    constexpr int kWords = (int)(sizeof(GBunch) / 8);
    static_assert(sizeof(GBunch) % 8 == 0, "descriptor of whole words");
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(s.b);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(T->b + s.first);
    for (int i = threadIdx.x; i < s.count * kWords; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kGPhases * (s.count + 1); i += blockDim.x) {
        const int ph = i / (s.count + 1), k = i - ph * (s.count + 1);
        T->start[ph][s.first + k] = s.start[ph][k];
    }
    if (threadIdx.x == 0) {
        T->nb = s.nb;
        T->accumulate = s.accumulate;
    }
}

// the bunch owning workgroup blk of a phase: the last b with start[b] <= blk
__device__ __forceinline__ int group_of(const unsigned *__restrict__ start, int nb, unsigned blk)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#define SC_GROUP_PROLOGUE(PH)                                    \
    const int b_ = group_of(T->start[PH], T->nb, blockIdx.x);    \
    const GBunch &d = T->b[b_];                                  \
    const unsigned lb = blockIdx.x - T->start[PH][b_];           \
    unsigned *cnt = d.counts;

template <typename PT, typename MT>
__global__ void __launch_bounds__(kBinThreads) k_bin_count_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGCount)
    (void)cnt;
    bin_count_body<PT, MT>(lb, (const PT *)d.x, (const PT *)d.y, (const PT *)d.z, d.sp, d.g, d.cmat,
                           0, d.g.tx * d.g.ty * d.g.tz);
}

__global__ void __launch_bounds__(256) k_bin_rowscan_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGRowscan)
    bin_rowscan_body(lb, d.cmat, d.sp.G, d.g.tx * d.g.ty * d.g.tz, cnt);
}

__global__ void __launch_bounds__(256) k_bin_scan_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGScan)
    (void)lb;
    bin_scan_body(cnt, cnt + kRange, cnt + 2 * kRange, d.chunks, cnt + 4 * kRange,
                  cnt + 3 * kRange, d.g.tx * d.g.ty * d.g.tz, d.chunk_log2);
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(kBinThreads) k_bin_scatter_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGScatter)
    bin_scatter_body<PT, MT, 1>(lb, (const PT *)d.x, (const PT *)d.y, (const PT *)d.z,
                                (const PT *)d.q, d.sp, d.g, d.cmat, cnt + kRange,
                                static_cast<Rec<PT> *>(d.sorted), cnt + 3 * kRange,
                                reinterpret_cast<unsigned long long *>(cnt + 4 * kRange + 4),
                                nullptr, nullptr, nullptr, nullptr, 0, d.g.tx * d.g.ty * d.g.tz);
}

template <typename PT, typename MT>
__global__ void __launch_bounds__(kDepThreads) k_deposit_tiles_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGDeposit)
    deposit_tiles_body<PT, MT>(lb, 0, static_cast<const Rec<PT> *>(d.sorted), nullptr, nullptr,
                               d.chunks, cnt + 4 * kRange, d.g, d.partials, 0,
                               cnt + 3 * kRange,
                               reinterpret_cast<const unsigned long long *>(cnt + 4 * kRange + 4),
                               d.sp.G, d.chunk_log2);
}

__global__ void __launch_bounds__(256) k_tile_reduce_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGReduce)
    tile_reduce_body(lb % kReduceBlocks, (int)(lb / kReduceBlocks), 0, d.partials, 0, cnt,
                     cnt + 2 * kRange, d.chunk_log2, 0);
}

template <typename MT>
__global__ void __launch_bounds__(256) k_gather_tiles_g(const GTable *__restrict__ T)
{
    SC_GROUP_PROLOGUE(kGGather)
    gather_tiles_body<MT>(lb, d.partials, cnt, d.g, static_cast<MT *>(d.rho), T->accumulate, 0);
}

#undef SC_GROUP_PROLOGUE

template <typename PT, typename MT>
void launch_groups(Ctx *ctx, GTable *T, const unsigned *total)
{
    hipStream_t st = ctx->stream;
    launch_begin(ctx, K_BIN_COUNT);
    hipLaunchKernelGGL((k_bin_count_g<PT, MT>), dim3(total[kGCount]), dim3(kBinThreads), 0, st,
                       (const GTable *)T);
    launch_end(ctx, K_BIN_COUNT);
    launch_begin(ctx, K_BIN_ROWSCAN);
    hipLaunchKernelGGL(k_bin_rowscan_g, dim3(total[kGRowscan]), dim3(256), 0, st, (const GTable *)T);
    launch_end(ctx, K_BIN_ROWSCAN);
    launch_begin(ctx, K_BIN_SCAN);
    hipLaunchKernelGGL(k_bin_scan_g, dim3(total[kGScan]), dim3(256), 0, st, (const GTable *)T);
    launch_end(ctx, K_BIN_SCAN);
    launch_begin(ctx, K_BIN_SCATTER);
    hipLaunchKernelGGL((k_bin_scatter_g<PT, MT>), dim3(total[kGScatter]), dim3(kBinThreads), 0, st,
                       (const GTable *)T);
    launch_end(ctx, K_BIN_SCATTER);
    launch_begin(ctx, K_DEPOSIT_TILES);
    hipLaunchKernelGGL((k_deposit_tiles_g<PT, MT>), dim3(total[kGDeposit]), dim3(kDepThreads), 0, st,
                       (const GTable *)T);
    launch_end(ctx, K_DEPOSIT_TILES);
    launch_begin(ctx, K_TILE_REDUCE);
    hipLaunchKernelGGL(k_tile_reduce_g, dim3(total[kGReduce]), dim3(256), 0, st, (const GTable *)T);
    launch_end(ctx, K_TILE_REDUCE);
    launch_begin(ctx, K_GATHER_TILES);
    hipLaunchKernelGGL((k_gather_tiles_g<MT>), dim3(total[kGGather]), dim3(256), 0, st,
                       (const GTable *)T);
    launch_end(ctx, K_GATHER_TILES);
}

}  // namespace

int ensure_group_table(Ctx *ctx, size_t bytes)
{
    if (ctx->gtable) return SC_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SC_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return SC_ERR_UNSUPPORTED;  // no allocation under capture
    if (hipMalloc(&ctx->gtable, bytes) != hipSuccess) {
        ctx->gtable = nullptr;
        return set_err(ctx, SC_ERR_NOMEM, "device allocation of the batch descriptor table failed");
    }
    return SC_OK;
}

size_t group_table_bytes();

// Qualifies: at most kGroupMax bunches, one mesh type, every bunch with
// particles (n in [1, 2^31)) and at most one binning range (kRange tiles).
int launch_deposit_grouped(Ctx *ctx, Mesh *const *ms, int nb, int pdt, const void *const *x,
                           const void *const *y, const void *const *z, const void *const *q,
                           const int64_t *n, const double *lo, const double *dl, int clear)
{
    if (!(ctx->grouped & SC_GROUP_PARTICLES) || nb < 1 || nb > kGroupMax) return SC_ERR_UNSUPPORTED;
    const int mdt = ms[0]->dtype;
    std::vector<Geo> geos(nb);
    for (int b = 0; b < nb; ++b) {
        const Mesh *m = ms[b];
        if (m->dtype != mdt || n[b] < 1 || n[b] > 0x7fffffffLL) return SC_ERR_UNSUPPORTED;
        Geo &g = geos[b];
        g.nx = (int)m->n[0];
        g.ny = (int)m->n[1];
        g.nz = (int)m->n[2];
        g.tx = (g.nx - 1 + kTile - 1) / kTile;
        g.ty = (g.ny - 1 + kTile - 1) / kTile;
        g.tz = (g.nz - 1 + kTile - 1) / kTile;
        if ((int64_t)g.tx * g.ty * g.tz > kRange) return SC_ERR_UNSUPPORTED;
        g.dev = nullptr;
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = lo[3 * b + a];
            g.dl[a] = dl[3 * b + a];
        }
    }
    if (int rc = ensure_group_table(ctx, group_table_bytes())) return rc;
    for (int b = 0; b < nb; ++b)
        if (int rc = ensure_bins(ms[b], n[b], pdt, geos[b].tx * geos[b].ty * geos[b].tz, 1)) return rc;
    // where each mesh's deposit starts on the stream (a later solve's Green
    // build on the side stream waits for it, as after launch_deposit)
    for (int b = 0; b < nb; ++b)
        if (int rc = mark_deposit_start(ms[b])) return rc;
    GTable *T = static_cast<GTable *>(ctx->gtable);
    // descriptors and block starts, stored kGroupStore at a time
    unsigned total[kGPhases] = {};
    GStore s;
    std::memset(&s, 0, sizeof(s));
    s.nb = nb;
    s.accumulate = clear ? 0 : 1;
    for (int b0 = 0; b0 < nb; b0 += kGroupStore) {
        s.first = b0;
        s.count = nb - b0 < kGroupStore ? nb - b0 : kGroupStore;
        for (int k = 0; k < s.count; ++k) {
            Mesh *m = ms[b0 + k];
            DepositBins &B = m->bins;
            const Geo &g = geos[b0 + k];
            const int ntiles = g.tx * g.ty * g.tz;
            const int64_t M = (int64_t)g.nx * g.ny * g.nz;
            GBunch &d = s.b[k];
            d.x = x[b0 + k];
            d.y = y[b0 + k];
            d.z = z[b0 + k];
            d.q = q[b0 + k];
            d.sp = split_of(n[b0 + k], m->slice_log2, ntiles);
            d.g = g;
            d.counts = B.counts;
            d.cmat = B.cmat;
            d.chunks = static_cast<int4 *>(B.chunks);
            d.sorted = B.sorted;
            d.partials = static_cast<double *>(B.partials);
            d.rho = m->rho;
            d.chunk_log2 = m->chunk_log2;
            const unsigned blocks[kGPhases] = {
                (unsigned)d.sp.G, (unsigned)div_up4(ntiles), 1u, (unsigned)d.sp.G,
                (unsigned)(ntiles + (n[b0 + k] >> m->chunk_log2) + 1),
                (unsigned)(kReduceBlocks * ntiles), (unsigned)((M + 255) / 256)};
            for (int ph = 0; ph < kGPhases; ++ph) {
                s.start[ph][k] = total[ph];
                total[ph] += blocks[ph];
            }
        }
        for (int ph = 0; ph < kGPhases; ++ph) s.start[ph][s.count] = total[ph];
        hipLaunchKernelGGL(k_group_store, dim3(1), dim3(256), 0, ctx->stream, s, T);
    }
    if (mdt == SC_F32 && pdt == SC_F32) launch_groups<float, float>(ctx, T, total);
    else if (mdt == SC_F64 && pdt == SC_F64) launch_groups<double, double>(ctx, T, total);
    else if (mdt == SC_F32) launch_groups<double, float>(ctx, T, total);
    else launch_groups<float, double>(ctx, T, total);
    SC_HIP(ctx, hipGetLastError());
    return SC_OK;
}

size_t group_table_bytes() { return sizeof(GTable); }

}  // namespace sc

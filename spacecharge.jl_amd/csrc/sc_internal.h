// sc_internal.h -- shared declarations of the gfx950 space-charge library.
//
// Host-side objects (context, mesh, workspace), error plumbing, the kernel
// registry used by the per-kernel timers, and the launch helpers that every
// .hip translation unit uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "spacecharge_mi355.h"

namespace sc {

// ---------------------------------------------------------------------------
// Kernel registry (indices are stable; names are what rocprofv3 shows).
enum KernelId : int {
    K_CLEAR = 0,
    K_DEPOSIT,
    K_INTERP,
    K_BOUNDS_PARTIAL,
    K_BOUNDS_FINAL,
    K_GREEN_X,
    K_GREEN_Y,
    K_GREEN_Z,
    K_RHO_FWD_X,
    K_RHO_FWD_Y,
    K_ZCONV,
    K_INV_Y,
    K_INV_X,
    K_BIN_COUNT,
    K_BIN_SCAN,
    K_BIN_SCATTER,
    K_DEPOSIT_TILES,
    K_INTERLEAVE,
    K_GATHER_TILES,
    K_GREEN_F,
    K_BIN_ROWSCAN,
    K_TILE_REDUCE,
    K_DEPOSIT_VJP,
    K_INTERP_VJP,
    K_COUNT
};
extern const char *const kKernelNames[K_COUNT];

// ---------------------------------------------------------------------------
// Mixed-radix plan for one transform length.  Radices are 2, 3, 4, 5, 7, 8
// (3, 5, 7 use the butterfly from the DFT definition).  Twiddles live in device memory,
// tw[m] = exp(-2 pi i m / n) in the transform's precision.
constexpr int kMaxStages = 24;
struct FftPlan {
    int n = 0;
    int nst = 0;
    int rad[kMaxStages] = {};
};

// Complex elements each lane holds in registers during one transform stage:
// a batch of nb pencils of length n in precision T needs
// nb*n <= fft_elems<T>() * kFftThreads.
constexpr int kFftThreads = 256;
template <typename T>
constexpr int fft_elems()
{
    return sizeof(T) == 4 ? 8 : 4;
}
template <typename T>
constexpr int fft_cap()
{
    return fft_elems<T>() * kFftThreads;
}
// Pencils longer than fft_cap (axes of more than fft_cap / 2 nodes) take
// twice the values per lane, one pencil per workgroup, and LDS buffers of
// up to 3 pencils (the z convolution: twiddles, spectrum, product), which
// the device's 160 KiB per workgroup holds up to fft_cap_long.
template <typename T>
constexpr int fft_elems_long()
{
    return 2 * fft_elems<T>();
}
template <typename T>
constexpr int fft_cap_long()
{
    return fft_elems_long<T>() * kFftThreads;
}
// Pencils longer than fft_cap_long run the same values per lane on
// workgroups of kFftThreadsXL lanes, up to fft_cap_xl (8192 Float32 or 4096
// Float64 points, 64 KiB per pencil buffer).  The z convolution then reads
// its twiddles from global memory: its two pencil buffers take 128 KiB.
constexpr int kFftThreadsXL = 512;
template <typename T>
constexpr int fft_cap_xl()
{
    return fft_elems_long<T>() * kFftThreadsXL;
}
// Calls f(std::integral_constant<int, E>{}, std::integral_constant<int, NT>{})
// with the values per lane and the workgroup size for pencils of length L.
template <typename T, typename F>
void with_fft_elems(int L, F &&f)
{
    using std::integral_constant;
    if (L <= fft_cap<T>())
        f(integral_constant<int, fft_elems<T>()>{}, integral_constant<int, kFftThreads>{});
    else if (L <= fft_cap_long<T>())
        f(integral_constant<int, fft_elems_long<T>()>{}, integral_constant<int, kFftThreads>{});
    else
        f(integral_constant<int, fft_elems_long<T>()>{}, integral_constant<int, kFftThreadsXL>{});
}
constexpr int kFftMaxRadix = 7;

bool make_plan(int n, FftPlan *p);  // false if n has a prime factor > kFftMaxRadix

// ---------------------------------------------------------------------------
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    struct Rec {
        int kid;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[K_COUNT] = {};
    int64_t launches[K_COUNT] = {};
    // bounds-reduction partials, allocated on first use
    double *red = nullptr;
    // device address of this GPU's fixed scratch array (sc_interpolate_noalloc)
    void *scratch = nullptr;
    // second stream (default priority, non-blocking), created on first use:
    // the Green-function build of a solve runs on it, concurrently with the
    // deposit still executing on `stream`
    hipStream_t side = nullptr;
    // batched calls: bunch b runs on bstreams[b % pool_size]; these fork
    // from `stream` (fork event) and join back (one event per stream)
    int pool_size = 4;
    std::vector<hipStream_t> bstreams;
    std::vector<hipEvent_t> bjoin;
    hipEvent_t bfork = nullptr;
    // set while a batched call launches: the solve keeps the Green build on
    // the bunch's own stream (the bunches overlap each other instead)
    bool in_batch = false;
    // grouped launches (one launch per phase for all bunches of a batched
    // deposit or interpolation): the per-bunch descriptor table in device
    // memory, allocated on first use; `grouped` off falls back to one
    // launch sequence per bunch on the stream pool
    void *gtable = nullptr;
    int grouped = 3;  // SC_GROUP_PARTICLES | SC_GROUP_SOLVE
    // grouped batched solves: per-bunch buffer table (device memory, first use)
    void *stable = nullptr;
};

// Device workspace of one mesh (lazy, like src/mesh.jl:49-72).
struct Workspace {
    bool ready = false;
    int dtype = 0;
    int64_t bytes = 0;
    // spectra in the mesh type T: A = rho after the x and y transforms,
    // B = three field spectra, each [nz][2ny][nx+1] complex, kx fastest.
    void *A = nullptr;
    void *B = nullptr;
    // transformed Green functions in T: free space [nz+1][ny+1][nx+1] real x3,
    // image (cathode) [2nz][ny+1][nx+1] complex x3.
    void *Gf = nullptr;
    void *Gi = nullptr;
    // Green build scratch: the antiderivative at the octant nodes (Float64,
    // free-space and image components) and the partially transformed cell
    // rows (T).
    double *F = nullptr;
    void *R = nullptr;
    // twiddle tables in T for the lengths 2nx, 2ny, 2nz
    void *twT[3] = {nullptr, nullptr, nullptr};
    FftPlan plan[3];
    // ordering of the Green build on the side stream: green_done after the
    // build (the z convolution waits for it), gf_free after the z
    // convolution (the next eager build waits for it; recorded by eager
    // calls only)
    hipEvent_t green_done = nullptr, gf_free = nullptr;
    bool gf_free_recorded = false;
    // recorded where each deposit into this mesh starts (mark_deposit_start):
    // the next build forks from it, so that it overlaps the deposit kernels
    // (latency bound) rather than the previous step's interpolation.
    // fork_ready: recorded by an eager deposit that no build has forked from
    // yet; start_recorded: an eager record exists (an older one is still a
    // safe fork point for an eager build, which also waits for gf_free after
    // the last eager z convolution, as long as no capture has used this
    // mesh: a graph's z convolutions record no gf_free); captured: a stream
    // capture has used this mesh's events; fork_capture: the id of the
    // capture whose deposit recorded the event (0: none).  A build with no
    // usable record forks from the end of the stream.
    hipEvent_t deposit_start = nullptr;
    bool fork_ready = false;
    bool start_recorded = false;
    bool captured = false;
    unsigned long long fork_capture = 0;
    // cached Green key for the opt-in reuse
    bool g_valid = false;
    double g_key[6] = {};
    int g_cath = 0;  // solve mode the cached spectra were built for
};

// Scratch of the tile-binned deposit, grown on demand.
struct DepositBins {
    void *sorted = nullptr;       // (x, y, z, q) records in tile order
    size_t sorted_bytes = 0;
    void *chunks = nullptr;       // int4 (tile, start, length, 0)
    void *partials = nullptr;     // one 17^3 sub-grid per chunk and channel
    int64_t chunk_cap = 0;        // chunks, and sub-grids per channel set of partials
    int partial_sets = 0;         // channel sets that partials holds (1 or 3)
    unsigned *counts = nullptr;   // per-tile totals | tile start | chunk_base | nchunks
    unsigned *cmat = nullptr;     // [tile][binning workgroup] counts, then offsets
    int64_t cmat_cap = 0;         // entries
    void *qs = nullptr;           // charges of channels 1, 2 in tile order (multi-channel deposit)
    size_t qs_bytes = 0;
};

struct Mesh {
    Ctx *ctx = nullptr;
    int dtype = 0;
    int64_t n[3] = {0, 0, 0};
    void *rho = nullptr;
    void *efield = nullptr;
    bool owns = false;
    bool green_reuse = false;
    bool generic_transforms = false;  // diagnostic: skip the register-resident engine
    Workspace ws;
    DepositBins bins;
    void *efi = nullptr;  // node-interleaved field for the interpolation gather
    int efi_pair = 1;     // entries per node in efi: 2 = node pairs (large Float64 meshes)
    int efi_layout = 0;   // requested efi_pair: 0 = by size (ensure_efi), 1 or 2 = forced
    // device-resident geometry (lo[3], hi[3], delta[3] as doubles holding
    // mesh-type values), written by sc_mesh_auto_bounds on the stream; the
    // step functions read it when called without bounds
    double *geom = nullptr;
    bool geom_set = false;
    // binned-deposit tiling: 2^chunk_log2 particles per deposit workgroup
    // (kChunkLog2Min..kChunkLog2Max) and about 2^slice_log2 particles per
    // binning workgroup (kSliceLog2Min..kSliceLog2Max); set by hand or by
    // the autotuner (sc_mesh_autotune_deposit)
    int chunk_log2 = 13;
    int slice_log2 = 12;
};

constexpr int kChunkLog2Min = 10, kChunkLog2Max = 14;
constexpr int kSliceLog2Min = 10, kSliceLog2Max = 13;

// ---------------------------------------------------------------------------
// Errors
int set_err(Ctx *ctx, int code, const std::string &msg);
int hip_err(Ctx *ctx, hipError_t e, const char *what);

#define SC_HIP(ctx, call)                                   \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return ::sc::hip_err((ctx), e_, #call); \
    } while (0)

// Launch bracketing for the per-kernel timers.
void launch_begin(Ctx *ctx, int kid);
void launch_end(Ctx *ctx, int kid);

// Kernel-launch entry points (defined in the .hip files).
int launch_clear(Ctx *ctx, void *p, int64_t bytes);
int launch_deposit(Mesh *m, int pdt, const void *x, const void *y, const void *z, const void *q,
                   int64_t n, const double lo[3], const double dl[3], int clear);
// The deposit of launch_deposit into `out` (nx*ny*nz of the mesh type)
// instead of m->rho.
int launch_deposit_to(Mesh *m, int pdt, const void *x, const void *y, const void *z, const void *q,
                      int64_t n, const double lo[3], const double dl[3], int clear, void *out);
// The deposit (launch_deposit) of nb bunches with one launch per phase,
// grouped over the bunches, on ctx->stream.  lo and dl hold 3 values per
// bunch.  Returns SC_ERR_UNSUPPORTED, having launched nothing, when the
// batch does not qualify (see the definition).
int launch_deposit_grouped(Ctx *ctx, Mesh *const *ms, int nb, int pdt, const void *const *x,
                           const void *const *y, const void *const *z, const void *const *q,
                           const int64_t *n, const double *lo, const double *dl, int clear);
constexpr int kGroupMax = 256;  // bunches per grouped call
constexpr int kGroupStore = 8;  // descriptors per store launch (kernel arguments stay small)
// The context's descriptor table of grouped launches (allocated once, of
// group_table_bytes(); SC_ERR_UNSUPPORTED when first needed under capture).
int ensure_group_table(Ctx *ctx, size_t bytes);
size_t group_table_bytes();
// Grouped batched solves (sc_solve.hip): whether a mesh can join a class
// (register-engine lengths, no Green reuse, no forced generic engine); the
// context's table (SC_ERR_UNSUPPORTED when first needed under capture);
// the field solve of k meshes of one shape and type on ctx->stream, table
// entries slot .. slot + k - 1, per-bunch bounds (3 values each) and gamma.
bool solve_groupable(const Mesh *m);
// Bytes of the spectra a solve of m streams between its passes; classes
// hold at most kGroupSolveBytes of them (larger meshes solve one by one,
// where each pass finds the previous one's output in the 256 MB cache).
int64_t solve_group_bytes(const Mesh *m);
constexpr int64_t kGroupSolveBytes = (int64_t)64 << 20;
int ensure_solve_table(Ctx *ctx);
int solve_class(Ctx *ctx, Mesh *const *ms, int k, const double *const *lo, const double *const *hi,
                const double *const *dl, const double *gamma, int cath, int slot);
// interpolate_field of nb bunches, grouped likewise (interleave, gather).
int launch_interp_grouped(Ctx *ctx, Mesh *const *ms, int nb, int pdt, const void *const *x,
                          const void *const *y, const void *const *z, const int64_t *n,
                          const double *lo, const double *dl, void *const *Ex, void *const *Ey,
                          void *const *Ez);
int launch_deposit_binned(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                          const void *q, int64_t n, const double lo[3], const double dl[3],
                          int clear, void *out);
// nch (1..3) deposits of the same particles with charges q[c] into out[c],
// binned once; SC_ERR_UNSUPPORTED (nothing launched) where the binned path
// does not apply.
int launch_deposit_binned_multi(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                                const void *const *q, int nch, int64_t n, const double lo[3],
                                const double dl[3], int clear, void *const *out);
// Tile-order permutation of the particles (perm[slot] = particle index).
int launch_sort_perm(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n,
                     const double lo[3], const double dl[3], int *perm);
// Vector-Jacobian products of deposit (g_rho -> g_q, g_x, g_y, g_z) and of
// interpolation (g_E -> g_efield, g_x, g_y, g_z) at the mesh's current
// efield; NULL outputs are skipped; position gradients are added to the
// outputs when accumulate != 0.
int launch_deposit_vjp(Mesh *m, int pdt, const void *x, const void *y, const void *z,
                       const void *q, int64_t n, const double lo[3], const double dl[3],
                       const void *g_rho, void *gq, void *gx, void *gy, void *gz, int accumulate);
int launch_interp_vjp(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n,
                      const double lo[3], const double dl[3], const void *gEx, const void *gEy,
                      const void *gEz, void *g_efield, void *gx, void *gy, void *gz, int accumulate);
void free_bins(Mesh *m);
int launch_interp(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n,
                  const double lo[3], const double dl[3], void *Ex, void *Ey, void *Ez);
// launch_interp with the node-interleaved field copy in the caller's buffer
// `efi` (P entries per node, nx*ny*nz*P*4 mesh-type values); allocates nothing.
int launch_interp_into(Mesh *m, void *efi, int P, int pdt, const void *x, const void *y,
                       const void *z, int64_t n, const double lo[3], const double dl[3], void *Ex,
                       void *Ey, void *Ez);
int launch_bounds(Ctx *ctx, int pdt, const void *x, const void *y, const void *z, int64_t n,
                  double mn[3], double mx[3]);
// The particle-extrema bounds of src/mesh.jl:95-174 computed on the device
// into m->geom, stream-ordered, without a host round trip.
int launch_bounds_geom(Mesh *m, int pdt, const void *x, const void *y, const void *z, int64_t n);
int solve_mesh(Mesh *m, const double lo[3], const double hi[3], const double dl[3], double gamma,
               int at_cathode);
// Scalar potential (free space) of m->rho into phi (nx*ny*nz of the mesh type).
int solve_potential_mesh(Mesh *m, const double lo[3], const double hi[3], const double dl[3],
                         double gamma, void *phi);
// Transpose of the field solve: g_rho (nx*ny*nz) = S^T g_efield (3 nx*ny*nz),
// both of the mesh type.
int solve_adjoint_mesh(Mesh *m, const double lo[3], const double hi[3], const double dl[3],
                       double gamma, int at_cathode, const void *g_efield, void *g_rho);
int ensure_workspace(Mesh *m);
int ensure_side_stream(Mesh *m);
// true while a stream capture is active on st (its id in *id, 0 otherwise)
bool stream_capturing(hipStream_t st, unsigned long long *id);
// Records where a deposit into m starts on the context stream, for the next
// solve's Green build on the side stream (no-op before the first solve).
int mark_deposit_start(Mesh *m);
void free_workspace(Mesh *m);

// Task of workgroup `id` in a one-dimensional launch of `total` workgroups.
// The dispatcher hands consecutive workgroups to the eight XCDs in turn, and
// each XCD has an L2 of its own.  This order gives XCD x one contiguous range
// of tasks, so the workgroups resident together on one XCD take neighbouring
// tasks and find the lines they share in that XCD's L2.
__device__ __forceinline__ int xcd_task(int id, int total)
{
    constexpr int kXcd = 8;
    const int q = total / kXcd, r = total % kXcd;
    const int x = id % kXcd, local = id / kXcd;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + local;
}

}  // namespace sc

struct sc_ctx {
    sc::Ctx c;
};
struct sc_mesh {
    sc::Mesh m;
};

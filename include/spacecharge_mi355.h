/*
 * spacecharge_mi355.h -- C ABI of the MI355X (gfx950) space-charge step.
 *
 * This is the drop-in boundary for SpaceCharge.jl's particle-in-cell path:
 * cloud-in-cell deposit, integrated-Green-function (IGF) solve by zero-padded
 * 2N FFT convolution (free space and cathode image), trilinear interpolation.
 * A host binding (Julia `ccall`, Python `ctypes`, C/C++) keeps the
 * reference's API -- Mesh3D, deposit!, clear_mesh!, solve!, interpolate_field
 * -- and forwards each call to one function below.
 *
 * Conventions
 *   - Plain C types only.  Every device pointer is a HIP device address
 *     (hipMalloc'd, or an allocation from a host framework such as torch).
 *   - Grids use the reference's layout (src/mesh.jl:19-34): rho[i,j,k] at
 *     i + nx*(j + ny*k) (x fastest); efield adds M*c for component c
 *     (M = nx*ny*nz; components slowest); particles are one array per
 *     coordinate.
 *   - dtype codes are element sizes: SC_F32 = 4 (Float32), SC_F64 = 8
 *     (Float64).  The mesh type T and the particle type may differ, as in
 *     the reference; arithmetic uses the wider of the two.
 *   - Mesh bounds, spacings and gamma cross as doubles that hold the mesh
 *     type's values (a Float32 mesh passes Float32-representable doubles).
 *   - All work is stream-ordered on the context's stream and asynchronous,
 *     like the reference's GPU backend (no synchronisation inside deposit!,
 *     solve! or interpolate_field).  Functions that return host values
 *     (sc_bounds_from_particles) synchronise the stream.
 *   - Calls on one context must come from one host thread at a time;
 *     different contexts are independent.
 *   - Status: 0 on success, otherwise an SC_ERR_* code; the message is
 *     available from sc_last_error(ctx) (thread-local when ctx is NULL).
 */
#ifndef SPACECHARGE_MI355_H
#define SPACECHARGE_MI355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SC_F32 4
#define SC_F64 8

#define SC_OK 0
#define SC_ERR_INVALID 1     /* bad argument (message says which)           */
#define SC_ERR_DEVICE 2      /* HIP runtime error                            */
#define SC_ERR_NOMEM 3       /* device allocation failed                     */
#define SC_ERR_UNSUPPORTED 4 /* valid request this build does not implement  */

typedef struct sc_ctx sc_ctx;
typedef struct sc_mesh sc_mesh;

/* ---- context --------------------------------------------------------------
 * No reference counterpart as a value: the reference selects its backend by
 * dispatch on the array type (src/deposition.jl:234-246) and launches on the
 * task-local stream.  A context pins one device and one stream. */
int sc_ctx_create(int device, sc_ctx **out);
int sc_ctx_destroy(sc_ctx *ctx);
/* Use an existing HIP stream (a hipStream_t; NULL = the null stream). */
int sc_ctx_set_stream(sc_ctx *ctx, void *stream);
void *sc_ctx_get_stream(sc_ctx *ctx);
/* Blocks until all work queued on the context's stream has finished
 * (the reference's callers use CUDA.synchronize(),
 * benchmark/full_pipeline_benchmark.jl:47). */
int sc_synchronize(sc_ctx *ctx);
const char *sc_last_error(sc_ctx *ctx);
int sc_device_count(int *count);

/* ---- device memory, for hosts without a device-array type ----------------
 * Replaces KernelAbstractions.zeros / similar on the GPU backend
 * (src/mesh.jl:159-160, src/interpolation.jl:110-112). */
int sc_alloc(sc_ctx *ctx, size_t bytes, void **out);
int sc_free(sc_ctx *ctx, void *ptr);
int sc_memcpy_h2d(sc_ctx *ctx, void *dst, const void *src, size_t bytes);
int sc_memcpy_d2h(sc_ctx *ctx, void *dst, const void *src, size_t bytes);
int sc_memcpy_d2d(sc_ctx *ctx, void *dst, const void *src, size_t bytes);
int sc_memset(sc_ctx *ctx, void *dst, int value, size_t bytes);

/* ---- mesh -----------------------------------------------------------------
 * Replaces Mesh3D{T,A,B} (src/mesh.jl:19-34) and its lazily created solver
 * workspace (_get_workspace, src/mesh.jl:49-72).  sc_mesh_create allocates
 * and zeroes rho (M elements) and efield (3M elements) like
 * KernelAbstractions.zeros (src/mesh.jl:159-160, 222-223);
 * sc_mesh_create_with_buffers wraps caller-owned device buffers of those
 * sizes instead.  The workspace is allocated on the first sc_solve and freed
 * by sc_mesh_destroy.  Errors: grid < 2 in any axis ("All elements of
 * grid_size must be at least 2.", src/mesh.jl:106-108). */
int sc_mesh_create(sc_ctx *ctx, int dtype, int64_t nx, int64_t ny, int64_t nz, sc_mesh **out);
int sc_mesh_create_with_buffers(sc_ctx *ctx, int dtype, int64_t nx, int64_t ny, int64_t nz,
                                void *rho, void *efield, sc_mesh **out);
int sc_mesh_destroy(sc_mesh *mesh);
void *sc_mesh_rho(sc_mesh *mesh);
void *sc_mesh_efield(sc_mesh *mesh);
/* Bytes of device workspace the mesh currently holds (0 before the first solve). */
int64_t sc_mesh_workspace_bytes(sc_mesh *mesh);
/* Opt-in reuse of the transformed Green function between solves whose
 * (delta, gamma, at_cathode, zmin+zmax) are bit-identical.  Off by default,
 * so every sc_solve rebuilds it as solve! does (src/solvers/free_space.jl:79-86). */
int sc_mesh_set_green_reuse(sc_mesh *mesh, int enable);
/* Diagnostic (no reference counterpart): enable != 0 runs every transform on
 * the generic mixed-radix LDS engine even for power-of-two lengths that the
 * register-resident engine handles, so both engines can be checked against
 * the same results. */
int sc_mesh_force_generic_transforms(sc_mesh *mesh, int enable);
/* Diagnostic (no reference counterpart): the layout of the node-interleaved
 * field copy that the interpolation and its adjoint gather from.  layout 1
 * holds one (Ex, Ey, Ez, 0) entry per node; layout 2 holds node pairs (nodes
 * i and i+1 along x per entry; Float64 meshes only), which the library
 * picks by itself for every Float64 mesh; 0 (the default) restores that
 * choice.
 * Results are the same bits in every layout.  Changing it synchronises the
 * context's stream and frees the current copy. */
int sc_mesh_set_field_layout(sc_mesh *mesh, int layout);

/* ---- deposit tiling and its autotuner (no reference counterpart: the
 * reference launches one thread per particle with global atomics,
 * src/deposition.jl:106-158) ---------------------------------------------
 * The binned deposit (sc_deposit, the interpolation adjoint's cotangent
 * deposit, sc_sort_permutation) splits each 16^3-cell tile's particles into
 * chunks of 2^chunk_log2 particles, one deposit workgroup each
 * (chunk_log2 in [10, 14], default 13), and bins the particles in slices of
 * about 2^slice_log2 particles, one binning workgroup each (slice_log2 in
 * [10, 13], default 12).  Results agree to the deposit's round-off for every
 * setting; Float32 meshes give the same bits for the same setting.
 * sc_mesh_autotune_deposit times the deposit of the given particles at every
 * setting (HIP events on the context's stream: one untimed call, then reps
 * timed calls), keeps the fastest on the mesh and deposits once more with
 * it, so rho ends holding that deposit.  times_ms (NULL allowed) receives
 * 5 x 4 mean times, row-major [chunk_log2 - 10][slice_log2 - 10].  It
 * synchronises and refuses to run under stream capture. */
int sc_mesh_set_deposit_tiling(sc_mesh *mesh, int chunk_log2, int slice_log2);
int sc_mesh_get_deposit_tiling(const sc_mesh *mesh, int *chunk_log2, int *slice_log2);
int sc_mesh_autotune_deposit(sc_mesh *mesh, int pdtype, const void *x, const void *y,
                             const void *z, const void *q, int64_t n, const double min_bounds[3],
                             const double delta[3], int reps, float *times_ms);

/* Replaces the particle-extrema constructor's bounds computation
 * (src/mesh.jl:95-174): device min/max of x, y, z, then
 *   delta = (max - min)/(n - 1) in the particle type,
 *   min -= 1e-6*delta, max += 1e-6*delta in Float64,
 *   delta = (max - min)/(n - 1), a zero delta becomes 1e-6,
 *   all three rounded to the mesh type.
 * Synchronises the stream.  Errors: grid < 2, n == 0 ("Particle arrays
 * cannot be empty."). */
int sc_bounds_from_particles(sc_ctx *ctx, int pdtype, const void *x, const void *y, const void *z,
                             int64_t n, const int64_t grid[3], int dtype, double min_bounds[3],
                             double max_bounds[3], double delta[3]);

/* Replaces extrema(particles_x) on a device array (src/mesh.jl:120-122), for
 * hosts that run the reference's particle-extrema constructor themselves
 * (the Julia binding's Base.extrema(::MIArray)): *mn, *mx = the smallest and
 * largest of x[0..n-1], as doubles holding particle-type values.
 * Synchronises the stream.  Errors: n == 0. */
int sc_extrema(sc_ctx *ctx, int pdtype, const void *x, int64_t n, double *mn, double *mx);

/* ---- device-resident bounds (no reference counterpart; serve tracking
 * loops that re-mesh every step without a host round trip) ----------------
 * sc_mesh_auto_bounds computes the bounds of Mesh3D(grid, x, y, z; T)
 * (src/mesh.jl:95-174, the same rules and bits as sc_bounds_from_particles)
 * on the device into the mesh's own geometry block, stream-ordered, without
 * synchronising.  Every step function then accepts NULL bounds (min_bounds
 * and delta NULL; for sc_solve and sc_solve_adjoint also max_bounds) and
 * reads that block instead, so the whole step, bounds included, can be
 * captured into a graph.  Without host spacings the solve rebuilds its
 * Green function every call (sc_mesh_set_green_reuse has no key to compare).
 * sc_mesh_read_bounds copies the block to the host (synchronises). */
int sc_mesh_auto_bounds(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                        int64_t n);
int sc_mesh_read_bounds(sc_mesh *mesh, double min_bounds[3], double max_bounds[3],
                        double delta[3]);

/* Replaces clear_mesh! (src/deposition.jl:10-12): rho = 0. */
int sc_clear_mesh(sc_mesh *mesh);

/* Replaces deposit! (src/deposition.jl:218-247) and its GPU kernel
 * (deposit_kernel!, :106-129; CIC core :28-86): for every particle, q times
 * the trilinear weights is added to the 8 surrounding nodes of rho.
 * clear != 0 zeroes rho first (:230-232).  Particles whose lower cell index
 * falls outside [0, n-2] are clamped to it (extrapolating weights, charge
 * still conserved) instead of writing outside the mesh. */
int sc_deposit(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
               const void *q, int64_t n, const double min_bounds[3], const double delta[3],
               int clear);

/* Replaces solve!(mesh; at_cathode) (src/solvers/free_space.jl:14-47) and
 * solve_freespace! (:56-101): efield[c] = FPEI * (rho (*) G_c) with the
 * integrated Green function of src/green_functions.jl:35-112 built for
 * (delta, gamma), as a linear convolution on the 2N zero-padded grid.  With
 * at_cathode != 0 the image charge -rho flipped in z, at z-offset
 * min_bounds[2] + max_bounds[2] (:34-45), is added.  The Green function is
 * evaluated and differenced in Float64 for both mesh types; the transforms
 * and products run in the mesh type.  Any axis of 2 to 4096 nodes on a
 * Float32 mesh (2 to 2048 on a Float64 mesh) is solved; beyond that the
 * call returns SC_ERR_UNSUPPORTED (one transform pencil per workgroup). */
int sc_solve(sc_mesh *mesh, const double min_bounds[3], const double max_bounds[3],
             const double delta[3], double gamma, int at_cathode);

/* Scalar potential in free space.  No reference solver computes it; its
 * Green function is potential_green_function (src/green_functions.jl:13-22),
 * the antiderivative of 1/r, used exactly as sc_solve uses the field's:
 * phi = FPEI * gamma/(dx dy dz gamma) * (rho (*) cell integral of 1/r) on
 * the same zero-padded linear convolution.  This is the lab-frame potential
 * of a bunch moving along z (gamma times the rest-frame one), so that
 * -dphi/dx and -dphi/dy approximate sc_solve's Ex and Ey, and
 * -dphi/dz / gamma^2 its Ez.  phi: caller-owned device array of nx*ny*nz
 * elements of the mesh type, in rho's order.  Uses (and may rebuild) the
 * mesh's solver workspace. */
int sc_solve_potential(sc_mesh *mesh, const double delta[3], double gamma, void *phi);

/* Replaces interpolate_field (src/interpolation.jl:100-128) and
 * interpolate_kernel! (:17-86): Ex/Ey/Ez[p] = sum over the 8 surrounding
 * nodes of the trilinear weight times efield, written in the particle type
 * (caller-allocated, n elements each).  Same clamp policy as sc_deposit. */
int sc_interpolate(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                   int64_t n, const double min_bounds[3], const double delta[3], void *Ex,
                   void *Ey, void *Ez);

/* No counterpart in the reference: sc_interpolate for callers that cannot
 * allocate device memory (code under stream or graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, same kernels and
 * the same bits as sc_interpolate, but it calls no hipMalloc, hipFree or
 * allocating library routine: the node-interleaved field copy lives in a
 * fixed 64 MiB array per GPU, held from library load until the process ends,
 * instead of the mesh's own buffer.  That limits the mesh to 4,194,304 nodes
 * (Float32, e.g. 160^3) or 2,097,152 nodes (Float64, e.g. 128^3); larger
 * meshes return SC_ERR_INVALID with nothing launched.  All calls on one GPU
 * share the array, and it keeps the last call's contents: calls that overlap
 * on one GPU (different streams or contexts, concurrent graphs) race and
 * must be ordered by the caller.  Stream-ordered on the context's stream. */
int sc_interpolate_noalloc(sc_mesh *mesh, int pdtype, const void *x, const void *y,
                           const void *z, int64_t n, const double min_bounds[3],
                           const double delta[3], void *Ex, void *Ey, void *Ez);

/* Tile order of the particles (no reference counterpart): perm[i] is the
 * index of the particle that goes to position i when the particles are
 * grouped by the 16^3-cell tile of the deposit that holds them (tiles in
 * x-fastest order; the order inside a tile is unspecified).  A tracking
 * code applies it to all its per-particle arrays every few steps: deposit
 * and interpolation on tile-ordered particles touch the mesh in local
 * order, which makes the interpolation's gathers and the deposit's scatter
 * cheaper.  The step's results do not depend on the particle order
 * (Float32 meshes bit for bit).  Bounds as for sc_deposit (NULL for the
 * mesh's device-resident ones).  perm: n int32 on the device. */
int sc_sort_permutation(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                        int64_t n, const double min_bounds[3], const double delta[3],
                        int32_t *perm);

/* ---- vector-Jacobian products (no reference counterpart: the reference has
 * no derivatives; these serve differentiable tracking, e.g. a ChainRulesCore
 * rrule per function) -----------------------------------------------------
 * Each returns the cotangents of its forward function's inputs given those
 * of its outputs, with the mesh geometry (min_bounds, delta, gamma, bounds)
 * held fixed.  The CIC weights are linear in the position inside a cell;
 * position derivatives hold the cell index fixed, so a particle on a cell
 * face takes the derivative of the cell it is deposited in.  Position
 * gradients (g_x, g_y, g_z: all NULL or all non-NULL, particle type) are
 * overwritten, or added to when accumulate != 0, so that the deposit's and
 * the interpolation's contributions can be summed in place. */

/* VJP of sc_deposit: g_q[p] = sum_8 w g_rho (g_rho: nx*ny*nz of the mesh
 * type) and g_pos[p] = q[p] * d/dpos sum_8 w g_rho.  g_q may be NULL. */
int sc_deposit_adjoint(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                       const void *q, int64_t n, const double min_bounds[3], const double delta[3],
                       const void *g_rho, void *g_q, void *g_x, void *g_y, void *g_z,
                       int accumulate);

/* VJP of sc_solve: g_rho = S^T g_efield, S the field solve's linear map from
 * rho to efield (with the cathode image when at_cathode != 0).  g_efield:
 * 3*nx*ny*nz, g_rho: nx*ny*nz, both of the mesh type, caller-owned.  Uses
 * the mesh's solver workspace (not rho or efield). */
int sc_solve_adjoint(sc_mesh *mesh, const double min_bounds[3], const double max_bounds[3],
                     const double delta[3], double gamma, int at_cathode, const void *g_efield,
                     void *g_rho);

/* VJP of sc_interpolate at the mesh's current efield: g_efield (3*nx*ny*nz
 * of the mesh type, overwritten; may be NULL) = the deposit of (g_Ex, g_Ey,
 * g_Ez) with the same weights, and g_pos[p] = sum_c g_Ec[p] * d/dpos
 * sum_8 w efield_c. */
int sc_interpolate_adjoint(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                           int64_t n, const double min_bounds[3], const double delta[3],
                           const void *g_Ex, const void *g_Ey, const void *g_Ez, void *g_efield,
                           void *g_x, void *g_y, void *g_z, int accumulate);

/* ---- many bunches per call (no reference counterpart: the reference runs
 * one mesh per call, e.g. benchmark/full_pipeline_benchmark.jl:25-47; these
 * serve the batched configuration of many independent bunches) -------------
 * Each performs its single-bunch counterpart (sc_deposit, sc_solve,
 * sc_interpolate) for nb bunches, bunch b on meshes[b].  The meshes must be
 * distinct and belong to one context.  The bunches run concurrently on the
 * context's stream pool (sc_ctx_set_pool_size, default 4 streams, default
 * priority): the pool streams wait for the work already on the context's
 * stream, bunch b is launched on pool stream b mod P, and the context's
 * stream then waits for all of them, so a batched call is ordered like a
 * sequence of its single-bunch counterparts and computes the same results.
 * Per-bunch arrays have nb entries; bounds, deltas have 3*nb (bunch-major).
 * Grouped launches (sc_ctx_set_grouped_launch, phases a combination of
 * SC_GROUP_PARTICLES and SC_GROUP_SOLVE, both by default):
 *  - SC_GROUP_PARTICLES: when every bunch has particles, all meshes share
 *    one type and nb <= 256, sc_deposit_batched and sc_interpolate_batched
 *    run each phase of the binned deposit (7 phases) and of the
 *    interpolation (2) as ONE launch over the workgroups of all bunches, on
 *    the context's stream;
 *  - SC_GROUP_SOLVE: sc_solve_batched solves each class of two or more
 *    meshes of one shape, type and cathode flag (transform lengths 32..512,
 *    no Green reuse, at most 64 MB of spectra per class, so small meshes)
 *    with one launch per Green and transform kernel; classes and the other
 *    bunches share the stream pool.
 * The per-bunch descriptors go to tables in the context's device memory
 * through store launches that carry them as kernel arguments, so a graph
 * capture replays them; grouped calls of one context must not overlap
 * (they share the tables).  The first grouped call allocates the tables and
 * must run outside stream capture (under capture it falls back to the pool).
 * Results are the same bits as the pool path for Float32 meshes. */
#define SC_GROUP_PARTICLES 1
#define SC_GROUP_SOLVE 2
int sc_ctx_set_pool_size(sc_ctx *ctx, int nstreams);
int sc_ctx_set_grouped_launch(sc_ctx *ctx, int phases);
int sc_deposit_batched(sc_mesh *const *meshes, int nb, int pdtype, const void *const *x,
                       const void *const *y, const void *const *z, const void *const *q,
                       const int64_t *n, const double *min_bounds, const double *delta, int clear);
int sc_solve_batched(sc_mesh *const *meshes, int nb, const double *min_bounds,
                     const double *max_bounds, const double *delta, const double *gamma,
                     const int *at_cathode);
int sc_interpolate_batched(sc_mesh *const *meshes, int nb, int pdtype, const void *const *x,
                           const void *const *y, const void *const *z, const int64_t *n,
                           const double *min_bounds, const double *delta, void *const *Ex,
                           void *const *Ey, void *const *Ez);

/* ---- per-kernel timing (no reference counterpart; the reference profiles
 * with CUDA.@profile, benchmark/full_pipeline_profile.jl:108-111) ----------
 * When enabled, every kernel launch on the context is bracketed by HIP
 * events on the context's stream.  sc_timing_read synchronises, folds the
 * recorded events into per-kernel totals and returns them. */
int sc_timing_enable(sc_ctx *ctx, int enable);
int sc_timing_reset(sc_ctx *ctx);
int sc_kernel_count(void);
const char *sc_kernel_name(int kernel_id);
int sc_timing_read(sc_ctx *ctx, int kernel_id, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* SPACECHARGE_MI355_H */

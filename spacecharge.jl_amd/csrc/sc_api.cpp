// sc_api.cpp -- the extern "C" boundary of libspacecharge_mi355.so.
//
// Validation, error strings, contexts, meshes and the per-kernel timers.
// Error messages that have a reference counterpart reuse its wording
// (src/mesh.jl:106-116,206-211; src/deposition.jl:226-228) so a host
// binding can re-raise them unchanged.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include <mutex>

#include "sc_internal.h"

namespace sc {

const char *const kKernelNames[K_COUNT] = {
    "clear",        "deposit",     "interpolate", "bounds_partial", "bounds_final",
    "green_x",      "green_y",     "green_z",     "rho_fwd_x",      "rho_fwd_y",
    "zconv",        "inv_y",       "inv_x",       "bin_count",      "bin_scan",
    "bin_scatter",  "deposit_tiles", "interleave",  "gather_tiles",   "green_f",
    "bin_rowscan",  "tile_reduce", "deposit_vjp", "interp_vjp",
};

namespace {
thread_local std::string g_err;  // errors without a context
}

int set_err(Ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    else g_err = msg;
    return code;
}

int hip_err(Ctx *ctx, hipError_t e, const char *what)
{
    std::string m = std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) +
                    ") in " + what;
    return set_err(ctx, e == hipErrorOutOfMemory ? SC_ERR_NOMEM : SC_ERR_DEVICE, m);
}

static hipEvent_t take_event(Ctx *ctx)
{
    if (!ctx->pool.empty()) {
        hipEvent_t e = ctx->pool.back();
        ctx->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void launch_begin(Ctx *ctx, int kid)
{
    if (!ctx->timing) return;
    Ctx::Rec r;
    r.kid = kid;
    r.a = take_event(ctx);
    r.b = take_event(ctx);
    if (!r.a || !r.b) return;
    (void)hipEventRecord(r.a, ctx->stream);
    ctx->pending.push_back(r);
}

void launch_end(Ctx *ctx, int kid)
{
    if (!ctx->timing || ctx->pending.empty()) return;
    Ctx::Rec &r = ctx->pending.back();
    if (r.kid != kid) return;
    (void)hipEventRecord(r.b, ctx->stream);
}

static void fold_timings(Ctx *ctx)
{
    if (ctx->pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ctx->total_ms[r.kid] += ms;
            ctx->launches[r.kid] += 1;
        }
        ctx->pool.push_back(r.a);
        ctx->pool.push_back(r.b);
    }
    ctx->pending.clear();
}

static bool dtype_ok(int dt) { return dt == SC_F32 || dt == SC_F64; }

static int check_dev(Ctx *ctx)
{
    SC_HIP(ctx, hipSetDevice(ctx->device));
    return SC_OK;
}

// ---- the allocation-free interpolation's scratch ----
// One fixed array per GPU, present from the moment the library's code is
// loaded on it until the process ends: it holds the node-interleaved field
// copy of sc_interpolate_noalloc.  In host code `scratch` names a host-side
// stand-in; the array's device address is looked up once per device
// (scratch_address, called by sc_ctx_create with the device current) and
// only that address is handed to kernels.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;  // per device index, nullptr until looked up

static int scratch_address(Ctx *ctx, int device, void **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_err(ctx, e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return set_err(ctx, SC_ERR_DEVICE, "scratch array has no device address");
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return SC_OK;
}

}  // namespace sc

using namespace sc;

extern "C" {

int sc_device_count(int *count)
{
    if (!count) return set_err(nullptr, SC_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_err(nullptr, e, "hipGetDeviceCount");
    *count = n;
    return SC_OK;
}

int sc_ctx_create(int device, sc_ctx **out)
{
    if (!out) return set_err(nullptr, SC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_err(nullptr, e, "hipGetDeviceCount");
    if (device < 0 || device >= n)
        return set_err(nullptr, SC_ERR_INVALID,
                       "device " + std::to_string(device) + " out of range (" + std::to_string(n) +
                           " visible)");
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_err(nullptr, e, "hipSetDevice");
    void *scr = nullptr;
    if (int rc = scratch_address(nullptr, device, &scr)) return rc;
    sc_ctx *c = new (std::nothrow) sc_ctx();
    if (!c) return set_err(nullptr, SC_ERR_NOMEM, "host allocation failed");
    c->c.device = device;
    c->c.stream = nullptr;
    c->c.scratch = scr;
    *out = c;
    return SC_OK;
}

int sc_ctx_destroy(sc_ctx *ctx)
{
    if (!ctx) return SC_OK;
    Ctx *c = &ctx->c;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    if (c->red) (void)hipFree(c->red);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    for (auto s : c->bstreams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (auto e : c->bjoin) (void)hipEventDestroy(e);
    if (c->bfork) (void)hipEventDestroy(c->bfork);
    if (c->gtable) (void)hipFree(c->gtable);
    if (c->stable) (void)hipFree(c->stable);
    delete ctx;
    return SC_OK;
}

int sc_ctx_set_stream(sc_ctx *ctx, void *stream)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // pending timer events belong to the old stream: fold them before switching
    if (s != ctx->c.stream && !ctx->c.pending.empty()) fold_timings(&ctx->c);
    ctx->c.stream = s;
    return SC_OK;
}

void *sc_ctx_get_stream(sc_ctx *ctx) { return ctx ? (void *)ctx->c.stream : nullptr; }

int sc_synchronize(sc_ctx *ctx)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    SC_HIP(&ctx->c, hipStreamSynchronize(ctx->c.stream));
    return SC_OK;
}

const char *sc_last_error(sc_ctx *ctx) { return ctx ? ctx->c.err.c_str() : g_err.c_str(); }

// ---- memory ----
int sc_alloc(sc_ctx *ctx, size_t bytes, void **out)
{
    if (!ctx || !out) return set_err(ctx ? &ctx->c : nullptr, SC_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (int rc = check_dev(&ctx->c)) return rc;
    if (bytes == 0) return SC_OK;
    SC_HIP(&ctx->c, hipMalloc(out, bytes));
    return SC_OK;
}

int sc_free(sc_ctx *ctx, void *ptr)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (!ptr) return SC_OK;
    SC_HIP(&ctx->c, hipStreamSynchronize(ctx->c.stream));
    SC_HIP(&ctx->c, hipFree(ptr));
    return SC_OK;
}

int sc_memcpy_h2d(sc_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (bytes == 0) return SC_OK;
    if (!dst || !src) return set_err(&ctx->c, SC_ERR_INVALID, "NULL pointer");
    SC_HIP(&ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    SC_HIP(&ctx->c, hipStreamSynchronize(ctx->c.stream));
    return SC_OK;
}

int sc_memcpy_d2h(sc_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (bytes == 0) return SC_OK;
    if (!dst || !src) return set_err(&ctx->c, SC_ERR_INVALID, "NULL pointer");
    SC_HIP(&ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    SC_HIP(&ctx->c, hipStreamSynchronize(ctx->c.stream));
    return SC_OK;
}

int sc_memcpy_d2d(sc_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (bytes == 0) return SC_OK;
    if (!dst || !src) return set_err(&ctx->c, SC_ERR_INVALID, "NULL pointer");
    SC_HIP(&ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
    return SC_OK;
}

int sc_memset(sc_ctx *ctx, void *dst, int value, size_t bytes)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (bytes == 0) return SC_OK;
    if (!dst) return set_err(&ctx->c, SC_ERR_INVALID, "NULL pointer");
    SC_HIP(&ctx->c, hipMemsetAsync(dst, value, bytes, ctx->c.stream));
    return SC_OK;
}

// ---- mesh ----
static int mesh_new(sc_ctx *ctx, int dtype, int64_t nx, int64_t ny, int64_t nz, void *rho,
                    void *efield, bool owns, sc_mesh **out)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    Ctx *c = &ctx->c;
    if (!out) return set_err(c, SC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!dtype_ok(dtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (nx <= 1 || ny <= 1 || nz <= 1)
        return set_err(c, SC_ERR_INVALID, "All elements of grid_size must be at least 2.");
    if (nx > 1 << 20 || ny > 1 << 20 || nz > 1 << 20 || nx * ny * nz > (int64_t)1 << 34)
        return set_err(c, SC_ERR_INVALID, "grid_size too large");
    if (int rc = check_dev(c)) return rc;
    sc_mesh *m = new (std::nothrow) sc_mesh();
    if (!m) return set_err(c, SC_ERR_NOMEM, "host allocation failed");
    Mesh &M = m->m;
    M.ctx = c;
    M.dtype = dtype;
    M.n[0] = nx;
    M.n[1] = ny;
    M.n[2] = nz;
    const size_t cells = (size_t)(nx * ny * nz);
    const size_t s = (size_t)dtype;
    if (owns) {
        if (hipMalloc(&M.rho, cells * s) != hipSuccess ||
            hipMalloc(&M.efield, 3 * cells * s) != hipSuccess) {
            if (M.rho) (void)hipFree(M.rho);
            delete m;
            return set_err(c, SC_ERR_NOMEM, "device allocation of rho/efield failed");
        }
        M.owns = true;
        int rc = launch_clear(c, M.rho, (int64_t)(cells * s));
        if (rc == SC_OK) rc = launch_clear(c, M.efield, (int64_t)(3 * cells * s));
        if (rc != SC_OK) {
            (void)hipFree(M.rho);
            (void)hipFree(M.efield);
            delete m;
            return rc;
        }
    } else {
        if (!rho || !efield) {
            delete m;
            return set_err(c, SC_ERR_INVALID, "rho and efield buffers must be non-NULL");
        }
        M.rho = rho;
        M.efield = efield;
        M.owns = false;
    }
    *out = m;
    return SC_OK;
}

int sc_mesh_create(sc_ctx *ctx, int dtype, int64_t nx, int64_t ny, int64_t nz, sc_mesh **out)
{
    return mesh_new(ctx, dtype, nx, ny, nz, nullptr, nullptr, true, out);
}

int sc_mesh_create_with_buffers(sc_ctx *ctx, int dtype, int64_t nx, int64_t ny, int64_t nz,
                                void *rho, void *efield, sc_mesh **out)
{
    return mesh_new(ctx, dtype, nx, ny, nz, rho, efield, false, out);
}

int sc_mesh_destroy(sc_mesh *mesh)
{
    if (!mesh) return SC_OK;
    Mesh &M = mesh->m;
    (void)hipSetDevice(M.ctx->device);
    (void)hipStreamSynchronize(M.ctx->stream);
    free_workspace(&M);
    free_bins(&M);
    if (M.efi) (void)hipFree(M.efi);
    if (M.geom) (void)hipFree(M.geom);
    if (M.owns) {
        (void)hipFree(M.rho);
        (void)hipFree(M.efield);
    }
    delete mesh;
    return SC_OK;
}

void *sc_mesh_rho(sc_mesh *mesh) { return mesh ? mesh->m.rho : nullptr; }
void *sc_mesh_efield(sc_mesh *mesh) { return mesh ? mesh->m.efield : nullptr; }
int64_t sc_mesh_workspace_bytes(sc_mesh *mesh) { return mesh ? mesh->m.ws.bytes : 0; }

int sc_mesh_set_green_reuse(sc_mesh *mesh, int enable)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    mesh->m.green_reuse = enable != 0;
    if (!enable) mesh->m.ws.g_valid = false;
    return SC_OK;
}

int sc_mesh_force_generic_transforms(sc_mesh *mesh, int enable)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    mesh->m.generic_transforms = enable != 0;
    return SC_OK;
}

int sc_mesh_set_field_layout(sc_mesh *mesh, int layout)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    if (layout < 0 || layout > 2)
        return set_err(M.ctx, SC_ERR_INVALID, "field layout must be 0 (by size), 1 or 2");
    if (layout == 2 && M.dtype != SC_F64)
        return set_err(M.ctx, SC_ERR_INVALID, "node pairs (field layout 2) need a Float64 mesh");
    if (layout == M.efi_layout) return SC_OK;
    if (M.efi) {
        // the copy is rebuilt on the next interpolation; wait for work that reads it
        SC_HIP(M.ctx, hipSetDevice(M.ctx->device));
        SC_HIP(M.ctx, hipStreamSynchronize(M.ctx->stream));
        SC_HIP(M.ctx, hipFree(M.efi));
        M.efi = nullptr;
    }
    M.efi_layout = layout;
    return SC_OK;
}

// ---- bounds (src/mesh.jl:95-174) ----
int sc_bounds_from_particles(sc_ctx *ctx, int pdtype, const void *x, const void *y, const void *z,
                             int64_t n, const int64_t grid[3], int dtype, double lo[3],
                             double hi[3], double delta[3])
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    Ctx *c = &ctx->c;
    if (!grid || !lo || !hi || !delta) return set_err(c, SC_ERR_INVALID, "NULL argument");
    if (!dtype_ok(pdtype) || !dtype_ok(dtype))
        return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    for (int a = 0; a < 3; ++a)
        if (grid[a] <= 1)
            return set_err(c, SC_ERR_INVALID, "All elements of grid_size must be at least 2.");
    if (n <= 0) return set_err(c, SC_ERR_INVALID, "Particle arrays cannot be empty.");
    if (!x || !y || !z) return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
    if (int rc = check_dev(c)) return rc;
    double mn[3], mx[3];
    if (int rc = launch_bounds(c, pdtype, x, y, z, n, mn, mx)) return rc;
    for (int a = 0; a < 3; ++a) {
        const double nm1 = (double)(grid[a] - 1);
        // first spacing in the particle type (src/mesh.jl:127-129)
        double d0;
        if (pdtype == SC_F32) d0 = (double)(((float)mx[a] - (float)mn[a]) / (float)nm1);
        else d0 = (mx[a] - mn[a]) / nm1;
        // Float64 padding (:132-137), spacing again (:140-142), zero patch (:145-149)
        const double l = mn[a] - 1e-6 * d0;
        const double h = mx[a] + 1e-6 * d0;
        double d = (h - l) / nm1;
        if (d == 0.0) d = 1e-6;
        // conversion to the mesh type (:150-156)
        if (dtype == SC_F32) {
            lo[a] = (double)(float)l;
            hi[a] = (double)(float)h;
            delta[a] = (double)(float)d;
        } else {
            lo[a] = l;
            hi[a] = h;
            delta[a] = d;
        }
    }
    return SC_OK;
}

int sc_extrema(sc_ctx *ctx, int pdtype, const void *x, int64_t n, double *mn, double *mx)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    Ctx *c = &ctx->c;
    if (!mn || !mx) return set_err(c, SC_ERR_INVALID, "NULL argument");
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n <= 0) return set_err(c, SC_ERR_INVALID, "collection must be non-empty");
    if (!x) return set_err(c, SC_ERR_INVALID, "NULL array pointer");
    if (int rc = check_dev(c)) return rc;
    double lo[3], hi[3];
    // the three-axis reduction with the one array on every axis
    if (int rc = launch_bounds(c, pdtype, x, x, x, n, lo, hi)) return rc;
    *mn = lo[0];
    *mx = hi[0];
    return SC_OK;
}

// ---- the step ----
static int check_geom(Ctx *c, const double lo[3], const double dl[3])
{
    if (!lo || !dl) return set_err(c, SC_ERR_INVALID, "NULL bounds or delta");
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(lo[a]) || !std::isfinite(dl[a]) || !(dl[a] > 0.0))
            return set_err(c, SC_ERR_INVALID, "bounds must be finite and delta positive");
    }
    return SC_OK;
}

// Host bounds, or none (lo and dl NULL) when the mesh holds device-resident ones.
static int check_geom_m(Ctx *c, const Mesh &M, const double lo[3], const double dl[3])
{
    if (!lo && !dl) {
        if (!M.geom_set)
            return set_err(c, SC_ERR_INVALID,
                           "NULL bounds: the mesh has no device-resident bounds (sc_mesh_auto_bounds)");
        return SC_OK;
    }
    return check_geom(c, lo, dl);
}

// The solve's bounds: all three given (and checked), or all NULL for the
// mesh's device-resident ones.
static int check_solve_geom(Ctx *c, const Mesh &M, const double lo[3], const double hi[3],
                            const double dl[3], int at_cathode)
{
    if (!lo && !hi && !dl) return check_geom_m(c, M, lo, dl);
    if (int rc = check_geom(c, lo, dl)) return rc;
    if (!hi) return set_err(c, SC_ERR_INVALID, "NULL max_bounds");
    if (at_cathode && !std::isfinite(hi[2]))
        return set_err(c, SC_ERR_INVALID, "max_bounds must be finite");
    return SC_OK;
}

int sc_mesh_auto_bounds(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                        int64_t n)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n <= 0) return set_err(c, SC_ERR_INVALID, "Particle arrays cannot be empty.");
    if (!x || !y || !z) return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
    if (int rc = check_dev(c)) return rc;
    return launch_bounds_geom(&M, pdtype, x, y, z, n);
}

int sc_mesh_read_bounds(sc_mesh *mesh, double lo[3], double hi[3], double delta[3])
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!lo || !hi || !delta) return set_err(c, SC_ERR_INVALID, "NULL argument");
    if (!M.geom_set)
        return set_err(c, SC_ERR_INVALID, "the mesh has no device-resident bounds (sc_mesh_auto_bounds)");
    if (int rc = check_dev(c)) return rc;
    double h[9];
    SC_HIP(c, hipMemcpyAsync(h, M.geom, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    SC_HIP(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) {
        lo[a] = h[a];
        hi[a] = h[3 + a];
        delta[a] = h[6 + a];
    }
    return SC_OK;
}

int sc_clear_mesh(sc_mesh *mesh)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    if (int rc = check_dev(M.ctx)) return rc;
    return launch_clear(M.ctx, M.rho, M.n[0] * M.n[1] * M.n[2] * M.dtype);
}

int sc_deposit(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
               const void *q, int64_t n, const double lo[3], const double dl[3], int clear)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
    if (n > 0 && (!x || !y || !z || !q)) return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    return launch_deposit(&M, pdtype, x, y, z, q, n, lo, dl, clear);
}

int sc_mesh_set_deposit_tiling(sc_mesh *mesh, int chunk_log2, int slice_log2)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    if (chunk_log2 < kChunkLog2Min || chunk_log2 > kChunkLog2Max || slice_log2 < kSliceLog2Min ||
        slice_log2 > kSliceLog2Max)
        return set_err(M.ctx, SC_ERR_INVALID,
                       "chunk_log2 must lie in [" + std::to_string(kChunkLog2Min) + ", " +
                           std::to_string(kChunkLog2Max) + "] and slice_log2 in [" +
                           std::to_string(kSliceLog2Min) + ", " + std::to_string(kSliceLog2Max) + "]");
    M.chunk_log2 = chunk_log2;
    M.slice_log2 = slice_log2;
    return SC_OK;
}

int sc_mesh_get_deposit_tiling(const sc_mesh *mesh, int *chunk_log2, int *slice_log2)
{
    if (!mesh || !chunk_log2 || !slice_log2) return set_err(nullptr, SC_ERR_INVALID, "NULL argument");
    *chunk_log2 = mesh->m.chunk_log2;
    *slice_log2 = mesh->m.slice_log2;
    return SC_OK;
}

// Times the binned deposit of these particles at every (chunk, slice)
// setting with HIP events on the context's stream (one untimed call, then
// `reps` timed calls each), keeps the fastest on the mesh and deposits once
// more with it, so rho ends holding this deposit.  times_ms (may be NULL)
// receives the mean per call, row-major [chunk_log2][slice_log2].
int sc_mesh_autotune_deposit(sc_mesh *mesh, int pdtype, const void *x, const void *y,
                             const void *z, const void *q, int64_t n, const double lo[3],
                             const double dl[3], int reps, float *times_ms)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n <= 0) return set_err(c, SC_ERR_INVALID, "the autotuner needs at least one particle");
    if (!x || !y || !z || !q) return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
    if (reps < 1 || reps > 1000) return set_err(c, SC_ERR_INVALID, "reps must lie in [1, 1000]");
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SC_HIP(c, hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return set_err(c, SC_ERR_INVALID, "the autotuner times kernels and cannot run under stream capture");
    hipEvent_t ev[2] = {nullptr, nullptr};
    SC_HIP(c, hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return set_err(c, SC_ERR_DEVICE, "hipEventCreate failed");
    }
    const int keep_c = M.chunk_log2, keep_s = M.slice_log2;
    int best_c = keep_c, best_s = keep_s, rc = SC_OK;
    float best = -1.f;
    for (int cl = kChunkLog2Min; cl <= kChunkLog2Max && rc == SC_OK; ++cl)
        for (int sl = kSliceLog2Min; sl <= kSliceLog2Max && rc == SC_OK; ++sl) {
            M.chunk_log2 = cl;
            M.slice_log2 = sl;
            rc = launch_deposit(&M, pdtype, x, y, z, q, n, lo, dl, 1);  // allocates, warms
            if (rc == SC_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = SC_ERR_DEVICE;
            for (int r = 0; r < reps && rc == SC_OK; ++r)
                rc = launch_deposit(&M, pdtype, x, y, z, q, n, lo, dl, 1);
            float ms = 0.f;
            if (rc == SC_OK && (hipEventRecord(ev[1], c->stream) != hipSuccess ||
                                hipEventSynchronize(ev[1]) != hipSuccess ||
                                hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
                rc = set_err(c, SC_ERR_DEVICE, "timing the deposit failed");
            if (rc != SC_OK) break;
            ms /= (float)reps;
            if (times_ms)
                times_ms[(cl - kChunkLog2Min) * (kSliceLog2Max - kSliceLog2Min + 1) + (sl - kSliceLog2Min)] = ms;
            if (best < 0.f || ms < best) {
                best = ms;
                best_c = cl;
                best_s = sl;
            }
        }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc != SC_OK) {
        M.chunk_log2 = keep_c;
        M.slice_log2 = keep_s;
        return rc;
    }
    M.chunk_log2 = best_c;
    M.slice_log2 = best_s;
    return launch_deposit(&M, pdtype, x, y, z, q, n, lo, dl, 1);
}

int sc_solve(sc_mesh *mesh, const double lo[3], const double hi[3], const double dl[3],
             double gamma, int at_cathode)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (int rc = check_solve_geom(c, M, lo, hi, dl, at_cathode)) return rc;
    if (!std::isfinite(gamma) || !(gamma > 0.0))
        return set_err(c, SC_ERR_INVALID, "gamma must be finite and positive");
    if (int rc = check_dev(c)) return rc;
    return solve_mesh(&M, lo, hi, dl, gamma, at_cathode);
}

int sc_solve_potential(sc_mesh *mesh, const double dl[3], double gamma, void *phi)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    static const double origin[3] = {0.0, 0.0, 0.0};  // free space: no bounds enter
    if (int rc = dl ? check_geom(c, origin, dl) : check_geom_m(c, M, nullptr, nullptr)) return rc;
    if (!std::isfinite(gamma) || !(gamma > 0.0))
        return set_err(c, SC_ERR_INVALID, "gamma must be finite and positive");
    if (!phi) return set_err(c, SC_ERR_INVALID, "NULL phi");
    if (int rc = check_dev(c)) return rc;
    return solve_potential_mesh(&M, dl ? origin : nullptr, dl ? origin : nullptr, dl, gamma, phi);
}

int sc_interpolate(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                   int64_t n, const double lo[3], const double dl[3], void *Ex, void *Ey, void *Ez)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
    if (n > 0 && (!x || !y || !z || !Ex || !Ey || !Ez))
        return set_err(c, SC_ERR_INVALID, "NULL particle or output pointer");
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    return launch_interp(&M, pdtype, x, y, z, n, lo, dl, Ex, Ey, Ez);
}

int sc_interpolate_noalloc(sc_mesh *mesh, int pdtype, const void *x, const void *y,
                           const void *z, int64_t n, const double lo[3], const double dl[3],
                           void *Ex, void *Ey, void *Ez)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
    if (n > 0 && (!x || !y || !z || !Ex || !Ey || !Ez))
        return set_err(c, SC_ERR_INVALID, "NULL particle or output pointer");
    // one (Ex, Ey, Ez, 0) entry of the mesh type per node
    const size_t need = (size_t)M.n[0] * (size_t)M.n[1] * (size_t)M.n[2] * 4 * (size_t)M.dtype;
    if (need > kScratchBytes)
        return set_err(c, SC_ERR_INVALID,
                       "mesh too large for the allocation-free interpolation: its field copy "
                       "needs " + std::to_string(need) + " bytes, the scratch holds " +
                           std::to_string(kScratchBytes));
    if (!c->scratch) return set_err(c, SC_ERR_DEVICE, "context has no scratch address");
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    return launch_interp_into(&M, c->scratch, 1, pdtype, x, y, z, n, lo, dl, Ex, Ey, Ez);
}

int sc_sort_permutation(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                        int64_t n, const double lo[3], const double dl[3], int32_t *perm)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
    if (n > 0 && (!x || !y || !z || !perm)) return set_err(c, SC_ERR_INVALID, "NULL pointer");
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    if (n == 0) return SC_OK;
    return launch_sort_perm(&M, pdtype, x, y, z, n, lo, dl, perm);
}

// ---- vector-Jacobian products (no reference counterpart) ----
static int check_pos_outputs(Ctx *c, const void *gx, const void *gy, const void *gz)
{
    if ((gx == nullptr) != (gy == nullptr) || (gx == nullptr) != (gz == nullptr))
        return set_err(c, SC_ERR_INVALID, "g_x, g_y, g_z must be all NULL or all non-NULL");
    return SC_OK;
}

int sc_deposit_adjoint(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                       const void *q, int64_t n, const double lo[3], const double dl[3],
                       const void *g_rho, void *g_q, void *g_x, void *g_y, void *g_z, int accumulate)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
    if (int rc = check_pos_outputs(c, g_x, g_y, g_z)) return rc;
    if (n > 0 && (g_q || g_x)) {
        if (!x || !y || !z) return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
        if (!g_rho) return set_err(c, SC_ERR_INVALID, "NULL g_rho");
        if (g_x && !q) return set_err(c, SC_ERR_INVALID, "position gradients need q");
    }
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    return launch_deposit_vjp(&M, pdtype, x, y, z, q, n, lo, dl, g_rho, g_q, g_x, g_y, g_z,
                              accumulate);
}

int sc_solve_adjoint(sc_mesh *mesh, const double lo[3], const double hi[3], const double dl[3],
                     double gamma, int at_cathode, const void *g_efield, void *g_rho)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (int rc = check_solve_geom(c, M, lo, hi, dl, at_cathode)) return rc;
    if (!std::isfinite(gamma) || !(gamma > 0.0))
        return set_err(c, SC_ERR_INVALID, "gamma must be finite and positive");
    if (!g_efield || !g_rho) return set_err(c, SC_ERR_INVALID, "NULL g_efield or g_rho");
    if (int rc = check_dev(c)) return rc;
    return solve_adjoint_mesh(&M, lo, hi, dl, gamma, at_cathode, g_efield, g_rho);
}

int sc_interpolate_adjoint(sc_mesh *mesh, int pdtype, const void *x, const void *y, const void *z,
                           int64_t n, const double lo[3], const double dl[3], const void *g_Ex,
                           const void *g_Ey, const void *g_Ez, void *g_efield, void *g_x, void *g_y,
                           void *g_z, int accumulate)
{
    if (!mesh) return set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
    Mesh &M = mesh->m;
    Ctx *c = M.ctx;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (n < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
    if (int rc = check_pos_outputs(c, g_x, g_y, g_z)) return rc;
    if (n > 0 && (g_efield || g_x)) {
        if (!x || !y || !z) return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
        if (!g_Ex || !g_Ey || !g_Ez) return set_err(c, SC_ERR_INVALID, "NULL field cotangent");
    }
    if (int rc = check_geom_m(c, M, lo, dl)) return rc;
    if (int rc = check_dev(c)) return rc;
    return launch_interp_vjp(&M, pdtype, x, y, z, n, lo, dl, g_Ex, g_Ey, g_Ez, g_efield, g_x, g_y,
                             g_z, accumulate);
}

// ---- many bunches per call ----
int sc_ctx_set_grouped_launch(sc_ctx *ctx, int phases)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (phases & ~(SC_GROUP_PARTICLES | SC_GROUP_SOLVE))
        return set_err(&ctx->c, SC_ERR_INVALID, "phases must combine SC_GROUP_PARTICLES and SC_GROUP_SOLVE");
    ctx->c.grouped = phases;
    return SC_OK;
}

int sc_ctx_set_pool_size(sc_ctx *ctx, int nstreams)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (nstreams < 1 || nstreams > 16)
        return set_err(&ctx->c, SC_ERR_INVALID, "pool size must be in [1, 16]");
    ctx->c.pool_size = nstreams;
    return SC_OK;
}

}  // extern "C"

namespace {

// Validates a batch of meshes (non-NULL, distinct, one context) and returns
// that context.
Ctx *batch_ctx(sc_mesh *const *meshes, int nb, int *rc)
{
    *rc = SC_OK;
    if (!meshes || nb < 1) {
        *rc = set_err(nullptr, SC_ERR_INVALID, "a batch needs at least one mesh");
        return nullptr;
    }
    if (!meshes[0]) {
        *rc = set_err(nullptr, SC_ERR_INVALID, "mesh is NULL");
        return nullptr;
    }
    Ctx *c = meshes[0]->m.ctx;
    for (int b = 0; b < nb; ++b) {
        if (!meshes[b]) {
            *rc = set_err(c, SC_ERR_INVALID, "mesh is NULL");
            return nullptr;
        }
        if (meshes[b]->m.ctx != c) {
            *rc = set_err(c, SC_ERR_INVALID, "all meshes of a batch must share one context");
            return nullptr;
        }
        // the bunches run concurrently: each needs its own mesh (and workspace)
        for (int a = 0; a < b; ++a) {
            if (meshes[a] == meshes[b]) {
                *rc = set_err(c, SC_ERR_INVALID, "a mesh appears twice in one batch");
                return nullptr;
            }
        }
    }
    return c;
}

// Runs op(b) for every bunch with c->stream set to its pool stream, between
// a fork from and a join back to the caller's stream.
template <typename Op>
int run_batched(Ctx *c, int nb, Op op)
{
    if (int rc = check_dev(c)) return rc;
    const int P = nb < c->pool_size ? nb : c->pool_size;
    while ((int)c->bstreams.size() < P) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        SC_HIP(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
        c->bstreams.push_back(s);
        SC_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->bjoin.push_back(e);
    }
    if (!c->bfork) SC_HIP(c, hipEventCreateWithFlags(&c->bfork, hipEventDisableTiming));
    hipStream_t home = c->stream;
    SC_HIP(c, hipEventRecord(c->bfork, home));
    for (int i = 0; i < P; ++i) SC_HIP(c, hipStreamWaitEvent(c->bstreams[i], c->bfork, 0));
    int rc = SC_OK;
    c->in_batch = true;
    for (int b = 0; b < nb && rc == SC_OK; ++b) {
        c->stream = c->bstreams[b % P];
        rc = op(b);
    }
    c->stream = home;
    c->in_batch = false;
    // join every used stream, also after an error, so the caller's stream
    // orders after whatever was launched
    for (int i = 0; i < P; ++i) {
        if (hipEventRecord(c->bjoin[i], c->bstreams[i]) != hipSuccess ||
            hipStreamWaitEvent(home, c->bjoin[i], 0) != hipSuccess)
            if (rc == SC_OK) rc = set_err(c, SC_ERR_DEVICE, "joining the batch streams failed");
    }
    return rc;
}

}  // namespace

extern "C" {

int sc_deposit_batched(sc_mesh *const *meshes, int nb, int pdtype, const void *const *x,
                       const void *const *y, const void *const *z, const void *const *q,
                       const int64_t *n, const double *min_bounds, const double *delta, int clear)
{
    int rc;
    Ctx *c = batch_ctx(meshes, nb, &rc);
    if (!c) return rc;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (!x || !y || !z || !q || !n || !min_bounds || !delta)
        return set_err(c, SC_ERR_INVALID, "NULL per-bunch array");
    for (int b = 0; b < nb; ++b) {
        if (n[b] < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
        if (n[b] > 0 && (!x[b] || !y[b] || !z[b] || !q[b]))
            return set_err(c, SC_ERR_INVALID, "NULL particle pointer");
        if (int e = check_geom(c, min_bounds + 3 * b, delta + 3 * b)) return e;
    }
    if (int e = check_dev(c)) return e;
    {
        // one launch per phase for all bunches when the batch qualifies
        std::vector<Mesh *> ms(nb);
        for (int b = 0; b < nb; ++b) ms[b] = &meshes[b]->m;
        const int g = launch_deposit_grouped(c, ms.data(), nb, pdtype, x, y, z, q, n, min_bounds,
                                             delta, clear);
        if (g != SC_ERR_UNSUPPORTED) return g;
    }
    return run_batched(c, nb, [&](int b) {
        return launch_deposit(&meshes[b]->m, pdtype, x[b], y[b], z[b], q[b], n[b], min_bounds + 3 * b,
                              delta + 3 * b, clear);
    });
}

int sc_solve_batched(sc_mesh *const *meshes, int nb, const double *min_bounds,
                     const double *max_bounds, const double *delta, const double *gamma,
                     const int *at_cathode)
{
    int rc;
    Ctx *c = batch_ctx(meshes, nb, &rc);
    if (!c) return rc;
    if (!min_bounds || !max_bounds || !delta || !gamma || !at_cathode)
        return set_err(c, SC_ERR_INVALID, "NULL per-bunch array");
    for (int b = 0; b < nb; ++b) {
        if (int e = check_geom(c, min_bounds + 3 * b, delta + 3 * b)) return e;
        if (!std::isfinite(gamma[b]) || !(gamma[b] > 0.0))
            return set_err(c, SC_ERR_INVALID, "gamma must be finite and positive");
        if (at_cathode[b] && !std::isfinite(max_bounds[3 * b + 2]))
            return set_err(c, SC_ERR_INVALID, "max_bounds must be finite");
    }
    if (int e = check_dev(c)) return e;
    int table = SC_ERR_UNSUPPORTED;
    if ((c->grouped & SC_GROUP_SOLVE) && nb > 1) {
        table = ensure_solve_table(c);  // SC_ERR_UNSUPPORTED: first needed under capture
        if (table != SC_OK && table != SC_ERR_UNSUPPORTED) return table;
    }
    if (table == SC_OK) {
        // classes of two or more same-shape bunches solve with one launch per
        // kernel; each class, and each remaining bunch, is one unit of the pool
        std::vector<std::vector<int>> units;
        std::vector<int> single;
        std::vector<bool> done(nb, false);
        for (int b = 0; b < nb; ++b) {
            if (done[b]) continue;
            const Mesh &mb = meshes[b]->m;
            std::vector<int> cls{b};
            done[b] = true;
            // a class's spectra stay within kGroupSolveBytes, so that each
            // grouped pass leaves them in the last-level cache for the next
            const int64_t per_bunch = solve_group_bytes(&mb);
            const int64_t cap = per_bunch > 0 ? kGroupSolveBytes / per_bunch : 0;
            if (solve_groupable(&mb) && cap >= 2)
                for (int o = b + 1; o < nb && (int64_t)cls.size() < cap; ++o) {
                    const Mesh &mo = meshes[o]->m;
                    if (!done[o] && solve_groupable(&mo) && mo.dtype == mb.dtype &&
                        mo.n[0] == mb.n[0] && mo.n[1] == mb.n[1] && mo.n[2] == mb.n[2] &&
                        (at_cathode[o] != 0) == (at_cathode[b] != 0)) {
                        cls.push_back(o);
                        done[o] = true;
                    }
                }
            units.push_back(cls);
        }
        std::vector<int> slot(units.size(), 0);
        int used = 0;
        for (size_t u = 0; u < units.size(); ++u) {
            if (units[u].size() < 2) continue;
            slot[u] = used;
            used += (int)units[u].size();
        }
        if (used <= kGroupMax)
            return run_batched(c, (int)units.size(), [&](int u) {
                const std::vector<int> &cls = units[u];
                if (cls.size() == 1) {
                    const int b = cls[0];
                    return solve_mesh(&meshes[b]->m, min_bounds + 3 * b, max_bounds + 3 * b,
                                      delta + 3 * b, gamma[b], at_cathode[b]);
                }
                const int k = (int)cls.size();
                std::vector<Mesh *> ms(k);
                std::vector<const double *> lo(k), hi(k), dl(k);
                std::vector<double> gm(k);
                for (int i = 0; i < k; ++i) {
                    ms[i] = &meshes[cls[i]]->m;
                    lo[i] = min_bounds + 3 * cls[i];
                    hi[i] = max_bounds + 3 * cls[i];
                    dl[i] = delta + 3 * cls[i];
                    gm[i] = gamma[cls[i]];
                }
                return solve_class(c, ms.data(), k, lo.data(), hi.data(), dl.data(), gm.data(),
                                   at_cathode[cls[0]] != 0, slot[u]);
            });
    }
    return run_batched(c, nb, [&](int b) {
        return solve_mesh(&meshes[b]->m, min_bounds + 3 * b, max_bounds + 3 * b, delta + 3 * b,
                          gamma[b], at_cathode[b]);
    });
}

int sc_interpolate_batched(sc_mesh *const *meshes, int nb, int pdtype, const void *const *x,
                           const void *const *y, const void *const *z, const int64_t *n,
                           const double *min_bounds, const double *delta, void *const *Ex,
                           void *const *Ey, void *const *Ez)
{
    int rc;
    Ctx *c = batch_ctx(meshes, nb, &rc);
    if (!c) return rc;
    if (!dtype_ok(pdtype)) return set_err(c, SC_ERR_INVALID, "dtype must be SC_F32 or SC_F64");
    if (!x || !y || !z || !n || !min_bounds || !delta || !Ex || !Ey || !Ez)
        return set_err(c, SC_ERR_INVALID, "NULL per-bunch array");
    for (int b = 0; b < nb; ++b) {
        if (n[b] < 0) return set_err(c, SC_ERR_INVALID, "negative particle count");
        if (n[b] > 0 && (!x[b] || !y[b] || !z[b] || !Ex[b] || !Ey[b] || !Ez[b]))
            return set_err(c, SC_ERR_INVALID, "NULL particle or output pointer");
        if (int e = check_geom(c, min_bounds + 3 * b, delta + 3 * b)) return e;
    }
    if (int e = check_dev(c)) return e;
    {
        std::vector<Mesh *> ms(nb);
        for (int b = 0; b < nb; ++b) ms[b] = &meshes[b]->m;
        const int g = launch_interp_grouped(c, ms.data(), nb, pdtype, x, y, z, n, min_bounds, delta,
                                            Ex, Ey, Ez);
        if (g != SC_ERR_UNSUPPORTED) return g;
    }
    return run_batched(c, nb, [&](int b) {
        return launch_interp(&meshes[b]->m, pdtype, x[b], y[b], z[b], n[b], min_bounds + 3 * b,
                             delta + 3 * b, Ex[b], Ey[b], Ez[b]);
    });
}

// ---- timers ----
int sc_timing_enable(sc_ctx *ctx, int enable)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (!enable) fold_timings(&ctx->c);
    ctx->c.timing = enable != 0;
    return SC_OK;
}

int sc_timing_reset(sc_ctx *ctx)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    fold_timings(&ctx->c);
    for (int k = 0; k < K_COUNT; ++k) {
        ctx->c.total_ms[k] = 0.0;
        ctx->c.launches[k] = 0;
    }
    return SC_OK;
}

int sc_kernel_count(void) { return K_COUNT; }

const char *sc_kernel_name(int kernel_id)
{
    if (kernel_id < 0 || kernel_id >= K_COUNT) return "";
    return kKernelNames[kernel_id];
}

int sc_timing_read(sc_ctx *ctx, int kernel_id, double *total_ms, int64_t *launches)
{
    if (!ctx) return set_err(nullptr, SC_ERR_INVALID, "ctx is NULL");
    if (kernel_id < 0 || kernel_id >= K_COUNT)
        return set_err(&ctx->c, SC_ERR_INVALID, "kernel id out of range");
    fold_timings(&ctx->c);
    if (total_ms) *total_ms = ctx->c.total_ms[kernel_id];
    if (launches) *launches = ctx->c.launches[kernel_id];
    return SC_OK;
}

}  // extern "C"

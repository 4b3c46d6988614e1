#!/usr/bin/env python3
"""Benchmark of the full space-charge step (deposit + solve + interpolate) on one MI355X.

Workload (the single-bunch GPU configuration): 4,000,000 particles, a 50/50
mix of a Gaussian (sigma = 1, 1, 3 mm) and a uniform ellipsoid (semi-axes 2,
2, 6 mm), 1 nC total in equal macro-charges, 128^3 mesh with bounds from the
particle extrema, gamma = 1, free space.  The mesh is built once, outside the
timing, as in the reference benchmark (benchmark/full_pipeline_benchmark.jl:
25-30); each timed step is deposit! + solve! + interpolate_field with inputs
resident in HBM.

--workload cathode runs the cathode-emission configuration instead: 2,000,000
particles, uniform disk of radius 1 mm with z ~ Exp(mean 0.2 mm), 100 pC,
64x64x128 mesh from the particle extrema, image charge on (solve!(mesh;
at_cathode=true)).

--workload batched runs 32 independent bunches (Np log-uniform in [1e4, 2e6],
meshes from 16^3 to 128^3) through the batched calls and reports, beside it,
the same bunches through a loop of single-bunch calls.

--workload adjoint runs the differentiable configuration: 1,000,000 Gaussian
particles (sigma = 1 mm), 64^3 mesh with fixed bounds (+-8 sigma); it times
one forward step and one vector-Jacobian product of the step (interpolation,
solve and deposit adjoints), and a 100-step drift-kick tracking run through
torch.autograd (segments of 10 steps checkpointed) whose loss, the final rms
emittance, is differentiated with respect to the initial positions and
charges.

--workload fftlib times the field solve of the bunch's mesh against its
FFT-library form (rocFFT through torch.fft on the 2n-padded grid).

Prints one JSON line.  `value` is particles/s of the whole job (one GPU).
The `roofline` object is for the kernel with the largest share of the step
time, measured with the library's HIP-event timers on the launching stream
over a second, instrumented pass of the same steps.  The `cpu_baseline`
object times the CPU oracle (a C restatement of the reference CPU path) on a
bounded sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "spacecharge.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def make_particles(n, dtype, device, seed=42):
    """50/50 Gaussian (1, 1, 3 mm) + uniform ellipsoid (2, 2, 6 mm), generated on the GPU."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    ng = n // 2
    nu = n - ng
    sig = torch.tensor([1e-3, 1e-3, 3e-3], dtype=torch.float64, device=device)
    gauss = torch.randn((ng, 3), generator=g, dtype=torch.float64, device=device) * sig
    semi = torch.tensor([2e-3, 2e-3, 6e-3], dtype=torch.float64, device=device)
    parts = []
    have = 0
    while have < nu:  # rejection sampling in the bounding box
        u = torch.rand((2 * (nu - have) + 1024, 3), generator=g, dtype=torch.float64,
                       device=device) * 2 - 1
        u = u[(u * u).sum(1) <= 1.0]
        parts.append(u)
        have += u.shape[0]
    ell = torch.cat(parts)[:nu] * semi
    p = torch.cat([gauss, ell]).to(dtype)
    x, y, z = (p[:, i].contiguous() for i in range(3))
    q = torch.full((n,), 1e-9 / n, dtype=dtype, device=device)
    return x, y, z, q


def make_cathode_particles(n, dtype, device, seed=7):
    """Cathode emission bunch: r = 1 mm sqrt(U) (uniform disk), z ~ Exp(mean 0.2 mm),
    100 pC total in equal macro-charges, generated on the GPU."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    u = torch.rand((n, 3), generator=g, dtype=torch.float64, device=device)
    r = 1e-3 * torch.sqrt(u[:, 0])
    th = 2 * torch.pi * u[:, 1]
    x = (r * torch.cos(th)).to(dtype).contiguous()
    y = (r * torch.sin(th)).to(dtype).contiguous()
    z = (-0.2e-3 * torch.log1p(-u[:, 2])).to(dtype).contiguous()
    q = torch.full((n,), 100e-12 / n, dtype=dtype, device=device)
    return x, y, z, q


def make_batch(dtype, device, nb=32, seed=11):
    """Many independent bunches: Np = floor(10^U(4, log10 2e6)) Gaussians with
    sigma ~ U(0.5, 3) mm per axis, each on a mesh drawn from 16^3, 32^3,
    64x64x128, 128^3 (seeded)."""
    import math
    import torch
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    grids = [(16, 16, 16), (32, 32, 32), (64, 64, 128), (128, 128, 128)]
    out = []
    for b in range(nb):
        u = torch.rand(5, generator=g, dtype=torch.float64)
        n = int(10 ** (4 + float(u[0]) * (math.log10(2e6) - 4)))
        sig = (0.5e-3 + 2.5e-3 * u[1:4]).to(device)
        grid = grids[int(float(u[4]) * 4) % 4]
        gd = torch.Generator(device=device)
        gd.manual_seed(seed * 1000 + b)
        p = (torch.randn((n, 3), generator=gd, dtype=torch.float64, device=device) * sig).to(dtype)
        x, y, z = (p[:, i].contiguous() for i in range(3))
        q = torch.full((n,), 1e-10 / n, dtype=dtype, device=device)
        out.append((grid, (x, y, z, q)))
    return out


def bench_batched(args, sc, torch, dev, T):
    """Batched step over many bunches against the loop of single-bunch calls."""
    bunches = make_batch(T, dev)
    ntot = sum(p[0].numel() for _, p in bunches)
    meshes = [sc.Mesh3D(grid, p[0], p[1], p[2], T=T, gamma=args.gamma) for grid, p in bunches]
    loop_meshes = [sc.Mesh3D(grid, p[0], p[1], p[2], T=T, gamma=args.gamma) for grid, p in bunches]
    parts = [p for _, p in bunches]

    def step_batched():
        sc.deposit_batched(meshes, *[[p[k] for p in parts] for k in range(4)])
        sc.solve_batched(meshes)
        return sc.interpolate_batched(meshes, *[[p[k] for p in parts] for k in range(3)])

    def step_loop():
        for m, p in zip(loop_meshes, parts):
            sc.deposit(m, *p)
            sc.solve(m)
            sc.interpolate_field(m, p[0], p[1], p[2])

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    if args.graph_only:
        # the default (grouped) step captured once and replayed: nothing else
        # runs, so a profile of this command is one kind of step
        for _ in range(2):
            step_batched()
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            step_batched()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step_batched()
        ms_graph = timed(graph.replay)
        return {
            "metric": "particles/s per full space-charge step (deposit+solve+interpolate), many bunches",
            "value": ntot / (ms_graph * 1e-3), "unit": "particles/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_graph,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
            "data": "synthetic (seeded Gaussian bunches generated on the GPU)",
            "config": {"workload": (f"{len(bunches)} bunches, {ntot} particles in total, graph replay "
                                    "of the grouped batched step"),
                       "particles": ntot, "bunches": len(bunches), "parallelism": "single GPU"},
        }
    ms = timed(step_batched)
    ms_loop = timed(step_loop)
    # the same batched step captured once into a graph and replayed (the
    # workspace, bins and scratch exist after the warmup above)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step_batched()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step_batched()
    ms_graph = timed(graph.replay)

    # deposit tilings picked per bunch by the autotuner, then the same
    # measurements (the graph is captured again: launch shapes changed)
    sc.autotune_batched(meshes, *[[p[k] for p in parts] for k in range(4)], reps=3)
    tilings = [list(m.deposit_tiling) for m in meshes]
    ms_tuned = timed(step_batched)
    with torch.cuda.stream(side):
        step_batched()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph_t = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph_t):
        step_batched()
    ms_graph_tuned = timed(graph_t.replay)

    # the same captured step with fewer grouped launches (same process):
    # particle phases only, then one launch sequence per bunch throughout
    by_mode = {}
    for name, particles, solve in (("particles_only", True, False), ("none", False, False)):
        sc.set_grouped_launch(particles, solve)
        with torch.cuda.stream(side):
            step_batched()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g_m = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g_m):
            step_batched()
        by_mode[name] = {"eager_ms_per_step": timed(step_batched), "graph_ms_per_step": timed(g_m.replay)}
    sc.set_grouped_launch(True, True)
    return {
        "metric": "particles/s per full space-charge step (deposit+solve+interpolate), many bunches",
        "value": ntot / (ms * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded Gaussian bunches generated on the GPU)",
        "config": {
            "workload": (f"{len(bunches)} bunches, {ntot} particles in total, Np log-uniform in "
                         "[1e4, 2e6], meshes from 16^3, 32^3, 64x64x128, 128^3, free space"),
            "particles": ntot,
            "bunches": len(bunches),
            "grids": [list(g) for g, _ in bunches],
            "gamma": args.gamma,
            "parallelism": "single GPU, bunches on a pool of streams",
        },
        "sequential_ms_per_step": ms_loop,
        "batched_speedup": ms_loop / ms,
        "graph_ms_per_step": ms_graph,
        "graph_particles_per_s": ntot / (ms_graph * 1e-3),
        "autotuned_ms_per_step": ms_tuned,
        "autotuned_graph_ms_per_step": ms_graph_tuned,
        "autotuned_graph_particles_per_s": ntot / (ms_graph_tuned * 1e-3),
        "deposit_tilings": tilings,
        "grouped_launches": "particle phases and same-shape small-mesh solves (default)",
        "other_grouping": by_mode,
    }


def bench_fftlib(args, sc, torch, dev, T):
    """The field solve of the headline mesh against its FFT-library form on
    the same GPU: torch.fft (rocFFT on ROCm) on the 2n-padded grid, one
    rfftn of rho, three spectral products and three irfftn (plus three rfftn
    when the Green spectra are transformed too), against sc.solve with the
    Green spectra rebuilt (as solve! does) and reused.  Both legs in the
    requested precision; the FFT-library legs use random data of the right
    shape (their cost does not depend on the values)."""
    n = args.grid
    L = 2 * n
    x, y, z, q = make_particles(args.particles, T, dev)
    mesh = sc.Mesh3D((n, n, n), x, y, z, T=T, gamma=args.gamma, total_charge=1e-9)
    sc.deposit(mesh, x, y, z, q)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    ms_rebuild = timed(lambda: sc.solve(mesh))
    mesh.set_green_reuse(True)
    ms_reuse = timed(lambda: sc.solve(mesh))
    mesh.set_green_reuse(False)
    rho = torch.zeros((L, L, L), dtype=T, device=dev)
    rho[:n, :n, :n] = torch.rand((n, n, n), dtype=T, device=dev)
    G = [torch.rand((L, L, L), dtype=T, device=dev) for _ in range(3)]
    Gh = [torch.fft.rfftn(g) for g in G]

    def lib_solve(green):
        spectra = [torch.fft.rfftn(g) for g in G] if green else Gh
        R = torch.fft.rfftn(rho)
        return [torch.fft.irfftn(R * g, s=(L, L, L))[:n, :n, :n] for g in spectra]

    ms_fft = timed(lambda: lib_solve(False))
    ms_fft_green = timed(lambda: lib_solve(True))
    return {
        "metric": "ms per field solve, library against an FFT-library (rocFFT via torch.fft) formulation",
        "value": ms_rebuild, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_rebuild, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": args.dtype, "data": "synthetic (seeded bunch on the GPU; random arrays for the FFT legs)",
        "config": {"workload": f"field solve of a {n}^3 mesh (transforms on {L}^3)", "grid": [n, n, n],
                   "parallelism": "single GPU"},
        "library_solve_ms_green_rebuilt": ms_rebuild,
        "library_solve_ms_green_reused": ms_reuse,
        "fftlib_solve_ms": ms_fft,
        "fftlib_solve_ms_with_green_transforms": ms_fft_green,
        "fftlib_rfftn_ms": timed(lambda: torch.fft.rfftn(rho)),
        "fftlib_irfftn_ms": timed(lambda: torch.fft.irfftn(Gh[0], s=(L, L, L))),
    }


def bench_adjoint(args, sc, torch, dev, T):
    """Forward step, its VJP, and a differentiated drift-kick tracking run."""
    import spacecharge.autograd as scad
    n = args.particles if args.particles != 4_000_000 else 1_000_000
    g = args.grid if args.grid != 128 else 64
    grid = (g, g, g)
    sig = 1e-3
    gen = torch.Generator(device=dev).manual_seed(17)
    p = torch.randn((n, 3), generator=gen, dtype=torch.float64, device=dev) * sig
    x, y, z = (p[:, i].contiguous().to(T) for i in range(3))
    q = torch.full((n,), 1e-9 / n, dtype=T, device=dev)
    mesh = sc.Mesh3D(grid, (-8 * sig,) * 3, (8 * sig,) * 3, T=T, gamma=args.gamma)
    gE = [torch.randn(n, generator=gen, dtype=torch.float64, device=dev).to(T) for _ in range(3)]

    def fwd():
        sc.deposit(mesh, x, y, z, q)
        sc.solve(mesh)
        return sc.interpolate_field(mesh, x, y, z)

    def vjp():
        gef, gx, gy, gz = sc.interpolate_adjoint(mesh, x, y, z, *gE)
        g_rho = sc.solve_adjoint(mesh, gef)
        return sc.deposit_adjoint(mesh, x, y, z, q, g_rho)

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms_fwd = timed(fwd, args.steps)
    fwd()
    ms_vjp = timed(vjp, args.steps)

    # drift-kick tracking (normalised units: p += kick E, x += dt p) through
    # the autograd layer, segments of 10 steps checkpointed, loss = final rms
    # emittance, gradients with respect to the initial positions and charges
    nsteps, every = 100, 10
    dt, kick = 1.0, 1e-3 * sig / 1e7  # a kick of ~1e-3 sigma per step at 1e7 V/m

    def track(grad, ckpt, recompute="trajectory"):
        xs, ys, zs, qs = (t.clone().requires_grad_(grad) for t in (x, y, z, q))
        zero = torch.zeros_like(xs)
        with torch.set_grad_enabled(grad):
            st = scad.track(mesh, xs, ys, zs, zero, zero, zero, qs, nsteps, dt, kick,
                            checkpoint_every=ckpt, recompute=recompute)
            loss = scad.rms_emittance(*st)
        if grad:
            loss.backward()
            return xs.grad, qs.grad
        return None, None

    def timed_track(grad, ckpt, recompute="trajectory", reps=3):
        # one untimed run, then the median of `reps` timed runs (the loop is
        # issued from Python, so single runs vary with the host)
        out = track(grad, ckpt, recompute)
        times = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = track(grad, ckpt, recompute)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        return sorted(times)[len(times) // 2], out

    ms_track_fwd, _ = timed_track(False, every)
    ms_track, (gx, gq) = timed_track(True, every)
    ms_track_all, _ = timed_track(True, 0)
    ms_track_steps, _ = timed_track(True, every, "step")
    return {
        "metric": "particles/s per differentiable space-charge step (forward + vector-Jacobian product)",
        "value": n / ((ms_fwd + ms_vjp) * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_fwd + ms_vjp,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded Gaussian bunch generated on the GPU)",
        "config": {
            "workload": (f"differentiable step, {n} particles, {grid[0]}^3 mesh with fixed bounds, "
                         f"{nsteps}-step drift-kick tracking through torch.autograd"),
            "particles": n,
            "grid": list(grid),
            "gamma": args.gamma,
            "parallelism": "single GPU",
        },
        "forward_ms": ms_fwd,
        "vjp_ms": ms_vjp,
        "vjp_over_forward": ms_vjp / ms_fwd,
        "tracking": {
            "steps": nsteps,
            "checkpoint_every": every,
            "loss": "final rms emittance (sum over planes)",
            "recompute": "particle trajectory from the stored per-step mesh fields",
            "forward_only_ms": ms_track_fwd,
            "forward_and_backward_ms": ms_track,
            "grad_over_forward": ms_track / ms_track_fwd,
            "forward_and_backward_ms_whole_steps_recomputed": ms_track_steps,
            "grad_over_forward_whole_steps_recomputed": ms_track_steps / ms_track_fwd,
            "forward_and_backward_ms_every_step_stored": ms_track_all,
            "grad_over_forward_every_step_stored": ms_track_all / ms_track_fwd,
            "grad_finite": bool(torch.isfinite(gx).all() and torch.isfinite(gq).all()),
        },
    }


def kernel_bytes(name, n, grid, s, ps):
    """Algorithmic HBM bytes of one launch of a step kernel (compulsory I/O):
    s = mesh type size, ps = particle type size."""
    nx, ny, nz = grid
    M = nx * ny * nz
    kx = nx + 1
    spec = nz * 2 * ny * kx * 2 * s  # one complex spectrum [nz][2ny][nx+1]
    if name == "deposit":
        return 4 * n * ps + M * s
    if name == "bin_count":
        return 3 * n * ps
    if name == "bin_scatter":
        return 8 * n * ps
    tiles = 1
    for g_ in grid:
        tiles *= (g_ - 1 + 15) // 16
    if name == "bin_rowscan":
        slices = min((n + 4095) // 4096, 2048)
        return 2 * 4 * tiles * slices
    if name == "deposit_tiles":
        return 4 * n * ps
    if name == "gather_tiles":
        return tiles * 17 ** 3 * 8 + M * s
    if name == "interleave":
        # reads the three components, writes 4 values per node, or 8 (node
        # pairs) for Float64 fields of 64 MiB and more, as the library picks
        pair = 2 if (s == 8 and M * 32 >= (64 << 20)) else 1
        return 3 * M * s + 4 * pair * M * s
    if name == "interpolate":
        # positions in, fields out, and the field once (6 Np s + 3 M s)
        return 6 * n * ps + 3 * M * s
    if name == "rho_fwd_x":
        return M * s + ny * nz * kx * 2 * s
    if name == "rho_fwd_y":
        return ny * nz * kx * 2 * s + spec
    if name == "zconv":
        return spec + 3 * (nz + 1) * (ny + 1) * kx * s + 3 * spec
    if name == "inv_y":
        return 3 * spec + 3 * ny * nz * kx * 2 * s
    if name == "inv_x":
        return 3 * ny * nz * kx * 2 * s + 3 * M * s
    return None


def measured_traffic(name, n, grid, dtype_name, cathode):
    """HBM bytes per launch of kernel `name` from a committed counter summary.

    The summaries under profiles/hbm_traffic_*.md carry a fenced ```json block
    with the workload they were measured on and, per kernel, FETCH_SIZE +
    WRITE_SIZE per dispatch scaled by the copy-kernel yardstick of the same
    runs.  Returns (bytes, file) for a summary of this exact workload, else
    (None, None).
    """
    import glob
    import re
    want = {"particles": n, "grid": list(grid), "dtype": dtype_name, "at_cathode": bool(cathode)}
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "hbm_traffic_*.md"))):
        with open(path) as f:
            text = f.read()
        for block in re.findall(r"```json\n(.*?)\n```", text, re.S):
            try:
                d = json.loads(block)
            except ValueError:
                continue
            if d.get("workload") == want and name in d.get("bytes_per_launch", {}):
                return d["bytes_per_launch"][name], os.path.relpath(path, REPO)
    return None, None


def _rel(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def accuracy(args, parts, grid, mesh, E_gpu, dtype_name):
    """The GPU step's field against the oracle on the same particles and the
    GPU mesh's own bounds, per component, at the mesh nodes and at the
    particles.  Float32 legs: against the Float64 oracle.  Float64 legs:
    against the same algorithm in extended precision (the Float64 oracle is
    itself only accurate to ~1e-11 at 128^3), with the Float64 oracle's own
    distance to it beside it."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    threads = os.cpu_count() or 1
    x, y, z, q = (t.cpu().numpy().astype(np.float64) for t in parts)
    lo, hi, dl = mesh.min_bounds, mesh.max_bounds, mesh.delta
    rho = orc.deposit(x, y, z, q, grid, lo, dl, np.float64)
    ef = orc.solve(rho, lo, hi, dl, float(mesh.gamma), bool(args.cathode), nthreads=threads)
    E_ref = orc.interpolate(ef, x, y, z, lo, dl, nthreads=threads)
    ef_gpu = mesh._ef.cpu().numpy().astype(np.float64).reshape(tuple(grid) + (3,), order="F")
    E_g = [e.cpu().numpy().astype(np.float64) for e in E_gpu]
    if dtype_name == "f32":
        return {
            "reference": "oracle Float64 step on the same particles and the GPU mesh's bounds",
            "mesh_field_rel_l2": [_rel(ef_gpu[..., c], ef[..., c]) for c in range(3)],
            "particle_field_rel_l2": [_rel(E_g[c], E_ref[c]) for c in range(3)],
        }
    tr = orc.solve_extended(rho, lo, hi, dl, float(mesh.gamma), bool(args.cathode),
                            nthreads=threads)
    E_tr = orc.interpolate(np.asfortranarray(tr), x, y, z, lo, dl, nthreads=threads)
    return {
        "reference": ("the oracle's algorithm in x87 extended precision on the same particles "
                      "and the GPU mesh's bounds"),
        "mesh_field_rel_l2": [_rel(ef_gpu[..., c], tr[..., c]) for c in range(3)],
        "particle_field_rel_l2": [_rel(E_g[c], E_tr[c]) for c in range(3)],
        "float64_oracle_mesh_field_rel_l2": [_rel(ef[..., c], tr[..., c]) for c in range(3)],
        "float64_oracle_particle_field_rel_l2": [_rel(E_ref[c], E_tr[c]) for c in range(3)],
    }


def cpu_baseline(args, parts, grid, dtype_name, max_seconds):
    """Time the CPU oracle on a bounded sample: whole steps of the same
    workload (the same particles, copied to the host), with the mesh's arrays
    and the solve workspace (padded arrays and transform plans) made once,
    as the reference keeps them on the mesh (src/mesh.jl:49-72); one untimed
    warm step, then at least 10 timed steps or max_seconds.  Deposit serial
    in particle order (src/deposition.jl:237-240); transforms, Green lattice
    and interpolation on all host cores.  Also times one padded 3-D transform
    of the oracle against scipy.fft at the same thread count, so the
    baseline's transforms are shown not to be a strawman."""
    import numpy as np
    import scipy.fft
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    npdt = np.float32 if dtype_name == "f32" else np.float64
    x, y, z, q = (np.ascontiguousarray(t.cpu().numpy()) for t in parts)
    n = x.shape[0]
    threads = os.cpu_count() or 1
    lo, hi, dl = orc.bounds_auto(x, y, z, grid, npdt)
    ws = orc.Workspace(grid, npdt)
    rho = np.zeros(tuple(grid), dtype=npdt, order="F")
    ef = np.zeros(tuple(grid) + (3,), dtype=npdt, order="F")

    def step():
        rho.fill(0)  # clear_mesh!
        orc.deposit(x, y, z, q, grid, lo, dl, npdt, rho=rho)
        ws.solve(rho, lo, hi, dl, args.gamma, bool(args.cathode), nthreads=threads, out=ef)
        orc.interpolate(ef, x, y, z, lo, dl, nthreads=threads)

    step()  # first solve!: warm caches, as after the reference's first call
    times = []
    t_start = time.perf_counter()
    while len(times) < 10 or time.perf_counter() - t_start < max_seconds:
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > 3 * max_seconds:
            break
    best = min(times)
    med = sorted(times)[len(times) // 2]
    # one padded transform: the oracle's (plans made per call) against scipy's pocketfft
    cdt = np.complex64 if npdt == np.float32 else np.complex128
    shape = tuple(2 * g for g in grid)
    rng = np.random.default_rng(0)
    a0 = np.asfortranarray((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(cdt))
    t_orc, t_sci = [], []
    for _ in range(3):
        a = a0.copy(order="F")
        t0 = time.perf_counter()
        orc.fft3d(a, -1, nthreads=threads)
        t_orc.append(time.perf_counter() - t0)
        b = a0.copy(order="F")
        t0 = time.perf_counter()
        b = scipy.fft.fftn(b, overwrite_x=True, workers=threads)
        t_sci.append(time.perf_counter() - t0)
    fft_err = _rel(a, b)
    return {
        "value": n / best,
        "unit": "particles/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{len(times)} full steps after one warm step ({dtype_name}, {n} particles, "
                   f"{grid[0]}x{grid[1]}x{grid[2]}, gamma={args.gamma}, "
                   f"{'cathode' if args.cathode else 'free space'}), mesh arrays and solve "
                   "workspace (padded arrays, transform plans) kept across steps; "
                   f"min {best * 1e3:.1f} ms, median {med * 1e3:.1f} ms; deposit single-threaded, "
                   f"transforms, Green lattice and interpolation on {threads} threads"),
        "ms_per_step_min": best * 1e3,
        "ms_per_step_median": med * 1e3,
        "fft_check": {
            "shape": list(shape),
            "dtype": str(np.dtype(cdt)),
            "threads": threads,
            "oracle_ms": min(t_orc) * 1e3,
            "scipy_fft_ms": min(t_sci) * 1e3,
            "oracle_over_scipy": min(t_orc) / min(t_sci),
            "rel_l2_oracle_vs_scipy": fft_err,
        },
    }


def run_leg(args, sc, torch, dev, T):
    """One precision leg of the single-bunch workload: builds the particles
    and the mesh (outside the timing), times exactly args.steps steps after
    args.warmup, then repeats them with the library's per-kernel HIP-event
    timers on the launching stream for the dominant kernel's roofline."""
    if args.workload == "cathode":
        n = args.particles if args.particles != 4_000_000 else 2_000_000
        grid = (64, 64, 128) if args.grid == 128 else (args.grid, args.grid, 2 * args.grid)
        x, y, z, q = make_cathode_particles(n, T, dev)
        total = 100e-12
    else:
        n = args.particles
        grid = (args.grid, args.grid, args.grid)
        x, y, z, q = make_particles(n, T, dev)
        total = 1e-9
    mesh = sc.Mesh3D(grid, x, y, z, T=T, gamma=args.gamma, total_charge=total)
    if args.sorted:
        perm = sc.sort_permutation(mesh, x, y, z)
        x, y, z, q = (torch.index_select(t, 0, perm) for t in (x, y, z, q))
    tuning = None
    if args.autotune:
        times = sc.autotune_deposit(mesh, x, y, z, q, reps=5)
        tuning = {"tiling": list(mesh.deposit_tiling), "deposit_ms": times[mesh.deposit_tiling],
                  "default_tiling_deposit_ms": times[(13, 12)]}

    def step():
        sc.deposit(mesh, x, y, z, q)
        sc.solve(mesh, at_cathode=args.cathode)
        return sc.interpolate_field(mesh, x, y, z)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: exactly `steps` steps, synchronised on both sides
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1e3 / args.steps

    # the same step captured once into a HIP graph and replayed: the same
    # launches and work (the Green build still forks onto the library's
    # second stream), without the host's launch path between the kernels
    E_eager = step()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        E_graph = step()
    graph.replay()
    torch.cuda.synchronize()
    # bitwise equal on Float32 meshes; Float64 deposits add in LDS atomics
    # whose order varies, so there the last bits may differ
    graph_diff = max(float((a.double() - b.double()).abs().max() / b.double().abs().max())
                     for a, b in zip(E_graph, E_eager))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - t2) * 1e3 / args.steps
    del graph, E_graph, E_eager

    # instrumented pass: per-kernel HIP events on the launching stream
    with sc.KernelTimer(dev) as kt:
        t1 = time.perf_counter()
        for _ in range(args.steps):
            E_last = step()
        torch.cuda.synchronize()
        ms_instr = (time.perf_counter() - t1) * 1e3 / args.steps
    per_kernel = kt.results()
    s = 4 if T == torch.float32 else 8
    dtype_name = "f32" if T == torch.float32 else "f64"
    tot = sum(v[0] for v in per_kernel.values())
    dom_name, (dom_ms, dom_n) = max(per_kernel.items(), key=lambda kv: kv[1][0])
    avg_s = dom_ms / dom_n * 1e-3
    nbytes = kernel_bytes(dom_name, n, grid, s, s)
    traffic, traffic_src = measured_traffic(dom_name, n, grid, dtype_name, args.cathode)
    roofline = {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": (nbytes / avg_s / 1e9) if nbytes else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (nbytes / avg_s / 1e9 / HBM_PEAK_GBS) if nbytes else None,
        "traffic": traffic,
        "traffic_source": traffic_src,
        "algorithmic_bytes": nbytes,
        "avg_launch_us": avg_s * 1e6,
    }
    breakdown = {k: round(v[0] / v[1] * 1e3, 2) for k, v in sorted(per_kernel.items(),
                                                                      key=lambda kv: -kv[1][0])}
    # whole step: compulsory I/O of the reference decomposition, 10*Np*s + 168*M*s
    # (deposit 4Np+M, interpolate 6Np+3M, solve 4M + 20P with P = 8M), cathode 339*M*s
    M = grid[0] * grid[1] * grid[2]
    b_step = 10 * n * s + (339 if args.cathode else 168) * M * s
    step_roofline = {
        "algorithmic_bytes": b_step,
        "achieved": b_step / (ms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": b_step / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    return {
        "n": n, "grid": grid, "dtype": dtype_name, "mesh": mesh, "parts": (x, y, z, q),
        "E_last": E_last, "ms": ms, "ms_instr": ms_instr, "ms_graph": ms_graph,
        "graph_diff": graph_diff, "roofline": roofline,
        "step_roofline": step_roofline, "breakdown": breakdown,
        "share": {k: round(v[0] / tot, 3) for k, v in per_kernel.items()}, "tuning": tuning,
    }


def leg_cpu(args, leg, budget):
    """Accuracy and CPU baseline of a leg (None when the CPU sample is off)."""
    if args.cpu_seconds <= 0:
        return None, None
    acc = accuracy(args, leg["parts"], leg["grid"], leg["mesh"], leg["E_last"], leg["dtype"])
    cpu = cpu_baseline(args, leg["parts"], leg["grid"], leg["dtype"], budget)
    return acc, cpu


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    # the GPU reaches its steady-state step time only after some tens of
    # steps (4M particles, 128^3: 0.412-0.415 ms over 20 steps after 5 of
    # warmup, 0.394-0.395 ms over 300 after 50, on one box); the timed work
    # stays well under a second
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--gamma", type=float, default=1.0)
    ap.add_argument("--cathode", action="store_true")
    ap.add_argument("--workload", choices=["bunch", "cathode", "batched", "adjoint", "fftlib"],
                    default="bunch",
                    help="bunch: Gaussian + uniform ellipsoid (defaults 4M particles, 128^3); "
                         "cathode: uniform disk with exponential z emitted from a cathode, "
                         "2M particles, 64x64x128, image charge on; batched: 32 bunches of "
                         "uneven size and mesh through the batched calls; adjoint: forward "
                         "step, its VJP and a differentiated 100-step tracking run, 1M "
                         "particles, 64^3; fftlib: the field solve of the bunch's mesh against "
                         "its rocFFT (torch.fft) formulation")
    ap.add_argument("--sorted", action="store_true",
                    help="put the particles in the deposit's tile order once before the warmup "
                         "(sort_permutation), as a tracking code does every few steps")
    ap.add_argument("--graph-only", action="store_true",
                    help="batched workload: time only the graph replay of the default step")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the deposit tiling by measurement (autotune_deposit) before "
                         "the warmup; the batched workload always reports both")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="time budget of each leg's CPU-oracle sample (0 disables the CPU "
                         "sample and the accuracy check)")
    ap.add_argument("--no-other-precision", action="store_true",
                    help="single-bunch workloads: skip the second leg in the other precision")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="launch a plain copy kernel (read N MiB, write N MiB) three times "
                         "before the warmup, as the yardstick of HBM counter runs")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not an implementation.", file=sys.stderr)
        sys.exit(2)

    import torch
    import spacecharge as sc

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    T = torch.float32 if args.dtype == "f32" else torch.float64
    if args.workload == "batched":
        print(json.dumps(bench_batched(args, sc, torch, dev, T)))
        return
    if args.workload == "adjoint":
        print(json.dumps(bench_adjoint(args, sc, torch, dev, T)))
        return
    if args.workload == "fftlib":
        print(json.dumps(bench_fftlib(args, sc, torch, dev, T)))
        return
    if args.workload == "cathode":
        args.cathode = True

    if args.calibration_copy_mib > 0:
        # elementwise negation: one read and one write of every byte, no reuse
        src = torch.ones(args.calibration_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    head = run_leg(args, sc, torch, dev, T)
    other = None
    if not args.no_other_precision:
        other = run_leg(args, sc, torch, dev, torch.float64 if T == torch.float32 else torch.float32)
    acc, cpu = leg_cpu(args, head, args.cpu_seconds)
    n, grid, ms = head["n"], head["grid"], head["ms"]
    if cpu is not None:
        cpu["accuracy"] = acc
    other_out = None
    if other is not None:
        o_acc, o_cpu = leg_cpu(args, other, args.cpu_seconds)
        other_out = {
            "dtype": other["dtype"],
            "value": other["n"] / (other["ms"] * 1e-3),
            "unit": "particles/s",
            "ms_per_step": other["ms"],
            "roofline": other["roofline"],
            "step_roofline": other["step_roofline"],
            "accuracy": o_acc,
            "cpu_baseline": o_cpu,
            "kernel_us_per_launch": other["breakdown"],
            "ms_per_step_instrumented": other["ms_instr"],
            "graph_ms_per_step": other["ms_graph"],
            "graph_max_rel_diff_vs_eager": other["graph_diff"],
        }

    out = {
        "metric": ("particles/s and ms per full space-charge step (4M particles, 128\u00b3), "
                   "FP32 vs FP64" if args.workload == "bunch" else
                   "particles/s and ms per full space-charge step, FP32 vs FP64"),
        "value": n / (ms * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": ("synthetic (seeded uniform-disk, exponential-z cathode bunch generated on the GPU)"
                 if args.workload == "cathode" else
                 "synthetic (seeded Gaussian + uniform-ellipsoid bunch generated on the GPU)"),
        "config": {
            "workload": (f"single bunch, {n} particles, {grid[0]}x{grid[1]}x{grid[2]} mesh, "
                         f"gamma={args.gamma}, {'cathode' if args.cathode else 'free space'}"),
            "particles": n,
            "grid": list(grid),
            "gamma": args.gamma,
            "at_cathode": bool(args.cathode),
            "particle_order": "tile order (sort_permutation)" if args.sorted else "as generated",
            "deposit_tiling": list(head["mesh"].deposit_tiling),
            "parallelism": "single GPU",
        },
        "autotune": head["tuning"],
        "roofline": head["roofline"],
        "step_roofline": head["step_roofline"],
        "cpu_baseline": cpu,
        "other_precision": other_out,
        "kernel_us_per_launch": head["breakdown"],
        "kernel_share": head["share"],
        "ms_per_step_instrumented": head["ms_instr"],
        "graph_ms_per_step": head["ms_graph"],
        "graph_particles_per_s": n / (head["ms_graph"] * 1e-3),
        "graph_max_rel_diff_vs_eager": head["graph_diff"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

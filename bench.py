#!/usr/bin/env python3
"""Benchmark of the SPH particle hot path on one MI355X.

Workload: the two-fluid large ball, 1,000,000 particles (the lattice at
dx = 0.0127 m filling a radius-1 m spherical volume-map boundary from the
bottom to just below the equator, split at x = 0 into rest densities 1000 and
1500), IISPHwOST + TakePredicted, CFL dt (lambda 0.4, max 1e-3), target
average density error 0.1 %. The fluid rests on the wall from the first step.

A step is one SPHSystem::step_forward_in_time on the device-resident state.
--settle untimed steps (default 60) let the fluid compress until every solve
carries its weight (about 22 Jacobi iterations per step, EQS1 + EQS2) and
let the code-shape tuner finish; W warm-up steps follow, then exactly K steps
are timed between device synchronizations. value = particles x K / wall time
(particle-steps/s).

The JSON line also carries
  roofline      the dominant kernel's compulsory HBM bytes per launch (single
                precision fields, F32_BYTES) divided by its mean launch
                duration (HIP events on the library's own stream, over the
                timed steps), against 8 TB/s; the same with this build's
                binary64 fields beside it;
  step_roofline the whole update's compulsory bytes per particle-step at the
                measured iteration count, times the particle-step rate;
  roofline_fp64 the dominant kernel's algorithmic binary64 operations per
                launch (FLOPS_PER_PAIR x neighbor pairs) over its mean launch
                duration, against the 78.6 TFLOP/s FP64 peak;
  cpu_baseline  the CPU oracle (binary64 restatement of the reference) timed
                on this host over a bounded sample of the same workload,
                started from the GPU's state at the timed window's start, on
                all host threads, and one step of it on one thread beside.

--workload serving runs the serving stream instead: seeded Poisson arrivals
of IISPH scene requests of mixed sizes served by continuous batching (one
world, HBM-budget admission) and, as the baseline, one request at a time with
the plain per-scene path; it reports throughput and p50/p99 step latency of
both.

--workload sensor_plane times sensor queries on the same 1M-particle state
instead: the six ScalarQuantity fields (one per step, in turn) on a
1024 x 1024 plane through the ball's center (get_quantity_at_positions).

--workload volume_map times the construction of the same workload's boundary
field instead: the volume map of the radius-1 m sphere container (h = 0.02,
cell 0.04) integrated from its SDF lattice with the degree-30 Gauss-Legendre
ball rule, in ball quadratures per second; a step is one whole build.

--workload mesh_sdf times the signed-distance lattice of that container built
from a triangle mesh (a geodesic sphere, 5,120 triangles), in lattice nodes
per second; a step is one whole build.

--workload contact times rigid-rigid contact: six dynamic cubes pressed into
a static slab, steps with contact and with it turned off.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-simulation_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic binary64 operations per neighbor pair (+, -, x, / and sqrt one
# each; maxima and selects not counted). grad W: r 3, |r|^2 5, sqrt 1, |r|/h 1,
# w'(q) 7, three components (gfac r_k / |r|) w' 9 = 26. Jacobi adds
# (a_i - a_j) . grad W 8 and dt^2 V_j and the sum 3; pressure acceleration adds
# V_j (p_i + p_j) c_i 3 and the three component updates 6.
FLOPS_PER_PAIR = {"jacobi": 37, "pressure_accel": 35}

# Algorithmic (compulsory) HBM bytes per particle per launch, binary64 SoA:
# the particle's own fields read once and its outputs written once; neighbor
# data is gathered from cache after the cell sort and the neighbor-index list
# is not counted. `per_body` is added per volume-map boundary (the particle's
# boundary-sample slot).
ALGO_BYTES = {
    "neighbor_count": {"base": 36, "per_body": 33},   # one sweep: x,y,z r; count, V w
    "pairs_volume": {"base": 32, "per_body": 33},     # x,y,z r; V w
    "boundary_sample": {"base": 81, "per_body": 0},   # x,y,z r; flag + 7 f64 w
    "nonpressure": {"base": 88, "per_body": 33},      # x,v,m,V r; a w
    "source_diag": {"base": 88, "per_body": 57},      # x,v*,m,V r; s,a_ii,p w
    "pressure_accel": {"base": 72, "per_body": 33},   # x,m,V,p r; a w
    "jacobi": {"base": 88, "per_body": 33},           # x,V,a_p,s,a_ii r; p r/w
}


# Compulsory HBM bytes per particle per launch with single-precision SoA
# fields (each field read once, outputs written once; neighbor data from cache
# after the cell sort; neighbor indices not counted). This is the roofline
# count; the binary64 count above (this build's storage) is reported beside it.
F32_BYTES = {
    "neighbor_count": 32,   # volume pass
    "pairs_volume": 32,
    "boundary_sample": 28,
    "nonpressure": 56,      # viscosity (gravity fused)
    "source_diag": 72,      # predicted velocity + source + diagonal + init
    "pressure_accel": 52,   # per Jacobi iteration
    "jacobi": 60,           # Ap + update + error
}


def step_bytes_f32(solver, iterations):
    """Compulsory bytes per particle-step of the whole update (single
    precision, static boundary): 384 fixed + 112 per Jacobi iteration, plus
    112 for IISPHwOST's second pressure acceleration and predicted x/v."""
    extra = 112 if solver == "IISPHwOST" else 0
    return 384 + 112 * iterations + extra


# Algorithmic binary64 operations per SDF evaluation in the volume-map ball
# quadrature (+, -, *, /, sqrt each one; common factors counted once): the
# serendipity interpolation 247 (cell and local coordinates 29, shape-function
# factors 26, 32 shape products 128, weighted sum 64), the cubic extension 13,
# the quadrature point 7, the weighted phi sum 2. degree^3 evaluations per node.
VM_FLOPS_PER_EVAL = 269
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(workload, particles):
    from sph_amd import scenes
    if workload == "two_fluid_ball":
        return scenes.two_fluid_ball(n=particles, solver="IISPHwOST")
    if workload == "two_fluid_ball_iisph":
        return scenes.two_fluid_ball(n=particles, solver="IISPH")
    if workload == "block_drop":
        return scenes.block_drop_iisph()
    if workload == "rigid_coupling":
        return scenes.rigid_coupling()
    if workload == "dam_break":
        return scenes.dam_break_sesph(side=20)
    raise SystemExit(f"unknown workload {workload}")


def traffic_from_profiles(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(scene, budget_s, threads, start_step=0):
    """The oracle on this host's cores over a bounded sample of the workload,
    started from `scene` (the state the GPU's timed window starts from)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    t0 = time.perf_counter()
    o = refcpu.OracleSystem(scene, threads=threads)
    setup = time.perf_counter() - t0
    steps, elapsed, iters, errs = 0, 0.0, [], []
    while steps < 50 and (steps == 0 or elapsed < budget_s):
        t = time.perf_counter()
        o.step(1)
        elapsed += time.perf_counter() - t
        steps += 1
        m = o.measure()
        iters.append(m["solver_iterations"] + m["eqs1_iterations"])
        errs.append(m["density_error"])
    where = (f"the GPU's state at step {start_step} (the timed window's start), restarted "
             f"as continue_from_checkpoint does" if start_step else "the initial state")
    return {
        "value": scene.n * steps / elapsed,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (binary64 C restatement of the reference CPU path, OpenMP) on "
                   f"{where}, {scene.n} particles: {steps} step(s), {elapsed:.1f} s of stepping "
                   f"({setup:.1f} s setup untimed); Jacobi iterations per step "
                   f"{np.mean(iters):.1f}, density error max {max(errs):.4f} %"),
    }


def serving_bench(args):
    """The serving stream, batched and sequential, on the same requests."""
    from sph_amd import serving, system
    t = time.perf_counter()
    stream = serving.make_stream(n_requests=args.requests, rate=args.rate,
                                 n_min=args.n_min, n_max=args.n_max, steps=args.serve_steps,
                                 seed=0,
                                 schemes=serving.MIXED_SCHEMES if args.mixed_solvers else None)
    log(f"[bench] serving stream: {len(stream)} requests, "
        f"{sum(r.scene.n for r in stream)} particles, built in {time.perf_counter() - t:.1f} s")
    free, total = system.device_memory()
    budget = int(args.budget_gb * 2**30) if args.budget_gb else int(0.8 * free)
    # warm the library and the device (module loads, first allocations)
    serving.BatchedServer(budget, max_slots=2).run(serving.make_stream(
        n_requests=2, rate=1000.0, n_min=2000, n_max=3000, steps=3, seed=99))
    log("[bench] batched run")
    bat = serving.BatchedServer(budget, max_slots=args.slots,
                                graphs=not args.no_graphs).run(stream).summary()
    log(f"[bench] batched: {bat['throughput_particle_steps_per_s']:.3e} particle-steps/s, "
        f"makespan {bat['makespan_s']:.1f} s")
    seq = bound = None
    if not args.no_sequential:
        log("[bench] sequential run")
        seq_rep = serving.SequentialServer().run(stream)
        seq = seq_rep.summary()
        bound = serving.arrival_bound(stream, seq_rep)
        bound["batched_fraction_of_ceiling"] = (bat["throughput_particle_steps_per_s"] /
                                                bound["throughput_ceiling_particle_steps_per_s"])
        log(f"[bench] sequential: {seq['throughput_particle_steps_per_s']:.3e} "
            f"particle-steps/s, makespan {seq['makespan_s']:.1f} s; arrival bound "
            f"{bound['makespan_lower_bound_s']:.1f} s")
    out = {
        "metric": "serving stream throughput (particle-steps/s, continuous batching); "
                  "p50/p99 step latency",
        "value": bat["throughput_particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded serving stream)",
        "config": {"workload": "serving_stream", "requests": args.requests,
                   "arrival_rate_per_s": args.rate, "particles": [args.n_min, args.n_max],
                   "steps_per_request": args.serve_steps,
                   "solver": ("IISPH, IISPHwOST, SESPH, SESPHwSplitting in turn"
                              if args.mixed_solvers else "IISPH"),
                   "hbm_budget_bytes": budget, "max_slots": args.slots,
                   "world_graphs": not args.no_graphs,
                   "device_memory_bytes": total},
        "p50_step_ms": bat["p50_step_ms"],
        "p99_step_ms": bat["p99_step_ms"],
        "batched": bat,
        "sequential": seq,
        "speedup_vs_sequential": (bat["throughput_particle_steps_per_s"] /
                                  seq["throughput_particle_steps_per_s"]) if seq else None,
        "busy_speedup_vs_sequential": (bat["busy_throughput_particle_steps_per_s"] /
                                       seq["busy_throughput_particle_steps_per_s"])
        if seq else None,
        # the arrivals' ceiling: no server finishes a request before its
        # arrival plus its service time alone
        "arrival_bound": bound,
    }
    print(json.dumps(out))


def sensor_plane_bench(args):
    """get_quantity_at_positions on a 1024^2 plane through the 1M-particle ball."""
    from sph_amd import system
    from sph_amd._native import QUANTITIES
    scene = build_scene("two_fluid_ball", args.particles)
    sim = system.System(scene)
    for _ in range(args.warmup):
        sim.step_forward_in_time(1)
    res = 1024
    u = np.linspace(-1.0, 1.0, res)
    gx, gy = np.meshgrid(u, u, indexing="xy")
    plane = np.stack([gx.ravel(), gy.ravel(), np.zeros(res * res)], axis=1).astype(np.float32)
    sim.get_quantity_at_positions("Density", plane)  # first-call allocations
    sim.profile(True)
    lat = []
    for k in range(args.steps):
        t = time.perf_counter()
        vals = sim.get_quantity_at_positions(QUANTITIES[k % 6], plane)
        lat.append(time.perf_counter() - t)
    prof = sim.profile_read()["quantity"]
    sim.profile(False)
    nq = len(plane)
    total = float(np.sum(lat))
    avg_s = prof["ms"] / 1e3 / max(prof["launches"], 1)
    per_query = 24 + 4  # f64 position read, f32 value written
    achieved = per_query * nq / avg_s / 1e9 if avg_s > 0 else 0.0
    out = {
        "metric": "sensor-plane queries/s (get_quantity_at_positions, 1M-particle state)",
        "value": nq * args.steps / total,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "p50_step_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_step_ms": float(np.percentile(lat, 99) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice + jitter, analytic volume-map boundary)",
        "config": {"workload": f"sensor_plane_{res}x{res}_{scene.name}", "particles": len(sim),
                   "queries_per_step": nq, "quantities": QUANTITIES,
                   "nonzero_fraction_last": float(np.count_nonzero(vals)) / nq},
        "kernel_queries_per_s": nq / avg_s if avg_s > 0 else 0.0,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_quantity",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": per_query * nq,
            "avg_launch_ms": avg_s * 1e3,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import refcpu
            refcpu.lib().ref_system_set_threads(threads)
            o = refcpu.OracleSystem(scene, threads=threads)
            sub = plane[::16]
            t0 = time.perf_counter()
            for k in range(6):
                o.quantity_at_positions(k, sub)
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {
                "value": 6 * len(sub) / dt, "unit": "queries/s", "cores": threads,
                "kind": "port",
                "sample": (f"oracle ref_system_quantity_at_positions (binary64 C restatement, "
                           f"OpenMP) on every 16th plane point ({len(sub)} queries), all six "
                           f"quantities, on the initial state (System::new): {dt:.1f} s"),
            }
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline unavailable: {e}")
    sim._ctx.close()
    print(json.dumps(out))


def volume_map_bench(args):
    """Volume-map construction of the two-fluid ball's sphere container."""
    from sph_amd import scenes, system
    degree = 30
    dx, radius = 0.01, 1.0
    h = 2 * dx
    c = np.zeros(3)
    sd = lambda q: scenes.sd_sphere_container(q, c, radius)
    sdf, x_min, _, n, cell = scenes.sdf_and_volume_map_grid(sd, c - radius, c + radius, h, dx)
    gx, gw = np.polynomial.legendre.leggauss(degree)
    sim = system.System(scenes.small_iisph_box())  # lends its stream and profiler
    stats = {}
    for _ in range(args.warmup):
        system.build_volume_map(sdf, h, x_min, n, cell, gx, gw, system=sim, stats=stats)
    sim.profile(True)
    lat = []
    for _ in range(args.steps):
        t = time.perf_counter()
        vals, pres = system.build_volume_map(sdf, h, x_min, n, cell, gx, gw, system=sim,
                                             stats=stats)
        lat.append(time.perf_counter() - t)
    prof = sim.profile_read()["vm_quadrature"]
    sim.profile(False)
    q = stats["quadrature_nodes"]
    total = float(np.sum(lat))
    avg_s = prof["ms"] / 1e3 / max(prof["launches"], 1)
    flops = VM_FLOPS_PER_EVAL * degree ** 3 * prof["units"] / max(prof["launches"], 1)
    achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
    out = {
        "metric": "volume-map construction (degree-30 ball quadrature), quadrature nodes/s",
        "value": q * args.steps / total,
        "unit": "quadrature nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "p50_step_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_step_ms": float(np.percentile(lat, 99) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic sphere SDF lattice)",
        "config": {"workload": "volume_map_sphere_r1_h0.02", "cells": [int(v) for v in n],
                   "nodes": int(vals.size), "quadrature_nodes": q, "degree": degree,
                   "present_nodes": int(pres.sum())},
        "roofline": {
            "bound": "valu_fp64",
            "kernel": "k_vm_quadrature",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic_flops_per_launch": flops,
            "avg_launch_ms": avg_s * 1e3,
        },
        "cpu_baseline": None,
    }
    sim._ctx.close()
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = vm_cpu_baseline(sdf, h, x_min, n, cell, gx, gw,
                                                  args.cpu_seconds, threads)
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline unavailable: {e}")
    print(json.dumps(out))


def vm_cpu_baseline(sdf, h, x_min, n, cell, gx, gw, budget_s, threads):
    """The oracle's volume-map build on a slab of cells through the sphere's
    equator, as thick as fits the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    refcpu.lib().ref_system_set_threads(threads)
    mid = int(n[2]) // 2
    cells, rate, q, dt = 1, 0.0, 0, 0.0
    while True:
        sub_min = np.array([x_min[0], x_min[1], x_min[2] + mid * cell])
        sub_n = np.array([n[0], n[1], cells])
        st = {}
        t = time.perf_counter()
        refcpu.vm_build(sdf, h, sub_min, sub_n, cell, gx, gw, stats=st)
        dt = time.perf_counter() - t
        q = st["quadrature_nodes"]
        rate = q / dt
        if dt > 0.3 * budget_s or cells >= int(n[2]) - mid:
            break
        cells = min(int(n[2]) - mid, max(cells + 1, int(cells * 0.8 * budget_s / max(dt, 1e-3))))
    return {
        "value": rate,
        "unit": "quadrature nodes/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle ref_vm_build (binary64 C restatement, OpenMP) on a slab of "
                   f"{cells} cell layer(s) through the sphere's equator: {q} quadrature "
                   f"nodes in {dt:.1f} s"),
    }


MESH_FLOPS_PER_PAIR = 78  # closest point on a triangle, face-region path, + squared distance


def mesh_sdf_bench(args):
    """Signed-distance lattice of the two-fluid ball's container built from a
    triangle mesh: a geodesic sphere of radius 1 m (5,120 triangles), h = 0.02."""
    from sph_amd import scenes, system
    dx = 0.01
    h = 2 * dx
    v, t = scenes.icosphere(1.0, subdivisions=4, inward=True)
    pad = 3.1 * h
    x_min, _, n = scenes.lattice_grid(v.min(axis=0), v.max(axis=0), pad, 4 * dx)
    cell = 4 * dx
    sim = system.System(scenes.small_iisph_box())  # lends its stream and profiler
    for _ in range(args.warmup):
        system.build_mesh_sdf(v, t, x_min, n, cell, pad, system=sim)
    sim.profile(True)
    lat = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        vals, pres = system.build_mesh_sdf(v, t, x_min, n, cell, pad, system=sim)
        lat.append(time.perf_counter() - t0)
    prof = sim.profile_read()["mesh_sdf"]
    sim.profile(False)
    nx, ny, nz = (3 * int(k) + 1 for k in n)
    ix, iy, iz = np.meshgrid(np.arange(nx) % 3, np.arange(ny) % 3, np.arange(nz) % 3,
                             indexing="ij")
    nodes = int((((ix != 0).astype(int) + (iy != 0) + (iz != 0)) <= 1).sum())
    total = float(np.sum(lat))
    avg_s = prof["ms"] / 1e3 / max(prof["launches"], 1)
    flops = MESH_FLOPS_PER_PAIR * nodes * len(t)
    achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
    out = {
        "metric": "signed-distance lattice from a triangle mesh, lattice nodes/s",
        "value": nodes * args.steps / total,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "p50_step_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_step_ms": float(np.percentile(lat, 99) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (geodesic sphere mesh, radius 1 m, 4 subdivisions)",
        "config": {"workload": "mesh_sdf_sphere_r1_h0.02", "triangles": int(len(t)),
                   "cells": [int(k) for k in n], "nodes": nodes,
                   "present_nodes": int(pres.sum())},
        "roofline": {
            "bound": "valu_fp64",
            "kernel": "k_mesh_sdf",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic_flops_per_launch": flops,
            "avg_launch_ms": avg_s * 1e3,
        },
        "cpu_baseline": None,
    }
    sim._ctx.close()
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import refcpu
            refcpu.lib().ref_system_set_threads(threads)
            m = refcpu.Mesh(v, t)
            rng = np.random.default_rng(0)
            k = 2000
            while True:
                pts = x_min + rng.uniform(0, 1, (k, 3)) * (np.asarray(n) * cell)
                t0 = time.perf_counter()
                m.signed_distance(pts)
                dt = time.perf_counter() - t0
                if dt > 0.3 * args.cpu_seconds or k >= nodes:
                    break
                k = min(nodes, max(2 * k, int(k * 0.8 * args.cpu_seconds / max(dt, 1e-3))))
            out["cpu_baseline"] = {
                "value": k / dt, "unit": "nodes/s", "cores": threads, "kind": "port",
                "sample": (f"oracle ref_mesh_sdf (binary64 C restatement, OpenMP) at {k} "
                           f"random lattice-box positions in {dt:.1f} s")}
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline unavailable: {e}")
    print(json.dumps(out))


def contact_bench(args):
    """Rigid-rigid contact: six dynamic cubes pressed into a static slab
    (scenes.contact_pile); steps timed with contact and with it turned off."""
    from sph_amd import scenes, system

    def run(degree):
        sc = scenes.contact_pile()
        sc.contact_degree = degree
        sim = system.System(sc)
        sim.step_forward_in_time(args.warmup)
        sim.profile(True)
        lat, cells, iters = [], 0, 0
        for _ in range(args.steps):
            t0 = time.perf_counter()
            sim.step_forward_in_time(1)
            lat.append(time.perf_counter() - t0)
            c, it = sim.contact_stats()
            cells += c
            iters += it
        prof = sim.profile_read()["contact"]
        sim.profile(False)
        sim._ctx.close()
        return sc, lat, cells, iters, prof

    sc, lat, cells, iters, prof = run(scenes.CONTACT_DEGREE)
    _, lat_off, _, _, _ = run(0)
    step_ms = float(np.mean(lat)) * 1e3
    off_ms = float(np.mean(lat_off)) * 1e3
    kern_ms = prof["ms"] / max(args.steps, 1)
    cand = prof["units"] / max(args.steps, 1)
    out = {
        "metric": "rigid-rigid contact: active contact cells per second of contact time",
        "value": cells / args.steps / max((step_ms - off_ms) / 1e3, 1e-9),
        "unit": "active cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "ms_per_step_contact_off": off_ms,
        "contact_ms_per_step": step_ms - off_ms,
        "contact_kernel_ms_per_step": kern_ms,
        "candidate_subcells_per_step": cand,
        "active_cells_per_step": cells / args.steps,
        "nncg_iterations_per_step": iters / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (box, slab and cube SDF lattices, seeded tilts)",
        "config": {"workload": sc.name, "bodies": len(sc.boundaries), "particles": sc.n,
                   "quadrature_degree": scenes.CONTACT_DEGREE},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import refcpu
            refcpu.lib().ref_system_set_threads(threads)
            o = refcpu.OracleSystem(scenes.contact_pile())
            o.step(1)
            k, t_on = 0, 0.0
            c_on = 0
            while t_on < 0.5 * args.cpu_seconds and k < 200:
                t0 = time.perf_counter()
                o.step(1)
                t_on += time.perf_counter() - t0
                c_on += o.contact_stats()[0]
                k += 1
            out["cpu_baseline"] = {
                "value": c_on / t_on, "unit": "active cells/s (whole step)", "cores": threads,
                "kind": "port",
                "sample": (f"oracle steps of the same scene with contact ({k} steps, "
                           f"{t_on:.1f} s; the fluid is 144 particles, so contact dominates)")}
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline unavailable: {e}")
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="two_fluid_ball")
    ap.add_argument("--particles", type=int, default=1_000_000,
                    help="two-fluid ball: particles")
    ap.add_argument("--settle", type=int, default=60,
                    help="untimed steps before the warm-up (the fluid settles into a loaded "
                         "state; the code-shape tuner finishes)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-one-thread", action="store_true",
                    help="skip the one-core step of the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain device copy of known size (counter calibration)")
    ap.add_argument("--shapes", default="",
                    help="code shapes of the volume, non-pressure, source/diagonal, "
                         "pressure-acceleration and Jacobi passes (five comma-separated indices: "
                         "the tuner is off; counter runs repeat a timed run's choice)")
    ap.add_argument("--requests", type=int, default=256, help="serving: requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: arrivals per second")
    ap.add_argument("--n-min", type=int, default=4096, help="serving: smallest scene")
    ap.add_argument("--n-max", type=int, default=262144, help="serving: largest scene")
    ap.add_argument("--serve-steps", type=int, default=200, help="serving: steps per request")
    ap.add_argument("--slots", type=int, default=64, help="serving: resident scenes at most")
    ap.add_argument("--budget-gb", type=float, default=0.0,
                    help="serving: HBM budget (default 80%% of free device memory)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="serving: skip the sequential baseline")
    ap.add_argument("--no-graphs", action="store_true",
                    help="serving: launch the world's Jacobi chunks directly")
    ap.add_argument("--mixed-solvers", action="store_true",
                    help="serving: requests cycle through the four pressure solvers "
                         "(one world per solver)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; the hot path's "
                          "multi-GPU decomposition (spatial slabs with halo exchanges) is designed, not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "serving":
        serving_bench(args)
        return
    if args.workload == "volume_map":
        volume_map_bench(args)
        return
    if args.workload == "mesh_sdf":
        mesh_sdf_bench(args)
        return
    if args.workload == "contact":
        contact_bench(args)
        return
    if args.workload == "sensor_plane":
        sensor_plane_bench(args)
        return

    import dataclasses
    from sph_amd import system
    scene = build_scene(args.workload, args.particles)
    log(f"[bench] scene {scene.name}: {scene.n} particles")
    sim = system.System(scene)
    pinned = bool(args.shapes)
    if pinned:
        from sph_amd import _native as N
        shapes = [int(x) for x in args.shapes.split(",")]
        if len(shapes) != len(N.PAIR_PASSES):
            raise SystemExit("--shapes takes one code shape per neighbor pass")
        for k, v in enumerate(shapes):
            sim.set_option(N.OPT_SHAPE_VOLUME + k, v)
    # settle: the first steps let the resting fluid compress until every
    # solve carries its weight, and the code-shape tuner finishes (one timed
    # step per shape); neither is part of the timed window, whatever --warmup is
    settled = 0
    while settled < args.settle or (not pinned and sim.tuning()["state"] != 2):
        sim.step_forward_in_time(1)
        settled += 1
        if settled > args.settle + 100:
            raise SystemExit("the code-shape tuner did not finish")
    for _ in range(args.warmup):
        sim.step_forward_in_time(1)
    start_step = settled + args.warmup
    start_state = sim.state() if not args.no_cpu else None
    sim.take_measurement()  # drains the stream

    def run(k):
        lat, iters, errs = [], [], []
        for _ in range(k):
            t = time.perf_counter()
            sim.step_forward_in_time(1)   # returns after the step's last device read-back
            lat.append(time.perf_counter() - t)
            m = sim.take_measurement()
            iters.append(m["solver_iterations"] + m["eqs1_iterations"])
            errs.append(m["density_error"])
        return lat, iters, errs

    # the timed window: the production path (Jacobi chunks replayed from HIP
    # graphs, no per-kernel events). Each step call returns after the step's
    # final device read-back, so the sum of the step latencies is the
    # device-synchronized time of the K steps.
    t_start = time.perf_counter()
    lat, iters, errs = run(args.steps)
    total = float(np.sum(lat))
    wall = time.perf_counter() - t_start
    # per-kernel durations: the next K steps again with HIP events around every
    # kernel launch (which launches the Jacobi chunks directly)
    sim.profile(True)
    plat, piters, _ = run(args.steps)
    prof = sim.profile_read()
    sim.profile(False)
    n = len(sim)
    value = n * args.steps / total

    nb = len(scene.boundaries)
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    units = d["units"] / max(d["launches"], 1)
    avg_s = d["ms"] / 1e3 / max(d["launches"], 1)
    bytes_per_launch = F32_BYTES[dom] * units
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    per_unit64 = ALGO_BYTES[dom]["base"] + ALGO_BYTES[dom]["per_body"] * nb
    achieved64 = per_unit64 * units / avg_s / 1e9 if avg_s > 0 else 0.0
    mean_iters = float(np.mean(iters))
    step_bytes = step_bytes_f32(scene.params.solver, mean_iters)

    out = {
        "metric": "particle-steps/sec (IISPH, 1M particles) at <=0.1% density error; p50/p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "wall_s_incl_measurements": wall,
        "p50_step_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_step_ms": float(np.percentile(lat, 99) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice + jitter, analytic volume-map boundary)",
        "config": {"workload": scene.name, "particles": n, "solver": scene.params.solver,
                   "integrator": scene.params.integrator, "dt": "CFL 0.4, max 1e-3",
                   "target_density_error_pct": scene.params.target_density_error,
                   "timed_steps": [start_step + 1, start_step + args.steps],
                   "parallelism": "single GPU"},
        "density_error_pct_max": float(max(errs)),
        "density_error_pct_last": float(errs[-1]),
        "jacobi_iterations_per_step_mean": mean_iters,
        "jacobi_iterations_per_step_range": [int(min(iters)), int(max(iters))],
        "step_roofline": {"bytes_per_particle_step": step_bytes,
                          "achieved": step_bytes * value / 1e9, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_bytes * value / 1e9 / HBM_PEAK_GBS},
        "code_shapes": sim.tuning(),
        "kernel_ms_per_step": {k: v["ms"] / args.steps for k, v in prof.items()},
        "profiled_window": {"ms_per_step": float(np.mean(plat)) * 1e3,
                            "jacobi_iterations_per_step_mean": float(np.mean(piters)),
                            "note": "the K steps after the timed window, with HIP events "
                                    "around every kernel (Jacobi chunks launched directly)"},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (traffic_from_profiles(dom)
                        if args.workload == "two_fluid_ball" and args.particles == 1_000_000
                        else None),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "bytes_per_particle": F32_BYTES[dom],
            "avg_launch_ms": avg_s * 1e3,
            "binary64_bytes_per_particle": per_unit64,
            "achieved_binary64": achieved64,
            "frac_binary64": achieved64 / HBM_PEAK_GBS,
        },
        "cpu_baseline": None,
    }
    if dom in FLOPS_PER_PAIR:
        # the same kernel against the FP64 vector peak: neighbor pairs per
        # particle estimated from 20,000 particles of the final state, without
        # the self pair (the pair passes do not evaluate it)
        from scipy.spatial import cKDTree
        xs = sim.state()["x"]
        pick = np.random.default_rng(0).choice(len(xs), size=min(20000, len(xs)), replace=False)
        pairs = float(np.mean(cKDTree(xs).query_ball_point(
            xs[pick], scene.params.kernel_support_radius, return_length=True))) - 1.0
        flops = FLOPS_PER_PAIR[dom] * pairs * units
        out["roofline_fp64"] = {
            "bound": "fp64_valu", "kernel": dom,
            "achieved": flops / avg_s / 1e12 if avg_s > 0 else 0.0,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": flops / avg_s / 1e12 / FP64_PEAK_TFLOPS if avg_s > 0 else 0.0,
            "flops_per_pair": FLOPS_PER_PAIR[dom], "pairs_per_particle": pairs,
            "note": "divisions and square roots count one operation each but issue as "
                    "~10 VALU instructions; the counter profile gives VALU busy"}
    if args.copy_probe:
        # cloning the system copies each binary64 particle array device-to-device
        # (one runtime copy kernel per array, 8 bytes x particles each); those
        # dispatches calibrate the HBM counters against a known byte count
        sim.clone()
        out["copy_probe_bytes"] = 8 * n
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        try:
            base = scene
            if not any(getattr(b, "rigid", None) is not None for b in scene.boundaries):
                base = dataclasses.replace(scene, x=start_state["x"], v=start_state["v"],
                                           m=start_state["m"], fluid_id=start_state["fluid_id"])
            else:
                start_step = 0
            out["cpu_baseline"] = cpu_baseline(base, args.cpu_seconds, threads, start_step)
            if threads > 1 and not args.no_one_thread:
                # the same state on one core: one step (a bounded sample)
                one = cpu_baseline(base, 0.0, 1, start_step)
                out["cpu_baseline"]["one_thread"] = {"value": one["value"], "cores": 1,
                                                     "sample": one["sample"]}
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline unavailable: {e}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

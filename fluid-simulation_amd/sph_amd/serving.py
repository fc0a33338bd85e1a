"""Continuous-batching service for many SPH scenes on one GPU.

The reference runs one simulation at a time on one worker thread and steps
it alone (``worker_loop``, sci-phi-backend/src/lib.rs:507, a thread spawned by
the UI, sci-phi/src/app/mod.rs:236: one ``Box<dyn SPHSystem>`` whose
``step_forward_in_time`` runs in a loop). Here a stream of scene
requests is served by one process and one device:

* a request is admitted when a slot and its device-memory estimate fit the
  HBM budget; its state (particles, neighbor and cell structures, volume-map
  lattices) then stays resident until it finishes;
* every admitted scene joins one ``World``: each time step of the world is one
  sequence of multi-scene launches that advances every resident scene by one
  of its own time steps (each keeps its own CFL step size and pressure-solve
  termination);
* requests that arrive while others run are admitted between world steps, and
  finished scenes leave between world steps, freeing their slot and memory.

Each scene's results are bit-identical to running it alone (the world's
per-scene arithmetic and reductions do not depend on the other scenes).

The sequential baseline serves the same stream one request at a time with
``System.step_forward_in_time``, the plain per-scene GPU path.
"""
from __future__ import annotations

import dataclasses
import hashlib
import math
import time
from collections import deque
from typing import Callable, Dict, List, Optional

import numpy as np

from . import scenes as S

# device bytes per particle while resident in a world: 49 binary64 and 4 u32
# per-particle fields of the scene, the world's per-particle sort scratch and
# 48 pair slots of 4 B (the world gives every member the same slot width, its
# largest neighbor count + 8: ~48 for the stream's scenes, which share one
# h / dx ratio) and one volume-map slot (flag + 7 binary64) per boundary, plus
# 64 B for the world's shared buffers, which grow in steps of 1.25x (the
# serving stream held 2.68 GB at its peak against 2.59 GB of estimates
# without it). The world's launch buffers only grow, and the library frees them once its
# residents need less than a quarter of them; servers report the bytes the
# library really holds (device_bytes) beside these estimates.
BYTES_PER_PARTICLE = 47 * 8 + 4 * 4 + 20 + 47 * 4 + 64   # + 64: the shared buffers' 1.25x growth
BYTES_PER_PARTICLE_BODY = 1 + 7 * 8
BYTES_PER_SCENE = 1 << 20          # stream, pinned staging, small tables


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                  # seconds after the stream starts
    scene: S.Scene
    steps: int
    # Jacobi iterations after which its pressure solve fails (None: the
    # library's default); a failing request is reported, the others go on
    max_iterations: Optional[int] = None


@dataclasses.dataclass
class Result:
    rid: int
    n: int
    steps: int
    arrival: float
    admitted: float                 # service started (before the scene's init)
    finished: float
    measurement: dict
    digest: str                     # sha256 of the final x, v (reference order)
    state: Optional[dict] = None
    error: Optional[str] = None     # set for a failed request (no valid state)


@dataclasses.dataclass
class Report:
    mode: str
    results: Dict[int, Result]
    step_ms: List[float]            # one entry per (request, step)
    makespan_s: float               # first arrival -> last completion
    busy_s: float                   # time inside stepping calls
    admit_s: float                  # time spent creating and initializing systems
    peak_resident: int
    peak_bytes: int
    peak_device_bytes: int = 0      # measured: what the library's buffers held

    @property
    def particle_steps(self) -> int:
        return sum(r.n * r.steps for r in self.results.values() if r.error is None)

    @property
    def failed(self) -> Dict[int, Result]:
        return {k: r for k, r in self.results.items() if r.error is not None}

    def summary(self) -> dict:
        lat = np.asarray(self.step_ms) if self.step_ms else np.zeros(1)
        ok = [r for r in self.results.values() if r.error is None]
        req = np.asarray([r.finished - r.arrival for r in ok] or [0.0])
        derr = [r.measurement["density_error"] for r in ok]
        perr = [r.measurement["predicted_density_error"] for r in ok]
        return {
            "mode": self.mode,
            "requests": len(self.results),
            "particle_steps": self.particle_steps,
            "makespan_s": self.makespan_s,
            "throughput_particle_steps_per_s": self.particle_steps / max(self.makespan_s, 1e-12),
            "busy_throughput_particle_steps_per_s": self.particle_steps / max(self.busy_s, 1e-12),
            "busy_s": self.busy_s,
            "admit_s": self.admit_s,
            "p50_step_ms": float(np.percentile(lat, 50)),
            "p99_step_ms": float(np.percentile(lat, 99)),
            "p50_request_s": float(np.percentile(req, 50)),
            "p99_request_s": float(np.percentile(req, 99)),
            "peak_resident_scenes": self.peak_resident,
            "peak_resident_bytes_est": self.peak_bytes,
            "peak_device_bytes_held": self.peak_device_bytes,
            # the last step's average density error after the position update
            # (measured) and the pressure solve's last predicted error
            "density_error_pct_max": float(max(derr)) if derr else 0.0,
            "density_error_pct_p50": float(np.percentile(derr, 50)) if derr else 0.0,
            "requests_measured_error_above_0p1_pct": int(sum(e > 0.1 for e in derr)),
            "predicted_density_error_pct_max": float(max(perr)) if perr else 0.0,
            "failed_requests": len(self.results) - len(ok),
        }


def arrival_bound(stream: List[Request], solo: Report) -> dict:
    """What the arrivals allow any server on this device: request r finishes
    no earlier than its arrival plus its service time alone (from the
    sequential run `solo`, system creation and init included), so the makespan
    is at least max_r(arrival_r + service_r) - first arrival; the throughput
    ceiling is the stream's particle-steps over that bound."""
    svc = {rid: r.finished - r.admitted for rid, r in solo.results.items() if r.error is None}
    served = [q for q in stream if q.rid in svc]
    if not served:
        return {"makespan_lower_bound_s": None, "throughput_ceiling_particle_steps_per_s": None}
    t0 = min(q.arrival for q in stream)
    last = max(served, key=lambda q: q.arrival + svc[q.rid])
    bound = last.arrival + svc[last.rid] - t0
    ps = sum(q.scene.n * q.steps for q in served)
    return {"makespan_lower_bound_s": bound,
            "last_arrival_s": max(q.arrival for q in stream) - t0,
            "binding_request": last.rid,
            "throughput_ceiling_particle_steps_per_s": ps / bound,
            "speedup_ceiling_vs_sequential": solo.makespan_s / bound}


def make_stream(n_requests: int = 256, rate: float = 20.0, n_min: int = 4096,
                n_max: int = 262144, steps: int = 200, seed: int = 0,
                dx: float = 0.02, schemes=None) -> List[Request]:
    """The seeded serving stream: Poisson arrivals at ``rate`` per second,
    particle counts log-uniform in [n_min, n_max], rest density U[800, 1500],
    viscosity U[0.005, 0.05], box or sphere container, ``steps`` IISPH steps.
    ``schemes``, a list of (solver, integrator) pairs, gives request k the
    pair k mod its length instead (the rest of the stream is unchanged)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    t -= t[0]
    ns = np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=n_requests)).astype(int)
    rho = rng.uniform(800.0, 1500.0, size=n_requests)
    nu = rng.uniform(0.005, 0.05, size=n_requests)
    shape = rng.integers(0, 2, size=n_requests)
    seeds = rng.integers(0, 2**31 - 1, size=n_requests)
    out = []
    for k in range(n_requests):
        sc = S.serving_scene(int(ns[k]), float(rho[k]), float(nu[k]),
                             "box" if shape[k] == 0 else "sphere", int(seeds[k]), dx=dx)
        if schemes:
            sc.params.solver, sc.params.integrator = schemes[k % len(schemes)]
        out.append(Request(k, float(t[k]), sc, int(steps)))
    return out


# the four pressure solvers, each with its integrator of the reference's configs
MIXED_SCHEMES = [("IISPH", "EulerCromer"), ("IISPHwOST", "TakePredicted"),
                 ("SESPH", "EulerCromer"), ("SESPHwSplitting", "EulerCromer")]


def lattice_bytes(lat: S.Lattice) -> int:
    return int(lat.values.size) * 8 + int(np.prod(lat.n)) + 256


def estimate_device_bytes(scene: S.Scene) -> int:
    """Device memory a scene holds while resident in a world (an estimate used
    for admission; the library allocates on demand)."""
    n = scene.n
    b = BYTES_PER_SCENE + n * BYTES_PER_PARTICLE
    for bd in scene.boundaries:
        if hasattr(bd, "sdf"):
            b += n * BYTES_PER_PARTICLE_BODY + lattice_bytes(bd.sdf) + lattice_bytes(bd.vm)
        else:
            b += int(len(bd.positions)) * 64
    return int(b)


def _digest(state: dict) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(state["x"]).tobytes())
    h.update(np.ascontiguousarray(state["v"]).tobytes())
    return h.hexdigest()


def _finish(req: Request, sim, admitted: float, now: float, keep_state: bool) -> Result:
    st = sim.state()
    return Result(req.rid, len(sim), req.steps, req.arrival, admitted, now,
                  sim.take_measurement(), _digest(st), st if keep_state else None)


def _failed(req: Request, sim, admitted: float, now: float, error: str) -> Result:
    return Result(req.rid, len(sim), req.steps, req.arrival, admitted, now, {}, "", None, error)


def _check_fits(req: Request, budget: int):
    need = estimate_device_bytes(req.scene)
    if need > budget:
        raise ValueError(f"request {req.rid} needs {need} B, more than the whole budget")
    return need


class BatchedServer:
    """Continuous batching: the library's server (sph_server_*) admits requests
    in arrival order under slots and the HBM budget and steps the admitted
    scenes together in one world per pressure solver; this loop feeds it
    arrivals, builds each
    admitted request's System and collects finished ones."""

    def __init__(self, hbm_budget_bytes: int, max_slots: int = 64,
                 clock: Callable[[], float] = time.perf_counter, keep_state: bool = False,
                 graphs: bool = True):
        from . import system  # loads the HIP library
        self._system = system
        self.budget = int(hbm_budget_bytes)
        self.max_slots = int(max_slots)
        self.clock = clock
        self.keep_state = keep_state
        self.graphs = graphs      # replay the world's Jacobi chunks from HIP graphs

    def run(self, stream: List[Request]) -> Report:
        system = self._system
        pending = deque(sorted(stream, key=lambda r: r.arrival))
        waiting: Dict[int, Request] = {}     # ticket -> submitted request
        admitted: Dict[int, list] = {}       # ticket -> [request, admitted at]
        results: Dict[int, Result] = {}
        step_ms: List[float] = []
        srv = system.Server(self.budget, self.max_slots)
        if not self.graphs:
            srv.set_option(system.N.OPT_GRAPHS, 0)
        busy = admit = 0.0
        peak_n = peak_b = peak_d = 0
        t0 = self.clock()
        try:
            while pending or waiting or admitted:
                now = self.clock() - t0
                while pending and pending[0].arrival <= now:
                    req = pending.popleft()
                    waiting[srv.submit(_check_fits(req, self.budget), req.steps)] = req
                while True:
                    t = srv.admit()
                    if t is None:
                        break
                    req = waiting.pop(t)
                    a0 = self.clock()
                    sim = system.System(req.scene, init=False)
                    if req.max_iterations is not None:
                        sim.set_max_iterations(req.max_iterations)
                    try:
                        sim.init()
                    except system.N.SphError as e:  # System::new's update failed
                        srv.cancel(t)
                        now = self.clock() - t0
                        results[req.rid] = _failed(req, sim, now, now, str(e))
                        sim._ctx.close()
                        continue
                    srv.attach(t, sim)
                    admit += self.clock() - a0
                    admitted[t] = [req, self.clock() - t0]
                st = srv.stats()
                peak_n = max(peak_n, st["resident"])
                peak_b = max(peak_b, st["bytes"])
                if st["resident"] == 0:
                    if pending:  # idle until the next arrival
                        time.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                    continue
                s0 = self.clock()
                peak_d = max(peak_d, srv.step()["device_bytes"])
                dt = self.clock() - s0
                busy += dt
                step_ms.extend([dt * 1e3] * st["resident"])
                while True:
                    got = srv.poll()
                    if got is None:
                        break
                    t, sim = got
                    req, t_adm = admitted.pop(t)
                    results[req.rid] = _finish(req, sim, t_adm, self.clock() - t0,
                                               self.keep_state)
                    sim._ctx.close()
                while True:
                    got = srv.poll_failed()
                    if got is None:
                        break
                    t, sim, status = got
                    req, t_adm = admitted.pop(t)
                    results[req.rid] = _failed(req, sim, t_adm, self.clock() - t0,
                                               f"status {status}: {sim.status()[1]}")
                    sim._ctx.close()
        finally:
            srv.close()
        makespan = max(r.finished for r in results.values()) - min(r.arrival for r in stream)
        return Report("batched", results, step_ms, makespan, busy, admit, peak_n, peak_b, peak_d)


class SequentialServer:
    """The baseline: one request at a time, each stepped alone."""

    def __init__(self, clock: Callable[[], float] = time.perf_counter, keep_state: bool = False):
        from . import system
        self._system = system
        self.clock = clock
        self.keep_state = keep_state

    def run(self, stream: List[Request]) -> Report:
        system = self._system
        results: Dict[int, Result] = {}
        step_ms: List[float] = []
        busy = admit = 0.0
        peak_d = 0
        t0 = self.clock()
        for req in sorted(stream, key=lambda r: r.arrival):
            wait = req.arrival - (self.clock() - t0)
            if wait > 0:
                time.sleep(wait)
            a0 = self.clock()
            t_adm = a0 - t0
            sim = system.System(req.scene, init=False)
            if req.max_iterations is not None:
                sim.set_max_iterations(req.max_iterations)
            try:
                sim.init()
                peak_d = max(peak_d, system.device_bytes_held())
                admit += self.clock() - a0
                for _ in range(req.steps):
                    s0 = self.clock()
                    sim.step_forward_in_time(1)
                    dt = self.clock() - s0
                    busy += dt
                    step_ms.append(dt * 1e3)
            except system.N.SphError as e:  # this request fails, the stream goes on
                now = self.clock() - t0
                results[req.rid] = _failed(req, sim, now, now, str(e))
                sim._ctx.close()
                continue
            results[req.rid] = _finish(req, sim, t_adm, self.clock() - t0, self.keep_state)
            sim._ctx.close()
        makespan = max(r.finished for r in results.values()) - min(r.arrival for r in stream)
        return Report("sequential", results, step_ms, makespan, busy, admit, 1,
                      max(estimate_device_bytes(r.scene) for r in stream), peak_d)

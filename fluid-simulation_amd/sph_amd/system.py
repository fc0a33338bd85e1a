"""Host-side mirror of the reference's plug-in API over the HIP library.

Names follow simulation_lib: ``SpatialHashing.find_samples`` returns a
``NeighborList`` (neighbor_search/mod.rs:18-161); ``System`` mirrors
``System<K,I,P,N,B>`` with ``step_forward_in_time`` / ``take_measurement``
(sph/mod.rs:40-184); the per-trait functions ``get_volume``,
``find_boundary_samples``, ``add_non_pressure_accelerations`` and
``solve_and_add_acceleration`` mirror the three traits a ``System`` is generic
over. Every call runs on the GPU through libsphhip.so.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as N
from .scenes import INTEGRATORS, SOLVERS, Params, Scene, contact_rule


def _params_struct(P: Params) -> N.SphParams:
    s = N.SphParams()
    s.time_increment = P.time_increment
    s.cfl = int(bool(P.cfl))
    s.max_time_increment = P.max_time_increment
    s.cfl_number = P.cfl_number
    s.kernel_support_radius = P.kernel_support_radius
    s.rest_density_grid_spacing = P.rest_density_grid_spacing
    s.disable_particles_below = P.disable_particles_below
    s.fluid_viscosity = P.fluid_viscosity
    s.boundary_viscosity = P.boundary_viscosity
    s.boundary_pressure_acceleration_weighting = P.boundary_pressure_acceleration_weighting
    s.boundary_rest_volume_weighting = P.boundary_rest_volume_weighting
    s.gravity_mode = int(P.gravity_mode)
    for d in range(3):
        s.gravity_center[d] = float(P.gravity_center[d])
    s.solver = SOLVERS[P.solver]
    s.integrator = INTEGRATORS[P.integrator]
    s.stiffness = P.stiffness
    s.target_density_error = P.target_density_error
    s.relaxation_factor = P.relaxation_factor
    s.min_diagonal_element = P.min_diagonal_element
    return s


def _lattice_struct(lat):
    s = N.SphLattice()
    for d in range(3):
        s.x_min[d] = float(lat.x_min[d])
        s.x_max[d] = float(lat.x_max[d])
        s.n[d] = int(lat.n[d])
    s.dx = float(lat.dx)
    vals = np.ascontiguousarray(lat.values, dtype=np.float64)
    pres = np.ascontiguousarray(lat.present, dtype=np.uint8)
    s.values = N.dptr(vals)
    s.present = pres.ctypes.data_as(N._u8p)
    s._keep = (vals, pres)
    return s


def _rigid_struct(rb):
    if rb is None:
        return None
    s = N.SphRigidInit()
    ib = np.asarray(rb.inertia_body, dtype=np.float64)
    s.mass = float(rb.mass)
    for k, v in enumerate(ib.ravel()):
        s.inertia_body[k] = float(v)
    for k, v in enumerate(np.linalg.inv(ib).ravel()):
        s.inertia_inv_body[k] = float(v)
    for k in range(3):
        s.com[k] = float(rb.com[k])
        s.lin_vel[k] = float(rb.lin_vel[k])
        s.ang_vel[k] = float(rb.ang_vel[k])
    for k in range(4):
        s.q[k] = float(rb.q[k])
    return s


class NeighborList:
    """CSR neighbor list (neighbor_search/mod.rs:49-161)."""

    def __init__(self, offsets, indices):
        self.offsets = offsets
        self.indices = indices

    def num_samples(self):
        return len(self.offsets) - 1

    def get_neighbors(self, i):
        return self.indices[self.offsets[i]:self.offsets[i + 1]]


class _Ctx:
    def __init__(self, params: Params):
        N.lib()
        self._p = _params_struct(params)
        h = C.c_void_p()
        N.check(N.lib().sph_ctx_create(C.byref(self._p), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            N.lib().sph_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpatialHashing:
    """``SpatialHashing::new(search_range)`` + ``find_samples`` on the GPU."""

    def __init__(self, search_range: float):
        self.cell_size = float(search_range)
        self._ctx = _Ctx(Params(kernel_support_radius=self.cell_size,
                                rest_density_grid_spacing=self.cell_size))

    def find_samples(self, search_range, positions, sample_positions) -> NeighborList:
        pos = N.f64(positions).reshape(-1, 3)
        smp = N.f64(sample_positions).reshape(-1, 3)
        n, m = len(pos), len(smp)
        off = np.zeros(n + 1, dtype=np.uint64)
        total = C.c_size_t(0)
        L = N.lib()
        cap = max(64 * n, 1)
        idx = np.zeros(cap, dtype=np.uint64)
        st = L.sph_neighbors_find(self._ctx.h, self.cell_size, float(search_range), N.dptr(pos),
                                  n, N.dptr(smp), m, off.ctypes.data_as(N._u64p),
                                  idx.ctypes.data_as(N._u64p), cap, C.byref(total))
        if st == -4:  # SPH_ERR_CAPACITY: retry with the exact size
            cap = int(total.value)
            idx = np.zeros(max(cap, 1), dtype=np.uint64)
            st = L.sph_neighbors_find(self._ctx.h, self.cell_size, float(search_range),
                                      N.dptr(pos), n, N.dptr(smp), m,
                                      off.ctypes.data_as(N._u64p), idx.ctypes.data_as(N._u64p),
                                      cap, C.byref(total))
        N.check(st)
        return NeighborList(off, idx[:int(total.value)].copy())


class System:
    """Device-resident ``System<K,I,P,N,B>`` for one scene (sph/mod.rs:75-97)."""

    def __init__(self, scene: Scene, init: bool = True):
        self.scene = scene
        self._ctx = _Ctx(scene.params)
        L = N.lib()
        self._keep = []
        for b in scene.boundaries:
            rig = _rigid_struct(getattr(b, "rigid", None))
            rp = C.byref(rig) if rig is not None else None
            if hasattr(b, "positions"):
                pos = N.f64(b.positions)
                self._keep.append(pos)
                N.check(L.sph_boundary_add_samples(self._ctx.h, N.dptr(pos), len(pos), rp))
            else:
                sdf, vm = _lattice_struct(b.sdf), _lattice_struct(b.vm)
                N.check(L.sph_boundary_add_volume_map(self._ctx.h, C.byref(sdf), C.byref(vm),
                                                      rp))
        for i, b in enumerate(scene.boundaries):
            if getattr(b, "aabb", None) is not None:
                c, hw = N.f64(b.aabb[0]), N.f64(b.aabb[1])
                N.check(L.sph_boundary_set_aabb(self._ctx.h, i, N.dptr(c), N.dptr(hw)))
        # rigid-rigid contact runs in every volume-map system of the reference
        degree = int(getattr(scene, "contact_degree", 0) or 0)
        if degree and all(not hasattr(b, "positions") for b in scene.boundaries):
            gx, gw = (N.f64(a) for a in contact_rule(degree))
            self._keep += [gx, gw]
            N.check(L.sph_contact_set_quadrature(self._ctx.h, N.dptr(gx), N.dptr(gw), degree))
        x, v, m = N.f64(scene.x), N.f64(scene.v), N.f64(scene.m)
        fid = np.ascontiguousarray(scene.fluid_id, dtype=np.uint32)
        N.check(L.sph_fluid_upload(self._ctx.h, len(x), N.dptr(x), N.dptr(v), N.dptr(m),
                                   fid.ctypes.data_as(N._u32p)))
        if init:
            N.check(L.sph_system_init(self._ctx.h))

    def init(self):
        """``System::new_boxed``'s update (after ``init=False``)."""
        N.check(N.lib().sph_system_init(self._ctx.h))

    @property
    def handle(self):
        return self._ctx.h

    def step_forward_in_time(self, n: int = 1):
        """n x SPHSystem::step_forward_in_time (sph/mod.rs:125-144). The wall
        clock of the last step is kept for recordings, as the reference keeps
        time_step_wall_clock_time (sph/mod.rs:143)."""
        n = int(n)
        if n > 1:
            N.check(N.lib().sph_step(self._ctx.h, n - 1, None))
        t0 = time.perf_counter()
        N.check(N.lib().sph_step(self._ctx.h, min(n, 1), None))
        if n >= 1:
            self._wall = time.perf_counter() - t0

    def __len__(self):
        n = C.c_size_t()
        N.check(N.lib().sph_fluid_len(self._ctx.h, C.byref(n)))
        return int(n.value)

    def take_measurement(self) -> dict:
        """``SPHSystem::take_measurement`` (sph/mod.rs:147-184): the library's
        sums and counters (also eqs1_iterations, steps and num_particles), and
        the reference's Measurement fields. The parameters enter through the
        solver's SolverMeasurementInfo: the state-equation solvers report their
        stiffness only, IISPH and IISPHwOST their target error and relaxation
        factor (sesph.rs:69-74, iisph.rs:98-106), the others 0. fluid_depth is
        never set by the reference (sph/mod.rs:738-746). The step's wall clock
        is left to recordings (time_step_info), so that two runs of one scene
        give equal measurements."""
        m = N.SphMeasurement()
        N.check(N.lib().sph_measure(self._ctx.h, C.byref(m)))
        out = m.as_dict()
        P = self.scene.params
        iterative = P.solver in ("IISPH", "IISPHwOST")
        out.update(stiffness=0.0 if iterative else P.stiffness,
                   fluid_viscosity=P.fluid_viscosity, boundary_viscosity=P.boundary_viscosity,
                   fluid_depth=0.0, rest_density_grid_spacing=P.rest_density_grid_spacing,
                   kernel_support_radius=P.kernel_support_radius,
                   target_density_error=P.target_density_error if iterative else 0.0,
                   relaxation_factor=P.relaxation_factor if iterative else 0.0)
        return out

    def state(self) -> dict:
        n = len(self)
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "a")}
        for k in ("V", "p", "m"):
            out[k] = np.zeros(n)
        out["fluid_id"] = np.zeros(n, dtype=np.uint32)
        N.check(N.lib().sph_fluid_download(self._ctx.h, N.dptr(out["x"]), N.dptr(out["v"]),
                                           N.dptr(out["a"]), N.dptr(out["V"]), N.dptr(out["p"]),
                                           N.dptr(out["m"]),
                                           out["fluid_id"].ctypes.data_as(N._u32p)))
        return out

    def get_quantity_at_positions(self, quantity, positions) -> np.ndarray:
        """``SPHSystem::get_quantity_at_positions`` (``sph/mod.rs:241-341``):
        ``quantity`` is a ScalarQuantity name ("Speed", "Volume", "Density",
        "DensityError", "Pressure", "KineticEnergy") or its index; positions
        are [n][3] f32. Returns n f32 values."""
        q = N.QUANTITIES.index(quantity) if isinstance(quantity, str) else int(quantity)
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        out = np.empty(len(pos), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        N.check(N.lib().sph_quantity_at_positions(self._ctx.h, q, len(pos),
                                                  pos.ctypes.data_as(fp),
                                                  out.ctypes.data_as(fp)))
        return out

    def error_history(self):
        cnt = C.c_size_t()
        buf = np.zeros(100000)
        N.check(N.lib().sph_error_history(self._ctx.h, N.dptr(buf), len(buf), C.byref(cnt)))
        return buf[:min(int(cnt.value), len(buf))].copy()

    def body_state(self, b: int) -> dict:
        out = np.zeros(22)
        N.check(N.lib().sph_body_state(self._ctx.h, int(b), N.dptr(out)))
        return {"com": out[0:3], "q": out[3:7], "lin_vel": out[7:10], "ang_mom": out[10:13],
                "force": out[13:16], "torque": out[16:19], "ang_vel": out[19:22]}

    def body_forces(self, b: int):
        out = np.zeros(6)
        N.check(N.lib().sph_body_forces(self._ctx.h, int(b), N.dptr(out)))
        return out

    def contact_stats(self):
        """(active contact cells, NNCG iterations) of the last step."""
        out = np.zeros(2, dtype=np.uint32)
        N.check(N.lib().sph_contact_stats(self._ctx.h, out.ctypes.data_as(N._u32p)))
        return int(out[0]), int(out[1])

    def contact_cells(self) -> np.ndarray:
        """The last step's active contact cells in solve order, [n][24]: body a,
        body b, volume, center, f_a, tau_a, f_b, tau_b, a_cc pressure, a_cc
        friction, p, lambda."""
        n = self.contact_stats()[0]
        out = np.zeros((n, 24))
        cnt = C.c_size_t()
        N.check(N.lib().sph_contact_cells(self._ctx.h, N.dptr(out), n, C.byref(cnt)))
        return out

    def profile(self, enable: bool = True):
        """Start (and clear) or stop HIP-event timing of the hot kernels."""
        N.check(N.lib().sph_profile(self._ctx.h, int(bool(enable))))

    def profile_read(self) -> dict:
        return _read_profile(N.lib().sph_profile_read, self._ctx.h)

    def set_cache_gradients(self, on: bool):
        """Read grad W from the pair cache (True) or recompute it (False, default)."""
        N.check(N.lib().sph_set_option(self._ctx.h, N.OPT_CACHE_GRADIENTS, int(bool(on))))

    def set_option(self, option: int, value: int):
        N.check(N.lib().sph_set_option(self._ctx.h, int(option), int(value)))

    def get_option(self, option: int) -> int:
        v = C.c_int()
        N.check(N.lib().sph_get_option(self._ctx.h, int(option), C.byref(v)))
        return int(v.value)

    def tuning(self) -> dict:
        """Tuner state and the code shape of each neighbor pass."""
        return _tuning(self.get_option)

    def set_max_iterations(self, cap: int):
        """Jacobi iterations after which a pressure solve that has not met its
        termination test fails (the reference has no cap)."""
        self.set_option(N.OPT_MAX_ITERATIONS, int(cap))

    def status(self):
        """(0, "") or (status, message) of the last step's failure."""
        v = C.c_int()
        N.check(N.lib().sph_system_status(self._ctx.h, C.byref(v)))
        msg = N.lib().sph_last_error() if v.value else b""
        return int(v.value), (msg.decode() if msg else "")

    # ---- the SPHSystem trait's accessors (sph/mod.rs:40-71, 107-238, 358-360) --
    def time(self) -> float:
        """``SPHSystem::time`` (sph/mod.rs:107-116)."""
        return float(self.take_measurement()["time"])

    def time_steps_propagated(self) -> int:
        """``SPHSystem::time_steps_propagated`` (sph/mod.rs:118-120)."""
        return int(self.take_measurement()["steps"])

    def get_fluid_ids(self) -> np.ndarray:
        """``SPHSystem::get_fluid_ids`` (sph/mod.rs:186-188)."""
        return self.state()["fluid_id"]

    def get_fluid_pos(self) -> np.ndarray:
        """``SPHSystem::get_fluid_pos`` (sph/mod.rs:190-196): positions as f32."""
        return self.state()["x"].astype(np.float32)

    def get_fluid_checkpoint(self):
        """``SPHSystem::get_fluid_checkpoint`` (sph/mod.rs:198-200)."""
        return self.get_checkpoint().fluid

    def get_boundary_checkpoint(self) -> list:
        """``SPHSystem::get_boundary_checkpoint`` (sph/mod.rs:358-360): one
        entry per boundary, the rigid state of a dynamic one, else None."""
        return self.get_checkpoint().dynamic_states

    def get_quantity_of_fluid_samples(self, quantity) -> np.ndarray:
        """``SPHSystem::get_quantity_of_fluid_samples`` (sph/mod.rs:202-238)."""
        q = N.QUANTITIES[quantity] if not isinstance(quantity, str) else quantity
        return self.quantity_of_fluid_samples(q)

    # ---- checkpoints (sph/mod.rs:362-378, 884-915) -------------------------
    def get_checkpoint(self):
        """``SystemCheckpoint`` of the current state (fluid in the reference's
        particle order, the integrator's slots, and one entry per boundary:
        the RigidBodyMotion state of a dynamic one, None for a static one)."""
        from . import state_io
        st = self.state()
        n = len(st["m"])
        npos, nvel = self.integrator_slot_counts()
        xp, vp = np.zeros((n, 3)), np.zeros((n, 3))
        ns = C.c_int()
        N.check(N.lib().sph_fluid_download_slots(self._ctx.h, C.byref(ns), N.dptr(xp),
                                                 N.dptr(vp)))
        pslots = [xp] if npos else []
        vslots = [vp] if nvel else []
        m = self.take_measurement()
        states = []
        for b, desc in enumerate(self.scene.boundaries):
            if getattr(desc, "rigid", None) is None:
                states.append(None)
                continue
            s = self.body_state(b)
            q = s["q"]  # (w, i, j, k) -> serialized as (i, j, k, w)
            states.append(state_io.RigidBodyCheckpoint(
                s["com"].copy(), np.array([q[1], q[2], q[3], q[0]]), s["lin_vel"].copy(),
                s["ang_mom"].copy(), s["force"].copy(), s["torque"].copy()))
        fluid = state_io.FluidCheckpoint(st["fluid_id"].copy(), st["x"], st["v"], st["m"],
                                         pslots, vslots)
        return state_io.SystemCheckpoint(int(m["steps"]), float(m["time"]), fluid, states)

    def integrator_slot_counts(self):
        """``IntegrationScheme::POSITION_SLOTS`` / ``VELOCITY_SLOTS``
        (integration_schemes/mod.rs:23-42) of this system's integrator."""
        npos, nvel = C.c_int(), C.c_int()
        N.check(N.lib().sph_fluid_slot_counts(self._ctx.h, C.byref(npos), C.byref(nvel)))
        return int(npos.value), int(nvel.value)

    def set_integrator_slots(self, positions=None, velocities=None):
        """The integrator's position and velocity slots, [n][3] in the
        reference's particle order (Verlet's position slot holds the position
        one step back; Fluid::resize_slots starts it at the origin)."""
        def arr(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
        xp, vp = arr(positions), arr(velocities)
        N.check(N.lib().sph_fluid_upload_slots(self._ctx.h, N.dptr(xp) if xp is not None else None,
                                               N.dptr(vp) if vp is not None else None))

    def continue_from_checkpoint(self, cp):
        """``SPHSystem::continue_from_checkpoint`` (sph/mod.rs:362-378) on this
        system: its scene's parameters and boundaries with the checkpoint's
        fluid, counters and rigid-body states, after the update() that resuming
        runs. The device context is built anew (as from_checkpoint does), so
        implementation options set on the old one are not carried over; not
        for a system attached to a world."""
        other = System.from_checkpoint(self.scene, cp)
        self._ctx, other._ctx = other._ctx, self._ctx
        self.scene, self._keep = other.scene, other._keep
        self.__dict__.pop("_wall", None)
        other._ctx.close()

    @classmethod
    def from_checkpoint(cls, scene: Scene, cp) -> "System":
        """``SPHSystem::continue_from_checkpoint``: a system of ``scene``'s
        parameters and boundaries carrying the checkpoint's fluid, counters
        and rigid-body states, after the update() that resuming runs."""
        import dataclasses
        f = cp.fluid
        sc = dataclasses.replace(scene, x=np.asarray(f.position, dtype=np.float64).reshape(-1, 3),
                                 v=np.asarray(f.velocity, dtype=np.float64).reshape(-1, 3),
                                 m=np.asarray(f.mass, dtype=np.float64).ravel(),
                                 fluid_id=np.asarray(f.fluid_id, dtype=np.uint32).ravel())
        sim = cls(sc, init=False)
        # the checkpoint's integrator slots are not read back: both of the
        # reference's resume paths re-size them (sph/mod.rs:368-373, 404-409)
        nb = len(cp.dynamic_states)
        present = np.zeros(max(nb, 1), dtype=np.uint8)
        bodies = np.zeros((max(nb, 1), 19))
        for b, d in enumerate(cp.dynamic_states):
            if d is None:
                continue
            present[b] = 1
            o = np.asarray(d.orientation, dtype=np.float64)  # (i, j, k, w)
            bodies[b] = np.concatenate([d.center_of_mass, [o[3], o[0], o[1], o[2]],
                                        d.linear_velocity, d.angular_momentum, d.force,
                                        d.torque])
        N.check(N.lib().sph_system_restore(sim._ctx.h, int(cp.time_steps_propagated),
                                           float(cp.current_time), nb,
                                           present.ctypes.data_as(N._u8p), N.dptr(bodies)))
        return sim

    def time_step_info(self, fluid=None, boundary=None) -> dict:
        """``TimeStepInfo::from_system`` (render_info.rs:80-89) for a recording:
        fluid = {"kind": "Samples", "coloring": {"kind": "Uniform" | "FluidId" |
        "QuantityGraded", "quantity": ...}} or {"kind": "SensorPlane",
        "quantity": ..., "planes": [{"positions", "rows", "cols"}]};
        boundary = {"kind": "Samples", "positions": [m][3], "coloring": ...}."""
        fluid = fluid or {"kind": "Samples", "coloring": {"kind": "Uniform"}}
        m = self.take_measurement()
        from .state_io import MEASUREMENT_FIELDS
        meas = {k: m[k] for k, _ in MEASUREMENT_FIELDS if k in m}
        meas["time_step_wall_clock_time"] = getattr(self, "_wall", 0.0)
        if fluid["kind"] == "Samples":
            st = self.state()
            col = fluid.get("coloring", {"kind": "Uniform"})
            if col["kind"] == "FluidId":
                col = {"kind": "FluidId", "id": st["fluid_id"],
                       "max_id": int(st["fluid_id"].max()) if len(st["fluid_id"]) else 0}
            elif col["kind"] == "QuantityGraded":
                col = {"kind": "QuantityGraded", "quantity": col["quantity"],
                       "data": self.quantity_of_fluid_samples(col["quantity"], st)}
            fv = {"kind": "Samples", "positions": st["x"].astype(np.float32), "coloring": col}
        else:
            planes = []
            for pl in fluid["planes"]:
                pos = np.asarray(pl["positions"], dtype=np.float32).reshape(-1, 3)
                planes.append({"positions": pos,
                               "data": self.get_quantity_at_positions(fluid["quantity"], pos),
                               "rows": pl["rows"], "cols": pl["cols"]})
            fv = {"kind": "SensorPlane", "quantity": fluid["quantity"], "planes": planes}
        bv = boundary or {"kind": "Samples", "positions": np.zeros((0, 3), dtype=np.float32),
                          "coloring": {"kind": "Uniform"}}
        return {"time_step_number": int(m["steps"]), "measurement": meas, "fluid": fv,
                "boundary": bv}

    def quantity_of_fluid_samples(self, quantity, st=None) -> np.ndarray:
        """``get_quantity_of_fluid_samples`` (sph/mod.rs:202-238): a per-particle
        quantity in the reference's particle order, as f32."""
        st = st if st is not None else self.state()
        V0 = self.scene.params.rest_density_grid_spacing ** 3
        v, V, m, p = st["v"], st["V"], st["m"], st["p"]
        if quantity == "Speed":
            out = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
        elif quantity == "Volume":
            out = V
        elif quantity == "Density":
            out = m / V
        elif quantity == "DensityError":
            out = np.where(V < V0, 100.0 * (V0 / V - 1.0), 0.0)
        elif quantity == "Pressure":
            out = p
        elif quantity == "KineticEnergy":
            out = (0.5 * m) * ((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
        else:
            raise ValueError(f"unknown quantity {quantity!r}")
        return np.asarray(out, dtype=np.float64).astype(np.float32)

    def clone(self) -> "System":
        """Deep copy into an independent device context (DynClone)."""
        h = C.c_void_p()
        N.check(N.lib().sph_ctx_clone(self._ctx.h, C.byref(h)))
        other = System.__new__(System)
        other.scene = self.scene
        other._keep = self._keep
        ctx = _Ctx.__new__(_Ctx)
        ctx._p = self._ctx._p
        ctx.h = h
        other._ctx = ctx
        return other

    # ---- per-trait entry points (host arrays in and out) -------------------
    def get_volume(self, x):
        x = N.f64(x).reshape(-1, 3)
        V = np.zeros(len(x))
        N.check(N.lib().sph_volume(self._ctx.h, len(x), N.dptr(x), N.dptr(V)))
        return V

    def find_boundary_samples(self, b: int, x):
        x = N.f64(x).reshape(-1, 3)
        out = np.zeros((len(x), 7))
        flags = np.zeros(len(x), dtype=np.uint8)
        N.check(N.lib().sph_boundary_find_samples(self._ctx.h, int(b), len(x), N.dptr(x),
                                                  N.dptr(out), flags.ctypes.data_as(N._u8p)))
        return out, flags

    def add_non_pressure_accelerations(self, x, v, m, V):
        x = N.f64(x).reshape(-1, 3)
        v = N.f64(v).reshape(-1, 3)
        m, V = N.f64(m), N.f64(V)
        a = np.zeros_like(x)
        ft = np.zeros(6 * max(len(self.scene.boundaries), 1))
        N.check(N.lib().sph_nonpressure_accel(self._ctx.h, len(x), N.dptr(x), N.dptr(v),
                                              N.dptr(m), N.dptr(V), N.dptr(a), N.dptr(ft)))
        return a, ft.reshape(-1, 6)

    def solve_and_add_acceleration(self, x, v, m, V, a, dt):
        x = N.f64(x).reshape(-1, 3)
        v = N.f64(v).reshape(-1, 3)
        m, V = N.f64(m), N.f64(V)
        a = N.f64(a).reshape(-1, 3).copy()
        p = np.zeros(len(x))
        ft = np.zeros(6 * max(len(self.scene.boundaries), 1))
        xp = np.zeros_like(x)
        vp = np.zeros_like(x)
        info = N.SphMeasurement()
        N.check(N.lib().sph_pressure_solve(self._ctx.h, len(x), N.dptr(x), N.dptr(v), N.dptr(m),
                                           N.dptr(V), float(dt), N.dptr(a), N.dptr(p),
                                           N.dptr(ft), N.dptr(xp), N.dptr(vp), C.byref(info)))
        return {"a": a, "p": p, "body_ft": ft.reshape(-1, 6), "x_pred": xp, "v_pred": vp,
                "info": info.as_dict()}


def _read_profile(fn, handle) -> dict:
    out = {}
    for k, name in enumerate(N.PROF_KERNELS):
        launches, units, ms = C.c_uint64(), C.c_uint64(), C.c_double()
        N.check(fn(handle, k, C.byref(launches), C.byref(units), C.byref(ms)))
        out[name] = {"launches": int(launches.value), "units": int(units.value),
                     "ms": float(ms.value)}
    return out


def _tuning(get) -> dict:
    out = {"state": get(N.OPT_TUNE_STATE), "steps_per_shape": get(N.OPT_AUTOTUNE)}
    for k, name in enumerate(N.PAIR_PASSES):
        out[name] = get(N.OPT_SHAPE_VOLUME + k)
    return out


class World:
    """Many systems advanced together by shared multi-system launches.

    The reference steps one ``System`` at a time on its worker thread
    (``worker_loop``, sci-phi-backend/src/lib.rs:507, spawned by
    sci-phi/src/app/mod.rs:236). A world steps every attached system by
    whole time steps in one sequence of launches; each system's results are
    identical to stepping it alone. Attached systems are stepped only through
    the world (their own ``step_forward_in_time`` raises until detached).
    """

    def __init__(self, systems=()):
        N.lib()
        h = C.c_void_p()
        N.check(N.lib().sph_world_create(C.byref(h)))
        self.h = h
        self.members = []
        for s in systems:
            self.attach(s)

    def attach(self, sim: "System"):
        N.check(N.lib().sph_world_attach(self.h, sim.handle))
        self.members.append(sim)

    def detach(self, sim: "System"):
        N.check(N.lib().sph_world_detach(self.h, sim.handle))
        self.members.remove(sim)

    def step(self, n: int = 1):
        N.check(N.lib().sph_world_step(self.h, int(n)))

    def size(self):
        ns, npart = C.c_size_t(), C.c_size_t()
        N.check(N.lib().sph_world_size(self.h, C.byref(ns), C.byref(npart)))
        return int(ns.value), int(npart.value)

    def set_cache_gradients(self, on: bool):
        N.check(N.lib().sph_world_set_option(self.h, N.OPT_CACHE_GRADIENTS, int(bool(on))))

    def set_option(self, option: int, value: int):
        N.check(N.lib().sph_world_set_option(self.h, int(option), int(value)))

    def get_option(self, option: int) -> int:
        v = C.c_int()
        N.check(N.lib().sph_world_get_option(self.h, int(option), C.byref(v)))
        return int(v.value)

    def tuning(self) -> dict:
        return _tuning(self.get_option)

    def profile(self, enable: bool = True):
        N.check(N.lib().sph_world_profile(self.h, int(bool(enable))))

    def profile_read(self) -> dict:
        return _read_profile(N.lib().sph_world_profile_read, self.h)

    def close(self):
        if getattr(self, "h", None):
            N.lib().sph_world_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_volume_map(sdf, h: float, x_min, n, dx: float, gl_x, gl_w, system=None, stats=None):
    """Volume-map node values and present flags on the lattice (x_min, n, dx)
    from the SDF lattice `sdf` (``DiscretizedBoundaryFields::new``,
    ``volume_map_boundary.rs:1593-1624``), with the Gauss-Legendre rule
    (gl_x, gl_w) on [-1, 1]; runs on the stream of `system` if given.
    `stats["quadrature_nodes"]` receives the number of ball quadratures run."""
    s = _lattice_struct(sdf)
    xm = np.ascontiguousarray(x_min, dtype=np.float64)
    nn = np.ascontiguousarray(n, dtype=np.int64)
    gx = np.ascontiguousarray(gl_x, dtype=np.float64)
    gw = np.ascontiguousarray(gl_w, dtype=np.float64)
    if xm.shape != (3,) or nn.shape != (3,) or gx.shape != gw.shape or gx.ndim != 1:
        raise ValueError("x_min and n need 3 entries; gl_x and gl_w one per node")
    shape = (3 * int(nn[2]) + 1, 3 * int(nn[1]) + 1, 3 * int(nn[0]) + 1)
    vals = np.empty(shape, dtype=np.float64)
    pres = np.empty(shape, dtype=np.uint8)
    handle = system._ctx.h if system is not None else None
    q = C.c_uint64()
    N.check(N.lib().sph_volume_map_build(handle, C.byref(s), float(h), N.dptr(xm),
                                         nn.ctypes.data_as(N._i64p), float(dx), N.dptr(gx),
                                         N.dptr(gw), len(gx), N.dptr(vals),
                                         pres.ctypes.data_as(N._u8p), C.byref(q)))
    if stats is not None:
        stats["quadrature_nodes"] = int(q.value)
    return vals, pres


class Server:
    """Continuous batching in the library (sph_server_*): FIFO admission under
    slots and a device-memory budget, admitted systems stepped together in one
    world per pressure solver, finished systems handed back. Admission and its bookkeeping touch
    no device; a request's System is built after it is admitted."""

    def __init__(self, budget_bytes: int, max_slots: int):
        h = C.c_void_p()
        N.check(N.lib().sph_server_create(int(budget_bytes), int(max_slots), C.byref(h)))
        self.h = h
        self._sims = {}

    def submit(self, device_bytes: int, steps: int) -> int:
        t = C.c_uint64()
        N.check(N.lib().sph_server_submit(self.h, int(device_bytes), int(steps), C.byref(t)))
        return int(t.value)

    def admit(self):
        """The next admitted ticket, or None while the queue's head does not fit."""
        t = C.c_uint64()
        st = N.lib().sph_server_admit(self.h, C.byref(t))
        N.check(st if st < 0 else 0)
        return int(t.value) if st == 1 else None

    def attach(self, ticket: int, sim: "System"):
        N.check(N.lib().sph_server_attach(self.h, int(ticket), sim.handle))
        self._sims[int(ticket)] = sim

    def cancel(self, ticket: int):
        """Give up an admitted ticket whose System could not be built."""
        N.check(N.lib().sph_server_cancel(self.h, int(ticket)))

    def step(self) -> dict:
        out = N.SphServerStats()
        N.check(N.lib().sph_server_step(self.h, C.byref(out)))
        return out.as_dict()

    def set_option(self, option: int, value: int):
        """An implementation option (N.OPT_*) of the server's worlds."""
        N.check(N.lib().sph_server_set_option(self.h, int(option), int(value)))

    def poll(self):
        """(ticket, System) of the oldest finished request, or None."""
        t, c = C.c_uint64(), C.c_void_p()
        st = N.lib().sph_server_poll(self.h, C.byref(t), C.byref(c))
        N.check(st if st < 0 else 0)
        if st != 1:
            return None
        sim = self._sims.pop(int(t.value), None)
        return int(t.value), sim

    def poll_failed(self):
        """(ticket, System, status) of the oldest failed request, or None; the
        System's state is not valid."""
        t, c, f = C.c_uint64(), C.c_void_p(), C.c_int()
        st = N.lib().sph_server_poll_failed(self.h, C.byref(t), C.byref(c), C.byref(f))
        N.check(st if st < 0 else 0)
        if st != 1:
            return None
        sim = self._sims.pop(int(t.value), None)
        return int(t.value), sim, int(f.value)

    def stats(self) -> dict:
        out = N.SphServerStats()
        N.check(N.lib().sph_server_read_stats(self.h, C.byref(out)))
        return out.as_dict()

    def close(self):
        if self.h:
            N.lib().sph_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_mesh_sdf(vertices, triangles, x_min, n, dx: float, prune: float, system=None):
    """Signed-distance node values and present flags on the lattice (x_min, n,
    dx) from a triangle mesh (``DiscretizedBoundaryFields::new``,
    ``volume_map_boundary.rs:1547-1591``): vertices [nv][3], triangles [nt][3]
    with outward winding; cells whose node values all lie beyond +-prune are
    pruned. Runs on the stream of `system` if given."""
    v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
    xm = np.ascontiguousarray(x_min, dtype=np.float64)
    nn = np.ascontiguousarray(n, dtype=np.int64)
    if xm.shape != (3,) or nn.shape != (3,):
        raise ValueError("x_min and n need 3 entries")
    shape = (3 * int(nn[2]) + 1, 3 * int(nn[1]) + 1, 3 * int(nn[0]) + 1)
    vals = np.empty(shape, dtype=np.float64)
    pres = np.empty(shape, dtype=np.uint8)
    handle = system._ctx.h if system is not None else None
    N.check(N.lib().sph_mesh_sdf_build(handle, N.dptr(v), len(v), t.ctypes.data_as(N._u32p),
                                       len(t), N.dptr(xm), nn.ctypes.data_as(N._i64p),
                                       float(dx), float(prune), N.dptr(vals),
                                       pres.ctypes.data_as(N._u8p)))
    return vals, pres


def device_bytes_held() -> int:
    """Device bytes the library's buffers hold now (every context, world, server)."""
    b = C.c_uint64()
    N.check(N.lib().sph_device_bytes_held(C.byref(b)))
    return int(b.value)


def device_memory():
    """(free, total) bytes of device memory."""
    f, t = C.c_size_t(), C.c_size_t()
    N.check(N.lib().sph_device_memory(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)

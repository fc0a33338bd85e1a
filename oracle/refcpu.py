"""ctypes binding of the CPU oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / baseline. The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefcpu.so")

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class RefLattice(C.Structure):
    _fields_ = [("x_min", C.c_double * 3), ("x_max", C.c_double * 3), ("dx", C.c_double),
                ("n", C.c_int64 * 3), ("values", _dp), ("present", _u8p)]


class RefParams(C.Structure):
    _fields_ = [("time_increment", C.c_double), ("cfl", C.c_int32),
                ("max_time_increment", C.c_double), ("cfl_number", C.c_double),
                ("kernel_support_radius", C.c_double), ("rest_density_grid_spacing", C.c_double),
                ("disable_particles_below", C.c_double), ("fluid_viscosity", C.c_double),
                ("boundary_viscosity", C.c_double),
                ("boundary_pressure_acceleration_weighting", C.c_double),
                ("boundary_rest_volume_weighting", C.c_double), ("gravity_mode", C.c_int32),
                ("gravity_center", C.c_double * 3), ("solver", C.c_int32),
                ("integrator", C.c_int32), ("stiffness", C.c_double),
                ("target_density_error", C.c_double), ("relaxation_factor", C.c_double),
                ("min_diagonal_element", C.c_double)]


class RefMeasurement(C.Structure):
    _fields_ = [("time", C.c_double), ("time_step_size", C.c_double), ("density", C.c_double),
                ("density_error", C.c_double), ("kinetic_energy", C.c_double),
                ("solver_iterations", C.c_uint32), ("eqs1_iterations", C.c_uint32),
                ("predicted_density_error", C.c_double), ("steps", C.c_uint64),
                ("num_particles", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefRigid(C.Structure):
    _fields_ = [("mass", C.c_double), ("inertia_inv_body", C.c_double * 9),
                ("com", C.c_double * 3), ("q", C.c_double * 4), ("lin_vel", C.c_double * 3),
                ("ang_mom", C.c_double * 3), ("force", C.c_double * 3),
                ("torque", C.c_double * 3), ("inertia_inv_world", C.c_double * 9),
                ("ang_vel", C.c_double * 3)]


# single reference functions and fields of the test hooks (refcpu.h)
FUNCTIONS = ["neighbors", "volume", "reset_acceleration", "gravity", "viscosity",
             "pressure_acceleration", "pressure_acceleration_overwrite",
             "pressure_acceleration_scratch", "sesph", "pred_vel", "source_vde", "source_vp",
             "diagonal", "initialize", "resolve", "sesph_splitting", "integrate"]
FIELDS = {"x": 3, "v": 3, "a": 3, "m": 1, "V": 1, "p": 1, "pred_v": 3, "pred_x": 3, "s": 1,
          "a_ii": 1, "pacc": 3, "dt": 0, "slot_x": 3}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ref_w.restype = C.c_double
        L.ref_w.argtypes = [C.c_double]
        L.ref_dqw.restype = C.c_double
        L.ref_dqw.argtypes = [C.c_double]
        L.ref_kernel_function.restype = C.c_double
        L.ref_kernel_function.argtypes = [_dp, C.c_double]
        L.ref_kernel_gradient.argtypes = [_dp, C.c_double, _dp]
        L.ref_get_cell.argtypes = [_dp, C.c_double, C.POINTER(C.c_int32)]
        L.ref_map_int_to_uint.restype = C.c_uint64
        L.ref_map_int_to_uint.argtypes = [C.c_int32]
        L.ref_hash.restype = C.c_uint64
        L.ref_hash.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        L.ref_find_samples_alloc.restype = C.c_void_p
        L.ref_find_samples_alloc.argtypes = [C.c_double, C.c_double, _dp, C.c_size_t, _dp,
                                             C.c_size_t]
        L.ref_nlist_total.restype = C.c_size_t
        L.ref_nlist_total.argtypes = [C.c_void_p]
        L.ref_nlist_copy.argtypes = [C.c_void_p, _u64p, _u64p]
        L.ref_nlist_destroy.argtypes = [C.c_void_p]
        L.ref_serendipity_function.restype = C.c_int
        L.ref_serendipity_function.argtypes = [C.POINTER(RefLattice), _dp, _dp]
        L.ref_serendipity_gradient.restype = C.c_int
        L.ref_serendipity_gradient.argtypes = [C.POINTER(RefLattice), _dp, _dp]
        L.ref_shape_functions.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.ref_reference_nodes.argtypes = [_dp]
        L.ref_node_offset.restype = C.c_int
        L.ref_node_offset.argtypes = [C.c_double]
        L.ref_quat_rotate.argtypes = [_dp, _dp, _dp]
        L.ref_quat_to_matrix.argtypes = [_dp, _dp]
        L.ref_system_create.restype = C.c_void_p
        L.ref_system_create.argtypes = [C.POINTER(RefParams), C.c_size_t, _dp, _dp, _dp, _u32p]
        L.ref_system_destroy.argtypes = [C.c_void_p]
        L.ref_system_add_static_samples.restype = C.c_int
        L.ref_system_add_static_samples.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp]
        L.ref_system_add_volume_map.restype = C.c_int
        L.ref_system_add_volume_map.argtypes = [C.c_void_p, C.POINTER(RefLattice),
                                                C.POINTER(RefLattice), _dp]
        L.ref_system_set_aabb.restype = C.c_int
        L.ref_system_set_aabb.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.ref_system_set_contact.restype = C.c_int
        L.ref_system_set_contact.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.ref_system_contact_stats.argtypes = [C.c_void_p, _u32p]
        L.ref_system_contact_cells.restype = C.c_size_t
        L.ref_system_contact_cells.argtypes = [C.c_void_p, _dp, C.c_size_t]
        L.ref_contact_solve.restype = C.c_uint32
        L.ref_contact_solve.argtypes = [_dp, C.c_int, _dp, C.c_size_t, C.c_double, _dp]
        L.ref_mesh_create.restype = C.c_void_p
        L.ref_mesh_create.argtypes = [_dp, C.c_size_t, _u32p, C.c_size_t]
        L.ref_mesh_destroy.argtypes = [C.c_void_p]
        L.ref_mesh_edges.restype = C.c_size_t
        L.ref_mesh_edges.argtypes = [C.c_void_p]
        L.ref_mesh_sdf.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp]
        L.ref_mesh_sdf_build.argtypes = [C.c_void_p, _dp, _i64p, C.c_double, C.c_double, _dp,
                                         _u8p]
        L.ref_system_init.restype = C.c_int
        L.ref_system_init.argtypes = [C.c_void_p]
        L.ref_system_step.restype = C.c_int
        L.ref_system_step.argtypes = [C.c_void_p, C.c_int]
        L.ref_system_len.restype = C.c_size_t
        L.ref_system_len.argtypes = [C.c_void_p]
        L.ref_system_get.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _u32p, _u64p]
        L.ref_system_measure.argtypes = [C.c_void_p, C.POINTER(RefMeasurement)]
        L.ref_system_error_history.restype = C.c_size_t
        L.ref_system_error_history.argtypes = [C.c_void_p, _dp, C.c_size_t]
        L.ref_system_body_state.restype = C.c_int
        L.ref_system_body_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref_system_body_forces.restype = C.c_int
        L.ref_system_body_forces.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref_system_fluid_nlist.restype = C.c_void_p
        L.ref_system_fluid_nlist.argtypes = [C.c_void_p]
        L.ref_system_boundary_samples.restype = C.c_size_t
        L.ref_system_boundary_samples.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_int,
                                                  _dp, C.c_size_t]
        L.ref_system_call.restype = C.c_int
        L.ref_system_call.argtypes = [C.c_void_p, C.c_int]
        L.ref_system_field.restype = C.c_int
        L.ref_system_field.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.ref_system_sample_volumes.restype = C.c_int
        L.ref_system_sample_volumes.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.ref_rigid_init.argtypes = [C.POINTER(RefRigid), C.c_double, _dp, _dp, _dp, _dp, _dp,
                                     _dp]
        L.ref_rigid_add_force.argtypes = [C.POINTER(RefRigid), _dp, _dp]
        L.ref_rigid_step.argtypes = [C.POINTER(RefRigid), C.c_double]
        L.ref_rigid_velocity_at_point.argtypes = [C.POINTER(RefRigid), _dp, _dp]
        L.ref_system_static_volume.restype = C.c_double
        L.ref_system_static_volume.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.ref_system_set_threads.argtypes = [C.c_int]
        L.ref_system_quantity_at_positions.restype = C.c_int
        L.ref_system_quantity_at_positions.argtypes = [C.c_void_p, C.c_int, C.c_size_t,
                                                       C.POINTER(C.c_float),
                                                       C.POINTER(C.c_float)]
        L.ref_vm_integrate_ball.restype = C.c_int
        L.ref_vm_integrate_ball.argtypes = [C.POINTER(RefLattice), _dp, C.c_double, _dp, _dp,
                                            C.c_int, _dp]
        L.ref_vm_node_value.restype = C.c_int
        L.ref_vm_node_value.argtypes = [C.POINTER(RefLattice), _dp, C.c_double, _dp, _dp,
                                        C.c_int, _dp]
        L.ref_vm_build.restype = C.c_uint64
        L.ref_vm_build.argtypes = [C.POINTER(RefLattice), C.c_double, _dp, _i64p, C.c_double,
                                   _dp, _dp, C.c_int, _dp, _u8p]
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(_dp)


def f64c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# -------------------------------------------------------------- primitives
def kernel_function(r, h):
    r = f64c(r)
    return lib().ref_kernel_function(dptr(r), h)


def kernel_gradient(r, h):
    r = f64c(r)
    out = np.zeros(3)
    lib().ref_kernel_gradient(dptr(r), h, dptr(out))
    return out


def get_cell(p, cs):
    p = f64c(p)
    out = (C.c_int32 * 3)()
    lib().ref_get_cell(dptr(p), cs, out)
    return tuple(out)


def find_samples(cell_size, rng, pos, samples):
    """SpatialHashing::find_samples -> (offsets, indices) as uint64 arrays."""
    pos = f64c(pos).reshape(-1, 3)
    samples = f64c(samples).reshape(-1, 3)
    L = lib()
    h = L.ref_find_samples_alloc(cell_size, rng, dptr(pos), len(pos), dptr(samples), len(samples))
    total = L.ref_nlist_total(h)
    off = np.zeros(len(pos) + 1, dtype=np.uint64)
    idx = np.zeros(max(total, 1), dtype=np.uint64)
    L.ref_nlist_copy(h, off.ctypes.data_as(_u64p), idx.ctypes.data_as(_u64p))
    L.ref_nlist_destroy(h)
    return off, idx[:total]


def lattice_struct(lat):
    s = RefLattice()
    for d in range(3):
        s.x_min[d] = float(lat.x_min[d])
        s.x_max[d] = float(lat.x_max[d])
        s.n[d] = int(lat.n[d])
    s.dx = float(lat.dx)
    vals = np.ascontiguousarray(lat.values, dtype=np.float64)
    pres = np.ascontiguousarray(lat.present, dtype=np.uint8)
    s.values = dptr(vals)
    s.present = pres.ctypes.data_as(_u8p)
    s._keep = (vals, pres)
    return s


def serendipity_function(lat, p):
    s = lattice_struct(lat)
    p = f64c(p)
    out = C.c_double()
    st = lib().ref_serendipity_function(C.byref(s), dptr(p), C.byref(out))
    return st, out.value


def serendipity_gradient(lat, p):
    s = lattice_struct(lat)
    p = f64c(p)
    out = np.zeros(3)
    st = lib().ref_serendipity_gradient(C.byref(s), dptr(p), dptr(out))
    return st, out


def vm_integrate_ball(sdf, center, h, gl_x, gl_w):
    """gauss_legendre_integrate_ball of the cubic extension of `sdf`; returns
    (status, value), status nonzero when an SDF evaluation failed."""
    s = lattice_struct(sdf)
    c, gx, gw = f64c(center), f64c(gl_x), f64c(gl_w)
    out = C.c_double()
    st = lib().ref_vm_integrate_ball(C.byref(s), dptr(c), float(h), dptr(gx), dptr(gw),
                                     len(gx), C.byref(out))
    return st, out.value


def vm_node_value(sdf, p, h, gl_x, gl_w):
    """SDFnWrapper::volume clamped to [0, ball volume]; (status, value)."""
    s = lattice_struct(sdf)
    c, gx, gw = f64c(p), f64c(gl_x), f64c(gl_w)
    out = C.c_double()
    st = lib().ref_vm_node_value(C.byref(s), dptr(c), float(h), dptr(gx), dptr(gw), len(gx),
                                 C.byref(out))
    return st, out.value


def vm_build(sdf, h, x_min, n, dx, gl_x, gl_w, stats=None):
    """Volume-map node values and present flags, shape (3n_z+1, 3n_y+1, 3n_x+1);
    `stats["quadrature_nodes"]` receives the number of ball quadratures run."""
    s = lattice_struct(sdf)
    xm, gx, gw = f64c(x_min), f64c(gl_x), f64c(gl_w)
    nn = np.ascontiguousarray(n, dtype=np.int64)
    shape = (3 * int(nn[2]) + 1, 3 * int(nn[1]) + 1, 3 * int(nn[0]) + 1)
    vals = np.zeros(shape, dtype=np.float64)
    pres = np.zeros(shape, dtype=np.uint8)
    q = lib().ref_vm_build(C.byref(s), float(h), dptr(xm), nn.ctypes.data_as(_i64p), float(dx),
                           dptr(gx), dptr(gw), len(gx), dptr(vals), pres.ctypes.data_as(_u8p))
    if stats is not None:
        stats["quadrature_nodes"] = int(q)
    return vals, pres


class Mesh:
    """An oriented triangle mesh with its pseudo-normals (ref_mesh)."""

    def __init__(self, vertices, triangles):
        self.v = f64c(np.asarray(vertices).reshape(-1, 3))
        self.t = np.ascontiguousarray(np.asarray(triangles).reshape(-1, 3), dtype=np.uint32)
        self.h = lib().ref_mesh_create(dptr(self.v), len(self.v),
                                       self.t.ctypes.data_as(_u32p), len(self.t))
        if not self.h:
            raise ValueError("triangle index out of range")

    def __del__(self):
        try:
            lib().ref_mesh_destroy(self.h)
        except Exception:
            pass

    @property
    def edges(self):
        return int(lib().ref_mesh_edges(self.h))

    def signed_distance(self, points):
        p = f64c(np.asarray(points).reshape(-1, 3))
        out = np.zeros(len(p))
        lib().ref_mesh_sdf(self.h, dptr(p), len(p), dptr(out))
        return out

    def sdf_lattice(self, x_min, n, dx, prune):
        """(values, present) of the dense (3n+1)^3 nodes, z-major."""
        x_min = f64c(x_min)
        n = np.ascontiguousarray(n, dtype=np.int64)
        shape = tuple(int(3 * k + 1) for k in n[::-1])
        values = np.zeros(shape)
        present = np.zeros(shape, dtype=np.uint8)
        lib().ref_mesh_sdf_build(self.h, dptr(x_min), n.ctypes.data_as(_i64p), float(dx),
                                 float(prune), dptr(values), present.ctypes.data_as(_u8p))
        return values, present


# contact-cell record: body a, body b, volume, center[3], f_a[3], tau_a[3],
# f_b[3], tau_b[3], a_cc pressure, a_cc friction, p, lambda[3]
CONTACT_REC = 24
# contact body record: dynamic, mass, com[3], v[3], L[3], F[3], tau[3], I^-1 world[9]
CONTACT_BODY = 26


def contact_solve(bodies, cells, dt):
    """solve_for_p_c after compute_diagonal on cell and body records; returns
    (cells with p and lambda, extra force/torque per body (nb, 6), iterations)."""
    b = f64c(np.asarray(bodies, dtype=np.float64).reshape(-1, CONTACT_BODY))
    c = f64c(np.array(cells, dtype=np.float64).reshape(-1, CONTACT_REC))
    extra = np.zeros((len(b), 6))
    it = lib().ref_contact_solve(dptr(b), len(b), dptr(c), len(c), float(dt), dptr(extra))
    return c, extra, int(it)


def params_struct(P):
    from_solver = {"SESPH": 0, "IISPH": 1, "IISPHwOST": 2, "SESPHwSplitting": 3}
    from_integ = {"EulerCromer": 0, "TakePredicted": 1, "ExplicitEuler": 2, "Verlet": 3}
    s = RefParams()
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
    s.solver = from_solver[P.solver]
    s.integrator = from_integ[P.integrator]
    s.stiffness = P.stiffness
    s.target_density_error = P.target_density_error
    s.relaxation_factor = P.relaxation_factor
    s.min_diagonal_element = P.min_diagonal_element
    return s


class OracleSystem:
    """The reference ``System<K,I,P,N,B>`` restated on the CPU (f64)."""

    def __init__(self, scene, threads=0):
        L = lib()
        if threads:
            L.ref_system_set_threads(int(threads))
        self._p = params_struct(scene.params)
        x, v, m = f64c(scene.x), f64c(scene.v), f64c(scene.m)
        fid = np.ascontiguousarray(scene.fluid_id, dtype=np.uint32)
        self.h = L.ref_system_create(C.byref(self._p), len(x), dptr(x), dptr(v), dptr(m),
                                     fid.ctypes.data_as(_u32p))
        self.nb = 0
        for b in scene.boundaries:
            rig = None
            if getattr(b, "rigid", None) is not None:
                rig = f64c(b.rigid.packed())
            rp = dptr(rig) if rig is not None else None
            if hasattr(b, "positions"):
                pos = f64c(b.positions)
                L.ref_system_add_static_samples(self.h, dptr(pos), len(pos), rp)
            else:
                sdf = lattice_struct(b.sdf)
                vm = lattice_struct(b.vm)
                L.ref_system_add_volume_map(self.h, C.byref(sdf), C.byref(vm), rp)
            self.nb += 1
        vms = [b for b in scene.boundaries if not hasattr(b, "positions")]
        for i, b in enumerate(scene.boundaries):
            if getattr(b, "aabb", None) is not None:
                c, hw = f64c(b.aabb[0]), f64c(b.aabb[1])
                L.ref_system_set_aabb(self.h, i, dptr(c), dptr(hw))
        degree = int(getattr(scene, "contact_degree", 0) or 0)
        if degree and len(vms) == len(scene.boundaries):
            gx, gw = np.polynomial.legendre.leggauss(degree)
            gx, gw = f64c(gx), f64c(gw)
            L.ref_system_set_contact(self.h, dptr(gx), dptr(gw), degree)
        L.ref_system_init(self.h)

    def __del__(self):
        try:
            lib().ref_system_destroy(self.h)
        except Exception:
            pass

    def step(self, n=1):
        lib().ref_system_step(self.h, int(n))

    def quantity_at_positions(self, quantity, positions):
        """get_quantity_at_positions: f32 values of quantity 0..5 at f32 positions."""
        q = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(len(q), dtype=np.float32)
        st = lib().ref_system_quantity_at_positions(
            self.h, int(quantity), len(q), q.ctypes.data_as(C.POINTER(C.c_float)),
            out.ctypes.data_as(C.POINTER(C.c_float)))
        if st:
            raise ValueError(f"quantity {quantity} refused")
        return out

    def contact_stats(self):
        """(active contact cells, NNCG iterations) of the last step."""
        out = np.zeros(2, dtype=np.uint32)
        lib().ref_system_contact_stats(self.h, out.ctypes.data_as(_u32p))
        return int(out[0]), int(out[1])

    def contact_cells(self):
        """The last step's active contact cells, CONTACT_REC doubles each."""
        n = int(lib().ref_system_contact_cells(self.h, None, 0))
        out = np.zeros((n, CONTACT_REC))
        if n:
            lib().ref_system_contact_cells(self.h, dptr(out), n)
        return out

    def __len__(self):
        return int(lib().ref_system_len(self.h))

    def state(self):
        n = len(self)
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "a")}
        for k in ("V", "p", "m"):
            out[k] = np.zeros(n)
        out["fluid_id"] = np.zeros(n, dtype=np.uint32)
        out["origin"] = np.zeros(n, dtype=np.uint64)
        lib().ref_system_get(self.h, dptr(out["x"]), dptr(out["v"]), dptr(out["a"]),
                             dptr(out["V"]), dptr(out["p"]), dptr(out["m"]),
                             out["fluid_id"].ctypes.data_as(_u32p),
                             out["origin"].ctypes.data_as(_u64p))
        return out

    def measure(self):
        m = RefMeasurement()
        lib().ref_system_measure(self.h, C.byref(m))
        return m.as_dict()

    def error_history(self):
        buf = np.zeros(4096)
        k = lib().ref_system_error_history(self.h, dptr(buf), 4096)
        return buf[:min(k, 4096)].copy()

    def body_state(self, b):
        out = np.zeros(22)
        lib().ref_system_body_state(self.h, b, dptr(out))
        return {"com": out[0:3], "q": out[3:7], "lin_vel": out[7:10], "ang_mom": out[10:13],
                "force": out[13:16], "torque": out[16:19], "ang_vel": out[19:22]}

    def body_forces(self, b):
        out = np.zeros(6)
        lib().ref_system_body_forces(self.h, b, dptr(out))
        return out

    def fluid_neighbors(self):
        L = lib()
        h = L.ref_system_fluid_nlist(self.h)
        total = L.ref_nlist_total(h)
        off = np.zeros(len(self) + 1, dtype=np.uint64)
        idx = np.zeros(max(total, 1), dtype=np.uint64)
        L.ref_nlist_copy(h, off.ctypes.data_as(_u64p), idx.ctypes.data_as(_u64p))
        return off, idx[:total]

    def boundary_samples(self, b, i, mode=0):
        buf = np.zeros(7 * 64)
        k = lib().ref_system_boundary_samples(self.h, b, i, mode, dptr(buf), 64)
        return buf[:7 * min(k, 64)].reshape(-1, 7)

    def call(self, fn):
        """Run one reference function on the current state (refcpu.h test hooks)."""
        return int(lib().ref_system_call(self.h, FUNCTIONS.index(fn)))

    def field(self, name):
        w = FIELDS[name]
        k = 1 if w == 0 else w * len(self)
        out = np.zeros(k)
        lib().ref_system_field(self.h, list(FIELDS).index(name), 0, dptr(out))
        return float(out[0]) if w == 0 else out.reshape(-1, 3) if w == 3 else out

    def set_field(self, name, values):
        w = FIELDS[name]
        buf = f64c(np.asarray(values, dtype=np.float64).reshape(-1))
        k = 1 if w == 0 else w * len(self)
        if len(buf) != k:
            raise ValueError(f"{name}: {len(buf)} values, need {k}")
        lib().ref_system_field(self.h, list(FIELDS).index(name), 1, dptr(buf))

    def sample_volumes(self, b, values=None):
        if values is None:
            out = np.zeros(4096)
            k = lib().ref_system_sample_volumes(self.h, b, 0, dptr(out))
            return out[:k].copy()
        buf = f64c(values)
        lib().ref_system_sample_volumes(self.h, b, 1, dptr(buf))

    def static_volume(self, b, k):
        return lib().ref_system_static_volume(self.h, b, k)


class Rigid:
    """RigidBodyMotion of the oracle (ref_rigid), for the rigid-body pins."""

    def __init__(self, mass, inertia, com, q=(1.0, 0.0, 0.0, 0.0), lin_vel=(0, 0, 0),
                 ang_vel=(0, 0, 0)):
        I = f64c(np.asarray(inertia, dtype=np.float64).reshape(9))
        Ii = f64c(np.linalg.inv(I.reshape(3, 3)).reshape(9))
        self.s = RefRigid()
        args = [f64c(a) for a in (com, q, lin_vel, ang_vel)]
        lib().ref_rigid_init(C.byref(self.s), float(mass), dptr(I), dptr(Ii),
                             *[dptr(a) for a in args])
        self._keep = (I, Ii, args)

    def add_force(self, f, loc):
        f, loc = f64c(f), f64c(loc)
        lib().ref_rigid_add_force(C.byref(self.s), dptr(f), dptr(loc))

    def step(self, dt):
        lib().ref_rigid_step(C.byref(self.s), float(dt))

    def velocity_at_point(self, p):
        p, out = f64c(p), np.zeros(3)
        lib().ref_rigid_velocity_at_point(C.byref(self.s), dptr(p), dptr(out))
        return out

    def __getattr__(self, k):
        return np.array(getattr(self.s, k))

"""Seeded scene construction for the SPH hot path.

The reference builds its scenes from triangle meshes (parry3d SDF queries,
mesh sampling, mass properties; ``simulation_lib/src/sph/setup/mod.rs:118-294``).
That setup path is outside the hot path, so scenes here are generated directly:

* fluid particles on a regular lattice at spacing ``dx`` with seeded uniform
  jitter (the reference's ``sample_volume_shifted`` also yields a lattice,
  ``utilities/sampling.rs:8-33``), mass ``rho0 * dx**3``
  (``sph/fluid.rs:87``);
* ``StaticSampleBoundary`` samples on the same lattice, as layers outside the
  container box (the reference samples mesh surfaces at ``0.7 dx``,
  ``static_sample_boundary.rs:63``; the boundary positions are therefore an
  input of the hot path here);
* ``VolumeMapBoundary`` fields on the reference's cubic-serendipity lattice rule
  (cell size ``4 dx``, SDF padding ``3.1 h`` pruned at +-padding, volume-map
  padding ``2 h + cell/6``; ``volume_map_boundary.rs:1120-1137, 1547-1636``;
  ``utilities/discretization.rs:99-245``), with nodal values taken from the
  analytic signed distance of a box or sphere and a volume map integrated
  exactly for the local half-space (the reference integrates the discretized
  SDF over the kernel ball with degree-30 Gauss-Legendre,
  ``volume_map_boundary.rs:1684-1714``);
* rigid-body mass properties in closed form for boxes and spheres.

Both the HIP path and the CPU oracle consume the identical arrays, so per-step
parity concerns the hot path alone.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

SOLVERS = {"SESPH": 0, "IISPH": 1, "IISPHwOST": 2, "SESPHwSplitting": 3}
INTEGRATORS = {"EulerCromer": 0, "TakePredicted": 1, "ExplicitEuler": 2, "Verlet": 3}

# to_offset() of the 32 reference nodes (utilities/discretization.rs:262-300):
# 8 corners, then the x-, y- and z-edge nodes
NODE_OFFSETS = np.array(
    [[0, 0, 0], [3, 0, 0], [0, 3, 0], [3, 3, 0], [0, 0, 3], [3, 0, 3], [0, 3, 3], [3, 3, 3],
     [1, 0, 0], [2, 0, 0], [1, 3, 0], [2, 3, 0], [1, 0, 3], [2, 0, 3], [1, 3, 3], [2, 3, 3],
     [0, 1, 0], [0, 2, 0], [3, 1, 0], [3, 2, 0], [0, 1, 3], [0, 2, 3], [3, 1, 3], [3, 2, 3],
     [0, 0, 1], [0, 0, 2], [3, 0, 1], [3, 0, 2], [0, 3, 1], [0, 3, 2], [3, 3, 1], [3, 3, 2]],
    dtype=np.int64)


@dataclasses.dataclass
class Params:
    """Mirror of the reference's ``[parameters]`` + ``[procedures]``
    (``sph/setup/input.rs:103-260``) restricted to what the hot path reads."""
    kernel_support_radius: float
    rest_density_grid_spacing: float
    time_increment: float = 1e-3
    cfl: bool = False
    max_time_increment: float = 1e-3
    cfl_number: float = 0.4
    disable_particles_below: float = -1e30
    fluid_viscosity: float = 0.01
    boundary_viscosity: float = 0.01
    boundary_pressure_acceleration_weighting: float = 1.0
    boundary_rest_volume_weighting: float = 1.0
    gravity_mode: int = 0          # 0 uniform -z, 1 radial
    gravity_center: tuple = (0.0, 0.0, 0.0)
    solver: str = "IISPH"
    integrator: str = "EulerCromer"
    stiffness: float = 1000.0
    target_density_error: float = 0.1   # percent
    relaxation_factor: float = 0.5
    min_diagonal_element: float = 1e-10


@dataclasses.dataclass
class Lattice:
    """Dense cubic-serendipity node lattice (``CubicSerendipityDiscretization``).

    ``values``/``present`` have shape ``(3*n[2]+1, 3*n[1]+1, 3*n[0]+1)`` (x
    fastest); ``present`` marks nodes that survive the reference's cell pruning.
    """
    x_min: np.ndarray
    x_max: np.ndarray
    dx: float
    n: np.ndarray
    values: np.ndarray
    present: np.ndarray


@dataclasses.dataclass
class Rigid:
    """Initial state of a ``RigidBodyMotion`` (``rigid_body_motion.rs:38-62``)."""
    mass: float
    inertia_body: np.ndarray      # 3x3
    com: np.ndarray
    q: np.ndarray = dataclasses.field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))
    lin_vel: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    ang_vel: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))

    def packed(self) -> np.ndarray:
        """mass, I_body[9], I_inv_body[9], com[3], q[4], v[3], omega[3]."""
        ib = np.asarray(self.inertia_body, dtype=np.float64)
        return np.concatenate([[self.mass], ib.ravel(), np.linalg.inv(ib).ravel(),
                               self.com, self.q, self.lin_vel, self.ang_vel]).astype(np.float64)


@dataclasses.dataclass
class StaticSamples:
    positions: np.ndarray                # (m,3) world frame, or body frame if rigid
    rigid: Optional[Rigid] = None


@dataclasses.dataclass
class VolumeMap:
    sdf: Lattice
    vm: Lattice
    rigid: Optional[Rigid] = None
    kind: str = "box"                    # description only
    # body-frame bounding box (center, half extents) of the boundary's mesh
    # (LocalAABB, volume_map_boundary.rs:1137-1150); None keeps the boundary
    # out of rigid-rigid contact
    aabb: Optional[tuple] = None


def aabb_of(lo, hi) -> tuple:
    """(center, half extents) of the box [lo, hi] (parry's Aabb::center and
    half_extents)."""
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    return ((lo + hi) * 0.5, (hi - lo) * 0.5)


# Gauss-Legendre degree of the contact-cell quadrature
# (VolumeMapBoundary::QUADRATURE_DEGREE, volume_map_boundary.rs:57)
CONTACT_DEGREE = 8


def contact_rule(degree: int = CONTACT_DEGREE):
    """Nodes (ascending on [-1, 1]) and weights of the contact quadrature."""
    return np.polynomial.legendre.leggauss(degree)


@dataclasses.dataclass
class Scene:
    name: str
    params: Params
    x: np.ndarray
    v: np.ndarray
    m: np.ndarray
    fluid_id: np.ndarray
    boundaries: List[object]
    # rigid-rigid contact quadrature degree of a volume-map system (the
    # reference always runs it; 0 turns it off)
    contact_degree: int = CONTACT_DEGREE

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


# --------------------------------------------------------------------------
# signed distance functions (positive on the fluid side)
# --------------------------------------------------------------------------
def sd_box_solid(p, lo, hi):
    """Exact signed distance of a solid box, positive outside."""
    c = (np.asarray(lo) + np.asarray(hi)) * 0.5
    half = (np.asarray(hi) - np.asarray(lo)) * 0.5
    q = np.abs(p - c) - half
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside


def sd_box_container(p, lo, hi):
    return -sd_box_solid(p, lo, hi)


def sd_sphere_solid(p, c, r):
    return np.linalg.norm(p - np.asarray(c), axis=-1) - r


def sd_sphere_container(p, c, r):
    return -sd_sphere_solid(p, c, r)


# --------------------------------------------------------------------------
# lattice construction (reference rule, analytic node values)
# --------------------------------------------------------------------------
def _node_positions(x_min, n, dx):
    axes = [x_min[d] + (np.arange(3 * n[d] + 1, dtype=np.float64) / 3.0) * dx for d in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x, y, z], axis=-1)


def _cell_node_values(values, n):
    """(32, nz, ny, nx) view of each cell's node values."""
    nx, ny, nz = (int(v) for v in n)
    out = np.empty((32, nz, ny, nx), dtype=values.dtype)
    for k, (ox, oy, oz) in enumerate(NODE_OFFSETS):
        out[k] = values[oz:oz + 3 * nz:3, oy:oy + 3 * ny:3, ox:ox + 3 * nx:3]
    return out


def _present_from_kept(keep, n):
    nx, ny, nz = (int(v) for v in n)
    present = np.zeros((3 * nz + 1, 3 * ny + 1, 3 * nx + 1), dtype=np.uint8)
    for ox, oy, oz in NODE_OFFSETS:
        sl = present[oz:oz + 3 * nz:3, oy:oy + 3 * ny:3, ox:ox + 3 * nx:3]
        sl |= keep.astype(np.uint8)
    return present


def lattice_grid(aabb_min, aabb_max, pad, dx):
    """x_min, x_max and cell counts of a lattice over the padded box."""
    x_min = np.asarray(aabb_min, dtype=np.float64) - pad
    x_max = np.asarray(aabb_max, dtype=np.float64) + pad
    n = np.array([int(round((x_max[d] - x_min[d]) / dx)) for d in range(3)], dtype=np.int64)
    return x_min, x_max, n


def _lattice(aabb_min, aabb_max, pad, dx, sdf, prune_lo, prune_hi, value_fn):
    x_min, x_max, n = lattice_grid(aabb_min, aabb_max, pad, dx)
    pts = _node_positions(x_min, n, dx)
    vals = value_fn(pts)
    cells = _cell_node_values(vals, n)
    keep = np.ones(cells.shape[1:], dtype=bool)
    if prune_lo is not None:
        keep &= ~np.all(cells < prune_lo, axis=0)
    if prune_hi is not None:
        keep &= ~np.all(cells > prune_hi, axis=0)
    present = _present_from_kept(keep, n)
    vals = np.where(present.astype(bool), vals, 0.0)
    return Lattice(x_min, x_max, float(dx), n, np.ascontiguousarray(vals, dtype=np.float64),
                   np.ascontiguousarray(present))


def kernel_w(q):
    """Cubic B-spline shape w(q) (``sph/kernel/cubic_spline.rs:12-13``)."""
    return np.maximum(1.0 - q, 0.0) ** 3 - 4.0 * np.maximum(0.5 - q, 0.0) ** 3


def _gamma(d, h):
    """Cubic extension of the SDF (``volume_map_boundary.rs:1716-1726``)."""
    g = kernel_w(np.abs(d) / h) / kernel_w(0.0)
    return np.where(d <= 0.0, 1.0, np.where(d >= h, 0.0, g))


_GL4_X, _GL4_W = np.polynomial.legendre.leggauss(4)


def halfspace_volume(sd, h):
    """Integral of gamma(sd(x)) over the kernel ball when sd varies linearly
    across it: V(sd) = int_{-h}^{h} gamma(sd + s) * pi (h^2 - s^2) ds.

    The integrand is a piecewise polynomial of degree <= 5 with breaks where
    sd + s in {0, h/2, h}; 4-point Gauss-Legendre per piece is exact."""
    sd = np.asarray(sd, dtype=np.float64)
    flat = sd.ravel()
    total = np.zeros_like(flat)
    brk = np.stack([np.full_like(flat, -h), -flat, 0.5 * h - flat, h - flat, np.full_like(flat, h)], -1)
    brk = np.clip(brk, -h, h)
    brk.sort(axis=-1)
    for k in range(4):
        a, b = brk[:, k], brk[:, k + 1]
        half = 0.5 * (b - a)
        mid = 0.5 * (a + b)
        for xg, wg in zip(_GL4_X, _GL4_W):
            s = mid + half * xg
            total += wg * half * _gamma(flat + s, h) * math.pi * (h * h - s * s)
    ball = 4.0 / 3.0 * math.pi * h ** 3
    out = np.where(flat >= 2.0 * h, 0.0, np.where(flat <= -h, ball, total))
    return np.clip(out, 0.0, ball).reshape(sd.shape)


def volume_map_fields(sd_fn, aabb_min, aabb_max, h, dx_particle):
    """SDF + volume-map lattices on the reference's grid rule."""
    cell = 4.0 * dx_particle                        # volume_map_boundary.rs:1120
    pad_sd = 3.1 * h                                 # :1562
    sdf = _lattice(aabb_min, aabb_max, pad_sd, cell, sd_fn, -pad_sd, pad_sd, sd_fn)
    pad_vm = 2.0 * h + cell / 6.0                    # :1595
    vm = _lattice(aabb_min, aabb_max, pad_vm, cell, sd_fn, 0.0, None,
                  lambda pts: halfspace_volume(sd_fn(pts), h))
    return sdf, vm


def sdf_and_volume_map_grid(sd_fn, aabb_min, aabb_max, h, dx_particle):
    """The SDF lattice (analytic node values) and the volume-map grid
    (x_min, x_max, n, cell size) of one boundary on the reference's rule."""
    cell = 4.0 * dx_particle
    pad_sd = 3.1 * h
    sdf = _lattice(aabb_min, aabb_max, pad_sd, cell, sd_fn, -pad_sd, pad_sd, sd_fn)
    x_min, x_max, n = lattice_grid(aabb_min, aabb_max, 2.0 * h + cell / 6.0, cell)
    return sdf, x_min, x_max, n, cell


def volume_map_quadrature(sd_fn, aabb_min, aabb_max, h, dx_particle, builder, degree=30):
    """SDF + volume-map lattices on the reference's grid rule, the volume map
    integrated from the SDF lattice with the degree-`degree` Gauss-Legendre
    ball rule (``volume_map_boundary.rs:1593-1624``).

    ``builder(sdf, h, x_min, n, dx, gl_x, gl_w) -> (values, present)`` runs the
    integration (``sph_amd.system.build_volume_map`` on the device). The
    reference takes its nodes from the gauss-quad crate; here they are numpy's
    ``leggauss`` nodes in ascending order."""
    sdf, x_min, x_max, n, cell = sdf_and_volume_map_grid(sd_fn, aabb_min, aabb_max, h,
                                                         dx_particle)
    gl_x, gl_w = np.polynomial.legendre.leggauss(degree)
    vals, pres = builder(sdf, h, x_min, n, cell, gl_x, gl_w)
    vm = Lattice(x_min, x_max, float(cell), n, np.ascontiguousarray(vals, dtype=np.float64),
                 np.ascontiguousarray(pres, dtype=np.uint8))
    return sdf, vm


# --------------------------------------------------------------------------
# triangle meshes (the reference builds its boundary fields from meshes)
# --------------------------------------------------------------------------
def box_mesh(lo, hi, inward=False):
    """The 8 vertices and 12 triangles of the box [lo, hi], wound counter-
    clockwise seen from outside (normals outward), or inward for a container
    whose fluid side is the box's interior."""
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    v = np.array([[hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]]
                  for i in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = []
    for a, b, c, d in quads:
        t += [(a, b, c), (a, c, d)]
    t = np.array(t, dtype=np.uint32)
    return v, (t[:, ::-1].copy() if inward else t)


def icosphere(radius, subdivisions=2, center=(0.0, 0.0, 0.0), inward=False):
    """A geodesic sphere: the icosahedron with each triangle split in four
    `subdivisions` times, new vertices pushed onto the sphere; outward winding
    unless `inward`."""
    g = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g),
         (0, 1, -g), (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    t = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[key] = len(v) - 1
            return mid[key]
        t = [tri for a, b, c in t
             for tri in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                         (m(a, b), m(b, c), m(c, a)))]
    verts = np.asarray(center, dtype=np.float64) + radius * np.array(v)
    tris = np.array(t, dtype=np.uint32)
    return verts, (tris[:, ::-1].copy() if inward else tris)


def mesh_fields(vertices, triangles, h, dx_particle, sdf_builder, vm_builder, degree=30):
    """DiscretizedBoundaryFields::new (volume_map_boundary.rs:1547-1636) from a
    triangle mesh: the signed-distance lattice over the mesh AABB padded by
    3.1 h, pruned outside +-3.1 h, then the volume map integrated from it with
    the degree-`degree` ball rule.

    ``sdf_builder(vertices, triangles, x_min, n, dx, prune) -> (values,
    present)`` evaluates the mesh distance (``sph_amd.system.build_mesh_sdf``
    on the device); ``vm_builder`` is as for volume_map_quadrature."""
    vertices = np.asarray(vertices, dtype=np.float64)
    cell = 4.0 * dx_particle                        # :1120
    lo, hi = vertices.min(axis=0), vertices.max(axis=0)
    pad_sd = 3.1 * h
    x_min, x_max, n = lattice_grid(lo, hi, pad_sd, cell)
    vals, pres = sdf_builder(vertices, triangles, x_min, n, cell, pad_sd)
    sdf = Lattice(x_min, x_max, float(cell), n, np.ascontiguousarray(vals, dtype=np.float64),
                  np.ascontiguousarray(pres, dtype=np.uint8))
    vx_min, vx_max, vn = lattice_grid(lo, hi, 2.0 * h + cell / 6.0, cell)
    gl_x, gl_w = np.polynomial.legendre.leggauss(degree)
    vvals, vpres = vm_builder(sdf, h, vx_min, vn, cell, gl_x, gl_w)
    vm = Lattice(vx_min, vx_max, float(cell), vn, np.ascontiguousarray(vvals, dtype=np.float64),
                 np.ascontiguousarray(vpres, dtype=np.uint8))
    return sdf, vm


# --------------------------------------------------------------------------
# fluid blocks
# --------------------------------------------------------------------------
def lattice_block(lo, counts, dx, rng, jitter):
    counts = [int(c) for c in counts]
    ix, iy, iz = np.meshgrid(*(np.arange(c) for c in counts), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=-1).astype(np.float64)
    pos = np.asarray(lo, dtype=np.float64) + (idx + 0.5) * dx
    if jitter > 0:
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    return pos


def _fluid(blocks, dx):
    xs, ms, fids = [], [], []
    for pos, rho0, fid in blocks:
        xs.append(pos)
        ms.append(np.full(len(pos), rho0 * dx ** 3))
        fids.append(np.full(len(pos), fid, dtype=np.uint32))
    x = np.ascontiguousarray(np.concatenate(xs), dtype=np.float64)
    return (x, np.zeros_like(x), np.concatenate(ms).astype(np.float64),
            np.concatenate(fids).astype(np.uint32))


def box_shell_samples(lo, hi, dx, layers):
    """Lattice samples of ``layers`` layers just outside the box [lo, hi]."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    n = np.round((hi - lo) / dx).astype(int)
    rngs = [np.arange(-layers, n[d] + layers) for d in range(3)]
    ix, iy, iz = np.meshgrid(*rngs, indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=-1)
    inside = np.all((idx >= 0) & (idx < n), axis=-1)
    idx = idx[~inside]
    return np.ascontiguousarray(lo + (idx + 0.5) * dx, dtype=np.float64)


# --------------------------------------------------------------------------
# the five workloads
# --------------------------------------------------------------------------
def dam_break_sesph(side=20, seed=0, f32_state=True) -> Scene:
    """SESPH dam break: side^3 particles at 0.02 m in a 1 x 0.6 x 0.6 m box of
    two layers of static boundary samples; h = 2 dx; rho0 1000, k 1000,
    nu 0.01; dt 5e-4; Euler-Cromer."""
    dx = 0.02
    rng = np.random.default_rng(seed)
    pos = lattice_block([0.0, 0.0, 0.0], [side] * 3, dx, rng, 1e-4)
    if f32_state:
        pos = pos.astype(np.float32).astype(np.float64)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    p = Params(kernel_support_radius=2 * dx, rest_density_grid_spacing=dx, time_increment=5e-4,
               fluid_viscosity=0.01, boundary_viscosity=0.01, solver="SESPH",
               integrator="EulerCromer", stiffness=1000.0)
    samples = box_shell_samples([0, 0, 0], [1.0, 0.6, 0.6], dx, 2)
    if f32_state:
        samples = samples.astype(np.float32).astype(np.float64)
    return Scene("sesph_dam_break", p, x, v, m, fid, [StaticSamples(samples)])


# Clearance (in particle spacings) between the fluid and a volume-map wall at
# which the wall's volume-map contribution leaves a lattice particle at about
# the rest volume: the cubic extension of the signed distance counts the band
# within h of the wall as boundary (volume_map_boundary.rs:1716-1726), so
# particles closer than ~1.2 dx start compressed (V0 / V = 2.1 at dx / 2).
WALL_CLEARANCE = 1.2


def block_drop_iisph(side=50, seed=0, dx=0.01, box=1.0, lift=0.0) -> Scene:
    """IISPH block drop: side^3 particles at dx in a box-shaped VolumeMapBoundary,
    the block centred in x and y and standing on the floor (``lift`` = 0: its
    lowest layer WALL_CLEARANCE spacings above it), so that the pressure solve
    carries the block's weight from the first steps while the block collapses
    sideways; relaxed Jacobi omega 0.5, target 0.1 %; CFL dt (0.4, max 1e-3)."""
    rng = np.random.default_rng(seed)
    edge = side * dx
    lo = [0.5 * (box - edge), 0.5 * (box - edge), lift + (WALL_CLEARANCE - 0.5) * dx]
    pos = lattice_block(lo, [side] * 3, dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver="IISPH", integrator="EulerCromer",
               target_density_error=0.1, relaxation_factor=0.5)
    lo_b, hi_b = np.zeros(3), np.full(3, box)
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    return Scene("iisph_block_drop", p, x, v, m, fid,
                 [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b))])


def lower_ball_fill(n, dx, radius, rng, jitter, clearance=WALL_CLEARANCE):
    """The n lowest points of the lattice ((i + 1/2) dx per axis, centred on
    the origin) lying at least clearance * dx inside the sphere of ``radius``:
    the sphere filled from the bottom, its top layer a disc around the
    vertical axis. Seeded uniform jitter of +-jitter per coordinate."""
    k = np.arange(-int(radius / dx) - 1, int(radius / dx) + 1, dtype=np.float64)
    c = (k + 0.5) * dx
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=-1)
    pos = pos[np.linalg.norm(pos, axis=-1) <= radius - clearance * dx]
    order = np.lexsort((np.hypot(pos[:, 0], pos[:, 1]), pos[:, 2]))
    if n > len(order):
        raise ValueError(f"{n} particles do not fit in the sphere at dx = {dx}")
    pos = pos[np.sort(order[:n])]
    if jitter > 0:
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    return pos


def two_fluid_ball(n=1_000_000, seed=0, dx=0.0127, radius=1.0, solver="IISPHwOST") -> Scene:
    """Two-fluid large ball: n particles (1M by default) filling a radius-1 m
    spherical VolumeMapBoundary from the bottom (dx = 0.0127 m puts 1M
    particles just below the equator), split at x = 0 into rest densities 1000
    and 1500; at rest on the container wall from the first step, so every
    pressure solve carries the fluid's weight while the heavier half slides
    under the lighter one. IISPHwOST + TakePredicted (or IISPH + EulerCromer);
    CFL dt (0.4, max 1e-3); uniform gravity; seeded +-1e-4 m jitter."""
    rng = np.random.default_rng(seed)
    pos = lower_ball_fill(n, dx, radius, rng, 1e-4)
    left = pos[:, 0] < 0.0
    x, v, m, fid = _fluid([(pos[left], 1000.0, 0), (pos[~left], 1500.0, 1)], dx)
    h = 2 * dx
    integ = "TakePredicted" if solver == "IISPHwOST" else "EulerCromer"
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver=solver, integrator=integ,
               target_density_error=0.1, relaxation_factor=0.5)
    c = np.zeros(3)
    sdf, vm = volume_map_fields(lambda q: sd_sphere_container(q, c, radius), c - radius,
                                c + radius, h, dx)
    return Scene(f"two_fluid_ball_{solver}", p, x, v, m, fid,
                 [VolumeMap(sdf, vm, None, "sphere", aabb_of(c - radius, c + radius))])


def rigid_coupling(counts=(158, 78, 40), seed=0, dx=0.0125) -> Scene:
    """Rigid-body coupling: a fluid layer in a 2 x 1 x 1 m box VolumeMapBoundary
    plus three two-way coupled dynamic bodies (cubes of edge 0.2 m with
    densities 500 and 800, a sphere of radius 0.15 m with density 1200) placed
    above the surface at seeded positions; IISPH; CFL dt."""
    rng = np.random.default_rng(seed)
    pos = lattice_block([dx, dx, dx], counts, dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver="IISPH", integrator="EulerCromer")
    lo_b, hi_b = np.zeros(3), np.array([2.0, 1.0, 1.0])
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    bounds = [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b))]
    surface = (counts[2] + 1) * dx
    bodies = [("cube", 0.2, 500.0), ("cube", 0.2, 800.0), ("sphere", 0.15, 1200.0)]
    for k, (kind, size, rho) in enumerate(bodies):
        cx = 0.4 + 0.6 * k + rng.uniform(-0.05, 0.05)
        cy = 0.5 + rng.uniform(-0.1, 0.1)
        cz = surface + 0.2 + rng.uniform(0.0, 0.1)
        if kind == "cube":
            half = 0.5 * size
            lo_l, hi_l = np.full(3, -half), np.full(3, half)
            mass = rho * size ** 3
            inertia = np.eye(3) * (mass * size * size / 6.0)
            fn = (lambda lo_l, hi_l: (lambda q: sd_box_solid(q, lo_l, hi_l)))(lo_l, hi_l)
        else:
            lo_l, hi_l = np.full(3, -size), np.full(3, size)
            mass = rho * 4.0 / 3.0 * math.pi * size ** 3
            inertia = np.eye(3) * (0.4 * mass * size * size)
            fn = (lambda r: (lambda q: sd_sphere_solid(q, np.zeros(3), r)))(size)
        s_l, v_l = volume_map_fields(fn, lo_l, hi_l, h, dx)
        rb = Rigid(mass=mass, inertia_body=inertia, com=np.array([cx, cy, cz]))
        bounds.append(VolumeMap(s_l, v_l, rb, kind, aabb_of(lo_l, hi_l)))
    return Scene("rigid_coupling", p, x, v, m, fid, bounds)


def small_iisph_box(side=12, seed=1, dx=0.02, box=0.6, solver="IISPH", jitter=1e-4,
                    height=None) -> Scene:
    """A small VolumeMap box scene for parity tests (CPU-oracle sized): a
    side x side x height (default side) block of particles."""
    rng = np.random.default_rng(seed)
    pos = lattice_block([0.5 * (box - side * dx), 0.5 * (box - side * dx), 0.02],
                        [side, side, height or side], dx, rng, jitter)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    integ = "TakePredicted" if solver == "IISPHwOST" else "EulerCromer"
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver=solver, integrator=integ)
    lo_b, hi_b = np.zeros(3), np.full(3, box)
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    return Scene(f"small_box_{solver}", p, x, v, m, fid,
                 [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b))])


def pool(width=0.24, layers=40, dx=0.01, seed=4, solver="IISPH", jitter=1e-4) -> Scene:
    """A pool at rest: a box container of the given width filled wall to wall
    (WALL_CLEARANCE spacings from every wall) with ``layers`` lattice layers,
    so that every pressure solve carries the column's weight from the first
    steps and nothing collapses sideways (CPU-oracle sized for long runs)."""
    rng = np.random.default_rng(seed)
    c = WALL_CLEARANCE * dx
    nxy = int(math.floor((width - 2 * c) / dx + 1e-9))
    off = 0.5 * (width - nxy * dx)
    pos = lattice_block([off, off, c - 0.5 * dx], [nxy, nxy, layers], dx, rng, jitter)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    integ = "TakePredicted" if solver == "IISPHwOST" else "EulerCromer"
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver=solver, integrator=integ)
    height = max(width, 2.0 * layers * dx)
    lo_b, hi_b = np.zeros(3), np.array([width, width, height])
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    return Scene(f"pool_{solver}", p, x, v, m, fid,
                 [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b))])


def small_rigid_scene(side=(14, 14, 8), seed=2, dx=0.02, solver="IISPH") -> Scene:
    """Small two-way coupling case: fluid layer + one dynamic cube."""
    rng = np.random.default_rng(seed)
    pos = lattice_block([dx, dx, dx], side, dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver=solver, integrator="EulerCromer")
    lo_b, hi_b = np.zeros(3), np.array([(side[0] + 2) * dx, (side[1] + 2) * dx, 0.5])
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    size = 0.1
    lo_l, hi_l = np.full(3, -0.5 * size), np.full(3, 0.5 * size)
    mass = 600.0 * size ** 3
    rb = Rigid(mass=mass, inertia_body=np.eye(3) * (mass * size * size / 6.0),
               com=np.array([0.5 * side[0] * dx + 0.013, 0.5 * side[1] * dx - 0.007,
                             (side[2] + 1) * dx + 0.5 * size + 2 * dx]),
               q=np.array([math.cos(0.1), math.sin(0.1) * 0.6, 0.0, math.sin(0.1) * 0.8]),
               ang_vel=np.array([0.3, -0.2, 0.5]))
    s_l, v_l = volume_map_fields(lambda q: sd_box_solid(q, lo_l, hi_l), lo_l, hi_l, h, dx)
    return Scene("small_rigid", p, x, v, m, fid,
                 [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b)),
                  VolumeMap(s_l, v_l, rb, "cube", aabb_of(lo_l, hi_l))])


def box_surface_samples(half, spacing):
    """Samples on the surface of the cube [-half, half]^3 (body frame): on each
    face a square grid of cell centers no wider than ``spacing`` apart."""
    n = max(1, int(math.ceil(2.0 * half / spacing - 1e-9)))
    t = -half + (np.arange(n) + 0.5) * (2.0 * half / n)
    u, w = np.meshgrid(t, t, indexing="ij")
    u, w = u.ravel(), w.ravel()
    faces = []
    for d in range(3):
        a, b = [k for k in range(3) if k != d]
        for sgn in (-1.0, 1.0):
            f = np.zeros((u.size, 3))
            f[:, d] = sgn * half
            f[:, a] = u
            f[:, b] = w
            faces.append(f)
    return np.ascontiguousarray(np.concatenate(faces), dtype=np.float64)


def sample_body_scene(side=(12, 12, 6), seed=6, dx=0.02, solver="IISPH") -> Scene:
    """Two-way coupling with a StaticSampleBoundary body
    (static_sample_boundary.rs:79-130): a fluid layer in a box of two layers
    of static samples, and a cube (edge 0.1 m, density 600) represented by
    samples on its surface at 0.7 dx (the reference's sampling spacing),
    tilted and spinning, released just above the layer."""
    rng = np.random.default_rng(seed)
    pos = lattice_block([0.0, 0.0, 0.0], side, dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    if solver in ("SESPH", "SESPHwSplitting"):
        p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, time_increment=2.5e-4,
                   solver=solver, integrator="EulerCromer", stiffness=1000.0)
    else:
        p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
                   max_time_increment=1e-3, cfl_number=0.4, solver=solver,
                   integrator="EulerCromer")
    hi = np.array([side[0] * dx, side[1] * dx, 0.4])
    walls = box_shell_samples([0.0, 0.0, 0.0], hi, dx, 2)
    size = 0.1
    mass = 600.0 * size ** 3
    rb = Rigid(mass=mass, inertia_body=np.eye(3) * (mass * size * size / 6.0),
               com=np.array([0.5 * side[0] * dx + 0.011, 0.5 * side[1] * dx - 0.006,
                             side[2] * dx + 0.5 * size + dx]),
               q=np.array([math.cos(0.1), math.sin(0.1) * 0.6, 0.0, math.sin(0.1) * 0.8]),
               lin_vel=np.array([0.0, 0.0, -0.3]), ang_vel=np.array([0.3, -0.2, 0.5]))
    body = box_surface_samples(0.5 * size, 0.7 * dx)
    return Scene(f"sample_body_{solver}", p, x, v, m, fid,
                 [StaticSamples(walls), StaticSamples(body, rb)])


def _rot_x(angle: float) -> np.ndarray:
    return np.array([math.cos(0.5 * angle), math.sin(0.5 * angle), 0.0, 0.0])


def contact_stack(seed=3, dx=0.02, depth=0.002, tilt=0.05, sphere=True,
                  solver="IISPH") -> Scene:
    """Rigid-rigid contact case: inside a 1 x 1 x 0.6 m box container, a static
    solid slab [0.2, 0.8]^2 x [0, 0.1], a dynamic cube (edge 0.2 m, density
    600, tilted by ``tilt`` about x) whose lowest edge sinks about ``depth``
    below the slab's top face and, with ``sphere``, a dynamic sphere (radius
    0.08 m, density 900, sliding at 0.3 m/s along x) that sinks ``depth`` into
    the cube's top face; a 4 x 4 x 3 fluid block sits in a far corner. IISPH;
    CFL dt."""
    rng = np.random.default_rng(seed)
    pos = lattice_block([0.86, 0.86, dx], [4, 4, 3], dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    integ = "TakePredicted" if solver == "IISPHwOST" else "EulerCromer"
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver=solver, integrator=integ)
    lo_b, hi_b = np.zeros(3), np.array([1.0, 1.0, 0.6])
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    bounds = [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b))]
    lo_s, hi_s = np.array([0.2, 0.2, 0.0]), np.array([0.8, 0.8, 0.1])
    s_s, v_s = volume_map_fields(lambda q: sd_box_solid(q, lo_s, hi_s), lo_s, hi_s, h, dx)
    bounds.append(VolumeMap(s_s, v_s, None, "slab", aabb_of(lo_s, hi_s)))
    half = 0.1
    lo_c, hi_c = np.full(3, -half), np.full(3, half)
    drop = half * (math.cos(tilt) + abs(math.sin(tilt)))  # com to lowest edge
    com_c = np.array([0.5 + rng.uniform(-0.01, 0.01), 0.5 + rng.uniform(-0.01, 0.01),
                      0.1 + drop - depth])
    mass_c = 600.0 * (2 * half) ** 3
    s_c, v_c = volume_map_fields(lambda q: sd_box_solid(q, lo_c, hi_c), lo_c, hi_c, h, dx)
    bounds.append(VolumeMap(s_c, v_c, Rigid(mass=mass_c, inertia_body=np.eye(3) * (
        mass_c * (2 * half) ** 2 / 6.0), com=com_c, q=_rot_x(tilt)), "cube",
        aabb_of(lo_c, hi_c)))
    if sphere:
        r = 0.08
        top = com_c[2] + half / math.cos(tilt)  # the top face above the com
        com_s = np.array([com_c[0] + 0.02, com_c[1] - 0.01, top + r - depth])
        mass_s = 900.0 * 4.0 / 3.0 * math.pi * r ** 3
        s_r, v_r = volume_map_fields(lambda q: sd_sphere_solid(q, np.zeros(3), r),
                                     np.full(3, -r), np.full(3, r), h, dx)
        bounds.append(VolumeMap(s_r, v_r, Rigid(mass=mass_s, inertia_body=np.eye(3) * (
            0.4 * mass_s * r * r), com=com_s, lin_vel=np.array([0.3, 0.0, 0.0])), "sphere",
            aabb_of(np.full(3, -r), np.full(3, r))))
    return Scene(f"contact_stack_{solver}", p, x, v, m, fid, bounds)


def contact_pile(cubes=(2, 3), dx=0.01, depth=0.002, seed=5, solver="IISPH") -> Scene:
    """Rigid-rigid contact: cubes[0] x cubes[1] dynamic cubes (edge 0.1 m,
    density 600, seeded tilts up to 0.05 rad) on a 0.15 m grid, each sinking
    about ``depth`` into a static slab that covers the floor of a box
    container (height 0.5 m); a 6 x 6 x 4 fluid block falls from a corner
    above the cubes. Two static boundaries and six cubes fill the eight
    boundary slots of a system. IISPH; CFL dt."""
    rng = np.random.default_rng(seed)
    sx, sy = cubes[0] * 0.15 + 0.05, cubes[1] * 0.15 + 0.05
    top = 0.05
    pos = lattice_block([0.01, 0.01, 0.3], [6, 6, 4], dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, 1000.0, 0)], dx)
    h = 2 * dx
    integ = "TakePredicted" if solver == "IISPHwOST" else "EulerCromer"
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, solver=solver, integrator=integ)
    lo_b, hi_b = np.zeros(3), np.array([sx, sy, 0.5])
    sdf, vm = volume_map_fields(lambda q: sd_box_container(q, lo_b, hi_b), lo_b, hi_b, h, dx)
    bounds = [VolumeMap(sdf, vm, None, "box", aabb_of(lo_b, hi_b))]
    lo_s, hi_s = np.zeros(3), np.array([sx, sy, top])
    s_s, v_s = volume_map_fields(lambda q: sd_box_solid(q, lo_s, hi_s), lo_s, hi_s, h, dx)
    bounds.append(VolumeMap(s_s, v_s, None, "slab", aabb_of(lo_s, hi_s)))
    half = 0.05
    lo_c, hi_c = np.full(3, -half), np.full(3, half)
    s_c, v_c = volume_map_fields(lambda q: sd_box_solid(q, lo_c, hi_c), lo_c, hi_c, h, dx)
    mass = 600.0 * (2 * half) ** 3
    for i in range(cubes[0]):
        for j in range(cubes[1]):
            tilt = rng.uniform(-0.05, 0.05)
            drop = half * (math.cos(tilt) + abs(math.sin(tilt)))
            com = np.array([0.1 + 0.15 * i, 0.1 + 0.15 * j, top + drop - depth])
            rb = Rigid(mass=mass, inertia_body=np.eye(3) * (mass * (2 * half) ** 2 / 6.0),
                       com=com, q=_rot_x(tilt))
            bounds.append(VolumeMap(s_c, v_c, rb, "cube", aabb_of(lo_c, hi_c)))
    return Scene(f"contact_pile_{cubes[0] * cubes[1]}", p, x, v, m, fid, bounds)


def halfspace_contact(depth, dx=0.02, half=0.1, offset=(0.013, -0.007), dt=1e-3,
                      floor_half=0.3) -> Scene:
    """Planar rigid-rigid contact: a static floor (solid below z = 0.5, AABB
    [0.5 -+ floor_half]^2 x [0.4, 0.5]) and a dynamic block whose signed
    distance is that of its bottom face alone (solid above it; AABB a cube of
    half edge ``half``), the face ``depth`` below the floor's top. Both signed
    distances are linear, so the lattices reproduce them exactly. Four fluid
    particles sit far outside both lattices; fixed dt."""
    h = 2 * dx
    zt = 0.5
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=False,
               time_increment=dt, solver="IISPH", integrator="EulerCromer")
    x = np.array([[3.0 + 0.02 * i, 3.0, 3.0] for i in range(4)])
    v = np.zeros_like(x)
    m = np.full(len(x), 1000.0 * dx ** 3)
    fid = np.zeros(len(x), dtype=np.uint32)
    lo_f = np.array([0.5 - floor_half, 0.5 - floor_half, zt - 0.1])
    hi_f = np.array([0.5 + floor_half, 0.5 + floor_half, zt])
    s_f, v_f = volume_map_fields(lambda q: q[..., 2] - zt, lo_f, hi_f, h, dx)
    lo_b, hi_b = np.full(3, -half), np.full(3, half)
    s_b, v_b = volume_map_fields(lambda q: -half - q[..., 2], lo_b, hi_b, h, dx)
    mass = 500.0 * (2 * half) ** 3
    com = np.array([0.5 + offset[0], 0.5 + offset[1], zt - depth + half])
    rb = Rigid(mass=mass, inertia_body=np.eye(3) * (mass * (2 * half) ** 2 / 6.0), com=com)
    return Scene("halfspace_contact", p, x, v, m, fid,
                 [VolumeMap(s_f, v_f, None, "floor", aabb_of(lo_f, hi_f)),
                  VolumeMap(s_b, v_b, rb, "block", aabb_of(lo_b, hi_b))])


# --------------------------------------------------------------------------
# serving stream scenes
# --------------------------------------------------------------------------
_CONTAINER_CACHE = {}


def container_fields(shape: str, size: float, h: float, dx: float):
    """Volume-map fields of a box ([0, size]^3) or sphere (radius ``size``
    centred at the origin) container, cached per (shape, size, h, dx): the
    field construction is per-scene setup, and requests share containers."""
    key = (shape, round(size, 9), h, dx)
    if key not in _CONTAINER_CACHE:
        if shape == "box":
            lo, hi = np.zeros(3), np.full(3, size)
            f = volume_map_fields(lambda q: sd_box_container(q, lo, hi), lo, hi, h, dx)
        else:
            c = np.zeros(3)
            f = volume_map_fields(lambda q: sd_sphere_container(q, c, size), c - size,
                                  c + size, h, dx)
        _CONTAINER_CACHE[key] = f
    return _CONTAINER_CACHE[key]


def serving_scene(n: int, rho0: float, nu: float, shape: str, seed: int, dx: float = 0.02,
                  steps_hint: int = 200) -> Scene:
    """One request of the serving stream: about ``n`` IISPH particles of rest
    density ``rho0`` and viscosity ``nu`` as a block in a box or sphere
    volume-map container (sizes quantized so that containers are shared);
    CFL dt (0.4, max 1e-3), target 0.1 %."""
    rng = np.random.default_rng(seed)
    side = max(2, int(round(n ** (1.0 / 3.0))))
    nz = max(1, int(round(n / (side * side))))
    counts = [side, side, nz]
    ext = np.array(counts, dtype=np.float64) * dx
    h = 2 * dx
    if shape == "box":
        step = 8 * dx
        size = math.ceil((max(ext) + 4 * dx) / step) * step
        lo = [0.5 * (size - ext[0]), 0.5 * (size - ext[1]), 2 * dx]
    else:
        step = 4 * dx
        size = math.ceil((0.5 * float(np.linalg.norm(ext)) + 3 * dx) / step) * step
        lo = list(-0.5 * ext)
    pos = lattice_block(lo, counts, dx, rng, 1e-4)
    x, v, m, fid = _fluid([(pos, rho0, 0)], dx)
    p = Params(kernel_support_radius=h, rest_density_grid_spacing=dx, cfl=True,
               max_time_increment=1e-3, cfl_number=0.4, fluid_viscosity=nu,
               boundary_viscosity=nu, solver="IISPH", integrator="EulerCromer",
               target_density_error=0.1, relaxation_factor=0.5)
    sdf, vm = container_fields(shape, size, h, dx)
    box = aabb_of(np.zeros(3), np.full(3, size)) if shape == "box" else \
        aabb_of(np.full(3, -size), np.full(3, size))
    return Scene(f"serve_{shape}_{len(x)}", p, x, v, m, fid,
                 [VolumeMap(sdf, vm, None, shape, box)])

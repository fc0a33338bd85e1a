"""System state and recordings in the reference's on-disk formats.

Two formats, both written and read here without the Rust crates:

* **Checkpoints.** ``SerSystemCheckpoint`` (sph/mod.rs:884-915) with
  ``SerFluidCheckpoint`` (fluid.rs:399-448), ``SerBoundaryCheckpoint``
  (boundary_handling/mod.rs:147-174) and ``SerRigidBodyMotionCheckpoint``
  (rigid_body_motion.rs:247-289). The reference saves them as RON text and
  resumes from that text (``--state``, setup/mod.rs:145); the struct also
  derives bincode ``Encode``/``Decode``, so both encodings are provided.
* **Recordings.** The recorder's file (sci-phi-backend/src/recording.rs:
  146-211): a sequence of records, each a u64 little-endian byte length
  followed by a bincode-encoded struct: first one ``SimulationParameters``
  header (render_info.rs:15-67), then one ``TimeStepInfo`` per recorded step
  (render_info.rs:69-107). Re-recording a step number truncates the file back
  to that step's record.

bincode is version 2.0.1 with ``config::standard()`` (render_info.rs:57):
little-endian, variable-length integers (a value below 251 is one byte; 251,
252 and 253 prefix a u16, u32 and u64), zigzag for signed integers, floats as
their raw little-endian bytes, ``bool`` and ``Option`` tags as one byte, enum
variants as a varint u32 index, sequences as a varint length followed by the
elements, fixed-size arrays and tuples as their elements with no length.

RON is the text serde produces through ron 0.12: structs as
``(field:value,...)``, fixed-size arrays as tuples ``(a,b,c)``, sequences as
``[...]``, ``Option`` as ``Some(x)``/``None``. The reader also accepts the
pretty-printed form (whitespace, trailing commas, comments) and the
``Name(...)`` struct prefix.

The fluid checkpoint holds the particle arrays in the reference's particle
order: the library keeps its particles sorted by cell on the device and
scatters them back to the reference's index order when it downloads them, so
a checkpoint written here lists the particles as the reference's would.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# ----------------------------------------------------------------------------
# bincode 2 (config::standard) primitives
# ----------------------------------------------------------------------------


class DecodeError(ValueError):
    pass


class Writer:
    """bincode 2 standard-config encoder into a bytearray."""

    def __init__(self):
        self.b = bytearray()

    def varint(self, v: int):
        if v < 0:
            raise ValueError("unsigned varint of a negative value")
        if v < 251:
            self.b.append(v)
        elif v < 1 << 16:
            self.b.append(251)
            self.b += struct.pack("<H", v)
        elif v < 1 << 32:
            self.b.append(252)
            self.b += struct.pack("<I", v)
        elif v < 1 << 64:
            self.b.append(253)
            self.b += struct.pack("<Q", v)
        else:
            self.b.append(254)
            self.b += v.to_bytes(16, "little")

    def zigzag(self, v: int):
        self.varint((v << 1) if v >= 0 else ((-v) << 1) - 1)

    def u8(self, v: int):
        self.b.append(v & 0xFF)

    def boolean(self, v: bool):
        self.b.append(1 if v else 0)

    def f32(self, v: float):
        self.b += struct.pack("<f", v)

    def f64(self, v: float):
        self.b += struct.pack("<d", v)

    def f64_array(self, a):
        """Elements of a fixed-size array or a packed run of f64 (no length)."""
        self.b += np.ascontiguousarray(a, dtype="<f8").tobytes()

    def f32_array(self, a):
        self.b += np.ascontiguousarray(a, dtype="<f4").tobytes()

    def u32_seq(self, a):
        a = np.asarray(a, dtype=np.uint64).ravel()
        self.varint(len(a))
        for v in a.tolist():
            self.varint(int(v))

    def string(self, s: str):
        raw = s.encode("utf-8")
        self.varint(len(raw))
        self.b += raw


class Reader:
    """bincode 2 standard-config decoder over a bytes object."""

    def __init__(self, data: bytes, pos: int = 0):
        self.d = memoryview(data)
        self.p = pos

    def _take(self, n: int) -> memoryview:
        if self.p + n > len(self.d):
            raise DecodeError(f"unexpected end of data at byte {self.p} (+{n})")
        out = self.d[self.p:self.p + n]
        self.p += n
        return out

    def varint(self, limit_bits: int = 64) -> int:
        tag = self._take(1)[0]
        if tag < 251:
            v = tag
        elif tag == 251:
            v = struct.unpack("<H", self._take(2))[0]
        elif tag == 252:
            v = struct.unpack("<I", self._take(4))[0]
        elif tag == 253:
            v = struct.unpack("<Q", self._take(8))[0]
        elif tag == 254:
            v = int.from_bytes(self._take(16), "little")
        else:
            raise DecodeError(f"invalid varint tag {tag}")
        if v >> limit_bits:
            raise DecodeError(f"integer {v} exceeds {limit_bits} bits")
        return v

    def zigzag(self, limit_bits: int = 64) -> int:
        v = self.varint(limit_bits)
        return (v >> 1) if not v & 1 else -((v + 1) >> 1)

    def u8(self) -> int:
        return self._take(1)[0]

    def boolean(self) -> bool:
        v = self.u8()
        if v > 1:
            raise DecodeError(f"invalid bool byte {v}")
        return bool(v)

    def f32(self) -> float:
        return struct.unpack("<f", self._take(4))[0]

    def f64(self) -> float:
        return struct.unpack("<d", self._take(8))[0]

    def f64_array(self, n: int) -> np.ndarray:
        return np.frombuffer(bytes(self._take(8 * n)), dtype="<f8").astype(np.float64)

    def f32_array(self, n: int) -> np.ndarray:
        return np.frombuffer(bytes(self._take(4 * n)), dtype="<f4").astype(np.float32)

    def u32_seq(self) -> np.ndarray:
        n = self.varint()
        return np.array([self.varint(32) for _ in range(n)], dtype=np.uint32)

    def string(self) -> str:
        n = self.varint()
        return bytes(self._take(n)).decode("utf-8")

    def option_tag(self) -> bool:
        t = self.u8()
        if t > 1:
            raise DecodeError(f"invalid Option tag {t}")
        return t == 1

    def done(self) -> bool:
        return self.p == len(self.d)


def _vec3_seq_write(w: Writer, a):
    """Vec<[f64; 3]>: a length, then 3 f64 per element."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    w.varint(len(a))
    w.f64_array(a)


def _vec3_seq_read(r: Reader) -> np.ndarray:
    n = r.varint()
    return r.f64_array(3 * n).reshape(n, 3)


# ----------------------------------------------------------------------------
# checkpoint structures
# ----------------------------------------------------------------------------


@dataclass
class RigidBodyCheckpoint:
    """SerRigidBodyMotionCheckpoint (rigid_body_motion.rs:247-289); the
    orientation is the quaternion as (i, j, k, w)."""
    center_of_mass: np.ndarray
    orientation: np.ndarray
    linear_velocity: np.ndarray
    angular_momentum: np.ndarray
    force: np.ndarray
    torque: np.ndarray

    FIELDS = ("center_of_mass", "orientation", "linear_velocity", "angular_momentum",
              "force", "torque")
    SIZES = (3, 4, 3, 3, 3, 3)


@dataclass
class FluidCheckpoint:
    """SerFluidCheckpoint (fluid.rs:399-448)."""
    fluid_id: np.ndarray
    position: np.ndarray
    velocity: np.ndarray
    mass: np.ndarray
    integrator_position_slots: List[np.ndarray] = field(default_factory=list)
    integrator_velocity_slots: List[np.ndarray] = field(default_factory=list)


@dataclass
class SystemCheckpoint:
    """SerSystemCheckpoint (sph/mod.rs:884-915) with its boundary part
    (SerBoundaryCheckpoint: one entry per boundary, None for a static one)."""
    time_steps_propagated: int
    current_time: float
    fluid: FluidCheckpoint
    dynamic_states: List[Optional[RigidBodyCheckpoint]]


def encode_checkpoint(cp: SystemCheckpoint) -> bytes:
    """bincode 2 standard encoding of SerSystemCheckpoint (field order of the
    struct definitions)."""
    w = Writer()
    w.varint(int(cp.time_steps_propagated))
    w.f64(cp.current_time)
    f = cp.fluid
    w.u32_seq(f.fluid_id)
    _vec3_seq_write(w, f.position)
    _vec3_seq_write(w, f.velocity)
    m = np.asarray(f.mass, dtype=np.float64).ravel()
    w.varint(len(m))
    w.f64_array(m)
    for slots in (f.integrator_position_slots, f.integrator_velocity_slots):
        w.varint(len(slots))
        for s in slots:
            _vec3_seq_write(w, s)
    w.varint(len(cp.dynamic_states))
    for d in cp.dynamic_states:
        if d is None:
            w.u8(0)
        else:
            w.u8(1)
            for name, n in zip(RigidBodyCheckpoint.FIELDS, RigidBodyCheckpoint.SIZES):
                a = np.asarray(getattr(d, name), dtype=np.float64).ravel()
                if len(a) != n:
                    raise ValueError(f"{name} needs {n} values")
                w.f64_array(a)
    return bytes(w.b)


def decode_checkpoint(data: bytes) -> SystemCheckpoint:
    r = Reader(data)
    steps = r.varint()
    t = r.f64()
    fid = r.u32_seq()
    pos = _vec3_seq_read(r)
    vel = _vec3_seq_read(r)
    m = r.f64_array(r.varint())
    slots = []
    for _ in range(2):
        slots.append([_vec3_seq_read(r) for _ in range(r.varint())])
    states = []
    for _ in range(r.varint()):
        if r.option_tag():
            states.append(RigidBodyCheckpoint(*[r.f64_array(n) for n in RigidBodyCheckpoint.SIZES]))
        else:
            states.append(None)
    if not r.done():
        raise DecodeError(f"{len(r.d) - r.p} trailing bytes")
    return SystemCheckpoint(steps, t, FluidCheckpoint(fid, pos, vel, m, slots[0], slots[1]),
                            states)


# ----------------------------------------------------------------------------
# RON
# ----------------------------------------------------------------------------


def _ron_f64(v: float) -> str:
    """A binary64 as RON text that parses back to the same bits: the shortest
    round-trip digits, always with a '.' or an exponent so that it stays a
    float literal."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    s = repr(v)
    if "." not in s and "e" not in s and "E" not in s:
        s += ".0"
    return s


def _ron_tuple(a) -> str:
    return "(" + ",".join(_ron_f64(x) for x in np.asarray(a, dtype=np.float64).ravel()) + ")"


def _ron_vec3_seq(a) -> str:
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    return "[" + ",".join(_ron_tuple(row) for row in a) + "]"


def checkpoint_to_ron(cp: SystemCheckpoint) -> str:
    """SerSystemCheckpoint as compact RON (the shape ron::to_string writes)."""
    f = cp.fluid
    parts = [
        "fluid_id:[" + ",".join(str(int(x)) for x in np.asarray(f.fluid_id).ravel()) + "]",
        "position:" + _ron_vec3_seq(f.position),
        "velocity:" + _ron_vec3_seq(f.velocity),
        "mass:[" + ",".join(_ron_f64(x) for x in np.asarray(f.mass).ravel()) + "]",
        "integrator_position_slots:[" + ",".join(_ron_vec3_seq(s) for s in
                                                 f.integrator_position_slots) + "]",
        "integrator_velocity_slots:[" + ",".join(_ron_vec3_seq(s) for s in
                                                 f.integrator_velocity_slots) + "]",
    ]
    fluid = "(" + ",".join(parts) + ")"
    states = []
    for d in cp.dynamic_states:
        if d is None:
            states.append("None")
        else:
            body = ",".join(f"{name}:{_ron_tuple(getattr(d, name))}"
                            for name in RigidBodyCheckpoint.FIELDS)
            states.append(f"Some(({body}))")
    boundary = "(dynamic_states:[" + ",".join(states) + "])"
    return (f"(time_steps_propagated:{int(cp.time_steps_propagated)},"
            f"current_time:{_ron_f64(cp.current_time)},fluid:{fluid},boundary:{boundary})")


class _RonParser:
    """A small RON reader for the value shapes of the checkpoint: structs,
    tuples, sequences, Option, integers and floats."""

    def __init__(self, text: str):
        self.s = text
        self.p = 0

    def error(self, msg):
        line = self.s.count("\n", 0, self.p) + 1
        return DecodeError(f"RON: {msg} (line {line}, offset {self.p})")

    def ws(self):
        s, n = self.s, len(self.s)
        while self.p < n:
            c = s[self.p]
            if c in " \t\r\n":
                self.p += 1
            elif s.startswith("//", self.p):
                e = s.find("\n", self.p)
                self.p = n if e < 0 else e + 1
            elif s.startswith("/*", self.p):
                e = s.find("*/", self.p + 2)
                if e < 0:
                    raise self.error("unterminated comment")
                self.p = e + 2
            else:
                break

    def peek(self) -> str:
        self.ws()
        return self.s[self.p] if self.p < len(self.s) else ""

    def expect(self, c: str):
        if self.peek() != c:
            raise self.error(f"expected {c!r}")
        self.p += 1

    def ident(self) -> str:
        self.ws()
        st = self.p
        while self.p < len(self.s) and (self.s[self.p].isalnum() or self.s[self.p] == "_"):
            self.p += 1
        return self.s[st:self.p]

    def number(self) -> str:
        self.ws()
        st = self.p
        while self.p < len(self.s) and self.s[self.p] in "+-0123456789.eEinfNa_":
            self.p += 1
        tok = self.s[st:self.p].replace("_", "")
        if not tok:
            raise self.error("expected a number")
        return tok

    def f64(self) -> float:
        tok = self.number()
        try:
            return float(tok)
        except ValueError:
            raise self.error(f"bad float {tok!r}") from None

    def uint(self) -> int:
        tok = self.number()
        try:
            v = int(tok)
        except ValueError:
            raise self.error(f"bad integer {tok!r}") from None
        if v < 0:
            raise self.error("negative value for an unsigned field")
        return v

    def seq(self, item):
        self.expect("[")
        out = []
        while self.peek() != "]":
            out.append(item())
            if self.peek() == ",":
                self.p += 1
            elif self.peek() != "]":
                raise self.error("expected ',' or ']'")
        self.p += 1
        return out

    def tup(self, n: int) -> List[float]:
        c = self.peek()
        close = ")" if c == "(" else "]"
        if c not in "([":
            raise self.error("expected a tuple")
        self.p += 1
        out = []
        for k in range(n):
            out.append(self.f64())
            if self.peek() == ",":
                self.p += 1
            elif k != n - 1:
                raise self.error("expected ','")
        self.expect(close)
        return out

    def struct(self, fields):
        """Parses `Name?(field: value, ...)` with `fields` = {name: parser}."""
        if self.peek().isalpha():
            self.ident()  # optional struct name
        self.expect("(")
        out = {}
        while self.peek() != ")":
            name = self.ident()
            if name not in fields:
                raise self.error(f"unknown field {name!r}")
            self.expect(":")
            out[name] = fields[name]()
            if self.peek() == ",":
                self.p += 1
            elif self.peek() != ")":
                raise self.error("expected ',' or ')'")
        self.p += 1
        missing = [k for k in fields if k not in out]
        if missing:
            raise self.error(f"missing fields {missing}")
        return out

    def option(self, item):
        name = self.ident()
        if name == "None":
            return None
        if name != "Some":
            raise self.error("expected Some(...) or None")
        self.expect("(")
        v = item()
        self.expect(")")
        return v

    def end(self):
        self.ws()
        if self.p != len(self.s):
            raise self.error("trailing characters")


def checkpoint_from_ron(text: str) -> SystemCheckpoint:
    P = _RonParser(text)

    def vec3_seq():
        rows = P.seq(lambda: P.tup(3))
        return np.array(rows, dtype=np.float64).reshape(-1, 3)

    def fluid():
        d = P.struct({
            "fluid_id": lambda: P.seq(P.uint),
            "position": vec3_seq,
            "velocity": vec3_seq,
            "mass": lambda: P.seq(P.f64),
            "integrator_position_slots": lambda: P.seq(vec3_seq),
            "integrator_velocity_slots": lambda: P.seq(vec3_seq),
        })
        ids = d["fluid_id"]
        if any(v >> 32 for v in ids):
            raise P.error("fluid id exceeds u32")
        return FluidCheckpoint(np.array(ids, dtype=np.uint32), d["position"], d["velocity"],
                               np.array(d["mass"], dtype=np.float64),
                               d["integrator_position_slots"], d["integrator_velocity_slots"])

    def rigid():
        d = P.struct({name: (lambda n=n: np.array(P.tup(n), dtype=np.float64))
                      for name, n in zip(RigidBodyCheckpoint.FIELDS, RigidBodyCheckpoint.SIZES)})
        return RigidBodyCheckpoint(**d)

    def boundary():
        return P.struct({"dynamic_states": lambda: P.seq(lambda: P.option(rigid))})

    d = P.struct({"time_steps_propagated": P.uint, "current_time": P.f64, "fluid": fluid,
                  "boundary": boundary})
    P.end()
    return SystemCheckpoint(d["time_steps_propagated"], d["current_time"], d["fluid"],
                            d["boundary"]["dynamic_states"])


def save_checkpoint(cp: SystemCheckpoint, path: str, fmt: str = "ron"):
    if fmt == "ron":
        data = checkpoint_to_ron(cp).encode("utf-8")
    elif fmt == "bincode":
        data = encode_checkpoint(cp)
    else:
        raise ValueError(f"unknown checkpoint format {fmt!r}")
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(data)
    os.replace(tmp, path)


def load_checkpoint(path: str, fmt: Optional[str] = None) -> SystemCheckpoint:
    with open(path, "rb") as fh:
        data = fh.read()
    if fmt is None:
        head = data.lstrip()[:1]
        fmt = "ron" if head in (b"(", b"S", b"/") else "bincode"
    if fmt == "ron":
        return checkpoint_from_ron(data.decode("utf-8"))
    if fmt == "bincode":
        return decode_checkpoint(data)
    raise ValueError(f"unknown checkpoint format {fmt!r}")


# ----------------------------------------------------------------------------
# recordings (render_info.rs, measurement.rs, recording.rs)
# ----------------------------------------------------------------------------

SCALAR_QUANTITIES = ["Speed", "Volume", "Density", "DensityError", "Pressure", "KineticEnergy"]

MEASUREMENT_FIELDS = [  # measurement.rs:52-75, in declaration order
    ("time", "f64"), ("density", "f64"), ("density_error", "f64"), ("kinetic_energy", "f64"),
    ("stiffness", "f64"), ("fluid_viscosity", "f64"), ("boundary_viscosity", "f64"),
    ("fluid_depth", "f64"), ("rest_density_grid_spacing", "f64"),
    ("kernel_support_radius", "f64"), ("time_step_size", "f64"),
    ("target_density_error", "f64"), ("solver_iterations", "u32"),
    ("relaxation_factor", "f64"), ("time_step_wall_clock_time", "f64"),
    ("predicted_density_error", "f64"),
]


@dataclass
class SimulationParameters:
    """render_info.rs:15-29."""
    particle_diameter: float
    light_position: tuple
    buffer_length_limit: int
    is_measured: bool
    is_recorded: bool
    explicitly_sampled_boundary: bool


def encode_simulation_parameters(sp: SimulationParameters) -> bytes:
    w = Writer()
    w.f32(sp.particle_diameter)
    w.f32_array(np.asarray(sp.light_position, dtype=np.float32).ravel()[:3])
    w.varint(int(sp.buffer_length_limit))
    w.boolean(sp.is_measured)
    w.boolean(sp.is_recorded)
    w.boolean(sp.explicitly_sampled_boundary)
    return bytes(w.b)


def decode_simulation_parameters(data: bytes) -> SimulationParameters:
    r = Reader(data)
    sp = SimulationParameters(r.f32(), tuple(r.f32_array(3).tolist()), r.varint(), r.boolean(),
                              r.boolean(), r.boolean())
    if not r.done():
        raise DecodeError("trailing bytes after SimulationParameters")
    return sp


def _write_measurement(w: Writer, m: dict):
    for name, kind in MEASUREMENT_FIELDS:
        v = m.get(name, 0)
        if kind == "u32":
            w.varint(int(v))
        else:
            w.f64(float(v))


def _read_measurement(r: Reader) -> dict:
    return {name: (r.varint(32) if kind == "u32" else r.f64()) for name, kind in MEASUREMENT_FIELDS}


def _write_f32x3_seq(w: Writer, a):
    a = np.asarray(a, dtype=np.float32).reshape(-1, 3)
    w.varint(len(a))
    w.f32_array(a)


def _read_f32x3_seq(r: Reader) -> np.ndarray:
    n = r.varint()
    return r.f32_array(3 * n).reshape(n, 3)


def _write_f32_seq(w: Writer, a):
    a = np.asarray(a, dtype=np.float32).ravel()
    w.varint(len(a))
    w.f32_array(a)


def _read_f32_seq(r: Reader) -> np.ndarray:
    return r.f32_array(r.varint())


def encode_time_step_info(info: dict) -> bytes:
    """TimeStepInfo (render_info.rs:69-78) with a `Samples` or `SensorPlane`
    fluid visualization and a `Samples` boundary visualization.

    info = {"time_step_number": int, "measurement": {...},
            "fluid": {"kind": "Samples", "positions": (n,3),
                      "coloring": {"kind": "Uniform"} | {"kind": "FluidId", "id": ..., "max_id": ...}
                                  | {"kind": "QuantityGraded", "quantity": str, "data": ...}}
                   | {"kind": "SensorPlane", "quantity": str,
                      "planes": [{"positions": (k,3), "data": (k,), "rows": r, "cols": c}, ...]},
            "boundary": {"kind": "Samples", "positions": (m,3),
                         "coloring": {"kind": "Uniform"} | {"kind": "BoundaryId", "ids": ..., "max_id": ...}}}
    """
    w = Writer()
    w.varint(int(info["time_step_number"]))
    _write_measurement(w, info.get("measurement", {}))
    f = info["fluid"]
    if f["kind"] == "Samples":  # FluidVisualization variant 1
        w.varint(1)
        _write_f32x3_seq(w, f["positions"])
        c = f.get("coloring", {"kind": "Uniform"})
        if c["kind"] == "Uniform":
            w.varint(0)
        elif c["kind"] == "FluidId":
            w.varint(1)
            w.u32_seq(c["id"])
            w.varint(int(c["max_id"]))
        elif c["kind"] == "QuantityGraded":
            w.varint(2)
            w.varint(SCALAR_QUANTITIES.index(c["quantity"]))
            _write_f32_seq(w, c["data"])
        else:
            raise ValueError(f"unknown sample coloring {c['kind']!r}")
    elif f["kind"] == "SensorPlane":  # variant 2
        w.varint(2)
        w.varint(SCALAR_QUANTITIES.index(f["quantity"]))
        w.varint(len(f["planes"]))
        for pl in f["planes"]:
            _write_f32x3_seq(w, pl["positions"])
            _write_f32_seq(w, pl["data"])
            w.varint(int(pl["rows"]))
            w.varint(int(pl["cols"]))
    else:
        raise ValueError("fluid visualization must be Samples or SensorPlane (surface meshes "
                         "are not produced by this library)")
    b = info.get("boundary", {"kind": "Samples", "positions": np.zeros((0, 3))})
    if b["kind"] != "Samples":
        raise ValueError("boundary visualization must be Samples")
    w.varint(1)  # BoundaryVisualization::Samples
    _write_f32x3_seq(w, b["positions"])
    c = b.get("coloring", {"kind": "Uniform"})
    if c["kind"] == "Uniform":
        w.varint(0)
    elif c["kind"] == "BoundaryId":
        w.varint(1)
        w.u32_seq(c["ids"])
        w.varint(int(c["max_id"]))
    else:
        raise ValueError(f"unknown boundary coloring {c['kind']!r}")
    return bytes(w.b)


def decode_time_step_info(data: bytes) -> dict:
    r = Reader(data)
    out = {"time_step_number": r.varint(), "measurement": _read_measurement(r)}
    kind = r.varint(32)
    if kind == 1:
        pos = _read_f32x3_seq(r)
        ck = r.varint(32)
        if ck == 0:
            col = {"kind": "Uniform"}
        elif ck == 1:
            col = {"kind": "FluidId", "id": r.u32_seq(), "max_id": r.varint(32)}
        elif ck == 2:
            q = r.varint(32)
            if q >= len(SCALAR_QUANTITIES):
                raise DecodeError(f"invalid ScalarQuantity {q}")
            col = {"kind": "QuantityGraded", "quantity": SCALAR_QUANTITIES[q],
                   "data": _read_f32_seq(r)}
        else:
            raise DecodeError(f"invalid FluidSampleColoring {ck}")
        out["fluid"] = {"kind": "Samples", "positions": pos, "coloring": col}
    elif kind == 2:
        q = r.varint(32)
        if q >= len(SCALAR_QUANTITIES):
            raise DecodeError(f"invalid ScalarQuantity {q}")
        planes = []
        for _ in range(r.varint()):
            planes.append({"positions": _read_f32x3_seq(r), "data": _read_f32_seq(r),
                           "rows": r.varint(), "cols": r.varint()})
        out["fluid"] = {"kind": "SensorPlane", "quantity": SCALAR_QUANTITIES[q], "planes": planes}
    elif kind == 0:
        raise DecodeError("TriangleMesh fluid visualizations are not supported")
    else:
        raise DecodeError(f"invalid FluidVisualization {kind}")
    bk = r.varint(32)
    if bk != 1:
        raise DecodeError("only Samples boundary visualizations are supported")
    bpos = _read_f32x3_seq(r)
    ck = r.varint(32)
    if ck == 0:
        col = {"kind": "Uniform"}
    elif ck == 1:
        col = {"kind": "BoundaryId", "ids": r.u32_seq(), "max_id": r.varint(32)}
    else:
        raise DecodeError(f"invalid BoundarySampleColoring {ck}")
    out["boundary"] = {"kind": "Samples", "positions": bpos, "coloring": col}
    if not r.done():
        raise DecodeError("trailing bytes after TimeStepInfo")
    return out


class Recording:
    """The recorder's file (recording.rs:109-211): a SimulationParameters
    header record, then TimeStepInfo records, each prefixed with its u64 LE
    byte length. Appending a step number already written truncates the file
    back to that step's record first (a re-produced step overwrites it)."""

    def __init__(self, path: str, params: SimulationParameters):
        if os.path.exists(path):
            raise FileExistsError(path)
        parent = os.path.dirname(os.path.abspath(path))
        os.makedirs(parent, exist_ok=True)
        self.path = path
        self.offsets = {}
        with open(path, "ab") as fh:
            raw = encode_simulation_parameters(params)
            fh.write(struct.pack("<Q", len(raw)))
            fh.write(raw)

    def append_time_step(self, info: dict):
        step = int(info["time_step_number"])
        raw = encode_time_step_info(info)
        with open(self.path, "r+b") as fh:
            if step in self.offsets:
                off = self.offsets[step]
                fh.truncate(off)
                self.offsets = {k: v for k, v in self.offsets.items() if k < step}
            else:
                off = fh.seek(0, os.SEEK_END)
            fh.seek(off)
            fh.write(struct.pack("<Q", len(raw)))
            fh.write(raw)
        self.offsets[step] = off


def read_recording(path: str):
    """(SimulationParameters, [TimeStepInfo dict, ...]) of a recording file."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    records = []
    while pos < len(data):
        if pos + 8 > len(data):
            raise DecodeError("truncated length prefix")
        n = struct.unpack_from("<Q", data, pos)[0]
        pos += 8
        if pos + n > len(data):
            raise DecodeError("truncated record")
        records.append(data[pos:pos + n])
        pos += n
    if not records:
        raise DecodeError("empty recording")
    return (decode_simulation_parameters(records[0]),
            [decode_time_step_info(rec) for rec in records[1:]])

"""Checkpoint and recording formats (CPU).

The encodings are pinned by the published rules of bincode 2.0.1's standard
configuration (varint thresholds 251/252/253/254, zigzag for signed values,
raw little-endian floats, one-byte Option and bool tags, no length for
fixed-size arrays) with hand-assembled byte strings in the field order of the
reference's struct declarations, and by the round-trip behaviour the
reference's own tests require of SerBoundaryCheckpoint
(tests/sph/boundary_handling/mod.rs:160-255). The byte-exact output of the
Rust crates themselves is parity unpinned: neither can be built here.
"""
import os
import struct

import numpy as np
import pytest

from sph_amd import state_io as S


def _vec(v):
    w = S.Writer()
    w.varint(v)
    return bytes(w.b)


@pytest.mark.parametrize("value,expected", [
    (0, b"\x00"), (1, b"\x01"), (250, b"\xfa"),
    (251, b"\xfb\xfb\x00"), (0xFFFF, b"\xfb\xff\xff"),
    (0x10000, b"\xfc\x00\x00\x01\x00"), (0xFFFFFFFF, b"\xfc\xff\xff\xff\xff"),
    (1 << 32, b"\xfd\x00\x00\x00\x00\x01\x00\x00\x00"),
    ((1 << 64) - 1, b"\xfd" + b"\xff" * 8),
])
def test_varint_thresholds(value, expected):
    assert _vec(value) == expected
    assert S.Reader(expected).varint() == value


@pytest.mark.parametrize("value,zz", [(0, 0), (-1, 1), (1, 2), (-2, 3), (2, 4),
                                      (-126, 251), (125, 250)])
def test_zigzag(value, zz):
    w = S.Writer()
    w.zigzag(value)
    assert bytes(w.b) == _vec(zz)
    assert S.Reader(bytes(w.b)).zigzag() == value


def test_reader_rejects_bad_input():
    with pytest.raises(S.DecodeError):
        S.Reader(b"\xff").varint()
    with pytest.raises(S.DecodeError):
        S.Reader(b"\xfc\x00").varint()
    with pytest.raises(S.DecodeError):
        S.Reader(b"\x02").boolean()
    with pytest.raises(S.DecodeError):
        S.Reader(_vec(1 << 33)).varint(32)


def _tiny_checkpoint():
    fluid = S.FluidCheckpoint(np.array([0, 3], dtype=np.uint32),
                              np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]),
                              np.array([[0.5, -0.0, 1e-300], [7.0, 8.0, 9.0]]),
                              np.array([1e-3, 2e-3]),
                              [np.array([[1.5, 2.5, 3.5], [4.5, 5.5, 6.5]])],
                              [np.array([[0.0, 0.0, -1.0], [1.0, 1.0, 1.0]])])
    rb = S.RigidBodyCheckpoint(np.array([7.0, 8.0, 9.0]), np.array([0.0, 0.0, 0.0, 1.0]),
                               np.array([0.1, 0.2, 0.3]), np.array([1.0, 2.0, 3.0]),
                               np.array([-1.0, 0.0, 1.0]), np.array([0.0, 5e-17, 0.0]))
    return S.SystemCheckpoint(300, 0.123456789, fluid, [None, rb, None])


def _f64s(*vals):
    return b"".join(struct.pack("<d", v) for v in vals)


def test_checkpoint_bincode_bytes_in_declaration_order():
    """SerSystemCheckpoint (sph/mod.rs:884-894) -> SerFluidCheckpoint
    (fluid.rs:399-408) -> SerBoundaryCheckpoint (boundary_handling/mod.rs:
    147-151) -> SerRigidBodyMotionCheckpoint (rigid_body_motion.rs:247-256)."""
    cp = _tiny_checkpoint()
    want = (
        b"\xfb" + struct.pack("<H", 300)          # time_steps_propagated: u64
        + _f64s(0.123456789)                      # current_time: f64
        + b"\x02\x00\x03"                         # fluid_id: Vec<u32>
        + b"\x02" + _f64s(1, 2, 3, 4, 5, 6)       # position: Vec<[f64; 3]>
        + b"\x02" + _f64s(0.5, -0.0, 1e-300, 7, 8, 9)
        + b"\x02" + _f64s(1e-3, 2e-3)             # mass: Vec<f64>
        + b"\x01\x02" + _f64s(1.5, 2.5, 3.5, 4.5, 5.5, 6.5)   # integrator_position_slots
        + b"\x01\x02" + _f64s(0, 0, -1, 1, 1, 1)               # integrator_velocity_slots
        + b"\x03"                                 # dynamic_states: Vec<Option<..>>
        + b"\x00"
        + b"\x01" + _f64s(7, 8, 9, 0, 0, 0, 1, 0.1, 0.2, 0.3, 1, 2, 3, -1, 0, 1, 0, 5e-17, 0)
        + b"\x00")
    got = S.encode_checkpoint(cp)
    assert got == want
    back = S.decode_checkpoint(got)
    assert S.encode_checkpoint(back) == want
    with pytest.raises(S.DecodeError):
        S.decode_checkpoint(got + b"\x00")
    with pytest.raises(S.DecodeError):
        S.decode_checkpoint(got[:-3])


def _same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def _check_same(cp, back):
    assert back.time_steps_propagated == cp.time_steps_propagated
    assert struct.pack("<d", back.current_time) == struct.pack("<d", cp.current_time)
    f, g = cp.fluid, back.fluid
    assert np.array_equal(f.fluid_id, g.fluid_id)
    for k in ("position", "velocity", "mass"):
        assert _same(getattr(f, k), getattr(g, k)), k
    for k in ("integrator_position_slots", "integrator_velocity_slots"):
        assert len(getattr(f, k)) == len(getattr(g, k))
        for a, b in zip(getattr(f, k), getattr(g, k)):
            assert _same(a, b)
    assert len(cp.dynamic_states) == len(back.dynamic_states)
    for d, e in zip(cp.dynamic_states, back.dynamic_states):
        assert (d is None) == (e is None)
        if d is not None:
            for name in S.RigidBodyCheckpoint.FIELDS:
                assert _same(getattr(d, name), getattr(e, name)), name


def test_checkpoint_ron_round_trip_keeps_every_bit():
    rng = np.random.default_rng(7)
    cp = _tiny_checkpoint()
    n = 50
    cp.fluid = S.FluidCheckpoint(rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
                                 rng.standard_normal((n, 3)) * 10.0 ** rng.integers(-300, 300, (n, 3)),
                                 rng.standard_normal((n, 3)), rng.random(n), [], [])
    cp.fluid.velocity[0] = [-0.0, 5e-324, 1.7976931348623157e308]
    text = S.checkpoint_to_ron(cp)
    assert text.startswith("(time_steps_propagated:300,current_time:")
    _check_same(cp, S.checkpoint_from_ron(text))
    _check_same(cp, S.decode_checkpoint(S.encode_checkpoint(cp)))


def test_ron_reader_accepts_pretty_form():
    text = """// saved state
    SerSystemCheckpoint(
        time_steps_propagated: 12,
        current_time: 0.5,
        fluid: (
            fluid_id: [1, 2],
            position: [(0.0, 1.0, 2.0), (3.0, 4.0, 5e-3),],
            velocity: [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)],
            mass: [1.0, 2.0],
            integrator_position_slots: [],
            integrator_velocity_slots: [],
        ),
        boundary: (
            dynamic_states: [None, Some((center_of_mass: (7.0, 8.0, 9.0),
                orientation: (0.0, 0.0, 0.0, 1.0), linear_velocity: (0.0, 0.0, 0.0),
                angular_momentum: (0.0, 0.0, 0.0), force: (0.0, 0.0, 0.0),
                torque: (0.0, 0.0, 0.0)))],  /* trailing comment */
        ),
    )
    """
    cp = S.checkpoint_from_ron(text)
    assert cp.time_steps_propagated == 12 and cp.current_time == 0.5
    assert cp.fluid.position[1, 2] == 5e-3
    assert cp.dynamic_states[0] is None
    assert list(cp.dynamic_states[1].center_of_mass) == [7.0, 8.0, 9.0]
    for bad in ("(time_steps_propagated:1)", text.replace("mass", "masses"),
                text + "x", text.replace("fluid_id: [1, 2]", "fluid_id: [-1, 2]")):
        with pytest.raises(S.DecodeError):
            S.checkpoint_from_ron(bad)


def test_boundary_entries_keep_their_places():
    """Round trips of SerBoundaryCheckpoint keep None/Some at their indices,
    including an empty list (tests/sph/boundary_handling/mod.rs:160-255)."""
    for states in ([None, _tiny_checkpoint().dynamic_states[1], None], [None], []):
        cp = _tiny_checkpoint()
        cp.dynamic_states = states
        for back in (S.checkpoint_from_ron(S.checkpoint_to_ron(cp)),
                     S.decode_checkpoint(S.encode_checkpoint(cp))):
            assert [d is None for d in back.dynamic_states] == [d is None for d in states]


def test_checkpoint_files(tmp_path):
    cp = _tiny_checkpoint()
    for fmt in ("ron", "bincode"):
        p = str(tmp_path / f"state.{fmt}")
        S.save_checkpoint(cp, p, fmt)
        _check_same(cp, S.load_checkpoint(p))


def _info(step, n=3, kind="Samples"):
    rng = np.random.default_rng(step)
    meas = {name: float(k) + 0.25 for k, (name, _) in enumerate(S.MEASUREMENT_FIELDS)}
    meas["solver_iterations"] = 300 + step
    if kind == "Samples":
        fluid = {"kind": "Samples", "positions": rng.random((n, 3)).astype(np.float32),
                 "coloring": {"kind": "QuantityGraded", "quantity": "Pressure",
                              "data": rng.random(n).astype(np.float32)}}
    else:
        fluid = {"kind": "SensorPlane", "quantity": "Speed",
                 "planes": [{"positions": rng.random((4, 3)).astype(np.float32),
                             "data": rng.random(4).astype(np.float32), "rows": 2, "cols": 2}]}
    return {"time_step_number": step, "measurement": meas, "fluid": fluid,
            "boundary": {"kind": "Samples", "positions": np.zeros((1, 3), np.float32),
                         "coloring": {"kind": "BoundaryId", "ids": [7], "max_id": 7}}}


def test_time_step_info_bytes():
    info = _info(5, n=1)
    raw = S.encode_time_step_info(info)
    meas = b""
    for name, kind in S.MEASUREMENT_FIELDS:
        v = info["measurement"][name]
        meas += _vec(v) if kind == "u32" else struct.pack("<d", v)
    pos = info["fluid"]["positions"]
    data = info["fluid"]["coloring"]["data"]
    want = (b"\x05" + meas
            + b"\x01" + b"\x01" + pos.astype("<f4").tobytes()        # Samples { positions,
            + b"\x02" + b"\x04" + b"\x01" + data.astype("<f4").tobytes()  # QuantityGraded(Pressure)
            + b"\x01" + b"\x01" + b"\x00" * 12                       # boundary Samples
            + b"\x01" + b"\x01\x07" + b"\x07")                       # BoundaryId
    assert raw == want
    back = S.decode_time_step_info(raw)
    assert S.encode_time_step_info(back) == raw


def test_recording_file_rewrites_reproduced_steps(tmp_path):
    path = str(tmp_path / "rec" / "run.bin")
    sp = S.SimulationParameters(0.02, (1.0, 2.0, 3.0), 100, True, True, False)
    rec = S.Recording(path, sp)
    for k in range(4):
        rec.append_time_step(_info(k, kind="Samples" if k % 2 else "SensorPlane"))
    size4 = os.path.getsize(path)
    # the simulation rewound: step 2 is produced again and replaces 2 and 3
    rec.append_time_step(_info(2))
    sp2, steps = S.read_recording(path)
    assert sp2 == S.SimulationParameters(np.float32(0.02).item(), (1.0, 2.0, 3.0), 100, True,
                                         True, False)
    assert [s["time_step_number"] for s in steps] == [0, 1, 2]
    assert os.path.getsize(path) < size4
    assert steps[2]["fluid"]["kind"] == "Samples"
    assert np.array_equal(steps[0]["fluid"]["planes"][0]["data"],
                          _info(0, kind="SensorPlane")["fluid"]["planes"][0]["data"])
    with pytest.raises(FileExistsError):
        S.Recording(path, sp)
    head = open(path, "rb").read(8)
    assert struct.unpack("<Q", head)[0] == len(S.encode_simulation_parameters(sp))

"""Checkpoint / resume and point-cloud output through the C-ABI
(fg_sim_save_state, fg_sim_load_state, fg_sim_write_ply).

  resume            a run stopped after k steps, saved, restored into a new
                    simulator of the same scene and stepped on gives the same
                    bits (particles, body state, FP64 sums) as the run that
                    never stopped: the scene has inactive reserve slots
                    emitted by an inflow, a dynamic torus and a static
                    cylinder, so every carried-over buffer matters
  round trip        save -> load -> save writes the same bytes
  refusals          a file of another scene, a damaged byte (checksum) and a
                    truncated file are refused with FG_ERR_INVALID_ARGUMENT
                    and leave the simulator's state as it was
  PLY output        the point cloud holds exactly the active particles in
                    particle order, their velocities and the (rho, p) of the
                    last step looked up through the sorted ids (bit-exact)
"""
import numpy as np
import pytest

from fluidgrid_amd import scenes

pytestmark = pytest.mark.gpu

H = 0.04


def _scene(nx=16):
    s = scenes.dam_break((nx, 10, 16), H, (1.0, 1.0, 1.0), 4e-4)
    top = float(s.pos[:, 1].max())
    s = scenes.with_inflow(s, (0.5, 0.2), (0.62, 0.3), speed=0.02 / (3.5 * 4e-4), layers=4)
    tor = scenes.torus_spheres(0.1, 0.03, 0.5 * H) + np.array([0.2, top + 0.03, 0.2])
    cyl = scenes.cylinder_spheres(0.05, 0.3, 0.5 * H) + np.array([0.6, 0.0, 0.6])
    b = scenes.make_bodies([tor], [1.0], H, s.common, static_sets=[cyl], st_buckets=(32, 32, 32))
    return s, b


def _sim(fglib, s, b, graph=True):
    sim = fglib.Simulator(s, fused=True, use_graph=graph)
    sim.add_static(b)
    sim.add_bodies(b)
    sim.set_inflow(s.info["inflow"])
    return sim


def _snapshot(sim):
    pos, vel = sim.state()
    bs = sim.body_state()
    c = sim.conservation()
    return pos, vel, bs, c


def _same(a, b):
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for k in a[2]:
        assert np.array_equal(a[2][k], b[2][k]), k
    for k in a[3]:
        # (the internal energy of a simulator that has not stepped is NaN:
        # its density buffer is still zero)
        assert np.array_equal(np.asarray(a[3][k]), np.asarray(b[3][k]), equal_nan=True), k


@pytest.mark.parametrize("graph", [True, False])
def test_resume_is_bit_identical(fglib, tmp_path, graph):
    s, b = _scene()
    ck = tmp_path / "state.fgs"
    sim = _sim(fglib, s, b, graph)
    try:
        sim.step(6)  # two of the four reserve layers are emitted by then
        sim.save_state(ck)
        active_at_save = int(sim.conservation()["count"])
        sim.step(20)
        want = _snapshot(sim)
    finally:
        sim.close()
    assert want[3]["count"] > active_at_save  # layers were emitted after the checkpoint
    sim = _sim(fglib, s, b, graph)
    try:
        sim.load_state(ck)
        sim.step(20)
        got = _snapshot(sim)
    finally:
        sim.close()
    _same(want, got)
    assert np.isfinite(got[0]).all()


def test_save_load_save_same_bytes(fglib, tmp_path):
    s, b = _scene()
    a, c = tmp_path / "a.fgs", tmp_path / "c.fgs"
    sim = _sim(fglib, s, b)
    try:
        sim.step(7)
        sim.save_state(a)
    finally:
        sim.close()
    sim = _sim(fglib, s, b)
    try:
        sim.load_state(a)
        sim.save_state(c)
    finally:
        sim.close()
    assert a.read_bytes() == c.read_bytes()


def test_refusals_leave_state_unchanged(fglib, tmp_path):
    s, b = _scene()
    good = tmp_path / "good.fgs"
    sim = _sim(fglib, s, b)
    try:
        sim.step(5)
        sim.save_state(good)
    finally:
        sim.close()
    data = good.read_bytes()
    damaged = tmp_path / "damaged.fgs"
    buf = bytearray(data)
    buf[len(buf) // 2] ^= 0x10
    damaged.write_bytes(bytes(buf))
    truncated = tmp_path / "truncated.fgs"
    truncated.write_bytes(data[:-100])
    # a different scene: one more column of particles
    s2, b2 = _scene(nx=17)
    sim = _sim(fglib, s2, b2)
    try:
        before = _snapshot(sim)
        with pytest.raises(fglib.FGError, match="rc=1 .*particle count|rc=1 .*parameters"):
            sim.load_state(good)
        _same(before, _snapshot(sim))
    finally:
        sim.close()
    sim = _sim(fglib, s, b)
    try:
        sim.step(3)
        before = _snapshot(sim)
        with pytest.raises(fglib.FGError, match="checksum"):
            sim.load_state(damaged)
        with pytest.raises(fglib.FGError, match="truncated"):
            sim.load_state(truncated)
        with pytest.raises(fglib.FGError, match="cannot open"):
            sim.load_state(tmp_path / "missing.fgs")
        _same(before, _snapshot(sim))
    finally:
        sim.close()


def _read_ply(path):
    data = open(path, "rb").read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().splitlines()
    assert head[0] == "ply" and head[1] == "format binary_little_endian 1.0"
    n = int(next(h for h in head if h.startswith("element vertex")).split()[-1])
    props = [h.split()[-1] for h in head if h.startswith("property float")]
    rec = np.frombuffer(data[end:], np.float32).reshape(n, len(props))
    return props, rec


def test_ply_point_cloud(fglib, tmp_path):
    """fg_sim_write_ply: the active particles in particle order with their
    velocities and the density/pressure of the last step."""
    s, b = _scene()
    sim = _sim(fglib, s, b)
    try:
        sim.step(8)
        p = tmp_path / "cloud.ply"
        n = sim.write_ply(p)
        pos, vel = sim.state()
        bufs = sim.buffers
        ids = fglib.download_raw(bufs.particleId, (s.n,), np.uint32)
        rp = fglib.download_raw(bufs.densityPressure, (s.n, 2), np.float32)
    finally:
        sim.close()
    props, rec = _read_ply(p)
    assert props == ["x", "y", "z", "vx", "vy", "vz", "density", "pressure"]
    act = pos[:, 3] > 0
    assert n == act.sum() == len(rec) and n < s.n  # reserve slots are left out
    by_particle = np.zeros((s.n, 2), np.float32)
    by_particle[ids] = rp
    assert np.array_equal(rec[:, :3], pos[act, :3]) and np.array_equal(rec[:, 3:6], vel[act, :3])
    assert np.array_equal(rec[:, 6:], by_particle[act])
    assert np.all(rec[:, 6] > 0)

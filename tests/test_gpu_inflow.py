"""GPU parity of the inflow emitter (fg_emit_inflow, fg_sim_set_inflow)
against the oracle.

Tolerances:
  emission on the same arrays    bit-exact positions and velocities, step
                                 counter exact (pure data placement)
  trajectory with inflow         10x the oracle's own divergence under a
                                 reversed bucket walk, at least 1e-4 h;
                                 active count exact
  eager and graph-replayed       bit-identical
  dt beyond the layer period     refused with FG_ERR_INVALID_ARGUMENT before
                                 any launch
"""
import ctypes as C

import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _scene():
    s = scenes.dam_break((16, 10, 16), 0.04, (1.0, 1.0, 1.0), 4e-4)
    return scenes.with_inflow(s, (0.5, 0.2), (0.62, 0.3), speed=0.02 / (3.5 * 4e-4), layers=4)


def test_emit_inflow_matches_oracle(fglib, orc):
    L = fglib.load()
    rng = np.random.default_rng(5)
    s = _scene()
    d = s.info["inflow"]
    rot = abi.InflowDesc.from_buffer_copy(d)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    u = scenes.quat_rotate(q, np.array([[0.0, 1.0, 0.0]]))[0]
    v = scenes.quat_rotate(q, np.array([[0.0, 0.0, 1.0]]))[0]
    rot.axisU, rot.axisV = abi.f3(*u), abi.f3(*v)
    dt = np.float32(s.common.deltaTimeIntegration)
    # the reference scene's simulation scale: the period carries the factor
    scaled = abi.InflowDesc.from_buffer_copy(d)
    cp = abi.CommonFluidParams.from_buffer_copy(s.common)
    for desc, scale in ((d, 1.0), (rot, 1.0), (scaled, 0.009)):
        cp.simulation_scale = scale
        L.gInitializeCommonFluidParams(cp)
        if scale != 1.0:
            desc.speed = float(np.float32(d.speed * scale))
        assert L.fg_inflow_period(C.byref(desc), scale) == orc.inflow_period(desc, scale)
        pos = s.pos.copy()
        vel = rng.normal(size=s.vel.shape).astype(np.float32)
        dp, dv = fglib.DeviceArray.from_numpy(pos), fglib.DeviceArray.from_numpy(vel)
        step = fglib.DeviceArray.from_numpy(np.zeros(1, np.uint32))
        for k in range(24):
            fglib.check(L.fg_emit_inflow(dp.ptr, dv.ptr, step.ptr, desc, float(dt)), "emit")
            orc.emit_inflow(pos, vel, k, desc, dt, scale)
        assert step.numpy()[0] == 24
        assert np.array_equal(dp.numpy(), pos) and np.array_equal(dv.numpy(), vel)
        assert np.count_nonzero(pos[:, 3] > 0) == s.info["active"] + desc.numLayers * desc.nu * desc.nv


def _run(fglib, s, steps, graph):
    sim = fglib.Simulator(s, fused=True, use_graph=graph)
    try:
        sim.set_inflow(s.info["inflow"])
        sim.step(steps)
        pos, vel = sim.state()
        c = sim.conservation()
    finally:
        sim.close()
    return pos, vel, c


def test_inflow_trajectory_matches_oracle(fglib, orc):
    s = _scene()
    d = s.info["inflow"]
    steps = 30
    pos, vel, c = _run(fglib, s, steps, True)
    runs = []
    for rev in (False, True):
        orc.set_reverse_order(rev)
        try:
            p, v, rho = s.pos.copy(), s.vel.copy(), np.zeros((s.n, 2), np.float32)
            for k in range(steps):
                orc.emit_inflow(p, v, k, d, np.float32(s.common.deltaTimeIntegration), s.common.simulation_scale)
                r = orc.step(p, v, s.fluid, s.common, s.kernels, s.key_mode, rho_p_prev=rho)
                p, v, rho = r.pos, r.vel, r.rho_p
        finally:
            orc.set_reverse_order(False)
        runs.append(p)
    fwd, rev = runs
    active = np.count_nonzero(fwd[:, 3] > 0)
    assert active > s.info["active"] + 2 * d.nu * d.nv  # several layers were emitted
    assert np.count_nonzero(pos[:, 3] > 0) == active == int(c["count"])
    h = s.common.kernel_radius
    div = np.max(np.abs(fwd[:, :3] - rev[:, :3])) / h
    err = np.max(np.abs(pos[:, :3] - fwd[:, :3])) / h
    print(f"inflow {steps} steps: max|dx|/h gpu {err:.3e}, order divergence {div:.3e}, active {active}")
    assert err <= max(10 * div, 1e-4)


def test_inflow_graph_equals_eager(fglib):
    s = _scene()
    a = _run(fglib, s, 20, False)
    b = _run(fglib, s, 20, True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_dt_beyond_period_refused(fglib):
    L = fglib.load()
    s = _scene()
    d = s.info["inflow"]
    L.gInitializeCommonFluidParams(s.common)
    period = L.fg_inflow_period(C.byref(d), s.common.simulation_scale)
    pos = fglib.DeviceArray.from_numpy(s.pos)
    vel = fglib.DeviceArray.from_numpy(s.vel)
    step = fglib.DeviceArray.from_numpy(np.zeros(1, np.uint32))
    L.fg_last_error()
    rc = L.fg_emit_inflow(pos.ptr, vel.ptr, step.ptr, d, float(1.5 * period))
    assert rc == abi.ERR_INVALID_ARGUMENT and L.fg_last_error() == abi.ERR_INVALID_ARGUMENT
    assert np.array_equal(pos.numpy(), s.pos) and step.numpy()[0] == 0

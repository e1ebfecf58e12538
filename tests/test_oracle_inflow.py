"""Pins of the oracle's inflow emission (oracle.emit_inflow) and of the scene
helper that reserves its slots.

  schedule        after K steps, min(floor(K dt / period), numLayers) layers
                  are active, period = spacing * scale / speed
  layer distance  with simulation_scale != 1, consecutive layers carried by
                  the integration x += v dt / scale lie `spacing` world units
                  apart, within one step's travel; dt > period is refused
  lattice         emitted slot firstSlot + l nu nv + iu + nu iv sits at
                  origin + iu s U + iv s V (FP32), w = 1, velocity speed * n
  reserve         parked one per cell (no two reserve slots share a cell),
                  inactive, outside the fluid's reach
"""
import math

import numpy as np
import pytest

from fluidgrid_amd import scenes


def _scene():
    s = scenes.dam_break((8, 6, 8), 0.04, (1.0, 1.0, 1.0), 4e-4)
    # one layer every 3.5 steps
    return scenes.with_inflow(s, (0.3, 0.2), (0.42, 0.3), speed=0.02 / (3.5 * 4e-4), layers=3)


def test_schedule_and_lattice(orc):
    s = _scene()
    d, dt = s.info["inflow"], np.float32(s.common.deltaTimeIntegration)
    pos, vel = s.pos.copy(), s.vel.copy()
    n0 = s.info["active"]
    period = float(np.float32(d.spacing)) / float(np.float32(d.speed))
    assert orc.inflow_period(d, 1.0) == period
    per = d.nu * d.nv
    assert per == 7 * 6 and d.firstSlot == n0
    for k in range(30):
        orc.emit_inflow(pos, vel, k, d, dt, 1.0)
        want_layers = min(math.floor((k + 1) * float(dt) / period), d.numLayers)
        assert np.count_nonzero(pos[:, 3] > 0) == n0 + want_layers * per
    for layer in range(d.numLayers):
        for t in range(per):
            iu, iv = t % d.nu, t // d.nu
            p = pos[d.firstSlot + layer * per + t]
            assert p[0] == np.float32(d.origin.x)
            assert p[1] == np.float32(np.float32(d.origin.y) + np.float32(iu) * np.float32(d.spacing))
            assert p[2] == np.float32(np.float32(d.origin.z) + np.float32(iv) * np.float32(d.spacing))
            assert p[3] == 1.0
            assert np.array_equal(vel[d.firstSlot + layer * per + t], np.array([d.speed, 0, 0, 0], np.float32))


def test_reserve_is_parked_one_per_cell(orc):
    s = _scene()
    d = s.info["inflow"]
    res = s.pos[d.firstSlot:]
    assert np.all(res[:, 3] == -1.0) and np.all(s.vel[d.firstSlot:] == 0)
    key, _ = orc.hash_keys(res, s.num_buckets, tuple(s.common.cell_size), s.common, s.key_mode)
    assert len(np.unique(key)) == len(key)
    fluid_top = s.pos[:d.firstSlot, 1].max()
    assert res[:, 1].min() > fluid_top + 10 * s.common.kernel_radius


def test_layer_distance_with_scale(orc):
    """Layers move by v dt / scale per step (K_FluidSim.cu:433-438); with the
    reference scene's scale 0.009 the period must carry that factor, or the
    layers end up spacing / scale apart."""
    s = _scene()
    d = s.info["inflow"]
    scale = np.float32(0.009)
    dt = np.float32(4e-4)
    d.speed = float(np.float32(0.02 * scale / (3.5 * float(dt))))  # one layer every 3.5 steps
    period = orc.inflow_period(d, scale)
    assert abs(period / float(dt) - 3.5) < 1e-5
    pos, vel = s.pos.copy(), s.vel.copy()
    per = d.nu * d.nv
    q = np.float32(dt / scale)
    for k in range(12):
        orc.emit_inflow(pos, vel, k, d, dt, scale)
        act = pos[:, 3] > 0
        act[:d.firstSlot] = False
        pos[act, :3] = pos[act, :3] + vel[act, :3] * q  # free flight of the emitted layers
    layers = [pos[d.firstSlot + l * per:d.firstSlot + (l + 1) * per] for l in range(3)]
    assert all(np.all(L[:, 3] > 0) for L in layers)
    step_travel = float(np.float32(d.speed)) * float(q)
    for a, b in zip(layers, layers[1:]):
        gap = float(np.mean(a[:, 0] - b[:, 0]))
        assert abs(gap - d.spacing) <= step_travel + 1e-6, (gap, d.spacing)


def test_dt_beyond_period_refused(orc):
    s = _scene()
    d = s.info["inflow"]
    period = orc.inflow_period(d, 1.0)
    with pytest.raises(ValueError):
        orc.emit_inflow(s.pos.copy(), s.vel.copy(), 0, d, np.float32(1.5 * period), 1.0)

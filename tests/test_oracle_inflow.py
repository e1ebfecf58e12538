"""Pins of the oracle's inflow emission (oracle.emit_inflow) and of the scene
helper that reserves its slots.

  schedule        after K steps, min(floor(K dt / period), numLayers) layers
                  are active, period = spacing / speed
  lattice         emitted slot firstSlot + l nu nv + iu + nu iv sits at
                  origin + iu s U + iv s V (FP32), w = 1, velocity speed * n
  reserve         parked one per cell (no two reserve slots share a cell),
                  inactive, outside the fluid's reach
"""
import math

import numpy as np

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
    per = d.nu * d.nv
    assert per == 7 * 6 and d.firstSlot == n0
    for k in range(30):
        orc.emit_inflow(pos, vel, k, d, dt)
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

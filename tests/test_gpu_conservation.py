"""GPU FP64 conservation sums (fg_conservation_f64, fg_sim_conservation)
against the oracle's numpy FP64 restatement.

Tolerances:
  sums on the same state          rtol 1e-12 (FP64 sums in another order);
                                  counts exact; bit-identical when repeated
  free fall of isolated particles momentum = N m g t and kinetic energy
                                  = N m (g t)^2 / 2 within rtol 1e-5 (FP32
                                  velocity updates); count and mass exact
"""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu


def test_conservation_matches_numpy(fglib, orc):
    L = fglib.load()
    s = scenes.config_scene(0)
    rng = np.random.default_rng(3)
    n = 100_003
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, 1, (n, 3))
    pos[:, 3] = np.where(rng.uniform(size=n) < 0.9, 1.0, -1.0)
    vel = rng.normal(0, 0.5, (n, 4)).astype(np.float32)
    spos = pos[rng.permutation(n)]
    rho_p = np.stack([rng.uniform(900, 1100, n), np.zeros(n)], 1).astype(np.float32)
    fglib.init_params(s.common, s.kernels, s.key_mode)
    d = {k: fglib.DeviceArray.from_numpy(v) for k, v in dict(p=pos, v=vel, sp=spos, r=rho_p).items()}
    ws = fglib.DeviceArray((L.fg_conservation_workspace_size(),), np.uint8)
    out = fglib.DeviceArray((8,), np.float64)
    m, r0, k = s.fluid.mass, s.fluid.restDensity, s.fluid.gasConstant
    res = []
    for _ in range(2):
        fglib.check(L.fg_conservation_f64(d["p"].ptr, d["v"].ptr, d["sp"].ptr, d["r"].ptr, n, m, r0, k, out.ptr,
                                          ws.ptr, ws.nbytes), "conservation")
        res.append(out.numpy())
    assert np.array_equal(res[0], res[1])
    g = tuple(s.common.global_external_force)
    want = orc.conservation(pos, vel, spos, rho_p, np.float32(m), g, s.common.simulation_scale, np.float32(r0),
                            np.float32(k))
    got = res[0]
    assert got[0] == want[0] and got[1] == pytest.approx(want[1], rel=1e-14)
    assert np.allclose(got[2:], want[2:], rtol=1e-12, atol=1e-12 * np.abs(want[2:]).max())


def test_free_fall_momentum_and_energy(fglib):
    h = 0.02
    s = scenes.dam_break((8, 8, 8), h, (1.0, 1.0, 1.0), 2e-4)
    g = np.arange(8) * 1.6 * h + 0.3  # spacing 1.6 h: no neighbours, away from the walls
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    s.pos = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size)], 1).astype(np.float32)
    s.vel = np.zeros_like(s.pos)
    s.fluid.max_particles = s.n
    steps = 40
    sim = fglib.Simulator(s, fused=True, use_graph=True)
    try:
        c0 = sim.conservation()
        sim.step(steps)
        c1 = sim.conservation()
    finally:
        sim.close()
    n, m = s.n, np.float32(s.fluid.mass)
    gy = s.common.global_external_force.y
    t = steps * float(np.float32(s.common.deltaTimeIntegration))
    assert c0["count"] == c1["count"] == n and c0["mass"] == c1["mass"]
    assert np.allclose(c0["momentum"], 0) and c0["kinetic"] == 0
    assert c1["momentum"][1] == pytest.approx(n * float(m) * gy * t, rel=1e-5)
    assert c1["momentum"][0] == 0 and c1["momentum"][2] == 0
    assert c1["kinetic"] == pytest.approx(0.5 * n * float(m) * (gy * t) ** 2, rel=1e-5)
    # the fall lowers the potential energy by what the kinetic energy gains (to first order in dt)
    dpe = c1["potential"] - c0["potential"]
    assert dpe < 0 and abs(dpe + c1["kinetic"]) < 0.05 * c1["kinetic"]

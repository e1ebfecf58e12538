"""The HIP step against the committed fixtures (tests/golden/): grid arrays
bit-exact, density within 1e-5 relative, positions after the step within
dt * 1e-5 * |v| of the fixture (the force tolerance carried through one
integration)."""
import os

import numpy as np
import pytest

from fluidgrid_amd import scenes

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"cfg0_step1": lambda: scenes.config_scene(0), "appmain_step1": scenes.appmain_scene}


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_step_matches_fixture(fglib, name):
    s = CASES[name]()
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    sim = fglib.Simulator(s, fused=True, use_graph=False)
    try:
        sim.step(1)
        pos, vel = sim.state()
        g = sim.arrays()
    finally:
        sim.close()
    assert np.array_equal(g["skey"], f["skey"])
    assert np.array_equal(g["spid"], f["spid"])
    occ = f["occupied"]
    assert np.array_equal(np.nonzero(g["start"] != 0xFFFFFFFF)[0], occ)
    assert np.array_equal(g["start"][occ], f["start"]) and np.array_equal(g["end"][occ], f["end"])
    act = g["spos"][:, 3] > 0
    assert np.all(np.abs(g["rho_p"][act, 0] - f["rho_p"][act, 0]) <= 1e-5 * f["rho_p"][act, 0])
    dt = s.common.deltaTimeIntegration
    q = dt / s.common.simulation_scale
    dv = np.abs(vel[:, :3] - f["vel"][:, :3])
    # velocity differences come only from FP32 rounding of the pair sums
    assert np.max(dv) <= 1e-3 * max(1.0, float(np.abs(f["vel"][:, :3]).max()))
    dx = np.abs(pos[:, :3] - f["pos"][:, :3])
    tol = dv * q + 2 * np.spacing(np.abs(f["pos"][:, :3])) + np.spacing(np.abs(f["vel"][:, :3] * q))
    bad = np.argwhere(dx > tol)
    for i, c in bad[:5]:
        print(f"particle {i} axis {c}: gpu x {pos[i, c]!r} v {vel[i, c]!r}; fixture x {f['pos'][i, c]!r} "
              f"v {f['vel'][i, c]!r}; start x {s.pos[i, c]!r} w {s.pos[i, 3]!r}")
    assert len(bad) == 0

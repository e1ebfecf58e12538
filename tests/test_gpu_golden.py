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
    assert np.all(np.abs(pos[:, :3] - f["pos"][:, :3]) <= dv * q + 2 * np.spacing(np.abs(f["pos"][:, :3])))

"""GPU parity of whole steps driven through the C++ HeterogeneousSimulator
(include/fluidgrid_amd_sim.h): one step against the oracle step, fused vs
unfused and graph vs eager bit-identity (including graph replays interleaved
with eager work), and 100-step trajectories.

Trajectory tolerance: the oracle is run twice, walking every bucket forward
and backward; both are valid FP32 evaluations.  Positions after 100 steps
must agree with the forward oracle within max(10 * that divergence, 1e-4 h).
"""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu


def test_one_step_matches_oracle(fglib, orc):
    s = scenes.config_scene(0)
    sim = fglib.Simulator(s, fused=True, use_graph=True)
    try:
        sim.step(1)
        pos, vel = sim.state()
        g = sim.arrays()
    finally:
        sim.close()
    st = orc.step(s.pos, s.vel, s.fluid, s.common, s.kernels, s.key_mode, want_acc=True)
    assert np.array_equal(g["skey"], st.skey)
    assert np.array_equal(g["spid"], st.spid)
    assert np.array_equal(g["start"], st.start)
    occ = st.start != 0xFFFFFFFF
    assert np.array_equal(g["end"][occ], st.end[occ])
    assert np.array_equal(g["spos"], st.spos)
    assert np.array_equal(g["svel"], st.svel)
    assert np.all(np.abs(g["rho_p"][:, 0] - st.rho_p[:, 0]) <= 1e-5 * st.rho_p[:, 0])
    dt = s.common.deltaTimeIntegration
    _, _, mag = orc.forces(st.spos, st.svel, st.rho_p.astype(np.float64), st.start, st.end, st.spid, s.fluid,
                           s.common, s.kernels, s.key_mode, f64=True, want_mag=True)
    a_ref = np.zeros((s.n, 1))
    a_ref[st.spid, 0] = np.linalg.norm(st.acc, axis=1) + mag
    tol_v = dt * 1e-5 * a_ref
    assert np.all(np.abs(vel[:, :3] - st.vel[:, :3]) <= tol_v)
    assert np.all(np.abs(pos[:, :3] - st.pos[:, :3]) <= tol_v * dt + 1e-7)


@pytest.mark.parametrize("which", ["cfg0", "appmain"])
def test_fused_graph_bit_identical_to_unfused_eager(fglib, which):
    s = scenes.config_scene(0) if which == "cfg0" else scenes.appmain_scene()
    results = {}
    for fused, graph in ((False, False), (True, True), (True, False), (False, True)):
        sim = fglib.Simulator(s, fused=fused, use_graph=graph)
        try:
            sim.step(5)
            results[(fused, graph)] = sim.state()
        finally:
            sim.close()
    base = results[(False, False)]
    for k, (p, v) in results.items():
        assert np.array_equal(p, base[0]), k
        assert np.array_equal(v, base[1]), k


def test_graph_replays_interleaved_with_eager_work(fglib):
    s = scenes.config_scene(0)
    sim = fglib.Simulator(s, fused=True, use_graph=False)
    try:
        sim.step(1)
        ref_states = []
        for _ in range(4):
            sim.step(2)
            ref_states.append(sim.state())
    finally:
        sim.close()
    sim = fglib.Simulator(s, fused=True, use_graph=True)
    try:
        sim.set_options(fused=True, use_graph=False)
        sim.step(1)  # eager
        got = []
        for k in range(4):
            sim.set_options(fused=True, use_graph=True)
            sim.step(2)  # replays of a freshly captured graph
            sim.synchronize()
            got.append(sim.state())
            # eager host work between replays
            _ = sim.arrays()
        for (p, v), (rp, rv) in zip(got, ref_states):
            assert np.array_equal(p, rp) and np.array_equal(v, rv)
    finally:
        sim.close()


@pytest.mark.parametrize("math_mode", [abi.MATH_FAST, abi.MATH_REFERENCE])
def test_trajectory_100_steps_cfg0(fglib, orc, math_mode):
    s = scenes.config_scene(0)
    steps = 100
    sim = fglib.Simulator(s, math_mode=math_mode, fused=True, use_graph=True)
    try:
        sim.step(steps)
        pos, vel = sim.state()
    finally:
        sim.close()
    p_fwd, v_fwd, _ = orc.run_steps(s, steps)
    p_rev, _, _ = orc.run_steps(s, steps, reverse=True)
    h = s.common.kernel_radius
    div = np.max(np.abs(p_fwd[:, :3] - p_rev[:, :3])) / h
    err = np.max(np.abs(pos[:, :3] - p_fwd[:, :3])) / h
    tol = max(10 * div, 1e-4)
    print(f"100-step max|dx|/h: gpu-vs-oracle {err:.3e}, oracle order divergence {div:.3e}, tol {tol:.3e}")
    assert err <= tol
    assert np.array_equal(pos[:, 3], p_fwd[:, 3])
    # mass is exact (fixed particle count and mass); momentum drift reported
    m = s.fluid.mass
    print("momentum gpu", (m * vel[:, :3].astype(np.float64)).sum(0), "oracle",
          (m * v_fwd[:, :3].astype(np.float64)).sum(0))


def test_trajectory_100_steps_appmain_ref_xor(fglib, orc):
    s = scenes.appmain_scene()
    steps = 100
    sim = fglib.Simulator(s, fused=True, use_graph=True)
    try:
        sim.step(steps)
        pos, _ = sim.state()
    finally:
        sim.close()
    p_fwd, _, _ = orc.run_steps(s, steps)
    p_rev, _, _ = orc.run_steps(s, steps, reverse=True)
    h = s.common.kernel_radius / s.common.simulation_scale  # world units
    div = np.max(np.abs(p_fwd[:, :3] - p_rev[:, :3])) / h
    err = np.max(np.abs(pos[:, :3] - p_fwd[:, :3])) / h
    tol = max(10 * div, 1e-4)
    print(f"AppMain 100-step max|dx|/h: gpu-vs-oracle {err:.3e}, order divergence {div:.3e}")
    assert err <= tol


@pytest.mark.parametrize("which", ["cfg0", "compressed", "coupled", "ragged"])
def test_schedules_and_block_sizes_bit_identical(fglib, which):
    """The LDS-tiled passes and the flat and row-by-row schedules of the
    density and force passes, at every workgroup size, give the same bits
    after 10 graph steps ('compressed' has rows longer than the mask width,
    which the flat force schedule hands to the row loop and the tiled passes
    to their mask fallback, runs longer than the row loop's LDS slices, and
    tiles past the LDS capacity; 'coupled' runs the coupled force kernels
    with a torus resting in the surface; 'ragged' has a particle count that
    is not a multiple of 64, so the last wavefront is partial and reads its
    rows from memory, and a cluster of inactive particles, so full
    wavefronts stage rows for lanes that have none)."""
    bodies = None
    if which == "cfg0":
        s = scenes.config_scene(0)
    elif which == "coupled":
        s = scenes.small_dam_break(lattice=(24, 10, 24), h=0.02, box=(0.5, 0.5, 0.5))
        top = float(s.pos[:, 1].max())
        tor = scenes.torus_spheres(0.06, 0.02, 0.01) + np.array([0.12, top, 0.12])
        bodies = scenes.make_bodies([tor], [0.5], 0.02, s.common)
    elif which == "ragged":
        s = scenes.small_dam_break(lattice=(23, 9, 21), h=0.03, box=(0.8, 0.8, 0.8))
        assert s.n % 64 != 0
        # Every 11th particle inactive, moved into a cluster far (> 10 h)
        # from the fluid: an inactive particle's (rho, p) is never computed,
        # so one within h of the fluid would feed its neighbours the initial
        # zero density (non-finite forces, whatever the schedule).
        idx = np.arange(0, s.n, 11)
        g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
        s.pos[idx, :3] = 0.6 + 0.015 * g[: len(idx)]
        s.pos[idx, 3] = 0.0
    else:
        s = scenes.small_dam_break(lattice=(16, 16, 16), h=0.04)
        s.pos[:, :3] = 0.3 + (s.pos[:, :3] - 0.3) * 0.55
    L = fglib.load()
    ref = None
    try:
        for tiled, fd, ff, blk in ((0, 0, 0, 256), (0, 1, 0, 64), (0, 0, 1, 128), (0, 1, 1, 512), (0, 1, 1, 256),
                                   (1, 0, 0, 256)):
            L.fg_set_tiled(tiled)
            L.fg_set_schedule(fd, ff)
            L.fg_set_tuning(blk, 8)
            sim = fglib.Simulator(s, fused=True, use_graph=True)
            try:
                if bodies is not None:
                    sim.add_bodies(bodies)
                sim.step(10)
                pos, vel = sim.state()
            finally:
                sim.close()
            if ref is None:
                ref = (pos, vel)
            else:
                assert np.array_equal(pos, ref[0]) and np.array_equal(vel, ref[1]), (tiled, fd, ff, blk)
    finally:
        L.fg_set_tiled(0)
        L.fg_set_schedule(0, 0)
        L.fg_set_tuning(256, 8)


def test_single_fluid_tuning_changes_no_result_and_restores_state(fglib):
    """The single-fluid tuner times launch shapes from one saved state and
    puts the state back (particles, bodies); stepping on with its choice
    gives the bits of the default shape."""
    s = scenes.small_dam_break(lattice=(24, 10, 24), h=0.02, box=(0.5, 0.5, 0.5))
    top = float(s.pos[:, 1].max())
    tor = scenes.torus_spheres(0.06, 0.02, 0.01) + np.array([0.12, top, 0.12])
    bodies = scenes.make_bodies([tor], [0.5], 0.02, s.common)
    L = fglib.load()
    try:
        sim = fglib.Simulator(s, fused=True, use_graph=True)
        try:
            sim.add_bodies(bodies)
            sim.step(3)
            before, bbefore = sim.state(), sim.body_state()
            t = sim.autotune(trial_steps=2)
            after, bafter = sim.state(), sim.body_state()
            assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
            for k in bbefore:
                assert np.array_equal(bbefore[k], bafter[k]), k
            assert t["sph_block"] in (64, 128, 256, 512) and 6 <= t["sort_digit_bits"] <= 8
            assert len(t["trials"]) >= 8
            sim.step(7)
            tuned = sim.state()
        finally:
            sim.close()
        L.fg_set_tuning(256, 8)
        L.fg_set_schedule(0, 0)
        L.fg_set_tiled(0)
        sim = fglib.Simulator(s, fused=True, use_graph=True)
        try:
            sim.add_bodies(bodies)
            sim.step(10)
            ref = sim.state()
        finally:
            sim.close()
        assert np.array_equal(tuned[0], ref[0]) and np.array_equal(tuned[1], ref[1])
    finally:
        L.fg_set_tuning(256, 8)
        L.fg_set_schedule(0, 0)
        L.fg_set_tiled(0)

"""GPU parity of rigid bodies in the batched-scene form (fg_bsim_add_bodies):
every scene of a batch, with its bodies, advances bit for bit as the same
scene and bodies stepped alone through the single-fluid coupled simulator
(HeterogeneousSimulator + RigidBodySimulator), whose own parity against the
oracle is tested in test_gpu_rigid.py and test_gpu_bodies.py.

  mixed batch        scenes with 0, 1 and several bodies (floating at the
                     surface, neighbours touching) together: particles,
                     velocities and body states bit-identical to the
                     single-scene runs after 320 graph-replayed steps; the
                     coupling acts (each scene's fluid ends differently
                     from the same fluid without its bodies)
  eager and graph    bit-identical
  launch tuning      restores particle and body state; no result changes
  argument checks    bodies for a missing scene, a body without spheres, a
                     second set for a scene and bodies after the first step
                     are refused
  configuration 3    the full 256-scene ensemble with its 479 bodies for
                     500 steps: finite, active flags kept, the sorted ids a
                     permutation, densities at least 1, and three scenes
                     (several bodies, one body, the largest scene) bit for
                     bit as when stepped alone
"""
import numpy as np
import pytest

from fluidgrid_amd import scenes

pytestmark = pytest.mark.gpu

STEPS = 320


def _ensemble():
    # seed chosen so that the draw holds scenes with 0, 1 and 2+ bodies
    ens = scenes.ensemble(num_scenes=7, seed=5, n_range=(1500, 6000), bodies=True)
    counts = [0 if b is None else b.b for b in ens.bodies]
    assert 0 in counts and 1 in counts and max(counts) >= 2, counts
    return ens


def _batch(fglib, ens, steps, graph=True, overlap=True, tuning=None):
    b = fglib.BatchSimulator(ens, use_graph=graph)
    try:
        b.set_overlap(overlap)
        if tuning is not None:
            b.set_tuning(*tuning)
        for s, bs in enumerate(ens.bodies):
            if bs is not None:
                b.add_bodies(s, bs)
        b.step(steps)
        pos, vel = b.state()
        return pos, vel, b.body_state()
    finally:
        b.close()


def _alone(fglib, sc, bodies, steps):
    sim = fglib.Simulator(sc, fused=True, use_graph=True)
    try:
        if bodies is not None:
            sim.add_bodies(bodies)
        sim.step(steps)
        pos, vel = sim.state()
        bs = sim.body_state() if bodies is not None else None
        return pos, vel, bs
    finally:
        sim.close()


def test_every_scene_matches_its_single_fluid_run(fglib):
    ens = _ensemble()
    pos, vel, bst = _batch(fglib, ens, STEPS)
    off, boff = 0, 0
    for s, sc in enumerate(ens.scenes):
        p1, v1, b1 = _alone(fglib, sc, ens.bodies[s], STEPS)
        assert np.array_equal(pos[off:off + sc.n], p1), s
        assert np.array_equal(vel[off:off + sc.n], v1), s
        off += sc.n
        if b1 is not None:
            nb = ens.bodies[s].b
            for k in ("pos", "vel", "orient", "L"):
                assert np.array_equal(bst[k][boff:boff + nb], b1[k]), (s, k)
            assert np.any(b1["L"][:, :3] != 0.0)  # torques from the fluid (and neighbours) acted
            p0, _, _ = _alone(fglib, sc, None, STEPS)
            assert not np.array_equal(p0, p1), s  # the bodies pushed the fluid
            boff += nb
    assert off == len(pos) and boff == len(bst["pos"])
    assert np.isfinite(pos).all() and np.isfinite(bst["pos"]).all()


def test_eager_and_graph_identical(fglib):
    ens = _ensemble()
    a = _batch(fglib, ens, 40, graph=True)
    b = _batch(fglib, ens, 40, graph=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for k in a[2]:
        assert np.array_equal(a[2][k], b[2][k]), k


@pytest.mark.parametrize("graph", [True, False])
def test_overlapped_and_inline_body_stages_identical(fglib, graph):
    """Body stages on their own stream (the default) give the bits of the
    in-line order, replayed as a graph and eagerly."""
    ens = _ensemble()
    a = _batch(fglib, ens, 40, graph=graph, overlap=True)
    b = _batch(fglib, ens, 40, graph=graph, overlap=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for k in a[2]:
        assert np.array_equal(a[2][k], b[2][k]), k


@pytest.mark.parametrize("tuning", [(256, 8, 0, 0, 1, 1), (128, 8, 0, 1, 4, 1), (64, 8, 1, 1, 1, 0),
                                    (256, 8, 0, 0, 0, 0), (512, 8, 1, 0, 0, 0), (128, 7, 1, 1, 0, 1)])
def test_queue_and_deferred_contact_identical(fglib, tuning):
    """The persistent work-queue form of the density and force launches and
    the coupled force pass with or without deferred contact give the bits of
    the default launches (one workgroup per block, deferred contact)."""
    ens = _ensemble()
    L = fglib.load()
    try:
        a = _batch(fglib, ens, 30)
        b = _batch(fglib, ens, 30, tuning=tuning)
    finally:
        L.fg_set_schedule(0, 0)
        L.fg_set_tuning(256, 8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for k in a[2]:
        assert np.array_equal(a[2][k], b[2][k]), k


def test_tuning_restores_bodies(fglib):
    ens = _ensemble()
    ref = _batch(fglib, ens, 30)
    L = fglib.load()
    b = fglib.BatchSimulator(ens)
    try:
        for s, bs in enumerate(ens.bodies):
            if bs is not None:
                b.add_bodies(s, bs)
        b.autotune(trial_steps=2)
        b.step(30)
        pos, vel = b.state()
        bst = b.body_state()
    finally:
        b.close()
        L.fg_set_schedule(0, 0)
        L.fg_set_tuning(256, 8)
    assert np.array_equal(pos, ref[0]) and np.array_equal(vel, ref[1])
    for k in bst:
        assert np.array_equal(bst[k], ref[2][k]), k


def test_add_bodies_argument_checks(fglib):
    ens = _ensemble()
    s0 = next(s for s, b in enumerate(ens.bodies) if b is not None)
    L = fglib.load()
    b = fglib.BatchSimulator(ens)
    try:
        with pytest.raises(fglib.FGError, match="no such scene"):
            b.add_bodies(len(ens.scenes), ens.bodies[s0])
        bad = ens.bodies[s0]
        offs = bad.offsets.copy()
        bad.offsets = offs.copy()
        bad.offsets[1] = bad.offsets[0]  # a body without spheres
        try:
            with pytest.raises(fglib.FGError, match="without spheres"):
                b.add_bodies(s0, bad)
        finally:
            bad.offsets = offs
        b.add_bodies(s0, ens.bodies[s0])
        with pytest.raises(fglib.FGError, match="already has bodies"):
            b.add_bodies(s0, ens.bodies[s0])
        b.step(1)
        s1 = next(s for s, x in enumerate(ens.bodies) if x is not None and s != s0)
        with pytest.raises(fglib.FGError, match="before the first step"):
            b.add_bodies(s1, ens.bodies[s1])
        L.fg_last_error()
    finally:
        b.close()


def test_configuration_3_full_ensemble_500_steps(fglib):
    ens = scenes.ensemble(bodies=True)
    steps = 500
    b = fglib.BatchSimulator(ens)
    try:
        for sc, bs in enumerate(ens.bodies):
            if bs is not None:
                b.add_bodies(sc, bs)
        b.step(steps)
        pos, vel = b.state()
        bst = b.body_state()
        rp, pid = b.density()
    finally:
        b.close()
    w0 = np.concatenate([sc.pos[:, 3] for sc in ens.scenes])
    assert np.array_equal(pos[:, 3], w0)
    assert np.isfinite(pos).all() and np.isfinite(vel).all() and np.isfinite(bst["pos"]).all()
    assert np.array_equal(np.sort(pid), np.arange(ens.n, dtype=np.uint32))
    assert np.all(rp[:, 0] >= 1.0)
    counts = [0 if x is None else x.b for x in ens.bodies]
    sizes = [sc.n for sc in ens.scenes]
    picks = {counts.index(max(counts)), counts.index(1), int(np.argmax(sizes))}
    p_off = np.concatenate([[0], np.cumsum(sizes)])
    b_off = np.concatenate([[0], np.cumsum(counts)])
    for s in sorted(picks):
        p1, v1, b1 = _alone(fglib, ens.scenes[s], ens.bodies[s], steps)
        assert np.array_equal(pos[p_off[s]:p_off[s + 1]], p1), s
        assert np.array_equal(vel[p_off[s]:p_off[s + 1]], v1), s
        for k in ("pos", "vel", "orient", "L") if b1 is not None else ():
            assert np.array_equal(bst[k][b_off[s]:b_off[s + 1]], b1[k]), (s, k)

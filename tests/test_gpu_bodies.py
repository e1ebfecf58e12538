"""GPU parity of body-body and body-static contact (K_RigidBodySim.cu:18-61)
through the C-ABI against the oracle.

Tolerances:
  per-body bounds                      bit-exact (min/max of the centres)
  contact forces and torques           |d| <= 1e-5 |F| + 1e-6 max|F| (the same
                                        IEEE operations in the same order; the
                                        dot and cross products may contract
                                        into FMAs)
  with and without the broad phase     bit-identical
  coupled trajectory                   10x the oracle's own divergence under a
                                        reversed bucket walk, at least 1e-4 h
"""
import ctypes as C

import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu

CELL, NB = 0.012, 4096


def _dev(lib, a):
    return lib.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _bodies(seed=4, nbody=3):
    rng = np.random.default_rng(seed)
    sets = []
    for b in range(nbody):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        c = np.array([0.5 + 0.045 * b, 0.5 + 0.01 * b, 0.5])
        sets.append(scenes.quat_rotate(q, scenes.box_spheres(0.06, 0.009)) + c)
    world = np.concatenate([np.concatenate([s, np.full((len(s), 1), b)], 1) for b, s in enumerate(sets)])
    offs = np.concatenate([[0], np.cumsum([len(s) for s in sets])]).astype(np.uint32)
    plate = np.stack(np.meshgrid(np.arange(0.4, 0.7, 0.009), [0.455], np.arange(0.4, 0.6, 0.009)), -1).reshape(-1, 3)
    st = np.concatenate([plate, np.zeros((len(plate), 1))], 1)
    vel = rng.normal(0, 0.3, (nbody, 4)).astype(np.float32)
    vel[:, 3] = 0
    cm = np.zeros((nbody, 4), np.float32)
    for b, s in enumerate(sets):
        cm[b, :3] = s.mean(0)
    return world.astype(np.float32), offs, st.astype(np.float32), vel, cm


def _grid(orc, spheres):
    key, pid = orc.hash_keys(spheres, NB, (CELL,) * 3, abi.CommonFluidParams(), abi.KEY_REF_XOR)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, pos, _ = orc.cell_offsets(sk, sp, spheres, None, NB)
    return pos, start, end, orc.grid_of_bodies(CELL, NB)


def _body_grid(d, prefix, m, nbodies, radius, ks, kd, bounds=None):
    g = abi.BodyGrid()
    g.orderedSpheres = d[prefix + "pos"].ptr
    g.bodyVelocities = d["vel"].ptr if nbodies else None
    g.gridStart, g.gridEnd, g.bounds = d[prefix + "start"].ptr, d[prefix + "end"].ptr, bounds
    g.cellSize = abi.Float3(CELL, CELL, CELL)
    g.numBuckets, g.relation = NB, abi.Int3(1, 1, 1)
    g.numSpheres, g.numBodies = m, nbodies
    g.fluidRadius, g.sphereRadius, g.kSpring, g.kDamping = 0.0, radius, ks, kd
    return g


def _close(g, o):
    scale = np.abs(o[:, :3]).max()
    return np.all(np.abs(g[:, :3].astype(np.float64) - o[:, :3]) <= 1e-5 * np.abs(o[:, :3]) + 1e-6 * scale)


def test_body_bounds_match_numpy(fglib):
    L = fglib.load()
    world, offs, _, _, _ = _bodies()
    d = {k: _dev(fglib, v) for k, v in dict(w=world, o=offs).items()}
    nb = len(offs) - 1
    out = fglib.DeviceArray((2 * nb, 4), np.float32)
    margin = np.float32(0.0101)
    fglib.check(L.fg_body_bounds(d["w"].ptr, d["o"].ptr, nb, float(margin), out.ptr), "bounds")
    got = out.numpy()
    for b in range(nb):
        s = world[offs[b]:offs[b + 1], :3]
        assert np.array_equal(got[2 * b, :3], s.min(0) - margin)
        assert np.array_equal(got[2 * b + 1, :3], s.max(0) + margin)


def test_body_forces_from_bodies_match_oracle(fglib, orc):
    L = fglib.load()
    world, offs, st, vel, cm = _bodies()
    nb, m = len(offs) - 1, len(world)
    r = 0.0045
    ks, kd = -2000.0, -50.0
    dpos, dstart, dend, dgrid = _grid(orc, world)
    spos, sstart, send, sgrid = _grid(orc, st)
    of, ot = orc.body_forces_from_bodies(world, vel, cm, r, ks, kd, dyn=(dpos, dstart, dend, dgrid, r),
                                         static=(spos, sstart, send, sgrid, r))
    assert np.count_nonzero(np.linalg.norm(of[:, :3], axis=1)) > 100
    d = {k: _dev(fglib, v) for k, v in dict(w=world, o=offs, vel=vel, cm=cm, dpos=dpos, dstart=dstart, dend=dend,
                                           spos=spos, sstart=sstart, send=send).items()}
    bb = fglib.DeviceArray((2 * nb, 4), np.float32)
    fglib.check(L.fg_body_bounds(d["w"].ptr, d["o"].ptr, nb, (2 * r + 0.001) * 1.01, bb.ptr), "bounds")
    sb = fglib.DeviceArray((2, 4), np.float32)
    fglib.check(L.fg_sphere_bounds(d["spos"].ptr, len(st), (2 * r + 0.001) * 1.01, sb.ptr), "static bounds")
    gd = _body_grid(d, "d", m, nb, r, ks, kd)
    results = []
    for use_bounds in (False, True):
        gs = _body_grid(d, "s", len(st), 0, r, ks, kd, bounds=sb.ptr if use_bounds else None)
        f = fglib.DeviceArray((m, 4), np.float32)
        t = fglib.DeviceArray((m, 4), np.float32)
        fglib.check(L.fg_body_forces_from_bodies(d["w"].ptr, d["vel"].ptr, d["cm"].ptr, f.ptr, t.ptr, m, nb, r,
                                                 C.byref(gd), bb.ptr if use_bounds else None, C.byref(gs), 0),
                    "body forces")
        results.append((f.numpy(), t.numpy()))
    (gf, gt), (gf2, gt2) = results
    assert np.array_equal(gf, gf2) and np.array_equal(gt, gt2)  # the broad phase changes nothing
    assert _close(gf, of) and _close(gt, ot)
    assert np.array_equal(gf[:, 3], of[:, 3]) and np.array_equal(gt[:, 3], ot[:, 3])
    print("bit-exact fraction (forces):", np.mean(gf == of))
    # accumulate onto given arrays
    rng = np.random.default_rng(2)
    f0 = rng.normal(size=(m, 4)).astype(np.float32)
    t0 = rng.normal(size=(m, 4)).astype(np.float32)
    fa, ta = _dev(fglib, f0), _dev(fglib, t0)
    gs = _body_grid(d, "s", len(st), 0, r, ks, kd)
    fglib.check(L.fg_body_forces_from_bodies(d["w"].ptr, d["vel"].ptr, d["cm"].ptr, fa.ptr, ta.ptr, m, nb, r,
                                             C.byref(gd), bb.ptr, C.byref(gs), 1), "accumulate")
    assert np.array_equal(fa.numpy()[:, :3], f0[:, :3] + gf[:, :3])
    assert np.array_equal(ta.numpy()[:, :3], t0[:, :3] + gt[:, :3])


def test_reference_launcher_matches_oracle(fglib, orc):
    """gCalculateRigidBodyForcesFromRB: the sorted sphere array is both the
    query set and the grid, radius and margin 0.001 as in the reference."""
    L = fglib.load()
    world, offs, st, vel, cm = _bodies(seed=7)
    m = len(world)
    r = 0.0045
    ks, kd = -2000.0, -50.0
    dpos, dstart, dend, dgrid = _grid(orc, world)
    spos, sstart, send, sgrid = _grid(orc, st)
    of, ot = orc.body_forces_from_bodies(dpos, vel, cm, r, ks, kd, dyn=(dpos, dstart, dend, dgrid, r),
                                         static=(spos, sstart, send, sgrid, r))
    d = {k: _dev(fglib, v) for k, v in dict(dpos=dpos, vel=vel, cm=cm, dstart=dstart, dend=dend, spos=spos,
                                           sstart=sstart, send=send).items()}
    f = fglib.DeviceArray((m, 4), np.float32)
    t = fglib.DeviceArray((m, 4), np.float32)
    L.gCalculateRigidBodyForcesFromRB(d["dpos"].ptr, d["vel"].ptr, d["cm"].ptr, t.ptr, f.ptr, d["dstart"].ptr,
                                      d["dend"].ptr, m, abi.Float3(CELL, CELL, CELL), r, NB, NB, d["spos"].ptr,
                                      d["sstart"].ptr, d["send"].ptr, ks, kd, 1)
    fglib.check(None, "gCalculateRigidBodyForcesFromRB")
    gf, gt = f.numpy(), t.numpy()
    assert np.count_nonzero(np.linalg.norm(of[:, :3], axis=1)) > 100
    assert _close(gf, of) and _close(gt, ot)


def _stack_scene(h=0.02):
    """A small fluid block in one corner, a static plate, one box resting on
    the plate and a second box pressed onto the first."""
    s = scenes.dam_break((12, 8, 12), h, (1.0, 1.0, 1.0), 2e-4)
    d = 0.5 * h
    plate = np.stack(np.meshgrid(np.arange(0.4, 0.8, d), [0.2], np.arange(0.4, 0.8, d)), -1).reshape(-1, 3)
    box = scenes.box_spheres(0.1, d)
    lo = box + np.array([0.6, 0.2 + 0.05 + 0.012, 0.6])
    hi = box + np.array([0.62, 0.2 + 0.15 + 0.022, 0.6])
    b = scenes.make_bodies([lo, hi], [0.7, 0.7], h, s.common, static_sets=[plate], st_buckets=(32, 32, 32))
    return s, b


def test_stacked_boxes_trajectory_matches_oracle(fglib, orc):
    s, b = _stack_scene()
    steps = 30
    sim = fglib.Simulator(s, fused=True, use_graph=True)
    try:
        sim.add_static(b)
        sim.add_bodies(b)
        sim.step(steps)
        pos, _ = sim.state()
        bs = sim.body_state()
        sums = sim.body_sums()
    finally:
        sim.close()
    runs = []
    for rev in (False, True):
        orc.set_reverse_order(rev)
        try:
            st = orc.CoupledState(s, b)
            for _ in range(steps):
                st = orc.coupled_step(st)
        finally:
            orc.set_reverse_order(False)
        runs.append(st)
    fwd, rev = runs
    h = s.common.kernel_radius
    # both boxes are in contact (plate/box and box/box): non-zero sums
    assert np.all(np.linalg.norm(fwd.sum_f, axis=1) > 0)
    div = np.max(np.abs(fwd.pos[:, :3] - rev.pos[:, :3])) / h
    err = np.max(np.abs(pos[:, :3] - fwd.pos[:, :3])) / h
    assert err <= max(10 * div, 1e-4)
    bdiv = np.max(np.abs(fwd.cm[:, :3] - rev.cm[:, :3])) / h
    berr = np.max(np.abs(bs["pos"][:, :3] - fwd.cm[:, :3])) / h
    print(f"stacked boxes {steps} steps: body |dx|/h gpu {berr:.3e}, order divergence {bdiv:.3e}; "
          f"sum F_y {sums[0][:, 1]}")
    assert berr <= max(10 * bdiv, 1e-5)
    assert np.allclose(sums[0][:, :3], fwd.sum_f, rtol=1e-3, atol=1e-3 * np.abs(fwd.sum_f).max())

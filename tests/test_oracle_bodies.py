"""Pins of the oracle's body-body and body-static contact
(K_RigidBodySim.cu:18-61, orc_body_forces_from_bodies).

  two spheres of different bodies       closed form
        F = (ks*l - kd*dot(n, v_j - v_i)) n within r_i + r_j + 0.001, with
        the body velocities; nothing beyond; nothing between spheres of one
        body (the w-group skip)
  static sphere                         closed form with v_j = 0
  random bodies + static spheres        O(M^2) FP64 brute force over all
        sphere pairs of different bodies and all sphere/static pairs, each
        pair counted as often as the hashed stencil visits the bucket of the
        other sphere; torque = (x - cm) x F exactly
  accumulate                            adds exactly the written result to
        the given forces and torques (float adds)
"""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

CELL, NB = 0.012, 4096


def _grid(orc, spheres, cell=CELL, nb=NB):
    key, pid = orc.hash_keys(spheres, nb, (cell,) * 3, abi.CommonFluidParams(), abi.KEY_REF_XOR)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, pos, _ = orc.cell_offsets(sk, sp, spheres, None, nb)
    return pos, start, end, orc.grid_of_bodies(cell, nb)


def _visits(orc, pos, cell=CELL, nb=NB):
    """The 27 buckets the reference stencil visits for each query sphere."""
    c32 = np.float32(cell)
    out = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                q = pos.copy()
                for ax, k in ((0, dx), (1, dy), (2, dz)):
                    q[:, ax] = pos[:, ax] + np.float32(c32 * np.float32(k))
                out.append(orc.hash_keys(q, nb, (cell,) * 3, abi.CommonFluidParams(), abi.KEY_REF_XOR)[0])
    return np.stack(out, 1)


@pytest.mark.parametrize("dist", [0.004, 0.0085, 0.0104, 0.02])
def test_two_body_contact_closed_form(orc, dist):
    r = 0.0045
    ks, kd = -800.0, -30.0
    vel = np.array([[0.3, -0.1, 0.0, 0.0], [-0.2, 0.4, 0.1, 0.0]], np.float32)
    sph = np.array([[0.5, 0.5, 0.5, 0.0], [0.5 + dist, 0.5, 0.5, 1.0]], np.float32)
    cm = np.array([[0.45, 0.5, 0.5, 0.0], [0.6, 0.5, 0.5, 0.0]], np.float32)
    dyn = (*_grid(orc, sph), r)
    f, t = orc.body_forces_from_bodies(sph, vel, cm, r, ks, kd, dyn=dyn)
    l = float(np.float32(0.5 + dist)) - float(np.float32(0.5))
    if l < 2 * r + 0.001:
        n = np.array([1.0, 0.0, 0.0])
        expect = (ks * l - kd * np.dot(n, vel[1, :3].astype(np.float64) - vel[0, :3])) * n
    else:
        expect = np.zeros(3)
    assert np.allclose(f[0, :3], expect, rtol=1e-5, atol=1e-9)
    assert f[0, 3] == 0.0 and f[1, 3] == 1.0 and t[1, 3] == 1.0
    # the same two spheres as one body: no force
    one = sph.copy()
    one[:, 3] = 0.0
    f1, _ = orc.body_forces_from_bodies(one, vel, cm, r, ks, kd, dyn=(*_grid(orc, one), r))
    assert np.all(f1[:, :3] == 0.0)


def test_static_contact_closed_form(orc):
    r, rs = 0.0045, 0.005
    ks, kd = -800.0, -30.0
    vel = np.array([[0.3, -0.6, 0.1, 0.0]], np.float32)
    sph = np.array([[0.5, 0.5, 0.5, 0.0]], np.float32)
    cm = np.array([[0.5, 0.55, 0.5, 0.0]], np.float32)
    st = np.array([[0.5, 0.5 - 0.007, 0.5, 0.0]], np.float32)
    f, t = orc.body_forces_from_bodies(sph, vel, cm, r, ks, kd, static=(*_grid(orc, st), rs))
    l = float(np.float32(0.5)) - float(np.float32(0.5 - 0.007))
    n = np.array([0.0, -1.0, 0.0])
    expect = (ks * l - kd * np.dot(n, -vel[0, :3].astype(np.float64))) * n
    assert np.allclose(f[0, :3], expect, rtol=1e-5)
    F = f[0, :3]
    d = (sph[0, :3] - cm[0, :3]).astype(np.float32)
    exact = np.array([d[1] * F[2] - d[2] * F[1], d[2] * F[0] - d[0] * F[2], d[0] * F[1] - d[1] * F[0]], np.float32)
    assert np.array_equal(t[0, :3], exact)


def _random_bodies(seed=4, nbody=3):
    rng = np.random.default_rng(seed)
    sets = []
    for b in range(nbody):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        c = np.array([0.5 + 0.045 * b, 0.5 + 0.01 * b, 0.5])  # overlapping neighbours
        sets.append(scenes.quat_rotate(q, scenes.box_spheres(0.06, 0.009)) + c)
    world = np.concatenate([np.concatenate([s, np.full((len(s), 1), b)], 1) for b, s in enumerate(sets)])
    plate = np.stack(np.meshgrid(np.arange(0.4, 0.7, 0.009), [0.455], np.arange(0.4, 0.6, 0.009)), -1).reshape(-1, 3)
    st = np.concatenate([plate, np.zeros((len(plate), 1))], 1)
    vel = rng.normal(0, 0.3, (nbody, 4)).astype(np.float32)
    vel[:, 3] = 0
    cm = np.zeros((nbody, 4), np.float32)
    for b, s in enumerate(sets):
        cm[b, :3] = s.mean(0)
    return world.astype(np.float32), st.astype(np.float32), vel, cm


def test_random_bodies_brute_force(orc):
    world, st, vel, cm = _random_bodies()
    r, rs = 0.0045, 0.0045
    ks, kd = -2000.0, -50.0
    dpos, dstart, dend, dgrid = _grid(orc, world)
    spos, sstart, send, sgrid = _grid(orc, st)
    f, t = orc.body_forces_from_bodies(world, vel, cm, r, ks, kd, dyn=(dpos, dstart, dend, dgrid, r),
                                       static=(spos, sstart, send, sgrid, rs))
    X = world[:, :3].astype(np.float64)
    b = world[:, 3].astype(int)
    vis = _visits(orc, world)

    def pairs(Y, wY, radius, keysY, vY):
        d = np.linalg.norm(X[:, None, :] - Y[None, :, :], axis=2)
        i, j = np.nonzero(d < float(np.float32(r) + np.float32(radius) + np.float32(0.001)))
        if wY is not None:
            keep = b[i] != wY[j]
            i, j = i[keep], j[keep]
        rel = Y[j] - X[i]
        l = np.linalg.norm(rel, axis=1)
        n = rel / l[:, None]
        dv = vY[j] - vel[b[i], :3].astype(np.float64)
        ff = (ks * l - kd * np.einsum("ij,ij->i", n, dv))[:, None] * n
        mult = (vis[i] == keysY[j][:, None]).sum(1)
        return i, ff * mult[:, None], mult

    kd_ = orc.hash_keys(world, NB, (CELL,) * 3, abi.CommonFluidParams(), abi.KEY_REF_XOR)[0]
    ks_ = orc.hash_keys(st, NB, (CELL,) * 3, abi.CommonFluidParams(), abi.KEY_REF_XOR)[0]
    i1, f1, m1 = pairs(X, b, r, kd_, vel[b, :3].astype(np.float64))
    i2, f2, m2 = pairs(st[:, :3].astype(np.float64), None, rs, ks_, np.zeros((len(st), 3)))
    assert len(i1) > 100 and len(i2) > 20  # bodies touch each other and the plate
    want = np.zeros_like(X)
    np.add.at(want, i1, f1)
    np.add.at(want, i2, f2)
    scale = max(np.abs(f1).max(), np.abs(f2).max())
    assert np.allclose(f[:, :3], want, rtol=1e-5, atol=1e-5 * scale)
    assert np.array_equal(f[:, 3], b.astype(np.float32)) and np.array_equal(t[:, 3], b.astype(np.float32))
    d = (world[:, :3] - cm[b, :3]).astype(np.float32)
    F = f[:, :3]
    cross = np.stack([d[:, 1] * F[:, 2] - d[:, 2] * F[:, 1], d[:, 2] * F[:, 0] - d[:, 0] * F[:, 2],
                      d[:, 0] * F[:, 1] - d[:, 1] * F[:, 0]], 1)
    assert np.array_equal(t[:, :3], cross)
    # accumulate: written result added in float to the given arrays
    rng = np.random.default_rng(9)
    f0 = rng.normal(size=f.shape).astype(np.float32)
    t0 = rng.normal(size=t.shape).astype(np.float32)
    fa, ta = orc.body_forces_from_bodies(world, vel, cm, r, ks, kd, dyn=(dpos, dstart, dend, dgrid, r),
                                         static=(spos, sstart, send, sgrid, rs), f_init=f0, t_init=t0)
    assert np.array_equal(fa[:, :3], f0[:, :3] + f[:, :3])
    assert np.array_equal(ta[:, :3], t0[:, :3] + t[:, :3])
    assert np.array_equal(fa[:, 3], f[:, 3])

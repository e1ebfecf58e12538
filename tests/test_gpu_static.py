"""GPU parity of static rigid bodies (RigidBodySimulator.cpp:146-214 grid,
K_FluidSim.cu:369-374 contact) through the C-ABI against the oracle.

Tolerances:
  static grid built by the simulator   bit-exact (sorted centres, cell ranges)
  coupled fluid forces                 |dv| <= dt*1e-5*(|a| + S + |F_ext|), as
                                        for the dynamic grid
  trajectories                         10x the oracle's own divergence under a
                                        reversed bucket walk (an equally valid
                                        FP32 summation order), at least 1e-4 h
  eager / two-stream / graph replay    bit-identical
"""
import ctypes as C

import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _dev(lib, a):
    return lib.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _scene(with_torus=True, h=0.02):
    """A fluid block with a static cylinder standing in it (the particles
    inside the cylinder and its contact shell removed), optionally a torus
    resting on the surface."""
    s = scenes.dam_break((24, 14, 24), h, (1.0, 1.0, 1.0), 2e-4)
    top = s.pos[:, 1].max()
    keep = np.hypot(s.pos[:, 0] - 0.12, s.pos[:, 2] - 0.12) > 0.05 + 0.5 * h
    s.pos, s.vel = s.pos[keep].copy(), s.vel[keep].copy()
    s.fluid.max_particles = s.n
    cyl = scenes.cylinder_spheres(0.05, top + 0.05, 0.5 * h) + np.array([0.12, 0.0, 0.12])
    sets, masses = [], []
    if with_torus:
        sets = [scenes.torus_spheres(0.06, 0.02, 0.5 * h) + np.array([0.3, top + 0.01, 0.3])]
        masses = [1.0]
    b = scenes.make_bodies(sets, masses, h, s.common, static_sets=[cyl], st_buckets=(32, 32, 32))
    return s, b


def _oracle_static(orc, s, b):
    rb = b.rb
    nst = rb.st_bucket_size.x * rb.st_bucket_size.y * rb.st_bucket_size.z
    return orc.static_grid(b.static, rb.particleRadius, nst, s.common), nst


def test_simulator_static_grid_bit_exact(fglib, orc):
    s, b = _scene(with_torus=False)
    sim = fglib.Simulator(s, fused=True, use_graph=False)
    try:
        sim.add_static(b)
        got = sim.static_grid()
    finally:
        sim.close()
    assert got is not None
    gpos, gstart, gend, nb, g = got
    (opos, ostart, oend, ogrid), nst = _oracle_static(orc, s, b)
    assert nb == nst and g.numBodies == 0 and not g.bodyVelocities
    assert g.sphereRadius == np.float32(b.sphere_radius)
    assert np.array_equal(gpos, opos)
    assert np.array_equal(gstart, ostart)
    occ = ostart != 0xFFFFFFFF
    assert np.array_equal(gend[occ], oend[occ])


@pytest.mark.parametrize("with_dynamic", [False, True])
def test_static_coupled_forces_match_oracle(fglib, orc, with_dynamic):
    """fg_sph_density_nbr + fg_sph_forces_coupled with a static grid (and a
    dynamic torus grid) against orc_forces_coupled on the same density."""
    L = fglib.load()
    s, b = _scene(with_torus=True)
    rng = np.random.default_rng(12)
    s.vel[:, :3] = rng.normal(0, 0.2, (s.n, 3)).astype(np.float32)
    fglib.init_params(s.common, s.kernels, s.key_mode, abi.MATH_FAST)
    nb = s.num_buckets
    key, pid = orc.hash_keys(s.pos, nb, tuple(s.common.cell_size), s.common, s.key_mode)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, spos, svel = orc.cell_offsets(sk, sp, s.pos, s.vel, nb)
    (stpos, ststart, stend, stgrid), nst = _oracle_static(orc, s, b)
    cell = b.rb.particleRadius
    nbb = b.rb.dn_bucket_size.x * b.rb.dn_bucket_size.y * b.rb.dn_bucket_size.z
    bgrid = orc.grid_of_bodies(cell, nbb)
    world = b.world.copy()
    world[:, 1] -= 0.04  # push the torus into the surface
    bkey, bpid = orc.hash_keys(world, nbb, (cell,) * 3, s.common, abi.KEY_REF_XOR)
    bsk, bsp = orc.sort_pairs(bkey, bpid)
    bstart, bend, bpos, _ = orc.cell_offsets(bsk, bsp, world, None, nbb)
    bvel = np.array([[0.05, -0.3, 0.02, 0.0]], np.float32)
    ks, kd = -b.spring, -b.damping
    d = {k: _dev(fglib, v) for k, v in dict(spos=spos, svel=svel, start=start, end=end, sp=sp,
                                           rp=np.zeros((s.n, 2), np.float32), bpos=bpos, bvel=bvel, bstart=bstart,
                                           bend=bend, stpos=stpos, ststart=ststart, stend=stend).items()}
    masks = fglib.DeviceArray((L.fg_neighbor_mask_bytes(s.n),), np.uint8)
    fglib.check(L.fg_sph_density_nbr(d["spos"].ptr, d["svel"].ptr, d["rp"].ptr, d["start"].ptr, d["end"].ptr, masks.ptr, s.n,
                                     s.fluid.gasConstant, s.fluid.restDensity, s.fluid.mass), "density")
    rp = d["rp"].numpy()
    st = (stpos, ststart, stend, stgrid, b.sphere_radius)
    dyn = (bpos, bvel, bstart, bend, bgrid) if with_dynamic else (None, None, None, None, None)
    args = (start, end, sp, s.fluid, s.common, s.kernels, s.key_mode, b.fluid_radius, (1, 1, 1), b.sphere_radius,
            ks, kd, *dyn)
    v_ref, _ = orc.forces_coupled(spos, svel, rp, *args, v_out_init=s.vel, static=st)
    _, acc64 = orc.forces_coupled(spos, svel, rp.astype(np.float64), *args, f64=True, want_acc=True, static=st)
    _, acc_un, mag = orc.forces(spos, svel, rp.astype(np.float64), start, end, sp, s.fluid, s.common, s.kernels,
                                s.key_mode, f64=True, want_acc=True, want_mag=True)
    ext = np.linalg.norm(acc64 - acc_un, axis=1)
    assert np.count_nonzero(ext > 0) > 50  # particles in contact with the cylinder

    gs = abi.BodyGrid()
    gs.orderedSpheres, gs.bodyVelocities = d["stpos"].ptr, None
    gs.gridStart, gs.gridEnd, gs.bounds = d["ststart"].ptr, d["stend"].ptr, None
    gs.cellSize = abi.Float3(cell, cell, cell)
    gs.numBuckets, gs.relation = nst, abi.Int3(1, 1, 1)
    gs.numSpheres, gs.numBodies = len(stpos), 0
    gs.fluidRadius, gs.sphereRadius, gs.kSpring, gs.kDamping = b.fluid_radius, b.sphere_radius, ks, kd
    gd = abi.BodyGrid()
    gd.orderedSpheres, gd.bodyVelocities = d["bpos"].ptr, d["bvel"].ptr
    gd.gridStart, gd.gridEnd, gd.bounds = d["bstart"].ptr, d["bend"].ptr, None
    gd.cellSize = abi.Float3(cell, cell, cell)
    gd.numBuckets, gd.relation = nbb, abi.Int3(1, 1, 1)
    gd.numSpheres, gd.numBodies = b.m, 1
    gd.fluidRadius, gd.sphereRadius, gd.kSpring, gd.kDamping = b.fluid_radius, b.sphere_radius, ks, kd
    vel = _dev(fglib, s.vel)
    fglib.check(L.fg_sph_forces_coupled(d["spos"].ptr, d["svel"].ptr, d["rp"].ptr, d["start"].ptr, d["end"].ptr,
                                        masks.ptr, vel.ptr, None, d["sp"].ptr, s.n, s.fluid.mass, s.fluid.viscosity,
                                        C.byref(gd) if with_dynamic else None, C.byref(gs)), "coupled forces")
    v_gpu = vel.numpy()
    scale = np.zeros(s.n)
    scale[sp] = np.linalg.norm(acc64, axis=1) + mag + ext
    tol = s.common.deltaTimeIntegration * 1e-5 * scale[:, None] + 2 * np.spacing(np.abs(v_ref[:, :3]))
    err = np.abs(v_gpu[:, :3].astype(np.float64) - v_ref[:, :3])
    assert np.all(err <= tol), f"{(err > tol).sum()} out of tolerance"


def test_forces_coupled_rejects_missing_grids(fglib):
    L = fglib.load()
    n = 64
    a = fglib.DeviceArray((n, 4), np.float32)
    r = fglib.DeviceArray((n, 2), np.float32)
    u = fglib.DeviceArray((n,), np.uint32)
    rc = L.fg_sph_forces_coupled(a.ptr, a.ptr, r.ptr, u.ptr, u.ptr, None, a.ptr, None, u.ptr, n, 1.0, 1.0, None,
                                 None)
    assert rc == 1  # FG_ERR_INVALID_ARGUMENT
    L.fg_last_error()
    gs = abi.BodyGrid()
    gs.numSpheres = 10  # spheres but no arrays
    rc = L.fg_sph_forces_coupled(a.ptr, a.ptr, r.ptr, u.ptr, u.ptr, None, a.ptr, None, u.ptr, n, 1.0, 1.0, None,
                                 C.byref(gs))
    assert rc == 1  # FG_ERR_INVALID_ARGUMENT
    L.fg_last_error()


def _trajectory(fglib, s, b, steps, graph=True, overlap=True):
    sim = fglib.Simulator(s, fused=True, use_graph=graph)
    try:
        sim.add_static(b)
        if b.b:
            sim.add_bodies(b)
            sim.set_overlap(overlap)
        sim.step(steps)
        pos, vel = sim.state()
        bs = sim.body_state() if b.b else None
    finally:
        sim.close()
    return pos, vel, bs


@pytest.mark.parametrize("with_dynamic", [False, True])
def test_static_trajectory_matches_oracle(fglib, orc, with_dynamic):
    """30 coupled steps with a static cylinder standing in the fluid (and a
    torus resting on the surface) against the oracle's coupled step."""
    s, b = _scene(with_torus=with_dynamic)
    steps = 30
    pos, vel, bs = _trajectory(fglib, s, b, steps)
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
    div = np.max(np.abs(fwd.pos[:, :3] - rev.pos[:, :3])) / h
    err = np.max(np.abs(pos[:, :3] - fwd.pos[:, :3])) / h
    print(f"static{'+dynamic' if with_dynamic else ''} {steps}-step max|dx|/h: gpu {err:.3e}, "
          f"order divergence {div:.3e}")
    assert err <= max(10 * div, 1e-4)
    # the cylinder holds the fluid back: no particle centre inside it
    r = np.hypot(pos[:, 0] - 0.12, pos[:, 2] - 0.12)
    assert np.all(r > 0.05 - 0.5 * h)
    if with_dynamic:
        bdiv = np.max(np.abs(fwd.cm[:, :3] - rev.cm[:, :3])) / h
        berr = np.max(np.abs(bs["pos"][:, :3] - fwd.cm[:, :3])) / h
        assert berr <= max(10 * bdiv, 1e-5)


def test_static_steps_bit_identical_eager_overlap_graph(fglib):
    s, b = _scene(with_torus=True)
    ref = _trajectory(fglib, s, b, 8, graph=False, overlap=False)
    for graph, overlap in ((False, True), (True, True)):
        got = _trajectory(fglib, s, b, 8, graph=graph, overlap=overlap)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
        for k in ref[2]:
            assert np.array_equal(ref[2][k], got[2][k]), k


@pytest.mark.parametrize("with_torus", [False, True])
def test_contact_deferred_bit_identical_to_inline(fglib, with_torus):
    """The coupled force pass with body contact deferred to a second kernel
    (fg_set_contact_defer(1), the default: particles inside a body grid's
    bounds finish there, their body-grid cells walked by two lanes) gives the
    bits of the inline contact walk, with the static cylinder alone and with
    the dynamic torus too, at two workgroup sizes."""
    s, b = _scene(with_torus=with_torus)
    L = fglib.load()
    try:
        L.fg_set_contact_defer(0)
        ref = _trajectory(fglib, s, b, 8)
        L.fg_set_contact_defer(1)
        for blk in (256, 64):
            L.fg_set_tuning(blk, 8)
            got = _trajectory(fglib, s, b, 8)
            assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]), blk
            if with_torus:
                for k in ref[2]:
                    assert np.array_equal(ref[2][k], got[2][k]), (blk, k)
    finally:
        L.fg_set_contact_defer(1)
        L.fg_set_tuning(256, 8)

"""GPU parity of the reference-signature launchers that the simulator itself
does not call, through the C-ABI with the reference's argument layout,
against the oracle:

  gSPHCalculateForcesWithSRB (K_FluidSim.cuh:46-78; 28 arguments remapped to
      a static and a dynamic body grid sharing one srb cell size), static
      grid only, dynamic grid only, and both, on a density from
      gCalcMassDensity
  gCalculateRigidBodyForcesFromFluids (K_RigidBodySim.cuh:27-50)
  gFindCellOffsetOrderingPos (K_HashGrid.cuh:37-56)

Tolerances:
  coupled fluid forces        |dv| <= dt*1e-5*(|a| + S + |F_ext|) + 2 ulp, as
                              for fg_sph_forces_coupled
  body forces/torques         <= 1 ulp (same IEEE operations, contraction off)
  cell ranges + reorder       bit-exact
"""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _dev(lib, a):
    return lib.DeviceArray.from_numpy(np.ascontiguousarray(a))


def _ulp_diff(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def _scene(h=0.02):
    """A fluid block with a static cylinder standing in it and a torus pushed
    into the surface (the grids of tests/test_gpu_static.py)."""
    s = scenes.dam_break((24, 14, 24), h, (1.0, 1.0, 1.0), 2e-4)
    rng = np.random.default_rng(31)
    top = s.pos[:, 1].max()
    keep = np.hypot(s.pos[:, 0] - 0.12, s.pos[:, 2] - 0.12) > 0.05 + 0.5 * h
    s.pos, s.vel = s.pos[keep].copy(), s.vel[keep].copy()
    s.vel[:, :3] = rng.normal(0, 0.2, (s.n, 3)).astype(np.float32)
    s.fluid.max_particles = s.n
    cyl = scenes.cylinder_spheres(0.05, top + 0.05, 0.5 * h) + np.array([0.12, 0.0, 0.12])
    torus = scenes.torus_spheres(0.06, 0.02, 0.5 * h) + np.array([0.3, top - 0.03, 0.3])
    b = scenes.make_bodies([torus], [1.0], h, s.common, static_sets=[cyl], st_buckets=(32, 32, 32))
    return s, b


def _grids(orc, s, b):
    nb = s.num_buckets
    key, pid = orc.hash_keys(s.pos, nb, tuple(s.common.cell_size), s.common, s.key_mode)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, spos, svel = orc.cell_offsets(sk, sp, s.pos, s.vel, nb)
    rb = b.rb
    cell = rb.particleRadius
    nst = rb.st_bucket_size.x * rb.st_bucket_size.y * rb.st_bucket_size.z
    stpos, ststart, stend, stgrid = orc.static_grid(b.static, cell, nst, s.common)
    nbb = rb.dn_bucket_size.x * rb.dn_bucket_size.y * rb.dn_bucket_size.z
    bkey, bpid = orc.hash_keys(b.world, nbb, (cell,) * 3, s.common, abi.KEY_REF_XOR)
    bsk, bsp = orc.sort_pairs(bkey, bpid)
    bstart, bend, bpos, _ = orc.cell_offsets(bsk, bsp, b.world, None, nbb)
    return dict(sp=sp, start=start, end=end, spos=spos, svel=svel, cell=cell, nst=nst, stpos=stpos,
                ststart=ststart, stend=stend, stgrid=stgrid, nbb=nbb, bgrid=orc.grid_of_bodies(cell, nbb),
                bpos=bpos, bstart=bstart, bend=bend, bsk=bsk, bsp=bsp)


@pytest.mark.parametrize("use_static,use_dynamic", [(True, False), (False, True), (True, True)])
def test_reference_coupled_launcher_matches_oracle(fglib, orc, use_static, use_dynamic):
    L = fglib.load()
    s, b = _scene()
    fglib.init_params(s.common, s.kernels, s.key_mode, abi.MATH_FAST)
    g = _grids(orc, s, b)
    bvel = np.array([[0.05, -0.3, 0.02, 0.0]], np.float32)
    ks, kd = -b.spring, -b.damping
    d = {k: _dev(fglib, v) for k, v in dict(spos=g["spos"], svel=g["svel"], start=g["start"], end=g["end"],
                                           sp=g["sp"], key=np.zeros(s.n, np.uint32),
                                           rp=np.zeros((s.n, 2), np.float32), bpos=g["bpos"], bvel=bvel,
                                           bstart=g["bstart"], bend=g["bend"], stpos=g["stpos"],
                                           ststart=g["ststart"], stend=g["stend"]).items()}
    # K_FluidSim.cuh:7-16: the hash table argument is accepted and unused
    L.gCalcMassDensity(d["spos"].ptr, d["rp"].ptr, d["key"].ptr, d["start"].ptr, d["end"].ptr, s.n,
                       s.fluid.gasConstant, s.fluid.restDensity, s.fluid.mass)
    fglib.check(None, "gCalcMassDensity")
    rp = d["rp"].numpy()
    rho_ref = orc.density(g["spos"], g["start"], g["end"], s.fluid, s.common, s.kernels, s.key_mode)
    act = g["spos"][:, 3] > 0
    assert np.all(np.abs(rp[act, 0] - rho_ref[act, 0]) <= 1e-5 * rho_ref[act, 0])

    st = (g["stpos"], g["ststart"], g["stend"], g["stgrid"], b.sphere_radius) if use_static else None
    dyn = ((g["bpos"], bvel, g["bstart"], g["bend"], g["bgrid"]) if use_dynamic
           else (None, None, None, None, None))
    args = (g["start"], g["end"], g["sp"], s.fluid, s.common, s.kernels, s.key_mode, b.fluid_radius, (1, 1, 1),
            b.sphere_radius, ks, kd, *dyn)
    v_ref, _ = orc.forces_coupled(g["spos"], g["svel"], rp, *args, v_out_init=s.vel, static=st)
    _, acc64 = orc.forces_coupled(g["spos"], g["svel"], rp.astype(np.float64), *args, f64=True, want_acc=True,
                                  static=st)
    _, acc_un, mag = orc.forces(g["spos"], g["svel"], rp.astype(np.float64), g["start"], g["end"], g["sp"],
                                s.fluid, s.common, s.kernels, s.key_mode, f64=True, want_acc=True, want_mag=True)
    ext = np.linalg.norm(acc64 - acc_un, axis=1)
    assert np.count_nonzero(ext > 0) > 20  # particles in contact with the bodies

    vel = _dev(fglib, s.vel)
    cell = g["cell"]
    L.gSPHCalculateForcesWithSRB(
        d["spos"].ptr, d["svel"].ptr, None, vel.ptr, d["rp"].ptr, d["sp"].ptr, d["start"].ptr, d["end"].ptr, s.n,
        s.fluid.mass, s.fluid.viscosity, b.fluid_radius, abi.Int3(1, 1, 1), b.sphere_radius,
        abi.Float3(cell, cell, cell), ks, kd,
        d["stpos"].ptr if use_static else None, d["ststart"].ptr if use_static else None,
        d["stend"].ptr if use_static else None, g["nst"],
        d["bpos"].ptr if use_dynamic else None, d["bvel"].ptr if use_dynamic else None,
        d["bstart"].ptr if use_dynamic else None, d["bend"].ptr if use_dynamic else None, g["nbb"],
        use_static, use_dynamic)
    fglib.check(None, "gSPHCalculateForcesWithSRB")
    v_gpu = vel.numpy()
    scale = np.zeros(s.n)
    scale[g["sp"]] = np.linalg.norm(acc64, axis=1) + mag + ext
    tol = s.common.deltaTimeIntegration * 1e-5 * scale[:, None] + 2 * np.spacing(np.abs(v_ref[:, :3]))
    err = np.abs(v_gpu[:, :3].astype(np.float64) - v_ref[:, :3])
    print(f"static={use_static} dynamic={use_dynamic}: max err/tol {np.max(err / tol):.3f}")
    assert np.all(err <= tol), f"{(err > tol).sum()} out of tolerance"


def test_reference_body_forces_launcher_matches_oracle(fglib, orc):
    L = fglib.load()
    s, b = _scene()
    fglib.init_params(s.common, s.kernels, s.key_mode)
    g = _grids(orc, s, b)
    bvel = np.array([[0.05, -0.3, 0.02, 0.0]], np.float32)
    cm = np.array([[*b.world[:, :3].mean(0), 0.0]], np.float32)
    ks, kd = -b.spring, -b.damping
    d = {k: _dev(fglib, v) for k, v in dict(w=b.world, bv=bvel, cm=cm, spos=g["spos"], svel=g["svel"],
                                           start=g["start"], end=g["end"], bstart=g["bstart"],
                                           bend=g["bend"]).items()}
    f = fglib.DeviceArray((b.m, 4), np.float32)
    t = fglib.DeviceArray((b.m, 4), np.float32)
    cell = g["cell"]
    L.gCalculateRigidBodyForcesFromFluids(d["w"].ptr, d["bv"].ptr, d["cm"].ptr, f.ptr, t.ptr, d["bstart"].ptr,
                                          d["bend"].ptr, b.m, abi.Float3(cell, cell, cell), b.sphere_radius,
                                          g["nbb"], d["spos"].ptr, d["svel"].ptr, d["start"].ptr, d["end"].ptr,
                                          b.fluid_radius, abi.Float3(*s.common.cell_size), s.num_buckets,
                                          abi.Int3(1, 1, 1), ks, kd)
    fglib.check(None, "gCalculateRigidBodyForcesFromFluids")
    of, ot = orc.body_forces_from_fluid(b.world, bvel, cm, b.sphere_radius, g["spos"], g["svel"], g["start"],
                                        g["end"], b.fluid_radius, orc.grid_of_fluid(s.common, s.key_mode), (1, 1, 1),
                                        ks, kd)
    gf, gt = f.numpy(), t.numpy()
    assert np.count_nonzero(np.linalg.norm(of[:, :3], axis=1)) > 10  # the torus is in contact
    assert _ulp_diff(gf, of).max() <= 1
    assert _ulp_diff(gt, ot).max() <= 1


def test_reference_cell_offsets_positions_only_bit_exact(fglib, orc):
    """gFindCellOffsetOrderingPos on the torus spheres' sorted body keys:
    cellStart (0xFFFFFFFF for empty buckets), cellEnd of occupied buckets and
    the gathered positions, against the oracle."""
    L = fglib.load()
    s, b = _scene()
    fglib.init_params(s.common, s.kernels, s.key_mode)
    g = _grids(orc, s, b)
    d = {k: _dev(fglib, v) for k, v in dict(key=g["bsk"], pid=g["bsp"], pos=b.world).items()}
    start = fglib.DeviceArray((g["nbb"],), np.uint32)
    end = fglib.DeviceArray((g["nbb"],), np.uint32)
    start.fill_u32(0x12345678)  # the launcher clears it
    out = fglib.DeviceArray((b.m, 4), np.float32)
    L.gFindCellOffsetOrderingPos(start.ptr, end.ptr, d["key"].ptr, d["pid"].ptr, d["pos"].ptr, out.ptr, b.m,
                                 g["nbb"])
    fglib.check(None, "gFindCellOffsetOrderingPos")
    gs, ge = start.numpy(), end.numpy()
    assert np.array_equal(gs, g["bstart"])
    occ = gs != 0xFFFFFFFF
    assert np.array_equal(ge[occ], g["bend"][occ])
    assert np.array_equal(out.numpy(), g["bpos"])
    assert int(np.sum(ge[occ] - gs[occ])) == b.m

"""Pins of the oracle's SPH stages (density, forces, walls, integrate).

Each function of oracle/fg_oracle.cpp used by the SPH stages is checked
against something the mathematics fixes, chosen so that a dropped term, a
wrong sign or a wrong neighbour set fails:

  density      isolated particle m*kPoly6*h^6 = 315*rho0/(512*pi); interior
               of a perfect 0.5h lattice from the lattice sum over integer
               offsets |n| < 2 (= 1009.775166894685 kg/m^3); O(N^2) brute
               force over all pairs (FP64, numpy)
  forces       two-particle closed form; exact momentum balance of the pair
               terms; zero force at a perfect-lattice interior point; O(N^2)
               brute force (FP64, numpy)
  walls        single-particle penalty force in closed form, clamp to
               velocity_limit, the zero-coefficient friction term's -v_t
  integrate    free fall over many steps
  constants    kPoly6/kSpiky/kVisc against 315/(64 pi h^9), 45/(pi h^6)
"""
import math

import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

RHO0 = 1000.0


def make_scene(pos, h=0.025, box=(2.0, 2.0, 2.0), vmin=(-1.0, -1.0, -1.0), g=0.0, vlimit=1e30, vel=None,
               k=3.0, mu=3.5, dt=4e-4, mode=abi.KEY_LINEAR):
    s = scenes.dam_break((1, 1, 1), h, box, dt)
    cp = s.common
    cp.volume_min = abi.f3(*vmin)
    cp.volume_max = abi.f3(vmin[0] + box[0], vmin[1] + box[1], vmin[2] + box[2])
    cp.global_external_force = abi.f3(0.0, g, 0.0)
    cp.velocity_limit = vlimit
    s.fluid.gasConstant = k
    s.fluid.viscosity = mu
    pos = np.asarray(pos, np.float32)
    s.pos = pos
    s.vel = np.zeros_like(pos) if vel is None else np.asarray(vel, np.float32)
    s.fluid.max_particles = len(pos)
    s.key_mode = mode
    return s


def run_grid(orc, s):
    nb = s.num_buckets
    key, pid = orc.hash_keys(s.pos, nb, tuple(s.common.cell_size), s.common, s.key_mode)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, spos, svel = orc.cell_offsets(sk, sp, s.pos, s.vel, nb)
    return sk, sp, start, end, spos, svel


def brute_density(pos, m, kpoly6, h, h2):
    p = pos[:, :3].astype(np.float64)
    d = p[:, None, :] - p[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    w = np.where(r < h, (h2 - r * r) ** 3, 0.0)
    return m * kpoly6 * w.sum(1)


def brute_forces(pos, vel, rho, pr, m, mu, kc, h, g):
    p = pos[:, :3].astype(np.float64)
    v = vel[:, :3].astype(np.float64)
    d = p[:, None, :] - p[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    n = len(p)
    mask = (r <= h) & ~np.eye(n, dtype=bool)
    rr = np.where(mask, r, 1.0)
    rho_ij = rho[:, None] * rho[None, :]
    press = -(d / rr[..., None]) * (kc.kSpikyGradient * (h - rr) ** 2 * (pr[:, None] + pr[None, :]) / rho_ij)[..., None]
    visc = (v[None, :, :] - v[:, None, :]) / rho_ij[..., None] * (kc.kViscosityLaplacian * (h - rr) * mu)[..., None]
    f = np.where(mask[..., None], press + visc, 0.0).sum(1)
    return f * m + np.array([0.0, g, 0.0])


def test_kernel_constants_closed_form(orc):
    for h in (0.025, 0.01, 0.005, 0.003, 0.0179199):
        kc = orc.kernels_cache(h)
        hf = float(np.float32(h))
        assert kc.kPoly6 == pytest.approx(315.0 / (64.0 * math.pi * hf ** 9), rel=1e-6)
        assert kc.kSpikyGradient == pytest.approx(-45.0 / (math.pi * hf ** 6), rel=1e-6)
        assert kc.kViscosityLaplacian == pytest.approx(45.0 / (math.pi * hf ** 6), rel=1e-6)
        py = scenes.kernels_cache(h)
        assert (py.kPoly6, py.kSpikyGradient, py.kViscosityLaplacian) == (
            kc.kPoly6, kc.kSpikyGradient, kc.kViscosityLaplacian)


@pytest.mark.parametrize("h", [0.025, 0.01, 0.005, 0.003])
def test_isolated_particle_density(orc, h):
    s = make_scene([[0.1, 0.2, 0.3, 1.0]], h=h)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    for f64 in (False, True):
        rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=f64)
        closed = 315.0 * RHO0 / (512.0 * math.pi)  # m*kPoly6*h^6 with m = rho0 (h/2)^3
        assert rp[0, 0] == pytest.approx(closed, rel=2e-6)
        assert rp[0, 1] == pytest.approx(3.0 * (rp[0, 0] - RHO0), rel=1e-6)


def lattice(nx, h, offset):
    i = np.arange(nx)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    pos = np.zeros((len(g), 4), np.float32)
    pos[:, :3] = offset + 0.5 * h * g
    pos[:, 3] = 1
    return pos, g


def lattice_sum():
    # sum over integer offsets n with |n|^2 < 4 of (1 - |n|^2/4)^3
    tot = 0.0
    for a in range(-2, 3):
        for b in range(-2, 3):
            for c in range(-2, 3):
                q = a * a + b * b + c * c
                if q < 4:
                    tot += (1 - q / 4.0) ** 3
    return tot


def test_lattice_interior_density_and_zero_force(orc):
    h = 0.03125  # a power of two: every lattice coordinate is exact in float32
    pos, g = lattice(10, h, -0.25 + 0.25 * h)
    s = make_scene(pos, h=h)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    closed = 315.0 * RHO0 / (512.0 * math.pi) * lattice_sum()
    assert closed == pytest.approx(1009.775166894685, rel=1e-12)
    interior_sorted = np.all((g[sp] >= 2) & (g[sp] <= 7), axis=1)
    # particles whose neighbours are all interior (equal pressures around them)
    core_sorted = np.all((g[sp] >= 3) & (g[sp] <= 6), axis=1)
    rp64 = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    rp32 = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=False)
    assert np.allclose(rp64[interior_sorted, 0], closed, rtol=1e-5)
    assert np.allclose(rp32[interior_sorted, 0], closed, rtol=2e-5)
    # Symmetric neighbourhood, zero velocity: no pressure or viscous force.
    v, acc, mag = orc.forces(spos, svel, rp64, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode,
                             f64=True, want_acc=True, want_mag=True)
    assert np.all(np.abs(acc[core_sorted]) <= 1e-12 * mag[core_sorted, None])
    assert np.all(mag[core_sorted] > 1.0)


def test_density_brute_force(orc):
    s = scenes.small_dam_break(lattice=(12, 10, 12), h=0.05)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    ref = brute_density(spos, s.fluid.mass, s.kernels.kPoly6, s.common.kernel_radius,
                        s.common.kernel_radius_squared)
    rp64 = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    rp32 = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=False)
    assert np.allclose(rp64[:, 0], np.maximum(1.0, ref), rtol=1e-12)
    assert np.allclose(rp32[:, 0], np.maximum(1.0, ref), rtol=1e-5)


def test_two_particle_closed_form(orc):
    h = 0.025
    r = 0.6 * h
    v1, v2 = np.array([0.1, -0.2, 0.05]), np.array([-0.3, 0.4, 0.2])
    pos = [[0.0, 0.0, 0.0, 1.0], [r, 0.0, 0.0, 1.0]]
    vel = [[*v1, 0.0], [*v2, 0.0]]
    s = make_scene(pos, h=h, vel=vel, mu=3.5, k=3.0)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    m, kc = float(np.float32(s.fluid.mass)), s.kernels
    hf = float(np.float32(h))
    h2 = float(np.float32(s.common.kernel_radius_squared))
    rf = float(np.float32(r))
    rho = m * kc.kPoly6 * ((h2) ** 3 + (h2 - rf * rf) ** 3)
    p = 3.0 * (rho - RHO0)
    assert np.allclose(rp[:, 0], rho, rtol=1e-9)
    vout, acc, _ = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode, f64=True,
                              v_out_init=s.vel, want_acc=True)
    # particle 0 at the origin, particle 1 at +x: d = x0 - x1 = (-r, 0, 0)
    press0 = -(np.array([-rf, 0, 0]) / rf) * kc.kSpikyGradient * (hf - rf) ** 2 * (2 * p) / rho ** 2
    visc0 = (v2 - v1) / rho ** 2 * kc.kViscosityLaplacian * (hf - rf) * 3.5
    a0 = m * (press0 + visc0)
    i0 = int(np.where(sp == 0)[0][0])
    i1 = int(np.where(sp == 1)[0][0])
    assert np.allclose(acc[i0], a0, rtol=1e-6, atol=1e-9)
    assert np.allclose(acc[i1], -a0, rtol=1e-6, atol=1e-9)  # equal masses, equal and opposite
    assert np.allclose(vout[0, :3], v1 + a0 * s.common.deltaTimeIntegration, rtol=1e-6, atol=1e-12)


def test_momentum_balance_of_pair_terms(orc):
    s = scenes.small_dam_break(lattice=(10, 8, 10), h=0.05)
    rng = np.random.default_rng(5)
    s.vel[:, :3] = rng.normal(0, 0.5, (s.n, 3))
    s.common.global_external_force = abi.f3(0, 0, 0)
    s.common.velocity_limit = 1e30
    s.common.volume_min = abi.f3(-5, -5, -5)
    s.common.volume_max = abi.f3(5, 5, 5)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    _, acc, mag = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode, f64=True,
                             want_acc=True, want_mag=True)
    total = (acc * s.fluid.mass).sum(0)
    assert np.all(np.abs(total) <= 1e-11 * (mag.sum() * s.fluid.mass))


def test_forces_brute_force(orc):
    s = scenes.small_dam_break(lattice=(10, 9, 11), h=0.05)
    rng = np.random.default_rng(7)
    s.vel[:, :3] = rng.normal(0, 0.3, (s.n, 3))
    s.common.velocity_limit = 1e30
    s.common.volume_min = abi.f3(-5, -5, -5)
    s.common.volume_max = abi.f3(5, 5, 5)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    _, acc, mag = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode, f64=True,
                             want_acc=True, want_mag=True)
    ref = brute_forces(spos, svel, rp[:, 0], rp[:, 1], s.fluid.mass, s.fluid.viscosity, s.kernels,
                       s.common.kernel_radius, s.common.global_external_force.y)
    assert np.all(np.abs(acc - ref) <= 1e-10 * (mag[:, None] + 1.0))
    # The FP32 restatement stays within 1e-5 of the magnitude scale.
    rp32 = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=False)
    _, acc32, _ = orc.forces(spos, svel, rp32, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode,
                             want_acc=True)
    assert np.all(np.abs(acc32 - acc) <= 1e-5 * (np.abs(acc) + mag[:, None]))


def test_penalty_wall_closed_form(orc):
    # One particle 1 mm above the floor moving down: only the y- face fires.
    bd = float(np.float32(0.003043174))
    y = 0.001
    vy = -0.2
    s = make_scene([[0.5, y, 0.5, 1.0]], h=0.025, box=(1, 1, 1), vmin=(0, 0, 0), g=-9.81,
                   vel=[[0.0, vy, 0.0, 0.0]], vlimit=600.0)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    _, acc, _ = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode, f64=True,
                           want_acc=True)
    diff = bd - float(np.float32(y))
    expect_y = -9.81 + (20000.0 * diff - 256.0 * vy)
    assert acc[0, 1] == pytest.approx(expect_y, rel=1e-6)
    assert abs(acc[0, 0]) < 1e-12 and abs(acc[0, 2]) < 1e-12
    # Deeper penetration exceeds velocity_limit: the acceleration is clamped to 600.
    s2 = make_scene([[0.5, -0.05, 0.5, 1.0]], h=0.025, box=(1, 1, 1), vmin=(0, 0, 0), g=-9.81, vlimit=600.0)
    sk, sp, start, end, spos, svel = run_grid(orc, s2)
    rp = orc.density(spos, start, end, s2.fluid, s2.common, s2.kernels, s2.key_mode, f64=True)
    _, acc, _ = orc.forces(spos, svel, rp, start, end, sp, s2.fluid, s2.common, s2.kernels, s2.key_mode, f64=True,
                           want_acc=True)
    assert np.linalg.norm(acc[0]) == pytest.approx(600.0, rel=1e-9)
    assert acc[0, 1] > 0


def test_fixed_box_friction_term(orc):
    # Within 0.003 of the x = -11 face of the fixed friction box: a = g - v_t
    # with v_t = v - v*(n.v) (K_Boundaries_Common.cu:59-71), n = +x.
    v = np.array([0.5, 1.0, 1.0])
    s = make_scene([[-10.999, 0.0, 0.0, 1.0]], h=0.5, box=(24, 24, 24), vmin=(-12, -12, -12), g=-9.81,
                   vel=[[*v, 0.0]], vlimit=600.0)
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    _, acc, _ = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode, f64=True,
                           want_acc=True)
    vt = v - v * v[0]
    assert np.allclose(acc[0], np.array([0, -9.81, 0]) - vt, rtol=1e-6, atol=1e-9)


def test_free_fall_integration(orc):
    s = make_scene([[0.3, 0.9, 0.3, 1.0]], h=0.025, box=(1, 1, 1), vmin=(0, 0, 0), g=-9.81, vlimit=600.0, dt=1e-3)
    pos, vel = s.pos.copy(), s.vel.copy()
    steps = 100
    for _ in range(steps):
        nb = s.num_buckets
        key, pid = orc.hash_keys(pos, nb, tuple(s.common.cell_size), s.common, s.key_mode)
        sk, sp = orc.sort_pairs(key, pid)
        start, end, spos, svel = orc.cell_offsets(sk, sp, pos, vel, nb)
        rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode)
        vel, _, _ = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode,
                               v_out_init=vel)
        pos = orc.integrate(pos, vel, s.common)
    dt = float(np.float32(1e-3))
    g = float(np.float32(-9.81))
    assert vel[0, 1] == pytest.approx(steps * g * dt, rel=1e-5)
    assert pos[0, 1] == pytest.approx(0.9 + dt * dt * g * steps * (steps + 1) / 2, rel=1e-5)
    assert pos[0, 0] == pytest.approx(0.3, abs=1e-7) and pos[0, 3] == 1.0


def test_inactive_particles_untouched(orc):
    s = scenes.small_dam_break(lattice=(6, 4, 6))
    s.pos[::5, 3] = -1.0
    s.pos[::5, :3] = 0.9  # parked away from the block
    sk, sp, start, end, spos, svel = run_grid(orc, s)
    init = np.full((s.n, 2), 7.0, np.float32)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, rho_p_init=init)
    inactive_sorted = spos[:, 3] <= 0
    assert np.all(rp[inactive_sorted] == 7.0)
    vout, _, _ = orc.forces(spos, svel, rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode,
                            v_out_init=s.vel)
    assert np.all(vout[::5] == s.vel[::5])
    newpos = orc.integrate(s.pos, vout, s.common)
    assert np.array_equal(newpos[::5], s.pos[::5])

"""GPU checks of the mixed-precision density pass (FG_PRECISION_MIXED): FP16
cell offsets written by the cell pass, packed-FP16 conservative culling,
exact FP32 survivors, Neumaier-compensated sums.

  row masks      bit-identical to the FP32 pass's on the same sorted inputs
                 (every pair the FP32 test accepts survives the cull, and the
                 survivors' test is the FP32 one), on lattice, jittered,
                 compressed, out-of-grid, anisotropic-cell, large-cell and
                 scaled (simulation_scale 0.009) scenes and at 1M particles
  neighbour records  position and velocity words bit-identical
  density        |rho - rho_64| <= 1e-6 rho_64 against the FP64 oracle (the
                 FP32 pass meets 1e-5)
  8M + 4 boxes   after 1 and 200 steps: max relative density error <= 1e-3
                 against the FP64 oracle step from the same state, with the
                 pressure, velocity and position errors reported
"""
import json
import math
import os

import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _scene(which):
    if which == "cfg0":
        return scenes.config_scene(0)
    if which == "compressed":
        s = scenes.small_dam_break(lattice=(16, 16, 16), h=0.04)
        s.pos[:, :3] = 0.3 + (s.pos[:, :3] - 0.3) * 0.55
        return s
    if which == "jittered":
        s = scenes.small_dam_break(lattice=(20, 14, 18), h=0.05)
        rng = np.random.default_rng(11)
        s.pos[:, :3] = rng.uniform(0.02, 0.6, (s.n, 3)).astype(np.float32)
        return s
    if which == "out_of_grid":
        s = scenes.small_dam_break(lattice=(16, 10, 14), h=0.05)
        rng = np.random.default_rng(12)
        k = rng.choice(s.n, s.n // 5, replace=False)
        # some beyond each face of the grid, clustered so that they have neighbours
        s.pos[k, 0] = -0.05 - rng.uniform(0, 0.04, k.size)
        s.pos[k[::2], 1] = 1.03 + rng.uniform(0, 0.04, k[::2].size)
        return s
    if which in ("aniso", "large_cells"):
        s = scenes.small_dam_break(lattice=(18, 12, 16), h=0.05)
        h = s.common.kernel_radius
        cs = (h, 1.3 * h, 1.7 * h) if which == "aniso" else (1.6 * h, 1.6 * h, 1.6 * h)
        cs = tuple(float(np.float32(c)) for c in cs)
        s.common.cell_size = abi.f3(*cs)
        ext = (1.0, 1.0, 1.0)
        dims = tuple(int(math.ceil(e / c - 1e-9)) for e, c in zip(ext, cs))
        s.common.bucket_size = abi.i3(*dims)
        s.common.keyBits = scenes.key_bits(dims[0] * dims[1] * dims[2])
        return s
    if which == "appmain_linear":
        s = scenes.appmain_scene()
        s.key_mode = abi.KEY_LINEAR
        return s
    if which == "cfg1":
        return scenes.config_scene(1)
    raise ValueError(which)


def _grid(fglib, s):
    """Hash, sort and cell pass with the FP16 offsets, on the GPU."""
    L = fglib.load()
    fglib.init_params(s.common, s.kernels, s.key_mode)
    n, nb = s.n, s.num_buckets
    d = dict(pos=fglib.DeviceArray.from_numpy(s.pos), vel=fglib.DeviceArray.from_numpy(s.vel),
             key=fglib.DeviceArray((n,), np.uint32), pid=fglib.DeviceArray((n,), np.uint32),
             start=fglib.DeviceArray((nb,), np.uint32), end=fglib.DeviceArray((nb,), np.uint32),
             spos=fglib.DeviceArray((n, 4), np.float32), svel=fglib.DeviceArray((n, 4), np.float32),
             rp=fglib.DeviceArray.from_numpy(np.zeros((n, 2), np.float32)),
             nbr=fglib.DeviceArray((L.fg_neighbor_mask_bytes(n) // 8 + 32,), np.uint64))
    L.gCalcHashKey(d["key"].ptr, d["pid"].ptr, d["pos"].ptr, n, nb, abi.Float3(*s.common.cell_size))
    L.gSort(d["key"].ptr, d["pid"].ptr, n)
    fglib.check(L.fg_cell_offsets_pos_vel(d["start"].ptr, d["end"].ptr, d["key"].ptr, d["pid"].ptr, d["pos"].ptr,
                                          d["spos"].ptr, d["vel"].ptr, d["svel"].ptr, n, nb, 1, d["nbr"].ptr),
                "cell pass")
    return d


def _density(fglib, s, d, mode):
    L = fglib.load()
    n = s.n
    fglib.check(L.fg_set_precision(mode), "precision")
    try:
        fglib.check(L.fg_sph_density_nbr(d["spos"].ptr, d["svel"].ptr, d["rp"].ptr, d["start"].ptr, d["end"].ptr,
                                         d["nbr"].ptr, n, s.fluid.gasConstant, s.fluid.restDensity, s.fluid.mass),
                    "density")
    finally:
        L.fg_set_precision(abi.PRECISION_FP32)
    masks = fglib.download_raw(d["nbr"].ptr, (9, n), np.uint64)
    rec_off = (9 * 8 * n + 255) & ~255
    rec = fglib.download_raw(d["nbr"].ptr + rec_off, (n, 2, 4), np.float32)
    return d["rp"].numpy().copy(), masks, rec


@pytest.mark.parametrize("which", ["cfg0", "compressed", "jittered", "out_of_grid", "aniso", "large_cells",
                                   "appmain_linear", "cfg1"])
def test_mixed_masks_match_fp32(fglib, orc, which):
    s = _scene(which)
    d = _grid(fglib, s)
    rp32, m32, rec32 = _density(fglib, s, d, abi.PRECISION_FP32)
    rpm, mm, recm = _density(fglib, s, d, abi.PRECISION_MIXED)
    nz = int(np.count_nonzero(m32))
    print(f"{which}: n={s.n:,}, {nz:,} non-zero row masks, {int(sum(bin(int(x)).count('1') for x in m32.ravel()[:200000])):,}"
          " bits in the first 200k")
    diff = np.nonzero(m32 != mm)
    assert diff[0].size == 0, f"{diff[0].size} row masks differ, first at row {diff[0][0]} particle {diff[1][0]}"
    assert np.array_equal(rec32[:, 0, :3], recm[:, 0, :3]) and np.array_equal(rec32[:, 1, :3], recm[:, 1, :3])
    spos, start, end = d["spos"].numpy(), d["start"].numpy(), d["end"].numpy()
    act = spos[:, 3] > 0
    if s.n <= 300_000:
        rho64 = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)[:, 0]
        e32 = np.abs(rp32[act, 0] - rho64[act]) / rho64[act]
        em = np.abs(rpm[act, 0] - rho64[act]) / rho64[act]
        print(f"  density vs FP64: FP32 max {e32.max():.2e} mean {e32.mean():.2e}; "
              f"mixed max {em.max():.2e} mean {em.mean():.2e}")
        assert em.max() <= 1e-6
    else:
        rel = np.abs(rpm[act, 0] - rp32[act, 0]) / rp32[act, 0]
        print(f"  density mixed vs FP32: max {rel.max():.2e}")
        assert rel.max() <= 5e-6


def test_mixed_needs_fp16_offsets(fglib):
    """A mixed density pass on a buffer without FP16 offsets for its particle
    count is refused, not run on stale data."""
    s = _scene("cfg0")
    L = fglib.load()
    d = _grid(fglib, s)
    other = fglib.DeviceArray((L.fg_neighbor_mask_bytes(s.n) // 8 + 32,), np.uint64)
    L.fg_set_precision(abi.PRECISION_MIXED)
    try:
        rc = L.fg_sph_density_nbr(d["spos"].ptr, d["svel"].ptr, d["rp"].ptr, d["start"].ptr, d["end"].ptr,
                                  other.ptr, s.n, s.fluid.gasConstant, s.fluid.restDensity, s.fluid.mass)
        assert rc == abi.ERR_NOT_INITIALIZED
        L.fg_last_error()
        rc = L.fg_sph_density_nbr(d["spos"].ptr, d["svel"].ptr, d["rp"].ptr, d["start"].ptr, d["end"].ptr,
                                  d["nbr"].ptr, s.n - 1, s.fluid.gasConstant, s.fluid.restDensity, s.fluid.mass)
        assert rc == abi.ERR_NOT_INITIALIZED
        L.fg_last_error()
        assert L.fg_set_precision(7) == abi.ERR_INVALID_ARGUMENT
        L.fg_last_error()
        assert L.fg_get_precision() == abi.PRECISION_MIXED
    finally:
        L.fg_set_precision(abi.PRECISION_FP32)


def test_mixed_simulator_graph_and_eager_agree(fglib):
    """The fused step in the mixed form: graph replay and eager launches give
    the same bits; the state stays finite and close to the FP32 form's."""
    s = _scene("cfg0")
    out = []
    for graph in (True, False):
        sim = fglib.Simulator(s, fused=True, use_graph=graph, precision=abi.PRECISION_MIXED)
        try:
            sim.step(10)
            out.append(sim.state())
        finally:
            sim.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    sim = fglib.Simulator(s, fused=True, use_graph=True)
    try:
        sim.step(10)
        p32, _ = sim.state()
    finally:
        sim.close()
    assert np.all(np.isfinite(out[0][0]))
    dx = np.abs(out[0][0][:, :3] - p32[:, :3]).max() / s.common.kernel_radius
    print(f"10 steps: max |x_mixed - x_fp32| = {dx:.2e} h")
    assert dx < 1e-4


def _f64_step(orc, s, pos, vel):
    nb = s.num_buckets
    key, pid = orc.hash_keys(pos, nb, tuple(s.common.cell_size), s.common, s.key_mode)
    skey, spid = orc.sort_pairs(key, pid)
    start, end, spos, svel = orc.cell_offsets(skey, spid, pos, vel, nb)
    rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode, f64=True)
    v_out, acc, mag = orc.forces(spos, svel, rp, start, end, spid, s.fluid, s.common, s.kernels, s.key_mode,
                                 f64=True, v_out_init=vel.astype(np.float64), want_acc=True, want_mag=True)
    q = float(s.common.deltaTimeIntegration) / float(s.common.simulation_scale)
    x = pos[:, :3].astype(np.float64).copy()
    act = pos[:, 3] > 0
    x[act] += v_out[act, :3] * q
    return dict(rho_p=rp, spid=spid, vel=v_out[:, :3], pos=x, acc=acc, mag=mag)


def test_mixed_8m_boxes_against_fp64_oracle(fglib, orc):
    """Configuration 2 (8,388,608 particles, four rigid boxes) in the mixed
    form: per-quantity error of one GPU step against one FP64 oracle step
    from the same state, after 1 and after 200 steps.  The boxes start
    0.15 m and more above the fluid and fall about 1 mm in 200 steps, so no
    sphere is within contact distance of the fluid (checked), and the FP64
    fluid step is the uncoupled one."""
    s = scenes.coupled_config_scene(2)
    b = s.info["bodies"]
    sim = fglib.Simulator(s, fused=True, use_graph=True, precision=abi.PRECISION_MIXED)
    report = {"workload": f"dam break, {s.n:,} particles, four rigid boxes ({b.m:,} spheres)",
              "form": "mixed: FP16 cull, exact FP32 survivors, compensated density"}
    try:
        sim.add_bodies(b)
        for k in (1, 200):
            if k > 1:
                sim.step(k - 2)  # reach state S_k (k-1 steps in total)
            pos0, vel0 = sim.state()
            bs = sim.body_state()
            sim.step(1)
            g = sim.arrays()
            pos1, vel1 = sim.state()
            top = float(pos0[pos0[:, 3] > 0, 1].max())
            reach = 2.0 * s.common.pointRadius + b.sphere_radius
            lowest = float(bs["pos"][:, 1].min()) - 0.05 * math.sqrt(3.0)  # half-diagonal of a 0.1 m box
            assert lowest > top + reach, "a box is within contact distance"
            ref = _f64_step(orc, s, pos0, vel0)
            act = ref["rho_p"][:, 0] > 0
            rel_rho = np.abs(g["rho_p"][:, 0] - ref["rho_p"][:, 0]) / ref["rho_p"][:, 0]
            dp = np.abs(g["rho_p"][:, 1] - ref["rho_p"][:, 1]) / (s.fluid.gasConstant * s.fluid.restDensity)
            dt = float(s.common.deltaTimeIntegration)
            scale = np.zeros(s.n)
            scale[ref["spid"]] = np.linalg.norm(ref["acc"], axis=1) + ref["mag"]
            ev = np.linalg.norm(vel1[:, :3] - ref["vel"], axis=1) / (dt * np.maximum(scale, 1e-30))
            ex = np.linalg.norm(pos1[:, :3] - ref["pos"], axis=1) / s.common.kernel_radius

            def st(e):
                return {"max": float(np.max(e)), "mean": float(np.mean(e))}
            report["step_1" if k == 1 else "after_200_steps"] = {"rho": st(rel_rho[act]), "p": st(dp[act]),
                                                                 "v": st(ev), "x": st(ex)}
    finally:
        sim.close()
    print(json.dumps(report, indent=1))
    dd = os.environ.get("FG_REPORT_DIR")
    if dd:
        os.makedirs(dd, exist_ok=True)
        with open(os.path.join(dd, "precision_mixed_vs_fp64.json"), "w") as f:
            json.dump(report, f, indent=1)
    for key in ("step_1", "after_200_steps"):
        assert report[key]["rho"]["max"] <= 1e-3
        assert report[key]["v"]["max"] <= 5e-5

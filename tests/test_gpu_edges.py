"""Edge cases of the reference's semantics through the C-ABI against the
oracle (the failure cases a caller can force from the host).

  coincident particles   two active particles at one position: their pair
                         has r = 0, and the pressure direction d/r is 0/0
                         (K_FluidSim.cu:117), so their velocities become NaN
                         in both the reference arithmetic and the fast path,
                         as in the oracle; every other particle within the
                         usual tolerance
  one bucket (REF_XOR)   numBuckets = 1: every particle hashes to bucket 0
                         and the 27-bucket stencil visits it 27 times, so each
                         neighbour counts 27 times; density and forces as the
                         oracle's, in both math modes
  no particles           every launcher of the step accepts N = 0 and
                         records no error
"""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

from test_gpu_sph import check_density, check_velocity, gpu_density, gpu_forces, grid_cpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("math_mode", [abi.MATH_FAST, abi.MATH_REFERENCE])
def test_coincident_particles_give_nan_like_the_reference(fglib, orc, math_mode):
    s = scenes.small_dam_break(lattice=(10, 6, 10))
    s.pos[41, :3] = s.pos[40, :3]
    sk, sp, start, end, spos, svel = grid_cpu(orc, s)
    ref_rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode)
    rp = gpu_density(fglib, s, spos, start, end, math_mode)
    check_density(rp, ref_rp, s)  # r = 0 is an ordinary density term
    v_ref, acc, _ = orc.forces(spos, svel, ref_rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode,
                               v_out_init=s.vel, want_acc=True)
    v_gpu, _ = gpu_forces(fglib, s, spos, svel, ref_rp, start, end, sp, math_mode)
    nan_ref = np.isnan(v_ref[:, :3]).any(1)
    nan_gpu = np.isnan(v_gpu[:, :3]).any(1)
    assert set(np.nonzero(nan_ref)[0]) == {40, 41}
    assert np.array_equal(nan_gpu, nan_ref)
    # NaN rows compared above; the rest within the usual tolerance
    v_gpu0, v_ref0 = v_gpu.copy(), v_ref.astype(np.float64)
    v_gpu0[nan_gpu, :3] = 0.0
    v_ref0[nan_ref, :3] = 0.0
    _, _, mag = orc.forces(spos, svel, ref_rp.astype(np.float64), start, end, sp, s.fluid, s.common, s.kernels,
                           s.key_mode, f64=True, want_mag=True)
    check_velocity(v_gpu0, v_ref0, np.nan_to_num(acc), np.nan_to_num(mag), sp, s.common.deltaTimeIntegration)


def test_single_bucket_counts_every_neighbour_27_times(fglib, orc):
    s = scenes.small_dam_break(lattice=(8, 4, 8), key_mode=abi.KEY_REF_XOR)
    s.common.bucket_size = abi.i3(1, 1, 1)
    assert s.num_buckets == 1
    sk, sp, start, end, spos, svel = grid_cpu(orc, s)
    assert np.all(sk == 0) and start[0] == 0 and end[0] == s.n
    ref_rp = orc.density(spos, start, end, s.fluid, s.common, s.kernels, s.key_mode)
    for mm in (abi.MATH_FAST, abi.MATH_REFERENCE):
        rp = gpu_density(fglib, s, spos, start, end, mm)
        check_density(rp, ref_rp, s)
        v_ref, acc, _ = orc.forces(spos, svel, ref_rp, start, end, sp, s.fluid, s.common, s.kernels, s.key_mode,
                                   v_out_init=s.vel, want_acc=True)
        _, _, mag = orc.forces(spos, svel, ref_rp.astype(np.float64), start, end, sp, s.fluid, s.common,
                               s.kernels, s.key_mode, f64=True, want_mag=True)
        v_gpu, _ = gpu_forces(fglib, s, spos, svel, ref_rp, start, end, sp, mm)
        check_velocity(v_gpu, v_ref.astype(np.float64), acc, mag, sp, s.common.deltaTimeIntegration)


def test_no_particles(fglib):
    L = fglib.load()
    s = scenes.small_dam_break(lattice=(2, 2, 2))
    fglib.init_params(s.common, s.kernels, s.key_mode, abi.MATH_FAST)
    nb = s.num_buckets
    start = fglib.DeviceArray((nb,), np.uint32)
    end = fglib.DeviceArray((nb,), np.uint32)
    one = fglib.DeviceArray((1, 4), np.float32)
    u = fglib.DeviceArray((1,), np.uint32)
    rp = fglib.DeviceArray((1, 2), np.float32)
    L.gCalcHashKey(u.ptr, u.ptr, one.ptr, 0, nb, abi.Float3(*s.common.cell_size))
    fglib.check(None, "hash")
    fglib.check(L.fg_sort_pairs(u.ptr, u.ptr, 0, 8), "sort")
    L.gCalcMassDensity(one.ptr, rp.ptr, None, start.ptr, end.ptr, 0, 3.0, 1000.0, 1.0)
    fglib.check(None, "density")
    L.gSPHCalculateForces(one.ptr, one.ptr, None, rp.ptr, start.ptr, end.ptr, one.ptr, u.ptr, 0, 1.0, 1.0)
    fglib.check(None, "forces")
    L.gSPHIntegrate(one.ptr, one.ptr, None, 0, rp.ptr, start.ptr, end.ptr)
    fglib.check(None, "integrate")
    fglib.check(L.fg_sph_density_nbr(one.ptr, one.ptr, rp.ptr, start.ptr, end.ptr, None, 0, 3.0, 1000.0, 1.0), "density_nbr")
    fglib.check(L.fg_sph_forces_integrate_nbr(one.ptr, one.ptr, rp.ptr, start.ptr, end.ptr, None, one.ptr, one.ptr,
                                              u.ptr, 0, 1.0, 1.0), "forces_nbr")

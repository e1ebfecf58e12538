"""The C-ABI library: loads without a GPU, exports every function declared in
include/*.h, and its POD layouts agree between the C header, the C++ mirror
and the ctypes mirror (no compute calls are made here)."""
import ctypes as C
import glob
import os
import re
import shutil
import subprocess
import textwrap

import numpy as np
import pytest

from fluidgrid_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
DECL = re.compile(r"^\s*(?:const\s+)?(?:void|int|size_t|char|unsigned int)\s*\*?\s*(\w+)\s*\(", re.M)


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names += DECL.findall(text)
    return sorted(set(names))


def test_header_declares_reference_launchers():
    names = declared_functions()
    for ref in ("gInitializeCommonFluidParams", "gInitializeKernelParams", "gCalcHashKey", "gSort",
                "gFindCellOffsetOrderingPosVel", "gFindCellOffsetOrderingPos", "gCalcMassDensity",
                "gSPHCalculateForces", "gSPHIntegrate", "gCalcHashKeyWithOffset"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    L = lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert missing == []


def test_every_c_header_is_checked():
    names = [os.path.basename(h) for h in HEADERS]
    for h in ("fluidgrid_amd.h", "fluidgrid_amd_sim.h", "fluidgrid_amd_batch.h", "fluidgrid_amd_mesh.h"):
        assert h in names


def test_bindings_cover_every_declared_symbol():
    assert sorted(set(declared_functions()) - set(lib.SIGNATURES)) == []


def test_library_reports_no_device_quietly():
    # Without a GPU the runtime reports zero devices; nothing aborts.
    assert lib.gpu_count() >= 0
    assert lib.load().fg_version().startswith(b"fluidgrid_amd")


def test_struct_layouts_match_c_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(textwrap.dedent(f"""
        #include <stdio.h>
        #include <stddef.h>
        #include "{ROOT}/include/fluidgrid_amd.h"
        int main(void) {{
          printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(fg_common_fluid_params), sizeof(fg_fluid_params),
                 sizeof(fg_kernels_cache), sizeof(fg_common_rigid_body_params), sizeof(fg_scene_desc),
                 sizeof(fg_inflow_desc), sizeof(fg_body_grid));
          printf("%zu %zu %zu %zu\\n", offsetof(fg_common_fluid_params, volume_min),
                 offsetof(fg_common_fluid_params, deltaTimeIntegration),
                 offsetof(fg_common_rigid_body_params, global_external_force),
                 offsetof(fg_common_rigid_body_params, particleRadius));
          return 0;
        }}"""))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = list(map(int, out))
    assert got[:5] == [C.sizeof(abi.CommonFluidParams), C.sizeof(abi.FluidParams), C.sizeof(abi.KernelsCache),
                       C.sizeof(abi.CommonRigidBodyParams), C.sizeof(abi.SceneDesc)] == [104, 28, 12, 112, 160]
    assert got[5:7] == [C.sizeof(abi.InflowDesc), C.sizeof(abi.BodyGrid)] and got[5] == 72
    assert got[7:] == [abi.CommonFluidParams.volume_min.offset, abi.CommonFluidParams.deltaTimeIntegration.offset,
                       abi.CommonRigidBodyParams.global_external_force.offset,
                       abi.CommonRigidBodyParams.particleRadius.offset]


def test_errors_are_recorded_not_fatal():
    L = lib.load()
    assert L.fg_set_key_mode(7) == 1
    assert L.fg_last_error() == 1
    assert b"unknown mode" in L.fg_last_error_string()
    assert L.fg_last_error() == 0


def test_new_entry_points_reject_bad_arguments_before_any_launch():
    """Argument checks of the extension entry points run on the host and
    record FG_ERR_INVALID_ARGUMENT without touching a device (null pointers,
    an empty inflow window, a short workspace, missing grids)."""
    import ctypes as C
    L = lib.load()
    bad = 1  # FG_ERR_INVALID_ARGUMENT
    d = abi.InflowDesc()
    d.nu, d.nv, d.spacing, d.speed, d.numLayers = 4, 4, 0.01, 1.0, 2
    cases = [
        lambda: L.fg_emit_inflow(None, None, None, C.byref(d), 1e-3),
        lambda: L.fg_emit_inflow(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None, 1e-3),
        lambda: L.fg_emit_inflow(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.byref(abi.InflowDesc()), 1e-3),
        lambda: L.fg_conservation_f64(None, None, None, None, 10, 1.0, 1000.0, 3.0, None, None, 0),
        lambda: L.fg_conservation_f64(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 10, 1.0,
                                      1000.0, 3.0, C.c_void_p(16), C.c_void_p(16), 8),
        lambda: L.fg_body_forces_from_bodies(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                                             C.c_void_p(16), 10, 1, 0.01, None, None, None, 0),
        lambda: L.fg_body_bounds(None, None, 3, 0.1, None),
        lambda: L.fg_sph_forces_coupled(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                                        C.c_void_p(16), None, C.c_void_p(16), None, C.c_void_p(16), 10, 1.0, 1.0,
                                        None, None),
    ]
    for k, call in enumerate(cases):
        assert call() == bad, k
        assert L.fg_last_error() == bad, k
        assert L.fg_last_error() == 0


def test_inflow_dt_beyond_period_refused_before_launch():
    """A step longer than the layer period spacing * scale / speed would owe
    several layers; fg_emit_inflow refuses it on the host."""
    import ctypes as C
    from fluidgrid_amd import scenes
    L = lib.load()
    s = scenes.with_inflow(scenes.small_dam_break(lattice=(4, 4, 4), h=0.04), (0.3, 0.2), (0.42, 0.3), speed=1.0)
    d = s.info["inflow"]
    cp = abi.CommonFluidParams.from_buffer_copy(s.common)
    cp.simulation_scale = 0.009
    L.gInitializeCommonFluidParams(cp)
    try:
        period = L.fg_inflow_period(C.byref(d), 0.009)
        assert abs(period - d.spacing * np.float32(0.009) / d.speed) <= 1e-12
        p = C.c_void_p(256)
        assert L.fg_emit_inflow(p, p, p, C.byref(d), float(np.float32(2.0 * period))) == 1
        assert b"exceeds the layer period" in L.fg_last_error_string()
        assert L.fg_last_error() == 1
    finally:
        L.gInitializeCommonFluidParams(s.common)


def test_checkpoint_and_mesh_entry_points_check_arguments():
    """Without a simulator, save/load report FG_ERR_NOT_INITIALIZED; the mesh
    entry points reject null pointers and empty meshes; nothing touches a
    device."""
    L = lib.load()
    L.fg_sim_shutdown()
    not_init = 4  # FG_ERR_NOT_INITIALIZED
    assert L.fg_sim_save_state(b"/nonexistent/x.fgs") == not_init
    assert L.fg_sim_load_state(b"/nonexistent/x.fgs") == not_init
    bad = 1
    for call in (lambda: L.fg_mesh_winding_lattice(None, 0, None, 3, abi.Float3(0, 0, 0), 0.1, 2, 2, 2, None),
                 lambda: L.fg_mesh_sample_spheres(None, 0, None, 0, 0.1, 0, None, 0, None),
                 lambda: L.fg_mesh_lattice(None, 0, 0.1, None, None),
                 lambda: L.fg_body_file_read(None, None, None, None, None, None, 0, None, 0)):
        assert call() == bad
        assert L.fg_last_error() == bad


def test_cells_smaller_than_h_are_refused():
    """The density and force passes walk 27 cells: with a cell edge below
    the kernel radius they record FG_ERR_UNSUPPORTED before any launch."""
    import ctypes as C
    from fluidgrid_amd import scenes
    L = lib.load()
    s = scenes.small_dam_break(lattice=(4, 4, 4), h=0.04)
    cp = abi.CommonFluidParams.from_buffer_copy(s.common)
    cp.cell_size = abi.f3(0.02, 0.04, 0.04)  # h/2 along x
    L.gInitializeCommonFluidParams(cp)
    L.gInitializeKernelParams(s.kernels)
    p = C.c_void_p(256)
    try:
        rc = L.fg_sph_density_nbr(p, p, p, p, p, p, 64, 3.0, 1000.0, 1.0)
        assert rc == 5  # FG_ERR_UNSUPPORTED
        assert b"smaller than the kernel radius" in L.fg_last_error_string()
        assert L.fg_last_error() == 5
        rc = L.fg_sph_forces_integrate_nbr(p, p, p, p, p, p, p, p, p, 64, 1.0, 1.0)
        assert rc == 5
        assert L.fg_last_error() == 5
    finally:
        L.gInitializeCommonFluidParams(s.common)


def test_batched_scene_with_cells_smaller_than_h_is_refused():
    """fg_bsim_create checks every scene's cells against its kernel radius
    on the host, before any allocation, and reports FG_ERR_UNSUPPORTED."""
    import numpy as np
    from fluidgrid_amd import scenes
    L = lib.load()
    a = scenes.small_dam_break(lattice=(4, 4, 4), h=0.04)
    b = scenes.small_dam_break(lattice=(4, 4, 4), h=0.04)
    commons = (abi.CommonFluidParams * 2)(a.common, b.common)
    commons[1].cell_size = abi.f3(0.04, 0.03, 0.04)  # below h along y
    fluids = (abi.FluidParams * 2)(a.fluid, b.fluid)
    counts = (C.c_uint * 2)(a.n, b.n)
    pos = np.ascontiguousarray(np.concatenate([a.pos, b.pos]))
    vel = np.ascontiguousarray(np.concatenate([a.vel, b.vel]))
    rc = L.fg_bsim_create(2, commons, fluids, counts, pos.ctypes.data, vel.ctypes.data)
    assert rc == abi.ERR_UNSUPPORTED
    assert b"scene 1 has cells smaller than the kernel radius" in L.fg_last_error_string()
    assert L.fg_last_error() == abi.ERR_UNSUPPORTED


def test_reference_prototypes_pass_arguments_intact(tmp_path):
    """A caller compiled against the reference's own prototypes (HIP
    float2/float3/float4/int3 and bool by value, its CommonFluidParams
    declaration, none of this library's headers) links unchanged: the
    parameter block round-trips bit for bit, and the validation errors echo
    every by-value scalar of gCalcHashKey (K_HashGrid.cuh:5-8) and of the
    28-argument gSPHCalculateForcesWithSRB (K_FluidSim.cuh:46-78), whose
    trailing arguments travel on the stack."""
    src = os.path.join(ROOT, "tests", "abi_probe", "reference_caller.cpp")
    exe = tmp_path / "reference_caller"
    libdir = os.path.join(ROOT, "fluidgrid_amd")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    subprocess.run([hipcc, "-std=c++17", "-O1", src, "-o", str(exe), "-L", libdir, "-lfluidgrid_amd",
                    f"-Wl,-rpath,{libdir}"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    lines = dict(l.split(" ", 1) for l in out.strip().splitlines())
    assert lines["sizeof"] == "104"
    assert lines["common_roundtrip"] == "rc=0 equal=1"
    assert lines["hash"].startswith("err=1 ")
    assert "cell_size (-0.5, 0.25, 1.5) must be positive (numParticles=4321 numBuckets=98765)" in lines["hash"]
    assert lines["srb"].startswith("err=1 ")
    assert ("(numParticles=12345 mass=0.25 kViscosity=3.5 sph_particle_radius=0.125 cell_size_relation=(1,2,3) "
            "srb_particle_radius=0.0625 srb_cell_size=(0.5,0.75,1.25) kspring=-2000 kdamping=-50 "
            "srb_num_buckets=4096 rb_num_buckets=777 useStaticRB=1 useDynamicRB=0)") in lines["srb"]


def test_contact_defer_switch_round_trips_on_the_host():
    """fg_set_contact_defer / fg_get_contact_defer: process-wide state, on by
    default, read back as set; a null output pointer is refused."""
    import ctypes as C
    L = lib.load()
    v = C.c_int(-1)
    assert L.fg_get_contact_defer(C.byref(v)) == 0 and v.value == 1
    try:
        assert L.fg_set_contact_defer(0) == 0
        assert L.fg_get_contact_defer(C.byref(v)) == 0 and v.value == 0
        assert L.fg_set_contact_defer(7) == 0
        assert L.fg_get_contact_defer(C.byref(v)) == 0 and v.value == 1
        assert L.fg_get_contact_defer(None) == 1  # FG_ERR_INVALID_ARGUMENT
        assert L.fg_last_error() == 1 and L.fg_last_error() == 0
    finally:
        L.fg_set_contact_defer(1)

"""ctypes bindings of libfluidgrid_amd.so (the C-ABI in include/*.h).

Loading fails with an ImportError-like RuntimeError when the library has not
been built: there is no fallback compute path.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfluidgrid_amd.so")

_lib = None

vp = C.c_void_p
u32 = C.c_uint
i32 = C.c_int
f32 = C.c_float
sz = C.c_size_t

# name -> (restype, argtypes) for every exported C function.
SIGNATURES = {
    # reference launchers
    "gInitializeCommonFluidParams": (None, [abi.CommonFluidParams]),
    "gInitializeKernelParams": (None, [abi.KernelsCache]),
    "gCalcHashKey": (None, [vp, vp, vp, u32, u32, abi.Float3]),
    "gCalcHashKeyWithOffset": (None, [vp, vp, vp, vp, vp, vp, u32, abi.Float3, i32]),
    "gSort": (None, [vp, vp, i32]),
    "gFindCellOffsetOrderingPosVel": (None, [vp, vp, vp, vp, vp, vp, vp, vp, u32, u32]),
    "gFindCellOffsetOrderingPos": (None, [vp, vp, vp, vp, vp, vp, u32, u32]),
    "gCalcMassDensity": (None, [vp, vp, vp, vp, vp, i32, f32, f32, f32]),
    "gSPHCalculateForces": (None, [vp, vp, vp, vp, vp, vp, vp, vp, u32, f32, f32]),
    "gSPHIntegrate": (None, [vp, vp, vp, u32, vp, vp, vp]),
    # extensions
    "fg_last_error": (i32, []),
    "fg_peek_error": (i32, []),
    "fg_last_error_string": (C.c_char_p, []),
    "fg_record_error": (None, [i32, C.c_char_p]),
    "fg_version": (C.c_char_p, []),
    "fg_set_stream": (i32, [vp]),
    "fg_get_stream": (vp, []),
    "fg_set_key_mode": (i32, [i32]),
    "fg_get_key_mode": (i32, []),
    "fg_set_math_mode": (i32, [i32]),
    "fg_get_math_mode": (i32, []),
    "fg_get_common_fluid_params": (i32, [C.POINTER(abi.CommonFluidParams)]),
    "fg_get_kernels_cache": (i32, [C.POINTER(abi.KernelsCache)]),
    "fg_sort_pairs": (i32, [vp, vp, u32, i32]),
    "fg_sort_pairs_iota": (i32, [vp, vp, u32, i32]),
    "fg_sort_workspace_size": (sz, [u32, i32]),
    "fg_sort_pairs_ws": (i32, [vp, vp, u32, i32, vp, sz, i32]),
    "fg_sort_pairs_noalloc": (i32, [vp, vp, u32, i32, i32]),
    "fg_calc_hash_key_only": (i32, [vp, vp, u32, u32, abi.Float3]),
    "fg_clear_occupied_cells": (i32, [vp, vp, u32, u32]),
    "fg_calc_hash_key_clear": (i32, [vp, vp, u32, u32, abi.Float3, vp]),
    "fg_cell_offsets_pos_vel": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, u32, u32, i32, vp]),
    "fg_sph_forces_integrate": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, u32, f32, f32]),
    "fg_neighbor_mask_bytes": (sz, [u32]),
    "fg_sph_density_nbr": (i32, [vp, vp, vp, vp, vp, vp, u32, f32, f32, f32]),
    "fg_sph_forces_integrate_nbr": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, f32, f32]),
    "fg_device_count": (i32, [C.POINTER(C.c_int)]),
    "fg_device_name": (i32, [i32, C.c_char_p, sz]),
    "fg_malloc": (i32, [C.POINTER(C.c_void_p), sz]),
    "fg_free": (i32, [vp]),
    "fg_memcpy_htod": (i32, [vp, vp, sz]),
    "fg_memcpy_dtoh": (i32, [vp, vp, sz]),
    "fg_memcpy_dtod": (i32, [vp, vp, sz]),
    "fg_memset_u32": (i32, [vp, u32, sz]),
    "fg_copy_bytes": (i32, [vp, vp, sz]),
    "fg_synchronize": (i32, []),
    "fg_event_create": (i32, [C.POINTER(C.c_void_p)]),
    "fg_event_destroy": (i32, [vp]),
    "fg_event_record": (i32, [vp]),
    "fg_event_record_stream": (i32, [vp, vp]),
    "fg_event_elapsed_ms": (i32, [vp, vp, C.POINTER(C.c_float)]),
    # simulator
    "fg_sim_init": (i32, [C.POINTER(abi.CommonFluidParams), i32, i32]),
    "fg_sim_add_fluid": (i32, [C.POINTER(abi.FluidParams), vp, vp]),
    "fg_sim_set_options": (i32, [i32, i32, i32]),
    "fg_sim_set_overlap": (i32, [i32]),
    "fg_sim_set_precision": (i32, [i32]),
    "fg_sim_autotune": (i32, [i32, vp, vp, i32, C.POINTER(C.c_int)]),
    "fg_sim_step": (i32, [i32]),
    "fg_sim_synchronize": (i32, []),
    "fg_sim_stream": (vp, []),
    "fg_sim_download": (i32, [i32, vp, vp]),
    "fg_sim_upload": (i32, [i32, vp, vp]),
    "fg_sim_fluid_buffers": (i32, [i32, vp]),
    "fg_sim_stage_times": (i32, [vp]),
    "fg_sim_conservation": (i32, [i32, vp]),
    "fg_sim_set_inflow": (i32, [i32, C.POINTER(abi.InflowDesc)]),
    "fg_emit_inflow": (i32, [vp, vp, vp, C.POINTER(abi.InflowDesc), f32]),
    "fg_inflow_period": (C.c_double, [C.POINTER(abi.InflowDesc), f32]),
    "fg_conservation_workspace_size": (sz, []),
    "fg_conservation_f64": (i32, [vp, vp, vp, vp, u32, f32, f32, f32, vp, vp, sz]),
    "fg_sim_shutdown": (i32, []),
    "fg_sim_save_state": (i32, [C.c_char_p]),
    "fg_sim_load_state": (i32, [C.c_char_p]),
    "fg_sim_write_ply": (i32, [i32, C.c_char_p, C.POINTER(C.c_uint)]),
    "fg_sim_add_bodies": (i32, [vp, vp, u32, vp, f32, C.POINTER(abi.CommonRigidBodyParams), f32, f32]),
    "fg_sim_add_static_spheres": (i32, [vp, u32, f32, C.POINTER(abi.CommonRigidBodyParams), f32, f32]),
    "fg_sim_static_grid": (i32, [C.POINTER(abi.BodyGrid)]),
    "fg_sim_body_state": (i32, [vp, vp, vp, vp]),
    "fg_sim_body_sums": (i32, [vp, vp]),
    "fg_sim_body_mass_properties": (i32, [vp, vp]),
    # coupling
    "gSPHCalculateForcesWithSRB": (None, [vp, vp, vp, vp, vp, vp, vp, vp, u32, f32, f32, f32, abi.Int3, f32,
                                          abi.Float3, f32, f32, vp, vp, vp, u32, vp, vp, vp, vp, u32, C.c_bool,
                                          C.c_bool]),
    "gCalculateRigidBodyForcesFromFluids": (None, [vp, vp, vp, vp, vp, vp, vp, u32, abi.Float3, f32, i32, vp, vp,
                                                   vp, vp, f32, abi.Float3, u32, abi.Int3, f32, f32]),
    "gCopyRBDataToParticle": (None, [vp, vp, vp, sz]),
    "gReorderForceTorque": (None, [vp, vp, vp, vp, vp, u32]),
    "gAddForceTorque": (None, [vp, vp, vp, vp, vp, vp, sz]),
    "gIntegrateRB": (None, [vp, vp, vp, vp, sz, f32, abi.Float4, abi.Float3, abi.Float3, vp, vp, vp, vp, vp, f32]),
    "fg_sph_forces_coupled": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, f32, f32, vp, vp]),
    "fg_body_forces_from_fluid": (i32, [vp, vp, vp, vp, vp, u32, u32, f32, vp, vp, vp, vp, u32, f32, abi.Float3,
                                        u32, abi.Int3, f32, f32]),
    "gCalculateRigidBodyForcesFromRB": (None, [vp, vp, vp, vp, vp, vp, vp, u32, abi.Float3, f32, i32, i32, vp, vp,
                                               vp, f32, f32, i32]),
    "fg_body_forces_from_bodies": (i32, [vp, vp, vp, vp, vp, u32, u32, f32, vp, vp, vp, i32]),
    "fg_body_bounds": (i32, [vp, vp, u32, f32, vp]),
    "fg_body_reduce_f64": (i32, [vp, vp, vp, u32, vp, vp, vp, vp]),
    "fg_body_mass_properties_f64": (i32, [vp, vp, u32, vp, vp, vp]),
    "fg_sphere_bounds": (i32, [vp, u32, f32, vp]),
    # meshes and body files (fluidgrid_amd_mesh.h)
    "fg_mesh_winding_lattice": (i32, [vp, u32, vp, u32, abi.Float3, f32, u32, u32, u32, vp]),
    "fg_mesh_load_obj": (i32, [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint), vp, vp]),
    "fg_mesh_lattice": (i32, [vp, u32, f32, C.POINTER(abi.Float3), vp]),
    "fg_mesh_winding_numbers": (i32, [vp, u32, vp, u32, f32, vp]),
    "fg_mesh_sample_spheres": (i32, [vp, u32, vp, u32, f32, i32, vp, u32, C.POINTER(C.c_uint)]),
    "fg_body_file_write": (i32, [C.c_char_p, i32, f32, f32, C.c_char_p, vp, u32]),
    "fg_body_file_read": (i32, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_uint), vp, u32, C.c_char_p, sz]),
    # tuning and batched scenes
    "fg_set_tuning": (i32, [i32, i32]),
    "fg_get_tuning": (i32, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fg_set_schedule": (i32, [i32, i32]),
    "fg_get_schedule": (i32, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fg_set_tiled": (i32, [i32]),
    "fg_set_contact_defer": (i32, [i32]),
    "fg_get_contact_defer": (i32, [C.POINTER(C.c_int)]),
    "fg_set_precision": (i32, [i32]),
    "fg_get_precision": (i32, []),
    "fg_get_tiled": (i32, [C.POINTER(C.c_int)]),
    "fg_batch_hash": (i32, [vp, u32, vp, vp, vp, u32]),
    "fg_batch_density_nbr": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, vp, u32]),
    "fg_batch_forces_integrate_nbr": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32]),
    "fg_bsim_create": (i32, [u32, vp, vp, vp, vp, vp]),
    "fg_bsim_stream": (vp, []),
    "fg_bsim_add_bodies": (i32, [u32, vp, vp, u32, vp, f32, C.POINTER(abi.CommonRigidBodyParams), f32, f32]),
    "fg_bsim_body_state": (i32, [C.POINTER(C.c_uint), vp, vp, vp, vp]),
    "fg_bsim_set_options": (i32, [i32, i32]),
    "fg_bsim_set_overlap": (i32, [i32]),
    "fg_bsim_step": (i32, [i32]),
    "fg_bsim_synchronize": (i32, []),
    "fg_bsim_download": (i32, [vp, vp]),
    "fg_bsim_download_density": (i32, [vp, vp]),
    "fg_bsim_stage_times": (i32, [vp]),
    "fg_bsim_info": (i32, [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_int)]),
    "fg_bsim_autotune": (i32, [i32, vp, vp, i32, C.POINTER(C.c_int)]),
    "fg_bsim_set_tuning": (i32, [i32, i32, i32, i32, i32, i32]),
    "fg_bsim_destroy": (i32, []),
}


class FluidBuffers(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("hashKey", "particleId", "gridStart", "gridEnd", "position",
                                          "orderedPosition", "velocity", "orderedVelocity", "densityPressure",
                                          "forces")] + [("numParticles", C.c_uint), ("numBuckets", C.c_uint)]


class FGError(RuntimeError):
    pass


def load():
    """Load libfluidgrid_amd.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FGError(f"{LIB_PATH} is missing: build it with `python -m fluidgrid_amd.build` "
                          "(there is no CPU fallback for the HIP path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc=None, what="fluidgrid_amd"):
    """Raise FGError if rc is non-zero or the thread's error slot is set."""
    L = load()
    err = L.fg_last_error()
    if (rc not in (None, 0)) or err:
        msg = L.fg_last_error_string().decode(errors="replace")
        raise FGError(f"{what}: rc={rc} err={err} {msg}")


def gpu_count():
    n = C.c_int(0)
    load().fg_device_count(C.byref(n))
    return n.value


def device_name(dev=0):
    buf = C.create_string_buffer(256)
    rc = load().fg_device_name(dev, buf, 256)
    check(rc, "fg_device_name")
    return buf.value.decode()


class DeviceArray:
    """A device allocation mirrored by a numpy dtype/shape."""

    def __init__(self, shape, dtype, data=None):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(load().fg_malloc(C.byref(p), max(self.nbytes, 16)), "fg_malloc")
        self.ptr = p.value
        if data is not None:
            self.upload(data)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        return cls(a.shape, a.dtype, a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(load().fg_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes), "htod")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(load().fg_synchronize(), "sync")
        check(load().fg_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "dtoh")
        return out

    def fill_u32(self, value):
        check(load().fg_memset_u32(self.ptr, value, self.nbytes // 4), "fill")

    def free(self):
        if self.ptr:
            load().fg_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class StreamTimer:
    """HIP events recorded on a given stream around a region of work; the
    elapsed milliseconds are read after the stop event completes."""

    def __init__(self, stream):
        L = load()
        self.stream = stream
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        check(L.fg_event_create(C.byref(self.e0)), "event")
        check(L.fg_event_create(C.byref(self.e1)), "event")

    def start(self):
        check(load().fg_event_record_stream(self.e0, self.stream), "event record")

    def stop_ms(self):
        L = load()
        check(L.fg_event_record_stream(self.e1, self.stream), "event record")
        ms = C.c_float()
        check(L.fg_event_elapsed_ms(self.e0, self.e1, C.byref(ms)), "elapsed")
        return float(ms.value)

    def close(self):
        L = load()
        for e in (self.e0, self.e1):
            if e:
                L.fg_event_destroy(e)
        self.e0 = self.e1 = None


def init_params(common, kernels, key_mode, math_mode=abi.MATH_FAST):
    L = load()
    L.gInitializeCommonFluidParams(common)
    L.gInitializeKernelParams(kernels)
    check(L.fg_set_key_mode(key_mode), "key mode")
    check(L.fg_set_math_mode(math_mode), "math mode")


def download_raw(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    check(load().fg_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes), "dtoh")
    return out


class Simulator:
    """The C++ HeterogeneousSimulator (one process-wide instance) driven
    through include/fluidgrid_amd_sim.h."""

    def __init__(self, scene, math_mode=abi.MATH_FAST, fused=True, use_graph=True, stage_timing=False,
                 precision=abi.PRECISION_FP32):
        L = load()
        self.scene = scene
        self.common = scene.common
        check(L.fg_sim_init(C.byref(scene.common), scene.key_mode, math_mode), "fg_sim_init")
        pos = np.ascontiguousarray(scene.pos, np.float32)
        vel = np.ascontiguousarray(scene.vel, np.float32)
        idx = L.fg_sim_add_fluid(C.byref(scene.fluid), pos.ctypes.data, vel.ctypes.data)
        if idx < 0:
            check(-idx, "fg_sim_add_fluid")
        self.idx = idx
        self.n = scene.n
        check(L.fg_sim_set_options(int(fused), int(use_graph), int(stage_timing)), "options")
        check(L.fg_sim_set_precision(int(precision)), "precision")
        self.buffers = FluidBuffers()
        check(L.fg_sim_fluid_buffers(idx, C.byref(self.buffers)), "buffers")

    def set_options(self, fused=True, use_graph=True, stage_timing=False):
        check(load().fg_sim_set_options(int(fused), int(use_graph), int(stage_timing)), "options")

    def autotune(self, trial_steps=10):
        """Launch-shape tuning of the single-fluid step (process-wide); the
        simulation state is restored.  Returns the choice and the trials."""
        nt = C.c_int()
        tun = np.zeros(5, np.int32)
        trials = np.zeros((32, 6), np.float32)
        check(load().fg_sim_autotune(int(trial_steps), tun.ctypes.data, trials.ctypes.data, 32, C.byref(nt)),
              "fg_sim_autotune")
        return dict(sph_block=int(tun[0]), sort_digit_bits=int(tun[1]), flat_density=bool(tun[2]),
                    flat_forces=bool(tun[3]), tiled=bool(tun[4]),
                    trials=[(int(a), int(b), int(c), int(d), int(e), float(m)) for a, b, c, d, e, m in
                            trials[:nt.value]])

    def set_precision(self, mode):
        """Density arithmetic of the fused step: abi.PRECISION_FP32 or abi.PRECISION_MIXED."""
        check(load().fg_sim_set_precision(int(mode)), "fg_sim_set_precision")

    def set_inflow(self, desc):
        """Attach an abi.InflowDesc: reserve slots are emitted from the next step."""
        check(load().fg_sim_set_inflow(self.idx, C.byref(desc)), "fg_sim_set_inflow")

    def conservation(self):
        """FP64 sums over active particles: dict of count, mass, momentum (3),
        kinetic, potential (gravity) and internal (EOS) energy."""
        out = np.zeros(8, np.float64)
        check(load().fg_sim_conservation(self.idx, out.ctypes.data), "fg_sim_conservation")
        return {"count": out[0], "mass": out[1], "momentum": out[2:5].copy(), "kinetic": out[5],
                "potential": out[6], "internal": out[7]}

    def set_overlap(self, on=True):
        check(load().fg_sim_set_overlap(int(on)), "fg_sim_set_overlap")

    def save_state(self, path):
        """Checkpoint of the whole simulator (fg_sim_save_state)."""
        check(load().fg_sim_save_state(str(path).encode()), "fg_sim_save_state")

    def write_ply(self, path):
        """Active particles as a binary PLY point cloud; returns the count."""
        n = C.c_uint(0)
        check(load().fg_sim_write_ply(self.idx, str(path).encode(), C.byref(n)), "fg_sim_write_ply")
        return n.value

    def load_state(self, path):
        """Resume from a checkpoint of the same scene (fg_sim_load_state)."""
        check(load().fg_sim_load_state(str(path).encode()), "fg_sim_load_state")

    def step(self, steps=1):
        check(load().fg_sim_step(int(steps)), "fg_sim_step")

    def synchronize(self):
        check(load().fg_sim_synchronize(), "fg_sim_synchronize")

    def state(self):
        pos = np.empty((self.n, 4), np.float32)
        vel = np.empty((self.n, 4), np.float32)
        check(load().fg_sim_download(self.idx, pos.ctypes.data, vel.ctypes.data), "download")
        return pos, vel

    def upload(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        check(load().fg_sim_upload(self.idx, pos.ctypes.data, vel.ctypes.data), "upload")

    def arrays(self):
        """Intermediate arrays of the last step (keys, ids, cells, sorted
        state, density) as numpy."""
        self.synchronize()
        b, n, nb = self.buffers, self.n, self.buffers.numBuckets
        return dict(
            skey=download_raw(b.hashKey, (n,), np.uint32),
            spid=download_raw(b.particleId, (n,), np.uint32),
            start=download_raw(b.gridStart, (nb,), np.uint32),
            end=download_raw(b.gridEnd, (nb,), np.uint32),
            spos=download_raw(b.orderedPosition, (n, 4), np.float32),
            svel=download_raw(b.orderedVelocity, (n, 4), np.float32),
            rho_p=download_raw(b.densityPressure, (n, 2), np.float32),
        )

    def stage_times(self):
        out = (C.c_float * 7)()
        check(load().fg_sim_stage_times(out), "stage times")
        return dict(zip(("hash", "sort", "cells", "density", "forces", "integrate", "bodies"), list(out)))

    def add_bodies(self, bodies):
        """Couple a scenes.BodySet to the fluid."""
        world = np.ascontiguousarray(bodies.world, np.float32)
        offs = np.ascontiguousarray(bodies.offsets, np.uint32)
        mass = np.ascontiguousarray(bodies.mass, np.float64)
        check(load().fg_sim_add_bodies(world.ctypes.data, offs.ctypes.data, bodies.b, mass.ctypes.data,
                                       bodies.sphere_radius, C.byref(bodies.rb), bodies.spring, bodies.damping),
              "fg_sim_add_bodies")
        self.nbodies = bodies.b

    def add_static(self, bodies):
        """Add the static spheres of a scenes.BodySet (bodies.static) as one
        static body the fluid collides with."""
        sph = np.ascontiguousarray(bodies.static, np.float32)
        check(load().fg_sim_add_static_spheres(sph.ctypes.data, sph.shape[0], bodies.sphere_radius,
                                               C.byref(bodies.rb), bodies.spring, bodies.damping),
              "fg_sim_add_static_spheres")

    def static_grid(self):
        """The simulator's static grid, downloaded: (sorted centres, start,
        end, numBuckets)."""
        g = abi.BodyGrid()
        check(load().fg_sim_static_grid(C.byref(g)), "fg_sim_static_grid")
        m, nb = g.numSpheres, g.numBuckets
        if m == 0:
            return None
        L = load()
        pos = np.empty((m, 4), np.float32)
        start = np.empty(nb, np.uint32)
        end = np.empty(nb, np.uint32)
        for a, p in ((pos, g.orderedSpheres), (start, g.gridStart), (end, g.gridEnd)):
            check(L.fg_memcpy_dtoh(a.ctypes.data, p, a.nbytes), "download")
        return pos, start, end, nb, g

    def body_state(self):
        b = self.nbodies
        out = [np.empty((b, 4), np.float32) for _ in range(4)]
        check(load().fg_sim_body_state(*[o.ctypes.data for o in out]), "body state")
        return dict(zip(("pos", "vel", "orient", "L"), out))

    def body_sums(self):
        f = np.empty((self.nbodies, 4), np.float64)
        t = np.empty((self.nbodies, 4), np.float64)
        check(load().fg_sim_body_sums(f.ctypes.data, t.ctypes.data), "body sums")
        return f, t

    def body_mass_properties(self):
        cm = np.empty((self.nbodies, 3), np.float64)
        inertia = np.empty((self.nbodies, 3, 3), np.float64)
        check(load().fg_sim_body_mass_properties(cm.ctypes.data, inertia.ctypes.data), "mass properties")
        return cm, inertia

    def stream(self):
        return load().fg_sim_stream()

    def close(self):
        load().fg_sim_shutdown()


class BatchSimulator:
    """The batched-scene simulator (include/fluidgrid_amd_batch.h) over a
    scenes.Ensemble: one launch per stage for every scene."""

    def __init__(self, ens, use_graph=True, stage_timing=False):
        L = load()
        S = len(ens.scenes)
        commons = (abi.CommonFluidParams * S)(*[sc.common for sc in ens.scenes])
        fluids = (abi.FluidParams * S)(*[sc.fluid for sc in ens.scenes])
        counts = np.ascontiguousarray([sc.n for sc in ens.scenes], np.uint32)
        pos = np.ascontiguousarray(np.concatenate([sc.pos for sc in ens.scenes]), np.float32)
        vel = np.ascontiguousarray(np.concatenate([sc.vel for sc in ens.scenes]), np.float32)
        check(L.fg_bsim_create(S, C.cast(commons, C.c_void_p), C.cast(fluids, C.c_void_p), counts.ctypes.data,
                               pos.ctypes.data, vel.ctypes.data), "fg_bsim_create")
        self.n = int(pos.shape[0])
        self.counts = counts
        self.set_options(use_graph, stage_timing)

    def set_overlap(self, on=True):
        """Body stages on a second stream (default) or in line."""
        check(load().fg_bsim_set_overlap(int(on)), "fg_bsim_set_overlap")

    def set_options(self, use_graph=True, stage_timing=False):
        check(load().fg_bsim_set_options(int(use_graph), int(stage_timing)), "fg_bsim_set_options")

    def add_bodies(self, scene, bodies):
        """Couple a scenes.BodySet to the fluid of scene `scene` (before the first step)."""
        world = np.ascontiguousarray(bodies.world, np.float32)
        offs = np.ascontiguousarray(bodies.offsets, np.uint32)
        mass = np.ascontiguousarray(bodies.mass, np.float64)
        check(load().fg_bsim_add_bodies(int(scene), world.ctypes.data, offs.ctypes.data, bodies.b, mass.ctypes.data,
                                        bodies.sphere_radius, C.byref(bodies.rb), bodies.spring, bodies.damping),
              "fg_bsim_add_bodies")

    def body_state(self):
        """Per body (scenes in order): dict of pos, vel, orient, L (float32 (B, 4))."""
        L = load()
        nb = C.c_uint(0)
        check(L.fg_bsim_body_state(C.byref(nb), None, None, None, None), "fg_bsim_body_state")
        out = [np.empty((nb.value, 4), np.float32) for _ in range(4)]
        check(L.fg_bsim_body_state(C.byref(nb), *[o.ctypes.data for o in out]), "fg_bsim_body_state")
        return dict(zip(("pos", "vel", "orient", "L"), out))

    def step(self, steps=1):
        check(load().fg_bsim_step(int(steps)), "fg_bsim_step")

    def stream(self):
        return load().fg_bsim_stream()

    def synchronize(self):
        check(load().fg_bsim_synchronize(), "fg_bsim_synchronize")

    def state(self):
        pos = np.empty((self.n, 4), np.float32)
        vel = np.empty((self.n, 4), np.float32)
        check(load().fg_bsim_download(pos.ctypes.data, vel.ctypes.data), "fg_bsim_download")
        return pos, vel

    def density(self):
        rp = np.empty((self.n, 2), np.float32)
        pid = np.empty(self.n, np.uint32)
        check(load().fg_bsim_download_density(rp.ctypes.data, pid.ctypes.data), "fg_bsim_download_density")
        return rp, pid

    def stage_times(self):
        out = (C.c_float * 5)()
        check(load().fg_bsim_stage_times(out), "stage times")
        return dict(zip(("hash", "sort", "cells", "density", "forces"), list(out)))

    def info(self):
        n, nb, kb = C.c_uint(), C.c_uint(), C.c_int()
        check(load().fg_bsim_info(C.byref(n), C.byref(nb), C.byref(kb)), "fg_bsim_info")
        return dict(particles=n.value, buckets=nb.value, key_bits=kb.value)

    def autotune(self, trial_steps=10):
        nt = C.c_int()
        tun = np.zeros(6, np.int32)
        trials = np.zeros((32, 7), np.float32)
        check(load().fg_bsim_autotune(int(trial_steps), tun.ctypes.data, trials.ctypes.data, 32, C.byref(nt)),
              "fg_bsim_autotune")
        return dict(sph_block=int(tun[0]), sort_digit_bits=int(tun[1]), flat_density=bool(tun[2]),
                    flat_forces=bool(tun[3]), queue_items=int(tun[4]), defer_contact=bool(tun[5]),
                    trials=[(int(a), int(b), int(c), int(d), int(q), int(dc), float(e))
                            for a, b, c, d, q, dc, e in trials[:nt.value]])

    def set_tuning(self, sph_block=256, sort_digit_bits=8, flat_density=False, flat_forces=False, queue_items=0,
                   defer_contact=True):
        check(load().fg_bsim_set_tuning(int(sph_block), int(sort_digit_bits), int(flat_density), int(flat_forces),
                                        int(queue_items), int(defer_contact)), "fg_bsim_set_tuning")

    def close(self):
        load().fg_bsim_destroy()

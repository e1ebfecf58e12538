"""numpy front end of the CPU oracle (oracle/fg_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The HIP product path never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from fluidgrid_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfg_oracle.so")
EMPTY = np.uint32(0xFFFFFFFF)

_lib = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        cpp = C.POINTER(abi.CommonFluidParams)
        kcp = C.POINTER(abi.KernelsCache)
        L.orc_kernels_cache.argtypes = [C.c_float, kcp]
        L.orc_hash.argtypes = [_f32p, C.c_uint32, C.c_uint32, abi.Float3, cpp, C.c_int, _u32p, _u32p]
        L.orc_sort_pairs.argtypes = [_u32p, _u32p, C.c_uint32]
        L.orc_cell_offsets.argtypes = [_u32p, _u32p, _u32p, _u32p, _f32p, _f32p, C.c_void_p, C.c_void_p,
                                       C.c_uint32, C.c_uint32]
        L.orc_density_f32.argtypes = [_f32p, _f32p, _u32p, _u32p, C.c_uint32, C.c_float, C.c_float,
                                      C.c_float, cpp, kcp, C.c_int]
        L.orc_density_f64.argtypes = [_f32p, _f64p, _u32p, _u32p, C.c_uint32, C.c_float, C.c_float,
                                      C.c_float, cpp, kcp, C.c_int]
        L.orc_forces_f32.argtypes = [_f32p, _f32p, _f32p, _u32p, _u32p, _f32p, _u32p, C.c_uint32, C.c_float,
                                     C.c_float, cpp, kcp, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_forces_f64.argtypes = [_f32p, _f32p, _f64p, _u32p, _u32p, _f64p, _u32p, C.c_uint32, C.c_float,
                                     C.c_float, cpp, kcp, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_integrate.argtypes = [_f32p, _f32p, C.c_uint32, cpp]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_set_reverse_order.argtypes = [C.c_int]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def kernels_cache(h):
    kc = abi.KernelsCache()
    lib().orc_kernels_cache(h, C.byref(kc))
    return kc


def hash_keys(pos, num_buckets, cell_size, common, mode):
    pos = _c(pos, np.float32)
    n = pos.shape[0]
    key = np.zeros(n, np.uint32)
    pid = np.zeros(n, np.uint32)
    lib().orc_hash(pos, n, num_buckets, abi.Float3(*cell_size), C.byref(common), mode, key, pid)
    return key, pid


def sort_pairs(key, val):
    key = _c(key, np.uint32).copy()
    val = _c(val, np.uint32).copy()
    lib().orc_sort_pairs(key, val, key.shape[0])
    return key, val


def cell_offsets(key, pid, pos, vel, num_buckets, end_init=None):
    key = _c(key, np.uint32)
    pid = _c(pid, np.uint32)
    pos = _c(pos, np.float32)
    n = key.shape[0]
    start = np.zeros(num_buckets, np.uint32)
    end = np.zeros(num_buckets, np.uint32) if end_init is None else _c(end_init, np.uint32).copy()
    spos = np.zeros((n, 4), np.float32)
    if vel is not None:
        vel = _c(vel, np.float32)
        svel = np.zeros((n, 4), np.float32)
        lib().orc_cell_offsets(start, end, key, pid, pos, spos, vel.ctypes.data, svel.ctypes.data, n, num_buckets)
    else:
        svel = None
        lib().orc_cell_offsets(start, end, key, pid, pos, spos, None, None, n, num_buckets)
    return start, end, spos, svel


def density(spos, start, end, fluid, common, kernels, mode, f64=False, rho_p_init=None):
    spos = _c(spos, np.float32)
    n = spos.shape[0]
    dt = np.float64 if f64 else np.float32
    out = np.zeros((n, 2), dt) if rho_p_init is None else _c(rho_p_init, dt).copy()
    fn = lib().orc_density_f64 if f64 else lib().orc_density_f32
    fn(spos, out, _c(start, np.uint32), _c(end, np.uint32), n, fluid.gasConstant, fluid.restDensity,
       fluid.mass, C.byref(common), C.byref(kernels), mode)
    return out


def forces(spos, svel, rho_p, start, end, pid, fluid, common, kernels, mode, f64=False, v_out_init=None,
           want_acc=False, want_mag=False):
    spos = _c(spos, np.float32)
    svel = _c(svel, np.float32)
    n = spos.shape[0]
    dt = np.float64 if f64 else np.float32
    rho_p = _c(rho_p, dt)
    v_out = np.zeros((n, 4), dt) if v_out_init is None else _c(v_out_init, dt).copy()
    acc = np.zeros((n, 3), dt) if want_acc else None
    mag = np.zeros(n, dt) if want_mag else None
    fn = lib().orc_forces_f64 if f64 else lib().orc_forces_f32
    fn(spos, svel, rho_p, _c(start, np.uint32), _c(end, np.uint32), v_out, _c(pid, np.uint32), n,
       fluid.mass, fluid.viscosity, C.byref(common), C.byref(kernels), mode,
       acc.ctypes.data if want_acc else None, mag.ctypes.data if want_mag else None)
    return v_out, acc, mag


def integrate(pos, vel, common):
    pos = _c(pos, np.float32).copy()
    vel = _c(vel, np.float32)
    lib().orc_integrate(pos, vel, pos.shape[0], C.byref(common))
    return pos


class Step:
    """Intermediate results of one oracle step (all numpy arrays)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def step(pos, vel, fluid, common, kernels, mode, rho_p_prev=None, end_init=None, want_acc=False,
         want_mag=False):
    """One FP32 oracle step, in the order of HeterogeneousSimulator::doTimeStep
    (HeterogeneousSimulator.cpp:54-127).  Returns every intermediate array."""
    nb = common.num_buckets()
    key, pid = hash_keys(pos, nb, tuple(common.cell_size), common, mode)
    skey, spid = sort_pairs(key, pid)
    start, end, spos, svel = cell_offsets(skey, spid, pos, vel, nb, end_init)
    rho_p = density(spos, start, end, fluid, common, kernels, mode, rho_p_init=rho_p_prev)
    v_new, acc, mag = forces(spos, svel, rho_p, start, end, spid, fluid, common, kernels, mode,
                             v_out_init=vel, want_acc=want_acc, want_mag=want_mag)
    p_new = integrate(pos, v_new, common)
    return Step(key=key, pid=pid, skey=skey, spid=spid, start=start, end=end, spos=spos, svel=svel,
                rho_p=rho_p, vel=v_new, pos=p_new, acc=acc, mag=mag)


def set_num_threads(t):
    lib().orc_set_num_threads(int(t))


def num_threads():
    return lib().orc_num_threads()


def set_reverse_order(flag):
    """Walk every bucket last-to-first (an alternative FP32 summation order)."""
    lib().orc_set_reverse_order(int(bool(flag)))


def run_steps(scene, steps, reverse=False, mode=None):
    """`steps` FP32 oracle steps of a scene; returns final (pos, vel, rho_p)."""
    mode = scene.key_mode if mode is None else mode
    pos, vel = scene.pos.copy(), scene.vel.copy()
    rho_p = np.zeros((scene.n, 2), np.float32)
    set_reverse_order(reverse)
    try:
        for _ in range(steps):
            st = step(pos, vel, scene.fluid, scene.common, scene.kernels, mode, rho_p_prev=rho_p)
            pos, vel, rho_p = st.pos, st.vel, st.rho_p
    finally:
        set_reverse_order(False)
    return pos, vel, rho_p

"""ctypes mirrors of the POD structs in include/fluidgrid_amd.h.

The layouts follow the reference's ApplicationData.cuh:23-112 (CommonFluidParams
104 bytes, FluidParams 28, KernelsCache 12, CommonRigidBodyParams 112) and are
checked against the C header by tests/test_abi.py.
"""
import ctypes as C


class Float2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class Float4(C.Structure):
    _pack_ = 16
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Double4(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("w", C.c_double)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class FluidParams(C.Structure):
    """ApplicationData.cuh:32-41"""
    _fields_ = [
        ("mass", C.c_float),
        ("restDensity", C.c_float),
        ("viscosity", C.c_float),
        ("stiffness", C.c_float),
        ("particle_radius", C.c_float),
        ("gasConstant", C.c_float),
        ("max_particles", C.c_int),
    ]


class CommonFluidParams(C.Structure):
    """ApplicationData.cuh:46-67"""
    _fields_ = [
        ("cell_size", Float3),
        ("bucket_size", Int3),
        ("keyBits", C.c_uint),
        ("kernel_radius", C.c_float),
        ("kernel_radius_squared", C.c_float),
        ("pointRadius", C.c_float),
        ("pointScale", C.c_float),
        ("velocity_limit", C.c_float),
        ("volume_min", Float3),
        ("volume_max", Float3),
        ("global_external_force", Float3),
        ("simulation_scale", C.c_float),
        ("deltaTimeIntegration", C.c_float),
        ("fluid_to_srb_relation", Int3),
    ]

    def num_buckets(self):
        b = self.bucket_size
        return b.x * b.y * b.z


class KernelsCache(C.Structure):
    """ApplicationData.cuh:107-112"""
    _fields_ = [("kPoly6", C.c_float), ("kSpikyGradient", C.c_float), ("kViscosityLaplacian", C.c_float)]


class _Float4Aligned(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class CommonRigidBodyParams(C.Structure):
    """ApplicationData.cuh:73-102 (float4 member aligned to 16 -> 12 bytes of padding)"""
    _fields_ = [
        ("rb_to_fluid_relation", Int3),
        ("st_bucket_size", Int3),
        ("dn_bucket_size", Int3),
        ("st_keyBits", C.c_uint),
        ("velocity_limit", C.c_float),
        ("volume_min", Float3),
        ("volume_max", Float3),
        ("_pad0", C.c_byte * 12),
        ("global_external_force", _Float4Aligned),
        ("simulation_scale", C.c_float),
        ("deltaTimeIntegration", C.c_float),
        ("particleRadius", C.c_float),
        ("_pad1", C.c_byte * 4),
    ]


class InflowDesc(C.Structure):
    """fg_inflow_desc (include/fluidgrid_amd.h)"""
    _fields_ = [("origin", Float3), ("axisU", Float3), ("axisV", Float3), ("normal", Float3), ("spacing", C.c_float),
                ("speed", C.c_float), ("nu", C.c_uint), ("nv", C.c_uint), ("firstSlot", C.c_uint),
                ("numLayers", C.c_uint)]


class BodyGrid(C.Structure):
    """fg_body_grid (include/fluidgrid_amd.h)"""
    _fields_ = [("orderedSpheres", C.c_void_p), ("bodyVelocities", C.c_void_p), ("gridStart", C.c_void_p),
                ("gridEnd", C.c_void_p), ("bounds", C.c_void_p), ("cellSize", Float3), ("numBuckets", C.c_uint),
                ("relation", Int3), ("numSpheres", C.c_uint), ("numBodies", C.c_uint), ("fluidRadius", C.c_float),
                ("sphereRadius", C.c_float), ("kSpring", C.c_float), ("kDamping", C.c_float)]


class SceneDesc(C.Structure):
    """fg_scene_desc (include/fluidgrid_amd.h): one scene of a batch."""
    _fields_ = [("common", CommonFluidParams), ("kernels", KernelsCache), ("mass", C.c_float),
                ("gasConstant", C.c_float), ("restDensity", C.c_float), ("viscosity", C.c_float),
                ("keyBase", C.c_uint), ("particleBegin", C.c_uint), ("particleCount", C.c_uint),
                ("reserved", C.c_uint * 4)]


KEY_REF_XOR = 0
KEY_LINEAR = 1
MATH_FAST = 0
MATH_REFERENCE = 1
PRECISION_FP32 = 0
PRECISION_MIXED = 1

# fg_status codes (include/fluidgrid_amd.h)
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_OUT_OF_MEMORY = 3
ERR_NOT_INITIALIZED = 4
ERR_UNSUPPORTED = 5

assert C.sizeof(CommonFluidParams) == 104
assert C.sizeof(FluidParams) == 28
assert C.sizeof(KernelsCache) == 12
assert C.sizeof(CommonRigidBodyParams) == 112
assert C.sizeof(SceneDesc) == 160


def f3(x, y, z):
    return Float3(x, y, z)


def i3(x, y, z):
    return Int3(x, y, z)

"""Triangle meshes for rigid bodies: generators of closed test meshes, an OBJ
writer, and the C-ABI of include/fluidgrid_amd_mesh.h (OBJ reading, the
sampling lattice, GPU winding numbers, sphere sampling, body particle files
as read by the reference's RigidBody::loadFromFile,
Simulators/RigidBody.cpp:55-108)."""
import ctypes as C

import numpy as np

from . import abi
from . import lib as _lib


# ---------------------------------------------------------------------------
# Closed, outward-oriented meshes (vertices float32 (V, 3), triangles
# uint32 (T, 3), corners counter-clockwise seen from outside)
# ---------------------------------------------------------------------------

def box_mesh(lo, hi):
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    v = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = [(a, b, c) for a, b, c, d in quads] + [(a, c, d) for a, b, c, d in quads]
    return v, np.array(t, np.uint32)


def icosphere_mesh(radius, levels, centre=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron with vertices on the sphere (20 * 4^levels
    triangles)."""
    g = (1 + 5 ** 0.5) / 2
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        f = [t for a, b, c in f for t in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                          (m(a, b), m(b, c), m(c, a)))]
    return (np.array(v) * radius + np.asarray(centre)).astype(np.float32), np.array(f, np.uint32)


def torus_mesh(major, minor, nu, nv, centre=(0.0, 0.0, 0.0)):
    """Torus around the y axis, nu segments around it and nv around the tube."""
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    r = major + minor * np.cos(W)
    v = np.stack([r * np.cos(U), minor * np.sin(W), r * np.sin(U)], -1).reshape(-1, 3)
    idx = lambda i, j: (i % nu) * nv + (j % nv)
    t = []
    for i in range(nu):
        for j in range(nv):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            t += [(a, d, c), (a, c, b)]
    return (v + np.asarray(centre)).astype(np.float32), np.array(t, np.uint32)


def write_obj(path, vertices, triangles, quads=False, corner_style="i", relative=False):
    """Writes a Wavefront OBJ file.  With quads, consecutive triangle pairs
    (a, b, c), (a, c, d) are written as one quad (a, b, c, d).  corner_style
    is one of "i", "i/t", "i//n", "i/t/n"; with relative, indices are
    negative (counted back from the last vertex)."""
    V = len(vertices)
    fmt = {"i": "{}", "i/t": "{}/1", "i//n": "{}//1", "i/t/n": "{}/1/1"}[corner_style]
    ref = (lambda k: str(k - V)) if relative else (lambda k: str(k + 1))
    with open(path, "w") as f:
        f.write("# mesh\n")
        for p in vertices:
            f.write(f"v {float(p[0]):.9g} {float(p[1]):.9g} {float(p[2]):.9g}\n")
        f.write("vt 0 0\nvn 0 1 0\n")
        t = np.asarray(triangles).reshape(-1, 3)
        i = 0
        while i < len(t):
            if quads and i + 1 < len(t) and t[i + 1][0] == t[i][0] and t[i + 1][1] == t[i][2]:
                cs = [t[i][0], t[i][1], t[i][2], t[i + 1][2]]
                i += 2
            else:
                cs = list(t[i])
                i += 1
            f.write("f " + " ".join(fmt.format(ref(int(k))) for k in cs) + "\n")


# ---------------------------------------------------------------------------
# C-ABI wrappers
# ---------------------------------------------------------------------------

def _arrays(vertices, triangles):
    v = np.ascontiguousarray(np.asarray(vertices, np.float32)[:, :3])
    t = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1))
    return v, t


def load_obj(path):
    L = _lib.load()
    nv, nt = C.c_uint(0), C.c_uint(0)
    _lib.check(L.fg_mesh_load_obj(str(path).encode(), C.byref(nv), C.byref(nt), None, None), "fg_mesh_load_obj")
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    _lib.check(L.fg_mesh_load_obj(str(path).encode(), C.byref(nv), C.byref(nt), v.ctypes.data, t.ctypes.data),
               "fg_mesh_load_obj")
    return v, t


def lattice(vertices, spacing):
    L = _lib.load()
    v, _ = _arrays(vertices, np.zeros(0))
    o = abi.Float3()
    dims = np.zeros(3, np.uint32)
    _lib.check(L.fg_mesh_lattice(v.ctypes.data, len(v), spacing, C.byref(o), dims.ctypes.data), "fg_mesh_lattice")
    return np.array([o.x, o.y, o.z], np.float32), tuple(int(d) for d in dims)


def winding_numbers(vertices, triangles, spacing):
    """GPU winding numbers of the mesh's sampling lattice (x fastest)."""
    L = _lib.load()
    v, t = _arrays(vertices, triangles)
    _, dims = lattice(v, spacing)
    w = np.zeros(dims[0] * dims[1] * dims[2], np.float32)
    _lib.check(L.fg_mesh_winding_numbers(v.ctypes.data, len(v), t.ctypes.data, len(t) // 3, spacing, w.ctypes.data),
               "fg_mesh_winding_numbers")
    return w


def sample_spheres(vertices, triangles, spacing, shell=False):
    L = _lib.load()
    v, t = _arrays(vertices, triangles)
    n = C.c_uint(0)
    _lib.check(L.fg_mesh_sample_spheres(v.ctypes.data, len(v), t.ctypes.data, len(t) // 3, spacing, int(shell),
                                        None, 0, C.byref(n)), "fg_mesh_sample_spheres")
    out = np.zeros((n.value, 4), np.float32)
    _lib.check(L.fg_mesh_sample_spheres(v.ctypes.data, len(v), t.ctypes.data, len(t) // 3, spacing, int(shell),
                                        out.ctypes.data, n.value, C.byref(n)), "fg_mesh_sample_spheres")
    return out[:n.value]


def write_body_file(path, spheres, radius, mass=1.0, static=False, mesh_name=""):
    L = _lib.load()
    s = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 4) if np.asarray(spheres).shape[-1] == 4
                             else np.concatenate([np.asarray(spheres, np.float32),
                                                  np.zeros((len(spheres), 1), np.float32)], 1))
    _lib.check(L.fg_body_file_write(str(path).encode(), int(static), mass, radius, mesh_name.encode(), s.ctypes.data,
                                    len(s)), "fg_body_file_write")


def read_body_file(path):
    """(spheres float32 (n, 4) relative to the body origin with w = radius,
    radius, mass, static, mesh name)."""
    L = _lib.load()
    st, mass, rad, n = C.c_int(0), C.c_float(0), C.c_float(0), C.c_uint(0)
    _lib.check(L.fg_body_file_read(str(path).encode(), C.byref(st), C.byref(mass), C.byref(rad), C.byref(n), None, 0,
                                   None, 0), "fg_body_file_read")
    out = np.zeros((n.value, 4), np.float32)
    name = C.create_string_buffer(4096)
    _lib.check(L.fg_body_file_read(str(path).encode(), C.byref(st), C.byref(mass), C.byref(rad), C.byref(n),
                                   out.ctypes.data, n.value, name, 4096), "fg_body_file_read")
    return out, rad.value, mass.value, bool(st.value), name.value.decode()

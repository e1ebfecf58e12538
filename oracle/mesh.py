"""CPU oracle of the mesh sampler (fluidgrid_amd/csrc/fg_mesh.hip and
host/MeshSampler.cpp): the sampling lattice, generalized winding numbers in
FP64 and the inside/shell selection.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
path never imports this module.

The reference has no sampler: its body particle files (read by
RigidBody::loadFromFile, Simulators/RigidBody.cpp:55-108) carry a mesh name
and sphere centres made outside it.  What is restated here is the published
method: the solid angle of a triangle (Van Oosterom & Strackee, IEEE Trans.
Biomed. Eng. 30(2):125-126, 1983) summed into a winding number (Jacobson,
Kavan & Sorkine-Hornung, ACM TOG 32(4), 2013).  Pins (tests/test_oracle_mesh.py):
the octant triangle subtends pi/2, a closed cube gives 1 inside and 0
outside, an open cube gives 5/6 at its centre, flipped faces negate w,
a fine sphere mesh classifies points by radius, brute-force inside tests of
a convex mesh agree.
"""
import numpy as np


def lattice(vertices, spacing):
    """Origin (float32[3]) and point counts of the lattice over the vertex
    bounds: n = max(1, ceil((hi - lo) / s)) per axis, origin = (lo + hi) / 2
    - (n - 1) s / 2 in FP64, rounded to FP32 (MeshSampler.cpp meshLattice)."""
    v = np.asarray(vertices, np.float32)[:, :3]
    s = float(np.float32(spacing))
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    n = np.maximum(1, np.ceil((hi - lo) / s)).astype(np.int64)
    origin = (0.5 * (lo + hi) - 0.5 * (n - 1) * s).astype(np.float32)
    return origin, tuple(int(k) for k in n)


def lattice_points(origin, spacing, dims):
    """Lattice points in x-fastest order: origin + i * spacing per axis,
    product and sum each rounded to FP32."""
    nx, ny, nz = dims
    s = np.float32(spacing)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    o = np.asarray(origin, np.float32)
    pts = np.stack([o[0] + ix.astype(np.float32) * s, o[1] + iy.astype(np.float32) * s,
                    o[2] + iz.astype(np.float32) * s], -1)
    return pts.reshape(-1, 3).astype(np.float32)


def solid_angles(points, a, b, c):
    """Signed solid angle of each triangle (a, b, c) seen from each point, FP64:
    2 atan2(a.(b x c), |a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|) with the
    corners relative to the point.  Shapes: points (P, 3), corners (T, 3);
    result (P, T)."""
    p = np.asarray(points, np.float64)[:, None, :]
    A = np.asarray(a, np.float64)[None] - p
    B = np.asarray(b, np.float64)[None] - p
    Cc = np.asarray(c, np.float64)[None] - p
    la, lb, lc = (np.sqrt(np.einsum("ptk,ptk->pt", X, X)) for X in (A, B, Cc))
    det = np.einsum("ptk,ptk->pt", A, np.cross(B, Cc))
    den = (la * lb * lc + np.einsum("ptk,ptk->pt", A, B) * lc + np.einsum("ptk,ptk->pt", B, Cc) * la
           + np.einsum("ptk,ptk->pt", Cc, A) * lb)
    return 2.0 * np.arctan2(det, den)


def winding(points, vertices, triangles, impl="c"):
    """FP64 winding number of each point: sum of solid angles / (4 pi).
    Triangles with a corner index beyond the vertex count add 0.  impl "c"
    runs orc_winding of the compiled oracle (OpenMP), "numpy" the
    vectorised restatement below; the pins check both."""
    if impl == "c":
        from oracle import oracle as _o
        L = _o.lib()
        pts = np.ascontiguousarray(np.asarray(points, np.float32)[:, :3])
        v = np.ascontiguousarray(np.asarray(vertices, np.float32)[:, :3])
        t = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1))
        out = np.zeros(len(pts))
        L.orc_winding(pts, len(pts), v, len(v), t, len(t) // 3, out)
        return out
    return _winding_numpy(points, vertices, triangles)


def _winding_numpy(points, vertices, triangles, chunk=2048):
    v = np.asarray(vertices, np.float32)[:, :3].astype(np.float64)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    t = t[np.all(t < len(v), axis=1)]
    pts = np.asarray(points, np.float32)
    out = np.zeros(len(pts))
    if len(t) == 0:
        return out
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    step = max(1, chunk * 256 // max(1, len(t)))
    for i in range(0, len(pts), step):
        out[i:i + step] = solid_angles(pts[i:i + step], a, b, c).sum(1) / (4.0 * np.pi)
    return out


def sample(vertices, triangles, spacing, shell=False, w=None):
    """Lattice points with winding number > 1/2 in lattice order (w = spacing
    / 2); with shell, only those with a face neighbour outside or beyond the
    lattice (MeshSampler.cpp sampleMeshSpheres).  `w` may supply the winding
    numbers of the lattice (x fastest) to classify instead of the FP64 ones."""
    origin, dims = lattice(vertices, spacing)
    pts = lattice_points(origin, spacing, dims)
    if w is None:
        w = winding(pts, vertices, triangles)
    nx, ny, nz = dims
    ins = (np.asarray(w) > 0.5).reshape(nz, ny, nx)
    keep = ins.copy()
    if shell:
        pad = np.pad(ins, 1, constant_values=False)
        interior = (pad[1:-1, 1:-1, :-2] & pad[1:-1, 1:-1, 2:] & pad[1:-1, :-2, 1:-1] & pad[1:-1, 2:, 1:-1]
                    & pad[:-2, 1:-1, 1:-1] & pad[2:, 1:-1, 1:-1])
        keep &= ~interior
    sel = pts[keep.reshape(-1)]
    return np.concatenate([sel, np.full((len(sel), 1), np.float32(0.5) * np.float32(spacing), np.float32)], 1)

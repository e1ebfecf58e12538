"""Pins of the mesh-sampler oracle (oracle/mesh.py) against closed forms.

  octant triangle          the triangle (e_x, e_y, e_z) seen from the origin
                           subtends 4 pi / 8 = pi / 2; reversed, -pi / 2
  closed cube              w = 1 at inside points, 0 outside (any distance)
  open cube (one face      w = 5/6 at the centre: each face subtends 4 pi / 6
    removed)
  flipped faces            w -> -w
  sphere mesh              a level-3 icosphere classifies points by radius
                           (inside the inradius: 1; outside the circumradius: 0)
  brute force              for a convex mesh, inside = on the inner side of
                           every face plane, checked on random points
  lattice                  counts and origin centred on the bounds, spacing
                           exact in FP32 steps
  shell                    a solid block's shell is its surface layer: the
                           inner points are exactly the block shrunk by one
  two implementations      the compiled FP64 oracle (orc_winding) and the numpy
                           restatement pass every pin above and agree to 1e-12
"""
import numpy as np
import pytest

from fluidgrid_amd import mesh as fm
from oracle import mesh as om


def test_octant_triangle_subtends_pi_over_2():
    e = np.eye(3)
    om_ = om.solid_angles(np.zeros((1, 3)), e[0:1], e[1:2], e[2:3])[0, 0]
    assert abs(om_ - np.pi / 2) < 1e-14
    rev = om.solid_angles(np.zeros((1, 3)), e[0:1], e[2:3], e[1:2])[0, 0]
    assert abs(rev + np.pi / 2) < 1e-14


IMPLS = pytest.mark.parametrize("impl", ["c", "numpy"])


@IMPLS
def test_closed_cube_inside_outside(impl):
    v, t = fm.box_mesh((-1, -2, -0.5), (2, 1, 0.75))
    rng = np.random.default_rng(1)
    inside = rng.uniform((-0.99, -1.99, -0.49), (1.99, 0.99, 0.74), (500, 3)).astype(np.float32)
    outside = rng.uniform(-10, 10, (3000, 3)).astype(np.float32)
    out_mask = np.any((outside < (-1, -2, -0.5)) | (outside > (2, 1, 0.75)), axis=1)
    outside = outside[out_mask]
    assert np.allclose(om.winding(inside, v, t, impl=impl), 1.0, atol=1e-12)
    assert np.allclose(om.winding(outside, v, t, impl=impl), 0.0, atol=1e-12)


@IMPLS
def test_open_cube_centre_is_five_sixths(impl):
    v, t = fm.box_mesh((-1, -1, -1), (1, 1, 1))
    keep = np.ones(len(t), bool)
    # drop the two triangles of the z = +1 face
    keep[[i for i, tri in enumerate(t) if np.all(v[tri][:, 2] == 1.0)]] = False
    assert keep.sum() == 10
    w = om.winding(np.zeros((1, 3), np.float32), v, t[keep], impl=impl)[0]
    assert abs(w - 5.0 / 6.0) < 1e-14


@IMPLS
def test_flipped_faces_negate(impl):
    v, t = fm.torus_mesh(1.0, 0.3, 24, 12)
    pts = np.random.default_rng(2).uniform(-1.5, 1.5, (400, 3)).astype(np.float32)
    w = om.winding(pts, v, t, impl=impl)
    wf = om.winding(pts, v, t[:, ::-1], impl=impl)
    assert np.allclose(wf, -w, atol=1e-12)
    assert np.any(w > 0.9) and np.any(np.abs(w) < 0.1)


@IMPLS
def test_icosphere_classifies_by_radius(impl):
    R = 0.7
    v, t = fm.icosphere_mesh(R, 3)
    assert len(t) == 20 * 4 ** 3
    # inradius of the mesh: smallest distance from the centre to a face plane
    a, b, c = (v[t[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    assert np.all(np.einsum("ij,ij->i", n, a) > 0)  # outward orientation
    r_in = np.einsum("ij,ij->i", n, a).min()
    pts = np.random.default_rng(3).uniform(-1, 1, (4000, 3)).astype(np.float32)
    r = np.linalg.norm(pts.astype(np.float64), axis=1)
    w = om.winding(pts, v, t, impl=impl)
    assert np.allclose(w[r < r_in * 0.999], 1.0, atol=1e-9)
    assert np.allclose(w[r > R * 1.001], 0.0, atol=1e-9)


@IMPLS
def test_convex_brute_force(impl):
    v, t = fm.icosphere_mesh(1.0, 1, centre=(0.2, -0.1, 0.3))
    a, b, c = (v[t[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    pts = np.random.default_rng(4).uniform(-1.2, 1.6, (3000, 3)).astype(np.float32)
    side = np.einsum("tk,ptk->pt", n, pts[:, None, :].astype(np.float64) - a[None])
    brute = np.all(side < 0, axis=1)
    margin = np.abs(side / np.linalg.norm(n, axis=1)).min(1) > 1e-6
    w = om.winding(pts, v, t, impl=impl)
    assert np.array_equal((w > 0.5)[margin], brute[margin])
    assert 200 < brute.sum() < 2800


def test_lattice_counts_and_origin():
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.25, 0.1], [0.5, -0.25, 0.3]], np.float32)
    origin, dims = om.lattice(v, 0.1)
    s = float(np.float32(0.1))
    assert dims == (int(np.ceil(1.0 / s)), int(np.ceil(float(np.float32(0.25) - np.float32(-0.25)) / s)),
                    int(np.ceil(float(np.float32(0.3)) / s)))
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    centre = 0.5 * (lo + hi)
    for a in range(3):
        assert abs(origin[a] + 0.5 * (dims[a] - 1) * s - centre[a]) < 1e-6
    pts = om.lattice_points(origin, 0.1, dims)
    assert pts.shape == (dims[0] * dims[1] * dims[2], 3)
    assert np.array_equal(pts[1, 0], np.float32(origin[0] + np.float32(s)))
    assert np.array_equal(pts[dims[0], 1], np.float32(origin[1] + np.float32(s)))
    # a flat mesh still gets one layer
    _, d = om.lattice(np.array([[0, 0, 0], [1, 1, 0]], np.float32), 0.3)
    assert d[2] == 1


def test_shell_of_a_block():
    v, t = fm.box_mesh((0, 0, 0), (1.0, 0.6, 0.8))
    s = 0.1
    solid = om.sample(v, t, s)
    shell = om.sample(v, t, s, shell=True)
    _, dims = om.lattice(v, s)
    assert len(solid) == dims[0] * dims[1] * dims[2]  # every lattice point is inside the block
    inner = (dims[0] - 2) * (dims[1] - 2) * (dims[2] - 2)
    assert len(shell) == len(solid) - inner
    assert np.all(solid[:, 3] == np.float32(0.05))


def test_c_and_numpy_agree_with_bad_triangles():
    v, t = fm.torus_mesh(0.5, 0.2, 16, 8)
    t = np.concatenate([t, [[0, 1, len(v)]], [[len(v) + 3, 0, 1]]]).astype(np.uint32)
    pts = np.random.default_rng(6).uniform(-0.8, 0.8, (700, 3)).astype(np.float32)
    a = om.winding(pts, v, t, impl="c")
    b = om.winding(pts, v, t, impl="numpy")
    assert np.abs(a - b).max() < 1e-12
    assert np.array_equal(a, om.winding(pts, v, t[:-2], impl="c"))

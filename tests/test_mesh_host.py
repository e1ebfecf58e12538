"""Host logic of the mesh sampler and the body particle file format through
the C-ABI (include/fluidgrid_amd_mesh.h); none of these calls touches a GPU.

  OBJ reader           the triangulation written by fluidgrid_amd.mesh.write_obj
                       reads back exactly, for every corner style, quads
                       (fan-split) and negative indices; bad corners, short
                       faces and missing files are errors
  lattice              equal to the oracle's FP64 restatement bit for bit
  body file            the reference's format (Simulators/RigidBody.cpp:55-108):
                       one item per line, "static" or a mass line, radius,
                       mesh name (spaces kept), count, "x y z" lines; w = radius;
                       CRLF line ends; a malformed number reads as atof does
                       (0); written files read back bit for bit
"""
import numpy as np
import pytest

from fluidgrid_amd import lib
from fluidgrid_amd import mesh as fm
from oracle import mesh as om


@pytest.mark.parametrize("style", ["i", "i/t", "i//n", "i/t/n"])
@pytest.mark.parametrize("relative", [False, True])
def test_obj_round_trip(tmp_path, style, relative):
    v, t = fm.torus_mesh(1.0, 0.25, 10, 6)
    p = tmp_path / "t.obj"
    fm.write_obj(p, v, t, quads=True, corner_style=style, relative=relative)
    lv, lt = fm.load_obj(p)
    assert np.array_equal(lv, v)
    assert np.array_equal(lt, t)


def test_obj_polygon_fan_and_ignored_statements(tmp_path):
    p = tmp_path / "poly.obj"
    p.write_text("o thing\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0.5 1.5 0\nv 0 1 0 1.0\nvn 0 0 1\ns off\n"
                 "usemtl m\nf 1 2 3 4 5\n")
    v, t = fm.load_obj(p)
    assert v.shape == (5, 3)
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nf 1 2 3\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n",
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "v 0 0\n", "v 0 0 0\nf 1 -2 x\n"])
def test_obj_errors(tmp_path, text):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(lib.FGError):
        fm.load_obj(p)


def test_obj_missing_file(tmp_path):
    with pytest.raises(lib.FGError, match="cannot open"):
        fm.load_obj(tmp_path / "none.obj")


@pytest.mark.parametrize("spacing", [0.01, 0.037, 0.25, 3.0])
def test_lattice_matches_oracle(spacing):
    v, _ = fm.torus_mesh(0.8, 0.2, 17, 9, centre=(0.31, -0.2, 1.7))
    o, d = fm.lattice(v, spacing)
    oo, od = om.lattice(v, spacing)
    assert d == od
    assert np.array_equal(o, oo)


def test_lattice_rejects_bad_spacing():
    v, _ = fm.box_mesh((0, 0, 0), (1, 1, 1))
    for s in (0.0, -1.0, float("nan")):
        with pytest.raises(lib.FGError):
            fm.lattice(v, s)
    with pytest.raises(lib.FGError):
        fm.lattice(v, 1e-6)  # over 65536 points per axis


@pytest.mark.parametrize("static", [False, True])
def test_body_file_round_trip(tmp_path, static):
    rng = np.random.default_rng(5)
    s = rng.normal(0, 0.3, (257, 3)).astype(np.float32)
    p = tmp_path / "body.txt"
    fm.write_body_file(p, s, radius=0.0125, mass=3.25, static=static, mesh_name="resources/my mesh.obj")
    got, rad, mass, st, name = fm.read_body_file(p)
    assert st == static and rad == np.float32(0.0125) and name == "resources/my mesh.obj"
    assert mass == (0.0 if static else np.float32(3.25))
    assert np.array_equal(got[:, :3], s)
    assert np.all(got[:, 3] == np.float32(0.0125))
    lines = p.read_text().splitlines()
    assert lines[0] == ("static" if static else "dynamic")
    assert lines[-1].count(" ") == 2 and len(lines) == (4 if static else 5) + 257


def test_body_file_reads_reference_layout(tmp_path):
    p = tmp_path / "ref.txt"
    p.write_bytes(b"dyn\r\n2.5\r\n0.01\r\nbunny.obj\r\n3\r\n0.1 0.2 0.3\r\n-1 0 1e-3\r\nabc 4 5\r\n")
    s, rad, mass, st, name = fm.read_body_file(p)
    assert not st and mass == np.float32(2.5) and rad == np.float32(0.01) and name == "bunny.obj"
    assert np.array_equal(s[:, :3], np.array([[0.1, 0.2, 0.3], [-1.0, 0.0, 1e-3], [0.0, 4.0, 5.0]], np.float32))
    p.write_text("static\n0.02\nplate.obj\n0\n")
    s, rad, mass, st, _ = fm.read_body_file(p)
    assert st and mass == 0.0 and len(s) == 0


def test_body_file_short_file_is_an_error(tmp_path):
    p = tmp_path / "short.txt"
    p.write_text("static\n0.02\nplate.obj\n5\n0 0 0\n")
    with pytest.raises(lib.FGError, match="sphere"):
        fm.read_body_file(p)
    with pytest.raises(lib.FGError, match="mesh name spans lines"):
        fm.write_body_file(tmp_path / "x.txt", np.zeros((1, 3)), 0.1, mesh_name="a\nb")

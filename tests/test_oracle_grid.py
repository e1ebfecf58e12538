"""Pins of the oracle's grid stage (hash, sort, cell offsets, reorder).

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned here by literal known answers, independent library
routines (numpy's stable argsort and unique) and invariants:

  orc_hash (REF_XOR)     literal keys computed by hand from K_DeviceFunctions.cu:17-26,
                         truncation toward zero, key range
  orc_hash (LINEAR)      cell coordinates vs numpy floor division, key range,
                         clamping of outside particles
  orc_sort_pairs         numpy argsort(kind="stable") on keys with ties, both
                         the counting path and the std::stable_sort path
  orc_cell_offsets       numpy unique(return_index, return_counts), cellEnd
                         left untouched for empty cells, gather through ids
"""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes


def _cp_for(cs, nb_dims=(64, 64, 64), vmin=(0.0, 0.0, 0.0)):
    cp = abi.CommonFluidParams()
    cp.cell_size = abi.f3(*cs)
    cp.bucket_size = abi.i3(*nb_dims)
    cp.volume_min = abi.f3(*vmin)
    return cp


# (position, cell size, numBuckets) -> key, computed by hand:
# xi = trunc((p + 5000) / cs) in float32, key = (xi*73856093 ^ yi*19349663 ^ zi*83492791) mod 2^32 % nb
REF_XOR_KNOWN = [
    ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 262144, 63272),
    ((-20.0, -20.0, -20.0), (1.9910525,) * 3, 262144, 14105),
    ((0.5, 85.0, 39.9), (1.9910525,) * 3, 262144, 164817),
    ((1.0, 2.0, 3.0), (0.025,) * 3, 64000, 4656),
    ((-5000.5, 0.0, 0.0), (1.0, 1.0, 1.0), 1000, 56),      # (-0.5) truncates to 0, not -1
    ((0.001, 0.999, 0.5), (0.025,) * 3, 64000, 35765),
]


@pytest.mark.parametrize("p,cs,nb,expected", REF_XOR_KNOWN)
def test_ref_xor_known_answers(orc, p, cs, nb, expected):
    pos = np.array([[*p, 1.0]], np.float32)
    key, pid = orc.hash_keys(pos, nb, cs, _cp_for(cs), abi.KEY_REF_XOR)
    assert int(key[0]) == expected
    assert int(pid[0]) == 0


def test_ref_xor_range_and_ids(orc):
    rng = np.random.default_rng(1)
    pos = np.zeros((5000, 4), np.float32)
    pos[:, :3] = rng.uniform(-50, 50, (5000, 3))
    for nb in (1, 7, 64000, 262144):
        key, pid = orc.hash_keys(pos, nb, (0.7, 0.7, 0.7), _cp_for((0.7,) * 3), abi.KEY_REF_XOR)
        assert key.max() < nb
        assert np.array_equal(pid, np.arange(5000, dtype=np.uint32))


def test_linear_cells_match_floor_division(orc):
    rng = np.random.default_rng(2)
    dims = (40, 30, 20)
    cs = 0.025
    n = 20000
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rng.uniform(0, dims[0] * cs, n)
    pos[:, 1] = rng.uniform(0, dims[1] * cs, n)
    pos[:, 2] = rng.uniform(0, dims[2] * cs, n)
    cp = _cp_for((cs,) * 3, dims)
    key, _ = orc.hash_keys(pos, dims[0] * dims[1] * dims[2], (cs,) * 3, cp, abi.KEY_LINEAR)
    c = np.floor(pos[:, :3].astype(np.float64) / np.float32(cs)).astype(np.int64)
    c = np.minimum(np.maximum(c, 0), np.array(dims) - 1)
    # Points within a few ulp of a cell face may round either way.
    frac = pos[:, :3].astype(np.float64) / cs
    safe = np.all(np.abs(frac - np.round(frac)) > 1e-4, axis=1)
    expect = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    assert np.array_equal(key[safe], expect[safe].astype(np.uint32))
    assert key.max() < dims[0] * dims[1] * dims[2]


def test_linear_clamps_outside_particles(orc):
    dims = (10, 10, 10)
    cp = _cp_for((0.1,) * 3, dims)
    pos = np.array([[-5, -5, -5, 1], [50, 50, 50, 1], [np.nan, 0.05, 0.05, 1], [0.05, np.inf, -np.inf, 1]],
                   np.float32)
    key, _ = orc.hash_keys(pos, 1000, (0.1,) * 3, cp, abi.KEY_LINEAR)
    assert key.tolist() == [0, 999, 0, 0 + 10 * 9]


@pytest.mark.parametrize("n,kmax", [(1, 10), (1000, 3), (4097, 1 << 16), (20000, 1 << 22), (5000, (1 << 31) + 5)])
def test_sort_matches_numpy_stable(orc, n, kmax):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, kmax, n, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    sk, sv = orc.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(sk, keys[order])
    assert np.array_equal(sv, vals[order])


def test_cell_offsets_against_unique(orc):
    rng = np.random.default_rng(3)
    n, nb = 3000, 500
    keys = np.sort(rng.integers(0, nb, n).astype(np.uint32))
    ids = rng.permutation(n).astype(np.uint32)
    pos = rng.standard_normal((n, 4)).astype(np.float32)
    vel = rng.standard_normal((n, 4)).astype(np.float32)
    end_init = np.full(nb, 12345, np.uint32)
    start, end, spos, svel = orc.cell_offsets(keys, ids, pos, vel, nb, end_init)
    u, first, counts = np.unique(keys, return_index=True, return_counts=True)
    exp_start = np.full(nb, 0xFFFFFFFF, np.uint32)
    exp_start[u] = first
    assert np.array_equal(start, exp_start)
    assert np.array_equal(end[u], (first + counts).astype(np.uint32))
    empty = np.setdiff1d(np.arange(nb), u)
    assert np.all(end[empty] == 12345)  # cellEnd is not cleared
    assert int((end[u] - start[u]).sum()) == n
    assert np.array_equal(spos, pos[ids])
    assert np.array_equal(svel, vel[ids])


def test_grid_stage_on_config0(orc):
    s = scenes.config_scene(0)
    nb = s.num_buckets
    key, pid = orc.hash_keys(s.pos, nb, tuple(s.common.cell_size), s.common, s.key_mode)
    sk, sp = orc.sort_pairs(key, pid)
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    same = sk[1:] == sk[:-1]
    assert np.all(sp[1:][same] > sp[:-1][same])  # ties keep index order
    assert np.array_equal(np.sort(sp), np.arange(s.n, dtype=np.uint32))


def test_appmain_scene_matches_reference_constants():
    s = scenes.appmain_scene()
    # AppMain.cpp:84-115: 24000 particles, all emitted active; spacing ~0.9956 world units.
    assert s.n == 24000
    assert np.all(s.pos[:, 3] == 1.0)
    assert abs(s.fluid.mass - 1.0922667e-3) < 1e-9
    assert abs(s.info["spacing"] - 0.99556) < 1e-4
    assert abs(s.common.kernel_radius - 0.0179199) < 1e-6
    assert s.common.num_buckets() == 64 ** 3
    # 41 x-values by 21 z-values per layer (x from -20 while < 20, z from -20 while < 0)
    layer = s.pos[s.pos[:, 1] == s.pos[0, 1]]
    assert len(layer) == 41 * 21

"""GPU parity of the grid stage through the C-ABI: keys, stable sort, cell
start/end and SoA reorder are bit-exact against the oracle (and numpy's stable
argsort for the sort)."""
import numpy as np
import pytest

from fluidgrid_amd import abi, scenes

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def _dev(lib, a):
    return lib.DeviceArray.from_numpy(a)


def _hash_gpu(lib, s):
    L = lib.load()
    lib.init_params(s.common, s.kernels, s.key_mode)
    dpos = _dev(lib, s.pos)
    dkey = lib.DeviceArray((s.n,), np.uint32)
    did = lib.DeviceArray((s.n,), np.uint32)
    L.gCalcHashKey(dkey.ptr, did.ptr, dpos.ptr, s.n, s.num_buckets, abi.Float3(*s.common.cell_size))
    lib.check(None, "gCalcHashKey")
    return dkey.numpy(), did.numpy()


@pytest.mark.parametrize("which", ["cfg0_linear", "appmain_ref_xor", "cfg0_ref_xor"])
def test_hash_bit_exact(fglib, orc, which):
    if which == "cfg0_linear":
        s = scenes.config_scene(0)
    elif which == "cfg0_ref_xor":
        s = scenes.config_scene(0, key_mode=abi.KEY_REF_XOR)
    else:
        s = scenes.appmain_scene()
    key, pid = _hash_gpu(fglib, s)
    okey, opid = orc.hash_keys(s.pos, s.num_buckets, tuple(s.common.cell_size), s.common, s.key_mode)
    assert np.array_equal(key, okey)
    assert np.array_equal(pid, opid)


def test_hash_outside_and_nan_positions(fglib, orc):
    s = scenes.small_dam_break()
    rng = np.random.default_rng(11)
    s.pos[:50, :3] = rng.uniform(-3, 4, (50, 3))
    s.pos[50, 0] = np.nan
    s.pos[51, 1] = np.inf
    s.pos[52, 2] = -np.inf
    for mode in (abi.KEY_LINEAR, abi.KEY_REF_XOR):
        s.key_mode = mode
        key, _ = _hash_gpu(fglib, s)
        okey, _ = orc.hash_keys(s.pos, s.num_buckets, tuple(s.common.cell_size), s.common, mode)
        assert np.array_equal(key, okey), mode


# Sizes up to 4,194,304 keys run 2048-element tiles, larger ones 4096-element tiles.
SORT_CASES = [(1, 1), (100, 8), (4095, 16), (4096, 16), (4097, 3), (65537, 22), (300_001, 25), (1_000_003, 27),
              (200_000, 32), (50_000, 9), (4_194_304, 23), (4_194_305, 23), (5_000_011, 25)]


@pytest.mark.parametrize("n,bits", SORT_CASES)
def test_sort_pairs_stable(fglib, n, bits):
    L = fglib.load()
    rng = np.random.default_rng(n * 31 + bits)
    hi = 1 << bits
    # Few distinct keys in part of the array so ties are common.
    k64 = rng.integers(0, hi, n, dtype=np.uint64)
    k64[: n // 3] = (k64[: n // 3] % 17) * np.uint64(hi // 17 if hi >= 17 else 1) % np.uint64(hi)
    keys = k64.astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    order = np.argsort(keys, kind="stable")
    dk, dv = fglib.DeviceArray.from_numpy(keys), fglib.DeviceArray.from_numpy(vals)
    fglib.check(L.fg_sort_pairs(dk.ptr, dv.ptr, n, bits), "fg_sort_pairs")
    assert np.array_equal(dk.numpy(), keys[order])
    assert np.array_equal(dv.numpy(), vals[order])
    # generic entry point (key width found on the device)
    dk.upload(keys)
    dv.upload(vals)
    L.gSort(dk.ptr, dv.ptr, n)
    fglib.check(None, "gSort")
    assert np.array_equal(dk.numpy(), keys[order])
    assert np.array_equal(dv.numpy(), vals[order])
    # values generated as 0..n-1
    dk.upload(keys)
    fglib.check(L.fg_sort_pairs_iota(dk.ptr, dv.ptr, n, bits), "iota")
    assert np.array_equal(dv.numpy(), order.astype(np.uint32))


def test_sort_all_equal_and_presorted(fglib):
    L = fglib.load()
    for keys in (np.zeros(70_000, np.uint32), np.arange(70_000, dtype=np.uint32),
                 np.arange(70_000, dtype=np.uint32)[::-1].copy()):
        n = len(keys)
        vals = np.arange(n, dtype=np.uint32)
        order = np.argsort(keys, kind="stable")
        dk, dv = fglib.DeviceArray.from_numpy(keys), fglib.DeviceArray.from_numpy(vals)
        fglib.check(L.fg_sort_pairs(dk.ptr, dv.ptr, n, 17), "sort")
        assert np.array_equal(dk.numpy(), keys[order]) and np.array_equal(dv.numpy(), vals[order])


def test_sort_with_caller_workspace(fglib):
    L = fglib.load()
    n, bits = 123_457, 20
    keys = np.random.default_rng(3).integers(0, 1 << bits, n).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    ws = fglib.DeviceArray((L.fg_sort_workspace_size(n, bits),), np.uint8)
    dk, dv = fglib.DeviceArray.from_numpy(keys), fglib.DeviceArray.from_numpy(vals)
    fglib.check(L.fg_sort_pairs_ws(dk.ptr, dv.ptr, n, bits, ws.ptr, ws.nbytes, 0), "ws")
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dv.numpy(), order.astype(np.uint32))
    # too small a workspace is an error, not a fault
    assert L.fg_sort_pairs_ws(dk.ptr, dv.ptr, n, bits, ws.ptr, 64, 0) != 0
    L.fg_last_error()


def _grid_gpu(lib, s):
    """hash -> sort -> cell offsets on the GPU; returns numpy arrays."""
    L = lib.load()
    lib.init_params(s.common, s.kernels, s.key_mode)
    nb = s.num_buckets
    dpos, dvel = lib.DeviceArray.from_numpy(s.pos), lib.DeviceArray.from_numpy(s.vel)
    dkey, did = lib.DeviceArray((s.n,), np.uint32), lib.DeviceArray((s.n,), np.uint32)
    dstart, dend = lib.DeviceArray((nb,), np.uint32), lib.DeviceArray((nb,), np.uint32)
    dend.fill_u32(0xABCDEF)
    dspos, dsvel = lib.DeviceArray((s.n, 4), np.float32), lib.DeviceArray((s.n, 4), np.float32)
    L.gCalcHashKey(dkey.ptr, did.ptr, dpos.ptr, s.n, nb, abi.Float3(*s.common.cell_size))
    L.gSort(dkey.ptr, did.ptr, s.n)
    L.gFindCellOffsetOrderingPosVel(dstart.ptr, dend.ptr, dkey.ptr, did.ptr, dpos.ptr, dspos.ptr, dvel.ptr,
                                    dsvel.ptr, s.n, nb)
    lib.check(None, "grid")
    return dict(skey=dkey.numpy(), spid=did.numpy(), start=dstart.numpy(), end=dend.numpy(), spos=dspos.numpy(),
                svel=dsvel.numpy())


@pytest.mark.parametrize("which", ["cfg0", "appmain", "one_cell", "n1", "n257"])
def test_cell_offsets_and_reorder_bit_exact(fglib, orc, which):
    if which == "cfg0":
        s = scenes.config_scene(0)
    elif which == "appmain":
        s = scenes.appmain_scene()
    elif which == "one_cell":
        s = scenes.small_dam_break(lattice=(6, 6, 6), h=0.05)
        s.pos[:, :3] = 0.5 + (s.pos[:, :3] - 0.5) * 0.05  # everything inside one cell
    elif which == "n1":
        s = scenes.small_dam_break(lattice=(1, 1, 1))
    else:
        s = scenes.small_dam_break(lattice=(257, 1, 1), h=0.002, box=(0.3, 0.01, 0.01))
    s.vel[:, :3] = np.random.default_rng(1).normal(0, 1, (s.n, 3))
    g = _grid_gpu(fglib, s)
    key, pid = orc.hash_keys(s.pos, s.num_buckets, tuple(s.common.cell_size), s.common, s.key_mode)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, spos, svel = orc.cell_offsets(sk, sp, s.pos, s.vel, s.num_buckets,
                                              np.full(s.num_buckets, 0xABCDEF, np.uint32))
    assert np.array_equal(g["skey"], sk)
    assert np.array_equal(g["spid"], sp)
    assert np.array_equal(g["start"], start)
    assert np.array_equal(g["end"], end)  # including untouched entries of empty cells
    assert np.array_equal(g["spos"], spos)
    assert np.array_equal(g["svel"], svel)


def test_grid_1m_properties(fglib, orc):
    """Config 1 size (1,048,576 particles, 150^3 buckets): keys and the stable
    permutation bit-exact against the oracle's keys and numpy's stable
    argsort; sortedness, ties by index, and cell ranges consistent."""
    s = scenes.config_scene(1)
    g = _grid_gpu(fglib, s)
    okey, _ = orc.hash_keys(s.pos, s.num_buckets, tuple(s.common.cell_size), s.common, s.key_mode)
    order = np.argsort(okey, kind="stable")
    assert np.array_equal(g["skey"], okey[order])
    assert np.array_equal(g["spid"], order.astype(np.uint32))
    sk, sp, start, end = g["skey"], g["spid"], g["start"], g["end"]
    n = s.n
    assert np.array_equal(np.sort(sp), np.arange(n, dtype=np.uint32))
    d = np.diff(sk.astype(np.int64))
    assert np.all(d >= 0)
    assert np.all(sp[1:][d == 0] > sp[:-1][d == 0])
    occ = np.nonzero(start != EMPTY)[0]
    assert int((end[occ].astype(np.int64) - start[occ]).sum()) == n
    assert np.all(sk[start[occ]] == occ)
    assert np.all(sk[end[occ] - 1] == occ)
    assert np.array_equal(g["spos"], s.pos[sp])


@pytest.mark.parametrize("which", ["cfg0", "appmain"])
@pytest.mark.parametrize("fused", [False, True])
def test_clear_occupied_then_no_fill_offsets(fglib, orc, which, fused):
    """Two grids on the same buffers: after the first, clearing the cells its
    sorted keys occupied (alone, or inside the next hash) leaves every cell
    start empty, and the second grid built without the bucket fill equals
    the oracle's (which fills)."""
    s = scenes.config_scene(0) if which == "cfg0" else scenes.appmain_scene()
    L = fglib.load()
    g0 = _grid_gpu(fglib, s)  # leaves init_params set for s
    nb = s.num_buckets
    dstart = fglib.DeviceArray.from_numpy(g0["start"])
    dkey = fglib.DeviceArray.from_numpy(g0["skey"])

    rng = np.random.default_rng(7)
    pos2 = s.pos.copy()
    pos2[:, :3] = pos2[rng.permutation(s.n), :3] + rng.normal(0, 0.3 * s.common.cell_size.x, (s.n, 3))
    vel2 = rng.normal(0, 1, (s.n, 4)).astype(np.float32)
    dpos, dvel = fglib.DeviceArray.from_numpy(pos2.astype(np.float32)), fglib.DeviceArray.from_numpy(vel2)
    did = fglib.DeviceArray.from_numpy(np.arange(s.n, dtype=np.uint32))
    dend = fglib.DeviceArray((nb,), np.uint32)
    dend.fill_u32(0xABCDEF)
    dspos, dsvel = fglib.DeviceArray((s.n, 4), np.float32), fglib.DeviceArray((s.n, 4), np.float32)
    cell = abi.Float3(*s.common.cell_size)
    if fused:
        assert L.fg_calc_hash_key_clear(dkey.ptr, dpos.ptr, s.n, nb, cell, dstart.ptr) == 0
        fglib.check(None, "hash+clear")
    else:
        assert L.fg_clear_occupied_cells(dstart.ptr, dkey.ptr, s.n, nb) == 0
        fglib.check(None, "clear")
        L.gCalcHashKey(dkey.ptr, did.ptr, dpos.ptr, s.n, nb, cell)
    assert np.all(dstart.numpy() == EMPTY)
    L.gSort(dkey.ptr, did.ptr, s.n)
    assert L.fg_cell_offsets_pos_vel(dstart.ptr, dend.ptr, dkey.ptr, did.ptr, dpos.ptr, dspos.ptr, dvel.ptr,
                                     dsvel.ptr, s.n, nb, 0, None) == 0
    fglib.check(None, "grid2")
    key, pid = orc.hash_keys(pos2.astype(np.float32), nb, tuple(s.common.cell_size), s.common, s.key_mode)
    sk, sp = orc.sort_pairs(key, pid)
    start, end, spos, svel = orc.cell_offsets(sk, sp, pos2.astype(np.float32), vel2, nb,
                                              np.full(nb, 0xABCDEF, np.uint32))
    assert np.array_equal(dkey.numpy(), sk)
    assert np.array_equal(dstart.numpy(), start)
    assert np.array_equal(dend.numpy(), end)
    assert np.array_equal(dspos.numpy(), spos)
    assert np.array_equal(dsvel.numpy(), svel)
    # null pointers are refused, not dereferenced
    assert L.fg_clear_occupied_cells(None, dkey.ptr, s.n, nb) != 0
    L.fg_last_error()
    assert L.fg_calc_hash_key_clear(dkey.ptr, dpos.ptr, s.n, nb, cell, None) != 0
    L.fg_last_error()

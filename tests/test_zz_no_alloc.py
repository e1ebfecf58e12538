"""The allocation-free sort entry point (fg_sort_pairs_noalloc) against the
entry point whose workspace the caller passes (fg_sort_pairs_ws): the same
stable permutation bit for bit, on two calls in a row (the second smaller,
so stale contents of the shared static workspace would show), with caller
values and with generated identity values, and the size limit refused before
any launch."""
import numpy as np
import pytest

from fluidgrid_amd import abi

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1048576


def _with_ws(fglib, keys, vals, bits, iota):
    L = fglib.load()
    k = fglib.DeviceArray.from_numpy(keys)
    v = fglib.DeviceArray.from_numpy(vals)
    nbytes = L.fg_sort_workspace_size(len(keys), bits)
    ws = fglib.DeviceArray((nbytes,), np.uint8)
    fglib.check(L.fg_sort_pairs_ws(k.ptr, v.ptr, len(keys), bits, ws.ptr, nbytes, int(iota)), "fg_sort_pairs_ws")
    return k.numpy(), v.numpy()


def _no_alloc(fglib, keys, vals, bits, iota):
    L = fglib.load()
    k = fglib.DeviceArray.from_numpy(keys)
    v = fglib.DeviceArray.from_numpy(vals)
    fglib.check(L.fg_sort_pairs_noalloc(k.ptr, v.ptr, len(keys), bits, int(iota)), "fg_sort_pairs_noalloc")
    return k.numpy(), v.numpy()


@pytest.mark.parametrize("iota", [False, True])
def test_noalloc_sort_matches_workspace_sort(fglib, iota):
    rng = np.random.default_rng(77 + iota)
    for n, bits in ((MAX_PAIRS, 25), (70001, 16)):  # the second call is smaller
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
        keys[: n // 3] = keys[0]  # long runs of ties
        vals = np.arange(n, dtype=np.uint32) if iota else rng.integers(0, 1 << 32, n, dtype=np.uint32)
        want = _with_ws(fglib, keys, vals, bits, iota)
        got = _no_alloc(fglib, keys, vals, bits, iota)
        assert np.array_equal(got[0], want[0]), (n, bits)
        assert np.array_equal(got[1], want[1]), (n, bits)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(got[0], keys[order]) and np.array_equal(got[1], vals[order])


def test_noalloc_sort_refuses_more_than_its_workspace(fglib):
    L = fglib.load()
    k = fglib.DeviceArray((MAX_PAIRS + 1,), np.uint32)
    v = fglib.DeviceArray((MAX_PAIRS + 1,), np.uint32)
    k.fill_u32(5)
    L.fg_last_error()
    assert L.fg_sort_pairs_noalloc(k.ptr, v.ptr, MAX_PAIRS + 1, 8, 0) == abi.ERR_INVALID_ARGUMENT
    assert b"static workspace" in L.fg_last_error_string()
    assert L.fg_last_error() == abi.ERR_INVALID_ARGUMENT
    assert np.all(k.numpy() == 5)

"""The committed build-generated fixtures (tests/golden/, made by
make_fixtures.py from the oracle): files match their SHA-256 manifest, and
the oracle still reproduces them bit for bit."""
import hashlib
import os

import numpy as np
import pytest

from fluidgrid_amd import scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"cfg0_step1": lambda: scenes.config_scene(0), "appmain_step1": scenes.appmain_scene}


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def manifest():
    out = {}
    for line in open(os.path.join(HERE, "MANIFEST.sha256")):
        h, key = line.split()
        out[key] = h
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_fixture_matches_manifest(name):
    m = manifest()
    arrays = load(name)
    assert set(f"{name}:{k}" for k in arrays) == {k for k in m if k.startswith(name + ":")}
    for k, v in arrays.items():
        assert hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() == m[f"{name}:{k}"], k


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(orc, name):
    s = CASES[name]()
    st = orc.step(s.pos, s.vel, s.fluid, s.common, s.kernels, s.key_mode)
    f = load(name)
    assert np.array_equal(st.skey, f["skey"]) and np.array_equal(st.spid, f["spid"])
    occ = np.nonzero(st.start != 0xFFFFFFFF)[0]
    assert np.array_equal(occ, f["occupied"])
    assert np.array_equal(st.start[occ], f["start"]) and np.array_equal(st.end[occ], f["end"])
    for k in ("rho_p", "vel", "pos"):
        assert np.array_equal(getattr(st, k), f[k]), k

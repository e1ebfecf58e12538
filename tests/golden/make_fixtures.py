"""Writes the build-generated fixtures in this folder.

The reference ships no golden vectors for the SPH step, and its code cannot
be built or run here, so these vectors are produced by this repository's
oracle (oracle/fg_oracle.cpp) and are labelled as such.  They freeze one
oracle step of two scenes, so that a later change to the oracle or to the
HIP path that alters a bit of the grid or a value of the density shows up:

  cfg0_step1.npz     16,384-particle dam break (LINEAR keys, 40^3 buckets)
  appmain_step1.npz  the reference application's scene (REF_XOR, 64^3 buckets)

Each file holds the sorted keys, the permutation, cellStart/cellEnd of the
occupied buckets, (rho, p) in sorted order and the velocities and positions
after the step.  MANIFEST.sha256 lists the SHA-256 of every array.

    python tests/golden/make_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fluidgrid_amd import scenes  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = {
    "cfg0_step1": lambda: scenes.config_scene(0),
    "appmain_step1": scenes.appmain_scene,
}


def fixture_arrays(scene):
    st = oracle.step(scene.pos, scene.vel, scene.fluid, scene.common, scene.kernels, scene.key_mode)
    occ = np.nonzero(st.start != 0xFFFFFFFF)[0].astype(np.uint32)
    return {
        "skey": st.skey, "spid": st.spid, "occupied": occ, "start": st.start[occ], "end": st.end[occ],
        "rho_p": st.rho_p, "vel": st.vel, "pos": st.pos,
    }


def digest(arrays):
    return {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in sorted(arrays.items())}


def main():
    lines = []
    for name, make in CASES.items():
        arrays = fixture_arrays(make())
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        for k, h in digest(arrays).items():
            lines.append(f"{h}  {name}:{k}")
    with open(os.path.join(HERE, "MANIFEST.sha256"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()

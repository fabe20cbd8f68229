"""Writes the committed golden fixtures under tests/golden/.

The reference ships no test vectors, decks or outputs, so these fixtures are
produced by the CPU oracle (itself pinned by the closed-form and brute-force
tests) on small seeded cases.  Each .npz holds the input deck (part1
records), the parameters, and the neighbour-search products and step fields
after the stated number of steps.  Regenerate with:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("urban-sph_a100_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import abi  # noqa: E402
import cases  # noqa: E402
import oracle  # noqa: E402

FIELDS = ["UX", "UY", "UZ", "PRES", "FTOTALX", "FTOTALZ", "FPX", "FPZ", "VIS_T", "SR", "FBX", "FMX",
          "INV_CM_XX", "INV_CM_XY"]


def case_defs():
    P, _ = cases.box(nf=6, walls=3)
    prm = cases.box_params(oracle.default_params(), tol=1e-12)
    yield "box6_bicgstab", P, None, prm, 2, 1e-4
    prm = cases.box_params(oracle.default_params())
    prm.update(ppe_solver=abi.PPE_REF_JACOBI)
    yield "box6_jacobi", P, None, prm, 2, 1e-4
    P, nrm, g = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(g, oracle.default_params(), tol=1e-12)
    yield "canyon_dx002", P, nrm, prm, 2, 5e-5


def main():
    for name, P, nrm, prm, steps, dt in case_defs():
        o = oracle.Oracle(prm, len(P))
        o.upload(P, None, nrm)
        for c in range(steps):
            rc = o.step(c, dt)
            assert rc == 0, rc
        keys, perm, gs, ge = o.nnps()
        out = dict(P1=P, params=np.frombuffer(bytes(prm), np.uint8), steps=steps, dt=dt, keys=keys, perm=perm,
                   cell_start=gs, cell_end=ge)
        if nrm is not None:
            out["normals"] = nrm
        for f in FIELDS:
            out["f_" + f] = o.field(f)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, len(P), "particles")


if __name__ == "__main__":
    main()

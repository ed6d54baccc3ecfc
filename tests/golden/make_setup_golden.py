"""Writes tests/golden/setup_digests.json: SHA-256 of the lumped mass and of the
seven canonical CSR triples (M, S1..S6) for small meshes, produced by the
REFERENCE's own compiled sources (oracle/_ref, built from /root/reference).
Run in the container that has /root/reference:  python tests/golden/make_setup_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

MESHES = [(4, 20, 20), (5, 10, 10), (2, 10, 10), (1, 10, 10)]


def digest(arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    names = ("M", "S1", "S2", "S3", "S4", "S5", "S6")
    out = {}
    for tc, nx, ny in MESHES:
        order = {1: 3, 2: 6, 4: 4, 5: 9}[tc]
        nn = (nx + 1) * (ny + 1) if tc in (1, 4) else (2 * nx + 1) * (2 * ny + 1)
        ne = nx * ny * (2 if tc in (1, 2) else 1)
        en, x, y = oracle.ref_mesh(tc, nx, ny, 0.0, 10.0 * nx, 0.0, 10.0 * ny, nn, ne, order)
        rho = np.random.default_rng(1).uniform(1500.0, 2500.0, nn)
        ref = oracle.ref_assemble(tc, en, x, y, rho)
        d = {"lump": digest([ref["lump"]])}
        for m in range(7):
            d[names[m]] = digest(oracle.ref_coo2csr(nn, *ref[m]))
        out[f"{tc}_{nx}_{ny}"] = d
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "setup_digests.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()

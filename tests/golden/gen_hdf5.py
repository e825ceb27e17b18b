"""Writes HDF5 fixtures in the layout of the reference's external golden
files, read by its Chapel tests:

  hdf5/construction/<model>.h5   /representatives  u64 [dimension]
      (chapel/test/TestStatesEnumeration.chpl:22-25)
  hdf5/matvec/<model>.h5         /x, /y            f64 [2][dimension]
      (chapel/test/TestMatrixVectorProduct.chpl:34,43; y = H x per row)

The reference's own files are downloaded from outside (chapel/Makefile:
132-150) and are not available here, so the expected values come from the
oracle (representatives by oracle enumeration, y by the oracle's push-form
product) on the model files in golden/models, and the files are written by
h5py, an HDF5 writer independent of the engine's: the conda interpreter
(PYTHON39, default /opt/conda/bin/python3.9) holds h5py 3.3 on libhdf5
1.10.6.  One matvec file is chunked and gzip-compressed, the rest are
contiguous, so both layouts are read by the tests.

Usage:  python tests/golden/gen_hdf5.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [REPO, os.path.join(REPO, "lattice-symmetries_amd")]

import lsamd  # noqa: E402  (model front-end only: host code, no device calls)
from oracle import oracle as O  # noqa: E402

MODELS = ["heisenberg_chain_10", "heisenberg_kagome_12", "heisenberg_kagome_12_symm"]
COMPRESSED = {"heisenberg_kagome_12"}
PYTHON39 = os.environ.get("PYTHON39", "/opt/conda/bin/python3.9")

WRITER = r"""
import sys, numpy as np, h5py
src, out, kind, compressed = sys.argv[1], sys.argv[2], sys.argv[3], sys.argv[4] == "1"
d = np.load(src)
with h5py.File(out, "w") as f:
    if kind == "construction":
        f.create_dataset("representatives", data=d["representatives"].astype("<u8"))
    else:
        kw = dict(chunks=(1, 64), compression="gzip", compression_opts=4) if compressed else {}
        f.create_dataset("x", data=d["x"].astype("<f8"), **kw)
        f.create_dataset("y", data=d["y"].astype("<f8"), **kw)
"""


def oracle_data(name: str, seed: int):
    m = lsamd.Model.from_file(os.path.join(HERE, "models", name + ".yaml"))
    n, si = m.number_sites, m.spin_inversion or 0
    if m.symmetries:
        S = lsamd.Symmetries([lsamd.Symmetry(p, s) for p, s in m.symmetries])
        og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    else:
        og = O.Group(n, np.arange(n, dtype=np.int64)[None, :], np.ones(1), np.zeros(1), si)
    lo = (1 << m.hamming_weight) - 1
    hi = lo << (n - m.hamming_weight - (1 if si else 0))
    reps = og.enumerate(lo, hi)
    diag, off = O.Terms(m.terms(-1, True)), O.Terms(m.terms(-1, False))
    x = np.random.default_rng(seed).standard_normal((2, reps.size))
    y = np.stack([O.matvec_push(og, n, si, reps, diag, off, x[k]) for k in range(2)])
    return reps, x, y


def main() -> None:
    for kind in ("construction", "matvec"):
        os.makedirs(os.path.join(HERE, "hdf5", kind), exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for i, name in enumerate(MODELS):
            reps, x, y = oracle_data(name, 100 + i)
            src = os.path.join(tmp, name + ".npz")
            np.savez(src, representatives=reps, x=x, y=y)
            for kind in ("construction", "matvec"):
                out = os.path.join(HERE, "hdf5", kind, name + ".h5")
                subprocess.run([PYTHON39, "-c", WRITER, src, out, kind, "1" if name in COMPRESSED else "0"],
                               check=True)
            print(f"{name}: {reps.size} representatives")


if __name__ == "__main__":
    main()

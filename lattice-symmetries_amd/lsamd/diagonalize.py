"""Exact diagonalisation of a model file with HDF5 output: the flow of the
reference's chapel/src/Diagonalize.chpl (main at :258-330, makeBasisStates
:226-244, saveEigenvectors :246-256), on one GPU.

The output file has the reference's layout:
  basis/representatives      u64 [dimension], sorted
  hamiltonian/eigenvectors   f64 [num_evals][dimension], row e for eigenvalue e
  hamiltonian/eigenvalues    f64 [num_evals], ascending
  hamiltonian/residuals      f64 [num_evals], || H v - lambda v ||
and the groups basis, hamiltonian and observables.  When the file already
holds basis/representatives they are read instead of enumerated
(Diagonalize.chpl:227-233), which is how a run resumes on a stored basis.

The reference solves with PRIMME (out of scope here); this driver uses the
engine's thick-restart block Lanczos (ls_amd_block_eigenvectors) on the
cached CSR.  Its parameters map as follows: eps -> relative residual
tolerance, max_block_size -> block size (raised to the next of 1, 2, 4, 8
that holds num_evals), max_basis_size -> Krylov blocks per restart cycle,
min_restart_size -> blocks kept at a restart.
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from . import DeviceArray, EngineError
from . import hdf5
from .model import Model

BLOCK_SIZES = (1, 2, 4, 8)


def _check_states(basis, states: np.ndarray, output: str) -> None:
    """A stored representative list must be strictly increasing, hold states
    of the basis's Hamming weight (spin bases) and, where the basis has
    permutation symmetries, only orbit representatives (checked on the GPU)."""
    def bad(i, why):
        raise EngineError(f"'{output}': basis/representatives[{i}] = {int(states[i])} {why}")
    if states.size > 1:
        up = states[1:] > states[:-1]
        if not np.all(up):
            bad(int(np.argmin(up)) + 1, "breaks the strictly increasing order")
    if states.size == 0:
        return
    if basis.number_bits < 64 and np.any(states >> np.uint64(basis.number_bits)):
        bad(int(np.argmax(states >> np.uint64(basis.number_bits) != 0)), f"has bits above {basis.number_bits}")
    h = getattr(basis, "hamming_weight", None)
    if h is not None and getattr(basis, "_bits", basis.number_spins) == basis.number_spins:
        pc = np.zeros(states.size, dtype=np.int64)
        for k in range(8):
            pc += np.array([bin(v).count("1") for v in range(256)], dtype=np.int64)[
                ((states >> np.uint64(8 * k)) & np.uint64(255)).astype(np.int64)]
        if np.any(pc != h):
            bad(int(np.argmax(pc != h)), f"does not have Hamming weight {h}")
    try:
        flags, _ = basis.is_representative(states)
    except EngineError as e:
        if "no permutation symmetries" in str(e):
            return
        raise
    if not np.all(flags == 1):
        bad(int(np.argmin(flags)), "is not a representative of this basis")


def make_basis_states(basis, output: str, check: bool = True) -> bool:
    """makeBasisStates (Diagonalize.chpl:226-244): reads basis/representatives
    from `output` if present, else enumerates and writes them.  Returns True
    when the stored list was used.  With check=True a stored list must be
    strictly increasing and every entry a representative of this basis (the
    reference adopts it unchecked)."""
    if hdf5.does_object_exist(output, "basis/representatives"):
        states = hdf5.read_dataset_as_blocks(output, "basis/representatives", rank=1, dtype=np.uint64)
        if check:
            _check_states(basis, states, output)
        basis.set_states(states)
        return True
    basis.build()
    hdf5.write_dataset_as_blocks(output, "basis/representatives", basis.states)
    return False


def diagonalize(input: str, output: str = "exact_diagonalization_output.h5", num_evals: int = 1,
                eps: float = 1e-6, max_basis_size: int = 0, min_restart_size: int = 0, max_block_size: int = 1,
                mode: str = "fp32", max_cycles: int = 80, seed: int = 1, check_representatives: bool = True) -> dict:
    """Lowest num_evals eigenpairs of the model file's Hamiltonian, written to
    `output` as Diagonalize.chpl does (config names input, kOutput, numEvals,
    kEps, kMaxBasisSize, kMinRestartSize, kMaxBlockSize at :164-172)."""
    if num_evals < 1:
        raise ValueError(f"invalid numEvals: {num_evals}")
    t0 = time.perf_counter()
    model = Model.from_file(input)
    if not model.has_hamiltonian:
        raise EngineError(f"'{input}' has no hamiltonian")
    hdf5.make_file(output)
    for g in ("basis", "hamiltonian", "observables"):
        hdf5.make_group(output, g)

    basis = model.device_basis()
    reused = make_basis_states(basis, output, check_representatives)
    n = basis.number_states
    t_basis = time.perf_counter()

    op = model.device_operator(basis)
    op.build_csr(96 if mode == "fp32" else 64)
    want = max(num_evals, max_block_size)
    block = next((b for b in BLOCK_SIZES if b >= want), None)
    if block is None:
        raise EngineError(f"numEvals and kMaxBlockSize must be at most {BLOCK_SIZES[-1]} (got {want})")
    cap = n // block - 1  # (max_blocks + 1) * block Krylov vectors must fit
    max_blocks = -(-max_basis_size // block) if max_basis_size > 0 else max(8, 48 // block)
    max_blocks = min(max_blocks, cap)
    if max_blocks < 4:
        raise EngineError(f"dimension {n} is too small for {num_evals} eigenvalues with block size {block}")
    keep = max(1, min_restart_size // block) if min_restart_size > 0 else 2
    keep = min(keep, max_blocks - 2)

    vecs = DeviceArray(num_evals * n * 8)
    r = op.block_eigensolve(mode, block_size=block, nev=num_evals, max_blocks=max_blocks, keep_blocks=keep,
                            max_cycles=max_cycles, tol=eps, switch_tol=max(eps, 1e-5), seed=seed,
                            eigenvectors=vecs)
    evals = np.asarray(r["eigenvalues"], dtype=np.float64)
    residuals = np.asarray(r["residuals"], dtype=np.float64) * np.maximum(1.0, np.abs(evals))
    t_solve = time.perf_counter()

    hdf5.write_device_dataset(output, "hamiltonian/eigenvectors", vecs, np.float64, (num_evals, n))
    hdf5.write_dataset(output, "hamiltonian/eigenvalues", evals)
    hdf5.write_dataset(output, "hamiltonian/residuals", residuals)
    t_end = time.perf_counter()
    return {
        "dimension": n,
        "eigenvalues": evals,
        "residuals": residuals,
        "converged": r["converged"],
        "cycles": r["cycles"],
        "representatives_reused": reused,
        "block_size": block,
        "max_blocks": max_blocks,
        "seconds": {"basis": t_basis - t0, "solve": t_solve - t_basis, "write": t_end - t_solve,
                    "total": t_end - t0},
    }


def main(argv: Optional[list] = None) -> int:
    """Command line in the reference's config style:
    python -m lsamd.diagonalize --input model.yaml --kOutput out.h5 --numEvals 2"""
    import argparse
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--input", default="data/heisenberg_chain_10.yaml")
    p.add_argument("--kOutput", default="exact_diagonalization_output.h5")
    p.add_argument("--numEvals", type=int, default=1)
    p.add_argument("--kEps", type=float, default=1e-6)
    p.add_argument("--kMaxBasisSize", type=int, default=0)
    p.add_argument("--kMinRestartSize", type=int, default=0)
    p.add_argument("--kMaxBlockSize", type=int, default=1)
    a = p.parse_args(argv)
    r = diagonalize(a.input, a.kOutput, a.numEvals, a.kEps, a.kMaxBasisSize, a.kMinRestartSize, a.kMaxBlockSize)
    print("Obtained eigenvalues: ", r["eigenvalues"])
    print("Residual norms:       ", r["residuals"])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

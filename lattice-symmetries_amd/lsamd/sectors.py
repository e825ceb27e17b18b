"""Many symmetry sectors of one Hamiltonian as a single device operator.

A chain of n sites with translations has n momenta k; with spin inversion
each splits into two sectors (eigenvalue +1 / -1).  Every sector gets its own
representatives (built on the device) and its own complex CSR cache; the
caches are then concatenated into one block-diagonal operator
(``DeviceSectorBatch``) whose product runs as one launch over all sectors.
"""
from typing import List, Optional, Sequence, Tuple

from . import DeviceBasis, DeviceOperator, DeviceSectorBatch, Symmetries, Symmetry
from . import models
from .expr import Expr

# Stabiliser sums at or below this are treated as exact zeros: with complex
# characters the reference's floating-point sum of a vanishing stabiliser
# character can be +-1e-16 instead of 0, which would admit states with a
# meaningless norm.
DEFAULT_NORM_TOLERANCE = 1e-9


def chain_sectors(n: int, spin_inversion: bool = True) -> List[Tuple[int, Optional[int]]]:
    """(momentum, spin inversion) pairs in the order k = 0..n-1, then +1, -1."""
    out = []
    for k in range(n):
        if spin_inversion:
            out.append((k, 1))
            out.append((k, -1))
        else:
            out.append((k, None))
    return out


def chain_sector_operators(n: int, hamming_weight: int, expr: Expr,
                           sectors: Sequence[Tuple[int, Optional[int]]],
                           norm_tolerance: float = DEFAULT_NORM_TOLERANCE):
    """Device bases and operators (complex CSR built) of the given chain sectors."""
    bases, ops = [], []
    for k, si in sectors:
        S = Symmetries([Symmetry(models.chain_translation(n), k)])
        b = DeviceBasis(n, hamming_weight, si, S)
        if norm_tolerance > 0:
            b.set_norm_tolerance(norm_tolerance)
        b.build()
        op = DeviceOperator(b, expr)
        op.build_csr(128)
        bases.append(b)
        ops.append(op)
    return bases, ops


def chain_sector_batch(n: int, hamming_weight: int, expr: Expr, spin_inversion: bool = True,
                       norm_tolerance: float = DEFAULT_NORM_TOLERANCE, keep_operators: bool = False):
    """All momentum (x spin inversion) sectors of a chain as one DeviceSectorBatch.

    The batch holds copies of the sector matrices, so the per-sector bases and
    operators are released unless keep_operators is set; with it, returns
    (batch, bases, ops)."""
    sectors = chain_sectors(n, spin_inversion)
    bases, ops = chain_sector_operators(n, hamming_weight, expr, sectors, norm_tolerance)
    batch = DeviceSectorBatch(ops)
    batch.sectors = sectors
    if keep_operators:
        return batch, bases, ops
    return batch

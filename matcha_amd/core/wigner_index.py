"""Index algebra of the Wigner-D layout used by sigma and D.

Layout ``[(l, m', m) for l for m' in -l..l for m in -l..l]``, as in the
reference's ``WignerDsize``/``WignerDindex`` (``src/matcha/core/
WignerDMatrices.py:48-97,969-1012``) and ``setup_utils.compute_size``
(``src/matcha/utils/setup_utils.py:35-37``).
"""
from __future__ import annotations


def Dsize(ell_max: int) -> int:
    """Entries with 0 <= l <= ell_max."""
    if ell_max < 0:
        return 0
    return (ell_max + 1) * (2 * ell_max + 1) * (2 * ell_max + 3) // 3


def Dindex(ell: int, mp: int, m: int) -> int:
    return Dsize(ell - 1) + (mp + ell) * (2 * ell + 1) + (m + ell)


def compute_size(n: int) -> int:
    """Entries with l < n (``compute_size`` of the reference)."""
    return Dsize(n - 1)

"""cuTFHE.utils (reference cuTFHE/utils.py:5-8)."""
from __future__ import annotations

import numpy as np


def all_binary_vectors(m: int) -> np.ndarray:
    """All 2^m binary vectors, meshgrid 'ij' order: row p is p in binary,
    most significant bit first (cuTFHE/utils.py:5-8)."""
    grids = np.meshgrid(*([np.array([0, 1], dtype=np.int32)] * m), indexing="ij")
    return np.stack(grids, axis=-1).reshape(-1, m)

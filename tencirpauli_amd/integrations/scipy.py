"""SciPy ``LinearOperator`` over a GPU MVP plan (py/integrations/scipy.py:13-47).

The eigensolver caller of the MVP: ``eigsh`` sees a matrix-free operator
whose matvec is one GPU apply on host complex128 vectors.
"""

from __future__ import annotations

from typing import Any, Optional

import numpy as np
from scipy.sparse.linalg import LinearOperator

from .._plan import DEFAULT_MAX_BYTES, validate_max_bytes


def to_scipy_linear_operator(plan: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> LinearOperator:
    """Wrap one immutable plan without materialising its matrix."""
    validate_max_bytes(max_bytes)
    dimension = int(plan.dimension)

    def matvec(vector: Any) -> np.ndarray:
        values = np.asarray(vector, dtype=np.complex128)
        if values.ndim == 2:
            if values.shape != (dimension, 1):
                raise ValueError(
                    f"LinearOperator vector must have shape ({dimension},) or ({dimension}, 1), got {values.shape}")
            values = values[:, 0]
        elif values.ndim != 1 or values.shape != (dimension,):
            raise ValueError(
                f"LinearOperator vector must have shape ({dimension},) or ({dimension}, 1), got {values.shape}")
        return np.asarray(plan.apply(np.ascontiguousarray(values), max_bytes=max_bytes), dtype=np.complex128)

    return LinearOperator(shape=(dimension, dimension), matvec=matvec, dtype=np.complex128)

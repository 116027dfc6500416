"""Gauss-Legendre rule on [0, 1] (src/BEM_functions/utils/functions.py:5-85).

The reference lists 16-digit node/weight literals for N in {1..8, 12}; the
same rule is produced here from numpy's Legendre roots, mapped with
(p + 1) / 2 and w / 2, with the same supported set and error."""
import numpy as np

_SUPPORTED = (1, 2, 3, 4, 5, 6, 7, 8, 12)


def get_gaussion_integration_points_and_weights(N: int = 2, np_type=np.float32):
    if N not in _SUPPORTED:
        raise NotImplementedError("Please refer N as 2~8 for gaussion interation")
    p, w = np.polynomial.legendre.leggauss(N)
    return [((p + 1.0) / 2.0).astype(np_type), (w / 2.0).astype(np_type)]

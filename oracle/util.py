"""Oracle (test infrastructure): random complex normal draws, reference pytenet/util.py:10-17."""

import numpy as np


def crandn(size=None, rng=None):
    """Standard complex normal samples: all real parts are drawn first, then all
    imaginary parts, and the sum is divided by sqrt(2) (pytenet/util.py:10-17).
    The draw order fixes seeded initial states bit for bit."""
    if rng is None:
        rng = np.random.default_rng()
    re = rng.normal(size=size)
    im = rng.normal(size=size)
    return (re + 1j * im) / np.sqrt(2)

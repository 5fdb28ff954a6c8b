"""
Generic utilities (reference pytenet/util.py:10-17).

Random initial states are drawn on the host with NumPy's Generator so that a
seed reproduces the reference's states bit for bit; they are uploaded to the
GPU once.
"""

import numpy as np

__all__ = ["crandn"]


def crandn(size=None, rng: np.random.Generator = None):
    """Standard complex normal samples (util.py:10-17): all real parts are drawn
    first, then all imaginary parts; the sum is divided by sqrt(2)."""
    if rng is None:
        rng = np.random.default_rng()
    re = rng.normal(size=size)
    im = rng.normal(size=size)
    return (re + 1j * im) / np.sqrt(2)

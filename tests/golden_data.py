"""Loader of the committed golden fixtures (tests/golden/*.npz)."""
import hashlib
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name="plane256"):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def dem_digest(dem):
    return hashlib.sha256(np.ascontiguousarray(dem, dtype=np.float64).tobytes()).hexdigest()

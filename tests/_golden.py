"""Loading of the committed golden fixtures (tests/golden/)."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixtures():
    with open(os.path.join(HERE, "reference_fixtures.json")) as f:
        return json.load(f)


def arrays():
    return np.load(os.path.join(HERE, "reference_arrays.npz"), allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

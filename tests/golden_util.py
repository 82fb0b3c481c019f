"""Loading of the golden fixtures (data written by tests/golden/make_golden.py)."""
import json
import os

import numpy as np

from overheatingproblem_amd.config import config_from_dict

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)


def cfg(data, key):
    return config_from_dict(json.loads(str(data[key])))


def keys_with_suffix(data, suffix):
    return sorted(k[: -len(suffix)] for k in data.files if k.endswith(suffix))


def same(a, b):
    """Bit-level equality that treats +0 == -0 and NaN == NaN (same positions)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.array_equal(a[~na], b[~nb]))

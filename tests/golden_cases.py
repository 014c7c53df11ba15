"""Shared loader for the golden fixtures in tests/golden/."""
import glob
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case_files():
    return sorted(glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load(path):
    g = dict(np.load(path))
    meta = json.loads(str(g.pop("meta_json")))
    return g, meta


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

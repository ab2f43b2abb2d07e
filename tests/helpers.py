"""Shared helpers for the parity tests: fixture loading and error norms."""
from __future__ import annotations

import json
import os

import numpy as np
import yaml

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MESHES = os.path.join(GOLDEN, "meshes")


def unit_golds() -> dict:
    with open(os.path.join(GOLDEN, "unit_golds.json")) as f:
        return json.load(f)


_golds = None


def fom_golds():
    global _golds
    if _golds is None:
        _golds = np.load(os.path.join(GOLDEN, "fom_golds.npz"))
    return _golds


def load_input(test: str) -> dict:
    with open(os.path.join(GOLDEN, "inputs", f"{test}.yaml")) as f:
        return yaml.safe_load(f)


def as_list(v):
    return list(v) if isinstance(v, (list, tuple)) else [v]


def realizations(sig: dict):
    """(depth, angle, period, delay) in the reference's nested order
    depth -> period -> angle -> delay (fom_problem_rank_two.hpp:141-162)."""
    out = []
    for d in as_list(sig["depth"]):
        for p in as_list(sig["period"]):
            for a in as_list(sig.get("angle", 0.0)):
                for dl in as_list(sig["delay"]):
                    out.append((float(d), float(a), float(p), float(dl)))
    return out


def rel_l2(x, ref) -> float:
    """Norm-wise relative error ||x - ref||_2 / ||ref||_2 (0 when both vanish)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    n = np.linalg.norm(ref.ravel())
    d = np.linalg.norm((x - ref).ravel())
    return 0.0 if n == 0 and d == 0 else (d / n if n > 0 else np.inf)


def rel_max(x, ref) -> float:
    """max|x - ref| / max|ref|."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    n = np.abs(ref).max() if ref.size else 0.0
    d = np.abs(x - ref).max() if ref.size else 0.0
    return 0.0 if d == 0 else (d / n if n > 0 else np.inf)

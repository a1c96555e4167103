"""Shared test helpers.  Tests marked ``gpu`` need an MI355X (gfx950); the rest
run on the CPU (oracle pinning, host code, ABI surface)."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden_meta() -> dict:
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


def golden_graph(name: str):
    z = np.load(os.path.join(GOLDEN, "graphs", f"{name}.npz"))
    return int(z["n"]), z["u"], z["v"]


def golden_result(tag: str):
    return np.load(os.path.join(GOLDEN, f"{tag}.npz"))


def survivors_of(off, nb, status):
    """Canonical surviving edges (u < v), the reference's dump format
    (Decomposition.h:419-452)."""
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    keep = (np.asarray(status)[: len(nb)] == ord("Y")) & (nb > src)
    return np.stack([src[keep], nb[keep]], 1).astype(np.uint32)


GOLDEN_TAGS = sorted(golden_meta().keys())

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU-only oracle runs")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def assert_fields_match(label, gpu, ora, fields=("ucont", "ucat", "p"), tol=1e-10):
    """Relative L2 of each field, GPU against the oracle; prints the
    achieved errors (pytest -s shows them) and asserts them below tol."""
    errs = {f: rel_l2(gpu.get(f), ora.get(f)) for f in fields}
    print(f"parity {label}: " + " ".join(f"{f} {e:.2e}" for f, e in errs.items()), flush=True)
    bad = {f: e for f, e in errs.items() if not e <= tol}
    assert not bad, (label, bad, tol)
    return errs

"""Shared pytest configuration: the `gpu` marker and fixture loading."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long-running test")


def load_golden(name: str) -> dict:
    from bwpm_amd.results import load_json

    return load_json((GOLDEN / f"{name}.result.json").read_text("utf-8"))[1]


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """Relative L2 error ‖a − b‖ / ‖b‖."""
    nb = float(np.linalg.norm(np.asarray(b).ravel()))
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel())) / (nb if nb > 0 else 1.0)

"""Host side of serving (no GPU needed): the served-mix generator is reproducible and stays inside
its size bounds, and the queue fails loudly where no device is present."""

from __future__ import annotations

import numpy as np
import pytest

from bwpm_amd import _core, serve


def test_served_requests_are_reproducible() -> None:
    a = serve.served_request(11, max_n=128, max_nz=240, born_n=32)
    b = serve.served_request(11, max_n=128, max_nz=240, born_n=32)
    assert a.name == b.name and a.iterations == b.iterations
    assert np.array_equal(a.label_volume(), b.label_volume()) and np.array_equal(a.field, b.field)
    assert a.fluctuation == b.fluctuation if not isinstance(a.fluctuation, np.ndarray) else np.array_equal(
        a.fluctuation, b.fluctuation)


def test_served_mix_bounds_and_kinds() -> None:
    kinds = set()
    for seed in range(60):
        p = serve.served_request(seed, max_n=192, max_nz=300, born_n=32)
        kinds.add(p.name.split("_")[1])
        lab = p.label_volume()
        assert lab.shape == (p.nz, p.ny, p.nx)
        assert p.field.shape == (p.ny, p.nx, 2)
        assert 1 <= p.iterations <= 3
        if p.nx == 1:
            assert p.ny in serve.SIZES_2D and 200 <= p.nz <= 1024
        elif p.fluctuation is not None:
            assert p.ny == 32 and p.nz <= 128
        else:
            assert p.ny in serve.SIZES_3D and p.ny <= 192 and 200 <= p.nz <= 300
        assert int(lab.max()) < len(p.indices)
    assert kinds == {"2d", "planar", "tilted", "silver", "born"}


def test_silver_layer_geometry() -> None:
    for geometry in (True, False):
        p = serve.planar_3d(32, 40, 1, 1.0, 1.0, serve.SILVER, layer=(10, 14), geometry=geometry)
        assert p.indices[1] == serve.SILVER
        lab = p.label_volume()
        assert np.all(lab[10:14] == 1) and np.all(lab[:10] == 0) and np.all(lab[14:] == 0)


@pytest.mark.skipif(_core.gpu_device_count() > 0, reason="checks the behaviour without a device")
def test_queue_needs_a_device() -> None:
    with pytest.raises(RuntimeError, match="no GPU device"):
        serve.Server()

"""Label volumes made inside the library from a geometry (bwpm_geometry_labels; host code, no GPU).

The rule is the reference's pixel-centre rule (src/bwpm/geometry.py:24-35): slab j, pixel (n, m)
takes the label of the body containing its centre x = (m − nx//2)·dx, y = (n − ny//2)·dy,
z = z0 + (j + ½)·dz. Checked bit for bit against the NumPy constructions of the same volumes: the
tilted interface of the 3D configuration and the fibre-facet volume of experiments/b3_scaling.py:
57-72 (cap layer, then core and cladding split by a radius).
"""

from __future__ import annotations

import numpy as np
import pytest

from bwpm_amd import _core, workloads
from bwpm_amd.specs import Above, Cylinder, Geometry, Layer


@pytest.mark.parametrize(("n", "nz", "tilt"), [(64, 40, 10.0), (96, 33, -7.5), (37, 12, 30.0)])
def test_tilted_interface_equals_the_numpy_volume(n: int, nz: int, tilt: float) -> None:
    a = workloads.tilted_3d(n=n, nz=nz, tilt_deg=tilt)
    g = workloads.tilted_3d(n=n, nz=nz, tilt_deg=tilt, geometry=True)
    got = _core.geometry_labels(g.labels, n, n, g.dy, g.dx, g.dz, 2)
    assert got.dtype == np.int32 and np.array_equal(got, a.labels)


def test_fibre_facet_volume_equals_the_numpy_construction() -> None:
    # b3_scaling.volume: z_in = front − 20 µm, cap (label 1) for −D < z ≤ 0, then core (2) where
    # r ≤ 1 µm and cladding (3) for z > 0; labels 0 (air) elsewhere
    ny, nx, d = 48, 40, 0.2e-3
    D = 2e-3
    z_in = -D - 20e-3 * 0.1
    nz = 60
    y = (np.arange(ny) - ny // 2) * d
    x = (np.arange(nx) - nx // 2) * d
    z = z_in + (np.arange(nz) + 0.5) * d
    rr = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)
    want = np.zeros((nz, ny, nx), dtype=np.int32)
    for j, zj in enumerate(z):
        if -D < zj <= 0:
            want[j] = 1
        elif zj > 0:
            want[j] = np.where(rr <= 1e-3, 2, 3)
    geo = Geometry(nz, (Layer(1, -D, 0.0), Layer(3, 0.0, np.inf), Cylinder(2, 1e-3, 0.0, np.inf)), 0, z_in)
    assert np.array_equal(_core.geometry_labels(geo, ny, nx, d, d, d, 4), want)
    with pytest.raises(ValueError, match="NaN"):
        _core.geometry_labels(Geometry(nz, (Layer(1, np.nan, 0.0),)), ny, nx, d, d, d, 4)


def test_later_primitives_win_and_bad_labels_are_refused() -> None:
    g = Geometry(4, (Above(1, -1.0), Above(2, 0.5e-3)), 0, 0.0)
    lab = _core.geometry_labels(g, 4, 4, 1e-3, 1e-3, 1e-3, 3)
    assert np.all(lab[0] == 1) and np.all(lab[1:] == 2)
    with pytest.raises(ValueError, match="index"):
        _core.geometry_labels(Geometry(2, (Above(5, 0.0),)), 4, 4, 1e-3, 1e-3, 1e-3, 3)
    with pytest.raises(ValueError, match="index"):
        _core.geometry_labels(Geometry(2, (), background=3), 4, 4, 1e-3, 1e-3, 1e-3, 3)

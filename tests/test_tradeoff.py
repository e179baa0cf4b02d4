"""Roughness (Regularization.Roughness, Observations.cs:488-516) against closed forms: a constant
field has none; a field linear along one axis has a zero Laplacian inside and -w, +w on the two faces
normal to that axis, so its roughness is w sqrt(2 Ny Nz); node-unit axis weights are (1, 1, w_v)."""
import math

import numpy as np

from tomostar_amd import tradeoff as T
from tomostar_amd.grid import GridDefinition


def _grid(nx=7, ny=5, nz=4):
    return GridDefinition(13.0, 13.3, 43.0, 43.2, 0.0, 6.0, nx, ny, nz)


def test_constant_field_is_smooth():
    g = _grid()
    assert T.roughness(g, np.full(g.count, 3.7), 0.1) < 1e-12


def test_linear_field_along_each_axis():
    g = _grid()
    k, j, i = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    wx, wy, wz = T.axis_weights(g, 0.1, False)
    for field, w, faces in ((i, wx, g.ny * g.nz), (j, wy, g.nx * g.nz), (k, wz, g.nx * g.ny)):
        r = T.roughness(g, field.ravel().astype(float), 0.1)
        assert math.isclose(r, w * math.sqrt(2 * faces), rel_tol=1e-12)


def test_axis_weights_in_nodes():
    assert T.axis_weights(_grid(), 0.25, True) == (1.0, 1.0, 0.25)
    wx, wy, wz = T.axis_weights(_grid(), 0.25, False)
    assert math.isclose(wx * wy, 1.0, rel_tol=1e-12)  # h^2 / (hx hy) with h the geometric mean

"""Cartesian sensor ('linear' interpolation) for the binary-mask contract (C1).

Pins of frontend.CartesianSensor that do not depend on the diagonal rule:
barycentric weights are non-negative, sum to one, sit on the enclosing cell's
corners, reproduce fields linear in x and y exactly, and give weight one to a
grid node.  A planted mistake (a transposed coordinate) is caught.  MATLAB's
own diagonal choice on the degenerate grid triangulation is not reproducible
without MATLAB: parity of C1's interpolation step with MATLAB is unpinned.
"""

import numpy as np
import pytest

import kwave_amd as kw

fe = kw.frontend
geo = kw.geometry


def _grid_coords(kgrid):
    x = (np.arange(kgrid.Nx) - kgrid.Nx // 2) * kgrid.dx
    y = (np.arange(kgrid.Ny) - kgrid.Ny // 2) * kgrid.dy
    return np.meshgrid(x, y, indexing="ij")


def _mask_values(sensor, field):
    """Field values at the mask points in find() (column-major) order."""
    return field.reshape(-1, order="F")[np.flatnonzero(sensor.mask.reshape(-1, order="F"))]


@pytest.mark.parametrize("N", [(128, 128), (64, 96), (65, 47)])
def test_weights_are_barycentric_and_exact_for_linear_fields(N):
    rng = np.random.default_rng(7)
    kgrid = fe.KWaveGrid(N[0], 1e-4, N[1], 1.3e-4)
    X, Y = _grid_coords(kgrid)
    pts = np.vstack([rng.uniform(X.min(), X.max(), 200), rng.uniform(Y.min(), Y.max(), 200)])
    s = fe.CartesianSensor(kgrid, pts)
    assert np.all(s.bc >= -1e-15) and np.allclose(s.bc.sum(1), 1.0, atol=1e-14)
    field = 3.0 + 2e3 * X - 7e2 * Y
    got = s.apply(_mask_values(s, field)[None])[0]
    assert np.allclose(got, 3.0 + 2e3 * pts[0] - 7e2 * pts[1], rtol=0, atol=1e-12)
    # vertices are corners of the enclosing cell
    lin = np.flatnonzero(s.mask.reshape(-1, order="F"))[s.tri]
    vx, vy = lin % kgrid.Nx, lin // kgrid.Nx
    gx = pts[0] / kgrid.dx + kgrid.Nx // 2
    gy = pts[1] / kgrid.dy + kgrid.Ny // 2
    assert np.all(np.abs(vx - gx[:, None]) <= 1 + 1e-9) and np.all(np.abs(vy - gy[:, None]) <= 1 + 1e-9)


def test_grid_node_gets_unit_weight():
    kgrid = fe.KWaveGrid(64, 1e-4, 64, 1e-4)
    X, Y = _grid_coords(kgrid)
    pts = np.array([[X[10, 20], X[33, 5]], [Y[10, 20], Y[33, 5]]])
    s = fe.CartesianSensor(kgrid, pts)
    field = np.random.default_rng(1).standard_normal(X.shape)
    got = s.apply(_mask_values(s, field)[None])[0]
    assert np.allclose(got, [field[10, 20], field[33, 5]], atol=1e-12)


def test_transposed_coordinates_are_caught():
    """Planted mistake: swapping x and y of the points breaks exactness on x-y asymmetric fields."""
    kgrid = fe.KWaveGrid(64, 1e-4, 64, 1e-4)
    X, Y = _grid_coords(kgrid)
    pts = geo.make_cart_circle(2e-3, 20)
    s = fe.CartesianSensor(kgrid, pts[::-1])
    field = 1e3 * X + 3e3 * Y
    got = s.apply(_mask_values(s, field)[None])[0]
    assert not np.allclose(got, 1e3 * pts[0] + 3e3 * pts[1], atol=1e-6)


def test_points_outside_grid_rejected():
    kgrid = fe.KWaveGrid(32, 1e-4, 32, 1e-4)
    with pytest.raises(ValueError):
        fe.CartesianSensor(kgrid, np.array([[5e-3], [0.0]]))


def test_c1_oracle_matches_closed_form_without_pml(oracle):
    """C1 with PMLAlpha = 0 against the exact homogeneous lossless solution, both interpolated."""
    f, s = kw.configs.c1(Nt=120)
    f["pml_x_alpha"] = np.zeros_like(f["pml_x_alpha"])
    f["pml_y_alpha"] = np.zeros_like(f["pml_y_alpha"])
    val = lambda k: float(np.asarray(f[k]).reshape(-1)[0])  # noqa: E731
    p = oracle.run(f, record=("p",))["p"]
    ex = oracle.exact_lossless_homogeneous(np.asarray(f["p0_source_input"], dtype=np.float64), val("c0"),
                                           val("dt"), (val("dx"), val("dy"), 1.0), 120)
    idx = np.asarray(f["sensor_mask_index"]).reshape(-1).astype(np.int64) - 1
    ex_mask = ex.reshape(120, -1)[:, idx]
    a, b = s.apply(p), s.apply(ex_mask)
    assert a.shape == (120, 50)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-12

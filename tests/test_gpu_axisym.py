"""Axisymmetric solver on the GPU against physics that does not go through the oracle.

testing/unit/kspaceFirstOrderAS_piston_analytical.m: the steady-state on-axis
amplitude of a CW piston must follow p(x) = P0 |2 sin(k (sqrt(x^2 + a^2) - x) / 2)|
within 1 % (L-inf, relative to the peak).  And the axisymmetric loop against the
3-D loop on a cylindrically symmetric field.
"""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def test_piston_on_axis_amplitude(kw):
    f, start, error_pct = cases.piston_axisym()
    print("[gpu axisym] piston 320x100 Nt=1920", flush=True)
    p = kw.run(f, record=("p",), record_start_index=start)["p"]
    err = error_pct(p)
    print(f"    L-inf error {err:.3f} %", flush=True)
    assert err < 1.0


def test_axisymmetric_matches_3d_on_gpu(kw):
    """The axisymmetric loop against the 3-D loop, both on the GPU, for a Gaussian
    on the axis (homogeneous, lossless, before the wave reaches a PML): on the
    axis and along a radius the two agree to the discretisation difference the
    FP64 oracle shows for the same pair (1.3e-3)."""
    fe = cases.fe
    print("[gpu axisym] axisymmetric vs 3-D, 64x32 / 64^3", flush=True)
    a = kw.run(cases.case_axisym(Nx=64, Ny=32, Nt=50), record=("p",))["p"]
    N, dx, c = 64, 1e-4, 32
    kg = fe.KWaveGrid(N, dx, N, dx, N, dx)
    kg.set_time(50, 0.3 * dx / 1500)
    g = (np.arange(N) - c) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p0 = 2.0 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * (2.5 * dx) ** 2))
    m = np.zeros((N, N, N))
    m[:, c, c] = 1
    m[c, c:, c] = 1
    f3 = fe.save_to_disk_dict(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": m},
                              pml_size=8, smooth_p0=False)
    b = kw.run(f3, record=("p",))["p"]
    e = cases.rel_l2(a, b)
    print(f"    rel L2 axisymmetric vs 3-D {e:.2e}", flush=True)
    assert e < 3e-3

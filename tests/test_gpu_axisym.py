"""Axisymmetric solver on the GPU against closed-form physics.

testing/unit/kspaceFirstOrderAS_piston_analytical.m: the steady-state on-axis
amplitude of a CW piston must follow p(x) = P0 |2 sin(k (sqrt(x^2 + a^2) - x) / 2)|
within 1 % (L-inf, relative to the peak).
"""

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

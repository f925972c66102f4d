"""The oracle's pins are discriminating: plant a plausible mistake in the
oracle (CPU only, via monkeypatch) and the pin that guards it must fail."""

import numpy as np
import pytest

import test_oracle as pins


def _fails(fn, *args):
    with pytest.raises(AssertionError):
        fn(*args)


def test_dropped_kappa_breaks_closed_form(kw, oracle, monkeypatch):
    monkeypatch.setattr(oracle, "sinc", lambda x: np.ones_like(np.asarray(x, dtype=float)))
    _fails(pins.test_lossless_homogeneous_matches_closed_form, kw, oracle, (32, 24))


def test_wrong_shift_sign_breaks_closed_form(kw, oracle, monkeypatch):
    orig = oracle.Operators.__init__

    def init(self, pb, full_complex=False):
        orig(self, pb, full_complex)
        self.d_pos = list(self.d_neg)

    monkeypatch.setattr(oracle.Operators, "__init__", init)
    _fails(pins.test_lossless_homogeneous_matches_closed_form, kw, oracle, (32, 24))


def test_absorption_sign_breaks_attenuation(kw, oracle, monkeypatch):
    orig = oracle.Operators.__init__

    def init(self, pb, full_complex=False):
        orig(self, pb, full_complex)
        self.absorb_tau = -self.absorb_tau

    monkeypatch.setattr(oracle.Operators, "__init__", init)
    _fails(pins.test_power_law_attenuation, kw, oracle)


def test_convective_factor_breaks_fubini(kw, oracle, monkeypatch):
    monkeypatch.setattr(oracle, "nonlinear_density", lambda srho, rho0: 1.0 * srho + rho0)
    _fails(pins.test_second_harmonic_follows_fubini, kw, oracle, 0.0, 0.05)


def test_unstaggered_pml_breaks_closed_form(oracle, monkeypatch):
    orig = oracle.get_pml
    monkeypatch.setattr(oracle, "get_pml", lambda n, d, dt, c, s, a, st: orig(n, d, dt, c, s, a, False))
    _fails(pins.test_pml_closed_form, oracle)


def test_source_scaling_factor_breaks_amplitude(kw, oracle, monkeypatch):
    fe = kw.frontend
    orig = fe.save_to_disk_dict

    def halved(*a, **k):
        f = orig(*a, **k)
        if "p_source_input" in f:
            f["p_source_input"] = f["p_source_input"] * 0.5
        return f

    monkeypatch.setattr(fe, "save_to_disk_dict", halved)
    _fails(pins.test_additive_source_scaling_against_finer_dt, kw, oracle)


@pytest.mark.parametrize("mistake", ["sign", "none"])
def test_wrong_non_staggered_shift_breaks_intensity_and_order(oracle, monkeypatch, mistake):
    """A shift of the wrong sign in u_non_staggered fails both its pins; no shift
    at all fails the intensity pin (the peak order then only ties, which the
    reference's order check accepts too)."""
    if mistake == "sign":
        monkeypatch.setattr(oracle, "x_shift_neg", lambda n, d: np.exp(1j * oracle.k_vec_unshifted(n, d) * d / 2))
    else:
        monkeypatch.setattr(oracle, "x_shift_neg", lambda n, d: np.ones(n, dtype=complex))
    _fails(pins.test_plane_wave_intensity_from_u_non_staggered, oracle)
    if mistake == "sign":
        _fails(pins.test_nonstaggered_velocity_peaks_are_ordered, oracle)


# ---- mistakes planted in a copy of the oracle's source ------------------------

def _mutated(oracle, old, new):
    """The oracle module with one statement replaced (a fresh module object)."""
    import inspect
    import types

    import sys

    src = inspect.getsource(oracle)
    assert src.count(old) == 1, old
    name = "kwave_oracle_mutant"
    mod = types.ModuleType(name)
    sys.modules[name] = mod  # (dataclasses look their module up by name)
    try:
        exec(compile(src.replace(old, new), name, "exec"), mod.__dict__)
    finally:
        del sys.modules[name]
    return mod


def test_dirichlet_p_into_one_component_breaks_pin(kw, oracle):
    import test_frontend_pins as fp

    m = _mutated(oracle, "rho[a].reshape(-1)[p_idx] = vals", "rho[0].reshape(-1)[p_idx] = vals * (a == 0)")
    _fails(fp.test_dirichlet_pressure_source_sets_the_pressure, kw, m, 3)


def test_additive_instead_of_dirichlet_u_breaks_pin(kw, oracle):
    import test_frontend_pins as fp

    m = _mutated(oracle, "                    uf[u_idx] = vals\n", "                    uf[u_idx] = uf[u_idx] + vals\n")
    _fails(fp.test_dirichlet_velocity_source_sets_the_velocity, kw, m, 2)


def test_running_max_as_min_breaks_all_outputs_pin(kw, oracle):
    import test_frontend_pins as fp

    m = _mutated(oracle, 'out["p_max_all"] = p.copy() if fi == 1 else np.maximum(out["p_max_all"], p)',
                 'out["p_max_all"] = p.copy() if fi == 1 else np.minimum(out["p_max_all"], p)')
    _fails(fp.test_all_and_final_outputs_equal_brute_force, kw, m, 2, 1)


def test_final_field_from_first_step_breaks_all_outputs_pin(kw, oracle):
    import test_frontend_pins as fp

    m = _mutated(oracle, 'out["u" + nm + "_final"] = u[a].copy()', 'out["u" + nm + "_final"] = 0.0 * u[a]')
    _fails(fp.test_all_and_final_outputs_equal_brute_force, kw, m, 3, 4)

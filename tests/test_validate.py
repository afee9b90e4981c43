"""Input cleaning and argument validation (src/gputls/validate.py:9-203,
src/gputls/helpers.py:16-59): these set the weights every chi2 uses, so they are pinned on
known answers here (the parity tests pass the same dy to both sides and could not see a
mistake in them)."""
import warnings

import numpy as np
import pytest

from gtls_amd.validate import cleaned_array, validate_args, validate_inputs


def test_cleaned_array_drops_invalid_samples():
    t = [1.0, 2.0, np.nan, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0]
    y = [1.0, 0.0, 1.0, np.inf, 1.0, -1.0, 1.0, None, 1.0]
    dy = [0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.0, 0.1, -np.inf]
    # kept: index 0 (all valid), 4 (valid), 6 only without dy (dy = 0 drops it)
    ct, cy, cdy = cleaned_array(t, y, dy)
    assert ct.tolist() == [1.0, 5.0] and cy.tolist() == [1.0, 1.0] and cdy.tolist() == [0.1, 0.2]
    ct, cy = cleaned_array(t, y)
    assert ct.tolist() == [1.0, 5.0, 7.0, 9.0] and cy.tolist() == [1.0, 1.0, 1.0, 1.0]
    assert ct.dtype == np.float64


def test_cleaned_array_lengths():
    # a longer t is cut to len(y) (the reference walks range(len(y))); a shorter one fails
    ct, cy = cleaned_array([1.0, 2.0, 3.0, 4.0], [1.0, 1.0, 1.0])
    assert ct.tolist() == [1.0, 2.0, 3.0]
    with pytest.raises(ValueError):
        cleaned_array([1.0, 2.0], [1.0, 1.0, 1.0])
    with pytest.raises(ValueError):
        cleaned_array(np.ones((2, 3)), np.ones((2, 3)))


def test_dy_normalised_to_mean_one():
    rng = np.random.default_rng(2)
    t = np.arange(100.0) + 1
    y = 1 + rng.normal(0, 1e-3, 100)
    dy = rng.uniform(1e-3, 3e-3, 100)
    _, _, d1 = validate_inputs(t, y, dy)
    assert np.mean(d1) == pytest.approx(1.0, abs=1e-15)
    assert np.allclose(d1 / d1[0], dy / dy[0], rtol=1e-15)
    # scale-free: any constant factor on dy gives the same weights
    _, _, d2 = validate_inputs(t, y, 7.5 * dy)
    assert np.allclose(d1, d2, rtol=1e-15, atol=0)


def test_missing_dy_is_the_std_of_the_flux():
    y = np.array([1.0, 1.002, 0.998, 1.001, 0.999, 1.0])   # std = sqrt(1e-6 * 10 / 6)
    t, y2, dy = validate_inputs(np.arange(6.0) + 1, y, None)
    assert np.all(dy == dy[0]) and dy.size == 6
    assert dy[0] == pytest.approx(np.sqrt(10e-6 / 6), rel=1e-9)


def test_input_errors_and_warning():
    with pytest.raises(ValueError, match="duration"):
        validate_inputs([1.0, 1.0, 1.0], [1.0, 1.0, 1.0], None)     # zero time span
    with pytest.raises(ValueError, match="duration"):
        validate_inputs([1.0], [1.0], None)                        # checked first
    with pytest.raises(ValueError, match="Too few"):
        validate_inputs([1.0, 2.0], [1.0, 1.0], None)
    with pytest.raises(ValueError, match="duration"):
        validate_inputs([1.0, 2.0, 3.0], [np.nan, -1.0, 0.0], None)  # nothing left
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        validate_inputs(np.arange(5.0) + 1, np.full(5, 1.05), None)
        assert any("normalized to 1" in str(x.message) for x in w)


def test_argument_defaults_and_checks():
    a = validate_args({})
    assert a.oversampling_factor == 3 and a.duration_grid_step == 1.1
    # src/gputls/constants.py:41,47,122
    assert a.T0_fit_margin == 0.125 and a.transit_depth_min == pytest.approx(1e-5, rel=1e-12)
    assert a.period_min == 0 and a.period_max == float("inf") and a.n_transits_min == 2
    assert a.mode == "fp64"
    assert validate_args({"T0_fit_margin": 0.5}).T0_fit_margin == 0.125   # clamped
    assert validate_args({"T0_fit_margin": -1}).T0_fit_margin == 0
    for bad in ({"R_star": 0}, {"R_star_min": 2.0, "R_star": 1.0}, {"M_star_max": 0.5},
                {"period_min": -1}, {"period_min": 5, "period_max": 4},
                {"n_transits_min": 0}, {"transit_template": "nope"}, {"mode": "fp16"}):
        with pytest.raises(ValueError):
            validate_args(bad)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        validate_args({"not_a_parameter": 1})
        assert any("Ignoring unknown parameter" in str(x.message) for x in w)

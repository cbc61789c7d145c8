"""The Python mirror's Step and Amplitude (model.rs:119-247) against the reference's own tests
(tests/plasticity.rs:88-127, tests/transient.rs:154-176): piecewise-linear amplitudes, their
typed refusals, and the Step's time step and theta range."""
import math

import pytest

from maille_amd import plasticity as P
from maille_amd import thermal as T


def test_amplitude_is_piecewise_linear_then_constant():
    """tests/plasticity.rs:88-104."""
    a = P.Amplitude([(0.0, 0.0), (1.0, 1.0), (3.0, -1.0)])
    for t, expected in [(0.0, 0.0), (0.25, 0.25), (1.0, 1.0), (2.0, 0.0), (2.5, -0.5),
                        (3.0, -1.0), (10.0, -1.0)]:
        assert abs(a.at(t) - expected) < 1e-15
    ramp = P.Amplitude.ramp(4.0)
    assert abs(ramp.at(1.0) - 0.25) < 1e-15 and abs(ramp.at(5.0) - 1.0) < 1e-15


@pytest.mark.parametrize("points,bad", [
    ([], 0),
    ([(0.0, 0.5), (1.0, 1.0)], 0),
    ([(0.1, 0.0), (1.0, 1.0)], 0),
    ([(0.0, 0.0), (1.0, 1.0), (1.0, 2.0)], 2),
    ([(0.0, 0.0), (1.0, math.nan)], 1),
    ([(0.0, 0.0), (math.inf, 1.0)], 1),
])
def test_amplitude_inputs_are_typed_errors(points, bad):
    """tests/plasticity.rs:106-127."""
    with pytest.raises(P.PlasticityInputError) as err:
        P.Amplitude(points)
    assert err.value.kind == "invalid_amplitude" and err.value.point == bad


def test_ramp_of_zero_duration_and_steps_without_increments_are_refused():
    with pytest.raises(P.PlasticityInputError):
        P.Amplitude.ramp(0.0)
    with pytest.raises(P.PlasticityInputError) as err:
        P.Step(1.0, 0)
    assert err.value.kind == "no_increments"
    with pytest.raises(P.PlasticityInputError):
        P.Step(-1.0, 3)


def test_transient_step_time_step_and_theta():
    """tests/transient.rs:154-160 and Step::with_theta (model.rs:152-159)."""
    step = T.Step(0.5, 4)
    assert abs(step.time_step - 0.125) < 1e-15 and step.theta == 1.0
    assert step.with_theta(0.5).theta == 0.5
    for bad in (0.49, 1.01):
        with pytest.raises(P.PlasticityInputError):
            step.with_theta(bad)

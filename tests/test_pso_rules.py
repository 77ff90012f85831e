"""Host-side PSO rules that need no GPU: the learning automata, the
parameter updaters and the configuration names (fail closed).

Reference: src/optimiser/learning_automata/learning_automata.f90:48-116,
src/optimiser/pso/pso.f90:638-825, src/config/optimiser/pso/pso.f90:92-157,
src/utils/str.f90:85-98 (sanitise).  Expected values are worked by hand.
"""
import numpy as np
import pytest

from ferebus_amd.pso import (AdaptiveSettings, LearningAutomata, ParameterUpdater, PsoSettings,
                             parameter_updater_name, sanitise, stopping_is_relative_change,
                             swarm_updater_name)


def test_automata_choose_reward_penalise():
    la = LearningAutomata(0.0, 1.0, 3, 0.1, 0.1)
    assert la.values == [0.0, 0.5, 1.0]
    assert la.choose(0.5) == 0.5 and la.selected == 1
    # a value on the boundary of two intervals selects the later action
    assert la.choose(1.0 / 3.0) == 0.5
    assert la.choose(0.0) == 0.0 and la.selected == 0
    assert la.choose(0.9) == 1.0
    la.selected = 1
    la.reward()                       # 1/3 + 0.1 (1 - 1/3); 1/3 (1 - 0.1)
    np.testing.assert_allclose(la.prob, [0.3, 0.4, 0.3], rtol=1e-15)
    la.penalise()                     # 0.4 (1 - 0.1); 0.1 / 2 + 0.9 * 0.3
    np.testing.assert_allclose(la.prob, [0.32, 0.36, 0.32], rtol=1e-15)
    # choose scales the uniform by the current sum of probabilities
    la.prob = [0.2, 0.2, 0.2]
    assert la.choose(0.5) == 0.5


class _Draws:
    def __init__(self, values):
        self.values = list(values)

    def random(self):
        return self.values.pop(0)


def test_adaptive_updater_integer_division():
    cfg = PsoSettings(parameter_updater="unified_adaptive",
                      adaptive=AdaptiveSettings(n_inertia_weight=3, n_cognitive_learning_rate=3,
                                                n_social_learning_rate=3, max_inertia_weight=1.0,
                                                max_cognitive_learning_rate=2.0, max_social_learning_rate=4.0,
                                                reward_step_length=0.1, penalty_step_length=0.1))
    costs0 = [5.0, 5.0, 5.0, 5.0]
    up = ParameterUpdater("unified_adaptive", cfg, _Draws([0.5, 0.0, 0.99]), costs0)
    assert up.per_swarm and not up.per_particle
    assert up.coef == [0.5, 0.0, 4.0]
    # three of four improved: 3 // 4 = 0 < 0.5, penalised (pso.f90:760)
    up.rng = _Draws([0.5, 0.5, 0.5])
    up.update([4.0, 4.0, 4.0, 6.0])
    for la in up.automata:
        np.testing.assert_allclose(sorted(la.prob), sorted([0.3, 0.35, 0.35]), rtol=1e-14)
    assert up.previous == [4.0, 4.0, 4.0, 6.0]
    # all four improved (strictly): rewarded
    before = [list(la.prob) for la in up.automata]
    sel = [la.selected for la in up.automata]
    up.rng = _Draws([0.1, 0.1, 0.1])
    up.update([3.0, 3.0, 3.0, 5.0])
    for la, b, s in zip(up.automata, before, sel):
        np.testing.assert_allclose(la.prob[s], b[s] + 0.1 * (1 - b[s]), rtol=1e-14)
    # ties do not count as improvements (cost < previous)
    up.rng = _Draws([0.1, 0.1, 0.1])
    before = [list(la.prob) for la in up.automata]
    sel = [la.selected for la in up.automata]
    up.update([3.0, 3.0, 3.0, 5.0])
    for la, b, s in zip(up.automata, before, sel):
        np.testing.assert_allclose(la.prob[s], b[s] * 0.9, rtol=1e-14)


def test_individual_adaptive_is_per_particle():
    cfg = PsoSettings(parameter_updater="IA")
    up = ParameterUpdater("individual_adaptive", cfg, _Draws([0.2] * 3), [1.0, 1.0])
    assert up.per_particle and not up.per_swarm
    up.rng = _Draws([0.2] * 3)
    up.update([0.5], first=1)        # one particle improved: 1 // 1 = 1, rewarded
    assert up.previous == [1.0, 0.5]


def test_random_updater_inequality():
    cfg = PsoSettings(parameter_updater="random")
    up = ParameterUpdater("random_unified", cfg, _Draws([0.5, 0.25, 0.5]), [1.0])
    w = 0.5
    b = 24 * (1 - w * w) / (7 - 5 * w)
    assert up.coef == [w, 0.25 * b, 0.5 * (b - 0.25 * b)]
    assert up.coef[1] + up.coef[2] < b


def test_names_fail_closed():
    assert sanitise("Relative Change") == "relative_change"
    assert stopping_is_relative_change("Relative Change") is True
    assert stopping_is_relative_change("none") is False
    for bad in ("potential_based", "Sign Test", "relative-change", "stall"):
        with pytest.raises(ValueError):
            stopping_is_relative_change(bad)
    assert swarm_updater_name("Default") == "global" and swarm_updater_name("FORCED") == "forced"
    with pytest.raises(ValueError):
        swarm_updater_name("local")
    assert parameter_updater_name("UAPSO") == "unified_adaptive"
    assert parameter_updater_name("random") == "random_unified"
    assert parameter_updater_name("ri") == "random_individual"
    with pytest.raises(ValueError):
        parameter_updater_name("adaptive")
    with pytest.raises(ValueError):
        PsoSettings(swarm_updater="nearest")

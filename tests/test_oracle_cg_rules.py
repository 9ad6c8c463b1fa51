"""The oracle's CG band-choice and reject rules (cg_module.f90:319-323, 361-364)
against a plane-wave model of one CG step (tests/planewave.py), in cases
built so that each rule must fire, or must not."""
import numpy as np
import pytest

from tests.planewave import PlaneWaveCell, cg_step_model, pinning_cases


@pytest.fixture(scope="module")
def run():
    from oracle.oracle import Oracle
    cell = PlaneWaveCell()
    cases = pinning_cases()
    X = np.stack([cell.vector(b, c) for b, c in cases])
    o = Oracle(cell.system, Ncg=1, mloop=3)
    x, e, res, cnt = o.conjugate_gradient(X, counts=True)
    models = [cg_step_model(cell, b, c) for b, c in cases]
    return cell, cases, X, x, e, cnt, models


def test_cases_are_built_as_intended(run):
    """The model says: band 0 swaps, band 1 is accepted with a ratio between
    1e2 and 1e8, band 2 is rejected (ratio above 1e8)."""
    *_, models = run
    m0, m1, m2 = models
    assert m0["swap"] and not m0["rejected"]
    assert abs(m0["W"][0] - m0["E0"]) > 0.3 and abs(m0["W"][1] - m0["E0"]) < 0.05
    assert not m1["swap"] and 1e3 < m1["ratio"] < 1e6 and not m1["rejected"]
    assert not m2["swap"] and m2["ratio"] > 1e10 and m2["rejected"]


def test_band_choice_keeps_the_upper_ritz_pair(run):
    cell, cases, X, x, e, cnt, models = run
    m = models[0]
    # the band stays at the excited level: E is the upper Ritz value, within
    # 2e-3 Ha of the (2,0,0) eigenvalue, not the lower one 0.67 Ha below
    assert abs(e[0] - m["E"]) < 1e-10
    assert abs(e[0] - cell.lam((2, 0, 0))) < 2e-3
    want = cell.vector(cases[0][0], m["coef"])
    assert np.max(np.abs(x[0] - want)) < 1e-7 * np.max(np.abs(want))


def test_step_below_reject_threshold_is_taken(run):
    cell, cases, X, x, e, cnt, models = run
    m = models[1]
    assert abs(e[1] - m["E"]) < 1e-10
    assert e[1] < m["E0"] - 0.05            # moved to the lower Ritz value
    want = cell.vector(cases[1][0], m["coef"])
    assert np.max(np.abs(x[1] - want)) < 1e-6 * np.max(np.abs(want))


def test_step_above_reject_threshold_is_undone(run):
    cell, cases, X, x, e, cnt, models = run
    m = models[2]
    assert cnt[1] == 1                       # one rejected update over the three bands
    assert abs(e[2] - m["E0"]) < 1e-12       # E restored to E1
    assert np.array_equal(x[2], X[2])        # x untouched

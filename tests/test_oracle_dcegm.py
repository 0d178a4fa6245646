"""Pins of the DC-EGM / FUES oracle (CPU).

* FUES rows of the reference's own unit tests (`tests/solution/test_fues.py`):
  analytic envelope of two crossing segments, the crossing inserted twice with
  both policies, node-aligned crossings, a three-branch chain, a dominated
  coincident pair, savings provenance, shift invariance, overflow.
* The Iskhakov et al. (2017) analytical value functions (the reference holds
  DC-EGM to 0.03 elementwise; the oracle lands within 3e-4).
* The independent `dcegm` package's taste-shock fixture and the two DC-EGM
  values the reference documents for the rows it excludes from that fixture
  (`tests/test_dcegm_fixture_crosscheck.py:128-136`).
* A fine consumption-grid search of the same stochastic model.
* Closed form of the last decision period: consume everything, V = u(w).
"""

import csv
from pathlib import Path

import numpy as np
import pytest

from oracle import dcegm as od
from pylcm_amd import examples as ex

GOLDEN = Path(__file__).parent / "golden"


def _interp(g, v, x):
    keep = ~np.isnan(g)
    return np.interp(x, g[keep], v[keep])


def _crossing():
    r_a = np.array([0.6, 1.2, 1.8, 2.4])
    r_b = np.array([0.8, 1.6, 2.6, 3.4])
    grid = np.concatenate([r_a, r_b])
    value = np.concatenate([1.0 + 0.4 * r_a, 0.1 + 0.8 * r_b])
    policy = np.concatenate([0.30 * r_a, 0.55 * r_b])
    return grid, policy, value


def test_fues_crossing_segments_analytic_envelope():
    g, p, v, n = od.refine_envelope(*_crossing(), 12)
    rq = np.linspace(0.7, 3.3, 53)
    np.testing.assert_allclose(_interp(g, v, rq), np.maximum(1.0 + 0.4 * rq, 0.1 + 0.8 * rq), atol=1e-12)
    at = np.isclose(g[~np.isnan(g)], 2.25, atol=1e-12)
    assert at.sum() == 2
    np.testing.assert_allclose(np.sort(p[~np.isnan(g)][at]), [0.3 * 2.25, 0.55 * 2.25], atol=1e-12)
    for dominated in (0.8, 1.6, 2.4):
        assert not np.any(np.isclose(g[~np.isnan(g)], dominated, atol=1e-8))
    raw = g[: n]
    assert np.all(np.diff(raw) >= 0) and np.all(np.isnan(g[n:]))


def test_fues_node_aligned_crossing_keeps_right_policy():
    g, p, v, _ = od.refine_envelope(np.array([9.0, 10, 12, 10, 11]), np.array([8.5, 9, 9, 4, 4.5]),
                                    np.array([4.6, 5, -100, 5, 5.8]), 12)
    assert _interp(g, p, 10.1) == pytest.approx(4.05, abs=1e-12)
    assert _interp(g, v, 10.1) == pytest.approx(5.08, abs=1e-12)


def test_fues_three_branch_chain():
    g, p, v, n = od.refine_envelope(
        np.array([9.0, 10, 10, 20, 20, 21]), np.array([8.0, 8, 4, 4, 2, 2]),
        np.array([4.875, 5, 5, 7.5, 7.5, 8]), 16, savings=np.array([1.0, 2, 6, 16, 18, 19]))
    assert n == 6
    assert _interp(g, p, 15.0) == pytest.approx(4.0, abs=1e-12)
    assert _interp(g, v, 15.0) == pytest.approx(6.25, abs=1e-12)


def test_fues_dominated_coincident_pair_and_shift_invariance():
    grid = np.array([9.0, 10, 9, 10, 11, 10, 11])
    pol = np.array([9.0, 9, 7, 7, 7, 4, 4])
    val = np.array([5 - 1 / 9, 5.0, 6 - 1 / 7, 6.0, 6 + 1 / 7, 5.0, 5.25])
    sav = np.array([0.0, 1, 2, 3, 4, 6, 7])
    g, p, v, n = od.refine_envelope(grid, pol, val, 16, savings=sav)
    assert n == 3
    assert _interp(g, v, 10.0) == pytest.approx(6.0, abs=1e-12)
    assert _interp(g, p, 10.0) == pytest.approx(7.0, abs=1e-12)
    g2, p2, v2, n2 = od.refine_envelope(grid, pol, val + 1e8, 16, savings=sav)
    assert n2 == n
    np.testing.assert_allclose(p2, p, atol=1e-8, equal_nan=True)


def test_fues_savings_rules_and_single_segments():
    gg = np.array([0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0])
    pp = gg - 0.01
    pp2 = pp.copy()
    pp2[4] = np.nextafter(pp[4], np.inf)
    assert od.refine_envelope(gg, pp, np.log(gg), 12)[3] == 8
    assert od.refine_envelope(gg, pp2, np.log(gg), 12)[3] == 8
    assert od.refine_envelope(gg, pp, np.log(gg), 12, savings=np.full(8, 0.01))[3] == 8
    lin = np.linspace(1.0, 10.0, 10)
    g, p, v, n = od.refine_envelope(lin, 0.3 * lin, np.log(lin), 14)
    assert n == 10
    np.testing.assert_allclose(g[:10], lin, atol=1e-12)
    steep = np.linspace(1.0, 5.0, 9)
    assert od.refine_envelope(steep, 5.0 - 0.9 * steep, np.log(steep), 12)[3] == 9
    g, _, _, _ = od.refine_envelope(np.array([1.0, 2, 3]), np.array([0.5, 1.8, 1.5]), np.array([1, 0.9, 2.0]), 6)
    assert not np.any(np.isclose(g[~np.isnan(g)], 2.0))
    # provenance: a cross-source decrease is dropped only with source savings
    grid = np.array([100000.0, 100001.0, 100002.0])
    pol = np.array([99998.0, 100000.0, 100001.125])
    val = np.array([10.0, 10.5, 10.55])
    assert od.refine_envelope(grid, pol, val, 8, savings=np.array([2.0, 1.0, 0.875]))[3] == 2


def test_fues_overflow_reported():
    assert od.refine_envelope(*_crossing(), 4)[3] > 4


def test_read_row_rules():
    xp = np.array([1.0, 2.0, 2.0, 4.0, np.nan])
    fp = np.array([1.0, 3.0, 5.0, 9.0, np.nan])
    got = od.read_row(np.array([0.0, 1.5, 2.0, 3.0, 4.0, 9.0, np.nan]), xp, fp)
    # secant extension below, right copy at a kink, clamp at/above the last
    np.testing.assert_allclose(got[:6], [-1.0, 2.0, 5.0, 7.0, 9.0, 9.0])
    assert np.isnan(got[6])
    assert od.read_row(np.array([5.0]), np.array([1.0, np.nan]), np.array([3.0, np.nan]))[0] == 3.0
    assert np.isnan(od.read_row(np.array([5.0]), np.array([np.nan, np.nan]), np.array([np.nan, np.nan]))[0])
    # Hermite with exact slopes reproduces a monotone cubic (the Fritsch-Carlson
    # limiter does not bind on x**3 over [1, 3]); the linear read does not
    x = np.linspace(1.0, 3.0, 9)
    q = np.linspace(1.05, 2.95, 40)
    assert np.max(np.abs(od.read_row(q, x, x**3, 3 * x**2) - q**3)) < 1e-12
    assert np.max(np.abs(od.read_row(q, x, x**3) - q**3)) > 1e-2


@pytest.mark.parametrize("case,n,dis", [("iskhakov_2017_five_periods", 6, 1.0),
                                        ("iskhakov_2017_low_delta", 4, 0.1)])
def test_iskhakov_analytical_solution(case, n, dis):
    m = ex.iskhakov_dcegm_model(n)
    p = ex.retirement_params(n, discount_factor=0.98, disutility_of_work=dis, interest_rate=0.0,
                             wage=20.0, log_utility=True)
    V, _ = od.solve(m, p)
    for kind, reg in (("worker", "working_life"), ("retired", "retirement")):
        an = np.genfromtxt(GOLDEN / "analytical_solution" / f"{case}__values_{kind}.csv", delimiter=",")
        num = np.stack([V[t][reg] for t in sorted(V)[:-1]])
        np.testing.assert_allclose(num, an, atol=1e-3)


@pytest.fixture(scope="module")
def twin_solution():
    m = ex.dcegm_twin_model()
    p = ex.dcegm_twin_params(0.2)
    return od.solve(m, p)[0]


def test_dcegm_package_fixture(twin_solution):
    V = twin_solution
    grid = np.array(ex.TWIN_WEALTH_POINTS)
    n = 0
    with open(GOLDEN / "dcegm_reference" / "ijrs_taste_shocks_reference.csv") as fh:
        for r in csv.DictReader(fh):
            lc, w = int(r["lagged_choice"]), float(r["wealth"])
            if lc == 1 and w <= 5.0:  # the rows the reference excludes for DC-EGM
                continue
            vals = np.array([float(r["value_work"]) if r["value_work"] else np.nan,
                             float(r["value_retire"]) if r["value_retire"] else np.nan])
            mx = np.nanmax(vals)
            emax = mx + 0.2 * np.log(np.nansum(np.exp((vals - mx) / 0.2)))
            i = int(np.searchsorted(grid, w))
            reg = "working_life" if lc == 0 else "retirement"
            np.testing.assert_allclose(V[int(r["period"])][reg][i], emax, rtol=2e-3, atol=2e-3)
            n += 1
    assert n > 80


def test_reference_documented_dcegm_values(twin_solution):
    """The reference's comment quotes its own DC-EGM at period 0, retirement:
    wealth 2 -> -23.468, wealth 5 -> -4.9236."""
    grid = np.array(ex.TWIN_WEALTH_POINTS)
    V = twin_solution[0]["retirement"]
    assert V[int(np.searchsorted(grid, 2.0))] == pytest.approx(-23.468, abs=6e-4)
    assert V[int(np.searchsorted(grid, 5.0))] == pytest.approx(-4.9236, abs=6e-5)


def test_last_decision_period_closed_form():
    """Only the stateless terminal follows, so E[mu'] = 0: every Euler candidate
    sits at (u')^{-1}(eps) far right, the constrained candidates are
    c_k = span 1e-4 ratio^k with span = max(w) - A_0 and value u(c_k), and the
    published V is floored by u(w) and the same for every income node."""
    m = ex.dcegm_retirement_model(3, n_savings=100, savings_max=50.0, n_wealth=40, wealth_range=(0.5, 50.0))
    p = ex.dcegm_params(3)
    V, C = od.solve(m, p)
    w = np.linspace(0.5, 50.0, 40)
    u_ret = w ** (1 - 2.0) / (1 - 2.0)
    ratio = 1e4 ** (1 / 19)
    ck = 50.0 * 1e-4 * ratio ** np.arange(20)
    for col in C[1]["retirement"]:
        assert col["n_kept"] == 120
        np.testing.assert_allclose(col["grid"][:20], ck, rtol=1e-14)
        np.testing.assert_allclose(col["policy"][:20], ck, rtol=1e-14)
        np.testing.assert_allclose(col["value"][:20], -1.0 / ck, rtol=1e-14)
        np.testing.assert_allclose(col["policy"][20:120], np.finfo(np.float64).eps ** -0.5, rtol=1e-14)
        assert np.all(col["V"] >= u_ret) and np.all(col["V"] - u_ret <= 1e-2 * np.abs(u_ret))
    np.testing.assert_allclose(V[1]["retirement"], np.broadcast_to(C[1]["retirement"][0]["V"], (5, 40)))


@pytest.mark.slow
def test_dcegm_against_fine_grid_search():
    from oracle import cpu_solve

    T = 5
    kw = dict(n_savings=300, savings_max=100.0, n_wealth=80, wealth_range=(0.05, 100.0))
    m = ex.dcegm_retirement_model(T, **kw)
    p = ex.dcegm_params(T)
    V, _ = od.solve(m, p)
    Vg, _ = cpu_solve.solve_c(ex.dcegm_retirement_model(T, n_consumption=4000, grid_search=True, **kw), p,
                              threads=4)
    for t in sorted(V):
        for r in ("working_life", "retirement"):
            if r not in V[t]:
                continue
            a, b = np.asarray(V[t][r])[..., 2:], np.asarray(Vg[t][r])[..., 2:]
            d = (a - b) / np.abs(b)
            assert d.min() > -1e-3  # the grid search is the coarser optimiser
            assert np.median(np.abs(d)) < 3e-3

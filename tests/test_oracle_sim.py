"""Pins of the simulation oracle: the Philox generator against the published
Random123 known-answer vectors, the draw transforms against their
distributions, and the simulated first-period choice against the solve."""

import numpy as np
import pytest

from oracle import gridsearch as og
from oracle import simulate as osim
from pylcm_amd import examples as ex

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_philox_known_answers(ctr, key, expected):
    out = osim.philox4x32_10([ctr], key)[0]
    assert tuple(int(x) for x in out) == expected


def test_normal_and_uniform_draws():
    b = osim.draws(np.arange(200_000), 3, 1, 12345)
    x = osim.normal_bm(b)
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1) < 0.01
    assert abs(np.mean(x**4) - 3) < 0.05
    u = osim.uniform53(b[:, 0], b[:, 1])
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.005


def test_categorical_draw_frequencies_and_zero_mass():
    p = np.tile([0.2, 0.0, 0.5, 0.3], (100_000, 1))
    u = osim.uniform53(*osim.draws(np.arange(100_000), 0, 2, 7)[:, :2].T)
    k = osim.draw_categorical(p, u)
    freq = np.bincount(k, minlength=4) / len(k)
    assert np.allclose(freq, [0.2, 0.0, 0.5, 0.3], atol=0.01)
    assert osim.draw_categorical(np.array([[0.0, 1.0, 0.0]]), np.array([0.999]))[0] == 1


def test_first_period_matches_solve_on_grid():
    """Agents starting on wealth grid points choose the solve's policy and get
    its value (the simulation's Q equals the solve's Q at on-grid states)."""
    m, p = ex.deterministic_workload()
    V, P = og.solve(m, p)
    w = og.linspace(1, 400, 100)
    n = 100
    ic = {"wealth": w, "regime_id": np.zeros(n, dtype=int), "age": np.zeros(n)}
    out = osim.simulate(m, p, V, ic, seed=1)
    assert np.array_equal(out["action"][0], P[0]["working_life"])
    assert np.allclose(out["value"][0], V[0]["working_life"], rtol=1e-14)
    # every agent is dead in the last period and was alive before
    assert np.all(out["regime"][-1] == 2) and np.all(out["regime"][0] == 0)


def test_stochastic_process_path_moments():
    """AR(1) income paths: z' = rho z + sigma eps with eps ~ N(0,1)."""
    T = 6
    m = ex.retirement_model(T, n_wealth=20, n_consumption=20, income_nodes=5)
    inc = {"rho": 0.9, "sigma": 0.1, "mu": 0.0}
    p = ex.retirement_params(T, income=inc)
    V, _ = og.solve(m, p)
    n = 3000
    ic = {"wealth": np.full(n, 50.0), "income": np.zeros(n), "regime_id": np.zeros(n, dtype=int)}
    out = osim.simulate(m, p, V, ic, seed=3)
    z1 = out["z"][1]
    assert abs(z1.mean()) < 0.01 and abs(z1.std() - 0.1) < 0.005


def test_taste_noise_is_mean_zero_gumbel():
    """scale * (Gumbel - Euler's gamma): mean 0, variance scale^2 pi^2 / 6."""
    n = 200_000
    noise = osim.taste_noise(np.arange(n), 3, 0, 12345, 0.7)
    assert abs(noise.mean()) < 0.01
    assert abs(noise.var() / (0.7 ** 2 * np.pi ** 2 / 6) - 1) < 0.02


def test_gumbel_max_choice_frequencies_are_logit():
    """The first argmax of Qc + noise picks cell d with probability
    softmax(Qc / scale)_d (McFadden), the closed form the smoothed solve uses."""
    n = 200_000
    Qc = np.array([0.3, 0.1, -0.2])
    agents = np.arange(n)
    noisy = Qc[None, :] + np.stack([osim.taste_noise(agents, 5, d, 99, 0.5) for d in range(3)], axis=1)
    freq = np.bincount(np.argmax(noisy, axis=1), minlength=3) / n
    logit = np.exp(Qc / 0.5) / np.exp(Qc / 0.5).sum()
    assert np.max(np.abs(freq - logit)) < 0.005

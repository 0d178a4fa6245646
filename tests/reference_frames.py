"""The reference's simulation frames as RNG-free checks (test infrastructure).

The reference pins its simulation with pickled DataFrames
(`tests/test_regression_test.py:69-140`), decoded to
`tests/golden/regression/{f32,f64}/*_simulation.npz`. Its random draws (JAX
threefry) cannot be reproduced, but every row in a non-terminal regime is a
state the reference visited: its value and chosen action there follow from the
solution alone, and the next row's wealth from the law of motion at that row's
income. `visited_states` turns those rows into initial conditions (one agent
per row, starting at the row's age), and `compare` checks a simulation of them
row by row at the reference's bar (`assert_frame_equal(atol=1e-5)`, whose
relative tolerance is also 1e-5).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from pylcm_amd import examples as ex
from tests import models as tm

GOLDEN = Path(__file__).resolve().parent / "golden" / "regression"
FRAMES = ("mortality_simulation", "precautionary_savings_simulation")
ATOL = RTOL = 1e-5


def load(prec: str, name: str) -> dict:
    z = np.load(GOLDEN / prec / f"{name}.npz")
    return {k: z[k] for k in z.files}


def model_and_params(name: str):
    """The model and params the reference's generator used
    (`tests/data/regression_tests/generate_benchmark_data.py:29-73`)."""
    if name == "mortality_simulation":
        return ex.mortality_model(4), ex.mortality_params(4)
    return (tm.precautionary_savings_model(5, wealth_n_points=10, consumption_n_points=10),
            tm.precautionary_savings_params(sigma=0.2, rho=0.9))


def terminal_names(model):
    return {r for r in model.regime_names if model.layouts[r].terminal}


def visited_states(frame: dict, model):
    """(row indices, initial conditions) of the frame's rows in non-terminal regimes."""
    term = terminal_names(model)
    rows = np.nonzero(~np.isin(frame["regime_name"], list(term)))[0]
    ic = {"age": frame["age"][rows].astype(np.float64),
          "wealth": frame["wealth"][rows].astype(np.float64),
          "regime_id": np.array([model.regime_code(str(r)) for r in frame["regime_name"][rows]])}
    if "income" in frame:
        ic["income"] = frame["income"][rows].astype(np.float64)
    return rows, ic


def next_row(frame: dict, i: int):
    """Index of the same subject's next row, or None."""
    j = i + 1
    if j < len(frame["period"]) and frame["subject_id"][j] == frame["subject_id"][i] \
            and frame["period"][j] == frame["period"][i] + 1:
        return j
    return None


def compare(frame: dict, model, rows, first: dict, second: dict):
    """Check per visited row: the value, consumption and labor choice of the
    agent started there (`first`: arrays over the agents), and, where the
    reference's subject and the agent are both alive one period later, the next
    wealth (`second`: the agents' records one period later; NaN wealth where the
    agent is no longer in a non-terminal regime). Returns (n_rows, n_next)."""
    term = terminal_names(model)
    n_next = 0
    for k, i in enumerate(rows):
        for col in ("value", "consumption"):
            np.testing.assert_allclose(first[col][k], frame[col][i], rtol=RTOL, atol=ATOL,
                                       err_msg=f"{col} at row {i}")
        if "labor_supply" in frame and frame["labor_supply"][i] != "":
            assert first["labor_supply"][k] == frame["labor_supply"][i], f"labor_supply at row {i}"
        j = next_row(frame, i)
        if j is None or str(frame["regime_name"][j]) in term or not np.isfinite(second["wealth"][k]):
            continue
        np.testing.assert_allclose(second["wealth"][k], frame["wealth"][j], rtol=RTOL, atol=ATOL,
                                   err_msg=f"next wealth after row {i}")
        n_next += 1
    return len(rows), n_next

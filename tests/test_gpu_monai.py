"""Monai valley benchmark against the reference binary's own gauge output
(cases/monai/{hwfv1,mwdg2}-eps-{0,0.001}/res.txt: FP32, max_ref_lvl 9,
22.5 s, 3 gauges every 0.1 s). The reference produced these with its CUDA
build (FMA contraction, pow), so agreement is to a tolerance, not bit for
bit: the mean absolute depth error over all rows and gauges must stay below
MAE_TOL (depths are 0.003-0.03 m). Rounding moves each step's dt, so a row's
time (the first step at or past each 0.1 s mark) may differ by part of a
step (T_TOL).
The comparison is the one of the reference's cases/monai/run-functional-tests.py
(mean absolute error of the depth columns)."""
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "monai")
MAE_TOL = 5e-4
T_TOL = 0.01


@pytest.mark.parametrize("solver,eps", [("hwfv1", 0.0), ("hwfv1", 1e-3), ("mwdg2", 0.0), ("mwdg2", 1e-3)])
def test_monai_gauges_match_reference_output(case_dir, solver, eps):
    import mwdg2
    par = cases.monai(os.path.join(str(case_dir), f"monai_{solver}_{eps}"), FIXTURE, solver=solver, eps=eps)
    with mwdg2.Simulation(par, "f32") as s:
        c = s.run(max_steps=200000, batch=256)
        rows = s.stage_records()
    tag = "0" if eps == 0 else "0.001"
    ref = np.loadtxt(os.path.join(FIXTURE, f"res-{solver}-eps-{tag}.txt"), skiprows=9)
    assert c["done"] == 1
    assert rows.shape[0] == ref.shape[0], (rows.shape, ref.shape)
    assert np.max(np.abs(rows[:, 0] - ref[:, 0])) < T_TOL
    err = np.abs(rows[:, 1:4] - ref[:, 1:4])
    print(f"monai {solver} eps={eps}: steps {c['steps']}, mean |dh| {err.mean():.2e}, max |dh| {err.max():.2e}")
    assert err.mean() < MAE_TOL
    # gauge bed elevations as the reference printed them (%f)
    head = open(os.path.join(FIXTURE, f"res-{solver}-eps-{tag}.txt")).read().splitlines()[3:6]
    zref = np.array([float(l.split()[3]) for l in head])
    assert np.allclose(rows[0, 4:7], zref, atol=6e-7)

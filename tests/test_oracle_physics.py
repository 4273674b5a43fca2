"""Pins the oracle's operators against textbook results and invariants fixed
by the mathematics, not against a restatement of its own formulas.

- HLL flux (operators/flux_HLL.cuh:9-111): Toro's HLL with two-rarefaction
  wave-speed estimates, dry-bed speeds and the transverse velocity upwinded by
  the middle-wave speed (Toro, Shock-Capturing Methods for Free-Surface
  Shallow Flows, 2001, ch. 10), written here from the book; consistency
  F(U, U) = F(U) and rotational symmetry x <-> y.
- Friction (operators/apply_friction.cuh:7-53): Manning's law as dt -> 0 and
  the closed-form large-dt limit of the linearised implicit update.
- Well-balancedness: a lake at rest over a rough bed stays at rest (FP64,
  |q| <= 1e-10), both solvers (the reference's c-property check,
  cases/tests/test.py:105,396).
- Mass: closed walls, no friction, no sources, uniform grid (epsilon 0)
  conserve sum(h0 * dx^2). On an adaptive grid the reference's neighbour
  rule (a leaf sees the neighbour at its own level, find_neighbours.cu:87-108)
  makes the fluxes on either side of a resolution change differ, so mass is
  only approximately conserved there and no invariant is claimed.
- CFL: dt_next equals the exact minimum of the per-leaf CFL step.
"""
import math
import os

import numpy as np
import pytest

import cases
import oracle

G = 9.80665
TOL = 1e-3


def toro_hll_x(UL, UR, g=G, tol=TOL):
    hL, huL, hvL = UL
    hR, huR, hvR = UR
    if hL < tol and hR < tol:
        return np.zeros(3)
    uL, vL = (huL / hL, hvL / hL) if hL >= tol else (0.0, 0.0)
    uR, vR = (huR / hR, hvR / hR) if hR >= tol else (0.0, 0.0)
    hL_, hR_ = (hL if hL >= tol else 0.0), (hR if hR >= tol else 0.0)
    aL, aR = math.sqrt(g * hL_), math.sqrt(g * hR_)
    # two-rarefaction estimates of the star region
    h_star = (0.5 * (aL + aR) + 0.25 * (uL - uR)) ** 2 / g
    u_star = 0.5 * (uL + uR) + aL - aR
    a_star = math.sqrt(g * h_star)
    sL = uR - 2 * aR if hL_ < tol else min(uL - aL, u_star - a_star)
    sR = uL + 2 * aL if hR_ < tol else max(uR + aR, u_star + a_star)
    FL = np.array([hL_ * uL, hL_ * uL * uL + 0.5 * g * hL_ ** 2, hL_ * uL * vL])
    FR = np.array([hR_ * uR, hR_ * uR * uR + 0.5 * g * hR_ ** 2, hR_ * uR * vR])
    if sL >= 0:
        return FL
    if sR < 0:
        return FR
    UL_ = np.array([hL_, hL_ * uL]); UR_ = np.array([hR_, hR_ * uR])
    F = (sR * FL[:2] - sL * FR[:2] + sL * sR * (UR_ - UL_)) / (sR - sL)
    s_mid = (sL * hR_ * (uR - sR) - sR * hL_ * (uL - sL)) / (hR_ * (uR - sR) - hL_ * (uL - sL))
    return np.array([F[0], F[1], F[0] * (vL if s_mid >= 0 else vR)])


STATES = [
    ((2.0, 0.0, 0.0), (0.5, 0.0, 0.0)),          # dam break, still water
    ((1.0, 2.5, 0.3), (1.2, -0.4, -0.7)),         # converging, transverse flow
    ((1.5, 6.0, 1.0), (1.4, 5.0, -1.0)),          # supercritical to the right
    ((1.0, -5.0, 0.2), (1.1, -6.0, 0.4)),         # supercritical to the left
    ((0.0, 0.0, 0.0), (1.0, -0.5, 0.25)),         # dry left bed
    ((1.0, 0.8, -0.3), (0.0, 0.0, 0.0)),          # dry right bed
    ((0.3, -0.2, 0.1), (0.4, 0.1, 0.0)),          # shallow rarefaction
]


@pytest.mark.parametrize("UL,UR", STATES)
def test_hll_matches_textbook(UL, UR):
    got = oracle.flux_hll(UL, UR, axis="x")
    exp = toro_hll_x(UL, UR)
    assert np.allclose(got, exp, rtol=1e-13, atol=1e-13), (got, exp)


@pytest.mark.parametrize("U", [(1.0, 0.5, -0.2), (2.0, 0.0, 0.0), (0.7, -3.0, 1.0)])
def test_hll_consistency(U):
    h, hu, hv = U
    phys = np.array([hu, hu * hu / h + 0.5 * G * h * h, hu * hv / h])
    assert np.allclose(oracle.flux_hll(U, U, axis="x"), phys, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("UL,UR", STATES)
def test_hll_y_is_rotated_x(UL, UR):
    # y flux of (h, qx, qy) equals the x flux of (h, qy, -qx), rotated back
    rl = (UL[0], UL[2], -UL[1]); rr = (UR[0], UR[2], -UR[1])
    fx = oracle.flux_hll(rl, rr, axis="x")
    fy = oracle.flux_hll(UL, UR, axis="y")
    assert np.allclose(fy, [fx[0], -fx[2], fx[1]], rtol=1e-14, atol=1e-14)


def test_friction_tends_to_manning_law():
    h, qx, qy, n = 0.8, 0.6, -0.25, 0.03
    u, v = qx / h, qy / h
    s = math.hypot(u, v)
    # Manning: dq/dt = -g n^2 q |q| / h^(7/3)
    rate = np.array([-G * n * n * u * s / h ** (1 / 3), -G * n * n * v * s / h ** (1 / 3)])
    for dt in (1e-4, 1e-5, 1e-6):
        q = oracle.friction(h, qx, qy, n, dt)
        est = (q - [qx, qy]) / dt
        assert np.allclose(est, rate, rtol=50 * dt, atol=0)


def test_friction_large_dt_limit():
    # the linearised implicit step tends to q_x u^2 / (2u^2 + v^2) as dt -> inf
    h, qx, qy, n = 0.5, 0.4, 0.3, 0.05
    u, v = qx / h, qy / h
    q = oracle.friction(h, qx, qy, n, 1e14)
    assert np.allclose(q, [qx * u * u / (2 * u * u + v * v), qy * v * v / (2 * v * v + u * u)], rtol=1e-9)
    q1 = oracle.friction(h, qx, 0.0, n, 1e14)
    assert abs(q1[0] - qx / 2) < 1e-9 and q1[1] == 0.0


def test_friction_below_thresholds_zeroes_discharge():
    assert np.array_equal(oracle.friction(5e-4, 0.3, 0.2, 0.03, 0.1), [0.0, 0.0])
    assert np.array_equal(oracle.friction(1.0, 1e-10, -1e-10, 0.03, 0.1), [0.0, 0.0])


def lake_case(d, L, solver, eta, seed=4, manning=0.0, t_end=5.0):
    os.makedirs(d, exist_ok=True)
    n = 1 << L
    z = cases.gaussian_random_field(n, seed, 0.5) + 1.0
    h = np.maximum(0.0, eta - z)
    cases.write_raster(os.path.join(d, "case.dem"), z, 0.0, 0.0, 1.0, fmt="%.17g")
    cases.write_raster(os.path.join(d, "case.start"), h, 0.0, 0.0, 1.0, fmt="%.17g")
    par = os.path.join(d, "case.par")
    cases.write_par(par, cases._par(solver, L, 1e-3, t_end, manning))
    return par, z


@pytest.mark.parametrize("solver", ["mwdg2", "hwfv1"])
def test_lake_at_rest_fully_wet(tmp_path, solver):
    par, z = lake_case(str(tmp_path / solver), 5, solver, eta=10.0)
    o = oracle.Oracle(par, "f64")
    try:
        for _ in range(8):
            o.step()
        n = int(o.scalar("length"))
        for m in (("0", "1x", "1y") if solver == "mwdg2" else ("0",)):
            assert np.max(np.abs(o.get(f"assem.qx{m}")[:n])) <= 1e-10
            assert np.max(np.abs(o.get(f"assem.qy{m}")[:n])) <= 1e-10
        eta0 = o.get("assem.h0")[:n] + o.get("assem.z0")[:n]
        assert np.max(np.abs(eta0 - 10.0)) <= 1e-10
    finally:
        o.close()


def total_mass(o):
    n = int(o.scalar("length"))
    L = int(o.scalar("L"))
    lev = o.get("assem.lev")[:n]
    h0 = o.get("assem.h0")[:n]
    dx = o.scalar("dx")
    return float(np.sum(h0 * (dx * (2.0 ** (L - lev))) ** 2))


@pytest.mark.parametrize("solver", ["mwdg2", "hwfv1"])
def test_mass_conserved_with_closed_walls(tmp_path, solver):
    par = cases.circular_dam_break(str(tmp_path / solver), L=6, eps=0.0, t_end=2.0, solver=solver,
                                   cfl=0.33 if solver == "mwdg2" else None)
    o = oracle.Oracle(par, "f64")
    try:
        o.step()
        m0 = total_mass(o)
        for _ in range(10):
            o.step()
        assert int(o.scalar("length")) == 4 ** 6
        # telescoping fluxes: each face's flux leaves one cell and enters the other
        assert abs(total_mass(o) - m0) <= 1e-12 * m0
    finally:
        o.close()


def test_cfl_min_is_exact_minimum(tmp_path):
    par = cases.random_case(str(tmp_path / "cfl"), L=5, seed=2, t_end=1.0)
    o = oracle.Oracle(par, "f64")
    try:
        for _ in range(3):
            o.step()
        n = int(o.scalar("length"))
        L = int(o.scalar("L"))
        lev = o.get("assem.lev")[:n]
        h = o.get("assem.h0")[:n]
        qx = o.get("assem.qx0")[:n]
        qy = o.get("assem.qy0")[:n]
        dxl = o.scalar("dx") * 2.0 ** (L - lev)
        cfl = o.scalar("CFL")
        wet = h >= TOL
        with np.errstate(divide="ignore", invalid="ignore"):
            c = np.sqrt(G * h)
            dtx = cfl * dxl / (np.abs(qx / h) + c)
            dty = cfl * dxl / (np.abs(qy / h) + c)
        cand = np.where(wet, np.minimum(dtx, dty), o.scalar("initial_tstep"))
        assert o.scalar("dt") == pytest.approx(float(cand.min()), rel=1e-14)
    finally:
        o.close()


def test_wavemaker_keeps_still_water_still():
    """Monai wavemaker (operators/non_reflective_wave.cuh:15-93): with no
    inlet elevation and the tank at rest at its still depth (0.13535 m), both
    Riemann invariants are those of the rest state, so the ghost is the rest
    state; a raised inlet raises the ghost depth and pushes water in."""
    hb = 0.13535
    h, u = oracle.wavemaker(0.0, 0.01, 0.014, hb, 0.0)
    assert abs(h - hb) < 1e-12 and abs(u) < 1e-12
    h1, u1 = oracle.wavemaker(0.005, 0.01, 0.014, hb, 0.0)
    assert h1 > hb and u1 > 0.0
    # a cell below the still depth draws water in through the boundary
    h2, u2 = oracle.wavemaker(0.0, 0.01, 0.014, 0.12, 0.0)
    assert 0.12 < h2 < hb and u2 > 0.0


def test_exp_within_two_ulp_of_libm():
    """The oracle's exp (Cody-Waite reduction + Taylor series, used for the
    Oregon wavemaker pulse) against numpy's exp, FP64 and FP32, over the
    pulse's range, the whole finite range and the reduction's edges."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-60.0, 0.0, 4000), rng.uniform(-745.0, 709.7, 4000),
                         np.linspace(-0.35, 0.35, 301), [0.0, 1e-12, -1e-12, 0.5 * np.log(2.0), -708.5, -744.0]])
    for x in xs:
        e, r = oracle.exp(float(x)), float(np.exp(x))
        if r < 2.3e-308:                                   # subnormal results: absolute half-ulp steps
            assert abs(e - r) <= 2 * 5e-324
        else:
            assert abs(e - r) <= 2 * np.spacing(r), (x, e, r)
    assert oracle.exp(0.0) == 1.0 and oracle.exp(-800.0) == 0.0 and np.isinf(oracle.exp(710.0))
    xf = np.concatenate([rng.uniform(-60.0, 0.0, 4000), rng.uniform(-87.0, 88.0, 4000)]).astype(np.float32)
    for x in xf:
        e, r = np.float32(oracle.exp(float(x), np.float32)), np.exp(np.float64(x))
        assert abs(float(e) - r) <= 2 * float(np.spacing(np.float32(r))), (x, e, r)
    assert oracle.exp(0.0, np.float32) == 1.0


def test_oregon_ghost_closed_form(tmp_path):
    """Oregon Seaside wavemaker (neighbours/add_ghost_cells.cu:471-481): at
    the finest west-border leaves of an HVAR span the ghost keeps the local
    depth and carries the discharge (2 s(t) - v) h with
    s(t) = 0.6 exp(-(t - 14.75)^2 / 4) and v = qx / h (0 below tol_h); the
    local depth is not reset to the inlet as for a plain HVAR boundary."""
    bci = "W 0 16 HVAR wave\n"
    bdy = "wave\n2 seconds\n5.0 0\n5.0 100\n"
    par = cases.random_case(str(tmp_path / "oregon"), L=4, seed=5, eps=0.0, t_end=30.0, bci=bci, bdy=bdy,
                            startq=True, wet_frac=0.8)
    with open(par, "a") as f:
        f.write("oregon-seaside\n")
    o = oracle.Oracle(par, "f64")
    try:
        o.step()
        n = int(o.scalar("length"))
        act = o.get("assem.act")[:n].astype(np.int64)
        lev = o.get("assem.lev")[:n].astype(np.int64)
        L = int(o.scalar("L"))
        code = act - np.array([oracle.lvl_idx(int(l)) for l in lev])
        x = np.array([oracle.compact(int(c)) for c in code])
        west = (x == 0) & (lev == L)
        assert west.sum() == 16
        for t in (10.0, 14.0, 14.75, 16.3):
            o.set_scalar("time", t)
            o.stage("load_soln")
            h = o.get("assem.h0")[:n].copy()
            qx = o.get("assem.qx0")[:n].copy()
            o.stage("ghost")
            assert np.array_equal(o.get("assem.h0")[:n], h)
            s = 0.6 * np.exp(-0.25 * (t - 14.75) ** 2)
            v = np.where(h > 1e-3, qx / np.where(h > 0, h, 1.0), 0.0)
            gh, gq = o.get("nb3.h0")[:n], o.get("nb3.qx0")[:n]
            assert np.array_equal(gh[west], h[west])
            assert np.allclose(gq[west], (2.0 * s - v[west]) * h[west], rtol=1e-13, atol=1e-15)
            assert np.any(h[west] > 1e-3)
            # the other west-border ghosts (coarser leaves: none here) and the
            # east border are closed walls: zero normal discharge
            east = o.get("nb1.act")[:n] == -1
            assert np.all(o.get("nb1.qx0")[:n][east] == 0.0)
    finally:
        o.close()


def _c_prop_1d(a):
    """Bed of the 1D c-property cases (synthetic/topo_c_prop_1D.cuh), from its
    piecewise definition: two ramps meeting at a = 25, a parabolic bump centred
    at a = 10 and a plateau on (39, 46.5), all scaled by 10."""
    if 22 <= a < 25:
        return 10 * (0.05 * a - 1.1)
    if 25 <= a <= 28:
        return 10 * (-0.05 * a + 1.4)
    if 8 < a < 12:
        return 10 * (0.2 - 0.05 * (a - 10) ** 2)
    if 39 < a < 46.5:
        return 3.0
    return 0.0


@pytest.mark.parametrize("tc,eta", [(1, 6.0), (2, 6.0), (19, 1.0), (21, 1.95), (22, 1.78)])
def test_synthetic_initial_state(tmp_path, tc, eta):
    """Built-in cases (synthetic/*.cuh, mra/modal_projections.cu:28-88): the
    c-property cases start as a lake at rest, eta = h + z constant (its
    slopes zero) and q = 0; cases 1 and 2 are the same bed along x and y,
    and a cell's bed mean is the mean of the bed at its 4 corners."""
    par = cases.synthetic(str(tmp_path / f"s{tc}"), tc, L=5)
    o = oracle.Oracle(par, "f64")
    try:
        N = 1 << 10
        fl = oracle.lvl_idx(5)
        e0 = o.get("eta0")[fl:fl + N]
        assert np.allclose(e0, eta, atol=1e-13, rtol=0)
        for f in ("eta1x", "eta1y", "qx0", "qy0", "qx1x", "qy1y"):
            assert np.allclose(o.get(f)[fl:fl + N], 0.0, atol=1e-13), f
        z0 = o.get("z0")[fl:fl + N]
        dx = 50.0 / 32
        if tc in (1, 2):
            for code in range(N):
                x, y = oracle.compact(code), oracle.compact(code >> 1)
                a = (x if tc == 1 else y)
                exp = 0.5 * (_c_prop_1d(a * dx) + _c_prop_1d((a + 1) * dx))
                assert abs(z0[code] - exp) < 1e-12, (code, z0[code], exp)
            zs = o.get("z1y" if tc == 1 else "z1x")[fl:fl + N]
            assert np.all(zs == 0.0)
        assert np.isclose(float(o.scalar("dx")), {1: 50, 2: 50, 19: 60, 21: 75, 22: 75}[tc] / 32, rtol=0, atol=1e-15)
    finally:
        o.close()


@pytest.mark.parametrize("solver", ["mwdg2", "hwfv1"])
@pytest.mark.parametrize("tc", [1, 2, 3, 4, 19, 21, 22])
def test_c_property_lake_at_rest(tmp_path, solver, tc):
    """The reference's c-property suite (cases/tests/test.py:105, plotted
    against 1e-10 at :396): a lake at rest over the 1D, three-cone,
    differentiable and non-differentiable beds, wet and wet/dry, stays at rest
    through adaptive steps (eps = 1e-3): every discharge coefficient stays
    below 1e-10 in FP64."""
    par = cases.synthetic(str(tmp_path / f"c{tc}{solver}"), tc, L=5, solver=solver, t_end=20.0)
    o = oracle.Oracle(par, "f64")
    try:
        steps = o.steps(60)
        assert steps >= 30
        n = int(o.scalar("length"))
        assert n < 1024                                    # the grid coarsened
        names = ["qx0", "qy0"] + (["qx1x", "qx1y", "qy1x", "qy1y"] if solver == "mwdg2" else [])
        worst = max(float(np.abs(o.get("assem." + f)[:n]).max()) for f in names)
        assert worst < 1e-10, worst
    finally:
        o.close()


def test_cube_root_against_libm():
    """The cube root that stands in for pow(h, 1/3) in the Manning term and the
    FREE ghost depth (operators/apply_friction.cuh:41,
    neighbours/add_ghost_cells.cu:135): within 4 ulp of numpy's cbrt in FP64
    and 2 ulp of the exact value in FP32, over depths from 1e-12 m to 1e4 m,
    subnormals, zero, and the error values."""
    rng = np.random.default_rng(11)
    xs = np.concatenate([10.0 ** rng.uniform(-12, 4, 6000), rng.uniform(1e-3, 3.0, 3000),
                         [1.0, 8.0, 27.0, 1e-310, 5e-324, 1e300]])
    for x in xs:
        got, ref = oracle.cbrt(float(x)), float(np.cbrt(x))
        assert abs(got - ref) <= 4 * np.spacing(ref), (x, got, ref)
        if x < 3e38:
            xf = np.float32(x)
            gf, rf = np.float32(oracle.cbrt(float(xf), np.float32)), np.cbrt(np.float64(xf))
            if xf > 0:
                assert abs(float(gf) - rf) <= 2 * float(np.spacing(np.float32(rf))), (x, gf, rf)
    assert oracle.cbrt(0.0) == 0.0 and np.isnan(oracle.cbrt(-1.0))
    assert oracle.cbrt(8.0) == 2.0 and oracle.cbrt(27.0, np.float32) == 3.0

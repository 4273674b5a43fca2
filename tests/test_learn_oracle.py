"""CPU checks of the learned-refinement restatements in oracle/mlp_oracle.py:
the network's backward against central finite differences of its forward
(float64, no rounding), the BF16 rounding against torch's bfloat16
conversion, the unit neighbour map against a 2-D grid, and the labels
against a direct loop."""
import numpy as np
import pytest

import mlp_oracle as mo


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, mo.N_PARAM)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(1)
    x = rng.normal(0, 1, (40, mo.NF))
    y = (rng.random(40) < 0.3).astype(np.float64)
    loss, g = mo.loss_grad(p, x, y, pos_weight=3.0, rounding=False)
    assert np.isfinite(loss) and loss > 0
    idx = np.concatenate([rng.choice(np.arange(mo.P_W1, mo.P_B1), 12), rng.choice(np.arange(mo.P_B1, mo.P_W2), 4),
                          rng.choice(np.arange(mo.P_W2, mo.P_B2), 12), rng.choice(np.arange(mo.P_B2, mo.P_W3), 4),
                          rng.choice(np.arange(mo.P_W3, mo.P_B3), 6), [mo.P_B3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, x, y, 3.0, False)[0] - mo.loss_grad(pm, x, y, 3.0, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_is_weighted_cross_entropy():
    """softplus(z) - y z equals -log sigmoid(z) for y = 1 and -log(1 - sigmoid(z)) for y = 0."""
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.normal(0, 1, (25, mo.NF))
    y = (rng.random(25) < 0.5).astype(np.float64)
    f = mo.forward(p, x, rounding=False)
    q = f["p"]
    ce = -(2.0 * y * np.log(q) + (1 - y) * np.log(1 - q))
    loss, _ = mo.loss_grad(p, x, y, pos_weight=2.0, rounding=False)
    assert abs(loss - ce.mean()) <= 1e-12 * max(1.0, abs(loss))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    a = np.concatenate([rng.normal(0, 1, 5000).astype(np.float32), rng.normal(0, 1e-20, 100).astype(np.float32),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -1.0 - 2 ** -8, 0.0, 65504.0])])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mo.bf16(a), ref)
    assert mo.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0          # tie to even
    assert mo.bf16(np.float32([1.0 + 3 * 2 ** -9]))[0] == 1.0 + 2 ** -7


def test_unit_neighbours_on_the_grid():
    n = 1 << mo.UNIT_LEVEL
    code = np.zeros((n, n), dtype=np.int64)        # code[y, x]
    for yy in range(n):
        for xx in range(n):
            c = 0
            for b in range(mo.UNIT_LEVEL):
                c |= ((xx >> b) & 1) << (2 * b) | ((yy >> b) & 1) << (2 * b + 1)
            code[yy, xx] = c
    for d, (dx, dy) in enumerate([(0, 1), (1, 0), (0, -1), (-1, 0)]):
        nb = mo.unit_neighbours(d)
        for yy in (0, 5, n - 1):
            for xx in (0, 17, n - 1):
                u = code[yy, xx]
                xn, yn = xx + dx, yy + dy
                exp = code[yn, xn] if 0 <= xn < n and 0 <= yn < n else u
                assert nb[u] == exp, (d, xx, yy)


def test_labels_direct_loop():
    rng = np.random.default_rng(11)
    chg = (rng.random((20, mo.N_UNITS)) < 0.05).astype(np.uint8)
    y = mo.labels(chg, 8)
    assert y.shape == (13, mo.N_UNITS)
    for t in (0, 6, 12):
        for u in (0, 100, 4095):
            assert y[t, u] == int(any(chg[t + k, u] for k in range(8)))

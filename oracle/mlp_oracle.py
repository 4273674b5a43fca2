"""CPU restatement of the learned-refinement pieces (test infrastructure only:
the tests may import it as the checker; the product never does).

The learned refinement has no counterpart in the reference, which always
re-encodes and decodes every parent of the previous tree
(mra/encode_flow.cu:3-92, mra/decoding.cu:3-67); these restatements follow
the HIP library's own definitions.

- `forward` / `loss_grad`: the 48 -> 64 -> 64 -> 1 network (ReLU, sigmoid),
  weighted binary cross-entropy w * (softplus(z3) - y z3) with w = pos_weight
  for y = 1, mean over the batch; `rounding=True` rounds to BF16 (nearest
  even) where the GPU kernels do (weights, inputs, layer-1 activations and the
  back-propagated layer-2 and layer-1 deltas), in float64 arithmetic
  otherwise. The gradient passes through the roundings unchanged.
  Pinned by central finite differences of the float64 loss with
  `rounding=False` (tests/test_learn_oracle.py), so the backward formulas are
  checked against the forward, not restated.
- `unit_features`: the 48 features of each level-6 unit from the hierarchy's
  unit-node coefficients, the per-unit encode statistics, tree-change flags
  and ages.
- `labels`: y[t][u] = OR of the unit's tree changes over steps t .. t+H-1.
"""
from __future__ import annotations

import numpy as np

NF, NH = 48, 64
P_W1 = 0
P_B1 = P_W1 + NH * NF
P_W2 = P_B1 + NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NH
N_PARAM = P_B3 + 1
UNIT_LEVEL = 6
N_UNITS = 1 << (2 * UNIT_LEVEL)


def bf16(a) -> np.ndarray:
    """float32 -> nearest-even BF16, returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p, dtype=np.float32):
    p = np.asarray(p, dtype=dtype)
    return (p[P_W1:P_B1].reshape(NH, NF), p[P_B1:P_W2], p[P_W2:P_B2].reshape(NH, NH), p[P_B2:P_W3],
            p[P_W3:P_B3], p[P_B3])


def forward(p, x, rounding: bool = True) -> dict:
    dt = np.float32 if rounding else np.float64
    w1, b1, w2, b2, w3, b3 = unpack(p, dt)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=dt).reshape(-1, NF)
    xb = r(x).astype(np.float64)
    w1b, w2b, w3b = (r(w).astype(np.float64) for w in (w1, w2, w3))
    z1 = xb @ w1b.T + b1.astype(np.float64)
    a1 = np.maximum(z1, 0.0)
    a1b = r(a1.astype(np.float32)).astype(np.float64) if rounding else a1
    z2 = a1b @ w2b.T + b2.astype(np.float64)
    a2 = np.maximum(z2, 0.0)
    z3 = a2 @ w3b + float(b3)
    p3 = 1.0 / (1.0 + np.exp(-z3))
    return {"xb": xb, "w1b": w1b, "w2b": w2b, "w3b": w3b, "z1": z1, "a1b": a1b, "z2": z2, "a2": a2, "z3": z3,
            "p": p3}


def loss_grad(p, x, y, pos_weight: float = 1.0, rounding: bool = True):
    """Mean weighted BCE and its gradient (N_PARAM float64)."""
    f = forward(p, x, rounding)
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    w = np.where(y > 0.5, pos_weight, 1.0)
    z3 = f["z3"]
    sp = np.maximum(z3, 0.0) + np.log1p(np.exp(-np.abs(z3)))
    loss = float(np.sum(w * (sp - y * z3)) / n)
    g3 = w * (f["p"] - y) / n
    r = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if rounding else (lambda a: a)
    gw3 = f["a2"].T @ g3
    gb3 = g3.sum()
    dz2 = np.where(f["z2"] > 0, g3[:, None] * f["w3b"][None, :], 0.0)
    dz2b = r(dz2)
    gw2 = dz2b.T @ f["a1b"]
    gb2 = dz2.sum(axis=0)
    da1 = dz2b @ f["w2b"]
    dz1 = np.where(f["z1"] > 0, da1, 0.0)
    dz1b = r(dz1)
    gw1 = dz1b.T @ f["xb"]
    gb1 = dz1.sum(axis=0)
    g = np.zeros(N_PARAM)
    g[P_W1:P_B1] = gw1.ravel()
    g[P_B1:P_W2] = gb1
    g[P_W2:P_B2] = gw2.ravel()
    g[P_B2:P_W3] = gb2
    g[P_W3:P_B3] = gw3
    g[P_B3] = gb3
    return loss, g


def _compact(c):
    c = np.asarray(c, dtype=np.int64) & 0x55555555
    c = (c | (c >> 1)) & 0x33333333
    c = (c | (c >> 2)) & 0x0F0F0F0F
    c = (c | (c >> 4)) & 0x00FF00FF
    c = (c | (c >> 8)) & 0x0000FFFF
    return c


def _dilate(v):
    v = np.asarray(v, dtype=np.int64) & 0xFFFF
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    v = (v | (v << 1)) & 0x55555555
    return v


def unit_neighbours(direction: int) -> np.ndarray:
    """Unit code of the N (0), E (1), S (2), W (3) neighbour unit (the unit
    itself on the domain border)."""
    u = np.arange(N_UNITS)
    x, y = _compact(u), _compact(u >> 1)
    dx, dy = [(0, 1), (1, 0), (0, -1), (-1, 0)][direction]
    xn, yn = x + dx, y + dy
    n = 1 << UNIT_LEVEL
    inside = (xn >= 0) & (yn >= 0) & (xn < n) & (yn < n)
    return np.where(inside, _dilate(np.clip(xn, 0, n - 1)) | (_dilate(np.clip(yn, 0, n - 1)) << 1), u)


def unit_features(coef: dict, zc: dict, ustat, chg, age, maxes, L: int) -> np.ndarray:
    """coef[name] / zc[name]: the unit nodes' coefficients (4096 values each,
    names eta0 eta1x eta1y qx0 ... qy1y and z0 z1x z1y); maxes as the step's
    encode uses them (eta, h, qx, qy, z). Returns float32 (4096, 48) before
    the BF16 rounding."""
    f32 = np.float32
    mh, mq, mr, mz = (f32(maxes[k]) for k in (1, 2, 3, 4))
    st = np.asarray(ustat, dtype=np.float32).reshape(N_UNITS, 6)
    chg = np.asarray(chg).astype(np.float32)
    age = np.asarray(age, dtype=np.int64)
    F = np.zeros((N_UNITS, NF), dtype=np.float32)
    lg = lambda r: np.log2(f32(1) + r.astype(np.float32)).astype(np.float32)
    for k in range(4):
        F[:, k] = lg(st[:, k])
    modes = ("0", "1x", "1y")
    for m in range(3):
        z = zc["z" + modes[m]].astype(np.float32)
        F[:, 4 + m] = (coef["eta" + modes[m]].astype(np.float32) - z) / mh
        F[:, 7 + m] = coef["qx" + modes[m]].astype(np.float32) / mq
        F[:, 10 + m] = coef["qy" + modes[m]].astype(np.float32) / mr
        F[:, 13 + m] = z / mz
    for d in range(4):
        un = unit_neighbours(d)
        z = zc["z0"].astype(np.float32)[un]
        F[:, 16 + 4 * d] = (coef["eta0"].astype(np.float32)[un] - z) / mh
        F[:, 17 + 4 * d] = coef["qx0"].astype(np.float32)[un] / mq
        F[:, 18 + 4 * d] = coef["qy0"].astype(np.float32)[un] / mr
        F[:, 19 + 4 * d] = z / mz
        F[:, 32 + 2 * d] = lg(st[un, 0])
        F[:, 33 + 2 * d] = lg(st[un, 1])
        F[:, 44 + d] = chg[un]
    parents = ((1 << (2 * (L - UNIT_LEVEL))) - 1) // 3
    F[:, 40] = st[:, 4] / f32(parents)
    F[:, 41] = st[:, 5] / f32(1 << (2 * (L - 1 - UNIT_LEVEL)))
    F[:, 42] = chg
    F[:, 43] = np.minimum(age, 64).astype(np.float32) / f32(64)
    return F


def labels(chg, horizon: int) -> np.ndarray:
    """chg: (steps, 4096) 0/1 -> labels (steps - horizon + 1, 4096)."""
    c = np.asarray(chg).astype(bool)
    T = c.shape[0]
    out = np.zeros((T - horizon + 1, c.shape[1]), dtype=np.uint8)
    for t in range(T - horizon + 1):
        out[t] = c[t:t + horizon].any(axis=0)
    return out

"""Helpers shared by the GPU parity tests: run the HIP product and the CPU
oracle on the same input files and compare their states after each step."""
from __future__ import annotations

import numpy as np

import oracle


def lvl_idx(l: int) -> int:
    return ((1 << (2 * l)) - 1) // 3 + 3


def oracle_leaf_state(o: oracle.Oracle, mw: bool):
    """Leaf values of the oracle in the hierarchy's eta form (reinsert_assem_sol)."""
    n = int(o.scalar("length"))
    out = {}
    modes = ("0", "1x", "1y") if mw else ("0",)
    for m in modes:
        h = o.get(f"assem.h{m}")[:n]
        z = o.get(f"assem.z{m}")[:n]
        out[f"eta{m}"] = h + z                  # same IEEE add as the reinsert
        out[f"qx{m}"] = o.get(f"assem.qx{m}")[:n]
        out[f"qy{m}"] = o.get(f"assem.qy{m}")[:n]
    return out


def compare_step(o: oracle.Oracle, s, mw: bool, tol: float = 0.0, label: str = ""):
    """Returns a dict of mismatch counts / max errors between oracle and GPU."""
    rep = {}
    leaves_o = o.leaves()
    leaves_g = s.get("leaves")
    rep["n_leaves"] = (len(leaves_o), len(leaves_g))
    assert len(leaves_o) == len(leaves_g), f"{label}: leaf count {len(leaves_o)} vs {len(leaves_g)}"
    assert np.array_equal(leaves_o, leaves_g), f"{label}: leaf lists differ at {np.flatnonzero(leaves_o != leaves_g)[:10]}"
    nb_o = o.neighbours()
    for k in range(4):
        g = s.get(f"nbr{k}")
        bad = np.flatnonzero(nb_o[k] != g)
        assert bad.size == 0, f"{label}: neighbour {k} differs at {bad[:10]}: {nb_o[k][bad[:5]]} vs {g[bad[:5]]}"
    D = int(o.scalar("D"))
    sig_o = o.get("sig")[3:D]          # entries 0..2 are the index pad
    sig_g = s.get("sig")[3:D]
    assert np.array_equal(sig_o, sig_g), f"{label}: masks differ at {np.flatnonzero(sig_o != sig_g)[:10]}"
    st = oracle_leaf_state(o, mw)
    worst = 0.0
    for name, ov in st.items():
        gv = s.get("A." + name)[leaves_g]
        if tol == 0.0:
            bad = np.flatnonzero(ov != gv)
            assert bad.size == 0, (f"{label}: {name} differs at {bad[:5]}: oracle {ov[bad[:3]]} gpu {gv[bad[:3]]}")
        else:
            err = float(np.max(np.abs(ov.astype(np.float64) - gv.astype(np.float64)))) if ov.size else 0.0
            worst = max(worst, err)
            assert err <= tol, f"{label}: {name} max |diff| {err} > {tol}"
    rep["max_err"] = worst
    c = s.counters()
    dt_o = o.scalar("dt")
    if tol == 0.0:
        assert c["dt_next"] == dt_o, f"{label}: dt {c['dt_next']!r} vs oracle {dt_o!r}"
    else:
        assert abs(c["dt_next"] - dt_o) <= tol * max(1.0, abs(dt_o)), f"{label}: dt {c['dt_next']} vs {dt_o}"
    assert abs(c["time"] - o.scalar("time")) <= (0 if tol == 0 else tol), f"{label}: time"
    return rep


def run_pair(par: str, precision: str, steps: int, tol: float = 0.0, check_every: int = 1,
             t0: float | None = None):
    """Steps both and compares after every check_every steps; t0 sets the
    start time of both (for time-dependent boundaries)."""
    import mwdg2
    o = oracle.Oracle(par, precision)
    s = mwdg2.Simulation(par, precision)
    mw = int(o.scalar("solver")) == 2
    try:
        L = int(o.scalar("L"))
        fl = lvl_idx(L)
        # initial finest level and z hierarchy
        for name in ("eta0", "qx0", "qy0") + (("eta1x", "eta1y") if mw else ()):
            a = o.get(name)[fl:]
            b = s.get("A." + name)[fl:]
            assert np.array_equal(a, b), f"initial {name} differs"
        for name in ("z0",) + (("z1x", "z1y") if mw else ()):
            assert np.array_equal(o.get(name), s.get(name)), f"z hierarchy {name} differs"
        assert np.array_equal(o.get("preflag")[3: int(o.scalar("D"))], s.get("preflag")[3: int(o.scalar("D"))]), "preflags differ"
        if t0 is not None:
            o.set_scalar("time", t0)
            s.set_scalar("time", t0)
        reps = []
        for k in range(steps):
            more = o.step()
            s.step(1)
            if not more:
                break
            if (k + 1) % check_every == 0:
                reps.append(compare_step(o, s, mw, tol, label=f"step {k + 1}"))
        return reps
    finally:
        s.close()
        o.close()

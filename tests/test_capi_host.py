"""The C-ABI library without a GPU: it loads, exports every entry point that
include/mwdg2.h declares, and its host-side input validation (the .par /
raster / .bci / .bdy / .stage readers, which run before any device call)
reports bad inputs as MW_ERR_INPUT / MW_ERR_ARG with a message instead of
exiting, unlike the reference readers (input/*.cpp, classes/Boundary.h)."""
import ctypes
import os
import re
import subprocess

import pytest

import cases
import mwdg2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mwdg2.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mw_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("mw_create", "mw_destroy", "mw_step", "mw_run_until", "mw_encode", "mw_decode",
              "mw_extract_leaves", "mw_update_stage", "mw_cfl_min", "mw_download", "mw_run_simulation"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = mwdg2.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", mwdg2.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared_functions()) <= exported


def test_version_and_error_string():
    lib = mwdg2.lib()
    assert lib.mw_version().decode().startswith("mwdg2")
    assert isinstance(lib.mw_last_error(), bytes)


def _create(par, prec=8, argv=()):
    args = (ctypes.c_char_p * max(1, len(argv)))(*[a.encode() for a in argv])
    cfg = mwdg2.Config(par.encode() if par is not None else None, prec, len(argv), args)
    h = ctypes.c_void_p()
    rc = mwdg2.lib().mw_create(ctypes.byref(cfg), ctypes.byref(h))
    return rc, mwdg2.lib().mw_last_error().decode()


def test_argument_errors():
    rc, msg = _create(None)
    assert rc == 3
    rc, msg = _create("x.par", prec=2)
    assert rc == 3
    assert mwdg2.lib().mw_step(None, 1) == 3
    assert mwdg2.lib().mw_sort_pairs(-1, None, None, None, None) == 3


def test_learned_and_ensemble_argument_errors():
    """The network, learned-step and ensemble entry points refuse bad
    arguments before touching the device."""
    import ctypes
    lib = mwdg2.lib()
    assert lib.mw_mlp_param_count() == 64 * 48 + 64 + 64 * 64 + 64 + 64 + 1
    assert lib.mw_mlp_create(1, None) == 3
    assert lib.mw_mlp_forward(None, None, 1, None) == 3
    assert lib.mw_mlp_gradient(None, None, None, 1, ctypes.c_float(1.0), None, None) == 3
    assert lib.mw_mlp_record(None, None, 1) == 3
    assert lib.mw_mlp_build_labels(None, 0, None) == 3
    assert lib.mw_mlp_fit(None, 1, 64, ctypes.c_float(1e-3), ctypes.c_float(1.0), None) == 3
    assert lib.mw_mlp_eval(None, ctypes.c_float(0.5), None) == 3
    assert lib.mw_set_learned(None, None, 8, 32, ctypes.c_float(0.5)) == 3
    assert lib.mw_learn_counters(None, None) == 3
    assert lib.mw_set_sweep_groups(None, 64) == 3
    assert lib.mw_get_sweep_groups(None, None) == 3
    h = ctypes.c_void_p()
    assert lib.mw_ensemble_create(None, 0, 0, ctypes.byref(h)) == 3
    assert lib.mw_ensemble_step(None, 1) == 3
    assert lib.mw_ensemble_sync(None) == 3
    assert lib.mw_ensemble_destroy(None) == 0
    assert lib.mw_mlp_destroy(None) == 0


def _par(tmp_path, name, edit=None, files=None, **kw):
    par = cases.random_case(str(tmp_path / name), L=4, seed=1, **kw)
    if edit:
        txt = open(par).read()
        for a, b in edit:
            txt = txt.replace(a, b)
        open(par, "w").write(txt)
    for fname, content in (files or {}).items():
        with open(os.path.join(os.path.dirname(par), fname), "w") as f:
            f.write(content)
    return par


@pytest.mark.parametrize("name,edit,files,needle", [
    ("missing_par", None, None, "cannot open"),
    ("level", [("max_ref_lvl   4", "max_ref_lvl   15")], None, "max_ref_lvl"),
    ("testcase", [("test_case     0", "test_case     24")], None, "[0, 23]"),
    ("testcase20", [("test_case     0", "test_case     20")], None, "non-square"),
    ("solver", [("mwdg2\n", "\n")], None, "solver"),
    ("dem", [("case.dem", "nothere.dem")], None, "cannot open"),
    ("bigdem", None, {"case.dem": "ncols 40\nnrows 4\nxllcorner 0\nyllcorner 0\ncellsize 1\nNODATA_value -9999\n"
                      + "0 " * 160}, "fit the mesh"),
    ("shortdem", None, {"case.dem": "ncols 16\nnrows 16\nxllcorner 0\nyllcorner 0\ncellsize 1\n"
                        "NODATA_value -9999\n1 2 3\n"}, "short raster"),
])
def test_input_errors(tmp_path, name, edit, files, needle):
    par = str(tmp_path / "nope.par") if name == "missing_par" else _par(tmp_path, name, edit, files)
    rc, msg = _create(par)
    assert rc == 1, (rc, msg)
    assert needle in msg, msg


def test_boundary_and_source_errors(tmp_path):
    par = _par(tmp_path, "bdy", bci="W 0 15 QVAR inflow\n")
    rc, msg = _create(par)
    assert rc == 1 and "bdyfile" in msg
    par = _par(tmp_path, "bdy2", bci="W 0 15 QVAR other\n", bdy="inflow\n2 seconds\n0 0\n1 1\n")
    rc, msg = _create(par)
    assert rc == 1 and "not in" in msg
    par = _par(tmp_path, "src", bci="P 99.5 1.5 QFIX 1.0\n")
    rc, msg = _create(par)
    assert rc == 1 and "outside the mesh" in msg

import glob
import hashlib
import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
PKG_DIR = os.path.join(REPO, "inria-internship-project-sw2d-enzyme_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, GOLDEN)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def load_pkg():
    if "sw2d_amd" in sys.modules:
        return sys.modules["sw2d_amd"]
    spec = importlib.util.spec_from_file_location("sw2d_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["sw2d_amd"] = m
    spec.loader.exec_module(m)
    return m


@pytest.fixture(scope="session")
def sw():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


def golden_names():
    """Fixtures of the forward solver made by make_golden.py from the compiled
    reference (the e_* ones belong to the 4D-Var layer, sp_toce to its own tests)."""
    from make_golden import golden_cases
    names = set(golden_cases())
    return sorted(n for n in (os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
                  if n in names)


def e_golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "e_*.npz")))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def regen_case(sw, name, folder):
    """Rebuild the inputs of a golden case and check they are byte-identical."""
    from make_golden import e_golden_cases, golden_cases
    gen, kw = (golden_cases()[name] if name in golden_cases() else e_golden_cases()[name][:2])
    from sw2d_amd import cases
    getattr(cases, gen)(folder, **kw)
    g = load_golden(name)
    for fname, dig in zip(g["digest_names"], g["digest_values"]):
        with open(os.path.join(folder, "input", str(fname)), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == str(dig), f"{name}: {fname} differs"
    return g


def time_of(result_file):
    """Time (s) encoded in a ResultHydro_pDDDDdHHhMMmSSsMMM.txt name (swIO.cpp:1194-1226)."""
    s = result_file.split("_p")[1]
    d = int(s[0:4]); h = int(s[5:7]); m = int(s[8:10]); sec = int(s[11:13]); ms = int(s[14:17])
    return d * 86400 + h * 3600 + m * 60 + sec + ms / 1000.0


def map_time(g, k):
    """Time of map k of a golden fixture: from its file name, or from the
    fixture's own record where the name cannot carry it (sub-millisecond
    runs: every map of those is named for t = 0)."""
    if "times" in g:
        return float(g["times"][k])
    return time_of(str(g["files"][k]))


def rel_err(a, b):
    """max |a-b| / max |b| (field-normalised), 0 when both are 0."""
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def load_sp_toce(sw):
    """The sp_toce fixture (tests/golden/make_sp_toce.py): an in-memory case in
    reference order plus its observation frames, assimilation settings and the
    reference's first descent-log row."""
    from sw2d_amd import _native as N
    from sw2d_amd.cases import MemCase
    g = dict(np.load(os.path.join(GOLDEN, "sp_toce.npz"), allow_pickle=False))
    arrays = {k[2:]: g[k] for k in g if k.startswith("a_")}
    p = N.Params()
    for k in g:
        if k.startswith("p_"):
            setattr(p, k[2:], g[k].item())
    case = MemCase(arrays, p, int(g["n_series"]), int(g["n_times"]))
    return case, g


def sp_toce_folder(sw, folder, probe_variant=False, frames=True):
    """Rebuild the reference's functional case sp_toce in `folder`: input/
    from tests/golden/sp_toce_input.npz (the case's files byte for byte) and
    reference/ResultHydro_p*.txt observation frames from sp_toce.npz (the
    reference's frames in the assimilation window, re-printed at their 11
    significant digits, which reproduces the values exactly).  probe_variant:
    the native 2 s probe run of make_sp_toce_case.py instead (no frames)."""
    g = np.load(os.path.join(GOLDEN, "sp_toce_input.npz"), allow_pickle=False)
    inp = os.path.join(folder, "input")
    os.makedirs(inp, exist_ok=True)
    for k in g.files:
        if k.startswith("in_"):
            with open(os.path.join(inp, k[3:]), "wb") as f:
                f.write(g[k].tobytes())
    if probe_variant:
        from make_sp_toce_case import probe_variant as pv
        pv(folder)
        return folder
    if frames:
        t = dict(np.load(os.path.join(GOLDEN, "sp_toce.npz"), allow_pickle=False))
        case = sw.Case(folder)
        ref = os.path.join(folder, "reference")
        os.makedirs(ref, exist_ok=True)
        for tk, fr in zip(t["obs_t"], t["frames"]):
            case.write_result(os.path.join(ref, sw.result_name(float(tk))), fr[0], fr[1], fr[2],
                              n_decimal=10)
    return folder

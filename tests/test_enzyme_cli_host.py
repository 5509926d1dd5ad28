"""Argument handling of the drop-in `sw2d_enzyme_gpu` that happens before any
device call (assim/main_enzyme.cpp:37-59): usage, help, version, unknown
modes, and the refused classic gen-ref (the reference's classic generation
stores no reference map, its native time loop never writes one,
swSpE_minimize.cpp:111-121)."""
import os
import subprocess

from conftest import PKG_DIR

EXE = os.path.join(PKG_DIR, "lib", "sw2d_enzyme_gpu")


def run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


def test_usage_help_and_version():
    r = run()
    assert r.returncode != 0 and "usage" in r.stderr
    assert run("-h").returncode == 0
    r = run("--version")
    assert r.returncode == 0 and "gfx950" in r.stdout
    assert run("/nonexistent/folder").returncode != 0


def test_unknown_mode_and_classic_gen_ref_are_refused(tmp_path):
    d = str(tmp_path)
    r = run(d, "bogus")
    assert r.returncode != 0 and "unknown mode" in r.stderr
    r = run(d, "gen-ref")
    assert r.returncode != 0 and "gen-ref-e" in r.stderr

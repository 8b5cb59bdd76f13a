"""The reference-side binding (meraxes-devs_qyxnew_amd/shim/mxr_meraxes.c) compiles as strict C99
with every warning an error against declarations of the run_globals fields it reads (the
stand-ins in tests/shim_env keep the reference's names and types), links against libmxr.so, and
its host-side entry point RtoM returns the library's value.  No GPU is used."""
import os
import subprocess

import mxr

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(REPO, "meraxes-devs_qyxnew_amd", "shim", "mxr_meraxes.c")
ENV = os.path.join(REPO, "tests", "shim_env")


def _cc(args, timeout=120):
    r = subprocess.run(["gcc", *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r


def test_shim_builds_strict_c99_both_name_sets(tmp_path):
    inc = ["-I", ENV, "-I", os.path.join(REPO, "include")]
    for extra in ([], ["-DMXR_REFERENCE_NAMES"]):
        _cc(["-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", *inc, *extra, "-c", SHIM, "-o",
             str(tmp_path / "shim.o")])


def test_shim_links_and_rtom_matches(tmp_path):
    inc = ["-I", ENV, "-I", os.path.join(REPO, "include")]
    _cc(["-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", *inc, "-DMXR_REFERENCE_NAMES", "-c", SHIM, "-o",
         str(tmp_path / "shim.o")])
    _cc(["-std=gnu99", "-Wall", "-Wextra", "-Werror", *inc, "-c", os.path.join(ENV, "host.c"), "-o",
         str(tmp_path / "host.o")])
    libdir = os.path.dirname(mxr.LIB_PATH)
    libname = os.path.basename(mxr.LIB_PATH)[3:-3]
    exe = tmp_path / "host"
    _cc([str(tmp_path / "host.o"), str(tmp_path / "shim.o"), "-L", libdir, "-l" + libname, "-Wl,-rpath," + libdir,
         "-o", str(exe)])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "entry points 1" in r.stdout

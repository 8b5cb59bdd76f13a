"""include/mxr.h is the boundary a C host (the reference is C99) compiles against: a C99
translation unit including it builds warning-free with gcc, links against the library and calls
entry points that need no GPU (no GPU is used)."""
import os
import subprocess

import mxr

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

C_SRC = r"""
#include <stdio.h>
#include "mxr.h"

int main(void)
{
  const int n = mxr_ps_length(64, 67.8, 0.678);
  const char* s = mxr_strerror(MXR_ERR_INVALID_ARG);
  printf("%d %s\n", n, s);
  return n > 0 && s != NULL ? 0 : 1;
}
"""


def test_c99_program_builds_and_links(tmp_path):
    src = tmp_path / "host.c"
    src.write_text(C_SRC)
    exe = tmp_path / "host"
    libdir = os.path.dirname(mxr.LIB_PATH)
    libname = os.path.basename(mxr.LIB_PATH)[3:-3]  # libmxr.so -> mxr
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(REPO, "include"),
           str(src), "-L", libdir, "-l" + libname, "-Wl,-rpath," + libdir, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n, s = r.stdout.split(" ", 1)
    assert int(n) == mxr.ps_length(64, 67.8, 0.678)
    assert s.strip() == mxr.lib().mxr_strerror(mxr.MXR_ERR_INVALID_ARG).decode()

"""The C ABI: the library loads, exports every entry point declared in
include/rbfx.h, and the ctypes mirror of rbfx_problem has the C layout.
No compute calls (these tests run without a GPU)."""
import os
import re
import subprocess
import tempfile

import pytest

from updes_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rbfx.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_python_binds():
    assert set(_declared()) <= set(L.EXPORTED), set(_declared()) - set(L.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("librbfx.so not built")
    lib = L.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.rbfx_abi_version() == 3
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rbfx_\w+)", out))
    assert set(_declared()) <= exported


def test_problem_struct_layout_matches_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "rbfx.h"
#define F(f) printf(#f " %zu\n", offsetof(rbfx_problem, f));
int main(void){
  printf("size %zu\n", sizeof(rbfx_problem));
  F(N) F(M) F(kernel) F(flags) F(param) F(d_x) F(d_y) F(Ni) F(Nd) F(Nn) F(Nr) F(Npd) F(d_nx) F(d_ny)
  F(d_robin_normal_x) F(d_robin_normal_y) F(d_beta) F(d_pd_i1) F(d_pd_i2) F(op_rbf_const) F(op_mon_const)
  F(d_op_rbf_row) F(d_op_mon_row) F(support_cols) F(d_support)
  printf("irsize %zu\n", sizeof(rbfx_ir_stats));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    vals = dict(line.split() for line in out.strip().splitlines())
    import ctypes
    assert int(vals.pop("size")) == ctypes.sizeof(L.Problem)
    assert int(vals.pop("irsize")) == ctypes.sizeof(L.IrStats)
    for name, off in vals.items():
        assert getattr(L.Problem, name).offset == int(off), name

"""The C++ API of include/ (op_seq.h -> op_lib.h -> op2_hip.h) from a C++
program: the reference's declaration-time checks (common/op_lib.h:160-190
type checks, common/op_lib_core.c:94-146 set and map checks) end the program
with the reference's messages before any device work, so no GPU is needed."""

import os
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "oplus2_matlab_amd", "lib")

PROGRAM = r"""
#include <cstring>
#include "op_seq.h"

int main(int argc, char **argv) {
  op_init_core(argc, argv, 0);  // op_init would also claim the device
  const char *c = argc > 1 ? argv[1] : "";
  if (!strcmp(c, "set_size")) op_decl_set(0, "empty");
  op_set nodes = op_decl_set(4, "nodes");
  op_set edges = op_decl_set(3, "edges");
  if (!strcmp(c, "map_range")) {
    int pp[6] = {0, 1, 1, 2, 2, 4};
    op_decl_map(edges, nodes, 2, pp, "ppedge");
  }
  if (!strcmp(c, "dat_type")) {
    double x[4] = {0, 1, 2, 3};
    op_decl_dat(nodes, 1, "float", x, "x");
  }
  if (!strcmp(c, "const_type")) {
    float alpha = 1.0f;
    op_decl_const2("alpha", 1, "double", &alpha);
  }
  printf("declarations passed\n");
  return 0;
}
"""


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB, "libop2hip.so")):
        pytest.skip("libop2hip.so not built")
    d = tmp_path_factory.mktemp("cxx")
    src = d / "decl.cpp"
    src.write_text(PROGRAM)
    exe = d / "decl"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", str(src), "-o", str(exe),
                    f"-L{LIB}", "-lop2hip", f"-Wl,-rpath,{LIB}"], check=True)
    return str(exe)


@pytest.mark.parametrize("case, code, message", [
    ("set_size", 255, "op_decl_set error -- negative/zero size for set: empty"),
    ("map_range", 255, "op_decl_map error -- invalid data for map ppedge"),
    ("dat_type", 1, 'incorrect type specified for dataset "x"'),
    ("const_type", 1, 'incorrect type specified for constant "alpha"'),
])
def test_declaration_checks(program, case, code, message):
    out = subprocess.run([program, case], capture_output=True, text=True, timeout=60)
    assert out.returncode == code, out.stdout + out.stderr
    assert message in out.stdout
    assert "declarations passed" not in out.stdout


def test_valid_declarations_pass(program):
    out = subprocess.run([program, "none"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "declarations passed" in out.stdout

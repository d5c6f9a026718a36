"""The C++ mirror header compiles against include/lsx.h and links liblsx.so; host-only calls
(option defaults, argument errors surfacing as lsx_exception) run without a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "ls-stag-3d_parallel_amd")

PROGRAM = r"""
#include <cstdio>
#include "lsx_iterative_solver.hpp"
int main() {
  lsx_options o = lsx_cpp::mg_options(LSX_KRYLOV_FGMRES, 30, 1000, 1e-8, 3, LSX_PREC_FP64);
  if (o.restart != 30 || o.mg != LSX_MG_CASCADE || o.n_post != 3) return 1;
  lsx_options m = lsx_cpp::mg_options(LSX_KRYLOV_BICGSTAB, 30, 1000, 1e-8, 3, LSX_PREC_MIXED);
  if (m.mg != LSX_MG_VCYCLE || m.n_pre != 2) return 2;
  try {
    lsx_cpp::check(lsx_set_csr(nullptr, 4, 4, 4, nullptr, nullptr, nullptr, 0, LSX_MEM_HOST));
    return 3;
  } catch (const lsx_cpp::lsx_exception& e) {
    if (e.status != LSX_ERR_ARG) return 4;
  }
  std::printf("ok\n");
  return 0;
}
"""


def test_cpp_mirror_compiles_links_and_runs(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "liblsx.so")):
        pytest.fail("liblsx.so not built")
    src = tmp_path / "mirror.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "mirror"
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", LIBDIR, "-llsx",
                        f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)

"""Host-side launch sizing of libbtpm (binarytpm_amd/csrc/runtime.h), on the CPU.

A dispatch's grid size is a 32-bit count of work-items per dimension, so the
runtime's grid helpers must never hand back a grid that silently wraps:
grid1() refuses counts that could pass 2^32 work-items, and grid_stride()
caps the workgroups of loops that cover arbitrarily long series.  A tiny host
program built against the header checks both, with hipcc's host pass (no GPU
code is run).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "binarytpm_amd", "csrc")

PROGRAM = r"""
#include "runtime.h"
#include <cstdint>
#include <cstdio>
using btpm::grid1;
using btpm::grid_stride;
static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s\n", #c); ++fails; } } while (0)
int main() {
  CHECK(grid1(1, 256) == 1);
  CHECK(grid1(256, 256) == 1);
  CHECK(grid1(257, 256) == 2);
  CHECK(grid1(0, 256) == 0);
  // the largest accepted count still fits 2^32 work-items after rounding up
  const size_t top = (size_t)UINT32_MAX - 1024u;
  CHECK((size_t)grid1(top, 256) * 256 <= (size_t)UINT32_MAX);
  bool threw = false;
  try { (void)grid1(top + 1, 256); } catch (const std::length_error&) { threw = true; }
  CHECK(threw);
  threw = false;
  // a 81920-facet x 100 x 656-step crater series, one work-item per element
  try { (void)grid1((size_t)81920 * 100 * 656, 256); } catch (const std::length_error&) { threw = true; }
  CHECK(threw);
  CHECK(grid_stride(0, 256) == 1);
  CHECK(grid_stride(1000, 256) == 4);
  CHECK(grid_stride((size_t)81920 * 100 * 656, 256) == 16384);
  CHECK((size_t)grid_stride((size_t)1 << 40, 256) * 256 <= (size_t)UINT32_MAX);
  std::printf("%s\n", fails ? "FAILED" : "OK");
  return fails ? 1 : 0;
}
"""


def test_grid_helpers_never_wrap(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    src = tmp_path / "grid.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "grid"
    rocm_include = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc))), "include")
    # host-only C++: the HIP headers need the platform named when not compiling HIP
    b = subprocess.run([hipcc, "-x", "c++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__",
                        "-I", rocm_include, "-I", CSRC, str(src), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr

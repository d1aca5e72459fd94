"""Host logic of the reference tables (csrc/hdg_tables.cpp): which local
elimination order the condensation kernel may use.  Natural order (no pivot
search) needs an SPD q block, i.e. positive cell weights and a
positive-definite reference mass; the reference's default rules give that at
k = 1 (Strang 3 points) and k = 3 (degree 6, 12 points) but not at k = 2
(degree 3: six points on a conic) or k = 4 (degree 6: 12 points for 15
functions), nor with the negative weights of the Grundmann-Moeller rule at
degree 7.  Compiled with g++; no device call."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include "hdg_tables.h"
int main() {
  const int cases[][2] = {{1, 2}, {1, 4}, {2, 3}, {2, 4}, {3, 4}, {3, 6}, {4, 6}, {4, 7}};
  for (auto& c : cases) {
    hdg::RefTables t;
    if (!t.build(c[0], c[1])) { std::printf("k%d q%d none\n", c[0], c[1]); continue; }
    std::printf("k%d q%d %d\n", c[0], c[1], t.mass_spd() ? 1 : 0);
  }
  return 0;
}
"""

EXPECT = {"k1 q2": "1", "k1 q4": "1", "k2 q3": "0", "k2 q4": "1", "k3 q4": "0", "k3 q6": "1",
          "k4 q6": "0", "k4 q7": "0"}


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_natural_order_only_where_the_mass_is_spd(tmp_path):
    src = tmp_path / "tables_check.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "tables_check"
    csrc = os.path.join(REPO, "hdiscontinuousgalerkin.jl_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", csrc, str(src), os.path.join(csrc, "hdg_tables.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout
    got = {" ".join(l.split()[:2]): l.split()[2] for l in out.strip().splitlines()}
    assert got == EXPECT, got

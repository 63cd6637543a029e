"""The hot-path CIC division X/dx = fma(fma(-q0, dx, X), 1/dx, q0) with q0 = X*(1/dx) (one Markstein
correction of a correctly rounded reciprocal, csrc/common.hpp div_dx) is the IEEE quotient: checked
bit for bit on the CPU (gcc, -ffp-contract=off, libm fma) on random positions, cell centres and
cell edges and their neighbouring doubles, for box sizes and meshes of the workloads."""
import shutil
import subprocess

import pytest

SRC = r"""
#include <stdio.h>
#include <math.h>
#include <stdint.h>
static uint64_t s = 88172645463325252ULL;
static double urand(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (s >> 11) * 0x1p-53; }
int main(void) {
  const double Ls[] = {1.0, 19.595917942265423, 3.0, 2.5, 7.3, 1e-3, 123.456};
  const int Ns[] = {8, 16, 32, 48, 64, 96, 128, 256, 512};
  long bad = 0, tot = 0;
  for (int a = 0; a < 7; ++a) for (int b = 0; b < 9; ++b) {
    const double L = Ls[a], dx = L / Ns[b], y = 1.0 / dx;
    for (long k = 0; k < 300000; ++k) {
      double x;
      double m = floor(urand() * Ns[b]);
      switch (k % 4) {
        case 0: x = urand() * L; break;
        case 1: x = (m + 0.5) * dx; x = (k & 8) ? nextafter(x, 0) : ((k & 16) ? nextafter(x, 1e9) : x); break;
        case 2: x = m * dx; if (k & 8) x = nextafter(x, 0); break;
        default: x = urand() * L * 1e-6;
      }
      const double q0 = x * y;
      const double q = fma(fma(-q0, dx, x), y, q0);
      if (q != x / dx) ++bad;
      ++tot;
    }
  }
  printf("%ld %ld\n", tot, bad);
  return bad != 0;
}
"""


def test_fma_division_is_ieee_quotient(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "div.c"
    src.write_text(SRC)
    exe = tmp_path / "div"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    tot, bad = map(int, out.stdout.split())
    assert tot == 7 * 9 * 300000 and bad == 0

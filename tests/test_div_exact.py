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


KICK_SRC = r"""
#include <stdio.h>
#include <math.h>
#include <stdint.h>
static uint64_t s = 0x9E3779B97F4A7C15ULL;
static double urand(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (s >> 11) * 0x1p-53; }
int main(void) {
  /* the divisors of the KDK schedule (pm3d.py:118-127): a*H = H_I a^(-1/2), a_mid^3 H_mid =
     H_I a_mid^(3/2), for a from 1e-3 to past 1e3; numerators of every sign and a wide exponent range */
  long bad = 0, tot = 0;
  for (int j = 0; j < 1500; ++j) {
    const double a = 1e-3 * pow(1.02, j * 0.47), am = a + 0.5 * 0.0125 * a;
    const double ys[3] = {a * (5.0 * pow(a, -1.5)), pow(am, 3) * (5.0 * pow(am, -1.5)), (a + 0.0125 * a) * (5.0 * pow(a + 0.0125 * a, -1.5))};
    for (int t = 0; t < 3; ++t) {
      const double y = ys[t], r = 1.0 / y;
      for (int k = 0; k < 4000; ++k) {
        double x = (urand() - 0.5) * ldexp(1.0, (int)(urand() * 120) - 90);
        if (k % 5 == 0) x = nextafter(x, -x);
        const double q0 = x * r;
        const double q = x == 0.0 ? q0 : fma(fma(-q0, y, x), r, q0);
        if (q != x / y) ++bad;
        ++tot;
      }
    }
  }
  printf("%ld %ld\n", tot, bad);
  return bad != 0;
}
"""


def test_reciprocal_division_of_kick_scalars_is_ieee_quotient(tmp_path):
    """The particle kernels' qdiv (csrc/common.hpp): x/y as q0 = x*r, r = 1/y correctly rounded on the
    host, plus one Markstein correction, bit for bit the IEEE quotient for the step's divisors."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "kdiv.c"
    src.write_text(KICK_SRC)
    exe = tmp_path / "kdiv"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    tot, bad = map(int, out.stdout.split())
    assert tot == 1500 * 3 * 4000 and bad == 0


MAGIC_SRC = r"""
#include <stdio.h>
#include <stdint.h>
/* csrc/common.hpp div32_make / div32: q = (umulhi(n, m) + n) >> s */
static void make(uint32_t d, uint32_t* m, uint32_t* sh) {
  uint32_t s = 0; while ((1ull << s) < d) ++s;
  *m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1); *sh = s;
}
int main(void) {
  const uint32_t Ns[] = {1, 2, 3, 8, 16, 32, 48, 64, 96, 128, 256, 512, 812};
  long bad = 0, tot = 0;
  for (int k = 0; k < 13; ++k)
    for (int sq = 0; sq < 2; ++sq) {
      const uint32_t d = sq ? Ns[k] * Ns[k] : Ns[k];
      uint32_t m, sh; make(d, &m, &sh);
      for (uint64_t n = 0; n < (1ull << 31); n += (n < (1u << 22) ? 1 : 997)) {
        const uint32_t q = (uint32_t)((((uint64_t)n * m) >> 32) + n) >> sh;
        if (q != n / d) ++bad;
        ++tot;
      }
    }
  printf("%ld %ld\n", tot, bad);
  return bad != 0;
}
"""


def test_magic_division_of_lattice_sites(tmp_path):
    """div32 (csrc/common.hpp), the particle index -> lattice site division of the particle kernels,
    equals integer division for every divisor Np and Np^2 of the supported sizes, on all indices
    below 2^22 and a stride of the rest of [0, 2^31)."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "magic.c"
    src.write_text(MAGIC_SRC)
    exe = tmp_path / "magic"
    subprocess.run([gcc, "-O2", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    tot, bad = map(int, out.stdout.split())
    assert tot > 0 and bad == 0


JAC_SRC = r"""
#include <stdio.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
static uint64_t s = 0xD1B54A32D192ED03ULL;
static double urand(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (s >> 11) * 0x1p-53; }
/* the four-site FP64 Jacobian's entry (csrc/particles.hpp min_image + qdiv) against the one-site
   kernel's (v - L*rint(v/L)) / dq, compared as bit patterns (the sign of zero included) */
static double min_image(double v, double L, double lo, double mid, double hi) {
  const double a = fabs(v);
  double k;
  if (a < lo) k = 0.0;
  else if (a > mid && a < hi) k = 1.0;
  else k = fabs(rint(v / L));
  return v - L * copysign(k, v);
}
int main(void) {
  const double Ls[] = {1.0, 19.595917942265423, 3.0, 2.5, 7.3, 1e-3, 123.456};
  const int Ns[] = {8, 16, 32, 48, 64, 96, 128, 256, 512, 812};
  long bad = 0, tot = 0;
  for (int a = 0; a < 7; ++a) for (int b = 0; b < 10; ++b) {
    const double L = Ls[a], dq = L / Ns[b], rdq = 1.0 / dq;
    const double lo = 0.49 * L, mid = 0.51 * L, hi = 1.49 * L;
    for (long k = 0; k < 200000; ++k) {
      double x1 = urand() * L, x2;
      switch (k % 6) {
        case 0: x2 = urand() * L; break;                                    /* any pair */
        case 1: x2 = fmod(x1 + dq * (1.0 + 0.3 * (urand() - 0.5)), L); break; /* lattice neighbours */
        case 2: x2 = x1 + dq * 1e-9 * (urand() - 0.5); break;                /* nearly coincident */
        case 3: x1 = urand() * dq; x2 = L - urand() * dq; break;             /* across the boundary */
        case 4: x2 = x1; break;                                              /* zero difference */
        default: {                                                           /* the band edges */
          const double e[6] = {lo, mid, hi, 0.5 * L, L, 1.5 * L};
          double v = e[k % 6 == 5 ? (k / 6) % 6 : 0] * (k & 64 ? -1.0 : 1.0);
          v = (k & 128) ? nextafter(v, 0) : ((k & 256) ? nextafter(v, 1e300 * v) : v);
          x1 = 0.0; x2 = v;
        }
      }
      const double v = x2 - x1;
      const double ref = (v - L * rint(v / L)) / dq;
      const double w = min_image(v, L, lo, mid, hi);
      const double q0 = w * rdq;
      const double got = w == 0.0 ? q0 : fma(fma(-q0, dq, w), rdq, q0);
      if (memcmp(&ref, &got, sizeof(double)) != 0) ++bad;
      ++tot;
    }
  }
  printf("%ld %ld\n", tot, bad);
  return bad != 0;
}
"""


def test_jacobian_entry_without_divisions_is_bitwise(tmp_path):
    """The FP64 Jacobian's division-free minimum image and reciprocal quotient equal the plain
    expression bit for bit (sign of zero included) over positions of every workload box and mesh."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "jac.c"
    src.write_text(JAC_SRC)
    exe = tmp_path / "jac"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    tot, bad = map(int, out.stdout.split())
    assert tot == 7 * 10 * 200000 and bad == 0


WRAP_SRC = r"""
#include <stdio.h>
/* the mixed Jacobian's minimum image of a cell difference (csrc/particles_mixed.hpp det_mx): the
   power-of-two form against the compare form, every dc in (-N, N) */
int main(void) {
  long bad = 0, tot = 0;
  for (int k = 1; k <= 10; ++k) {
    const int N = 1 << k, sh = 32 - k;
    for (int dc = -N + 1; dc < N; ++dc) {
      int r = dc;
      if (r > N / 2) r -= N;
      if (r < -N / 2) r += N;
      const int w = (int)((unsigned)dc << sh) >> sh;
      const int p = dc == N / 2 ? dc : w;
      if (p != r) ++bad;
      ++tot;
    }
  }
  printf("%ld %ld\n", tot, bad);
  return bad != 0;
}
"""


def test_mixed_minimum_image_power_of_two_form(tmp_path):
    """Sign extension of the low log2 N bits with the +N/2 tie kept equals the two-compare minimum
    image for every cell difference of meshes 2..1024."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "wrap.c"
    src.write_text(WRAP_SRC)
    exe = tmp_path / "wrap"
    subprocess.run([gcc, "-O2", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    tot, bad = map(int, out.stdout.split())
    assert tot == sum(2 * (1 << k) - 1 for k in range(1, 11)) and bad == 0

"""The FP64 kernels replace a / b by q = a*y and two Markstein corrections
(r = fma(-b, q, a); q = fma(r, y, q)) with y = RN(1/b).  This checks, on the
CPU with the same algorithm (C, fma from libm), that the result is bitwise
equal to IEEE division for the kernels' constant divisors, for divisors near
the densities and relaxation times they meet, for the window counts 1..64 of
the dynamic model's box means, and for random normal operands.
Plain multiplication by the reciprocal is shown to differ, so the test has
teeth."""
import math
import os
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ULL;
static uint64_t nx(void){ s ^= s >> 12; s ^= s << 25; s ^= s >> 27; return s * 0x2545F4914F6CDD1DULL; }
static double u01(void){ return (double)(nx() >> 11) / 9007199254740992.0; }
static double div2(double a, double b, double y){ double q = a*y; double r = fma(-b,q,a); q = fma(r,y,q); r = fma(-b,q,a); return fma(r,y,q); }
static double div0(double a, double b, double y){ (void)b; return a*y; }
static int same(double x, double y){ uint64_t a,b; memcpy(&a,&x,8); memcpy(&b,&y,8); return a==b || (x==0.0 && y==0.0); }
int main(int argc, char** argv){
  long n = atol(argv[1]);
  const double cs2 = 1.0/3.0;
  const double dens[4] = {cs2, 2.0*cs2*cs2, 2.0*cs2, cs2*cs2};
  long bad2 = 0, bad1 = 0;
  for (long i = 0; i < n; ++i) {
    double a = (u01()*2.0 - 1.0) * ldexp(1.0, (int)(nx() % 80) - 60);
    double b;
    int m = (int)(i % 8);
    if (m < 4) b = dens[m];
    else if (m == 4) b = 0.9 + 0.2*u01();
    else if (m == 5) b = 0.5 + 40.0*u01();
    else if (m == 6) b = (0.5 + u01()) * ldexp(1.0, (int)(nx() % 200) - 100);
    else b = (double)(1 + nx() % 64);  /* box-filter window counts */
    double y = 1.0 / b;
    double q = a / b;
    if (!same(div2(a,b,y), q)) ++bad2;
    if (!same(div0(a,b,y), q)) ++bad1;
  }
  printf("%ld %ld\n", bad2, bad1);
  return 0;
}
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    d = tmp_path_factory.mktemp("div")
    c = d / "div.c"
    c.write_text("#include <stdlib.h>\n" + SRC)
    exe = d / "div"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(c), "-o", str(exe), "-lm"], check=True)
    return str(exe)


def test_two_step_fma_division_is_ieee(prog):
    out = subprocess.run([prog, "4000000"], capture_output=True, text=True, check=True).stdout
    bad2, bad1 = map(int, out.split())
    assert bad2 == 0
    assert bad1 > 1000  # plain reciprocal multiplication is not IEEE division


# ---------------------------------------------------------------------------
# One correction for the constant divisors cs2, 2 cs2^2, 2 cs2, cs2^2 and 3
# (the dynamic model's trace).
#
# For these b the rounded reciprocal y = RN(1/b) is 3, 4.5, 1.5, 9, RN(1/3) and
# y*b = 1 - 2^-54 exactly, so a*y = (a/b)(1 - 2^-54) lies within a relative
# 2^-54 below a/b.  Then q0 = RN(a*y) is a faithful quotient: either a*y is
# already within the two floating-point neighbours of a/b, or it lies below
# the lower neighbour F1 by less than |a/b| 2^-54 < the half spacing below F1
# (when F1 = 2^e, a*y >= 2^e - 2^(e-54), the midpoint, which ties to the even
# 2^e).  With y = RN(1/b) and q0 faithful, Markstein's theorem makes one
# correction r = fma(-b, q0, a), q1 = fma(r, y, q0) the correctly rounded
# a/b.  Checked here exactly (the reciprocal error) and by exhaustive-style
# sampling including quotients next to rounding midpoints and binade
# boundaries.

SRC1 = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s = 0xD1B54A32D192ED03ULL;
static uint64_t nx(void){ s ^= s >> 12; s ^= s << 25; s ^= s >> 27; return s * 0x2545F4914F6CDD1DULL; }
static double u01(void){ return (double)(nx() >> 11) / 9007199254740992.0; }
static double div1(double a, double b, double y){ double q = a*y; double r = fma(-b,q,a); return fma(r,y,q); }
static int same(double x, double y){ uint64_t a,b; memcpy(&a,&x,8); memcpy(&b,&y,8); return a==b || (x==0.0 && y==0.0); }
int main(int argc, char** argv){
  long n = atol(argv[1]);
  const double cs2 = 1.0/3.0;
  const double dens[5] = {cs2, 2.0*cs2*cs2, 2.0*cs2, cs2*cs2, 3.0};
  long bad = 0, bad0 = 0, cases = 0;
  for (int d = 0; d < 5; ++d) {
    const double b = dens[d], y = 1.0 / b;
    for (long i = 0; i < n; ++i) {
      double a;
      const int kind = (int)(i % 4);
      const int e = (int)(nx() % 120) - 60;
      if (kind == 0) {                      /* random dividend */
        a = (u01()*2.0 - 1.0) * ldexp(1.0, e);
      } else if (kind == 1) {               /* quotient next to a midpoint of its binade */
        const uint64_t m = (1ULL << 52) | (nx() >> 12);
        const double mid = ldexp((double)m + 0.5, e - 52);  /* exact: 54 bits */
        a = nextafter(mid * b, (nx() & 1) ? INFINITY : -INFINITY);
        for (int k = (int)(nx() % 4); k > 0; --k) a = nextafter(a, (nx() & 1) ? INFINITY : -INFINITY);
      } else if (kind == 2) {               /* quotient next to a power of two */
        a = ldexp(1.0, e) * b;
        for (int k = (int)(nx() % 6); k > 0; --k) a = nextafter(a, (nx() & 1) ? INFINITY : -INFINITY);
      } else {                              /* dividend with a short significand */
        a = ldexp((double)(nx() % 4096) + 1.0, e - 12) * ((nx() & 1) ? 1.0 : -1.0);
      }
      if (!isfinite(a) || fabs(a) < 0x1p-900) continue;
      ++cases;
      const double q = a / b;
      if (!same(div1(a,b,y), q)) ++bad;
      if (!same(a*y, q)) ++bad0;
    }
  }
  printf("%ld %ld %ld\n", bad, bad0, cases);
  return 0;
}
"""


def test_constant_divisors_reciprocal_error_is_2_pow_minus_54():
    from fractions import Fraction as Fr

    cs2 = 1.0 / 3.0
    for b in (cs2, 2.0 * cs2 * cs2, 2.0 * cs2, cs2 * cs2, 3.0):
        y = 1.0 / b
        # y = RN(1/b): within half an ulp of 1/b (ulp of y, y not a power of 2)
        half_ulp = Fr(abs(math.nextafter(y, math.inf) - y)) / 2
        assert abs(Fr(y) - 1 / Fr(b)) <= half_ulp
        assert 1 - Fr(y) * Fr(b) == Fr(1, 2 ** 54)   # y*b = 1 - 2^-54 exactly


def test_one_correction_division_is_ieee_for_the_constant_divisors(tmp_path):
    c = tmp_path / "div1.c"
    c.write_text(SRC1)
    exe = tmp_path / "div1"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(c), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "5000000"], capture_output=True, text=True, check=True).stdout
    bad, bad0, cases = map(int, out.split())
    assert cases > 24_000_000
    assert bad == 0
    assert bad0 > 100_000  # a*y alone is not IEEE division

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


# The device's FP64 division, reciprocal and square root (Arith<double>::div,
# rcp_f64, Arith<double>::sqrtp in lbm_device.hpp): the hardware's first
# approximation of 1/b or 1/sqrt(x), refined by FMA steps, with no operand
# scaling.  Here the first approximation is the exact value perturbed by a
# random relative error of up to 2^-20 (AMD documents the FP64 reciprocal and
# reciprocal square root to 2^29 ulp, a relative 2^-23; the GPU test measures
# them), and the refined results must equal IEEE division, 1/b and sqrt for
# operands and quotients of magnitude in [2^-900, 2^900], including quotients
# next to a rounding midpoint (read from stdin).  The one exception is the
# known one of the reciprocal iteration: divisors whose significand is all
# ones, 2^k (2 - 2^-52), where 1/b lies 2^-106 relative above the midpoint
# next to a power of two; from a perturbed first approximation the iteration
# lands on that power of two and its last correction ties to it.  There the
# sequence is exact only from a correctly rounded first approximation, which
# is checked here, and the hardware's is checked on the GPU
# (test_gpu_arith.py).  Mutants with one refinement step less must differ.
SRC2 = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ULL;
static uint64_t nx(void){ s ^= s >> 12; s ^= s << 25; s ^= s >> 27; return s * 0x2545F4914F6CDD1DULL; }
static double u01(void){ return (double)(nx() >> 11) / 9007199254740992.0; }
static double approx(double v, double emax){ return v * (1.0 + (2.0*u01() - 1.0) * emax); }
static double div_dev(double a, double b, double y, int newton){
  double e = fma(-b,y,1.0); y = fma(y,e,y);
  if (newton > 1) { e = fma(-b,y,1.0); y = fma(y,e,y); }
  double q = a*y; double r = fma(-b,q,a); return fma(r,y,q);
}
static double rcp_dev(double b, double y, int newton){
  double e = fma(-b,y,1.0); y = fma(y,e,y);
  if (newton > 1) { e = fma(-b,y,1.0); y = fma(y,e,y); }
  double r = fma(-b,y,1.0); return fma(r,y,y);
}
static double sqrt_dev(double x, double y, int corr){
  double g = x*y; double h = y*0.5;
  if (corr >= 0) { double r = fma(-h,g,0.5); g = fma(g,r,g); h = fma(h,r,h); }
  else corr = 2;  /* mutant: no Goldschmidt step */
  double d = fma(-g,g,x); g = fma(d,h,g);
  if (corr > 1) { d = fma(-g,g,x); g = fma(d,h,g); }
  return (x == 0.0 || isinf(x)) ? x : g;
}
static int all_ones(double b){ uint64_t u; memcpy(&u, &b, 8); return (u & 0xFFFFFFFFFFFFFULL) == 0xFFFFFFFFFFFFFULL; }
static double mag(double lo, double hi){ return exp2(lo + (hi - lo) * u01()); }
int main(int argc, char** argv){
  long n = atol(argv[1]);
  const double emax = 0x1p-20;
  long bd = 0, br = 0, bs = 0, md = 0, mr = 0, ms = 0, cd = 0, cs = 0, ones = 0;
  for (long i = 0; i < n; ++i) {
    double a, b, x;
    switch (i % 4) {
      case 0: a = mag(-900, 900); b = mag(-900, 900); x = mag(-900, 900); break;
      case 1: a = mag(-8, 8); b = mag(-8, 8); x = mag(-140, 8); break;  /* the kernels' ranges */
      case 2:  /* short significands: exact quotients and roots */
        a = (double)(nx() >> 40) * ldexp(1.0, (int)(nx() % 40) - 20);
        b = (double)((nx() >> 44) | 1) * ldexp(1.0, (int)(nx() % 40) - 20);
        x = (double)(nx() >> 38) * (double)(nx() >> 38) * ldexp(1.0, 2*((int)(nx() % 40) - 20));
        break;
      default:  /* significands near 1 and near 2 */
        b = ldexp(1.0 + (double)(nx() >> 52) * 0x1p-52, (int)(nx() % 60) - 30);
        if (nx() & 1) b = ldexp(2.0 - (double)(nx() >> 52) * 0x1p-52, (int)(nx() % 60) - 30);
        a = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, (int)(nx() % 60) - 30);
        x = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, (int)(nx() % 200) - 100);
    }
    if (nx() & 1) a = -a;
    if (nx() & 1) b = -b;
    if (a == 0.0 || x == 0.0) continue;
    double lq = log2(fabs(a)) - log2(fabs(b));
    if (lq < 900.0 && lq > -900.0 && !all_ones(b)) {
      ++cd;
      double y0 = approx(1.0 / b, emax);
      if (div_dev(a, b, y0, 2) != a / b) ++bd;
      if (div_dev(a, b, y0, 1) != a / b) ++md;
      if (rcp_dev(b, y0, 2) != 1.0 / b) ++br;
      if (rcp_dev(b, y0, 1) != 1.0 / b) ++mr;
    }
    ++cs;
    double z0 = approx(1.0 / sqrt(x), emax);
    if (sqrt_dev(x, z0, 2) != sqrt(x)) ++bs;
    if (sqrt_dev(x, z0, -1) != sqrt(x)) ++ms;
  }
  /* stdin: "d a b", quotients next to a rounding midpoint; "s x", roots next
     to one, refined from a first approximation within 2^-24 */
  long np_ = 0, bp = 0, bp0 = 0, mp = 0, nr = 0, brt = 0, mrt = 0;
  double a, b;
  char tag;
  while (scanf(" %c", &tag) == 1) {
    if (tag == 'd' && scanf("%la %la", &a, &b) == 2) {
      ++np_;
      /* an all-ones divisor needs a correctly rounded first approximation */
      double y0 = all_ones(b) ? 1.0 / b : approx(1.0 / b, emax);
      if (div_dev(a, b, y0, 2) != a / b) ++bp;
      if (div_dev(a, b, y0, 1) != a / b) ++mp;
      if (a * (1.0 / b) != a / b) ++bp0;
    } else if (tag == 's' && scanf("%la", &a) == 1) {
      ++nr;
      double z0 = approx(1.0 / sqrt(a), 0x1p-24);
      if (sqrt_dev(a, z0, 2) != sqrt(a)) ++brt;
      if (sqrt_dev(a, z0, 1) != sqrt(a)) ++mrt;
    } else {
      return 2;
    }
  }
  /* all-ones significands: exact from RN(1/b), not from a perturbed start */
  long ob = 0, op = 0;
  for (int e = -60; e <= 60; ++e) {
    for (int sg = -1; sg <= 1; sg += 2) {
      const double bo = sg * ldexp(2.0 - 0x1p-52, e);
      ++ones;
      if (rcp_dev(bo, 1.0 / bo, 2) != 1.0 / bo || div_dev(1.0, bo, 1.0 / bo, 2) != 1.0 / bo) ++ob;
      for (int t = 0; t < 16; ++t)
        if (rcp_dev(bo, approx(1.0 / bo, emax), 2) != 1.0 / bo) { ++op; break; }
    }
  }
  /* +0 and infinity go through the square root's select; +0 / b is +0 */
  if (sqrt_dev(0.0, approx(1.0, emax), 2) != 0.0 || !isinf(sqrt_dev(INFINITY, 0.0, 2))) ++bs;
  if (div_dev(0.0, 3.5, approx(1.0 / 3.5, emax), 2) != 0.0 || signbit(div_dev(0.0, 3.5, 1.0 / 3.5, 2))) ++bd;
  printf("%ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld\n", bd, br, bs, md, mr, ms,
         cd, cs, np_, bp, bp0, ones, ob, op, mp, nr, brt, mrt);
  return 0;
}
"""


def hard_operands_text(count, seed):
    """stdin of the program above: quotients a/b within 2^-54/B of a rounding
    midpoint (significands A < B with 2^54 A = (2Q+1) B +- 1) and square
    roots within |t| 2^-109 of one (X 2^55 = s^2 - t, 1024 <= |t|), both
    scaled by random powers of two."""
    import random

    from tests.test_gpu_arith import near_midpoint_pairs, near_midpoint_roots

    r = random.Random(seed)
    lines = []
    for a, b in near_midpoint_pairs(count, seed):
        fa = math.ldexp(float(a), r.randrange(-40, 40) - 52)
        fb = math.ldexp(float(b), r.randrange(-40, 40) - 52)
        lines.append(f"d {fa.hex()} {fb.hex()}")
    for sg in (1, -1):
        for x, _ in near_midpoint_roots(count // 2, sg, t_start=1025):
            lines.append(f"s {math.ldexp(float(x), 2 * r.randrange(-40, 40) - 53).hex()}")
    return "\n".join(lines) + "\n"


def test_unscaled_division_reciprocal_and_sqrt_are_ieee(tmp_path):
    c = tmp_path / "div2.c"
    c.write_text(SRC2)
    exe = tmp_path / "div2"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(c), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "8000000"], input=hard_operands_text(20000, 5),
                         capture_output=True, text=True, check=True).stdout
    (bd, br, bs, md, mr, ms, cd, cs, npairs, bp, bp0, ones, ob, op, mp, nroots, brt,
     mrt) = map(int, out.split())
    assert cd > 7_000_000 and cs > 7_900_000 and npairs == 20000 and nroots == 20000
    assert ones == 242
    assert (bd, br, bs, bp, brt) == (0, 0, 0, 0, 0)
    assert bp0 > 5000  # the pairs are hard: a * (1/b) misses half of them
    # all-ones significands: exact from a correctly rounded first approximation,
    # wrong from a perturbed one (so the GPU check of the hardware's matters)
    assert ob == 0 and op > 200
    # one refinement step less is not enough: the reciprocal without its
    # second Newton step and the root without its Goldschmidt step miss on
    # random operands; the quotient without the second Newton step (its
    # correction absorbs most of the reciprocal's error) and the root without
    # its second correction miss next to midpoints
    assert mr > 100_000 and ms > 10_000 and mp > 1000 and mrt > 100

"""The FP64 kernels replace a / b by q = a*y and two Markstein corrections
(r = fma(-b, q, a); q = fma(r, y, q)) with y = RN(1/b).  This checks, on the
CPU with the same algorithm (C, fma from libm), that the result is bitwise
equal to IEEE division for the kernels' constant divisors, for divisors near
the densities and relaxation times they meet, and for random normal operands.
Plain multiplication by the reciprocal is shown to differ, so the test has
teeth."""
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
    int m = (int)(i % 7);
    if (m < 4) b = dens[m];
    else if (m == 4) b = 0.9 + 0.2*u01();
    else if (m == 5) b = 0.5 + 40.0*u01();
    else b = (0.5 + u01()) * ldexp(1.0, (int)(nx() % 200) - 100);
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

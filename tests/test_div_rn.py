"""The quotient the kernels use for stencil indices (div_rn in
libhybridgrains_amd/csrc/hg_mpm_kernels.h): q = a * inv_h followed by two
corrections q += fma(-q, h, a) * inv_h (fused), with inv_h = RN(1 / h).  The
reference floors and ceils the IEEE quotient (x - min) / h
(uGIMPLinearBasisFunctions.cpp:68-90, LinearBasisFunctions.cpp:48-60,
CubicBasisFunctions.cpp:61-75), so the two must agree bit for bit.  The same
operation sequence is run here in C (C's fma is the IEEE fused multiply-add,
as v_fma_f64 is) against the division over cell widths of the shipped
scenes, random widths and widths whose significand is all ones, for
quotients near integers (where floor and ceil could flip) and elsewhere.
Zero quotients are left out: the sequence may give +0 for -0, which floor,
ceil and every comparison treat alike.
"""
import shutil
import subprocess

import pytest

C_SOURCE = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double u01(void) { return (xr() >> 11) * (1.0 / 9007199254740992.0); }
static double div_rn(double a, double h, double ih) {
  double q = a * ih;
  q = fma(fma(-q, h, a), ih, q);
  return fma(fma(-q, h, a), ih, q);
}
int main(int argc, char **argv) {
  const long per = atol(argv[1]);
  long bad = 0, n = 0;
  for (int hi = 0; hi < 40; ++hi) {
    double h;
    if (hi == 0) h = 0.001;
    else if (hi == 1) h = 0.0025;
    else if (hi == 2) h = 1.0 / 3.0;
    else if (hi == 3) h = 0.1;
    else if (hi == 4) h = 0.02;
    else h = ldexp(1.0 + u01(), -(int)(xr() % 14));
    if (hi >= 5 && hi < 15) {  /* significands close to all ones */
      uint64_t b; memcpy(&b, &h, 8); b |= 0x000FFFFFFFFFFFF0ull; memcpy(&h, &b, 8);
    }
    const double ih = 1.0 / h;
    for (long k = 0; k < per; ++k) {
      double a;
      const int m = (int)(xr() % 3);
      if (m == 0) {
        a = (u01() * 4.0 - 2.0) * (1.0 + (double)(xr() % 2000));
      } else {
        a = floor((u01() * 4.0 - 2.0) * 3000.0) * h;
        const int nud = (int)(xr() % 9) - 4;
        for (int z = 0; z < abs(nud); ++z) a = nextafter(a, nud > 0 ? INFINITY : -INFINITY);
        if (m == 2) a *= h;
      }
      const double ref = a / h, got = div_rn(a, h, ih);
      if (ref == 0.0) continue;
      ++n;
      if (memcmp(&ref, &got, 8) != 0) ++bad;
    }
  }
  printf("%ld %ld\n", n, bad);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_rn_is_the_ieee_quotient(tmp_path):
    src = tmp_path / "div_rn.c"
    src.write_text(C_SOURCE)
    exe = tmp_path / "div_rn"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "100000"], check=True, capture_output=True, text=True,
                         timeout=120).stdout.split()
    n, bad = int(out[0]), int(out[1])
    assert n > 3_000_000
    assert bad == 0

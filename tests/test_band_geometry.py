"""Host band geometry of the linear plans (poyd_amd/csrc/poyg_band.h): the
closed-form widest row and cell count of a band equal the per-row walk of
lin_row_host (src/algn.c:872-968) over every row, for all three band cases.
The header is plain C++; the test compiles a small checker with g++."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "poyd_amd", "csrc", "poyg_band.h")

CHECKER = r'''
#include <cstdio>
#include "poyg_band.h"
using namespace poyg;
int main() {
  long bad = 0, tot = 0, cases[3] = {0, 0, 0};
  for (int n = 1; n <= 200; ++n)
    for (int m = n; m <= n + 600; m += (m < n + 40 ? 1 : 9))
      for (int dw = 0; dw <= 150; dw += 5) {
        const LinGeom g = lin_geom(m, n, dw);
        int mx = 0;
        int64_t c = 0;
        for (int i = 0; i < m; ++i) {
          int lo, hi;
          lin_row_host(i, m, n, g, lo, hi);
          if (lo < 0 || hi >= n || lo > hi) ++bad;
          mx = std::max(mx, hi - lo + 1);
          c += hi - lo + 1;
        }
        int mx2;
        int64_t c2;
        lin_band_extent(m, n, g, mx2, c2);
        ++tot;
        ++cases[g.cs];
        if (mx != mx2 || c != c2) ++bad;
      }
  printf("%ld %ld %ld %ld %ld\n", tot, bad, cases[0], cases[1], cases[2]);
  return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_band_extent_closed_form_equals_row_walk(tmp_path):
    src = tmp_path / "band_check.cpp"
    exe = tmp_path / "band_check"
    src.write_text(CHECKER)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    tot, bad, full, case2, case3 = map(int, out)
    assert bad == 0
    assert tot > 100000
    assert full > 0 and case2 > 0 and case3 > 0   # every band case is exercised


LIMIT_CHECKER = r'''
#include <cstdio>
#include "poyg_band.h"
using namespace poyg;
int main(int argc, char **argv) {
  int m, n, dw;
  while (scanf("%d %d %d", &m, &n, &dw) == 3) {
    const LinGeom g = lin_geom(m, n, dw);
    int mx;
    int64_t c;
    lin_band_extent(m, n, g, mx, c);
    printf("%d %lld\n", lin_dm_fits(m, mx) ? 1 : 0, (long long)m * lin_dm_stride(mx));
  }
  return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_linear_direction_band_limit(tmp_path):
    """The linear plans refuse (POYG_EUNSUPPORTED) a pair whose direction
    band reaches the kernels' 32-bit "out of range" offset 0x7ffffff0: the
    header's test agrees with the byte count for pairs on both sides of it."""
    src = tmp_path / "limit.cpp"
    exe = tmp_path / "limit"
    src.write_text(LIMIT_CHECKER)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)],
                   check=True)
    cases = [(16385, 16385, 2000), (5001, 5001, 600), (60000, 40000, 100), (40000, 39000, 5000),
             (70000, 40000, 10)]
    inp = "".join(f"{m} {n} {dw}\n" for m, n, dw in cases)
    out = subprocess.run([str(exe)], input=inp, check=True, capture_output=True, text=True).stdout.split("\n")
    got = [tuple(map(int, line.split())) for line in out if line.strip()]
    for (m, n, dw), (fits, nbytes) in zip(cases, got):
        assert fits == (nbytes < 0x7ffffff0), (m, n, dw)
    assert got[0][0] == 1 and got[1][0] == 1            # reference-sized sequences (<= 16384) fit
    assert got[2][0] == 0 and got[4][0] == 0            # a full plane of 60k x 40k does not

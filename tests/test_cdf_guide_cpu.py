"""The guide-table light selection (sample_cdf_guided, csrc/ljx_device.h)
returns the index of lajolla's TableDist1D sampling (upper bound over the
CDF, src/table_dist.cpp:27-33, restated as sample_cdf) for every u: random
CDFs with zero-weight entries, one to 5,000 entries, guide tables coarser and
finer than the CDF, u at every cell boundary k / m, just below it, and at
random.  Host code built with hipcc from the device header; no GPU needed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lajolla_public_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROG = r"""
#include "ljx_device.h"
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
using namespace ljx;
int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long checked = 0, bad = 0;
    const int sizes[] = {1, 2, 3, 7, 64, 1000, 1025, 4999};
    for (int n : sizes) {
        for (int trial = 0; trial < 4; trial++) {
            std::vector<double> w(n), cdf(n + 1, 0.0);
            for (int i = 0; i < n; i++) w[i] = (trial == 1 && i % 3 == 0) ? 0.0 : std::pow(U(rng), trial == 2 ? 8.0 : 1.0);
            if (trial == 3) for (int i = 0; i < n; i++) w[i] = 1.0;
            double tot = 0;
            for (int i = 0; i < n; i++) tot += w[i];
            if (tot <= 0) { w[0] = 1; tot = 1; }
            for (int i = 0; i < n; i++) cdf[i + 1] = cdf[i] + w[i] / tot;
            cdf[n] = 1.0;
            for (int m = 1; m <= 8192; m <<= 2) {
                std::vector<int> g(m + 1);
                cdf_guide(cdf.data(), n + 1, m, g.data());
                std::vector<double> us;
                for (int k = 0; k < m; k++) {
                    us.push_back(double(k) / m);
                    if (k > 0) us.push_back(std::nextafter(double(k) / m, 0.0));
                }
                for (int i = 0; i <= n; i++) {
                    us.push_back(cdf[i]);
                    us.push_back(std::nextafter(cdf[i], 0.0));
                    us.push_back(std::nextafter(cdf[i], 1.0));
                }
                for (int r = 0; r < 20000; r++) us.push_back(U(rng));
                for (double u : us) {
                    if (!(u >= 0.0 && u < 1.0)) continue;
                    checked++;
                    if (sample_cdf(cdf.data(), n, u) != sample_cdf_guided(cdf.data(), n, u, g.data(), m)) bad++;
                }
            }
        }
    }
    std::printf("checked %lld mismatches %lld\n", checked, bad);
    return bad == 0 && checked > 0 ? 0 : 1;
}
"""


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_guided_cdf_sampling_equals_binary_search(tmp_path):
    src = tmp_path / "cdf_guide.hip"
    src.write_text(PROG)
    exe = tmp_path / "cdf_guide"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I" + CSRC, str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr

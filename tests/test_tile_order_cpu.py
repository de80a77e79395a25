"""The accumulation's workgroup -> tile order (csrc/tile_order.h) covers every tile exactly once.

The header is plain C++ shared by the kernels and the host; it is compiled here with g++ into a
checker that walks every workgroup of every tile-grid size up to 160 x 90 tiles, for both
band orientations, and for xcd_rank over list lengths up to 5000.
"""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "halide-sar-app_amd", "csrc")

CHECKER = r"""
#include "tile_order.h"
#include <cstdio>
#include <vector>
int main() {
    long bad = 0;
    for (int tx = 1; tx <= 160; ++tx)
        for (int ty = 1; ty <= 90; ++ty)
            for (int cols = 0; cols < 2; ++cols) {
                std::vector<int> seen(tx * ty, 0);
                for (int id = 0; id < tx * ty; ++id) {
                    int x = -1, y = -1;
                    sarbp::xcd_band_tile(id, tx, ty, cols != 0, x, y);
                    if (x < 0 || x >= tx || y < 0 || y >= ty) ++bad;
                    else seen[y * tx + x]++;
                }
                for (int v : seen) bad += v != 1;
            }
    for (int n = 1; n <= 5000; ++n) {
        std::vector<int> seen(n, 0);
        for (int id = 0; id < n; ++id) {
            int t = sarbp::xcd_rank(id, n);
            if (t < 0 || t >= n) ++bad;
            else seen[t]++;
        }
        for (int v : seen) bad += v != 1;
        // XCD k (workgroups id % 8 == k) takes one contiguous run of the list
        for (int k = 0; k < 8 && k < n; ++k) {
            int prev = -1;
            for (int id = k; id < n; id += 8) {
                int t = sarbp::xcd_rank(id, n);
                if (prev >= 0 && t != prev + 1) ++bad;
                prev = t;
            }
        }
    }
    std::printf("bad %ld\n", bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tile_order_is_a_bijection(tmp_path):
    src = tmp_path / "check.cpp"
    src.write_text(CHECKER)
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "bad 0", r.stdout + r.stderr

"""The workgroup -> tile maps of the launches (taa-star_amd/csrc/taa_tiles.hpp), compiled for the
host: for every tile count up to 4,100 and at the sizes of 4K and 8K frames and of large batches,
the single-launch map is a bijection of the n workgroups onto the n tiles that gives each XCD
(workgroup index mod 8) one contiguous range of tiles, taken in workgroup order; the batched map
covers every tile exactly once over the 8 * ceil(total / 8) workgroups the host launches, gives
XCD k the k-th eighth, and leaves only the rounding's workgroups without a tile (-1)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "taa-star_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "taa_tiles.hpp"
using taastar::batch_tile_of;
using taastar::xcd_tile_of;
static long long bad = 0, checked = 0;
static void check_single(int n) {
    std::vector<int> seen(n, 0), lo(8, -1), hi(8, -1), last(8, -1);
    for (int b = 0; b < n; ++b) {
        const int t = xcd_tile_of(b, n), x = b & 7;
        ++checked;
        if (t < 0 || t >= n) { ++bad; continue; }
        ++seen[t];
        if (last[x] >= 0 && t != last[x] + 1) ++bad;  // an XCD's tiles follow its workgroup order
        last[x] = t;
        if (lo[x] < 0) lo[x] = t;
        hi[x] = t;
    }
    for (int t = 0; t < n; ++t) if (seen[t] != 1) ++bad;
    for (int x = 1; x < 8; ++x)                        // the ranges tile [0, n) in XCD order
        if (lo[x] >= 0 && lo[x - 1] >= 0 && lo[x] != hi[x - 1] + 1) ++bad;
}
static void check_batch(int total) {
    const int per = (total + 7) / 8, grid = 8 * per;
    std::vector<int> seen(total, 0);
    for (int b = 0; b < grid; ++b) {
        const int t = batch_tile_of(b, total), x = b & 7, s = b >> 3;
        ++checked;
        if (t == -1) { if (x * per + s < total) ++bad; continue; }  // only the rounding is idle
        if (t < 0 || t >= total) { ++bad; continue; }
        if (t != x * per + s) ++bad;                   // XCD x takes the x-th eighth in order
        ++seen[t];
    }
    for (int t = 0; t < total; ++t) if (seen[t] != 1) ++bad;
}
int main() {
    for (int n = 1; n <= 4100; ++n) { check_single(n); check_batch(n); }
    const int big[] = {9216, 10240, 36864, 73728, 159744, 1 << 20};
    for (int n : big) { check_single(n); check_batch(n); }
    std::printf("%lld checked, %lld bad\n", checked, bad);
    return bad == 0 && checked > 0 ? 0 : 1;
}
"""


def test_tile_maps_on_the_host(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "tiles.cpp"
    exe = tmp_path / "tiles"
    src.write_text(PROGRAM)
    cmd = ["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-I" + CSRC,
           str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 bad" in r.stdout

"""The Catmull-Rom footprint rules the kernels use (taa-star_amd/csrc/taa_footprint.hpp), compiled
for the host and checked exhaustively on small images and at the int range's ends: the interior
test of a pre-clamped base agrees with the raw form for every int; an interior footprint's four
taps lie in the image; the pre-clamp never changes which texel a tap reads once clamped to the
edge; forming k + 2 of a pre-clamped base cannot overflow (for every extent the host accepts: an
image row of 2^28 RGBA32F texels already reaches the 4 GiB limit)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "taa-star_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

PROGRAM = r"""
#include <climits>
#include <cstdint>
#include <cstdio>
#include "taa_footprint.hpp"
using taastar::cr_base;
using taastar::cr_inside;
using taastar::cr_inside_raw;
static long long clamp64(long long v, long long lo, long long hi) { return v < lo ? lo : (v > hi ? hi : v); }
int main() {
    long long checked = 0, bad = 0;
    const int ns[] = {1, 2, 3, 4, 5, 6, 7, 8, 31, 640, 3840, 7680, 1 << 20, 1 << 28};
    for (int n : ns) {
        long long ks[200];
        int m = 0;
        for (int d = -40; d <= 40; ++d) ks[m++] = d;
        for (int d = -40; d <= 40; ++d) ks[m++] = (long long)n + d;
        const long long ends[] = {INT_MIN, INT_MIN + 1, INT_MIN + 2, INT_MAX - 2, INT_MAX - 1, INT_MAX};
        for (long long e : ends) ks[m++] = e;
        for (int j = 0; j < m; ++j) {
            if (ks[j] < INT_MIN || ks[j] > INT_MAX) continue;
            const int kr = (int)ks[j];
            const int k = cr_base(kr, n);
            ++checked;
            if ((long long)k + 2 > INT_MAX) ++bad;                       // k + 2 of a pre-clamped base fits
            if (cr_inside(k, n) != cr_inside_raw(kr, n)) ++bad;          // both forms of the test agree
            if (cr_inside(k, n) && (k - 1 < 0 || (long long)k + 2 > n - 1)) ++bad;  // interior taps in the image
            for (int i = 0; i < 4; ++i)                                   // the same texel after the edge clamp
                if (clamp64((long long)k - 1 + i, 0, n - 1) != clamp64((long long)kr - 1 + i, 0, n - 1)) ++bad;
        }
    }
    std::printf("%lld checked, %lld bad\n", checked, bad);
    return bad == 0 && checked > 0 ? 0 : 1;
}
"""


def test_footprint_rules_on_the_host(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "footprint.cpp"
    exe = tmp_path / "footprint"
    src.write_text(PROGRAM)
    cmd = ["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-I" + CSRC,
           str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 bad" in r.stdout

"""Host check of the fixed-divisor division used by the binning passes.

The wavefront -> (baseline, time block, channel block) mapping of the binning
kernels divides by counts fixed at launch with ngc::FastDiv (multiply-high and
shifts).  This builds a small host program against the library's internal
header and compares FastDiv with plain integer division, including a divisor
of 0, which an empty time or channel axis produces and which FastDiv takes as 1.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")

PROGRAM = r"""
#include "ngc_internal.hpp"
int main() {
    long bad = 0, n_checked = 0;
    auto check = [&](uint32_t d, uint32_t n) {
        const ngc::FastDiv f = ngc::FastDiv::of(d);
        const uint32_t dd = d ? d : 1;
        bad += f.div(n) != n / dd;
        n_checked++;
    };
    for (uint32_t d = 0; d < 2048; d++) {
        for (uint32_t n = 0; n < 4096; n++) check(d, n);
        for (uint32_t n : {0xffffffffu, 0xfffffffeu, 0x80000000u, 0x7fffffffu, 2147483649u})
            check(d, n);
    }
    // large divisors, and n around multiples of d
    uint64_t x = 88172645463325252ull;
    for (int i = 0; i < 200000; i++) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        const uint32_t d = (uint32_t)(x >> 32) >> (i % 32);
        const uint32_t n = (uint32_t)x;
        check(d, n);
        if (d > 1) {
            const uint32_t k = n / d;
            check(d, k * d);
            check(d, k * d - 1);
        }
    }
    printf("%ld %ld\n", n_checked, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fastdiv_matches_integer_division(tmp_path):
    src = tmp_path / "fastdiv_check.cpp"
    exe = tmp_path / "fastdiv_check"
    src.write_text(PROGRAM)
    r = subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "cngi_prototype_amd", "csrc"),
                        "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    n_checked, bad = map(int, out.stdout.split())
    assert n_checked > 8_000_000
    assert bad == 0 and out.returncode == 0

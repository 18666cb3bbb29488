"""The multiply-high division that the render kernels use for pixel and
block indices (voxvision_amd/csrc/fastdiv.h), compiled for the host with
g++ and checked against the % and / operators: every divisor up to 5000,
the powers of two and their neighbours, and large random divisors, each at
the multiples around it, the top of the range and random numerators."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHECK = r"""
#define __host__
#define __device__
#include "fastdiv.h"
#include <stdio.h>
#include <stdlib.h>

static uint64_t rng = 88172645463325252ull;
static uint32_t next(void) {
    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
    return (uint32_t)(rng >> 11);
}

static int check(uint32_t d) {
    FastDiv f = make_fastdiv(d);
    const uint32_t top = 0x7fffffffu;
    uint32_t ns[64];
    int k = 0;
    ns[k++] = 0; ns[k++] = 1; ns[k++] = top; ns[k++] = top - 1;
    ns[k++] = d - 1; ns[k++] = d; ns[k++] = top / d * d;
    ns[k++] = top / d * d - 1;
    while (k < 64) {
        uint32_t q = next() % (top / d + 1);
        uint32_t n = q * d + (next() % 3) - 1;  /* around a multiple */
        ns[k++] = (n > top) ? (next() & top) : n;
    }
    for (int i = 0; i < k; i++) {
        uint32_t n = ns[i];
        if (f.div(n) != n / d) {
            printf("FAIL d=%u n=%u got %u want %u\n", d, n, f.div(n), n / d);
            return 1;
        }
    }
    return 0;
}

int main(void) {
    int bad = 0;
    for (uint32_t d = 1; d <= 5000; d++) bad |= check(d);
    for (int l = 1; l < 31; l++) {
        bad |= check(1u << l);
        bad |= check((1u << l) - 1);
        bad |= check((1u << l) + 1);
    }
    for (int i = 0; i < 20000; i++) bad |= check(1 + next() % 0x7ffffffeu);
    /* render shapes: blocks per row and per frame */
    static const uint32_t wh[][2] = {{1920, 1080}, {3840, 2160}, {640, 480},
                                     {320, 240}, {160, 96}, {16, 4}};
    for (unsigned i = 0; i < sizeof wh / sizeof wh[0]; i++) {
        bad |= check(wh[i][0] / 4);
        bad |= check(wh[i][0] / 4 * (wh[i][1] / 4));
    }
    if (!bad) printf("ok\n");
    return bad;
}
"""


def test_fastdiv_matches_integer_division():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "check.cpp")
        exe = os.path.join(d, "check")
        open(src, "w").write(CHECK)
        r = subprocess.run(
            ["g++", "-O2", "-I",
             os.path.join(ROOT, "voxvision_amd", "csrc"), src, "-o", exe],
            capture_output=True, text=True)
        if r.returncode:
            pytest.fail(r.stderr)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        assert r.stdout.strip() == "ok"

"""The division by 6 of the CG kernels' interior fast path, as a method (CPU, no GPU needed).

diag(K) is 6 at every INTERIOR cell, so the tile kernels replace the division sequence there with
q = x c (c = 1/6 rounded), one fused correction q + (x - 6q) c, the sign of x for a zero, and the
division itself for non-zero |x| < 2^-124, infinities and NaN (Markstein's method for a constant
divisor).  This compiles that method in C with the host's fused multiply-add and compares it with
the IEEE quotient on all 2^32 single-precision inputs; the device code is checked the same way by
test_gpu_parity.test_div6_equals_division_on_every_float.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float div6(float x) {
    const float c6 = 1.0f / 6.0f;
    const float q = x * c6;
    const float r = fmaf(-q, 6.0f, x);
    float y = copysignf(fmaf(r, c6, q), x);
    const float ax = fabsf(x);
    if (!(ax >= 0x1p-124f && ax <= FLT_MAX) && x != 0.0f) y = x / 6.0f;
    return y;
}
static float plain(float x) {   /* the method without the fallback: must differ somewhere */
    const float c6 = 1.0f / 6.0f, q = x * c6;
    return fmaf(fmaf(-q, 6.0f, x), c6, q);
}
int main(void) {
    long bad = 0, bad_plain = 0;
#pragma omp parallel for reduction(+ : bad, bad_plain) schedule(static)
    for (long i = 0; i < (1L << 32); ++i) {
        uint32_t u = (uint32_t)i, a, b, p;
        float x, ref, y, z;
        memcpy(&x, &u, 4);
        ref = x / 6.0f;
        y = div6(x);
        z = plain(x);
        memcpy(&a, &y, 4); memcpy(&b, &ref, 4); memcpy(&p, &z, 4);
        if (!(isnan(ref) && isnan(y)) && a != b) bad++;
        if (!(isnan(ref) && isnan(z)) && p != b) bad_plain++;
    }
    printf("%ld %ld\n", bad, bad_plain);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div6_method_exact_on_every_float():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "div6.c"), os.path.join(d, "div6")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-mfma", "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    bad, bad_plain = int(out[0]), int(out[1])
    assert bad == 0
    # the fallback is needed: the bare correction differs on subnormal quotients and at -0
    assert bad_plain > 0

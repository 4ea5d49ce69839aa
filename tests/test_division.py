"""The kernels divide positions by the node spacing h as x * (1/h) corrected
by one fused multiply-add of the residual (div_by in
squishy_volumes_amd/csrc/device_math.hpp).  By Markstein's theorem that is
the correctly rounded quotient; the reference divides
(position_to_shift_quadratic, rust/crates/cpu/src/kernels.rs:46-49), so the
cell and the in-cell offset of every particle must agree bit for bit.

This test checks the identity on the CPU with IEEE binary32 arithmetic and a
true fused multiply-add (C's fmaf), over spacings that are and are not powers
of two and positions across the whole range the grid allows.
"""
import shutil
import subprocess

import pytest

C_SOURCE = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9e3779b97f4a7c15ull;
static uint32_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (uint32_t)s; }
static float fbits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static float div_by(float x, float h, float y) {
    const float q = x * y;
    const float r = fmaf(-q, h, x);
    return fmaf(r, y, q);
}
int main(void) {
    const float fixed[] = {0.1f, 1.0f / 3.0f, 1.0f / 7.0f, 1.0f / 256.0f, 0.5f, 0.015625f, 1.0f / 300.0f, 0.0123f};
    long bad = 0, n = 0;
    for (int hb = 0; hb < 400; ++hb) {
        const float h = hb < 8 ? fixed[hb] : fbits(0x3a000000u + rnd() % 0x0a000000u); /* about 5e-4 .. 0.5 */
        const float y = 1.0f / h;
        for (int k = 0; k < 50000; ++k) {
            float x;
            switch (k % 3) {
            case 0: x = ((float)(int32_t)(rnd() % 4000001u) - 2000000.0f) * 1e-6f; break;          /* |x| <= 2 */
            case 1: x = fbits(0x30000000u + rnd() % 0x18000000u) * ((rnd() & 1u) ? 1.0f : -1.0f); break;
            default: x = ((float)(rnd() % 2048u) + 0.5f) * h; break;                             /* cell faces */
            }
            ++n;
            if (div_by(x, h, y) != x / h) {
                if (bad < 5) printf("x=%a h=%a\n", x, h);
                ++bad;
            }
        }
    }
    printf("%ld %ld\n", bad, n);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_reciprocal_division_is_correctly_rounded(tmp_path):
    src = tmp_path / "div.c"
    exe = tmp_path / "div"
    src.write_text(C_SOURCE)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    bad, n = int(out[-2]), int(out[-1])
    assert n == 400 * 50000
    assert bad == 0, out

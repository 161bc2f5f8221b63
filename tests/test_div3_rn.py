"""The FP64 MRT kernel divides by 3! and 4! without a division instruction:
x / 6 = RN(q + rem r) on x / 2 and x / 24 the same on x / 8, with
r = RN(1/3), q = RN(x r), rem = x - 3q formed exactly by a fused multiply-add
(div3_rn in lbm_kernels.h).  The method is checked here on the CPU against
IEEE division, bit for bit, over random doubles of every exponent the
populations reach, the mantissas next to the multiples of 3 and the
power-of-two boundaries.  The GPU parity tests then compare the kernel's
populations with the oracle's plain divisions."""
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static double div3_rn(double x) {
    const double r = 1.0 / 3.0;
    const double q = x * r;
    const double rem = fma(-q, 3.0, x);
    return fma(rem, r, q);
}
static uint64_t bits(double v) { uint64_t u; memcpy(&u, &v, 8); return u; }
static uint64_t s = 88172645463325252ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static long bad = 0, n = 0;
static void check(double x) {
    for (int k = 0; k < 3; ++k) {
        const double scale = k == 0 ? 1.0 : (k == 1 ? 0.5 : 0.125);   /* 3, 3!, 4! */
        const double y = x * scale;
        const double want = k == 0 ? x / 3.0 : (k == 1 ? x / 6.0 : x / 24.0);
        const double got = div3_rn(y);
        ++n;
        if (bits(got) != bits(want) && !(got == 0 && want == 0)) {
            if (bad < 5) printf("x=%a k=%d got %a want %a\n", x, k, got, want);
            ++bad;
        }
    }
}
int main(void) {
    /* random mantissas at exponents 2^-900 .. 2^900, both signs */
    for (long i = 0; i < 4000000; ++i) {
        uint64_t m = next() & 0x000FFFFFFFFFFFFFull;
        uint64_t e = 123 + next() % 1800;
        uint64_t sg = next() & 1;
        uint64_t u = (sg << 63) | (e << 52) | m;
        double x; memcpy(&x, &u, 8);
        check(x);
    }
    /* the magnitudes the MRT sums take: s^n d with d of 1e-12 .. 1e2 */
    for (long i = 0; i < 2000000; ++i) {
        double x = ldexp((double)(next() >> 11) * 0x1p-53, -40 + (int)(next() % 48));
        check(x);
        check(-x);
    }
    /* next to exact multiples of 3 and to the binade edges */
    for (long j = 1; j < 200000; ++j) {
        double base = 3.0 * (double)j;
        double x = base;
        for (int d = 0; d < 4; ++d) { check(x); check(-x); x = nextafter(x, INFINITY); }
        x = base;
        for (int d = 0; d < 4; ++d) { check(x); x = nextafter(x, 0.0); }
    }
    for (int e = -1000; e <= 1000; ++e) {
        double p = ldexp(1.0, e), x = p;
        for (int d = 0; d < 8; ++d) { check(x); x = nextafter(x, INFINITY); }
        x = p;
        for (int d = 0; d < 8; ++d) { check(x); x = nextafter(x, 0.0); }
        check(1.5 * p);
        check(nextafter(1.5 * p, 0.0));
        check(nextafter(1.5 * p, INFINITY));
    }
    check(0.0);
    printf("%ld %ld\n", n, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_div3_rn_equals_ieee_division(tmp_path):
    c = tmp_path / "div3.c"
    exe = tmp_path / "div3"
    c.write_text(SRC)
    # -mfma: the hardware fused multiply-add; -ffp-contract=off: no other fusion
    flags = ["-O2", "-ffp-contract=off"]
    probe = subprocess.run(["gcc", "-mfma", "-x", "c", "-", "-fsyntax-only"], input=b"",
                           capture_output=True)
    if probe.returncode == 0:
        flags.append("-mfma")
    subprocess.run(["gcc", *flags, str(c), "-o", str(exe), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    n, bad = map(int, r.stdout.split()[-2:])
    assert n > 10 ** 7 and bad == 0, r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_div3_rn_check_catches_a_missing_correction(tmp_path):
    """the same check on the plain product x * RN(1/3) (the correction step
    dropped) reports mismatches, so the check above can fail"""
    c = tmp_path / "div3_bad.c"
    exe = tmp_path / "div3_bad"
    c.write_text(SRC.replace("return fma(rem, r, q);", "return q + 0.0 * rem;"))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(c), "-o", str(exe), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    n, bad = map(int, r.stdout.split()[-2:])
    assert r.returncode != 0 and bad > n // 10, r.stdout

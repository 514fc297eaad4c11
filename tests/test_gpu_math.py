"""The kernels' binary64 shortcuts for log, exp and pow against the
double-float routines they stand for (MI355X), on every binary32 input.

`fast_log` / `fast_exp` / `fast_pow` (csrc/te_device.hpp) round a binary64
evaluation to binary32 and defer to the restated double-float routine
(`rr_log` / `rr_exp` / `rr_pow`, pressure_tendency.wgsl:114-227) when the
value lies within 2^-40 of a rounding boundary.  Replacing one by the other
is only allowed if no input tells them apart: the self-test runs all 2^32
bit patterns through both on the device (pow at the exponents the kernels
use: cp/rd, rd/cv, 1/3) and counts differing bits.

`recip` (the tridiagonal recurrences' pivot reciprocal: hardware estimate
plus one Newton-Raphson step with an FMA residual) goes through the same test
against the IEEE division 1.0f / x; it defers to the division exactly for
|x| outside [2^-64, 2^64] (2^31 - 2 bit patterns, NaNs and zeros included).
"""
import sys

import pytest

import te_host

pytestmark = pytest.mark.gpu


def test_shortcuts_equal_the_restated_routines_on_every_input(gpu):
    r = te_host.selftest_math()
    print(r, flush=True)
    sys.stdout.flush()
    assert set(r) == {"log", "exp", "pow_cpdrd", "pow_rddcv", "pow_third", "recip"}
    for name, v in r.items():
        assert v["mismatch"] == 0, (name, v)
        if name == "recip":
            assert v["deferred"] == (1 << 31) - 2, v
        else:
            assert v["deferred"] < 1 << 25, (name, v)  # the shortcut decides for > 99 % of the inputs

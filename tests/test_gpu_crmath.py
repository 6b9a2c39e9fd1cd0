"""The short correctly rounded float32 division and square root of the kernels
(geoswe_amd/csrc/gs_crmath.h) equal the compiler's general IEEE expansions on the device:
every float32 input for 1/b and sqrt(x), 2^32 seeded pairs and four hard divisors against
every numerator for a/b. The host check (test_crmath_host.py) proves the refinement steps
for every value a 1-ulp hardware approximation can take; this one pins what the hardware
returns."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = {0: "1/b, every b", 1: "sqrt(x), every x", 2: "a/b, 2^32 seeded pairs", 3: "a/b, every a, hard divisors"}


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_short_cr_ops_equal_ieee(gpu, which):
    from geoswe_amd import _lib as L
    counts = L.DeviceArray((4,), np.uint64, zero=True)
    L.call("gs_crmath_check", which, counts.ptr, None)
    L.sync(None)
    c = counts.download()
    assert c[0] == 0, f"{NAMES[which]}: {int(c[0])} mismatches, last offending index {int(c[1]) - 1:#x}"
    assert c[2] == 0, f"{NAMES[which]} (short form): {int(c[2])} mismatches, last offending index {int(c[3]) - 1:#x}"

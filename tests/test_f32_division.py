"""The FP32 histogram binning evaluates the reference GPU kernel's IEEE
single-precision quotient sqrtf(s) / bin as (float)((double)r * RN(1/bin)).
The double product lies within 2^-51.9 of r / bin, and a quotient of two
floats lies at least 2^-49 (relative) from every midpoint between floats, so
rounding the product to float must give the IEEE quotient.  Checked here
against numpy's IEEE float32 division, bit for bit, on random quotients over
many binades and on quotients adjacent to float midpoints, for the bin widths
the tests and benches use."""
import numpy as np
import pytest

BINS = [0.5, 0.25, 0.1, 0.02, 0.01, 0.005, 0.004, 0.002, 0.001, 0.0002, 1.0 / 3.0, 0.0123]


def via_double(r, b):
    return (r.astype(np.float64) * (1.0 / np.float64(b))).astype(np.float32)


@pytest.mark.parametrize("bin_", BINS)
def test_product_in_double_equals_ieee_float_division(bin_):
    rng = np.random.default_rng(int(bin_ * 1e6) + 7)
    b = np.float32(bin_)
    # random r over 30 binades (distances 1e-4 .. 1e5 A)
    r = (10.0 ** rng.uniform(-4, 5, 2_000_000)).astype(np.float32)
    want = r / b
    assert want.dtype == np.float32
    assert np.array_equal(via_double(r, b).view(np.uint32), want.view(np.uint32))
    # quotients next to float midpoints: r = nearest floats to m * b for
    # midpoints m between consecutive floats, and their neighbours
    q = (10.0 ** rng.uniform(-2, 7, 500_000)).astype(np.float32)
    m = (q.astype(np.float64) + np.nextafter(q, np.float32(np.inf)).astype(np.float64)) / 2
    r2 = (m * np.float64(b)).astype(np.float32)
    for d in (-2, -1, 0, 1, 2):
        rr = r2.view(np.int32) + np.int32(d)
        rr = rr[(rr > 0)].view(np.float32)
        assert np.array_equal(via_double(rr, b).view(np.uint32), (rr / b).view(np.uint32))

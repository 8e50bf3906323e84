"""The FP32 histogram binning evaluates the reference GPU kernel's IEEE
single-precision quotient sqrtf(s) / bin as (float)((double)r * RN(1/bin)).
The double product lies within 2^-51.9 of r / bin, and a quotient of two
floats lies at least 2^-49 (relative) from every midpoint between floats, so
rounding the product to float must give the IEEE quotient.  Checked here
against numpy's IEEE float32 division, bit for bit, on random quotients over
many binades and on quotients adjacent to float midpoints, for the bin widths
the tests and benches use.

For bin widths that pass a check over all 2^23 significands on the device,
the kernels instead take q0 = RN(r yb), yb = RN(1/bin), and two FMAs
(residual r - q0 bin, correction q0 + e yb).  That sequence is evaluated here
exactly (each FMA rounded once from its exact value) and compared with IEEE
division over the same 2^23 significands, for several bin widths; the same
comparison fails for the uncorrected q0."""
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


def round_sum_to_f32(a, p):
    """RN to float32 of the exact sum a + p of two doubles: TwoSum gives
    s + err = a + p exactly; RN_f32(s) differs from RN_f32(a + p) only when
    s is exactly half way between two floats and err breaks the tie."""
    s = a + p
    bv = s - a
    err = (a - (s - bv)) + (p - bv)
    f = s.astype(np.float32)
    up = np.nextafter(f, np.float32(np.inf))
    dn = np.nextafter(f, np.float32(-np.inf))
    f64 = f.astype(np.float64)
    f = np.where((s == (f64 + up.astype(np.float64)) / 2) & (err > 0), up, f)
    f = np.where((s == (f64 + dn.astype(np.float64)) / 2) & (err < 0), dn, f)
    return f


def quot_fma(r, b):
    """The kernels' FP32 quotient form, each FMA rounded once from its exact
    value.  r - q0 b is exact in double: q0 b has 48 bits and lies within
    2^-22 of r; e yb has 48 bits."""
    b = np.float32(b)
    yb = np.float32(1.0) / b
    q0 = r * yb
    e = (r.astype(np.float64) - q0.astype(np.float64) * np.float64(b)).astype(np.float32)
    return round_sum_to_f32(q0.astype(np.float64), e.astype(np.float64) * np.float64(yb))


def significand_mismatches(form, b):
    """Number of r in [1, 2), all 2^23 floats, where form(r, b) != r / b."""
    bad = 0
    for c in range(8):
        r = (np.arange(c << 20, (c + 1) << 20, dtype=np.uint32) | np.uint32(0x3F800000)).view(np.float32)
        bad += int(np.count_nonzero(form(r, b).view(np.uint32) != (r / np.float32(b)).view(np.uint32)))
    return bad


QUOT_BINS = [0.005, 0.01, 0.004, 0.0002, 1.0 / 3.0, 0.0123] + \
    list(np.random.default_rng(5).uniform(1e-4, 0.5, 3))


@pytest.mark.parametrize("bin_", QUOT_BINS)
def test_fma_quotient_equals_ieee_float_division_on_every_significand(bin_):
    assert significand_mismatches(quot_fma, bin_) == 0


def test_significand_comparison_detects_the_uncorrected_quotient():
    bad = [significand_mismatches(lambda r, b: r * (np.float32(1.0) / np.float32(b)), b)
           for b in QUOT_BINS[:4]]
    assert all(x > 0 for x in bad), bad

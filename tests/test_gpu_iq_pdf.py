"""H3 and H4 parity.  Histogram -> I(q): within 1e-5 relative of the FP64
oracle on the same histogram (pointwise where I >= 1e-3 max I).  PDF: equal
to the FP64 oracle bit for bit, for every type and both radius
conventions."""
import numpy as np
import pytest

import xs_oracle as O
from conftest import golden, rel_err
from xansons_hip import structures as st

pytestmark = pytest.mark.gpu

TOL = 1e-5   # relative error bar for intensities vs FP64


def check_iq(ctx, hist, ec, bin_, q, ff=None, sl=None, label=""):
    Sg = ctx.hist_pairsum(hist, bin_, q)
    So = O.hist_pairsum(hist, bin_, q)
    Ig = O.intensity_1d(Sg, ec, q, ff=ff, sl=sl)
    Io = O.intensity_1d(So, ec, q, ff=ff, sl=sl)
    pw, l2 = rel_err(Ig, Io)
    print(f"{label}: I(q) max rel {pw:.3e}  L2 rel {l2:.3e}", flush=True)
    assert pw <= TOL and l2 <= TOL
    return pw, l2


def test_iq_fcc_c1(ctx):
    g = golden("fcc1985")
    x64 = g["xyz"].astype(np.float64)
    nh = O.hist_size(x64, 0.01)
    h, _ = O.hist_f64(x64, g["el_count"], 0.01, nh)
    check_iq(ctx, h, g["el_count"], 0.01, g["q"], ff=g["ff"], label="c1")


def test_iq_nacl_small(ctx):
    g = golden("nacl_small")
    x64 = g["xyz"].astype(np.float64)
    nh = O.hist_size(x64, 0.005)
    h, _ = O.hist_f64(x64, g["el_count"], 0.005, nh)
    check_iq(ctx, h, g["el_count"], 0.005, g["q"], sl=g["sl"], label="nacl neutron")
    check_iq(ctx, h, g["el_count"], 0.005, g["q"], ff=g["ff"], label="nacl xray")


def test_iq_q_zero_and_empty(ctx):
    h = np.zeros((1, 50), dtype=np.uint64)
    h[0, 10] = 7
    S = ctx.hist_pairsum(h, 0.1, np.array([0.0, 1.0]))
    assert S[0, 0] == 0.0
    assert abs(S[0, 1] - 7 * np.sin(1.05) / 1.05) < 1e-6 * 7
    S0 = ctx.hist_pairsum(np.zeros((3, 100), dtype=np.uint64), 0.1, np.linspace(0.1, 5, 9))
    assert not S0.any()


def test_iq_rocksalt_million(ctx, tmp_path):
    """Full-size histogram -> I(q) on 4096 q over 0.2-25 A^-1."""
    ens = st.rocksalt_cube()
    nh = O.hist_size(ens.xyz64(), 0.005)
    h, _ = ctx.hist(ens.xyz, ens.el_count, 0.005, nh)
    q = st.q_mesh(0.2, 25.0, 4096)
    ff = st.gaussian_ff(q, st.gaussian_ff_coeffs(2, seed=5))
    check_iq(ctx, h, ens.el_count, 0.005, q, ff=ff, label="c2 xray")


@pytest.mark.parametrize("ptype", [0, 1, 2])
@pytest.mark.parametrize("cpu_radius", [True, False])
def test_pdf_bit_exact(ctx, ptype, cpu_radius):
    g = golden("nacl_small")
    x64 = g["xyz"].astype(np.float64)
    nh = O.hist_size(x64, 0.005)
    h, _ = O.hist_f64(x64, g["el_count"], 0.005, nh)
    pg = ctx.pdf(h, g["el_count"], 0.005, ptype, 0.04, g["sl"], cpu_radius=cpu_radius)
    po = O.pdf(h, g["el_count"], 0.005, ptype, 0.04, g["sl"], cpu_radius_offset=cpu_radius)
    assert np.array_equal(pg.view(np.uint64), po.view(np.uint64))

"""The oracle against the reference's own executable (golden vectors written
by tests/golden/make_golden.py).  Printed reference values carry 7
significant digits, so agreement is checked at rtol 1e-6; histogram counts
recovered from the printed RDFs must match exactly."""
import json
import os

import numpy as np
import pytest

import xs_oracle as O
from conftest import golden

RTOL = 1e-6


def close(a, b, rtol=RTOL, atol=0.0):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol)


@pytest.fixture(scope="module")
def fcc():
    return golden("fcc1985")


@pytest.fixture(scope="module")
def nacl():
    return golden("nacl_small")


def test_debye_xray_fcc(fcc):
    xyz = fcc["xyz"].astype(np.float64)
    S = O.debye_pairsum(xyz, fcc["el_count"], fcc["q"])
    I = O.intensity_1d(S, fcc["el_count"], fcc["q"], ff=fcc["ff"])
    close(I, fcc["I_debye"])


def test_hist_f64_exact_fcc(fcc):
    xyz = fcc["xyz"].astype(np.float64)
    nh = O.hist_size(xyz, float(fcc["bin"]))
    h, ov = O.hist_f64(xyz, fcc["el_count"], float(fcc["bin"]), nh)
    n = int(fcc["el_count"].sum())
    rec = np.rint(fcc["rdf"] * float(fcc["bin"]) * n / 2).astype(np.uint64)
    assert ov == 0
    assert np.array_equal(h[0, :rec.size], rec)
    assert h[0, rec.size:].sum() == 0
    assert h.sum() == n * (n - 1) // 2


def test_hist_intensity_fcc(fcc):
    xyz = fcc["xyz"].astype(np.float64)
    nh = O.hist_size(xyz, float(fcc["bin"]))
    h, _ = O.hist_f64(xyz, fcc["el_count"], float(fcc["bin"]), nh)
    S = O.hist_pairsum(h, float(fcc["bin"]), fcc["q"])
    close(O.intensity_1d(S, fcc["el_count"], fcc["q"], ff=fcc["ff"]), fcc["I_hist"])


def test_density_and_pdfs_fcc(fcc):
    xyz = fcc["xyz"].astype(np.float64)
    rho = O.atomic_density(xyz)
    assert abs(rho / float(fcc["rho0_printed"]) - 1) < 1e-5   # printed with 6 digits
    nh = O.hist_size(xyz, float(fcc["bin"]))
    h, _ = O.hist_f64(xyz, fcc["el_count"], float(fcc["bin"]), nh)
    rp = O.pdf(h, fcc["el_count"], float(fcc["bin"]), 2, rho, fcc["sl"])
    ref = fcc["rpdf_auto"]
    close(rp[0, :ref.size], ref, atol=1e-5 * np.abs(ref).max())
    p = O.pdf(h, fcc["el_count"], float(fcc["bin"]), 1, 0.059, fcc["sl"])
    ref = fcc["pdf_0059"]
    close(p[0, :ref.size], ref)


def test_debye_neutron_and_polar_nacl(nacl):
    xyz = nacl["xyz"].astype(np.float64)
    ec, q = nacl["el_count"], nacl["q"]
    S = O.debye_pairsum(xyz, ec, q)
    close(O.intensity_1d(S, ec, q, sl=nacl["sl"]), nacl["I_debye_neut"])
    close(O.intensity_1d(S, ec, q, ff=nacl["ff"], polar=True, lam=float(nacl["lam"])),
          nacl["I_debye_xray_polar"])


def test_partial_hists_exact_nacl(nacl):
    """Per-pair-type counts.  The reference names partial files by the
    alphabetical element map (IO.cpp:74-106) while the data are in table
    order, so 'Cl_Cl' holds type (0,0) = Na-Na."""
    xyz = nacl["xyz"].astype(np.float64)
    ec, b = nacl["el_count"], float(nacl["bin"])
    n = int(ec.sum())
    h, ov = O.hist_f64(xyz, ec, b, O.hist_size(xyz, b))
    assert ov == 0
    for key, p in (("Cl_Cl", 0), ("Cl_Na", 1), ("Na_Na", 2)):
        rec = np.rint(nacl["partial_rdf_" + key] * b * n / 2).astype(np.uint64)
        assert np.array_equal(h[p, :rec.size], rec)
        assert h[p, rec.size:].sum() == 0
    assert h[0].sum() == ec[0] * (ec[0] - 1) // 2
    assert h[1].sum() == ec[0] * ec[1]
    assert h[2].sum() == ec[1] * (ec[1] - 1) // 2


def test_hist_intensity_and_rpdf_nacl(nacl):
    xyz = nacl["xyz"].astype(np.float64)
    ec, q, b = nacl["el_count"], nacl["q"], float(nacl["bin"])
    h, _ = O.hist_f64(xyz, ec, b, O.hist_size(xyz, b))
    S = O.hist_pairsum(h, b, q)
    close(O.intensity_1d(S, ec, q, sl=nacl["sl"]), nacl["I_hist_neut"])
    close(O.intensity_1d(S, ec, q, ff=nacl["ff"]), nacl["I_hist_xray"])
    ref = nacl["rpdf_total"]
    rp = O.pdf(h, ec, b, 2, 0.04, nacl["sl"], cpu_radius_offset=True)
    close(rp[0, :ref.size], ref, atol=1e-5 * np.abs(ref).max())
    # without the CPU loop's radius offset the total differs visibly
    rp_gpu = O.pdf(h, ec, b, 2, 0.04, nacl["sl"], cpu_radius_offset=False)
    assert np.max(np.abs(rp_gpu[0, :ref.size] - ref)) > 1.0
    rd = O.pdf(h, ec, b, 0, 0.04, nacl["sl"])
    m = nacl["rdf_total"] != 0
    close(rd[0, :m.size][m], nacl["rdf_total"][m])


def test_partial_debye_two_blocks():
    g = golden("partial_2blocks")
    xyz = g["xyz"].astype(np.float64)
    S = O.partial_debye_pairsum(xyz, g["el_count"], g["blk_count"], 2, g["q"])
    I = O.partial_intensity_1d(S, g["el_count"], g["blk_count"], 2, g["q"], sl=g["sl"])
    close(I[0], g["I_total"])
    for k, key in enumerate(("00", "01", "11")):
        close(I[1 + k], g["I_part_" + key])


@pytest.mark.parametrize("conv", range(12))
def test_2d_all_conventions(conv):
    g = golden("amp2d_small")
    xyz = g["xyz"].astype(np.float64)
    emin = g["euler_min_deg"] / 180. * np.pi
    emax = g["euler_max_deg"] / 180. * np.pi
    lam = 4. * np.pi / 6.0            # default wavelength 4 pi / qmax (ReadXML.cpp:219)
    I2, I1 = O.calc_int2d(xyz, g["el_count"], g["q"], int(g["nfi"]), emin, emax, g["euler_n"],
                          conv, lam, sl=g["sl"])
    close(I2, g[f"I2D_{conv}"])
    close(I1, g[f"I1D_{conv}"])


def test_2d_xray_polar():
    g = golden("amp2d_small")
    xyz = g["xyz"].astype(np.float64)
    emin = g["euler_min_deg"] / 180. * np.pi
    emax = g["euler_max_deg"] / 180. * np.pi
    I2, I1 = O.calc_int2d(xyz, g["el_count"], g["q"], int(g["nfi"]), emin, emax, g["euler_n"],
                          5, 1.0, polar=True, ff=g["ff"])
    close(I2, g["I2D_xray_polar"])
    close(I1, g["I1D_xray_polar"])


def test_cpu_calibration_record():
    """The oracle's C restatement against the reference's own OpenMP
    executable, side by side on the same inputs at 8 threads in the build
    container (tests/golden/make_cpu_calibration.py): times within 15% of
    each other, outputs equal to the reference's printed 7 digits."""
    with open(os.path.join(os.path.dirname(__file__), "golden", "cpu_calibration.json")) as f:
        cal = json.load(f)
    assert cal["threads"] == 8
    for case in ("debye_fcc1985", "hist_rocksalt64000"):
        assert 0.85 <= cal[case]["port_over_reference"] <= 1.15, case
    assert cal["debye_fcc1985"]["I_max_rel_diff"] < 1e-6
    h = cal["hist_rocksalt64000"]
    assert h["pair_total_port"] == h["pairs"]
    assert h["rdf_max_rel_diff"] < 1e-6

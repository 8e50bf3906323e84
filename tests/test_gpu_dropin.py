"""End to end: the reference's own driver (main.cpp, ReadXML.cpp, IO.cpp)
linked with the HIP drop-in (oracle/_ref/XaNSoNS_HIP, built by
`make -C oracle hipref` where the reference sources exist) reproduces the
reference CPU executable's printed outputs on the golden inputs.  Both sides
print 7 significant digits; RDF-derived histogram counts must match exactly."""
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from conftest import REPO, golden, rel_err

sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import make_golden as mg  # noqa: E402
from xansons_hip import structures as st  # noqa: E402

pytestmark = pytest.mark.gpu
EXE = os.path.join(REPO, "oracle", "_ref", "XaNSoNS_HIP")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in executable not built")


def run(tmp, xml_text, env=None):
    with open(os.path.join(tmp, "start.xml"), "w") as f:
        f.write(xml_text)
    out = subprocess.run([EXE, "start.xml"], cwd=tmp, capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, **(env or {})))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


def col(tmp, name):
    return np.loadtxt(os.path.join(tmp, name), ndmin=2)[:, 1]


@pytest.fixture
def tmp():
    d = tempfile.mkdtemp(prefix="xs_dropin_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def ens_of(g, names):
    return st.Ensemble(g["xyz"], [int(c) for c in g["el_count"]], names)


@needs_exe
def test_fcc_debye_hist_pdf(tmp):
    g = golden("fcc1985")
    ens = ens_of(g, ["Au"])
    mg.write_atoms(os.path.join(tmp, "atoms.txt"), ens)
    mg.write_ff(os.path.join(tmp, "ff.txt"), ens.names, g["q"], g["ff"])
    mg.write_sl(os.path.join(tmp, "sl.txt"), ens.names, g["sl"])
    qel = '<q min="0.5" max="10" N="1024"/>'
    run(tmp, mg.xml(dict(name="A1", source="xray", scenario="Debye", FFfilename="ff.txt"), [qel],
                    ["atoms.txt"]))
    assert rel_err(col(tmp, "A1_xray_1D.txt"), g["I_debye"])[0] < 2e-6
    run(tmp, mg.xml(dict(name="A2", source="xray", scenario="Debye_hist", FFfilename="ff.txt",
                         hist_bin="0.01"), [qel], ["atoms.txt"]))
    assert rel_err(col(tmp, "A2_xray_1D.txt"), g["I_hist"])[0] < 2e-6
    run(tmp, mg.xml(dict(name="A3", source="neutron", scenario="PDFonly", FFfilename="sl.txt",
                         hist_bin="0.01"), ['<PDF type="RDF" density="0.05"/>'], ["atoms.txt"]))
    assert np.array_equal(col(tmp, "A3_neut_RDF.txt"), g["rdf"])


@needs_exe
def test_nacl_debyepdf_partials(tmp):
    g = golden("nacl_small")
    ens = ens_of(g, ["Na", "Cl"])
    mg.write_atoms(os.path.join(tmp, "atoms.txt"), ens)
    mg.write_sl(os.path.join(tmp, "sl.txt"), ens.names, g["sl"])
    qel = '<q min="0.2" max="25" N="512"/>'
    run(tmp, mg.xml(dict(name="B3", source="neutron", scenario="DebyePDF", FFfilename="sl.txt",
                         hist_bin="0.005"),
                    [qel, '<PDF type="RDF" density="0.04" PrintPartial="yes"/>'], ["atoms.txt"]))
    assert rel_err(col(tmp, "B3_neut_1D.txt"), g["I_hist_neut"])[0] < 2e-6
    assert np.array_equal(col(tmp, "B3_neut_RDF.txt"), g["rdf_total"])
    for key in ("Cl_Cl", "Cl_Na", "Na_Na"):
        assert np.array_equal(col(tmp, f"B3_neut_partial_RDF_{key}.txt"), g["partial_rdf_" + key])
    run(tmp, mg.xml(dict(name="B4", source="neutron", scenario="DebyePDF", FFfilename="sl.txt",
                         hist_bin="0.005"),
                    [qel, '<PDF type="rPDF" density="0.04" PrintPartial="yes"/>'], ["atoms.txt"]))
    assert np.array_equal(col(tmp, "B4_neut_rPDF.txt"), g["rpdf_total"])


@needs_exe
def test_partial_intensity_two_blocks(tmp):
    g = golden("partial_2blocks")
    x = g["xyz"]
    ni0, ni1, no0, no1 = (int(v) for v in g["blk_count"])
    b0 = st.Ensemble(np.ascontiguousarray(np.concatenate([x[:, :ni0], x[:, ni0 + ni1:ni0 + ni1 + no0]], 1)),
                     [ni0, no0], ["Ni", "O"])
    b1 = st.Ensemble(np.ascontiguousarray(np.concatenate([x[:, ni0:ni0 + ni1], x[:, ni0 + ni1 + no0:]], 1)),
                     [ni1, no1], ["Ni", "O"])
    mg.write_atoms(os.path.join(tmp, "b0.txt"), b0)
    mg.write_atoms(os.path.join(tmp, "b1.txt"), b1)
    mg.write_sl(os.path.join(tmp, "sl.txt"), ["Ni", "O"], g["sl"])
    run(tmp, mg.xml(dict(name="C", source="neutron", scenario="Debye", FFfilename="sl.txt",
                         PartialIntensity="yes"), ['<q min="0.3" max="12" N="256"/>'],
                    ["b0.txt", "b1.txt"]))
    assert rel_err(col(tmp, "C_neut_1D.txt"), g["I_total"])[0] < 2e-6
    for key in ("00", "01", "11"):
        assert rel_err(col(tmp, f"C_neut_{key[0]}-{key[1]}_1D.txt"), g["I_part_" + key])[0] < 2e-6


@needs_exe
@pytest.mark.parametrize("conv", list(range(12)))
def test_2d(tmp, conv):
    g = golden("amp2d_small")
    ens = ens_of(g, ["Si", "O"])
    mg.write_atoms(os.path.join(tmp, "atoms.txt"), ens)
    mg.write_sl(os.path.join(tmp, "sl.txt"), ens.names, g["sl"])
    cv = mg.CONVENTIONS[conv]
    run(tmp, mg.xml(dict(name="D", source="neutron", scenario="2D", FFfilename="sl.txt"),
                    ['<q min="0.5" max="6" N="10" Nfi="12"/>',
                     f'<Euler convention="{cv}" min="10 20 30" max="100 80 200" N="2 2 3"/>'],
                    ["atoms.txt"]))
    I2 = np.loadtxt(os.path.join(tmp, "D_neut_2D.txt"), ndmin=2)
    assert rel_err(I2, g[f"I2D_{conv}"])[0] < 2e-6
    assert rel_err(col(tmp, "D_neut_1D.txt"), g[f"I1D_{conv}"])[0] < 2e-6


@needs_exe
def test_2d_xray_polar(tmp):
    """The 2D polarisation branch (x-ray form factors, PolarFactor with a
    given wavelength) against the reference's 2D and 1D files."""
    g = golden("amp2d_small")
    ens = ens_of(g, ["Si", "O"])
    mg.write_atoms(os.path.join(tmp, "atoms.txt"), ens)
    mg.write_ff(os.path.join(tmp, "ff.txt"), ens.names, g["q"], g["ff"])
    run(tmp, mg.xml(dict(name="DX", source="xray", scenario="2D", FFfilename="ff.txt",
                         PolarFactor="yes", wavelength="1.0"),
                    ['<q min="0.5" max="6" N="10" Nfi="12"/>',
                     '<Euler convention="ZXZ" min="10 20 30" max="100 80 200" N="2 2 3"/>'],
                    ["atoms.txt"]))
    I2 = np.loadtxt(os.path.join(tmp, "DX_xray_2D.txt"), ndmin=2)
    assert rel_err(I2, g["I2D_xray_polar"])[0] < 2e-6
    assert rel_err(col(tmp, "DX_xray_1D.txt"), g["I1D_xray_polar"])[0] < 2e-6


@needs_exe
@pytest.mark.parametrize("case", ["rocksalt", "hex_copies", "shared_sites", "single_cell"])
def test_atoms_file_equals_reference(tmp, case):
    """The drop-in's CalcAndPrintAtoms (GPU builder) prints the same
    <name>_atoms.xyz as the reference's CPU builder, byte for byte."""
    g = np.load(os.path.join(REPO, "tests", "golden", f"atoms_{case}.npz"))
    with open(os.path.join(tmp, "sl.txt"), "w") as f:
        f.write(str(g["sl"]))
    out = run(tmp, str(g["xml"]))
    assert "Total calculation time" in out
    with open(os.path.join(tmp, f"{case}_atoms.xyz")) as f:
        assert f.read() == str(g["xyz_text"])


@needs_exe
@pytest.mark.parametrize("case", ["nacl", "xray_partial", "cluster"])
def test_rearrangement_equals_reference(tmp, case):
    """Debye + RearrangementInt (main.cpp:179): the drop-in's printed
    intensity files against the reference executable's (7 digits)."""
    import json
    g = np.load(os.path.join(REPO, "tests", "golden", f"rearr_{case}.npz"))
    files = json.loads(str(g["files_json"]))
    for fn, text in files.items():
        if fn != "start.xml":
            with open(os.path.join(tmp, fn), "w") as f:
                f.write(text)
    run(tmp, files["start.xml"])
    printed = json.loads(str(g["printed_json"]))
    assert printed
    for fn, text in printed.items():
        want = np.loadtxt(text.splitlines(), ndmin=2)[:, 1]
        got = col(tmp, fn)
        assert rel_err(got, want)[0] < 2e-6, fn


OMP_EXE = os.path.join(REPO, "oracle", "_ref", "XaNSoNS_OMP")


@needs_exe
@pytest.mark.skipif(not os.path.exists(OMP_EXE), reason="reference CPU executable not built")
def test_histogram_stage_against_reference_cpu(tmp):
    """The reference's own driver with the drop-in and the reference CPU
    executable (16 threads) on one input: a rock-salt cube of 36^3
    conventional cells (373,248 atoms, 7.0e10 pairs) built by the
    reference's ensemble builder from an 8-atom cell with Uiso = 0 (so both
    programs place the same atoms), RDF at 0.005 A.  The printed RDFs agree
    to their 7 digits; both programs' own stage timers are printed.  The
    drop-in's histogram stage is its process's first histogram call: the
    host copy of the coordinates, the copies over PCIe, the tiling and the
    first launches of the histogram kernels (the GPU context already exists:
    the drop-in's ensemble builder made it)."""
    fcc = [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]
    atoms = "".join(f'      <Atom name="{nm}" r="{f[0] + dx} {f[1]} {f[2]}"/>\n'
                    for nm, dx in (("Na", 0.0), ("Cl", 0.5)) for f in fcc)
    with open(os.path.join(tmp, "sl.txt"), "w") as f:
        f.write("Na 3.63\nCl 9.577\n")
    xml = ('<?xml version="1.0"?>\n<Config>\n  <Calculation name="H" source="neutron" '
           'scenario="PDFonly" FFfilename="sl.txt" hist_bin="0.005">\n'
           '    <PDF type="RDF" density="0.05"/>\n  </Calculation>\n'
           '  <Block fractional="Yes">\n    <Atoms>\n' + atoms + '    </Atoms>\n'
           '    <CellVectors a="5.64 0 0" b="0 5.64 0" c="0 0 5.64" N="36 36 36"/>\n'
           '  </Block>\n</Config>\n')
    pat = r"(?:Histogram|Total) calculation time: ([0-9.eE+-]+) s"
    out = run(tmp, xml)
    got = col(tmp, "H_neut_RDF.txt")
    t_drop = [float(v) for v in re.findall(pat, out)]
    os.remove(os.path.join(tmp, "H_neut_RDF.txt"))
    threads = max(1, min(16, os.cpu_count() or 1))
    ref = subprocess.run([OMP_EXE, "start.xml", str(threads)], cwd=tmp, capture_output=True, text=True,
                         timeout=600, env=dict(os.environ, OMP_PROC_BIND="close"))
    assert ref.returncode == 0, ref.stdout[-2000:] + ref.stderr[-2000:]
    want = col(tmp, "H_neut_RDF.txt")
    t_ref = [float(v) for v in re.findall(pat, ref.stdout)]
    print(f"373,248 atoms, 6.966e+10 pairs; histogram / total stage seconds: drop-in {t_drop}, "
          f"reference ({threads} threads) {t_ref}", flush=True)
    assert got.shape == want.shape
    m = want != 0
    assert np.array_equal(got == 0, want == 0)
    assert np.max(np.abs(got[m] - want[m]) / np.abs(want[m])) < 1e-6
    assert t_drop and t_ref

"""Golden vectors for RearrangementInt (CalcFunctions.cpp:469-545) from the
reference's own code.

Run in the build container only (it needs the reference sources at
/root/reference to build oracle/_ref/ref_atoms and oracle/_ref/XaNSoNS):

    python tests/golden/make_rearr_golden.py

For each case the script writes a start XML with <Rearrangement> blocks, the
neighbour file(s) and the SL / FF table, then
  * runs oracle/_ref/ref_atoms --rearrangement: the reference's ReadConfig,
    CalcAndPrintAtoms and RearrangementInt on I0[k] = 1 + k/64, printing the
    parsed blocks (with neighbour lists), q, SL/FF and the resulting I
    (hex floats, exact);
  * runs the reference executable oracle/_ref/XaNSoNS on the same input and
    keeps its printed intensity files (Debye + rearrangement, 7 digits).
Both are stored in rearr_<case>.npz.  The fixtures are data: inputs and the
reference's outputs.
"""
from __future__ import annotations

import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import make_atoms_golden as mag  # noqa: E402
import make_golden as mg  # noqa: E402

REF_EXE = os.path.join(REPO, "oracle", "_ref", "XaNSoNS")


def neighb_text(lists):
    """Neighbour file (block::readRearrangement, IO.cpp:198-227): per cell
    atom 'iAtom NeighbMax', then one 'name x y z' line per neighbour."""
    out = []
    for i, nb in enumerate(lists):
        out.append(f"{i} {len(nb)}")
        out.extend(f"{n} {mag.vec(r)}" for n, r in nb)
    return "\n".join(out) + "\n"


def run_case(case, calc_attrs, children, blocks_xml, files):
    tmp = tempfile.mkdtemp(prefix="xs_rearr_golden_")
    try:
        attrs = " ".join(f'{k}="{v}"' for k, v in calc_attrs.items())
        xml = ('<?xml version="1.0"?>\n<Config>\n' + f'  <Calculation name="{case}" {attrs}>\n' +
               "".join("    " + c + "\n" for c in children) + "  </Calculation>\n" +
               "".join(blocks_xml) + "</Config>\n")
        files = dict(files, **{"start.xml": xml})
        for fn, text in files.items():
            with open(os.path.join(tmp, fn), "w") as f:
                f.write(text)
        subprocess.run([mag.HARNESS, "start.xml", "out.json", "--rearrangement"], cwd=tmp,
                       check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(tmp, "out.json")) as f:
            d = json.load(f)
        os.remove(os.path.join(tmp, "out.json"))
        for p in glob.glob(os.path.join(tmp, "*_1D.txt")) + glob.glob(os.path.join(tmp, "*_atoms.xyz")):
            os.remove(p)
        subprocess.run([REF_EXE, "start.xml"], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        printed = {os.path.basename(p): open(p).read()
                   for p in sorted(glob.glob(os.path.join(tmp, "*_1D.txt")))}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    h = lambda a: np.array([float.fromhex(v) for v in a])  # noqa: E731
    np.savez_compressed(
        os.path.join(HERE, f"rearr_{case}.npz"),
        files_json=np.array(json.dumps(files)), blocks_json=np.array(json.dumps(d["blocks"])),
        nel=np.array(d["nel"]), el_count=np.array(d["el_count"], np.uint32),
        block_el_count=np.array(d["block_el_count"], np.uint32).reshape(len(d["blocks"]), d["nel"]),
        source=np.array(d["source"]), partial=np.array(d["partial"]), polar=np.array(d["polar"]),
        lam=np.array(float.fromhex(d["lambda"])), q=h(d["q"]), sl=h(d["sl"]),
        ff=np.array([h(r) for r in d["ff"]]) if d["ff"] else np.zeros((0, len(d["q"]))),
        rearr_I=h(d["rearr_I"]), printed_json=np.array(json.dumps(printed)))
    print(f"{case}: atoms {d['el_count']}, RearrangementInt {d['rearr_seconds']:.4f} s, "
          f"printed {sorted(printed)}")


def jitter(rng, r, s):
    return tuple(np.asarray(r) + rng.uniform(-s, s, 3))


def case_nacl():
    """Rock-salt conventional cell listed Na, Cl alternately (sortAtoms
    reorders the atoms but not the neighbour lists, as the reference does),
    2x2x2 cells, each atom's six nearest neighbours moved by up to 0.04 A."""
    rng = np.random.default_rng(101)
    a = 5.64
    na = [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]
    atoms = []
    for f in na:
        atoms.append(("Na", np.array(f) * a))
        atoms.append(("Cl", (np.array(f) + [0.5, 0, 0]) * a))
    nb = []
    for name, r in atoms:
        other = "Cl" if name == "Na" else "Na"
        nb.append([(other, jitter(rng, r + d * a / 2, 0.04))
                   for d in np.vstack([np.eye(3), -np.eye(3)])])
    blk = mag.block_xml([(n, r, 1.0, 0.0) for n, r in atoms],
                        cell=((a, 0, 0), (0, a, 0), (0, 0, a)), ncell=(2, 2, 2),
                        extra=['<Rearrangement filename="nb.txt" Rcorr="0.15"/>'])
    run_case("nacl", dict(source="neutron", scenario="Debye", FFfilename="sl.txt"),
             ['<q min="0.5" max="8" N="64"/>'], [blk],
             {"sl.txt": "Na 3.63\nCl 9.577\n", "nb.txt": neighb_text(nb)})


def case_xray_partial():
    """X-ray form factors, polarisation factor, two blocks with partial
    intensities: block 0 (60 atoms, 2 cells along a) with random neighbour
    lists of 0-5 entries, block 1 (20 atoms) without rearrangement."""
    rng = np.random.default_rng(102)
    names = ["Si", "O"]
    r0 = rng.uniform(0, 9.0, (60, 3))
    e0 = rng.integers(0, 2, 60)
    nb = [[(names[rng.integers(0, 2)], jitter(rng, r0[i], 1.8)) for _ in range(rng.integers(0, 6))]
          for i in range(60)]
    b0 = mag.block_xml([(names[e], r, 1.0, 0.0) for e, r in zip(e0, r0)],
                       cell=((9.5, 0, 0), (0, 9.5, 0), (0, 0, 9.5)), ncell=(2, 1, 1),
                       extra=['<Rearrangement filename="nb.txt" Rcorr="0.4"/>'])
    r1 = rng.uniform(20, 28, (20, 3))
    b1 = mag.block_xml([(names[e], r, 1.0, 0.0) for e, r in zip(rng.integers(0, 2, 20), r1)])
    q = np.array([0.3 + k * (9.0 - 0.3) / 48 for k in range(48)])
    ff = np.array([14.0 * np.exp(-0.02 * q ** 2) + 0.5, 8.0 * np.exp(-0.03 * q ** 2) + 0.2])
    tmp = tempfile.mkdtemp()
    try:
        mg.write_ff(os.path.join(tmp, "ff.txt"), names, q, ff)
        ff_text = open(os.path.join(tmp, "ff.txt")).read()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    run_case("xray_partial",
             dict(source="xray", scenario="Debye", FFfilename="ff.txt", PolarFactor="Yes",
                  wavelength="1.54", PartialIntensity="Yes"),
             ['<q min="0.3" max="9.0" N="48"/>'], [b0, b1],
             {"ff.txt": ff_text, "nb.txt": neighb_text(nb)})


def case_cluster():
    """A 500-atom three-element cell; each atom's neighbours are its actual
    neighbours within 2.7 A, moved by up to 0.1 A."""
    rng = np.random.default_rng(103)
    names = ["Na", "Cl", "Au"]
    r = rng.uniform(0, 20.9, (500, 3))
    e = rng.integers(0, 3, 500)
    d = np.linalg.norm(r[:, None] - r[None], axis=-1)
    nb = [[(names[e[j]], jitter(rng, r[j], 0.1)) for j in np.nonzero((d[i] < 2.7) & (d[i] > 0))[0]]
          for i in range(500)]
    blk = mag.block_xml([(names[k], p, 1.0, 0.0) for k, p in zip(e, r)],
                        extra=['<Rearrangement filename="nb.txt" Rcorr="0.3"/>'])
    run_case("cluster", dict(source="neutron", scenario="Debye", FFfilename="sl.txt"),
             ['<q min="0.5" max="10" N="32"/>'], [blk],
             {"sl.txt": "Na 3.63\nCl 9.577\nAu 7.63\n", "nb.txt": neighb_text(nb)})


def main():
    mag.build_harness()
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    case_nacl()
    case_xray_partial()
    case_cluster()


if __name__ == "__main__":
    main()

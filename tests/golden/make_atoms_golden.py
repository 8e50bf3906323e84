"""Golden vectors for the ensemble builder (block::calcAtoms,
CalcFunctions.cpp:251-397) from the reference's own code.

Run in the build container only (it needs the reference sources at
/root/reference to build oracle/_ref/ref_atoms):

    python tests/golden/make_atoms_golden.py

For each case the script writes a start XML (and the SL table it names),
runs oracle/_ref/ref_atoms -- the reference's ReadConfig and
CalcAndPrintAtoms linked with a small dump harness -- and stores in
atoms_<case>.npz:
  * the parsed blocks as the builder receives them (JSON with hex floats,
    exact), the XML and SL text;
  * deterministic cases (every occupancy decided, no displacements): the
    placed atoms per element (float64, exact), per-block counts and the
    reference's printed <name>_atoms.xyz;
  * random cases: the reference's per-element counts and displacement /
    orientation statistics of its atoms (its builder is clock-seeded, so only
    distributions can be compared).
The fixtures are data: inputs and the reference's outputs.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_atoms")

SL = {"Na": 3.63, "Cl": 9.577, "Zn": 5.68, "O": 5.803, "Au": 7.63}


def build_harness():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "refatoms"], check=True)


def calc_node(name, scenario="Debye"):
    return (f'  <Calculation name="{name}" source="neutron" scenario="{scenario}" '
            f'FFfilename="sl.txt" PrintAtoms="Yes">\n    <q min="0.5" max="5" N="4"/>\n'
            '  </Calculation>\n')


def vec(v):
    return " ".join(repr(float(x)) for x in v)


def block_xml(atoms, cell=((1, 0, 0), (0, 1, 0), (0, 0, 1)), ncell=(1, 1, 1), symm=(),
              copies=(), rcut=None, rcut_copies=None, attrs=None, extra=()):
    """atoms: (name, r, occ, Uiso); symm: (r, R1, R2, R3); copies: (r, euler);
    extra: further child elements (XML lines)."""
    a = " ".join(f'{k}="{v}"' for k, v in (attrs or {}).items())
    out = [f"  <Block {a}>" if a else "  <Block>", "    <Atoms>"]
    for name, r, occ, uiso in atoms:
        out.append(f'      <Atom name="{name}" r="{vec(r)}" occ="{occ!r}" Uiso="{uiso!r}"/>')
    out.append("    </Atoms>")
    for r, r1, r2, r3 in symm:
        out.append(f'    <SymmEqPos r="{vec(r)}" R1="{vec(r1)}" R2="{vec(r2)}" R3="{vec(r3)}"/>')
    if rcut is not None or rcut_copies is not None:
        c = []
        if rcut is not None:
            c.append(f'Rcut="{rcut!r}"')
        if rcut_copies is not None:
            c.append(f'RcutCopies="{rcut_copies!r}"')
        out.append("    <CutOff " + " ".join(c) + "/>")
    if copies:
        out.append('    <Copies convention="ZYX">')
        for r, eu in copies:
            out.append(f'      <Copy r="{vec(r)}" Euler="{vec(eu)}"/>')
        out.append("    </Copies>")
    out.extend("    " + e for e in extra)
    out.append(f'    <CellVectors a="{vec(cell[0])}" b="{vec(cell[1])}" c="{vec(cell[2])}" '
               f'N="{" ".join(str(int(n)) for n in ncell)}"/>')
    out.append("  </Block>")
    return "\n".join(out) + "\n"


def run_harness(name, blocks_xml, names):
    tmp = tempfile.mkdtemp(prefix="xs_atoms_golden_")
    try:
        xml = '<?xml version="1.0"?>\n<Config>\n' + calc_node(name) + "".join(blocks_xml) + "</Config>\n"
        sl = "".join(f"{n} {SL[n]!r}\n" for n in names)
        with open(os.path.join(tmp, "start.xml"), "w") as f:
            f.write(xml)
        with open(os.path.join(tmp, "sl.txt"), "w") as f:
            f.write(sl)
        subprocess.run([HARNESS, "start.xml", "out.json"], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL)
        with open(os.path.join(tmp, "out.json")) as f:
            d = json.load(f)
        with open(os.path.join(tmp, f"{name}_atoms.xyz")) as f:
            xyz_text = f.read()
        return d, xml, sl, xyz_text
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def atoms_of(d):
    """Per-element (n, 3) float64 arrays from the harness's hex dump."""
    return [np.array([float.fromhex(v) for v in a], dtype=np.float64).reshape(-1, 3)
            for a in d["atoms"]]


def save_deterministic(case, d, xml, sl, xyz_text):
    per_el = atoms_of(d)
    np.savez_compressed(
        os.path.join(HERE, f"atoms_{case}.npz"),
        blocks_json=np.array(json.dumps(d["blocks"])), xml=np.array(xml), sl=np.array(sl),
        xyz_text=np.array(xyz_text), nel=np.array(d["nel"]),
        el_count=np.array(d["el_count"], np.uint32),
        block_el_count=np.array(d["block_el_count"], np.uint32).reshape(len(d["blocks"]), d["nel"]),
        atoms=np.concatenate(per_el) if per_el else np.zeros((0, 3)))
    print(f"{case}: {sum(d['el_count'])} atoms, el_count {d['el_count']}, "
          f"reference builder {d['seconds']:.4f} s")


# --- deterministic cases ---------------------------------------------------

def case_rocksalt():
    """Fractional rock-salt basis, fcc centring translations and an
    inversion: every Cl image under the inversion lands on another image
    one cell away, so half of the sites merge."""
    I = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    symm = [((0.5, 0.5, 0), *I), ((0.5, 0, 0.5), *I), ((0, 0.5, 0.5), *I),
            ((0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1))]
    a = 5.64
    blk = block_xml([("Na", (0, 0, 0), 1.0, 0.0), ("Cl", (0.5, 0, 0), 1.0, 0.0)],
                    cell=((a, 0, 0), (0, a, 0), (0, 0, a)), ncell=(3, 3, 3), symm=symm,
                    attrs={"fractional": "Yes"})
    save_deterministic("rocksalt", *run_harness("rocksalt", [blk], ["Na", "Cl"]))


def case_hex_copies():
    """Wurtzite-type cell (non-orthogonal axes), centred, cut to a sphere,
    three rotated copies and a cut-off around the copies' centre."""
    a, c, u = 3.25, 5.21, 0.382
    cell = ((a, 0, 0), (-a / 2, a * np.sqrt(3) / 2, 0), (0, 0, c))
    atoms = [("Zn", (1 / 3, 2 / 3, 0), 1.0, 0.0), ("Zn", (2 / 3, 1 / 3, 0.5), 1.0, 0.0),
             ("O", (1 / 3, 2 / 3, u), 1.0, 0.0), ("O", (2 / 3, 1 / 3, 0.5 + u), 1.0, 0.0)]
    copies = [((0, 0, 0), (0, 0, 0)), ((21.5, -3.0, 4.0), (30, 45, 60)),
              ((-6.0, 19.0, -11.0), (-75, 10, 135))]
    blk = block_xml(atoms, cell=cell, ncell=(8, 8, 5), copies=copies, rcut=9.5, rcut_copies=24.0,
                    attrs={"fractional": "Yes", "centered": "Yes"})
    save_deterministic("hex_copies", *run_harness("hex_copies", [blk], ["Zn", "O"]))


def case_shared_sites():
    """Two blocks.  Block 0: a Na site 0.2 A from a Cl of occupancy 0 (after
    sortAtoms the Na site comes first and lists the Cl with cumulative
    occupancy 1 + 0), a same-element pair 0.28 A apart and a mirror plane
    that maps the x = 0 Na onto itself.  Block 1: a simple cubic Cl lattice,
    appended after block 0's atoms of each element."""
    mirror = ((0, 0, 0), (-1, 0, 0), (0, 1, 0), (0, 0, 1))
    b0 = block_xml([("Cl", (1.0, 1.0, 1.0), 0.0, 0.0), ("Na", (1.0, 1.0, 1.2), 1.0, 0.0),
                    ("Na", (0.0, 2.5, 2.5), 1.0, 0.0), ("Na", (0.2, 2.5, 2.7), 1.0, 0.0),
                    ("Cl", (2.2, 0.4, 3.1), 1.0, 0.0)],
                   cell=((4.0, 0, 0), (0, 4.0, 0), (0, 0, 4.0)), ncell=(4, 3, 2), symm=[mirror])
    b1 = block_xml([("Cl", (0.5, 0.5, 0.5), 1.0, 0.0)],
                   cell=((3.0, 0, 0), (0, 3.0, 0), (0, 0, 3.0)), ncell=(4, 4, 4))
    save_deterministic("shared_sites", *run_harness("shared_sites", [b0, b1], ["Na", "Cl"]))


def case_single_cell():
    """One large 'cell' (Ncell = 1, the input format of a finished structure
    file): 800 random atoms of three elements plus 80 near-duplicates
    0.05-0.4 A from an earlier atom, which the builder must drop."""
    rng = np.random.default_rng(20161)
    base = rng.uniform(0, 24.0, size=(800, 3))
    names = np.array(["Na", "Cl", "Au"])[rng.integers(0, 3, size=800)]
    idx = rng.choice(800, size=80, replace=False)
    dirs = rng.normal(size=(80, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dup = base[idx] + dirs * rng.uniform(0.05, 0.4, size=(80, 1))
    dnames = np.array(["Na", "Cl", "Au"])[rng.integers(0, 3, size=80)]
    atoms = [(n, r, 1.0, 0.0) for n, r in zip(names, base)] + \
            [(n, r, 1.0, 0.0) for n, r in zip(dnames, dup)]
    blk = block_xml(atoms)
    save_deterministic("single_cell", *run_harness("single_cell", [blk], ["Na", "Cl", "Au"]))


# --- random cases: the reference's statistics -------------------------------

def nearest_lattice_offsets(x, a):
    return x - np.round(x / a) * a


def case_random_occupancy():
    """Simple cubic a = 4 A, 24^3 cells, Na (occ 0.35) and Cl (occ 0.40)
    sharing the site, Uiso 0.01 (sd 0.1 A): counts are binomial and each
    atom's displacement is |N(0, 0.1)| along a direction with phi uniform in
    [0, 2 pi) and theta uniform in [0, pi) (CalcFunctions.cpp:359-365)."""
    a = 4.0
    blk = block_xml([("Na", (0, 0, 0), 0.35, 0.01), ("Cl", (0, 0, 0), 0.40, 0.01)],
                    cell=((a, 0, 0), (0, a, 0), (0, 0, a)), ncell=(24, 24, 24))
    d, xml, sl, _ = run_harness("random_occ", [blk], ["Na", "Cl"])
    allat = np.concatenate(atoms_of(d))
    disp = nearest_lattice_offsets(allat, a)
    r = np.linalg.norm(disp, axis=1)
    stats = dict(mean_r=float(r.mean()), mean_r2=float((r ** 2).mean()),
                 mean_cos2=float(((disp[:, 2] / r) ** 2).mean()),
                 mean_z=float(disp[:, 2].mean()), n=int(r.size))
    np.savez_compressed(os.path.join(HERE, "atoms_random_occ.npz"),
                        blocks_json=np.array(json.dumps(d["blocks"])), xml=np.array(xml),
                        sl=np.array(sl), el_count=np.array(d["el_count"], np.uint32),
                        ncells=np.array(24 ** 3), stats_json=np.array(json.dumps(stats)))
    print("random_occ:", d["el_count"], stats)


def case_molecules():
    """Rigid two-atom molecules on a simple cubic lattice (a = 6 A, 20^3
    cells) with random molecule rotations and a molecule displacement
    (mol_Uiso 0.04, sd 0.2 A): the Na atom at the molecule origin moves by
    the molecule displacement only; the Na->Cl vector is the rotated bond."""
    a = 6.0
    bond = (1.2, 0.5, 0.3)
    blk = block_xml([("Na", (0, 0, 0), 1.0, 0.0), ("Cl", bond, 1.0, 0.0)],
                    cell=((a, 0, 0), (0, a, 0), (0, 0, a)), ncell=(20, 20, 20),
                    attrs={"mol_rotation": "Yes", "mol_Uiso": "0.04"})
    d, xml, sl, _ = run_harness("molecules", [blk], ["Na", "Cl"])
    na, cl = atoms_of(d)
    disp = nearest_lattice_offsets(na, a)
    r = np.linalg.norm(disp, axis=1)
    b = cl - na
    stats = dict(mean_r=float(r.mean()), mean_r2=float((r ** 2).mean()),
                 mean_cos2=float(((disp[:, 2] / r) ** 2).mean()),
                 bond_len_max_dev=float(np.abs(np.linalg.norm(b, axis=1) - np.linalg.norm(bond)).max()),
                 bond_mom2=(b.T @ b / len(b)).tolist(), bond_mean=b.mean(axis=0).tolist(),
                 n=int(len(b)))
    np.savez_compressed(os.path.join(HERE, "atoms_molecules.npz"),
                        blocks_json=np.array(json.dumps(d["blocks"])), xml=np.array(xml),
                        sl=np.array(sl), el_count=np.array(d["el_count"], np.uint32),
                        bond=np.array(bond), stats_json=np.array(json.dumps(stats)))
    print("molecules:", d["el_count"], stats)


def main():
    build_harness()
    case_rocksalt()
    case_hex_copies()
    case_shared_sites()
    case_single_cell()
    case_random_occupancy()
    case_molecules()


if __name__ == "__main__":
    main()

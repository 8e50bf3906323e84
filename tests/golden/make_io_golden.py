"""Raw output files of the reference's writers (IO.cpp:51-181), for
byte-level checks of xansons_hip.io.

Run in the build container only (it needs the reference sources at
/root/reference to build oracle/_ref/XaNSoNS):

    python tests/golden/make_io_golden.py

Runs the reference executable on a 2-element cluster (neutron DebyePDF with
partial PDFs, once per PDF type) and on a small 2D case, and stores every
printed file's text in io_files.npz together with the inputs.  The fixture
is data: the reference's outputs.
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
import make_golden as mg  # noqa: E402

REF_EXE = os.path.join(REPO, "oracle", "_ref", "XaNSoNS")


def run(tmp, xml_text):
    with open(os.path.join(tmp, "start.xml"), "w") as f:
        f.write(xml_text)
    subprocess.run([REF_EXE, "start.xml"], cwd=tmp, check=True, stdout=subprocess.DEVNULL)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    rng = np.random.default_rng(404)
    # Na listed first in the SL table, so element 0 is Na while the map
    # (and the partial file labels) sort Cl first
    r = rng.uniform(0, 14.0, (3, 240))
    el_count = [130, 110]
    atoms = "".join(f"{'Na' if j < el_count[0] else 'Cl'} {float(r[0, j])!r} {float(r[1, j])!r} "
                    f"{float(r[2, j])!r}\n" for j in range(240))
    sl = "Na 3.63\nCl 9.577\n"
    files = {}
    tmp = tempfile.mkdtemp(prefix="xs_io_golden_")
    try:
        for fn, text in (("atoms.txt", atoms), ("sl.txt", sl)):
            with open(os.path.join(tmp, fn), "w") as f:
                f.write(text)
        for t in ("RDF", "PDF", "rPDF"):
            run(tmp, mg.xml(dict(name=f"io{t}", source="neutron", scenario="DebyePDF",
                                 FFfilename="sl.txt", hist_bin="0.02"),
                            ['<q min="0.5" max="8" N="16"/>',
                             f'<PDF type="{t}" density="0.08" PrintPartial="Yes"/>'],
                            ["atoms.txt"]))
        run(tmp, mg.xml(dict(name="io2D", source="neutron", scenario="2D", FFfilename="sl.txt"),
                        ['<q min="0.5" max="6" N="8" Nfi="6"/>',
                         '<Euler convention="ZXZ" min="0 0 0" max="0 0 0" N="1 1 1"/>'],
                        ["atoms.txt"]))
        for p in sorted(glob.glob(os.path.join(tmp, "io*.txt"))):
            files[os.path.basename(p)] = open(p).read()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(os.path.join(HERE, "io_files.npz"), files_json=np.array(json.dumps(files)),
                        names=np.array(["Na", "Cl"]), bin=np.array(0.02))
    print(sorted(files))


if __name__ == "__main__":
    main()

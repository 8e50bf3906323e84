"""Time the reference's own CPU executable and the oracle's C restatement side
by side, on the same inputs and thread count, and record both times and the
agreement of their outputs.

Run in the build container only (it needs oracle/_ref/XaNSoNS_OMP, built from
the reference sources by `make -C oracle ref`):

    python tests/golden/make_cpu_calibration.py

Cases (8 threads each):
- the 1,985-atom fcc sphere (a = 4.078 A, radius 20 A), direct Debye I(q) on
  the reference mesh of 1,024 q over 0.5-10 A^-1 (x-ray, 4-Gaussian form
  factors): the reference's "1D pattern calculation time"
  (CalcFunctions.cpp:548-549, 631-632) against `or_debye_pairsum` +
  `or_intensity_1d`;
- a rock-salt cube of 20^3 conventional cells (64,000 atoms, 2.0e9 pairs,
  a = 5.64 A), histogram at 0.005 A: the reference's "Histogram calculation
  time" (CalcFunctions.cpp:917-918) against `or_hist_f64`.
The reference builds its ensembles with a clock-seeded RNG, so every atom is
given with Uiso = 0 (the positions are then exactly the file's).

Writes cpu_calibration.json next to this script: both times, their ratio,
and the output agreement (Debye I(q), and the total RDF of the histogram:
largest relative difference against the printed 7 digits).
"""
from __future__ import annotations

import json
import os
import re
import sys
import tempfile
import shutil
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "oracle"))
import make_golden as G  # noqa: E402
import xs_oracle as O  # noqa: E402
from xansons_hip import structures as st  # noqa: E402

THREADS = 8


def stage_time(stdout, label):
    return float(re.search(label + r" calculation time: ([0-9.eE+-]+) s", stdout).group(1))


def debye_case(tmp):
    ens = st.fcc_sphere()
    q = st.q_mesh(0.5, 10.0, 1024)
    ff = st.gaussian_ff(q, st.gaussian_ff_coeffs(1, seed=2))
    G.write_atoms(os.path.join(tmp, "atoms.txt"), ens)
    G.write_ff(os.path.join(tmp, "ff.txt"), ens.names, q, ff)
    out = G.run_ref(tmp, G.xml(dict(name="D", source="xray", scenario="Debye", FFfilename="ff.txt"),
                               ['<q min="0.5" max="10" N="1024"/>'], ["atoms.txt"]), THREADS)
    t_ref = stage_time(out, "1D pattern")
    I_ref = G.load_cols(os.path.join(tmp, "D_xray_1D.txt"))[:, 1]
    x64 = ens.xyz.astype(np.float64)
    t0 = time.perf_counter()
    S = O.debye_pairsum(x64, ens.el_count, q, THREADS)
    I = O.intensity_1d(S, ens.el_count, q, ff=ff)
    t_port = time.perf_counter() - t0
    return {"workload": f"fcc sphere, {ens.n} atoms, direct Debye, 1,024 q (0.5-10 A^-1)",
            "reference_s": t_ref, "port_s": round(t_port, 4),
            "port_over_reference": round(t_port / t_ref, 3),
            "I_max_rel_diff": float(np.max(np.abs(I - I_ref) / np.abs(I_ref)))}


def hist_case(tmp, ncell=20, bin_=0.005):
    ens = st.rocksalt_cube(ncell=ncell, vacancy=0.0, sigma=0.08, seed_vac=21, seed_jit=22)
    G.write_atoms(os.path.join(tmp, "atoms.txt"), ens)
    sl = np.array([3.63, 9.577])
    G.write_sl(os.path.join(tmp, "sl.txt"), ens.names, sl)
    out = G.run_ref(tmp, G.xml(dict(name="H", source="neutron", scenario="PDFonly", FFfilename="sl.txt",
                                    hist_bin=str(bin_)), ['<PDF type="RDF" density="0.05"/>'],
                               ["atoms.txt"]), THREADS)
    t_ref = stage_time(out, "Histogram")
    rdf = G.load_cols(os.path.join(tmp, "H_neut_RDF.txt"))
    x64 = ens.xyz.astype(np.float64)
    nh = O.hist_size(x64, bin_)
    t0 = time.perf_counter()
    h, _ = O.hist_f64(x64, ens.el_count, bin_, nh, THREADS)
    t_port = time.perf_counter() - t0
    pairs = ens.n * (ens.n - 1) // 2
    return {"workload": f"rock-salt cube {ncell}^3 cells, {ens.n} atoms, {pairs:.3e} pairs, "
                        f"histogram at {bin_} A",
            "reference_s": t_ref, "port_s": round(t_port, 4),
            "port_over_reference": round(t_port / t_ref, 3),
            "pair_total_port": int(h.sum()), "pairs": pairs,
            "rdf_columns": int(rdf.shape[1]), "rdf_rows": int(rdf.shape[0]),
            "rdf_max_rel_diff": rdf_diff(rdf, h, ens.el_count, bin_, sl)}


def rdf_diff(rdf, h, el_count, bin_, sl):
    """Largest relative difference between the printed total RDF (7
    significant digits) and the oracle's RDF of the port's counts
    (or_pdf, RDF type: sum over pair types of 2 h / (bin N) sl_a sl_b / <sl>^2,
    CalcFunctions.cpp:924-954), over the non-zero printed values."""
    want = O.pdf(h, el_count, bin_, 0, 0.05, sl)[0]
    got = rdf[:, 1]
    w = want[:got.size]
    m = got != 0
    return float(np.max(np.abs(got[m] - w[m]) / np.abs(got[m])))


def main():
    if not os.path.exists(G.REF_EXE):
        G.build_ref()
    res = {"threads": THREADS, "host": None}
    try:
        with open("/proc/cpuinfo") as f:
            res["host"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                               None)
    except OSError:
        pass
    for name, fn in (("debye_fcc1985", debye_case), ("hist_rocksalt64000", hist_case)):
        tmp = tempfile.mkdtemp(prefix="xs_cal_")
        try:
            res[name] = fn(tmp)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        print(name, res[name], flush=True)
    with open(os.path.join(HERE, "cpu_calibration.json"), "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the histogram path on one MI355X.

Workload (one step): the 980,219-atom two-element rock-salt cube
(a = 5.64 A, 50^3 cells, 2% vacancies, sigma = 0.08 A):
  1. three pair-type distance histograms, bin 0.005 A (97,251 bins, uint64)
  2. histogram -> pair sums on 4,096 q over 0.2-25 A^-1
  3. rPDF G(r) over all bins (0-483 A)
Inputs are resident in HBM before the timed region; outputs stay on the GPU.
value = pairs per second of the whole step (Gpairs/s).

A second object, "ensemble", times the batched path on 1,000 nanoparticles
(500-200,000 atoms, 1-4 elements, bin 0.01 A, 2,048 q; histograms, pair sums
and rPDF of every particle): one plan, every run replayed as a hipGraph,
against the same kernels called structure by structure.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                  [--ensemble N] [--copy-probe]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "softx-d-16-00083_amd"))

import numpy as np  # noqa: E402

from xansons_hip import structures as st  # noqa: E402

METRIC = "Gpairs/s histogram+Debye; batched patterns/s vs plain at ≤1e-5 rel err"
BIN = 0.005
NQ = 4096
FP64_PEAK_TFLOPS = 78.6           # MI355X published FP64 vector peak (binning is FP64)
VALU_ISSUE_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12   # non-packed lane-ops/s (39.3 T)
OPS_PER_PAIR = 11                 # 3 sub, 3 mul, 2 add, sqrt, div, cvt
# VALU issue cycles per 64 pairs of the histogram kernel, from rocprofv3 SQ
# counters on this workload: 11.68 VALU wave-instructions per 64 pairs, 1.04
# of them v_rsq_f64, which occupies the issue port four times as long as
# v_fma_f64 (3 extra cycles); SQ_ACTIVE_INST_VALU reads 14.79.  A CU
# issues one 64-lane VALU instruction per cycle (4 SIMDs x 16 lanes), so the
# fraction of the CUs' issue cycles at the 2.4 GHz peak clock the kernel
# spends issuing VALU work is SLOTS_PER_PAIR * pairs / 64 / (256 * 2.4e9 * t).
SLOTS_PER_PAIR = 14.79
CU_ISSUE_PER_S = 256 * 2.4e9
# FP32 binning kernel (hist_kernel_f32<32768, true>) on the same workload:
# SQ_ACTIVE_INST_VALU reads 23.89 wave-instructions per 64 pairs.  Its
# roofline fraction is the algorithmic one of the FP64 line: 11 ops per pair
# x pairs / t against MI355X's FP32 vector peak (157.3 TFLOP/s).  Beside it,
# the issue rate it reaches in wave-instructions per CU-cycle at 2.4 GHz:
# non-packed v_fma_f32 issues at 1.93x the rate of v_fma_f64 on this GPU
# (measured), i.e. up to two FP32 wave-instructions per CU-cycle against one
# for FP64, so each kernel's issue rate is read against the peak of its own
# instruction type (2 for this one, 1 for the FP64 kernel).
F32_ACTIVE_PER_64_PAIRS = 23.89
FP32_PEAK_TFLOPS = 157.3
# Direct Debye and 2D units (one op per sin, cos, add, multiply): 4 ops per
# pair-q term (sin, q r, x 1/(q r), accumulate) and 7 per complex-exponential
# term (sin, cos, 3 for q.r, 2 accumulations).  Both kernels compute in FP64,
# so the peak is the FP64 vector one.
DEBYE_OPS_PER_PAIR_Q = 4
AMP2D_OPS_PER_TERM = 7
# HBM bytes per launch of hist_kernel_f64<32768, true> on this workload, from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (KiB per dispatch), scaled by
# the counters' reading for a 1 GiB device-to-device copy (FETCH_SIZE reads
# 524,308 KiB for the 1,048,576 KiB copied, WRITE_SIZE 1,048,576 KiB).
HIST_FETCH_KIB, HIST_WRITE_KIB = 7746906, 9982160
FETCH_SCALE, WRITE_SCALE = 1048576 / 524308, 1.0
HIST_TRAFFIC_BYTES = int((HIST_FETCH_KIB * FETCH_SCALE + HIST_WRITE_KIB * WRITE_SCALE) * 1024)


def rel_errors(got, ref, floor=1e-3):
    """Error of a result against its FP64 evaluation: the largest pointwise
    relative error over the points where |ref| >= floor * max |ref|, and the
    L2-relative error ||got - ref|| / ||ref|| over all points."""
    got, ref = np.ravel(got).astype(np.float64), np.ravel(ref).astype(np.float64)
    m = np.abs(ref) >= floor * np.abs(ref).max()
    return {"max_rel": float((np.abs(got - ref)[m] / np.abs(ref[m])).max()),
            "l2_rel": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))}


def fp64_pairsums(h, bin_, q, chunk=32):
    """Histogram pair sums in FP64 numpy, S_p(q) = sum_k h_p[k] sin(q r_k) /
    (q r_k + 1e-8) with r_k = (k + 1/2) bin (CalcFunctions.cpp:856-862)."""
    h = np.atleast_2d(h)
    S = np.zeros((h.shape[0], q.size))
    for p in range(h.shape[0]):
        k = np.nonzero(h[p])[0]
        r = (k + 0.5) * bin_
        w = h[p, k].astype(np.float64)
        for i in range(0, q.size, chunk):
            qr = np.outer(q[i:i + chunk], r)
            S[p, i:i + chunk] = (np.sin(qr) / (qr + 1e-8)) @ w
    return S


def intensity_two_elements(S, el_count, sl):
    """I(q) of a two-element structure from its three pair sums (types 00, 01,
    11) and scattering lengths: (sum_e n_e sl_e^2 + 2 sum_ab sl_a sl_b S_ab) / N."""
    n = sum(el_count)
    return (el_count[0] * sl[0] ** 2 + el_count[1] * sl[1] ** 2
            + 2 * (sl[0] * sl[0] * S[0] + sl[0] * sl[1] * S[1] + sl[1] * sl[1] * S[2])) / n


def gpu_step_setup(ctx, ens, nh, q, f64):
    import torch
    dev = torch.device("cuda", 0)
    xyz = torch.from_numpy(np.ascontiguousarray(ens.xyz64() if f64 else ens.xyz)).to(dev)
    P = ens.n_el * (ens.n_el + 1) // 2
    hist = torch.empty(P * nh, dtype=torch.int64, device=dev)
    ovf = torch.empty(1, dtype=torch.int64, device=dev)
    qd = torch.from_numpy(q).to(dev)
    S = torch.empty(P * q.size, dtype=torch.float64, device=dev)
    pdf = torch.empty((1 + P) * nh, dtype=torch.float64, device=dev)
    sl = np.array([3.63, 9.577])
    rho0 = float(ens.n) / (50 * 5.64) ** 3

    def step():
        ctx.hist_dev(xyz.data_ptr(), ens.n, ens.el_count, BIN, nh, hist.data_ptr(), ovf.data_ptr(),
                     f64=f64)
        ctx.hist_pairsum_dev(hist.data_ptr(), P, nh, BIN, qd.data_ptr(), q.size, S.data_ptr())
        ctx.pdf_dev(hist.data_ptr(), ens.el_count, nh, BIN, "rPDF", rho0, sl, pdf.data_ptr())

    return step, hist, ovf, S


def time_steps(ctx, step, steps, warmup):
    import torch
    for _ in range(warmup):
        step()
    ctx.synchronize()
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern = {k: ctx.profile_read(k) for k in ("hist", "hist_prep", "hist_pairsum", "pdf")}
    ctx.profile(False)
    return (t1 - t0) * 1e3 / steps, kern


def host_cpu():
    """The host's CPU model and the CPU counts the baseline ran beside."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(m=36, a=5.64, sigma=0.08):
    """Time the reference's own CPU histogram (XaNSoNS_OMP built from the
    reference sources, its calcHist stage timer) on a bounded sample of the
    workload: a rock-salt cube of m^3 conventional cells (the workload's
    crystal, a = 5.64 A, displacements of sd 0.08 A; 8 m^3 atoms, all sites
    occupied so the pair count is exact), built by the reference's own
    ensemble builder from an 8-atom cell; m = 36 gives 373,248 atoms and
    7.0e10 pairs, about 11 s on 16 host threads.  Falls back to the oracle
    port where that executable is absent."""
    n = 8 * m ** 3
    pairs = n * (n - 1) // 2
    threads = max(1, min(16, os.cpu_count() or 1))
    host = host_cpu()
    sample = (f"rock-salt cube of {m}^3 conventional cells ({n} atoms, {pairs:.3e} pairs; the "
              f"workload's crystal, a = {a} A, displacement sd {sigma} A), bin {BIN} A, "
              "histogram stage only")
    fcc = [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]
    exe = os.path.join(REPO, "oracle", "_ref", "XaNSoNS_OMP")
    if os.path.exists(exe):
        tmp = tempfile.mkdtemp(prefix="xs_cpu_")
        try:
            atoms = "".join(f'      <Atom name="{nm}" r="{f[0] + dx} {f[1]} {f[2]}" Uiso="{sigma ** 2!r}"/>\n'
                            for nm, dx in (("Na", 0.0), ("Cl", 0.5)) for f in fcc)
            with open(os.path.join(tmp, "sl.txt"), "w") as f:
                f.write("Na 3.63\nCl 9.577\n")
            with open(os.path.join(tmp, "start.xml"), "w") as f:
                f.write('<?xml version="1.0"?>\n<Config>\n  <Calculation name="cpu" source="neutron" '
                        f'scenario="PDFonly" FFfilename="sl.txt" hist_bin="{BIN}">\n'
                        '    <PDF type="RDF" density="0.05"/>\n  </Calculation>\n'
                        '  <Block fractional="Yes">\n    <Atoms>\n' + atoms + '    </Atoms>\n'
                        f'    <CellVectors a="{a} 0 0" b="0 {a} 0" c="0 0 {a}" N="{m} {m} {m}"/>\n'
                        '  </Block>\n</Config>\n')
            env = dict(os.environ, OMP_PROC_BIND="close")
            out = subprocess.run([exe, "start.xml", str(threads)], cwd=tmp, capture_output=True,
                                 text=True, timeout=600, check=True, env=env).stdout
            t = float(re.search(r"Histogram calculation time: ([0-9.eE+-]+) s", out).group(1))
            return dict(value=pairs / t / 1e9, unit="Gpairs/s", cores=threads, kind="reference",
                        sample=sample + "; reference XaNSoNS_OMP (calcHist, FP64), its own stage timer, "
                        "OMP_PROC_BIND=close", seconds=t, host=host)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import xs_oracle as O
    r = np.arange(m, dtype=np.float64)
    cells = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    f = np.array(fcc, dtype=np.float64)
    rng = np.random.default_rng(34)
    na = ((cells[:, None] + f[None]).reshape(-1, 3)) * a
    cl = ((cells[:, None] + (f + [0.5, 0, 0])[None]).reshape(-1, 3)) * a
    xyz = np.concatenate([na, cl]).T + rng.normal(0, sigma, (3, n))
    nh = O.hist_size(xyz, BIN)
    t0 = time.perf_counter()
    O.hist_f64(np.ascontiguousarray(xyz), [n // 2, n // 2], BIN, nh, threads)
    t = time.perf_counter() - t0
    return dict(value=pairs / t / 1e9, unit="Gpairs/s", cores=threads, kind="port",
                sample=sample + "; C restatement of calcHist (FP64, zig-zag rows)", seconds=t,
                host=host)


def copy_probe(nbytes=1 << 30):
    """One device-to-device copy of a known size, to scale the HBM counters."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}))


def ensemble_leg(ctx, count, steps, warmup, use_graph=True):
    """Batched (one plan, graph replays) vs structure-by-structure runs of the
    same kernels on `count` nanoparticles: histograms (FP64 binning), pair
    sums on 2,048 q and rPDF per particle.  Both write the same layout, which
    is compared after the timed runs.  A fresh ensemble costs its plan (tiled
    on the GPU from the device coordinates) plus one run; replays of the same
    coordinates cost the run alone; both are reported, and both again with
    the pair sums on the int8 matrix cores (XS_PAIRSUM_MFMA).
    use_graph=False replays eagerly (for kernel traces)."""
    import torch
    import xansons_hip as xh
    from xansons_hip import structures as st2
    bin_, nq = 0.01, 2048
    t0 = time.perf_counter()
    parts = st2.particle_ensemble(count)
    t_gen = time.perf_counter() - t0
    pairs = sum(p.n * (p.n - 1) // 2 for p in parts)
    q = st2.q_mesh(0.2, 25.0, nq)
    # rho0 as GetAtomicDensity (CalcFunctions.cpp:212-223): N over the sphere
    # around the centroid that reaches the farthest atom
    def density(x):
        r = np.sqrt(((x - x.mean(axis=1, keepdims=True)) ** 2).sum(axis=0)).max()
        return x.shape[1] / (4.0 / 3.0 * np.pi * max(r, 1e-3) ** 3)

    rho = np.array([density(p.xyz64()) for p in parts])
    sl = np.concatenate([np.linspace(2.0, 9.0, p.n_el) for p in parts])
    bctx = xh.Context(0)          # own non-default stream (graph capture)
    dev = torch.device("cuda", 0)
    # coordinates on the device before any timing, for both forms
    xyz_d = [torch.from_numpy(np.ascontiguousarray(p.xyz64())).to(dev) for p in parts]
    torch.cuda.synchronize()
    # a plan tiled on the host (worker threads, then upload), for reference
    structs = [(np.ascontiguousarray(p.xyz64()), p.el_count) for p in parts]
    t0 = time.perf_counter()
    xh.Batch(bctx, structs, bin_, f64=True).close()
    t_plan_host = time.perf_counter() - t0
    del structs
    # the plan of a fresh ensemble, tiled on the GPU from the device
    # coordinates, is part of its first pattern
    t0 = time.perf_counter()
    b = xh.Batch(bctx, [(x.data_ptr(), p.el_count) for x, p in zip(xyz_d, parts)], bin_, f64=True,
                 device=True)
    t_plan = time.perf_counter() - t0
    qd = torch.from_numpy(q).to(dev)
    hist = torch.empty(b.hist_len, dtype=torch.int64, device=dev)
    S = torch.empty(b.ntypes_total * nq, dtype=torch.float64, device=dev)
    pdf = torch.empty(b.pdf_len, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def batched():
        b.run_dev(hist.data_ptr(), qd.data_ptr(), nq, S.data_ptr(), "rPDF", rho, sl,
                  pdf.data_ptr(), graph=use_graph)

    # a fresh ensemble: plan + first run (eager launches, as a new plan's
    # first run would be), timed end to end
    t0 = time.perf_counter()
    b.run_dev(hist.data_ptr(), qd.data_ptr(), nq, S.data_ptr(), "rPDF", rho, sl, pdf.data_ptr())
    bctx.synchronize()
    t_first = time.perf_counter() - t0
    for _ in range(warmup):
        batched()
    bctx.synchronize()
    bctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        batched()
    bctx.synchronize()
    t_b = (time.perf_counter() - t0) / steps
    g_ms = bctx.profile_read("batch_graph")
    bctx.profile(False)
    # kernel split of one eager run (same launches, outside the graph)
    bctx.profile(True)
    b.run_dev(hist.data_ptr(), qd.data_ptr(), nq, S.data_ptr(), "rPDF", rho, sl, pdf.data_ptr())
    bctx.synchronize()
    split_ms = {k: round(bctx.profile_read(k)[0], 3) for k in ("hist", "batch_pairsum", "batch_pdf")}
    bctx.profile(False)

    # the same plan with the pair sums on the int8 matrix cores
    # (XS_PAIRSUM_MFMA): its first run (column groups planned and uploaded),
    # then replays, then the kernel split of one eager run
    Sm = torch.empty_like(S)
    torch.cuda.synchronize()

    def batched_mfma(graph):
        b.run_dev(hist.data_ptr(), qd.data_ptr(), nq, Sm.data_ptr(), "rPDF", rho, sl,
                  pdf.data_ptr(), graph=graph, mfma=True)

    t0 = time.perf_counter()
    batched_mfma(False)
    bctx.synchronize()
    t_first_m = time.perf_counter() - t0
    for _ in range(warmup):
        batched_mfma(use_graph)
    bctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batched_mfma(use_graph)
    bctx.synchronize()
    t_bm = (time.perf_counter() - t0) / steps
    bctx.profile(True)
    batched_mfma(False)
    bctx.synchronize()
    pairsum_mfma_ms = round(bctx.profile_read("batch_pairsum")[0], 3)
    bctx.profile(False)
    torch.cuda.synchronize()
    sm_rel = float(((Sm - S).abs().max() / S.abs().max()).item())
    # I(q) of every particle from both pair sums (the reference's combination:
    # sum_e N_e sl_e^2 + 2 sum_{a<=b} sl_a sl_b S_ab, over N), pointwise
    # where I >= 1e-3 max I
    S_h, Sm_h = S.cpu().numpy(), Sm.cpu().numpy()
    im_rel, im_l2, e0 = 0.0, 0.0, 0
    for s_i, p in enumerate(parts):
        sle, ec = sl[e0:e0 + p.n_el], p.el_count
        e0 += p.n_el
        t0_ = int(b.type_first[s_i])
        w = [sle[a_] * sle[b_] for a_ in range(p.n_el) for b_ in range(a_, p.n_el)]
        self_ = sum(ec[a_] * sle[a_] ** 2 for a_ in range(p.n_el))

        def inten(SS):
            rows = SS[t0_ * nq:(t0_ + len(w)) * nq].reshape(len(w), nq)
            return (self_ + 2.0 * (np.asarray(w)[:, None] * rows).sum(axis=0)) / p.n
        e_ = rel_errors(inten(Sm_h), inten(S_h))
        im_rel, im_l2 = max(im_rel, e_["max_rel"]), max(im_l2, e_["l2_rel"])

    # structure by structure, same kernels, outputs in the batch layout
    hist2 = torch.empty_like(hist)
    S2 = torch.empty_like(S)
    pdf2 = torch.empty_like(pdf)
    torch.cuda.synchronize()

    def plain():
        e0 = 0
        for s, p in enumerate(parts):
            P = p.n_el * (p.n_el + 1) // 2
            nh = int(b.nhist[s])
            hp = hist2.data_ptr() + 8 * int(b.hist_off[s])
            bctx.hist_dev(xyz_d[s].data_ptr(), p.n, p.el_count, bin_, nh, hp, None, f64=True)
            bctx.hist_pairsum_dev(hp, P, nh, bin_, qd.data_ptr(), nq,
                                  S2.data_ptr() + 8 * int(b.type_first[s]) * nq)
            bctx.pdf_dev(hp, p.el_count, nh, bin_, "rPDF", rho[s], sl[e0:e0 + p.n_el],
                         pdf2.data_ptr() + 8 * int(b.pdf_off[s]), cpu_radius=True)
            e0 += p.n_el

    plain()
    bctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain()
    bctx.synchronize()
    t_p = (time.perf_counter() - t0) / steps
    torch.cuda.synchronize()
    hist_eq = bool(torch.equal(hist, hist2))
    pdf_eq = bool(torch.equal(pdf, pdf2))
    s_rel = float(((S - S2).abs().max() / S2.abs().max()).item())

    # structure by structure with the reference GPU binning (FP32 coordinates,
    # IEEE float sqrt and division), the plain FP32 form: one timed pass
    xyz32_d = [torch.from_numpy(np.ascontiguousarray(p.xyz)).to(dev) for p in parts]
    hist3 = torch.empty_like(hist)

    def plain_f32():
        e0 = 0
        for s, p in enumerate(parts):
            P = p.n_el * (p.n_el + 1) // 2
            nh = int(b.nhist[s])
            hp = hist3.data_ptr() + 8 * int(b.hist_off[s])
            bctx.hist_dev(xyz32_d[s].data_ptr(), p.n, p.el_count, bin_, nh, hp, None, f64=False)
            bctx.hist_pairsum_dev(hp, P, nh, bin_, qd.data_ptr(), nq,
                                  S2.data_ptr() + 8 * int(b.type_first[s]) * nq)
            bctx.pdf_dev(hp, p.el_count, nh, bin_, "rPDF", rho[s], sl[e0:e0 + p.n_el],
                         pdf2.data_ptr() + 8 * int(b.pdf_off[s]), cpu_radius=True)
            e0 += p.n_el

    plain_f32()
    bctx.synchronize()
    t0 = time.perf_counter()
    plain_f32()
    bctx.synchronize()
    t_p32 = time.perf_counter() - t0
    torch.cuda.synchronize()
    moved32 = int((hist - hist3).abs().sum().item()) // 2
    ovf = int(b.overflow().sum())
    b.close()
    bctx.close()
    hist_ms = split_ms["hist"]
    t_fresh = t_plan + t_first
    return {
        "workload": f"{count} nanoparticles, N log-uniform 500-200,000 (mean {sum(p.n for p in parts) // count}), "
                    "1-4 elements, fcc/bcc sphere, cube, core-shell; bin 0.01 A; 2,048 q 0.2-25 A^-1; "
                    "histograms + pair sums + rPDF per particle",
        "pairs": pairs,
        "patterns_per_s_batched": round(count / t_fresh, 2),
        "patterns_per_s_plain": round(count / t_p, 2),
        "speedup": round(t_p / t_fresh, 3),
        "patterns_per_s_plain_f32": round(count / t_p32, 2),
        "speedup_vs_plain_f32": round(t_p32 / t_fresh, 3),
        "batched_note": "fresh ensemble: plan (tiled on the GPU from the device coordinates) + first "
                        "run; both forms start from coordinates already on the device",
        "ms_plan": round(t_plan * 1e3, 2),
        "ms_plan_host_tiled": round(t_plan_host * 1e3, 2),
        "ms_first_run": round(t_first * 1e3, 2),
        "patterns_per_s_batched_replay": round(count / t_b, 2),
        "speedup_replay": round(t_p / t_b, 3),
        "replay_note": "the same coordinates again: graph replays of the finished plan" if use_graph
                       else "the same coordinates again: eager runs of the finished plan",
        "plain_f32_note": "structure by structure with FP32 coordinates and the reference GPU binning "
                          "(IEEE float sqrt and division); the batched form bins in FP64 (the reference "
                          "CPU expression)",
        "mfma_pairsums": {
            "patterns_per_s_batched": round(count / (t_plan + t_first_m), 2),
            "speedup": round(t_p / (t_plan + t_first_m), 3),
            "patterns_per_s_batched_replay": round(count / t_bm, 2),
            "speedup_replay": round(t_p / t_bm, 3),
            "ms_first_run": round(t_first_m * 1e3, 2),
            "ms_batched": round(t_bm * 1e3, 2),
            "pairsum_ms": pairsum_mfma_ms,
            "S_max_rel_diff_vs_fp64_pairsums": sm_rel,
            "I_max_rel_diff_vs_fp64_pairsums": im_rel,
            "I_l2_rel_diff_vs_fp64_pairsums": im_l2,
            "note": "the same plan and runs with XS_PAIRSUM_MFMA: pair sums of all pair types as "
                    "int8 matrix-core products (types grouped by count slices and bin count); "
                    "histograms and PDFs unchanged"},
        "ms_batched": round(t_b * 1e3, 2),
        "ms_plain": round(t_p * 1e3, 2),
        "gpairs_per_s_batched_replay": round(pairs / t_b / 1e9, 2),
        "graph_ms": round(g_ms[0] / max(g_ms[1], 1), 3) if use_graph else None,
        "kernels_ms_eager_run": split_ms,
        "hist_roofline": {"bound": "valu", "achieved": round(OPS_PER_PAIR * pairs / (hist_ms * 1e-3) / 1e12, 3),
                          "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(OPS_PER_PAIR * pairs / (hist_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)},
        "checks": {"hist_equal_plain": hist_eq, "pdf_equal_plain": pdf_eq,
                   "pairsum_max_rel_diff_vs_plain": s_rel, "overflow": ovf,
                   "f32_binning_moved_pairs": moved32},
        "setup_s": {"generate": round(t_gen, 2)},
    }


def pairsum_matrix_leg(ctx, steps):
    """Histogram -> pair sums for many histograms on one bin grid: 256
    two-element particles (768 histograms, 16,384 bins of 0.004 A), 8,192 q
    over 0.2-25 A^-1.  The FP64 vector kernel against the int8 matrix-core
    product (XS_PAIRSUM_MFMA); errors against an FP64 product on the CPU."""
    import torch
    import xansons_hip as xh
    nh, bin_, nq = 16384, 0.004, 8192
    parts = st.c4_particles()
    b = xh.Batch(ctx, [(p.xyz64(), p.el_count) for p in parts], bin_, nhist=[nh] * len(parts))
    dev = torch.device("cuda", 0)
    hist = torch.empty(b.hist_len, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    b.run_dev(hist.data_ptr())
    ctx.synchronize()
    ovf = int(b.overflow().sum())
    ncols = b.ntypes_total
    b.close()
    q = st.q_mesh(0.2, 25.0, nq)
    qd = torch.from_numpy(q).to(dev)
    Sv = torch.empty(ncols * nq, dtype=torch.float64, device=dev)
    Sm = torch.empty_like(Sv)
    torch.cuda.synchronize()

    def timed(fn, names):
        fn()
        ctx.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        t = (time.perf_counter() - t0) / steps
        k = {n: ctx.profile_read(n)[0] / steps for n in names}
        ctx.profile(False)
        return t, k

    tv, kv = timed(lambda: ctx.hist_pairsum_dev(hist.data_ptr(), ncols, nh, bin_, qd.data_ptr(), nq,
                                                Sv.data_ptr()), ["hist_pairsum"])
    tm, km = timed(lambda: ctx.hist_pairsum_dev(hist.data_ptr(), ncols, nh, bin_, qd.data_ptr(), nq,
                                                Sm.data_ptr(), mfma=True),
                   ["pairsum_slices", "pairsum_mfma"])
    Sf = torch.empty_like(Sv)
    tf, kf = timed(lambda: ctx.hist_pairsum_dev(hist.data_ptr(), ncols, nh, bin_, qd.data_ptr(), nq,
                                                Sf.data_ptr(), f32=True), ["hist_pairsum"])
    H = hist.view(ncols, nh).cpu().numpy().view(np.uint64)
    Svh = Sv.view(ncols, nq).cpu().numpy()
    Smh = Sm.view(ncols, nq).cpu().numpy()
    Sfh = Sf.view(ncols, nq).cpu().numpy()
    # FP64 reference on the first 16 particles (48 columns)
    cols = 48
    r = (np.arange(nh) + 0.5) * bin_
    qr = np.outer(q, r)
    A = np.sin(qr) / (qr + 1e-8)
    del qr
    Se = (A @ H[:cols].astype(np.float64).T).T
    del A

    def s_err(S):
        return float((np.abs(S[:cols] - Se).max(axis=1) / np.abs(Se).max(axis=1)).max())

    def i_errs(S):
        """Worst pointwise and worst L2-relative error of I(q) over the particles."""
        worst = {"max_rel": 0.0, "l2_rel": 0.0}
        for p_i in range(cols // 3):
            ec = parts[p_i].el_count
            sl = (3.63, 9.577)
            e = rel_errors(intensity_two_elements(S[3 * p_i:3 * p_i + 3], ec, sl),
                           intensity_two_elements(Se[3 * p_i:3 * p_i + 3], ec, sl))
            worst = {k: max(worst[k], e[k]) for k in worst}
        return worst

    def i_err(S):
        return i_errs(S)["max_rel"]

    def i_l2(S):
        return i_errs(S)["l2_rel"]

    cmax = int(H.max())
    nb = 1
    while cmax >> (7 * nb):
        nb += 1
    flop = 2.0 * nq * nh * ncols
    mm_ms = km["pairsum_mfma"]
    int8_ops = 5 * nb * flop
    return {
        "workload": f"256 two-element particles (fcc / core-shell spheres, 531-7,803 atoms), "
                    f"{ncols} histograms x {nh} bins (0.004 A), {nq} q 0.2-25 A^-1",
        "columns": ncols, "nhist": nh, "nq": nq, "max_count": cmax, "count_slices": nb,
        "overflow": ovf,
        "ms_fp64_vector": round(tv * 1e3, 3), "ms_mfma": round(tm * 1e3, 3),
        "ms_plain_fp32": round(tf * 1e3, 3),
        "kernel_ms": {"hist_pairsum": round(kv["hist_pairsum"], 3),
                      "pairsum_slices": round(km["pairsum_slices"], 3),
                      "pairsum_mfma": round(mm_ms, 3),
                      "hist_pairsum_f32": round(kf["hist_pairsum"], 3)},
        "speedup": round(tv / tm, 2),
        "speedup_vs_plain_fp32": round(tf / tm, 2),
        "plain_fp32_note": "FP32 terms with IEEE sinf and FP32 sums over 2048-bin chunks, one thread "
                           "per q (the arithmetic of the reference GPU kernels)",
        "fp64_equiv_tflops_mfma": round(flop / tm / 1e12, 2),
        "roofline_mfma": {"bound": "mfma", "achieved": round(int8_ops / (mm_ms * 1e-3) / 1e12, 1),
                          "peak": 5000.0, "unit": "TOP/s (int8)",
                          "frac": round(int8_ops / (mm_ms * 1e-3) / 1e12 / 5000.0, 4),
                          "int8_products": 5 * nb},
        "max_rel_err_S_vs_fp64": {"fp64_vector": s_err(Svh), "mfma": s_err(Smh),
                                  "plain_fp32": s_err(Sfh)},
        "max_rel_err_I_vs_fp64": {"fp64_vector": i_err(Svh), "mfma": i_err(Smh),
                                  "plain_fp32": i_err(Sfh)},
        "l2_rel_err_I_vs_fp64": {"fp64_vector": i_l2(Svh), "mfma": i_l2(Smh),
                                 "plain_fp32": i_l2(Sfh)},
        "error_sample": f"first {cols // 3} particles ({cols} histograms), all {nq} q; "
                        "S: max |dS| / max |S| per histogram; I: pointwise where I >= 1e-3 max I, "
                        "and ||dI|| / ||I|| (worst particle)",
    }


def amp2d_leg(ctx, nrot=256, nrot_list=16):
    """2D pattern throughput: the 50,000-atom three-element disordered solid,
    a 512 x 512 Cartesian q plane (|q_x|, |q_y| <= 8 A^-1, q_z = 0) and
    all 256 seeded uniform rotations; complex-exponential terms per second of
    the grid kernel (xs_amp2d_grid) over the whole workload, with the list
    kernel (xs_amp2d, same points as an explicit list) timed beside it on
    the first `nrot_list` rotations.  Checks:
    Friedel symmetry I(q) = I(-q) over the plane, grid vs list, and a sample
    of 64 q points x 2 rotations against an FP64 numpy sum."""
    import torch
    ens = st.disordered_cubic()
    x64 = ens.xyz64()
    n = 512
    qv = st.q_plane(8.0, n)
    nqp = qv.shape[0]
    # q_plane's point k = i n + j has q = (g[i], g[j], 0): rows along q_y
    step = 2.0 * 8.0 / (n - 1)
    q0, du, dv = np.array([qv[0, 0], qv[0, 1], 0.0]), np.array([0.0, step, 0.0]), \
        np.array([step, 0.0, 0.0])
    rot = st.uniform_rotations(256, seed=8)[:nrot].reshape(nrot, 9)
    ff = st.gaussian_ff(np.linalg.norm(qv, axis=1), st.gaussian_ff_coeffs(3, seed=9))
    dev = torch.device("cuda", 0)
    xd = torch.from_numpy(x64).to(dev)
    qd = torch.from_numpy(qv).to(dev)
    rd = torch.from_numpy(np.ascontiguousarray(rot)).to(dev)
    fd = torch.from_numpy(np.ascontiguousarray(ff)).to(dev)
    Id = torch.empty(nqp, dtype=torch.float64, device=dev)
    Il = torch.empty(nqp, dtype=torch.float64, device=dev)
    Ig = torch.empty(nqp, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def grid():
        ctx.amp2d_grid_dev(xd.data_ptr(), ens.n, ens.el_count, q0, du, dv, n, n, rd.data_ptr(),
                           nrot, fd.data_ptr(), Id.data_ptr())

    def listed():
        ctx.amp2d_dev(xd.data_ptr(), ens.n, ens.el_count, qd.data_ptr(), nqp, rd.data_ptr(), nrot_list,
                      fd.data_ptr(), Il.data_ptr())

    def grid_first():
        ctx.amp2d_grid_dev(xd.data_ptr(), ens.n, ens.el_count, q0, du, dv, n, n, rd.data_ptr(),
                           nrot_list, fd.data_ptr(), Ig.data_ptr())

    def timed(fn):
        fn()
        ctx.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        fn()
        ctx.synchronize()
        t = time.perf_counter() - t0
        k = ctx.profile_read("amp2d")[0]
        ctx.profile(False)
        return t, k

    t, kms = timed(grid)
    _, kms_list = timed(listed)
    grid_first()
    ctx.synchronize()
    I = Id.cpu().numpy()
    fried = float(np.max(np.abs(I - I[::-1]) / np.maximum(np.abs(I), 1e-300)))
    IL = Il.cpu().numpy()
    IG = Ig.cpu().numpy()
    vs_list = float(np.max(np.abs(IG - IL)) / np.abs(IL).max())
    # FP64 numpy evaluation of 64 q points x 2 rotations
    idx = np.random.Generator(np.random.PCG64(1)).choice(nqp, 64, replace=False)
    Ie = np.zeros(64)
    for r in range(2):
        qr = qv[idx] @ rot[r].reshape(3, 3)          # (R^T q) . x = q . (R x)
        amp = np.zeros(64, dtype=np.complex128)
        off = 0
        for e, c in enumerate(ens.el_count):
            ph = qr @ x64[:, off:off + c]
            amp += ff[e, idx] * np.exp(1j * ph).sum(axis=1)
            off += c
        Ie += np.abs(amp) ** 2
    ctx.amp2d_grid_dev(xd.data_ptr(), ens.n, ens.el_count, q0, du, dv, n, n, rd.data_ptr(), 2,
                       fd.data_ptr(), Id.data_ptr())
    ctx.synchronize()
    err = float(np.max(np.abs(Id.cpu().numpy()[idx] - Ie) / Ie))
    err_l2 = rel_errors(Id.cpu().numpy()[idx], Ie)["l2_rel"]
    terms = ens.n * nqp * nrot
    terms_list = ens.n * nqp * nrot_list
    achieved = AMP2D_OPS_PER_TERM * terms / (kms * 1e-3) / 1e12
    return {
        "workload": f"{ens.n}-atom 3-element disordered solid, 512 x 512 q plane (|q| components <= 8), "
                    f"all {nrot} seeded uniform rotations",
        "terms": terms, "ms": round(t * 1e3, 2), "kernel_ms": round(kms, 2),
        "gterms_per_s": round(terms / (kms * 1e-3) / 1e9, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "ops_per_term": AMP2D_OPS_PER_TERM, "kernel": "amp2d_grid_kernel"},
        "list_kernel_rotations": nrot_list,
        "list_kernel_ms": round(kms_list, 2),
        "list_gterms_per_s": round(terms_list / (kms_list * 1e-3) / 1e9, 1),
        "checks": {"friedel_max_rel": fried, "grid_vs_list_max_rel": vs_list,
                   "fp64_sample_max_rel": err, "fp64_sample_l2_rel": err_l2,
                   "fp64_sample": "64 q points x 2 rotations, FP64 numpy sum",
                   "grid_vs_list": f"first {nrot_list} rotations, all q points"},
    }


def debye_leg(ctx):
    """Direct Debye pair sums (every pair, every q, FP64 terms): the 1,985-atom
    fcc sphere on the reference q mesh (1,024 q over 0.5-10 A^-1; 2.0e9
    pair-q), the user manual's NaCl 10x10x10 example (8,000 atoms, two
    elements, 1,024 q from 0.5 A^-1 in steps of 6.5/1,024; 3.3e10 pair-q)
    and the 50,000-atom three-element disordered solid on 1,024 q over
    0.5-10 A^-1 (1.28e12 pair-q).  Gpair-q/s from the kernel's HIP events;
    the small case is checked against an FP64 numpy sum over all pairs."""
    import torch
    dev = torch.device("cuda", 0)
    out = {}
    for name, ens, q in (("fcc_1985", st.fcc_sphere(), st.q_mesh(0.5, 10.0, 1024)),
                         ("nacl_10x10x10_8000", st.rocksalt_cube(ncell=10, vacancy=0.0, sigma=0.0),
                          st.q_mesh(0.5, 7.0, 1024)),
                         ("disordered_50000", st.disordered_cubic(), st.q_mesh(0.5, 10.0, 1024))):
        x64 = ens.xyz64()
        xd = torch.from_numpy(x64).to(dev)
        qd = torch.from_numpy(q).to(dev)
        P = ens.n_el * (ens.n_el + 1) // 2
        S = torch.empty(P * q.size, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.debye_pairsum_dev(xd.data_ptr(), ens.n, ens.el_count, qd.data_ptr(), q.size, S.data_ptr())
        ctx.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        ctx.debye_pairsum_dev(xd.data_ptr(), ens.n, ens.el_count, qd.data_ptr(), q.size, S.data_ptr())
        ctx.synchronize()
        t = time.perf_counter() - t0
        kms = ctx.profile_read("debye")[0]
        ctx.profile(False)
        pairs = ens.n * (ens.n - 1) // 2
        achieved = DEBYE_OPS_PER_PAIR_Q * pairs * q.size / (kms * 1e-3) / 1e12
        rec = {"atoms": ens.n, "nq": int(q.size), "pair_q": pairs * int(q.size), "ms": round(t * 1e3, 3),
               "kernel_ms": round(kms, 3), "gpair_q_per_s": round(pairs * q.size / (kms * 1e-3) / 1e9, 1),
               "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                            "ops_per_pair_q": DEBYE_OPS_PER_PAIR_Q, "kernel": "debye_kernel<true>"}}
        if ens.n <= 2000:
            d = np.sqrt(((x64[:, :, None] - x64[:, None, :]) ** 2).sum(axis=0))
            iu = np.triu_indices(ens.n, 1)
            r = d[iu]
            Se = np.array([np.sum(np.sin(qq * r) / (qq * r + 1e-8)) for qq in q])
            Sg = S.cpu().numpy()
            rec["max_rel_err_vs_fp64"] = float(np.max(np.abs(Sg - Se)) / np.max(np.abs(Se)))
            rec["l2_rel_err_vs_fp64"] = rel_errors(Sg, Se)["l2_rel"]
        out[name] = rec
    return out


def progress(msg):
    """One line per finished leg on stderr (stdout carries only the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def atoms_workloads(n_cells=150, n_single=60000):
    """Ensemble-builder workloads: (name, blocks, n_el, files for the reference).

    rocksalt_sphere: rock-salt cell given as a 2-atom basis, three fcc
    centring translations and an inversion (10 images per cell, 8 after the
    0.5 A merge), n_cells^3 cells, cut to a sphere around the origin.
    single_cell: a finished disordered structure of n_single atoms plus 5 %
    near-duplicates given as one cell, the O(N^2) merge of the reference."""
    from xansons_hip.atoms import Block
    sl = "Na 3.63\nCl 9.577\n"
    calc = ('  <Calculation name="b" source="neutron" scenario="Debye" FFfilename="sl.txt">\n'
            '    <q min="0.5" max="5" N="4"/>\n  </Calculation>\n')
    a, half = 5.64, n_cells // 2
    rcut = half * a
    I = np.eye(3)
    symm_f = [((0.5, 0.5, 0.0), I), ((0.5, 0.0, 0.5), I), ((0.0, 0.5, 0.5), I), ((0.0, 0.0, 0.0), -I)]
    basis = np.array([[-half, -half, -half], [-half + 0.5, -half, -half]], dtype=np.float64) * a
    rs = Block(r_atom=basis, el=[0, 1],
               p_symm=np.array([I] + [P for _, P in symm_f]),
               r_symm=np.array([[0.0, 0.0, 0.0]] + [np.array(t) * a for t, _ in symm_f]),
               cell=np.eye(3) * a, ncell=(n_cells,) * 3, rcut=rcut)
    v = lambda x: " ".join(repr(float(c)) for c in x)  # noqa: E731
    xml_rs = ('<?xml version="1.0"?>\n<Config>\n' + calc + '  <Block>\n    <Atoms>\n' +
              f'      <Atom name="Na" r="{v(basis[0])}"/>\n      <Atom name="Cl" r="{v(basis[1])}"/>\n'
              '    </Atoms>\n' +
              "".join(f'    <SymmEqPos r="{v(t)}" R1="{v(P[0])}" R2="{v(P[1])}" R3="{v(P[2])}"/>\n'
                      for t, P in symm_f) +
              f'    <CutOff Rcut="{rcut!r}"/>\n'
              f'    <CellVectors a="{a!r} 0 0" b="0 {a!r} 0" c="0 0 {a!r}" N="{n_cells} {n_cells} {n_cells}"/>\n'
              '  </Block>\n</Config>\n')
    rng = np.random.default_rng(60000)
    nd = n_single // 20
    r = rng.uniform(0, (n_single / 0.06) ** (1 / 3), (n_single, 3))
    r = np.concatenate([r, r[rng.choice(n_single, nd, replace=False)] +
                        rng.normal(size=(nd, 3)) * 0.12])
    el = rng.integers(0, 2, n_single + nd)
    names = np.array(["Na", "Cl"])
    atoms_txt = "".join(f"{names[e]} {float(x)!r} {float(y)!r} {float(z)!r}\n"
                        for e, (x, y, z) in zip(el, r))
    order = np.argsort(el, kind="stable")   # block::sortAtoms groups by element, stably
    sc = Block(r_atom=r[order], el=el[order])
    xml_sc = ('<?xml version="1.0"?>\n<Config>\n' + calc +
              '  <Block>\n    <Atoms filename="atoms.txt"/>\n  </Block>\n</Config>\n')
    return [("rocksalt_sphere", [rs], 2, {"start.xml": xml_rs, "sl.txt": sl}),
            ("single_cell", [sc], 2, {"start.xml": xml_sc, "sl.txt": sl, "atoms.txt": atoms_txt})]


def reference_atoms_seconds(files):
    """The reference's CalcAndPrintAtoms (block::calcAtoms, one thread) on the
    same input through oracle/_ref/ref_atoms; (seconds, el_count) or None."""
    exe = os.path.join(REPO, "oracle", "_ref", "ref_atoms")
    if not os.path.exists(exe):
        return None
    tmp = tempfile.mkdtemp(prefix="xs_atoms_ref_")
    try:
        for fn, text in files.items():
            with open(os.path.join(tmp, fn), "w") as f:
                f.write(text)
        subprocess.run([exe, "start.xml", "out.json", "--no-atoms"], cwd=tmp, check=True,
                       stdout=subprocess.DEVNULL, timeout=600)
        with open(os.path.join(tmp, "out.json")) as f:
            d = json.load(f)
        return d["seconds"], [int(c) for c in d["el_count"]]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def atoms_leg(ctx, cpu=True, reps=3):
    """GPU ensemble builder (xs_atoms_build, host blocks in, element-major
    atoms on the device out) vs the reference's builder on the same input."""
    from xansons_hip.atoms import Atoms
    out = {}
    for name, blocks, n_el, files in atoms_workloads():
        Atoms(ctx, blocks, n_el, seed=1).close()  # warm-up (module load, allocations)
        ts = []
        for rep in range(reps):
            ctx.profile(True)
            t0 = time.perf_counter()
            a = Atoms(ctx, blocks, n_el, seed=1 + rep)
            ts.append(time.perf_counter() - t0)
            phases = {k: round(ctx.profile_read(k)[0], 3)
                      for k in ("atoms_merge", "atoms_place", "atoms_order")}
            ctx.profile(False)
            counts = [int(c) for c in a.el_count]
            a.close()
        t = float(np.median(ts))
        cand = sum(b.candidates() for b in blocks)
        n = sum(counts)
        rec = {"candidates": cand, "atoms": n, "ms": round(t * 1e3, 3), "phases_ms": phases,
               "matoms_per_s": round(n / t / 1e6, 1)}
        if cpu:
            ref = reference_atoms_seconds(files)
            if ref is not None:
                rec["reference_s"] = round(ref[0], 3)
                rec["reference_cores"] = 1
                rec["speedup"] = round(ref[0] / t, 1)
                rec["el_count_equal_reference"] = ref[1] == counts
        out[name] = rec
    return out


def rearr_workload(m=32, nq=1024):
    """RearrangementInt workload: one block given as a single cell of m^3
    atoms on a simple cubic lattice (2.67 A) moved by up to 0.4 A (so no two
    atoms merge), 3 elements element-major, each atom with its neighbours
    within 2.7 A moved by up to 0.1 A, Rcorr 0.3, neutron, nq q on
    0.5-10 A^-1.  Returns the RearrBlock, q, SL and the files for the
    reference."""
    from xansons_hip.rearrangement import RearrBlock
    rng = np.random.default_rng(2027)
    names = ["Na", "Cl", "Au"]
    n = m ** 3
    grid = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
    r = grid * 2.67 + rng.uniform(-0.4, 0.4, (n, 3))
    el = np.sort(rng.integers(0, 3, n))
    # neighbours within 2.7 A through a cell list
    cell = 2.7
    key = np.floor((r - r.min(0)) / cell).astype(np.int64)
    g = int(key.max()) + 1
    lin = (key[:, 0] * g + key[:, 1]) * g + key[:, 2]
    order = np.argsort(lin, kind="stable")
    starts = np.searchsorted(lin[order], np.arange(g ** 3 + 1))
    nbs = []
    for i in range(n):
        cand = []
        kx, ky, kz = key[i]
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    x, y, z = kx + dx, ky + dy, kz + dz
                    if 0 <= x < g and 0 <= y < g and 0 <= z < g:
                        c = (x * g + y) * g + z
                        cand.append(order[starts[c]:starts[c + 1]])
        c = np.concatenate(cand)
        d = np.linalg.norm(r[c] - r[i], axis=1)
        j = c[(d < 2.7) & (c != i)]
        nbs.append(np.column_stack([el[j], r[j] + rng.uniform(-0.1, 0.1, (len(j), 3))]))
    q = np.array([0.5 + k * (10.0 - 0.5) / nq for k in range(nq)])
    sl = np.array([3.63, 9.577, 7.63])
    lines = [f"{names[e]} {float(x)!r} {float(y)!r} {float(z)!r}" for e, (x, y, z) in zip(el, r)]
    nbl = []
    for i, nb in enumerate(nbs):
        nbl.append(f"{i} {len(nb)}")
        nbl.extend(f"{names[int(e)]} {float(x)!r} {float(y)!r} {float(z)!r}" for e, x, y, z in nb)
    xml = ('<?xml version="1.0"?>\n<Config>\n  <Calculation name="rr" source="neutron" '
           'scenario="Debye" FFfilename="sl.txt">\n'
           f'    <q min="0.5" max="10" N="{nq}"/>\n  </Calculation>\n  <Block>\n'
           '    <Atoms filename="atoms.txt"/>\n'
           '    <Rearrangement filename="nb.txt" Rcorr="0.3"/>\n  </Block>\n</Config>\n')
    files = {"start.xml": xml, "sl.txt": "Na 3.63\nCl 9.577\nAu 7.63\n",
             "atoms.txt": "\n".join(lines) + "\n", "nb.txt": "\n".join(nbl) + "\n"}
    blk = RearrBlock(r_atom=r, el=el.astype(np.uint32), neighbours=np.concatenate(nbs), rcorr=0.3,
                     coeff=1, nb_count=np.array([len(x) for x in nbs]))
    return blk, q, sl, files


def rearr_leg(ctx, cpu=True, reps=3):
    """RearrangementInt with GPU pair sums vs the reference's own function
    (oracle/_ref/ref_atoms --rearrangement, one thread) on the same input."""
    from xansons_hip.rearrangement import rearrangement_intensity
    blk, q, sl, files = rearr_workload()
    n = blk.r_atom.shape[0]
    I0 = 1.0 + np.arange(q.size) / 64.0
    args = (I0, q, [blk], 3, n, "neutron")
    rearrangement_intensity(*args, sl=sl, ctx=ctx)  # warm-up
    ts = []
    for _ in range(reps):
        ctx.profile(True)
        t0 = time.perf_counter()
        out = rearrangement_intensity(*args, sl=sl, ctx=ctx)
        ts.append(time.perf_counter() - t0)
        k = {name: round(ctx.profile_read(name)[0], 3) for name in ("cutoff_pairs", "pairlist_sums")}
        ctx.profile(False)
    t = float(np.median(ts))
    nnb = int(blk.nb_count.sum())
    rec = {"atoms": n, "nq": int(q.size), "neighbours": nnb, "reference_pair_tests": n * n,
           "ms": round(t * 1e3, 3), "kernels_ms": k}
    if cpu:
        exe = os.path.join(REPO, "oracle", "_ref", "ref_atoms")
        if os.path.exists(exe):
            tmp = tempfile.mkdtemp(prefix="xs_rearr_ref_")
            try:
                for fn, text in files.items():
                    with open(os.path.join(tmp, fn), "w") as f:
                        f.write(text)
                subprocess.run([exe, "start.xml", "out.json", "--no-atoms", "--rearrangement"],
                               cwd=tmp, check=True, stdout=subprocess.DEVNULL, timeout=900)
                with open(os.path.join(tmp, "out.json")) as f:
                    d = json.load(f)
            finally:
                shutil.rmtree(tmp, ignore_errors=True)
            ref = np.array([float.fromhex(v) for v in d["rearr_I"]])
            rec["reference_s"] = round(d["rearr_seconds"], 3)
            rec["reference_cores"] = 1
            rec["speedup"] = round(d["rearr_seconds"] / t, 1)
            rec["max_rel_err_vs_reference"] = float(np.abs(out - ref).max() / np.abs(ref).max())
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble", type=int, default=1000,
                    help="particles in the batched-ensemble leg (0 skips it)")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D pattern leg")
    ap.add_argument("--no-debye", action="store_true", help="skip the direct Debye leg")
    ap.add_argument("--no-matrix", action="store_true",
                    help="skip the many-histogram pair-sum leg (FP64 vector vs int8 matrix cores)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--no-atoms", action="store_true", help="skip the ensemble-builder leg")
    ap.add_argument("--no-rearr", action="store_true", help="skip the RearrangementInt leg")
    ap.add_argument("--only", choices=["matrix", "2d", "ensemble", "debye", "atoms", "rearr"],
                    help="run one extra leg alone and print its object (for profiling)")
    ap.add_argument("--no-graph", action="store_true",
                    help="ensemble leg: replay the plan with eager launches instead of a hipGraph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported: "
              "this backend has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.only:
        import torch
        import xansons_hip as xh
        torch.cuda.init()          # the legs allocate through torch
        ctx = xh.Context(0)
        leg = {"matrix": lambda: pairsum_matrix_leg(ctx, max(1, args.steps)),
               "2d": lambda: amp2d_leg(ctx),
               "debye": lambda: debye_leg(ctx),
               "atoms": lambda: atoms_leg(ctx, cpu=not args.no_cpu_baseline),
               "rearr": lambda: rearr_leg(ctx, cpu=not args.no_cpu_baseline),
               "ensemble": lambda: ensemble_leg(ctx, args.ensemble, max(1, args.steps // 2), 1,
                                                use_graph=not args.no_graph)}
        print(json.dumps({args.only: leg[args.only]()}))
        ctx.close()
        return

    import torch
    import xansons_hip as xh

    ens = st.rocksalt_cube()
    # bin count as GetHistSize (CalcFunctions.cpp:224-243): bounding-box diagonal / bin + 1
    nh = int((np.linalg.norm(ens.xyz64().max(axis=1) - ens.xyz64().min(axis=1)) / BIN)) + 1
    na, nb = ens.el_count
    pairs = ens.n * (ens.n - 1) // 2
    q = st.q_mesh(0.2, 25.0, NQ)

    ctx = xh.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    # headline: FP64 binning, the reference CPU path's expression (bit-exact counts)
    step, hist, ovf, S = gpu_step_setup(ctx, ens, nh, q, f64=True)
    ms, kern = time_steps(ctx, step, args.steps, args.warmup)
    h = hist.view(3, nh).cpu().numpy()
    totals_ok = [int(v) for v in h.sum(axis=1)] == [na * (na - 1) // 2, na * nb, nb * (nb - 1) // 2]
    # I(q) against its FP64 evaluation from the same histograms, every 16th q
    qs = slice(0, NQ, 16)
    Sg = S.view(3, NQ).cpu().numpy()[:, qs]
    Se = fp64_pairsums(h, BIN, q[qs])
    sl = (3.63, 9.577)
    iq_err = rel_errors(intensity_two_elements(Sg, ens.el_count, sl),
                        intensity_two_elements(Se, ens.el_count, sl))
    s_err = [rel_errors(Sg[p], Se[p], floor=0.0) for p in range(3)]
    del step
    # the same histogram from host buffers (the C ABI copies the coordinates
    # in and the counts out): one warm-up call, then two timed
    x_host = np.ascontiguousarray(ens.xyz64())
    ctx.hist(x_host, ens.el_count, BIN, nh, f64=True)
    t0 = time.perf_counter()
    for _ in range(2):
        h_host, _ = ctx.hist(x_host, ens.el_count, BIN, nh, f64=True)
    host_ms = (time.perf_counter() - t0) * 1e3 / 2
    host_equal = bool(np.array_equal(h_host, h))
    del x_host, h_host
    step32, hist32, _, S32 = gpu_step_setup(ctx, ens, nh, q, f64=False)
    ms32, kern32 = time_steps(ctx, step32, max(1, args.steps // 2), 1)
    iq32_err = rel_errors(intensity_two_elements(S32.view(3, NQ).cpu().numpy()[:, qs], ens.el_count, sl),
                          intensity_two_elements(Se, ens.el_count, sl))
    del step32, hist32, S32

    hist_ms = kern["hist"][0] / max(kern["hist"][1], 1)
    hist32_ms = kern32["hist"][0] / max(kern32["hist"][1], 1)
    achieved = OPS_PER_PAIR * pairs / (hist_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(pairs / (ms * 1e-3) / 1e9, 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded rock-salt cube)",
        "config": {"workload": "rock-salt cube 980,219 atoms, 2 elements, 3 pair-type histograms "
                               f"bin 0.005 A ({nh:,} bins), I(q) 4,096 q 0.2-25 A^-1, rPDF",
                   "atoms": ens.n, "pairs": pairs, "nhist": nh, "nq": NQ,
                   "binning": "FP64 (reference CPU expression, counts bit-exact)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": HIST_TRAFFIC_BYTES, "kernel": "hist_kernel_f64<32768, true>",
                     "ops_per_pair": OPS_PER_PAIR, "kernel_ms": round(hist_ms, 3),
                     "valu_issue_frac": round(achieved / VALU_ISSUE_TOPS, 4),
                     "valu_issue_cycles_per_64_pairs": SLOTS_PER_PAIR,
                     "valu_issue_cycle_frac": round(
                         SLOTS_PER_PAIR * pairs / 64 / (CU_ISSUE_PER_S * hist_ms * 1e-3), 4),
                     "traffic_note": "HBM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE, copy-calibrated); "
                                     "J records read per work item through the scalar cache, their "
                                     "L2 misses served from MALL/HBM, and histogram window flushes "
                                     "(device-scope atomics, 64 B each): ~26 GB at ~120 GB/s, 1.5% "
                                     "of HBM bandwidth"},
        "kernels_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in kern.items()},
        "host_buffers": {"ms": round(host_ms, 3), "gpairs_per_s": round(pairs / (host_ms * 1e-3) / 1e9, 3),
                         "counts_equal": host_equal,
                         "note": "xs_hist with host pointers (FP64 binning), PCIe copies of the "
                                 "23.5 MB of coordinates in and the 2.3 MB of counts out included; "
                                 "histogram only, not the pair sums and PDF of a step"},
        "f32_binning": {"ms_per_step": round(ms32, 3),
                        "gpairs_per_s": round(pairs / (ms32 * 1e-3) / 1e9, 3),
                        "hist_kernel_ms": round(hist32_ms, 3),
                        "kernel": "hist_kernel_f32<32768, true>",
                        "roofline": {"bound": "valu",
                                     "achieved": round(OPS_PER_PAIR * pairs / (hist32_ms * 1e-3) / 1e12, 3),
                                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                     "frac": round(OPS_PER_PAIR * pairs / (hist32_ms * 1e-3) / 1e12
                                                   / FP32_PEAK_TFLOPS, 4),
                                     "ops_per_pair": OPS_PER_PAIR},
                        "valu_active_per_64_pairs": F32_ACTIVE_PER_64_PAIRS,
                        "valu_wave_instr_per_cu_cycle": round(
                            F32_ACTIVE_PER_64_PAIRS * pairs / 64 / (CU_ISSUE_PER_S * hist32_ms * 1e-3), 4),
                        "valu_issue_peak_wave_instr_per_cu_cycle": 2,
                        "I_vs_fp64": {k: float(f"{v:.3e}") for k, v in iq32_err.items()}},
        "checks": {"pair_totals_exact": totals_ok, "overflow": int(ovf.item()),
                   "I_vs_fp64": {k: float(f"{v:.3e}") for k, v in iq_err.items()},
                   "S_vs_fp64_per_type": [{k: float(f"{v:.3e}") for k, v in e.items()} for e in s_err],
                   "error_sample": f"{NQ // 16} q (every 16th); FP64 numpy pair sums of the same "
                                   "histograms; max_rel: pointwise where |ref| >= 1e-3 max |ref| "
                                   "(all points for S), l2_rel: ||got - ref|| / ||ref||"},
    }
    progress(f"headline done: {line['value']} Gpairs/s")
    if not args.no_matrix:
        line["pairsum_matrix"] = pairsum_matrix_leg(ctx, max(1, args.steps))
        progress("pair-sum matrix leg done")
    if not args.no_2d:
        line["pattern_2d"] = amp2d_leg(ctx)
        progress("2D leg done")
    if not args.no_debye:
        line["debye"] = debye_leg(ctx)
        progress("Debye leg done")
    if args.ensemble > 0:
        del hist, ovf
        torch.cuda.empty_cache()
        line["ensemble"] = ensemble_leg(ctx, args.ensemble, max(1, args.steps // 2), 1,
                                        use_graph=not args.no_graph)
        progress("ensemble leg done")
    if not args.no_atoms:
        line["atoms_builder"] = atoms_leg(ctx, cpu=not args.no_cpu_baseline)
        progress("ensemble-builder leg done")
    if not args.no_rearr:
        line["rearrangement"] = rearr_leg(ctx, cpu=not args.no_cpu_baseline)
        progress("RearrangementInt leg done")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
        progress("CPU baseline done")
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Poses scored per second of the FFT translational-rotational search on one
MI355X, on the single-ligand throughput workload: 64^3 x 9 channel FP32
potential grids (3000-atom pseudo-protein, 0.5 A), one 40-atom ligand with 100
torsion-perturbed conformers x 2000 quaternions = 200,000 poses.

A step is one pass of the search over all 200,000 poses with the inputs
already resident in HBM (tables uploaded once before the timed region):
prep -> spread + forward FFT passes -> conj-multiply / channel sum -> inverse
FFT passes -> windowed argmin, replayed as one hipGraph, plus the copy of the
per-pose results (energy, translation index, minimum corner) to the host.

Beside it: the reference-dataflow engine run as the reference drives it (one
launch sequence per batch of 50 rotations of one conformer, full-grid stages)
on a conformer subset, per-kernel device time from HIP events on the launch
stream (for the roofline), and the literal cpuNaive dense correlation timed on
the host cores over a bounded sample of the same poses.

The serving leg streams docking requests (15-80 atoms, 1-50 conformers,
500/2000/5000 quaternions each, against resident 64^3 and 96^3 x 9 grids) into
the continuous-batching server: first all at once to find its capacity, then
as a Poisson stream at a fixed fraction of that capacity; it reports p50/p99
submit-to-result latency and the served poses/s.

--workload batched runs the uneven-batching workload instead (64 ligands of
15-60 atoms, 10-100 conformers, 1000 quaternions, one hipGraph-captured launch
set); --workload large the 128^3 x 9 grids at 0.375 A (50-atom ligand, 50
conformers x 3000 quaternions, top-10 poses per conformer, 1% of the rotations
re-scored in FP64 by the oracle).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "poses scored/sec at 64³×9ch FP32; p50/p99 request latency in serve loop"
PEAK_HBM_GBPS = 8000.0
PEAK_MFMA_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix cores
PEAK_MFMA_FP16_TFLOPS = 2500.0  # MI355X dense FP16 matrix cores
PEAK_VALU_FP32_TFLOPS = 78.6  # MI355X FP32 vector ALUs, unpacked: 256 CUs x 64 lanes x 2 flops x 2.4 GHz


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def fused_kernel_bytes(geoms, npch, N, mid_launches, columns=False, atoms=False):
    """Algorithmic HBM bytes of each fused kernel over a set of poses.

    geoms: [(box, win, nch)] per job, box/win [..., 3] per pose, nch the
    job's channel count.  plane = N*(N/2+1)*8 bytes (one complex
    half-spectrum plane), row = (N/2+1)*8 bytes (one z half-spectrum row).
    Plane form:
      forward planes: writes nch*box_x planes per pose (the pruned y/z spectra)
      mid lines:      reads those planes, writes win_x planes (x-inverse done),
                      and reads the protein spectra (npch*H*8) once per launch
      inverse+argmin: reads win_x planes per pose, writes one 8-byte key
    Atom form: no forward stage; the mid stage reads the atoms (nothing
      compared with the rest), writes win_x*win_y rows and reads the protein
      spectra once per launch; inverse z as the column form.
    Column form:
      forward z:      writes nch*box_x*box_y rows per pose (z spectra)
      mid kz:         reads those rows, writes win_x*win_y rows (x and y
                      inverses done), reads the protein spectra once per launch
      inverse z:      reads win_x*win_y rows per pose, writes one 8-byte key
    """
    plane = N * (N // 2 + 1) * 8
    row = (N // 2 + 1) * 8
    H8 = N * N * (N // 2 + 1) * 8
    fwd = mid = inv = 0.0
    for box, win, nch, *_ in geoms:
        b = box.reshape(-1, 3).astype(np.float64)
        w = win.reshape(-1, 3).astype(np.float64)
        valid = (w > 0).all(axis=1)
        b, w = b[valid], w[valid]
        if atoms:
            lig = np.zeros(len(b))
            out = w[:, 0] * w[:, 1] * row
        elif columns:
            lig = nch * b[:, 0] * b[:, 1] * row
            out = w[:, 0] * w[:, 1] * row
        else:
            lig = nch * b[:, 0] * plane
            out = w[:, 0] * plane
        fwd += float(lig.sum())
        mid += float((lig + out).sum())
        inv += float((out + 8).sum())
    mid += npch * H8 * mid_launches
    if atoms:
        return {"prep": 0.0, "mid_atoms": mid, "inv_z": inv}
    if columns:
        return {"prep": 0.0, "fwd_z": fwd, "mid_kz": mid, "inv_z": inv}
    return {"prep": 0.0, "fwd_planes": fwd, "mid_lines": mid, "inv_argmin": inv}


def mid_flops(geoms, N, columns):
    """FFT-convention FP32 flops of the mid kernel over a set of poses: a
    complex line transform of n points counts 5 n log2 n, a complex
    multiply-accumulate 8, a complex scale 2.
      column form, per pose and kz plane: y lines x < box_x and all N x
        lines per channel, N*N products per channel, then N inverse x lines,
        win_x inverse y lines and N*N scales;
      plane form, per pose: N*(N/2+1) x lines and their products per channel,
        then the inverse x lines and scales."""
    F = 5.0 * N * np.log2(N)
    ZH = N // 2 + 1
    total = 0.0
    for box, win, nch, *_ in geoms:
        b = box.reshape(-1, 3).astype(np.float64)
        w = win.reshape(-1, 3).astype(np.float64)
        valid = (w > 0).all(axis=1)
        b, w = b[valid], w[valid]
        if columns:
            per = ZH * (nch * ((b[:, 0] + N) * F + 8 * N * N) + (N + w[:, 0]) * F + 2 * N * N)
        else:
            per = np.full(len(b), N * ZH * (nch * (F + 8 * N) + F + 2 * N))
        total += float(per.sum())
    return total


def atom_mfma_flops(geoms, N):
    """Flops of the atom form's matrix-core products over a set of poses: per
    pose, kz plane and channel atom, a (N x 2) by (2 x 2N) real product
    (N kx rows; real and imaginary k slots; 2N columns for the real and
    imaginary parts of the N ky values), i.e. 8 N^2 flops.  nent (the fourth
    item of a geometry) counts the ligand's channel atoms: every atom is in
    one vdW channel and in the elec channel."""
    ZH = N // 2 + 1
    total = 0.0
    for box, win, nch, nent in geoms:
        w = win.reshape(-1, 3)
        nvalid = int((w > 0).all(axis=1).sum())
        total += float(nvalid) * ZH * nent * 8.0 * N * N
    return total


def reference_bytes(N, C):
    """B(N, C) = 8CG + 24CH + 16H + 8G: each stage boundary of the reference
    dataflow written and read once."""
    G = N ** 3
    H = N * N * (N // 2 + 1)
    return 8 * C * G + 24 * C * H + 16 * H + 8 * G


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_single_ligand_64.md")


def profiled_traffic(kernel, poses_per_launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the single-ligand workload (per-pose
    figures, copy-calibrated), scaled to this run's poses per launch; None if
    the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for ln in f:
                if ln.startswith("traffic-per-pose:"):
                    kv = dict(t.split("=") for t in ln.split(":", 1)[1].split())
                    if kv.get("kernel") == kernel:
                        return (float(kv["fetch"]) + float(kv["write"])) * poses_per_launch
    except OSError:
        pass
    return None


def profiled_issue(kernel, products):
    """The SIMD-cycle fractions (vector-instruction issue, matrix-core busy)
    of `kernel` with the atom form's product form `products` from the
    committed rocprofv3 SQ-counter summary of the single-ligand workload;
    None if absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for ln in f:
                if ln.startswith("issue-per-simd-cycle:"):
                    kv = dict(t.split("=") for t in ln.split(":", 1)[1].split())
                    if kv.get("kernel") == kernel and kv.get("products") == str(products):
                        return {"valu_issue": float(kv["valu_issue"]), "mfma_busy": float(kv["mfma_busy"]),
                                "source": "profiles/traffic_single_ligand_64.md (rocprofv3 SQ counters)"}
    except OSError:
        pass
    return None


def percentile(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


def serve_leg(ctx, cdk, workloads, nreq, load, max_batch, seed):
    """Continuous-batching server over resident 64^3 and 96^3 grids.

    Pass 1 submits every request at once; the makespan gives the server's
    capacity in requests/s.  Pass 2 submits the same requests as a Poisson
    stream at `load` x that capacity; latency is submit-to-result as the
    server measures it."""
    prots = []
    for name in ("p64", "p96"):
        g = workloads.protein_grids(name)
        prots.append(cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"]))
    rs = np.random.default_rng(seed)
    reqs = [workloads.serving_request(rs, seed + 2 * k) for k in range(nreq)]
    poses = np.array([r["quats"].shape[0] * r["quats"].shape[1] for r in reqs], np.float64)

    # The plain per-ligand GPU loop (the reference dataflow, batches of 50
    # rotations) on each resident protein, timed on the 40-atom single-ligand
    # ligand (its cost is the transforms', nearly independent of atom count):
    # the time it would take for the same requests, each on its own grid.
    L = workloads.config2(nconf=1, nquat=1000)["ligands"][0]
    plain, fused = [], []
    for p in prots:
        lig = cdk.Ligand(p, L["charge"], L["epsilon"], L["radius"])
        pb = cdk.Batch(ctx, p)
        pb.add(lig, L["conformers"][:1], L["quats"][:1])
        for engine, out in ((cdk.ENGINE_REFERENCE, plain), (cdk.ENGINE_FUSED, fused)):
            pb.run(engine)
            best = 1e30
            for _ in range(2):
                ts = time.perf_counter()
                pb.run(engine)
                best = min(best, time.perf_counter() - ts)
            out.append(L["quats"].shape[1] / best)
        pb.close()
        lig.close()
    plain_s = float(sum(n / plain[r["protein_index"]] for n, r in zip(poses, reqs)))
    # The same requests through the plain loop at its HBM roofline: the
    # reference dataflow's algorithmic bytes B(N, C) per pose (each stage
    # boundary written and read once) at 8 TB/s, C the request's channels.
    edges = [p.n for p in prots]
    chans = []
    for r in reqs:
        lg = cdk.Ligand(prots[r["protein_index"]], r["charge"], r["epsilon"], r["radius"])
        chans.append(lg.channels()[0] + 1)
        lg.close()
    roof_s = float(sum(n * reference_bytes(edges[r["protein_index"]], c) / (PEAK_HBM_GBPS * 1e9)
                       for n, r, c in zip(poses, reqs, chans)))
    # each request alone on the fused engine (its poses at that protein's rate)
    alone = np.array([n / fused[r["protein_index"]] for n, r in zip(poses, reqs)])
    budget = (8 << 30) + (8 << 30)  # context workspace + resident request state

    def run(arrivals):
        srv = cdk.Server(ctx, prots, budget, max_batch)
        t0 = time.perf_counter()
        sub = []
        ids = []
        for r, ta in zip(reqs, arrivals):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            sub.append(time.perf_counter() - t0)
            ids.append(srv.submit(r["protein_index"], r["charge"], r["epsilon"], r["radius"], r["conformers"],
                                  r["quats"]))
        lat = np.array([srv.result(i)[3] for i in ids])
        st = srv.stats()
        srv.close()
        span = float(np.max(np.asarray(sub) + lat))
        return lat, span, st

    log("serving: plain loop %s poses/s, fused alone %s poses/s; saturated pass" % (
        [round(v) for v in plain], [round(v) for v in fused]))
    lat0, span0, st0 = run(np.zeros(nreq))
    log("serving: saturated pass %.1f s for %d poses" % (span0, poses.sum()))
    cap_req = nreq / span0
    rate = load * cap_req
    arrivals = np.cumsum(rs.exponential(1.0 / rate, nreq))
    arrivals -= arrivals[0]
    lat, span, st = run(arrivals)
    log("serving: Poisson pass %.1f s" % span)
    for p in prots:
        p.close()
    out = {
        "requests": nreq,
        "poses": int(poses.sum()),
        "capacity_poses_per_s": round(poses.sum() / span0, 1),
        "capacity_req_s": round(cap_req, 3),
        "offered_rate_req_s": round(rate, 3),
        "load": load,
        "p50_s": round(percentile(lat, 50), 4),
        "p99_s": round(percentile(lat, 99), 4),
        "mean_s": round(float(lat.mean()), 4),
        "poses_per_s": round(poses.sum() / span, 1),
        "saturated_p50_s": round(percentile(lat0, 50), 4),
        "launch_sets": int(st["launch_sets"]),
        "mean_launch_set_poses": round(st["mean_batch_poses"], 1),
        "peak_admitted_bytes": int(st["peak_admitted_bytes"]),
        "max_batch_poses": max_batch,
        "proteins": "64^3 x 9 and 96^3 x 9 FP32 at 0.5 A, spectra resident",
        "slowdown_p50": round(percentile(lat / alone, 50), 2),
        "slowdown_p99": round(percentile(lat / alone, 99), 2),
        "slowdown_note": "latency / the request's own GPU time alone (its poses at the fused engine's rate on its "
                         "protein); request sizes span 500 to 250,000 poses on two grid sizes, so p99/p50 of "
                         "latency mostly reflects request size",
        "fused_alone_poses_per_s": {"64": round(fused[0], 1), "96": round(fused[1], 1)},
        "plain_loop_poses_per_s": {"64": round(plain[0], 1), "96": round(plain[1], 1)},
        "plain_loop_s": round(plain_s, 3),
        "capacity_vs_plain_loop": round(plain_s / span0, 2),
        "plain_loop_roofline_s": round(roof_s, 3),
        "capacity_vs_roofline_plain_loop": round(roof_s / span0, 2),
    }
    return out


def fused_steps(cdk, batch, steps, warmup, npose):
    """W untimed and K timed passes of the whole fused search (hipGraph
    replay + result copy-back), each bracketed by device synchronisation
    inside Batch.run."""
    for i in range(warmup):
        batch.run(cdk.ENGINE_FUSED)
        log("warmup %d done" % i)
    times = []
    for i in range(steps):
        ts = time.perf_counter()
        batch.run(cdk.ENGINE_FUSED)
        times.append(time.perf_counter() - ts)
        log("step %d: %.3f s (%.0f poses/s)" % (i, times[-1], npose / times[-1]))
    return float(sum(times))


def kernel_leg(ctx, cdk, batch, geoms, npch, N):
    """Per-kernel device time from HIP events on the launch stream (one eager
    pass), algorithmic GB/s per kernel, and the roofline of the dominant one."""
    ctx.set_option("time_stages", 1)
    ctx.stage_times(reset=True)
    batch.run(cdk.ENGINE_FUSED)
    ms, launches = ctx.stage_times(reset=True)
    ctx.set_option("time_stages", 0)
    group = ctx.get_option("last_mid_group")
    pipe = ctx.get_option("last_pipeline")
    columns, atoms = pipe == 1, pipe == 2
    if atoms:
        names = ["prep", None, "mid_atoms", "inv_z"]
    elif columns:
        names = ["prep", "fwd_z", "mid_kz", "inv_z"]
    else:
        names = ["prep", "fwd_planes", "mid_lines", "inv_argmin"]
    kb = fused_kernel_bytes(geoms, npch, N, int(launches[2]), columns, atoms)
    kernels = {}
    for i, nm in enumerate(names):
        if nm and launches[i]:
            kernels[nm] = {"ms_total": round(float(ms[i]), 3), "launches": int(launches[i]),
                           "avg_ms": round(float(ms[i] / launches[i]), 4),
                           "GBps": round(kb[nm] / (ms[i] * 1e-3) / 1e9, 1) if ms[i] > 0 else None}
    dom = max([nm for nm in names if nm], key=lambda k: kernels.get(k, {}).get("ms_total", 0))
    dom_avg_s = kernels[dom]["ms_total"] * 1e-3 / kernels[dom]["launches"]
    dom_bytes = kb[dom] / kernels[dom]["launches"]
    achieved = dom_bytes / dom_avg_s / 1e9
    log("kernels", json.dumps(kernels))
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None, "avg_launch_ms": round(dom_avg_s * 1e3, 4),
            "algorithmic_bytes_per_launch": round(dom_bytes)}
    if dom == "mid_atoms":
        # The ligand spectra are matrix-core products: their algorithmic flops
        # against the peak of the matrix cores that run them.  Split-FP16 form:
        # FP16 units, three FP16 products per FP32-accurate product (so a
        # third of the FP16 peak is that form's ceiling); FP32 form: FP32 units.
        tf = atom_mfma_flops(geoms, N) / kernels[dom]["launches"] / dom_avg_s / 1e12
        hbm = dict(roof)
        products = ctx.get_option("atom_products")
        f16 = products in (0, 2)
        nsplit = {0: 3, 2: 6}.get(products, 1)  # FP16 partial products per FP32-accurate product
        peak = PEAK_MFMA_FP16_TFLOPS if f16 else PEAK_MFMA_FP32_TFLOPS
        roof = {"bound": "mfma", "kernel": dom, "achieved": round(tf, 2), "peak": peak,
                "unit": "TFLOP/s", "frac": round(tf / peak, 4), "traffic": None,
                "products": {0: "split-f16 x2 (hi/lo parts, 3 FP16 partial products, FP32 accumulate, ~2^-21)",
                             2: "split-f16 x3 (exact hi/mid/lo parts, 6 FP16 partial products, FP32 accumulate, "
                                "~2^-32: FP32-accurate)",
                             1: "f32"}[products],
                "avg_launch_ms": hbm["avg_launch_ms"], "algorithmic_bytes_per_launch": hbm["algorithmic_bytes_per_launch"],
                "hbm": {"achieved": hbm["achieved"], "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": hbm["frac"]}}
        if f16:
            roof["frac_of_split_peak"] = round(nsplit * tf / peak, 4)
    if dom in ("mid_kz", "mid_lines"):
        # The transforms run on the vector ALUs in FP32 without packed
        # instructions: 78.6 TFLOP/s is that ceiling (half the listed
        # 157.3 TFLOP/s FP32 matrix-core peak).
        tf = mid_flops(geoms, N, columns) / kernels[dom]["launches"] / dom_avg_s / 1e12
        roof["valu"] = {"achieved": round(tf, 2), "peak": PEAK_VALU_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tf / PEAK_VALU_FP32_TFLOPS, 4)}
    if columns or atoms:
        kernels["tuned"] = {"pipeline": "atoms" if atoms else "columns", "mid_group": group,
                            "max_dyn_lds": ctx.get_option("max_dyn_lds")}
    else:
        kernels["tuned"] = {"pipeline": "planes",
                            "fwd_variant": ["block", "register"][ctx.get_option("last_fwd_variant")],
                            "inv_variant": ["block", "register"][ctx.get_option("last_inv_variant")],
                            "mid_group": group, "mid_depth": ctx.get_option("last_mid_depth"),
                            "mid_poses": ctx.get_option("last_mid_poses"),
                            "max_dyn_lds": ctx.get_option("max_dyn_lds")}
    return kernels, roof


def reference_leg(ctx, cdk, prot, lig, confs, quats, fused_idx):
    """The reference dataflow, one launch sequence per batch of 50 rotations
    of one conformer (the plain per-ligand GPU loop)."""
    pb = cdk.Batch(ctx, prot)
    pb.add(lig, confs, quats)
    pb.run(cdk.ENGINE_REFERENCE)
    ctx.set_option("time_stages", 1)
    ctx.stage_times(reset=True)
    ts = time.perf_counter()
    pb.run(cdk.ENGINE_REFERENCE)
    plain_s = time.perf_counter() - ts
    rms, rl = ctx.stage_times(reset=True)
    ctx.set_option("time_stages", 0)
    ts = time.perf_counter()
    pb.run(cdk.ENGINE_REFERENCE)
    plain_s = min(plain_s, time.perf_counter() - ts)
    n = confs.shape[0] * quats.shape[-2]
    Er, ir, _ = pb.results(0)
    pb.close()
    agree = float((ir == fused_idx).all(axis=-1).mean())
    ref_names = ["prep", "spread", "r2c", "conj_mult", "sum_grids", "c2r", "argmin"]
    stage = {nm: round(float(rms[i]), 2) for i, nm in enumerate(ref_names) if rl[i]}
    value = n / plain_s
    log("reference engine %.0f poses/s, stage ms %s" % (value, stage))
    return {"value": round(value, 1), "unit": "poses/s", "poses": n, "stage_ms": stage,
            "argmin_agreement": round(agree, 4)}


def host_cpu():
    """The host the CPU baseline ran on: logical CPUs, CPUs this process may
    use, and the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def cpu_threads_default():
    # every host core this process may use, at most the GPU box's CPU share
    # of 16 cores per GPU
    return max(1, min(16, host_cpu()["usable"]))


def cpu_leg(g, L, nsample, threads, N, budget_s=20.0):
    """The literal cpuNaive dense FP32 correlation (cpuNaive.cpp:184-222) on
    the host cores over the first rotations of conformer 0 (a bounded sample
    of the same poses: at most nsample rotations, stopping once budget_s has
    been spent), poses/s extrapolated from that sample."""
    from oracle import oracle as orc
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    ts = time.perf_counter()
    ns = 0
    for r in range(min(nsample, L["quats"].shape[1])):
        Lb, bx, wn, _ = prob.pose_box(L["conformers"][0], L["quats"][0, r])
        if (wn > 0).all() and (bx <= N).all():
            orc.correlate_window_full_timed(prob.Pused, Lb, wn, threads)
        ns += 1
        if time.perf_counter() - ts > budget_s:
            break
    cpu_s = time.perf_counter() - ts
    cpu = {"value": round(ns / cpu_s, 4), "unit": "poses/s", "cores": int(threads), "kind": "port",
           "host": host_cpu(), "build": orc.CPUBASE_FLAGS,
           "sample": "extrapolated from the first %d rotations of conformer 0 (%.1f s): the literal cpuNaive dense "
                     "FP32 loop (cpuNaive.cpp:194-216) over the ligand box and window, %d threads over x"
                     % (ns, cpu_s, threads)}
    log("cpu baseline", cpu)
    return cpu


def cpu_callers(cdk, workloads, protein, g, L, E, idx, mc, nsample=2000, ntrial=2000):
    """Beside the device figures of the callers object: the oracle's FP64
    restatements on one host thread, for the grid energy of explicit search
    poses (the rescoring's energy: trilinear grid values at the atoms,
    orc_sparse_energy_f64, over the first valid poses of conformer 0) and
    for random-torsion conformer trials of the same 40-atom chain
    (orc_conformers), and for the protein's potential grids built from its
    atoms (orc_build_grids, brute force)."""
    from oracle import oracle as orc
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    qs = np.flatnonzero(np.isfinite(E[0]))[:nsample]
    rot = orc.rotate(L["quats"][0, qs], L["conformers"][0])
    ts = time.perf_counter()
    for k, q in enumerate(qs):
        prob.sparse_energy(rot[k], mc[0, q], idx[0, q])
    dt = time.perf_counter() - ts
    out = {"cores": 1, "kind": "port",
           "rescore_energy": {"poses": int(qs.size), "s": round(dt, 4), "poses_per_s": round(qs.size / dt, 1),
                              "note": "one oracle call per pose from Python"}}
    x0, _, _, _ = cdk.synth_ligand(4242, 40)
    rotors, mv, ex = workloads.chain_topology(40)
    ts = time.perf_counter()
    orc.conformers(x0.astype(np.float64), rotors, mv, ex, ntrial, 11, 1.2)
    dt = time.perf_counter() - ts
    out["conformers"] = {"trials": ntrial, "s": round(dt, 4), "trials_per_s": round(ntrial / dt, 1)}
    # the protein's potential grids from its atoms (orc_build_grids, brute force FP64)
    natom_p, box_p, n, h = workloads.PROTEINS[protein]
    xyz, q, e, r = cdk.synth_protein_atoms(workloads.PROTEIN_SEED[protein], natom_p, box_p)
    gmin = np.asarray(g["mid"], np.float64) - np.asarray(g["length"], np.float64) / 2
    ts = time.perf_counter()
    orc.build_grids(xyz, q, e, r, g["radii"], n, gmin, h)
    dt = time.perf_counter() - ts
    out["grid_builder"] = {"atoms": natom_p, "grids": "%d^3 x %d" % (n, len(g["radii"]) + 1), "s": round(dt, 3)}
    return out


def arithmetic_note(roof):
    prod = str((roof or {}).get("products", ""))
    if prod.startswith("split-f16 x3"):
        return ("FP32 grids, transforms, protein products and sums; ligand-spectrum products on the FP16 matrix "
                "cores with every FP32 operand split exactly into three FP16 parts after power-of-two scaling "
                "(6 FP16 partial products per FP32 product, exact in FP32, FP32 accumulate: about 2^-32 per "
                "product, below FP32 rounding)")
    if prod.startswith("split-f16 x2"):
        return ("FP32 grids, transforms, protein products and sums; ligand-spectrum products on the FP16 matrix "
                "cores with hi/lo operand splits (3 FP16 partial products per product, FP32 accumulate, about "
                "2^-21 per product: narrower than FP32)")
    return "FP32"


def line(args, metric, value, total, workload, extra_config, roof, kernels, cpu, **more):
    out = {
        "metric": metric,
        "value": round(value, 1),
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "arithmetic": arithmetic_note(roof),
        "data": "synthetic (seeded pseudo-protein grids, random-walk ligands, torsion conformers, Shoemake quaternions)",
        "config": dict({"workload": workload}, **extra_config),
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    out.update(more)
    print(json.dumps(out), flush=True)


def callers_leg(ctx, cdk, workloads, prot, lig, L, E, idx, mc, nlow=10):
    """The search's callers on either side, each timed end to end through
    the C-ABI with host arrays in and out (staging, device buffers and PCIe
    included), best of a few calls: retention of the top-10 poses per
    conformer with their rebuilt coordinates (NoMini:466-487), FP64 grid
    rescoring of those poses with and without the gradient, rigid-body
    refinement of them (50 steps), the grid handoff to the host minimiser
    (every channel, x fastest, AddGridForcesToOpenMMSystem.cpp:30-50) and
    random-torsion conformer generation (GeneRandomConformations.cpp:14-111,
    steric filter)."""
    def best(f, reps=3):
        b, r = 1e30, None
        for _ in range(reps):
            t = time.perf_counter()
            r = f()
            b = min(b, time.perf_counter() - t)
        return b, r

    out = {}
    nconf = E.shape[0]
    t = time.perf_counter()
    kept = []
    for c in range(nconf):
        rot = ctx.Rotate(L["quats"][c], L["conformers"][c])
        _, pose = cdk.rank_poses(prot, rot, E[c], idx[c], mc[c], nlow)
        kept.append(pose)
    dt = time.perf_counter() - t
    poses = np.concatenate(kept)
    n = poses.shape[0]
    out["retention"] = {"conformers": nconf, "poses_kept": n, "s": round(dt, 4),
                        "conformers_per_s": round(nconf / dt, 1)}
    dt, (e, _, ins) = best(lambda: cdk.rescore_poses(prot, lig, poses))
    out["rescore"] = {"poses": n, "atoms": lig.natom, "s": round(dt, 5), "poses_per_s": round(n / dt, 1),
                      "on_grid": int((ins == lig.natom).sum())}
    dt, _ = best(lambda: cdk.rescore_poses(prot, lig, poses, want_grad=True))
    out["rescore_with_gradient"] = {"poses": n, "s": round(dt, 5), "poses_per_s": round(n / dt, 1)}
    dt, (_, e2) = best(lambda: cdk.refine_poses(prot, lig, poses, steps=50), reps=2)
    out["refine"] = {"poses": n, "steps": 50, "s": round(dt, 5), "poses_per_s": round(n / dt, 1),
                     "mean_energy_drop": round(float(np.mean(e - e2)), 4)}
    nch = prot.num_grids
    dt, _ = best(lambda: [cdk.grid_for_openmm(prot, ch) for ch in range(nch)])
    nbytes = nch * prot.n ** 3 * 8
    out["grid_for_openmm"] = {"channels": nch, "bytes_out": nbytes, "s": round(dt, 5),
                              "GBps_to_host": round(nbytes / dt / 1e9, 2)}
    x0, _, _, _ = cdk.synth_ligand(4242, 40)
    rot, mv, ex = workloads.chain_topology(40)
    ntrial = 20000
    dt, (confs, _) = best(lambda: cdk.generate_conformers(ctx, x0, rot, mv, ex, ntrial, 11, 1.2, max_keep=1000),
                          reps=2)
    out["conformers"] = {"atoms": 40, "rotors": int(rot.shape[0]), "trials": ntrial, "kept": int(confs.shape[0]),
                         "s": round(dt, 5), "trials_per_s": round(ntrial / dt, 1)}
    return out


def run_single(args, cdk, workloads):
    t0 = time.time()
    g = workloads.protein_grids(args.protein)
    w = workloads.config2(nconf=args.conformers, nquat=args.quaternions)
    L = w["ligands"][0]
    log("inputs ready in %.1fs" % (time.time() - t0))
    ctx = cdk.Context(0, use_graph=int(not args.no_graph), fused_chunk=args.chunk, mid_group=args.group, fwd_variant=args.fwd,
                      inv_variant=args.inv, pipeline=args.pipeline,
                      atom_products=args.atom_products, overlap_z=args.overlap_z)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    nu, used, _ = lig.channels()
    C = nu + 1
    N = g["n"]
    npose = args.conformers * args.quaternions

    batch = cdk.Batch(ctx, prot)
    batch.add(lig, L["conformers"], L["quats"])
    total = fused_steps(cdk, batch, args.steps, args.warmup, npose)
    value = npose * args.steps / total
    E, idx, mc = batch.results(0)
    box, win = batch.geometry(0)
    n_valid = int(np.isfinite(E).sum())
    kernels, roof = kernel_leg(ctx, cdk, batch, [(box, win, C, 2 * lig.natom)], prot.num_grids, N)
    tr = profiled_traffic(roof["kernel"], n_valid / kernels[roof["kernel"]]["launches"]) if args.protein == "p64" else None
    if tr is not None:
        roof["traffic"] = round(tr)
        roof["traffic_source"] = "profiles/traffic_single_ligand_64.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE per pose)"
    iss = (profiled_issue(roof["kernel"], ctx.get_option("atom_products"))
           if args.protein == "p64" and roof["kernel"] == "mid_atoms" else None)
    if iss is not None:
        # what binds the kernel: vector-instruction issue plus matrix-core
        # busy cycles, which barely overlap
        roof["issue"] = iss

    pc = max(1, min(args.plain_conformers, args.conformers))
    ref = reference_leg(ctx, cdk, prot, lig, L["conformers"][:pc], L["quats"][:pc], idx[:pc])
    ref["speedup_fused"] = round(value / ref["value"], 2)
    # the plain loop at its HBM roofline: B(N, C) algorithmic bytes per pose at 8 TB/s
    ref["roofline_poses_per_s"] = round(PEAK_HBM_GBPS * 1e9 / reference_bytes(N, C), 1)
    ref["speedup_fused_vs_roofline"] = round(value / ref["roofline_poses_per_s"], 2)
    f64 = fp64_check(args, g, L, E, idx, mc, args.fp64_sample, args.fp64_full) if args.fp64_sample > 0 else None
    cpu = cpu_leg(g, L, args.cpu_sample, args.cpu_threads, N) if args.cpu_sample > 0 else None
    if cpu is not None and args.callers:
        try:
            cpu["callers"] = cpu_callers(cdk, workloads, args.protein, g, L, E, idx, mc)
        except Exception as ex:  # a baseline's failure must not lose the search's line
            cpu["callers"] = {"error": "%s: %s" % (type(ex).__name__, ex)}
        log("cpu callers", cpu["callers"])

    # The same protein's grids built on the device from its atoms.
    natom_p, box_p, _, _ = workloads.PROTEINS[args.protein]
    xyz_p, q_p, e_p, r_p = cdk.synth_protein_atoms(workloads.PROTEIN_SEED[args.protein], natom_p, box_p)
    ts = time.perf_counter()
    built = cdk.Protein.build(ctx, xyz_p, q_p, e_p, r_p, g["radii"], N, g["spacing"], g["mid"], g["length"])
    build_s = time.perf_counter() - ts
    diff = float(np.abs(built.grids().astype(np.float64) - g["values"]).max())
    built.close()
    grid_builder = {"atoms": natom_p, "grids": "%d^3 x 9" % N, "device_s_incl_spectra": round(build_s, 4),
                    "max_abs_diff_vs_host_recipe": diff}
    log("grid builder", grid_builder)

    callers = None
    if args.callers:
        try:
            callers = callers_leg(ctx, cdk, workloads, prot, lig, L, E, idx, mc)
        except Exception as ex:  # a caller's failure must not lose the search's line
            callers = {"error": "%s: %s" % (type(ex).__name__, ex)}
        log("callers", json.dumps(callers))

    serving = None
    if args.serve_requests > 0:
        serving = serve_leg(ctx, cdk, workloads, args.serve_requests, args.serve_load, args.serve_batch, 7100)
        log("serving", json.dumps(serving))

    metric = METRIC if args.protein == "p64" else "poses scored/sec at %d³×9ch FP32" % N
    line(args, metric, value, total,
         "single-ligand throughput: %d^3 x 9ch FP32 grids at %g A, 40-atom ligand, "
         "%d conformers x %d quaternions = %d poses" % (N, g["spacing"], args.conformers, args.quaternions, npose),
         {"grid_edge": N, "channels_used": C, "poses": npose, "valid_poses": n_valid,
          "engine": "fused" + ("" if args.no_graph else "+hipGraph"), "translations_per_s": round(value * N ** 3, 1)},
         roof, kernels, cpu,
         pipeline_equiv={"bytes_per_pose": reference_bytes(N, C), "GBps": round(value * reference_bytes(N, C) / 1e9, 1),
                         "note": "poses/s x B(N,C) of the unfused reference dataflow"},
         reference_engine=ref, fp64_check=f64, serving=serving, grid_builder=grid_builder,
         callers=callers)
    batch.close()
    lig.close()
    prot.close()
    ctx.close()


def run_batched(args, cdk, workloads):
    """64 uneven ligands in one hipGraph-captured launch set sequence."""
    t0 = time.time()
    g = workloads.protein_grids("p64")
    w = workloads.config3()
    log("inputs ready in %.1fs" % (time.time() - t0))
    ctx = cdk.Context(0, use_graph=int(not args.no_graph), fused_chunk=args.chunk, mid_group=args.group, fwd_variant=args.fwd,
                      inv_variant=args.inv, pipeline=args.pipeline,
                      atom_products=args.atom_products, overlap_z=args.overlap_z)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    N = g["n"]
    ligs = [cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"]) for L in w["ligands"]]
    batch = cdk.Batch(ctx, prot)
    for lig, L in zip(ligs, w["ligands"]):
        batch.add(lig, L["conformers"], L["quats"])
    npose = batch.total_poses()
    total = fused_steps(cdk, batch, args.steps, args.warmup, npose)
    value = npose * args.steps / total
    geoms = []
    n_valid = 0
    for j, lig in enumerate(ligs):
        box, win = batch.geometry(j)
        geoms.append((box, win, lig.channels()[0] + 1, 2 * lig.natom))
        n_valid += int(np.isfinite(batch.results(j)[0]).sum())
    kernels, roof = kernel_leg(ctx, cdk, batch, geoms, prot.num_grids, N)

    # The same ligands searched one at a time must give identical results.
    same = True
    for j in range(min(4, len(ligs))):
        L = w["ligands"][j]
        E1, i1, m1 = cdk.search(ctx, prot, ligs[j], L["conformers"], L["quats"])
        E, i, m = batch.results(j)
        same &= bool(np.array_equal(E.view(np.uint32), E1.view(np.uint32)) and np.array_equal(i, i1))
    L0 = w["ligands"][0]
    pc = max(1, min(args.plain_conformers, L0["conformers"].shape[0]))
    ref = reference_leg(ctx, cdk, prot, ligs[0], L0["conformers"][:pc], L0["quats"][:pc], batch.results(0)[1][:pc])
    ref["speedup_fused"] = round(value / ref["value"], 2)
    cpu = cpu_leg(g, L0, args.cpu_sample, args.cpu_threads, N) if args.cpu_sample > 0 else None
    atoms = [L["charge"].shape[0] for L in w["ligands"]]
    confs = [L["conformers"].shape[0] for L in w["ligands"]]
    line(args, "poses scored/sec at 64³×9ch FP32, 64 uneven ligands batched", value, total,
         "uneven batching: 64^3 x 9ch FP32 at 0.5 A, 64 ligands (atoms %d-%d, conformers %d-%d) x 1000 "
         "quaternions = %d poses in one hipGraph" % (min(atoms), max(atoms), min(confs), max(confs), npose),
         {"grid_edge": N, "ligands": len(ligs), "poses": npose, "valid_poses": n_valid,
          "engine": "fused" + ("" if args.no_graph else "+hipGraph")},
         roof, kernels, cpu, reference_engine=ref, per_ligand_identical_first4=same)
    batch.close()
    for lig in ligs:
        lig.close()
    prot.close()
    ctx.close()


def fp64_check(args, g, L, E, idx, mc, nsample, nfull):
    """Sampled rotations re-scored by the FP64 oracle at the GPU's translation
    (relative energy error), and a few searched in full in FP64 (argmin)."""
    from oracle import oracle as orc
    orc.set_threads(args.cpu_threads)
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    nconf, nquat = E.shape
    rs = np.random.default_rng(55)
    flat = rs.choice(nconf * nquat, min(nsample, nconf * nquat), replace=False)
    max_rel = 0.0
    checked = 0
    for f in flat:
        c, i = divmod(int(f), nquat)
        if not np.isfinite(E[c, i]):
            continue
        rot = orc.rotate(L["quats"][c, i:i + 1], L["conformers"][c])[0]
        e64 = prob.sparse_energy(rot, mc[c, i], idx[c, i])
        max_rel = max(max_rel, abs(float(E[c, i]) - e64) / max(abs(e64), 1e-30))
        checked += 1
    full = {"poses": 0, "exact": 0}
    for f in flat[:nfull]:
        c, i = divmod(int(f), nquat)
        o = prob.search_pose(L["conformers"][c], L["quats"][c, i], double=True)
        full["poses"] += 1
        full["exact"] += int(o["valid"] and np.array_equal(o["index"], idx[c, i]))
    log("fp64 check: %d poses, max rel err %.3g; full searches %s" % (checked, max_rel, full))
    return {"sampled_rotations": checked, "max_rel_energy_err": float("%.3g" % max_rel), "full_search_argmin": full}


def run_large(args, cdk, workloads):
    """128^3 x 9 at 0.375 A; top-10 poses per conformer; 1% FP64 re-score."""
    t0 = time.time()
    g = workloads.protein_grids("p128")
    w = workloads.config5()
    L = w["ligands"][0]
    log("inputs ready in %.1fs" % (time.time() - t0))
    ctx = cdk.Context(0, use_graph=int(not args.no_graph), fused_chunk=args.chunk, mid_group=args.group, fwd_variant=args.fwd,
                      inv_variant=args.inv, pipeline=args.pipeline,
                      atom_products=args.atom_products, overlap_z=args.overlap_z)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    C = lig.channels()[0] + 1
    N = g["n"]
    nconf, nquat = L["quats"].shape[:2]
    npose = nconf * nquat
    batch = cdk.Batch(ctx, prot)
    batch.add(lig, L["conformers"], L["quats"])
    total = fused_steps(cdk, batch, args.steps, args.warmup, npose)
    value = npose * args.steps / total
    E, idx, mc = batch.results(0)
    box, win = batch.geometry(0)
    kernels, roof = kernel_leg(ctx, cdk, batch, [(box, win, C, 2 * lig.natom)], prot.num_grids, N)

    # Top-10 poses per conformer (sort_index + pose rebuild, NoMini:466-487).
    ts = time.perf_counter()
    top = []
    for c in range(nconf):
        rot = ctx.Rotate(L["quats"][c], L["conformers"][c])
        rank, pose = cdk.rank_poses(prot, rot, E[c], idx[c], mc[c], 10)
        top.append(rank)
    top_s = time.perf_counter() - ts

    # 1% of the rotations re-scored by the FP64 oracle at the GPU's translation,
    # and a few searched in full (FP64 dense) for the argmin.
    f64 = fp64_check(args, g, L, E, idx, mc, max(1, npose // 100), args.large_full_checks)
    cpu = cpu_leg(g, L, args.cpu_sample, args.cpu_threads, N) if args.cpu_sample > 0 else None
    line(args, "poses scored/sec at 128³×9ch FP32", value, total,
         "large grid: 128^3 x 9ch FP32 at 0.375 A, 50-atom ligand, %d conformers x %d quaternions = %d poses, "
         "top-10 per conformer" % (nconf, nquat, npose),
         {"grid_edge": N, "channels_used": C, "poses": npose, "valid_poses": int(np.isfinite(E).sum()),
          "engine": "fused" + ("" if args.no_graph else "+hipGraph")},
         roof, kernels, cpu,
         fp64_check=f64,
         top10={"conformers": nconf, "host_s": round(top_s, 3)})
    batch.close()
    lig.close()
    prot.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["single", "batched", "large"], default="single")
    ap.add_argument("--protein", choices=["p64", "p96", "p128"], default="p64",
                    help="grids of the single-ligand workload (the headline is p64)")
    ap.add_argument("--conformers", type=int, default=100)
    ap.add_argument("--quaternions", type=int, default=2000)
    ap.add_argument("--plain-conformers", type=int, default=10)
    ap.add_argument("--cpu-sample", type=int, default=200,
                    help="at most this many rotations timed on the CPU, within 20 s (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0 = every usable host core, at most the box's share of 16)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--chunk", type=int, default=0, help="poses per fused launch set (0 = from the HBM budget)")
    ap.add_argument("--group", type=int, default=0, help="poses per mid-kernel workgroup (0 = autotuned)")
    ap.add_argument("--pipeline", type=int, default=-1,
                    help="fused engine form: 0 planes, 1 columns (64^3), 2 atoms (64^3, 96^3, 128^3), -1 autotuned")
    ap.add_argument("--overlap-z", type=int, default=1,
                    help="atom form: overlap each launch set's z-inverse stage with the next set (1) or not (0)")
    ap.add_argument("--atom-products", type=int, default=2,
                    help="atom form products: 2 exact three-part FP16 splits (FP32-accurate), 0 two-part FP16 "
                         "splits, 1 FP32 matrix cores")
    ap.add_argument("--fwd", type=int, default=-1, help="forward plane form: 0 block, 1 register, -1 autotuned")
    ap.add_argument("--inv", type=int, default=-1, help="inverse plane form: 0 block, 1 register, -1 autotuned")
    ap.add_argument("--serve-requests", type=int, default=40, help="requests in the serving leg (0 = skip)")
    ap.add_argument("--callers", type=int, default=1,
                    help="time the callers either side of the search (retention, rescoring, refinement, grid handoff, conformers)")
    ap.add_argument("--serve-load", type=float, default=0.7, help="Poisson rate as a fraction of capacity")
    ap.add_argument("--serve-batch", type=int, default=4096, help="poses per server launch set")
    ap.add_argument("--large-full-checks", type=int, default=6, help="full FP64 searches in the large workload")
    ap.add_argument("--fp64-sample", type=int, default=400,
                    help="rotations re-scored by the FP64 oracle in the single-ligand workload (0 = skip)")
    ap.add_argument("--fp64-full", type=int, default=4, help="full FP64 searches in the single-ligand workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (a single MI355X here): the search would shard by independent poses, "
              "each GPU running this single-GPU pipeline on its range with its own protein copy; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.cpu_threads <= 0:
        args.cpu_threads = cpu_threads_default()
    import cdocker_amd as cdk
    from cdocker_amd import workloads
    {"single": run_single, "batched": run_batched, "large": run_large}[args.workload](args, cdk, workloads)


if __name__ == "__main__":
    main()

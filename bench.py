#!/usr/bin/env python3
"""Simulated ECGs/s at 128^3, 600 ms, 12 leads -- batched vs plain, on one MI355X.

One step = one batch of 256 parameter vectors (4 free AP coefficients per
endo/mid/epi layer + 2 electrode displacements per lead) simulated on the
cached 128^3 ellipsoid heart: AP tables, the chunk kernel (lead field, chunk
sums, per-voxel APs contracted on the matrix cores), the activation-class
terms and the FP64 reduction, with the per-request inputs already in HBM.
Beside it: the same batch with activation classes off (every voxel's AP
evaluated, the path of hearts whose activation times are distinct), the plain
drop-in (one ekg_simulate call per simulation, host buffers in and out), a
short continuous-batching burst (p50/p99 latency), accuracy against the FP64
oracle and the compiled reference, and the reference's CPU path timed on the
host cores.

Prints one JSON line on stdout; details go to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ekgsim_amd import synth as S  # noqa: E402  (no GPU use)

METRIC = ("simulated ECGs/sec at 128^3, 600 ms, 12 leads (batched vs plain) at "
          "≤1e-5 rel-L2; p99")
# MI355X dense FP64 matrix-core peak (the lead contraction's precision), FLOP/s;
# also the FP64 vector and the non-packed FP32 vector rate
MFMA_FP64_PEAK = 78.6e12
# Sigmoid window of the transition tiles in the exp2 domain x = c1 (t - a),
# c1 = -k1 log2(e): outside (SATC_X, SKIP_X) a voxel's AP takes the FP64
# closed form (sigmoid within 2^-24.5 of 1, as in the FP32 AP) or is below
# 2^-36 of its plateau, so only window voxel-steps need an evaluated AP.
SATC_X, SKIP_X = -24.5, 48.0


def window_voxel_steps(act, k1, start, dt, T):
    """Voxel-steps (u, j) with SATC_X < c1 (t_j - a_u) < SKIP_X, t_j = start + j dt."""
    c1 = -k1 * math.log2(math.e)
    lo = act + SKIP_X / c1          # t > a - 48/|c1|
    hi = act + SATC_X / c1          # t < a + 24.5/|c1|
    jlo = np.clip(np.floor((lo - start) / dt) + 1, 0, T)
    jhi = np.clip(np.ceil((hi - start) / dt), 0, T)
    return int(np.maximum(jhi - jlo, 0).sum())


def lead_field_cells(layer):
    """Lead-weight evaluations per lead of one request in the chunk kernel:
    each myocardial voxel evaluates its own cell and its myocardial 3D4
    (corner) neighbours."""
    myo = layer > 0
    P = np.pad(myo, 1)
    n0, n1, n2 = myo.shape
    cnt = np.ones(myo.shape, np.int64)
    for d0 in (-1, 1):
        for d1 in (-1, 1):
            for d2 in (-1, 1):
                cnt += P[1 + d0:1 + d0 + n0, 1 + d1:1 + d1 + n1, 1 + d2:1 + d2 + n2]
    return int(cnt[myo].sum())


# Algorithmic work, counted per unit.  Class-window kernel (the dominant one
# when the heart has activation classes):
#   per (class-window entry, lead): 8 FP64 matrix-core flops (|e|^2 as a
#     4-term product, 4 FMAs) and 3 FP32 flops (rsq 1, |e|^-3 2), the lead
#     weight shared by the entry's classes;
#   per (class cell, lead): 8 FP64 matrix-core flops (an FMA for each of the
#     coefficients D0, D1, D2, K).
# Chunk kernel (dominant without classes):
#   per lead-weight evaluation (voxel, lead, cell): 6 FP64 flops (|e|^2 as
#     one FMA from the voxel's own |e|^2 and the dipole sum, the dipole
#     offset, the FMA into G) and 3 FP32 flops (rsq 1, r^3 2);
#   per window voxel-step (the transition quads): 2L FP64 MFMA flops and 11
#     FP32 ops (9 AP assembly + 2 transcendental).
# Both count at full weight against 78.6 TFLOP/s: the build disables packed
# FP32, so an FP32 vector operation issues at the FP64 vector rate, which on
# MI355X equals the FP64 matrix rate.
FP32_PER_CLASS_ENTRY, FP64_PER_CLASS_ENTRY, FP64_PER_CLASS_CELL = 3, 8, 8
FP32_PER_CELL, FP64_PER_CELL = 3, 6
FP32_PER_WINDOW_STEP = 11
# Serving bursts: warm-up bursts before the measured ones, and measured bursts
# (the median is reported).
SERVE_WARM, SERVE_MEASURED = 5, 3
# Naive per-voxel-step model: every voxel at every step, 38 VALU lane-cycles
# at L = 12, against 256 CU x 64 lanes x 2.4 GHz.
NAIVE_LANE_CYCLES, LANE_PEAK = 38, 256 * 64 * 2.4e9

# HBM bytes per launch of the dominant kernel on this workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by its reading for a
# 256 MiB calibration copy: with activation classes
# (profiles/ecg_class_window_kernel_128cubed_batch256_hbm_traffic.md) and
# without (profiles/ecg_chunk_kernel_128cubed_batch256_hbm_traffic.md).
TRAFFIC_BYTES = {"ecg_class_window_kernel": 436566943, "ecg_chunk_kernel": 1131874216}
# Issue-pipe occupancy of the dominant kernel on this workload, from one
# rocprofv3 --pmc pass (SQ_VALU_MFMA_BUSY_CYCLES, SQ_INSTS_VALU minus the MFMA
# instructions at 4 cycles each, SQ_VALU_MFMA_COEXEC_CYCLES, GRBM_GUI_ACTIVE)
# over 1024 SIMDs: profiles/ecg_class_window_kernel_128cubed_batch256_pipes.md
# and profiles/ecg_chunk_kernel_128cubed_batch256_pipes.md.  The matrix cores
# and the vector ALU never overlap (COEXEC 0), so their sum is the fraction of
# the kernel's cycles spent issuing.
PIPES = {"ecg_class_window_kernel": {"mfma_busy": 0.563, "valu_issue": 0.294, "coexec": 0.0},
         "ecg_chunk_kernel": {"mfma_busy": 0.116, "valu_issue": 0.806, "coexec": 0.0}}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def max_abs_err(ecg, refs):
    """Largest |GPU - FP64 oracle| over the checked ECGs, absolute and
    relative to the largest |oracle value| of the same ECG (None when no ECG
    was checked)."""
    if not refs:
        return None
    ab = max(float(np.max(np.abs(ecg[r] - ref))) for r, ref in refs.items())
    rel = max(float(np.max(np.abs(ecg[r] - ref)) / np.max(np.abs(ref))) for r, ref in refs.items())
    return {"abs": ab, "rel_to_peak": rel}


def workload(n, L, B, pseed0=1000):
    h = S.make_heart(n, seed=1)
    ks, ps = [], []
    base = S.lead_positions(n, L)
    for r in range(B):
        x = S.random_params(3, L, pseed0 + r)
        ks.append(S.layer_k_from_params(x, 3))
        ps.append((x, base))
    return h, ks, ps


# The reference's shipped heart (testRun/model_24: 93 x 124 x 124, 24 layers,
# FP64 conduction_24, two measuring points), stored as a test fixture.
TESTRUN_DIR = os.path.join(ROOT, "tests", "golden", "cli", "model24")
TESTRUN_BASE = [0, 2.5, 100, 0.90, 0.1, 0.0010, 0.10, 0.10, 250]
TESTRUN_KMIN = [0, 1.5, 100, 0.85, 0.05, 0.0003, 0.01, 0.01, 300]
TESTRUN_KMAX = [0, 3.5, 100, 0.95, 0.2, 0.0010, 0.10, 0.10, 450]


def testrun_workload(B):
    """(layer, seed, conduction, layer_k [B][24][9], lead positions [B][2][3])
    for B parameter vectors drawn within the ini bounds (endo-mid-epi
    k5..k8, then two displacements per lead in [-50, 50])."""
    from ekgsim_amd import _capi as A
    z = np.load(os.path.join(TESTRUN_DIR, "heart.npz"), allow_pickle=False)
    with open(os.path.join(TESTRUN_DIR, "conduction.matrix")) as f:
        f.readline()
        hdr = f.readline().split()
        cond = np.array(f.read().split(), float).reshape(int(hdr[3]), int(hdr[1]))
    pts = []
    with open(os.path.join(TESTRUN_DIR, "points.txt")) as f:
        for ln in f:
            if ln.strip():
                x, y, zz = (float(v) for v in ln.split(","))
                pts.append([zz, y, x])
    pts = np.array(pts)
    rng = np.random.default_rng(24)
    ks, ps = [], []
    for _ in range(B):
        genes = rng.uniform([3e-4, 0.01, 0.01, 300] * 3, [1e-3, 0.1, 0.1, 450] * 3)
        disp = rng.uniform(-50, 50, size=(len(pts), 2))
        k, _ = A.layer_aps(np.stack([TESTRUN_BASE] * 3), TESTRUN_KMIN, TESTRUN_KMAX, [5, 6, 7, 8], 2, 0.6, 24,
                           genes)
        ks.append(k)
        ps.append(S.move_leads(pts, disp))
    return z["layer"], z["seed"], cond, np.stack(ks), np.stack(ps)


def poisson_burst(srv, reqs, rate, seed, keep=()):
    """Submit reqs [(heart id, k, pos, start, length, dt)] at Poisson arrival
    times of the given rate and take every result in place
    (ekg_server_wait_view, no copy); returns (ECGs/s, latencies in ms, {i:
    copy of result i for i in keep}).  The descriptors are built before the
    clock starts and the C entry points are called directly, so the client
    costs a few microseconds per request instead of bounding the measured
    rate.  The timed region ends when the last result is in host memory; the
    kept results are copied and every ticket released after it."""
    import ctypes as C
    from ekgsim_amd import _capi as A
    L = A.lib()
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    descs = [A._Desc(k, p, "3D4", st, ln, dt) for (_, k, p, st, ln, dt) in reqs]
    shapes = [(d.p.shape[0], A.num_steps(ln, dt)) for d, (_, _, _, _, ln, dt) in zip(descs, reqs)]
    tickets = [C.c_uint64() for _ in reqs]
    lats = [C.c_double() for _ in reqs]
    ptrs = [C.POINTER(C.c_double)() for _ in reqs]
    ns = [C.c_size_t() for _ in reqs]
    args_s = [(srv.h, int(r[0]), C.byref(d.s), C.byref(t)) for r, d, t in zip(reqs, descs, tickets)]
    submit, wait_view = L.ekg_server_submit, L.ekg_server_wait_view
    t0 = time.perf_counter()
    nxt = t0
    for a, gap in zip(args_s, gaps):
        nxt += gap
        while time.perf_counter() < nxt:
            pass
        A.check(submit(*a))
    for t, pt, n, lt in zip(tickets, ptrs, ns, lats):
        A.check(wait_view(srv.h, t, C.byref(pt), C.byref(n), C.byref(lt)))
    el = time.perf_counter() - t0
    kept = {}
    for i in keep:
        Li, Ti = shapes[i]
        assert ns[i].value == Li * Ti
        kept[i] = np.ctypeslib.as_array(ptrs[i], shape=(Li, Ti)).copy()
    for t in tickets:
        A.check(L.ekg_server_release(srv.h, t))
    return len(reqs) / el, [lt.value for lt in lats], kept


def stress(args):
    """The large configuration: a 256^3 heart (2,964,088 myocardial voxels),
    12 leads, 600 ms at dt 0.25 (2400 samples), 1024 parameter vectors served
    through the continuous-batching server in waves admitted by its device
    budget.  Prints one JSON line; two ECGs are checked against the FP64
    oracle on an activation window and a repolarisation window."""
    import torch
    import ekgsim_amd as E
    n, L, R, dt, length = 256, 12, args.stress_requests, 0.25, 600.0
    h = S.make_heart(n, seed=1)
    reqs = []
    for r in range(R):
        x = S.random_params(3, L, 4000 + r)
        reqs.append((S.layer_k_from_params(x, 3), S.move_leads(S.lead_positions(n, L), S.displacements(x, 3, L))))
    ctx = E.Context(0)
    H = E.Heart(ctx, h.layer, h.seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H.excite_delay(h.delay)
    ca_ms = (time.perf_counter() - t0) * 1e3
    budget = args.stress_budget_gib << 30
    srv = E.Server(ctx, hbm_budget_bytes=budget, max_batch=256, n_streams=2)
    hid = srv.add_heart(H)
    # One warm-up request (kernel loading).  The first stream of R requests
    # grows the server's workspaces and pinned result buffers as it goes
    # (cold); the second meets them grown (warm, the steady state of a server
    # that keeps running) and is the reported rate.  Both submit every request
    # at once and take the results in place; requests 0 and R-1 are copied
    # for the oracle check.
    srv.wait(srv.submit(hid, reqs[0][0], reqs[0][1], "3D4", 0.0, length, dt))
    sreqs = [(hid, k, p, 0.0, length, dt) for k, p in reqs]
    cold_rate, cold_lat, _ = poisson_burst(srv, sreqs, 1e12, 41)
    log(f"stress, cold stream: {cold_rate:.1f} ECGs/s, p50 {np.percentile(cold_lat, 50):.1f} ms")
    g0 = srv.stats()["groups"]
    rate, lat, kept = poisson_burst(srv, sreqs, 1e12, 42, keep=(0, R - 1))
    el = R / rate
    st = srv.stats()
    st["groups"] -= g0
    srv.close()
    got = {r: (kept[r],) for r in kept}
    log(f"stress: {R} requests in {el:.2f} s, {st}")
    err = None
    if not args.no_cpu:
        from oracle import oracle as O
        act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
        num = den = 0.0
        for r in (0, R - 1):
            k, pos = reqs[r]
            for (st_ms, ln) in ((40.0, 10.0), (300.0, 10.0)):
                ref = O.run(h.layer, act, k, pos, "3D4", st_ms, ln, dt)
                j0 = int(st_ms / dt)
                seg = got[r][0][:, j0:j0 + ref.shape[1]]
                num += float(np.sum((seg - ref) ** 2))
                den += float(np.sum(ref ** 2))
        err = math.sqrt(num / den)
        log(f"stress: rel-L2 vs FP64 oracle on the windows {err:.2e}")
    print(json.dumps({
        "metric": "simulated ECGs/sec at 256^3, 600 ms at 0.25 ms (2400 samples), 12 leads, served in "
                  "HBM-admitted waves",
        "value": round(R / el, 2), "unit": "ECGs/s", "n_gpus": 1, "requests": R,
        "seconds": round(el, 3), "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2), "groups": st["groups"],
        "cold_ecgs_per_s": round(cold_rate, 2), "cold_p50_ms": round(float(np.percentile(cold_lat, 50)), 2),
        "peak_device_bytes": st["peak_bytes"], "hbm_budget_bytes": budget, "excitation_ms": round(ca_ms, 2),
        "rel_l2_vs_fp64_windows": err,
        "config": {"workload": f"{n}^3 ellipsoid heart ({h.n_myo} myocardial voxels), {L} leads, "
                               f"{int(length / dt)} steps of {dt} ms, 3D4 dipoles, {R} parameter vectors"},
    }), flush=True)


# ---------------------------------------------------------------------------
# CPU baseline leg (runs before the GPU is touched; forked workers)
# ---------------------------------------------------------------------------
_CPU = {}


def _cpu_worker(_i):
    from oracle import oracle as O
    c = _CPU
    out = []
    t0 = time.perf_counter()
    fn = O.ref_run if c["kind"] == "reference" else (lambda *a, **k: O.run(*a, literal=True, **k))
    e1 = fn(c["layer"], c["act"], c["k"], c["pos"], "3D4", c["fixed_start"], 1.0, 1.0)
    t_fixed = time.perf_counter() - t0
    wins = []
    for s in c["windows"]:
        t0 = time.perf_counter()
        e = fn(c["layer"], c["act"], c["k"], c["pos"], "3D4", float(s), float(c["w"]), 1.0)
        wins.append(time.perf_counter() - t0)
        out.append(e)
    return t_fixed, wins, out, e1


def cpu_baseline(h, act, k0, pos0, T, cores, w, windows):
    from oracle import oracle as O
    kind = "reference" if O.have_ref() else "port"
    _CPU.update(kind=kind, layer=h.layer, act=act, k=k0, pos=pos0, w=w, windows=windows,
                fixed_start=float(windows[0]))
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, range(cores))
    wall = time.perf_counter() - t0
    fixed = float(np.median([r[0] for r in res]))
    per_step = float(np.median([(sum(r[1]) - len(windows) * fixed) / (len(windows) * w) for r in res]))
    per_ecg = fixed + T * per_step
    ecg_per_s = cores / per_ecg
    return {
        "kind": kind, "cores": cores, "per_ecg_s": per_ecg, "ecgs_per_s": ecg_per_s,
        "wall_s": wall, "windows": [(int(s), int(s + w)) for s in windows],
        "window_ecg": res[0][2], "fixed_s": fixed, "per_step_s": per_step,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--leads", type=int, default=12)
    ap.add_argument("--length", type=float, default=600.0)
    ap.add_argument("--plain-calls", type=int, default=40)
    ap.add_argument("--serve-requests", type=int, default=384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--cpu-window", type=int, default=5)
    ap.add_argument("--check", type=int, default=3, help="ECGs checked against the FP64 oracle")
    ap.add_argument("--kernel", default="auto", choices=["auto", "valu", "mfma"],
                    help="lead-contraction path of the fused kernel")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the batched steps (for rocprofv3 runs)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--tune", default=None,
                    help="fixed launch shape chunk,waves,phase,kernel[,class_min] (e.g. 256,4,2,2) "
                         "instead of autotuning")
    ap.add_argument("--no-general", action="store_true",
                    help="skip the timing with activation classes off")
    ap.add_argument("--no-graph", action="store_true", help="launch the timed steps eagerly")
    ap.add_argument("--stress", action="store_true",
                    help="time the large configuration (256^3, dt 0.25, 1024 vectors) instead")
    ap.add_argument("--stress-requests", type=int, default=1024)
    ap.add_argument("--stress-budget-gib", type=int, default=16)
    ap.add_argument("--no-testrun", action="store_true", help="skip the shipped 24-layer heart")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="one device-to-device copy of MIB MiB before the steps (scale for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path "
              "(requests would shard across GPUs by request, with no exchange between them).", file=sys.stderr)
        sys.exit(2)
    if args.stress:
        stress(args)
        return

    n, L, B = args.grid, args.leads, args.batch
    T = int(math.ceil(args.length / 1.0))
    h, ks, xs = workload(n, L, B)
    V = h.n_myo
    log(f"heart {n}^3: {V} myocardial voxels; batch {B} x {L} leads x {T} steps")

    # ---- CPU leg first: the FP64 oracle / compiled reference on the host ----
    cpu = None
    ref_full = {}
    if not args.no_cpu and not args.profile_only:
        from oracle import oracle as O
        act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
        pos_r = [S.move_leads(base, S.displacements(x, 3, L)) for (x, base) in xs]
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        windows = [0, 100, 200, 300, 400, 500]
        cpu = cpu_baseline(h, act, ks[0], pos_r[0], T, cores, args.cpu_window, windows)
        log(f"cpu baseline ({cpu['kind']}, {cores} cores): {cpu['per_ecg_s']:.1f} s per ECG, "
            f"{cpu['ecgs_per_s']:.4f} ECGs/s (wall {cpu['wall_s']:.1f} s)")
        for r in sorted({0, B // 2, B - 1})[:max(0, args.check)]:
            t0 = time.perf_counter()
            ref_full[r] = O.run(h.layer, act, ks[r], pos_r[r], "3D4", 0.0, args.length, 1.0)
            log(f"oracle FP64 ECG {r}: {time.perf_counter() - t0:.1f} s")

    # ---- GPU ----
    import torch
    import ekgsim_amd as E

    ctx = E.Context(0)
    ctx.set_kernel(args.kernel)
    heart = E.Heart(ctx, h.layer, h.seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    heart.excite_delay(h.delay)
    ca_ms = (time.perf_counter() - t0) * 1e3
    pos = np.stack([E.move_leads(base, S.displacements(x, 3, L)) for (x, base) in xs])
    kk = np.stack(ks)
    dev = torch.device("cuda", 0)
    d_k = torch.from_numpy(kk).to(dev)
    d_p = torch.from_numpy(pos).to(dev)
    d_e = torch.empty((B, L, T), dtype=torch.float64, device=dev)

    def step():
        heart.simulate_batch_device(d_k.data_ptr(), d_p.data_ptr(), d_e.data_ptr(), B, 3, L,
                                    "3D4", 0.0, args.length, 1.0)

    def timed_steps(graphs):
        ctx.set_graphs(graphs)
        step()
        ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        gs = ctx.graph_stats()
        ctx.set_graphs(False)
        return el, gs

    tuning = None
    if args.tune:
        tv = [int(v) for v in args.tune.split(",")]
        heart.set_tune(*tv)
        tuning = heart.get_tune()
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tuning = heart.autotune(ks[0], pos[0], "3D4", 0.0, args.length, 1.0, batch=B)
        tuning["seconds"] = round(time.perf_counter() - t0, 2)
        log(f"autotune: {tuning}")
    if args.calibrate_copy > 0:
        src = torch.zeros(args.calibrate_copy << 17, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    # fused-kernel time: HIP events on the context stream over eager steps
    ctx.set_profiling(True)
    ctx.reset_profiling()
    for _ in range(min(args.steps, 5)):
        step()
    ctx.synchronize()
    nl, kms = ctx.kernel_time("ecg_fused")
    ctx.set_profiling(False)
    # timed steps (replayed from a captured hipGraph unless --no-graph)
    el, gstats = timed_steps(not args.no_graph)
    ms_step = el * 1e3 / args.steps
    value = B * args.steps / el
    avg_kernel_s = kms / 1e3 / max(nl, 1)
    act_grid = heart.activation()
    act_myo = act_grid[h.layer > 0]
    info = heart.layout_info("3D4") if args.kernel != "valu" else None
    n_classes = info["classes"] if info and heart.get_tune()["class_min"] >= 0 else 0
    win = cells = ccells = centries = 0
    if n_classes > 0:
        # the class-window kernel: every entry's weight and every class cell's
        # contraction once per (request, lead)
        dom = "ecg_class_window_kernel"
        ccells = B * L * info["class_cells"]
        centries = B * L * info["class_entries"]
        f64 = FP64_PER_CLASS_CELL * ccells + FP64_PER_CLASS_ENTRY * centries
        f32 = FP32_PER_CLASS_ENTRY * centries
    else:
        dom = "ecg_chunk_kernel" if args.kernel != "valu" else "ecg_fused_kernel"
        win = sum(window_voxel_steps(act_myo, float(k[0][1]), 0.0, 1.0, T) for k in ks)
        cells = B * L * lead_field_cells(h.layer) if args.kernel != "valu" else 0
        f64 = 2.0 * L * win + FP64_PER_CELL * cells
        f32 = FP32_PER_WINDOW_STEP * win + FP32_PER_CELL * cells
    flops = f64 + f32
    achieved = flops / avg_kernel_s
    naive_rate = value * V * T
    log(f"batched: {ms_step:.2f} ms/step, {value:.1f} ECGs/s; {dom} {avg_kernel_s * 1e3:.2f} ms; "
        f"{achieved / 1e12:.2f} TFLOP/s FP64-equivalent; layout {info}")
    if args.profile_only:
        print(json.dumps({"ms_per_step": ms_step, "dominant_kernel_ms": avg_kernel_s * 1e3,
                          "calibration_copy_bytes": args.calibrate_copy << 20}))
        return

    ecg_gpu = d_e.cpu().numpy()

    # accuracy vs FP64 oracle (full ECGs) and vs the compiled reference (windows)
    errs = {}
    for r, ref in ref_full.items():
        errs[r] = float(np.linalg.norm(ecg_gpu[r] - ref) / np.linalg.norm(ref))
    rel_max = max(errs.values()) if errs else None
    max_err = max_abs_err(ecg_gpu, ref_full)
    ref_win_err = None
    if cpu is not None:
        num = den = 0.0
        for (s, e), we in zip(cpu["windows"], cpu["window_ecg"]):
            num += float(np.sum((ecg_gpu[0][:, s:e] - we) ** 2))
            den += float(np.sum(we ** 2))
        ref_win_err = math.sqrt(num / den)
        log(f"rel-L2 vs FP64 oracle {errs}; vs {cpu['kind']} on windows {ref_win_err:.2e}")

    # ---- criteria on the device for the whole batch (correlation, peaks) ----
    tg = ecg_gpu[0] / (ecg_gpu[0].max(axis=1, keepdims=True) - ecg_gpu[0].min(axis=1, keepdims=True))
    offs = 1.0 - tg.min(axis=1) / (tg.max(axis=1) - tg.min(axis=1))
    d_c = torch.empty((B, 2 * L), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.criteria_device("correlation", tg, offs, d_e.data_ptr(), B, L, T, d_c.data_ptr(), peaks=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        ctx.criteria_device("correlation", tg, offs, d_e.data_ptr(), B, L, T, d_c.data_ptr(), peaks=True)
    ctx.synchronize()
    crit_ms = (time.perf_counter() - t0) * 1e3 / 10
    log(f"criteria on the device: {crit_ms:.3f} ms per batch of {B} x {L} leads")

    # ---- the same batch with activation classes off: every voxel's AP ----
    general = None
    if not args.no_general and args.kernel != "valu" and n_classes > 0:
        saved = heart.get_tune()
        heart.set_tune(class_min=-1)
        gtune = None
        if args.tune:
            heart.set_tune(saved["chunk_voxels"], saved["mfma_waves"], saved["lead_phase"], saved["kernel"], -1)
        elif not args.no_autotune:
            gtune = heart.autotune(ks[0], pos[0], "3D4", 0.0, args.length, 1.0, batch=B)
        for _ in range(args.warmup):
            step()
        gel, _ = timed_steps(not args.no_graph)
        gecg = d_e.cpu().numpy()
        gerr = {r: float(np.linalg.norm(gecg[r] - ref) / np.linalg.norm(ref)) for r, ref in ref_full.items()}
        general = {"ecgs_per_s": round(B * args.steps / gel, 2), "ms_per_step": round(gel * 1e3 / args.steps, 3),
                   "autotuned": gtune, "rel_l2_max_vs_fp64": max(gerr.values()) if gerr else None,
                   "max_err_vs_fp64": max_abs_err(gecg, ref_full)}
        log(f"activation classes off: {general}")
        heart.set_tune(saved["chunk_voxels"], saved["mfma_waves"], saved["lead_phase"], saved["kernel"],
                       saved["class_min"])

    # ---- plain drop-in: one ekg_simulate per simulation, host in/out ----
    lat = []
    for i in range(args.plain_calls + 3):
        t1 = time.perf_counter()
        heart.simulate(ks[i % B], pos[i % B], "3D4", 0.0, args.length, 1.0)
        if i >= 3:
            lat.append(time.perf_counter() - t1)
    plain = len(lat) / sum(lat)
    log(f"plain drop-in: {plain:.1f} ECGs/s (p50 {np.percentile(lat, 50) * 1e3:.2f} ms)")
    # the same calls with the context's graph mode on (off by default): the
    # launches of the single-request group replay from a captured hipGraph
    ctx.set_graphs(True)
    lat_g = []
    for i in range(args.plain_calls + 3):
        t1 = time.perf_counter()
        heart.simulate(ks[i % B], pos[i % B], "3D4", 0.0, args.length, 1.0)
        if i >= 3:
            lat_g.append(time.perf_counter() - t1)
    ctx.set_graphs(False)
    plain_g = len(lat_g) / sum(lat_g)
    log(f"plain drop-in, graph mode: {plain_g:.1f} ECGs/s (p50 {np.percentile(lat_g, 50) * 1e3:.2f} ms)")

    # ---- serving burst: 4 cached hearts, Poisson arrivals, mixed shapes ----
    serve = None
    if args.serve_requests > 0:
        srv = E.Server(ctx, hbm_budget_bytes=48 << 30, max_batch=64, n_streams=3)
        hs = []
        for gi, g in enumerate((64, 96, 128, 160)):
            hh = S.make_heart(g, seed=gi + 1)
            H = E.Heart(ctx, hh.layer, hh.seed)
            H.excite_delay(hh.delay)
            hs.append((srv.add_heart(H), g, H))
        rng = np.random.default_rng(7)
        reqs = []
        for r in range(args.serve_requests):
            hid, g, _ = hs[rng.integers(0, 4)]
            nl = int(rng.integers(2, 13))
            dur = float(rng.integers(400, 801))
            x = S.random_params(3, nl, 5000 + r)
            reqs.append((hid, S.layer_k_from_params(x, 3),
                         E.move_leads(S.lead_positions(g, nl), S.displacements(x, 3, nl)), dur))
        # Poisson arrivals at a rate above the service rate (saturating).  The
        # server's workspaces and pinned buffers grow over the first few
        # bursts (group compositions differ from burst to burst); after
        # SERVE_WARM of them, the median of SERVE_MEASURED bursts is reported
        # (steady state)
        sreq = [(hid, k, p, 0.0, dur, 1.0) for hid, k, p, dur in reqs]
        first = poisson_burst(srv, sreq, 4.0 * value, 6)[0]
        for sd in range(SERVE_WARM - 1):
            poisson_burst(srv, sreq, 4.0 * value, 20 + sd)
        g0 = srv.stats()["groups"]
        runs = [poisson_burst(srv, sreq, 4.0 * value, 7 + sd) for sd in range(SERVE_MEASURED)]
        st = srv.stats()
        srv.close()
        rate_m, lats, _ = sorted(runs, key=lambda x: x[0])[SERVE_MEASURED // 2]
        serve = {"requests": len(reqs), "ecgs_per_s": rate_m,
                 "p50_ms": float(np.percentile(lats, 50)), "p99_ms": float(np.percentile(lats, 99)),
                 "groups": (st["groups"] - g0) / SERVE_MEASURED, "peak_inflight": st["peak_inflight"],
                 "bursts_ecgs_per_s": [x[0] for x in runs], "first_burst_ecgs_per_s": first}
        log(f"serving: {serve}")

    # ---- serving on the batch's own heart and shape, like for like with the
    # plain drop-in: every request 128^3 x 12 leads x 600 ms ----
    same = None
    if args.serve_requests > 0:
        srv = E.Server(ctx, hbm_budget_bytes=48 << 30, max_batch=128, n_streams=3)
        hid = srv.add_heart(heart)
        sreqs = [(hid, ks[r % B], pos[r % B], 0.0, args.length, 1.0) for r in range(2 * B)]
        first = poisson_burst(srv, sreqs, 4.0 * value, 9)[0]
        for sd in range(SERVE_WARM - 1):
            poisson_burst(srv, sreqs, 4.0 * value, 30 + sd)
        g0 = srv.stats()["groups"]
        chk = (0, 1, B - 1, B, 2 * B - 1)
        runs = [poisson_burst(srv, sreqs, 4.0 * value, 11 + sd, keep=chk) for sd in range(SERVE_MEASURED)]
        st = srv.stats()
        srv.close()
        same_bits = all(np.array_equal(outs[r], ecg_gpu[r % B]) for _, _, outs in runs for r in chk)
        rate_s, lats_s, _ = sorted(runs, key=lambda x: x[0])[SERVE_MEASURED // 2]
        same = {"requests": len(sreqs), "ecgs_per_s": rate_s, "p50_ms": float(np.percentile(lats_s, 50)),
                "p99_ms": float(np.percentile(lats_s, 99)), "groups": (st["groups"] - g0) / SERVE_MEASURED,
                "vs_plain": rate_s / plain, "served_equal_batched_bitwise": same_bits,
                "bursts_ecgs_per_s": [x[0] for x in runs], "first_burst_ecgs_per_s": first}
        log(f"serving, one heart and shape: {same}")

    # ---- the reference's shipped 24-layer heart (FP64 conduction: distinct
    # activation times, so no activation classes) ----
    testrun = None
    if not args.no_testrun and os.path.exists(os.path.join(TESTRUN_DIR, "heart.npz")):
        tl, tsd, tcond, tks, tps = testrun_workload(B)
        th = E.Heart(ctx, tl, tsd)
        th.excite_transfer(tcond)
        tlen, tstart = 400.0, 100.0
        tT = int(tlen)
        dk = torch.from_numpy(tks).to(dev)
        dp = torch.from_numpy(tps).to(dev)
        de = torch.empty((B, tps.shape[1], tT), dtype=torch.float64, device=dev)
        ttune = None
        if not args.no_autotune:
            ttune = th.autotune(tks[0], tps[0], "3D4", tstart, tlen, 1.0, batch=B)

        def tstep():
            th.simulate_batch_device(dk.data_ptr(), dp.data_ptr(), de.data_ptr(), B, 24, tps.shape[1], "3D4",
                                     tstart, tlen, 1.0)

        ctx.set_graphs(not args.no_graph)
        tstep()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tstep()
        ctx.synchronize()
        tel = time.perf_counter() - t0
        ctx.set_graphs(False)
        tecg = de.cpu().numpy()
        tlat = []
        for i in range(23):
            t1 = time.perf_counter()
            th.simulate(tks[i], tps[i], "3D4", tstart, tlen, 1.0)
            if i >= 3:
                tlat.append(time.perf_counter() - t1)
        terr = tbits = None
        if not args.no_cpu:
            from oracle import oracle as O
            tact = O.excite_fp64(tl, tsd, tcond)
            tbits = bool(np.array_equal(tact.view(np.uint64), th.activation().view(np.uint64)))
            tref = O.run(tl, tact, tks[0], tps[0], "3D4", tstart, tlen, 1.0)
            terr = float(np.linalg.norm(tecg[0] - tref) / np.linalg.norm(tref))
        tinfo = th.layout_info("3D4")
        testrun = {"ecgs_per_s": round(B * args.steps / tel, 2), "ms_per_batch": round(tel * 1e3 / args.steps, 3),
                   "plain_ecgs_per_s": round(len(tlat) / sum(tlat), 2),
                   "myocardial_voxels": int((tl > 0).sum()), "layers": 24, "leads": int(tps.shape[1]),
                   "steps": tT, "class_voxels": tinfo["class_voxels"], "autotuned": ttune,
                   "rel_l2_vs_fp64": terr, "activation_bit_exact_vs_oracle": tbits,
                   "reference_published_s_per_ecg": 439.131}
        log(f"shipped 24-layer heart: {testrun}")
        th.close()

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "ECGs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lead weights and per-voxel APs, f64 class APs and sums",
        "data": "synthetic (seeded ellipsoid heart, random parameter vectors)",
        "config": {
            "workload": (f"batched: {B} parameter vectors per step on one cached {n}^3 ellipsoid "
                         f"heart ({V} myocardial voxels), {L} leads, {T} steps of 1 ms, 3D4 dipoles"),
            "grid": n, "myocardial_voxels": V, "leads": L, "steps_per_ecg": T, "batch": B,
            "neighbourhood": "3D4", "excitation": "int32 per-voxel delay 1-3, 5 endocardial seeds",
            "launch": {"autotuned": tuning, "graph_replay": not args.no_graph, "graph": gstats},
            "layout": info,
        },
        "classes_off": general,
        "plain_ecgs_per_s": round(plain, 2),
        "plain_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "plain_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "plain_graph_ecgs_per_s": round(plain_g, 2),
        "plain_graph_p50_ms": round(float(np.percentile(lat_g, 50)) * 1e3, 3),
        "batched_vs_plain": round(value / plain, 2),
        "serving": serve,
        "serving_same_heart": same,
        "p99_ms": None if serve is None else round(serve["p99_ms"], 3),
        "testrun_model24": testrun,
        "rel_l2_max_vs_fp64": rel_max,
        "max_err_vs_fp64": max_err,
        "rel_l2_vs_reference_windows": ref_win_err,
        "excitation_ms": round(ca_ms, 2),
        "criteria_ms_per_batch": round(crit_ms, 4),
        "roofline": {
            "bound": "mfma",
            "kernel": dom,
            "achieved": round(achieved / 1e12, 4),
            "peak": MFMA_FP64_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / MFMA_FP64_PEAK, 5),
            "traffic": TRAFFIC_BYTES.get(dom),
            "pipes": PIPES.get(dom),
            "algorithmic_flops_per_launch": flops,
            "fp64_flops_per_launch": f64,
            "fp32_flops_per_launch": f32,
            "class_cell_leads_per_launch": ccells,
            "class_entry_leads_per_launch": centries,
            "lead_weight_evaluations_per_launch": cells,
            "window_voxel_steps_per_launch": win,
            "per_unit": {"class_entry_lead": {"fp32": FP32_PER_CLASS_ENTRY, "fp64": FP64_PER_CLASS_ENTRY},
                         "class_cell_lead": {"fp32": 0, "fp64": FP64_PER_CLASS_CELL},
                         "lead_weight_evaluation": {"fp32": FP32_PER_CELL, "fp64": FP64_PER_CELL},
                         "window_voxel_step": {"fp32": FP32_PER_WINDOW_STEP, "fp64": 2 * L},
                         "fp32_weight": 1.0},
            "naive_voxel_steps_per_s": naive_rate,
            "naive_model_frac": round(naive_rate * NAIVE_LANE_CYCLES / LANE_PEAK, 3),
            "kernel_ms": round(avg_kernel_s * 1e3, 4),
            "launches": nl,
        },
        "cpu_baseline": None if cpu is None else {
            "value": cpu["ecgs_per_s"], "unit": "ECGs/s", "cores": cpu["cores"], "kind": cpu["kind"],
            "sample": (f"{n}^3 heart, {L} leads, request 0: steps {cpu['windows']} of {T} "
                       f"plus a 1-step run for the per-call setup, one process per core; "
                       f"{cpu['per_ecg_s']:.1f} s per ECG per core"),
        },
        "vs_cpu": None if cpu is None else round(value / cpu["ecgs_per_s"], 1),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

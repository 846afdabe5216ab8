"""The batched and serving configurations at their stated sizes.

Batched (scope config 2): 256 parameter vectors on one cached 128^3 ellipsoid
heart (370,656 myocardial voxels), 12 leads, 600 steps of 1 ms, 3D4 dipoles,
integer per-voxel delays 1-3 -- the bench's own batch, run through the
device-resident batch entry point with graph replay on.  Checked: activation
times bit-exact against the oracle's CA; three ECGs of the batch within
rel-L2 1e-5 of the FP64 oracle (ekg_oracle.c, the restatement of
simulator.cpp:452-550); eight requests bitwise equal to their plain
one-call simulation; a replayed graph bitwise equal to its first launch.
Both with activation classes (the default) and without (every voxel's AP).

Serving (scope config 3): four cached hearts (64/96/128/160^3, seeds 1-4),
requests of 2-12 leads and 400-800 ms with Poisson arrivals into a server
under a device budget that bounds every group; each served ECG equals its
alone call bit for bit, and one request per heart is within 1e-5 of the
oracle.
"""
import threading
import time

import numpy as np
import pytest

from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5

_ORACLE = {}


def _oracle_act(h, key):
    if ("act",) + key not in _ORACLE:
        _ORACLE[("act",) + key] = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    return _ORACLE[("act",) + key]


def _oracle_ecg(h, key, r, k, pos, length):
    kk = ("ecg",) + key + (r,)
    if kk not in _ORACLE:
        t0 = time.perf_counter()
        _ORACLE[kk] = O.run(h.layer, _oracle_act(h, key), k, pos, "3D4", 0.0, length, 1.0)
        print(f"oracle ECG {key} request {r}: {time.perf_counter() - t0:.1f} s", flush=True)
    return _ORACLE[kk]


def _batch_inputs(n=128, L=12, B=256):
    h = S.make_heart(n, seed=1)
    base = S.lead_positions(n, L)
    ks, ps = [], []
    for r in range(B):
        x = S.random_params(3, L, 1000 + r)
        ks.append(S.layer_k_from_params(x, 3))
        ps.append(S.move_leads(base, S.displacements(x, 3, L)))
    return h, np.stack(ks), np.stack(ps)


@pytest.mark.parametrize("class_min", [0, -1], ids=["classes", "classes_off"])
def test_batched_256_on_128cubed_heart(class_min):
    import torch

    import ekgsim_amd as E
    n, L, B, T = 128, 12, 256, 600
    h, ks, ps = _batch_inputs(n, L, B)
    ctx = E.Context(0)
    heart = E.Heart(ctx, h.layer, h.seed)
    heart.excite_delay(h.delay)
    act = heart.activation()
    assert np.array_equal(act, _oracle_act(h, (n, 1)))
    heart.set_tune(class_min=class_min)
    info = heart.layout_info("3D4")
    print(f"batched 128^3 x {B}, class_min {class_min}: layout {info}", flush=True)
    assert info["voxels"] == 370656
    if class_min == 0:
        assert info["class_voxels"] > 0.99 * info["voxels"]
    else:
        assert info["class_voxels"] == 0
    dev = torch.device("cuda", 0)
    d_k = torch.from_numpy(ks).to(dev)
    d_p = torch.from_numpy(ps).to(dev)
    d_e = torch.full((B, L, T), float("nan"), dtype=torch.float64, device=dev)
    ctx.set_graphs(True)
    heart.simulate_batch_device(d_k.data_ptr(), d_p.data_ptr(), d_e.data_ptr(), B, 3, L, "3D4", 0.0, float(T), 1.0)
    ctx.synchronize()
    first = d_e.cpu().numpy()
    d_e.fill_(float("nan"))
    torch.cuda.synchronize()
    heart.simulate_batch_device(d_k.data_ptr(), d_p.data_ptr(), d_e.data_ptr(), B, 3, L, "3D4", 0.0, float(T), 1.0)
    ctx.synchronize()
    second = d_e.cpu().numpy()
    gs = ctx.graph_stats()
    ctx.set_graphs(False)
    assert gs["launches"] >= 2 and gs["captures"] >= 1
    assert np.all(np.isfinite(first))
    assert np.array_equal(first.view(np.uint64), second.view(np.uint64))
    for r in (0, 128, 255):
        e = rel_l2(first[r], _oracle_ecg(h, (n, 1), r, ks[r], ps[r], float(T)))
        print(f"request {r}: rel-L2 vs FP64 oracle {e:.2e}", flush=True)
        assert e <= TOL
    for r in (0, 1, 37, 100, 128, 200, 254, 255):
        alone = heart.simulate(ks[r], ps[r], "3D4", 0.0, float(T), 1.0)
        assert np.array_equal(alone.view(np.uint64), first[r].view(np.uint64)), r
    heart.close()
    ctx.close()


def test_serving_four_hearts_poisson_under_budget():
    import ekgsim_amd as E
    ctx = E.Context(0)
    hearts = []
    for gi, g in enumerate((64, 96, 128, 160)):
        hh = S.make_heart(g, seed=gi + 1)
        H = E.Heart(ctx, hh.layer, hh.seed)
        H.excite_delay(hh.delay)
        assert np.array_equal(H.activation(), _oracle_act(hh, (g, gi + 1)))
        hearts.append((hh, H, g))
    rng = np.random.default_rng(31)
    reqs = []
    for r in range(48):
        hi = r if r < 4 else int(rng.integers(0, 4))  # every heart is requested
        hh, H, g = hearts[hi]
        nl = int(rng.integers(2, 13))
        dur = float(rng.integers(400, 801))
        x = S.random_params(3, nl, 5000 + r)
        reqs.append((hi, S.layer_k_from_params(x, 3),
                     E.move_leads(S.lead_positions(g, nl), S.displacements(x, 3, nl)), dur))
    budget = 24 << 20
    srv = E.Server(ctx, hbm_budget_bytes=budget, max_batch=32, n_streams=3)
    ids = [srv.add_heart(H) for _, H, _ in hearts]
    gaps = rng.exponential(1.0 / 20000.0, size=len(reqs))
    tickets = [None] * len(reqs)

    def producer():
        nxt = time.perf_counter()
        for i, ((hi, k, p, dur), gap) in enumerate(zip(reqs, gaps)):
            nxt += gap
            while time.perf_counter() < nxt:
                pass
            tickets[i] = srv.submit(ids[hi], k, p, "3D4", 0.0, dur, 1.0)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    got = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    lat = [g[1] for g in got]
    print(f"serving: {st}; p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms", flush=True)
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= budget
    assert st["groups"] > 1 and st["peak_inflight"] > 1
    checked = set()
    for i, ((hi, k, p, dur), (ecg, _)) in enumerate(zip(reqs, got)):
        hh, H, g = hearts[hi]
        alone = H.simulate(k, p, "3D4", 0.0, dur, 1.0)
        assert np.array_equal(alone.view(np.uint64), ecg.view(np.uint64)), i
        if hi not in checked:
            e = rel_l2(ecg, _oracle_ecg(hh, (g, hi + 1), i, k, p, dur))
            print(f"heart {g}^3 request {i} ({k.shape[0]} layers, {p.shape[0]} leads, {dur:g} ms): "
                  f"rel-L2 {e:.2e}", flush=True)
            assert e <= TOL
            checked.add(hi)
    assert checked == {0, 1, 2, 3}
    for _, H, _ in hearts:
        H.close()
    ctx.close()

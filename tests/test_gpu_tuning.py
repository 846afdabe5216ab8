"""Launch tuning and graph replay keep the results.

The autotuner only changes launch shapes (chunk size, fused-kernel waves,
lead-field phase width, contraction path): every setting must give ECGs
within rel-L2 1e-5 of the FP64 oracle.  Graph replay must give the eager
result bit for bit, on every replay, with new inputs written into the same
device buffers between replays and eager work on the same context in
between (so a replay cannot be served by stale state).
"""
import numpy as np
import pytest

from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _setup(ctx, n=40, seed=9, L=6):
    from ekgsim_amd import Heart
    h = S.make_heart(n, seed=seed)
    g = Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    return h, g, act


def _req(n, L, s):
    x = S.random_params(3, L, s)
    return S.layer_k_from_params(x, 3), S.move_leads(S.lead_positions(n, L), S.displacements(x, 3, L))


def test_every_tuning_keeps_parity(ctx):
    h, g, act = _setup(ctx)
    k, pos = _req(40, 6, 31)
    ref = O.run(h.layer, act, k, pos, "3D4", 0.0, 400.0, 1.0)
    for chunk in (64, 256, 512):
        for waves in (2, 16):
            for phase in (2, 4):
                g.set_tune(chunk, waves, phase)
                got = g.simulate(k, pos, "3D4", 0.0, 400.0, 1.0)
                assert rel_l2(got, ref) <= TOL, (chunk, waves, phase)
    g.set_tune()
    assert g.get_tune()["chunk_voxels"] == 0


def test_autotune_chooses_a_valid_shape(ctx):
    h, g, act = _setup(ctx, n=48, seed=2)
    k, pos = _req(48, 8, 41)
    t = g.autotune(k, pos, "3D4", 0.0, 300.0, 1.0, batch=8)
    assert 64 <= t["chunk_voxels"] <= 512
    assert t["lead_phase"] in (2, 4) and 1 <= t["mfma_waves"] <= 16 and t["ms_per_batch"] > 0
    assert g.get_tune()["chunk_voxels"] == t["chunk_voxels"]
    ref = O.run(h.layer, act, k, pos, "3D4", 0.0, 300.0, 1.0)
    assert rel_l2(g.simulate(k, pos, "3D4", 0.0, 300.0, 1.0), ref) <= TOL
    g.set_tune()


def test_graph_replay_equals_eager(ctx):
    import torch
    h, g, act = _setup(ctx, n=32, seed=4)
    B, L, T = 6, 5, 300
    dev = torch.device("cuda", 0)
    d_k = torch.empty((B, 3, 9), dtype=torch.float64, device=dev)
    d_p = torch.empty((B, L, 3), dtype=torch.float64, device=dev)
    d_e = torch.empty((B, L, T), dtype=torch.float64, device=dev)

    def fill(s0):
        ks, ps = zip(*[_req(32, L, s0 + b) for b in range(B)])
        d_k.copy_(torch.from_numpy(np.stack(ks)))
        d_p.copy_(torch.from_numpy(np.stack(ps)))
        return ks, ps

    def run():
        g.simulate_batch_device(d_k.data_ptr(), d_p.data_ptr(), d_e.data_ptr(), B, 3, L, "3D4", 0.0, float(T), 1.0)
        ctx.synchronize()
        return d_e.cpu().numpy()

    ctx.set_graphs(False)
    eager = {}
    for s0 in (100, 200, 300):
        fill(s0)
        eager[s0] = run()
    ctx.set_graphs(True)
    before = ctx.graph_stats()
    try:
        for rep, s0 in enumerate((100, 200, 300, 100, 300)):
            fill(s0)
            got = run()
            assert np.array_equal(got, eager[s0]), (rep, s0)
            # eager work on the same context between replays
            k, pos = _req(32, 3, 900 + rep)
            g.simulate(k, pos, "3D4", 0.0, 120.0, 1.0)
        # back-to-back batches with new inputs in the same buffers replay the
        # captured graph without a new capture
        fill(200)
        run()
        c0 = ctx.graph_stats()["captures"]
        for s0 in (100, 300):
            fill(s0)
            assert np.array_equal(run(), eager[s0])
        assert ctx.graph_stats()["captures"] == c0
    finally:
        ctx.set_graphs(False)
    st = ctx.graph_stats()
    # 5 batches and 5 single calls in the loop, then 3 batches
    assert st["launches"] - before["launches"] == 13
    ks, ps = fill(300)
    ref = O.run(h.layer, act, ks[0], ps[0], "3D4", 0.0, float(T), 1.0)
    assert rel_l2(eager[300][0], ref) <= TOL

"""Continuous batching: every request's ECG is the same alone and in the mix.

Requests of several hearts and shapes are submitted together, while others
run, and under an HBM budget that forces small groups; each result must equal
the plain one-call simulation of the same request bit for bit (every
request's partial sums come from its own rows over the heart's fixed
chunking, so grouping cannot change them), and the plain call must be within
rel-L2 1e-5 of the FP64 oracle.
"""
import threading
import time

import numpy as np
import pytest

from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _hearts(ctx, specs):
    from ekgsim_amd import Heart
    out = []
    for n, seed in specs:
        h = S.make_heart(n, seed=seed)
        H = Heart(ctx, h.layer, h.seed)
        H.excite_delay(h.delay)
        out.append((h, H))
    return out


def _requests(hearts, n_req, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(n_req):
        hi = int(rng.integers(0, len(hearts)))
        h, _ = hearts[hi]
        L = int(rng.choice([2, 3, 6]))
        length = float(rng.choice([200.0, 300.0]))
        x = S.random_params(3, L, 7000 + seed * 100 + r)
        k = S.layer_k_from_params(x, 3)
        pos = S.move_leads(S.lead_positions(h.layer.shape[0], L), S.displacements(x, 3, L))
        reqs.append((hi, k, pos, length))
    return reqs


def _alone(hearts, req):
    hi, k, pos, length = req
    return hearts[hi][1].simulate(k, pos, "3D4", 0.0, length, 1.0)


def test_mixed_burst_equals_alone(ctx):
    from ekgsim_amd import Server
    hearts = _hearts(ctx, [(24, 1), (32, 2), (20, 3)])
    reqs = _requests(hearts, 48, 1)
    srv = Server(ctx, max_batch=8, n_streams=3)
    ids = [srv.add_heart(H) for _, H in hearts]
    tickets = [srv.submit(ids[hi], k, pos, "3D4", 0.0, length, 1.0) for hi, k, pos, length in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    srv.close()
    for req, g in zip(reqs, got):
        assert np.array_equal(g, _alone(hearts, req))
    assert st["completed"] == len(reqs)
    assert st["groups"] < len(reqs)      # requests were batched
    assert st["peak_inflight"] > 1       # and ran concurrently
    # the plain call itself is the checked drop-in
    hi, k, pos, length = reqs[0]
    h = hearts[hi][0]
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    assert rel_l2(got[0], O.run(h.layer, act, k, pos, "3D4", 0.0, length, 1.0)) <= 1e-5


def test_arrivals_while_running_equal_alone(ctx):
    """A producer thread keeps submitting while earlier groups are on the
    GPU; results are collected in completion order."""
    from ekgsim_amd import Server
    hearts = _hearts(ctx, [(28, 4), (24, 5)])
    reqs = _requests(hearts, 40, 2)
    srv = Server(ctx, max_batch=4, n_streams=2)
    ids = [srv.add_heart(H) for _, H in hearts]
    tickets = [None] * len(reqs)
    done = threading.Event()

    def produce():
        for i, (hi, k, pos, length) in enumerate(reqs):
            tickets[i] = srv.submit(ids[hi], k, pos, "3D4", 0.0, length, 1.0)
            if i % 5 == 4:
                time.sleep(0.002)
        done.set()

    th = threading.Thread(target=produce)
    th.start()
    got = {}
    while len(got) < len(reqs):
        for i, t in enumerate(tickets):
            if t is not None and i not in got and srv.poll(t):
                got[i] = srv.wait(t)[0]
        time.sleep(0.0005)
    th.join()
    assert done.is_set()
    srv.close()
    for i, req in enumerate(reqs):
        assert np.array_equal(got[i], _alone(hearts, req))


def test_hbm_budget_limits_groups(ctx):
    """A budget of about two requests' device memory: groups shrink to fit,
    in-use bytes never exceed the budget, results are unchanged; a request
    larger than the whole budget fails with ENOMEM instead of running."""
    from ekgsim_amd import EkgError, Server
    hearts = _hearts(ctx, [(24, 6)])
    reqs = [(0, k, pos, 300.0) for _, k, pos, _ in _requests(hearts, 12, 3)]
    probe = Server(ctx, max_batch=1, n_streams=1)
    hid = probe.add_heart(hearts[0][1])
    t = probe.submit(hid, reqs[0][1], reqs[0][2], "3D4", 0.0, 300.0, 1.0)
    probe.wait(t)
    one = probe.stats()["peak_bytes"]
    probe.close()
    assert one > 0
    budget = int(2.5 * one)
    srv = Server(ctx, hbm_budget_bytes=budget, max_batch=8, n_streams=3)
    hid = srv.add_heart(hearts[0][1])
    tickets = [srv.submit(hid, k, pos, "3D4", 0.0, length, 1.0) for _, k, pos, length in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    assert st["peak_bytes"] <= budget
    assert st["groups"] >= len(reqs) / 2
    for req, g in zip(reqs, got):
        assert np.array_equal(g, _alone(hearts, req))
    # one request alone larger than the budget
    big = _requests(hearts, 1, 4)[0]
    t = srv.submit(hid, big[1], big[2], "3D4", 0.0, 50000.0, 1.0)
    with pytest.raises(EkgError):
        srv.wait(t)
    srv.close()


def test_hearts_on_different_paths_equal_alone():
    """Two hearts tuned to different contraction paths (matrix cores and
    vector ALUs) in one server: requests of the two never share a launch
    group (each heart keeps the layout of its own path), and every served
    ECG equals the alone call bit for bit."""
    from ekgsim_amd import Context, Server
    c = Context(0)  # the context leaves the path to each heart's tuning
    hearts = _hearts(c, [(24, 11), (28, 12)])
    hearts[0][1].set_tune(kernel="mfma")
    hearts[1][1].set_tune(kernel="valu")
    reqs = _requests(hearts, 24, 3)
    srv = Server(c, max_batch=8, n_streams=2)
    ids = [srv.add_heart(H) for _, H in hearts]
    tickets = [srv.submit(ids[hi], k, pos, "3D4", 0.0, length, 1.0) for hi, k, pos, length in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    srv.close()
    for req, g in zip(reqs, got):
        assert np.array_equal(g, _alone(hearts, req))
    for _, H in hearts:
        H.close()
    c.close()


@pytest.mark.parametrize("window_us", [-1, 0, 3000])
def test_batching_window_equals_alone(ctx, window_us):
    """The batching window (off, the 1 ms default, 3 ms) changes only how
    requests are grouped, never a result; with the window on, a burst does
    not take more launch groups than without it."""
    from ekgsim_amd import Server
    hearts = _hearts(ctx, [(24, 13), (20, 14)])
    reqs = _requests(hearts, 40, 4)
    alone = [_alone(hearts, r) for r in reqs]

    def burst(w):
        srv = Server(ctx, max_batch=16, n_streams=2, batch_window_us=w)
        ids = [srv.add_heart(H) for _, H in hearts]
        tickets = [srv.submit(ids[hi], k, pos, "3D4", 0.0, length, 1.0) for hi, k, pos, length in reqs]
        got = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
        srv.close()
        return got, st["groups"]

    got, groups = burst(window_us)
    for g, a in zip(got, alone):
        assert np.array_equal(g, a)
    if window_us >= 0:
        _, groups_off = burst(-1)
        print(f"window {window_us} us: {groups} groups; no window: {groups_off} groups", flush=True)
        assert groups <= groups_off + 1


def test_result_views_equal_copies(ctx):
    """ekg_server_wait_view hands out a request's ECG in the server's pinned
    buffer without copying: the view equals the plain simulation bit for bit
    and stays valid while later groups run (its group's buffer is held until
    release), copies and views mix freely, and a ticket is released once."""
    from ekgsim_amd import EkgError, Server
    hearts = _hearts(ctx, [(24, 21), (20, 22)])
    reqs = _requests(hearts, 24, 5)
    alone = [_alone(hearts, r) for r in reqs]
    srv = Server(ctx, max_batch=6, n_streams=2)
    ids = [srv.add_heart(H) for _, H in hearts]
    tickets = [srv.submit(ids[hi], k, pos, "3D4", 0.0, length, 1.0) for hi, k, pos, length in reqs]
    views = {}
    for i, t in enumerate(tickets):
        if i % 2 == 0:
            views[i] = (t, srv.wait_view(t)[0])
        else:
            assert np.array_equal(srv.wait(t)[0], alone[i]), i
    # a second burst reuses the pinned pool while the first burst's views are held
    more = [srv.submit(ids[hi], k, pos, "3D4", 0.0, length, 1.0) for hi, k, pos, length in reqs[:8]]
    for i, t in enumerate(more):
        assert np.array_equal(srv.wait(t)[0], alone[i]), ("second burst", i)
    for i, (t, v) in views.items():
        assert not v.flags.writeable
        assert np.array_equal(v, alone[i]), i
        srv.release(t)
    with pytest.raises(EkgError) as ei:
        srv.release(views[0][0])
    assert ei.value.code == 7  # EKG_ENOTFOUND: released already
    srv.close()

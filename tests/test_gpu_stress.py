"""Large and worst-case shapes.

* The stress configuration: a 256^3 heart (2,964,088 myocardial voxels),
  12 leads, 600 ms at dt 0.25 (2400 samples), many parameter vectors served
  in HBM-admitted waves.  Activation times are bit-exact against the oracle
  at full size; two ECGs are checked against the FP64 oracle on a window in
  the activation phase and one in repolarisation (dt 0.25 makes the sample
  times exact binary fractions, so an oracle window starting at 40 ms
  reproduces samples 160.. of the full run), by the rel-L2 over both windows
  (the full ECG's rel-L2 is dominated by the activation phase in the same
  way; the repolarisation window alone is ~80x smaller, so its own relative
  error reads ~80x larger for the same absolute error); the two served ECGs
  equal the plain call bit for bit.
* The precision worst case of the small hearts: 20^3, six transmural
  layers, 26-neighbour dipoles, 16 leads, dt 0.5, on every chunk size of
  the matrix-core kernel's LDS rows.
"""
import numpy as np
import pytest

from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5
N_REQ = 40
WINDOWS = [(40.0, 10.0), (300.0, 10.0)]   # (start, length) in ms


@pytest.fixture(scope="module")
def big():
    h = S.make_heart(256, seed=1)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    reqs = []
    for r in range(N_REQ):
        x = S.random_params(3, 12, 4000 + r)
        reqs.append((S.layer_k_from_params(x, 3),
                     S.move_leads(S.lead_positions(256, 12), S.displacements(x, 3, 12))))
    refs = {}
    for r in (0, N_REQ - 1):
        k, pos = reqs[r]
        for (st, ln) in WINDOWS:
            refs[(r, st)] = O.run(h.layer, act, k, pos, "3D4", st, ln, 0.25)
    return h, act, reqs, refs


def test_stress_256_dt025_admitted_waves(ctx, big):
    from ekgsim_amd import Heart, Server
    h, act, reqs, refs = big
    H = Heart(ctx, h.layer, h.seed)
    H.excite_delay(h.delay)
    assert np.array_equal(H.activation(), act)
    # device bytes of one request, then a budget of about 6 requests
    probe = Server(ctx, max_batch=1, n_streams=1)
    hid = probe.add_heart(H)
    probe.wait(probe.submit(hid, reqs[0][0], reqs[0][1], "3D4", 0.0, 600.0, 0.25))
    one = probe.stats()["peak_bytes"]
    probe.close()
    budget = int(6.5 * one)
    srv = Server(ctx, hbm_budget_bytes=budget, max_batch=64, n_streams=2)
    hid = srv.add_heart(H)
    tickets = [srv.submit(hid, k, pos, "3D4", 0.0, 600.0, 0.25) for k, pos in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    srv.close()
    print(f"256^3 dt 0.25: {N_REQ} requests in {st['groups']} groups, peak {st['peak_bytes'] / 2**30:.1f} GiB "
          f"of a {budget / 2**30:.1f} GiB budget")
    assert st["completed"] == N_REQ
    assert st["peak_bytes"] <= budget
    assert st["groups"] >= N_REQ // 6
    assert all(g.shape == (12, 2400) for g in got)
    for r in (0, N_REQ - 1):
        k, pos = reqs[r]
        assert np.array_equal(got[r], H.simulate(k, pos, "3D4", 0.0, 600.0, 0.25))
        num = den = 0.0
        for (st_ms, ln) in WINDOWS:
            j0 = int(st_ms / 0.25)
            ref = refs[(r, st_ms)]
            seg = got[r][:, j0:j0 + ref.shape[1]]
            num += float(np.sum((seg - ref) ** 2))
            den += float(np.sum(ref ** 2))
            print(f"request {r} window {st_ms:.0f} ms: rel-L2 {rel_l2(seg, ref):.3e} "
                  f"(max |err| {np.abs(seg - ref).max():.2e}, max |ref| {np.abs(ref).max():.2e})")
        err = (num / den) ** 0.5
        print(f"request {r}: rel-L2 over both windows {err:.3e}")
        assert err <= TOL
    H.close()


@pytest.mark.parametrize("chunk", [0, 64, 512])
def test_six_layers_26_neighbours_16_leads(ctx, chunk):
    h = S.make_heart(20, seed=5, n_layers=6)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    x = S.random_params(6, 16, 321)
    k = S.layer_k_from_params(x, 6)
    pos = S.move_leads(S.lead_positions(20, 16), S.displacements(x, 6, 16))
    from ekgsim_amd import Heart
    g = Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    g.set_tune(chunk)
    got = g.simulate(k, pos, "3D8", 0.0, 500.0, 0.5)
    ref = O.run(h.layer, act, k, pos, "3D8", 0.0, 500.0, 0.5)
    err = rel_l2(got, ref)
    print(f"20^3 six layers 3D8 16 leads dt 0.5, chunk {chunk}: rel-L2 {err:.3e}")
    assert err <= TOL
    g.set_tune()
    g.close()

"""Activation classes on the matrix-core path.

Voxels of one layer with one activation time have the same AP, so their
dipole terms factor through the sums of their lead weights: a run of at
least `class_min` such voxels is summed as a class.  The ECG must stay within
rel-L2 1e-5 of the FP64 oracle whatever the threshold:

* integer-delay hearts (every activation time an integer, so runs of
  thousands of voxels), with classes at the default threshold, with every
  voxel in a class (class_min 1: layers with no chunk outside the classes),
  with a low threshold that leaves small runs outside, and with classes off;
  on both 3-D neighbourhoods, 16 leads and six layers, and on a one-plane
  heart with both 2-D neighbourhoods;
* a heart from the FP64 transfer-matrix excitation, whose times are mostly
  distinct, gives the same result with classes on or off;
* a request gives the same bits alone, in a uniform batch and in a server
  group mixing a heart with classes and one without.
"""
import numpy as np
import pytest

from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _int_heart(ctx, n, seed, n_layers=3):
    from ekgsim_amd import Heart
    h = S.make_heart(n, seed=seed, n_layers=n_layers)
    g = Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    return h, g, act


def _req(n, L, s, n_layers=3):
    x = S.random_params(n_layers, L, s)
    return (S.layer_k_from_params(x, n_layers),
            S.move_leads(S.lead_positions(n, L), S.displacements(x, n_layers, L)))


def _class_runs(h, act):
    lay = h.layer.ravel()
    m = lay > 0
    _, cnt = np.unique(np.stack([lay[m].astype(np.float64), act.ravel()[m]], 1), axis=0, return_counts=True)
    return cnt


@pytest.mark.parametrize("nhood", ["3D4", "3D8"])
def test_integer_heart_every_threshold(ctx, nhood):
    h, g, act = _int_heart(ctx, 40, 9)
    runs = _class_runs(h, act)
    assert runs.max() >= 64 and runs.min() < 16   # large classes and small runs both occur
    k, pos = _req(40, 6, 31)
    ref = O.run(h.layer, act, k, pos, nhood, 0.0, 400.0, 1.0)
    out = {}
    for cm in (0, 1, 16, -1):
        g.set_tune(class_min=cm)
        out[cm] = g.simulate(k, pos, nhood, 0.0, 400.0, 1.0)
        err = rel_l2(out[cm], ref)
        print(f"{nhood} class_min {cm}: rel-L2 {err:.3e}")
        assert err <= TOL, cm
    assert g.get_tune()["class_min"] == -1
    g.set_tune()
    assert g.get_tune()["class_min"] == 0
    g.close()


@pytest.mark.parametrize("nhood", ["2D4", "2D8"])
def test_two_dimensional_heart(ctx, nhood):
    from ekgsim_amd import Heart
    h = S.make_heart(64, seed=2, dims=(1, 64, 64))
    if h.seed.sum() == 0:
        h.seed.ravel()[np.flatnonzero(h.layer.ravel())[0]] = 1
    g = Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    k, pos = _req(64, 4, 77)
    ref = O.run(h.layer, act, k, pos, nhood, 0.0, 300.0, 1.0)
    for cm in (0, 1, -1):
        g.set_tune(class_min=cm)
        err = rel_l2(g.simulate(k, pos, nhood, 0.0, 300.0, 1.0), ref)
        print(f"2-D {nhood} class_min {cm}: rel-L2 {err:.3e}")
        assert err <= TOL
    g.set_tune()
    g.close()


def _window_counts(h, act, class_min, window):
    """(class, cell) pairs with a non-zero 3D4 dipole vector and (window,
    cell) entries, restated on the grid with numpy: classes are the runs of
    one layer and one activation time with at least class_min voxels, in
    layer and time order, taken `window` at a time; for a class S the vector
    of cell w is D(w) = [w in S] s_w + sum over the corners d with w - d in S
    and w myocardial of d, s_w the sum of the corners d with w + d
    myocardial."""
    lay = h.layer
    myo = lay > 0
    keys = {}
    for l in range(1, int(lay.max()) + 1):
        times, cnt = np.unique(act[lay == l], return_counts=True)
        for t, c in zip(times, cnt):
            if c >= class_min:
                keys[(l, t)] = len(keys)
    cls = np.full(lay.shape, -1, np.int64)
    for (l, t), i in keys.items():
        cls[(lay == l) & (act == t)] = i
    P = np.pad(myo, 1)
    C = np.pad(cls, 1, constant_values=-1)
    n0, n1, n2 = lay.shape
    corners = [(a, b, c) for a in (-1, 1) for b in (-1, 1) for c in (-1, 1)]
    s_w = np.zeros(lay.shape + (3,), np.int64)
    for d in corners:
        nb = P[1 + d[0]:1 + d[0] + n0, 1 + d[1]:1 + d[1] + n1, 1 + d[2]:1 + d[2] + n2]
        s_w += nb[..., None] * np.array(d)
    pairs = 0
    per_window = {}
    for i in range(len(keys)):
        D = np.where((cls == i)[..., None], s_w, 0)
        for d in corners:
            src = C[1 - d[0]:1 - d[0] + n0, 1 - d[1]:1 - d[1] + n1, 1 - d[2]:1 - d[2] + n2] == i
            D += (src & myo)[..., None] * np.array(d)
        nz = np.any(D != 0, axis=-1)
        pairs += int(nz.sum())
        per_window.setdefault(i // window, np.zeros(lay.shape, bool))
        per_window[i // window] |= nz
    return pairs, int(sum(m.sum() for m in per_window.values()))


def test_layout_info(ctx):
    h, g, act = _int_heart(ctx, 40, 9)
    runs = _class_runs(h, act)
    info = g.layout_info("3D4")
    print(info)
    assert info["voxels"] == int((h.layer > 0).sum())
    assert info["class_voxels"] == int(runs[runs >= 64].sum())
    assert info["classes"] == int((runs >= 64).sum())
    assert 0 < info["class_cells"] <= 9 * info["class_voxels"]
    assert info["class_chunks"] >= info["classes"] and info["chunks"] > 0
    pairs, entries = _window_counts(h, act, 64, 6)
    assert info["class_cells"] == pairs
    assert info["class_entries"] == entries
    assert -(-info["classes"] // 6) <= info["class_pieces"] <= entries // 16 + info["classes"]
    g.set_tune(class_min=-1)
    off = g.layout_info("3D8")
    assert off["classes"] == 0 and off["class_voxels"] == 0 and off["class_cells"] == 0
    g.set_tune()
    g.close()


def test_six_layers_16_leads_short_step(ctx):
    h, g, act = _int_heart(ctx, 24, 5, n_layers=6)
    k, pos = _req(24, 16, 321, n_layers=6)
    ref = O.run(h.layer, act, k, pos, "3D8", 0.0, 300.0, 0.5)
    for cm in (0, 1):
        g.set_tune(class_min=cm)
        err = rel_l2(g.simulate(k, pos, "3D8", 0.0, 300.0, 0.5), ref)
        print(f"six layers 3D8 16 leads dt 0.5, class_min {cm}: rel-L2 {err:.3e}")
        assert err <= TOL
    g.set_tune()
    g.close()


def test_transfer_heart_same_with_or_without_classes(ctx):
    from ekgsim_amd import Heart
    h = S.make_heart(36, seed=4)
    T = S.layered_transfer(3, 12, 1.0, 1.5)
    act = O.excite_fp64(h.layer, h.seed, T)
    g = Heart(ctx, h.layer, h.seed)
    g.excite_transfer(T)
    k, pos = _req(36, 4, 77)
    ref = O.run(h.layer, act, k, pos, "3D4", 0.0, 300.0, 1.0)
    on = g.simulate(k, pos, "3D4", 0.0, 300.0, 1.0)
    g.set_tune(class_min=-1)
    off = g.simulate(k, pos, "3D4", 0.0, 300.0, 1.0)
    g.set_tune()
    assert rel_l2(on, ref) <= TOL and rel_l2(off, ref) <= TOL
    assert rel_l2(on, off) <= 1e-9
    g.close()


def test_same_bits_alone_in_batch_and_in_mixed_groups(ctx):
    from ekgsim_amd import Heart, Server
    h1, g1, act1 = _int_heart(ctx, 32, 2)
    h2 = S.make_heart(28, seed=6)
    T = S.layered_transfer(3, 12, 1.0, 1.5)
    g2 = Heart(ctx, h2.layer, h2.seed)
    g2.excite_transfer(T)
    reqs1 = [_req(32, 3, 500 + i) for i in range(5)]
    reqs2 = [_req(28, 3, 600 + i) for i in range(5)]
    alone1 = [g1.simulate(k, p, "3D4", 0.0, 250.0, 1.0) for k, p in reqs1]
    alone2 = [g2.simulate(k, p, "3D4", 0.0, 250.0, 1.0) for k, p in reqs2]
    # a uniform batch of one heart runs chunk-major with classes, two requests
    # per class-chunk block (five: the last is paired with itself); alone, one
    # request per block with four cell groups per iteration
    batch = g1.simulate_batch(np.stack([k for k, _ in reqs1]), np.stack([p for _, p in reqs1]),
                              "3D4", 0.0, 250.0, 1.0)
    for i in range(5):
        assert np.array_equal(batch[i], alone1[i])
    srv = Server(ctx, max_batch=10, n_streams=1)
    i1, i2 = srv.add_heart(g1), srv.add_heart(g2)
    tickets = []
    for a, b in zip(reqs1, reqs2):
        tickets.append(srv.submit(i1, a[0], a[1], "3D4", 0.0, 250.0, 1.0))
        tickets.append(srv.submit(i2, b[0], b[1], "3D4", 0.0, 250.0, 1.0))
    got = [srv.wait(t)[0] for t in tickets]
    srv.close()
    for i in range(5):
        assert np.array_equal(got[2 * i], alone1[i])
        assert np.array_equal(got[2 * i + 1], alone2[i])
    ref = O.run(h1.layer, act1, reqs1[0][0], reqs1[0][1], "3D4", 0.0, 250.0, 1.0)
    assert rel_l2(alone1[0], ref) <= TOL
    g1.close()
    g2.close()


def test_batch_of_eight_same_bits_as_alone(ctx):
    """A uniform batch whose size is a multiple of 8 takes the XCD-ordered
    class-kernel blocks (each request's blocks on one XCD) and the paired
    class-chunk blocks; every request equals its single call bitwise."""
    h, g, act = _int_heart(ctx, 24, 4)
    reqs = [_req(24, 5, 700 + i) for i in range(16)]
    batch = g.simulate_batch(np.stack([k for k, _ in reqs]), np.stack([p for _, p in reqs]),
                             "3D4", 0.0, 300.0, 1.0)
    for i, (k, p) in enumerate(reqs):
        assert np.array_equal(batch[i], g.simulate(k, p, "3D4", 0.0, 300.0, 1.0)), i
    ref = O.run(h.layer, act, reqs[9][0], reqs[9][1], "3D4", 0.0, 300.0, 1.0)
    assert rel_l2(batch[9], ref) <= TOL
    g.close()


@pytest.mark.parametrize("L", [2, 7, 16])
def test_batch_lead_packs_same_bits_as_alone(ctx, L):
    """Paired class-chunk blocks at 4, 8 and 16 lead slots (2, 7 and 16
    leads; an odd batch, so the last request is paired with itself) give each
    request the bits of its single call, within 1e-5 of the FP64 oracle."""
    h, g, act = _int_heart(ctx, 22, 8)
    reqs = [_req(22, L, 40 * L + i) for i in range(3)]
    batch = g.simulate_batch(np.stack([k for k, _ in reqs]), np.stack([p for _, p in reqs]),
                             "3D8", 0.0, 250.0, 1.0)
    for i, (k, p) in enumerate(reqs):
        assert np.array_equal(batch[i], g.simulate(k, p, "3D8", 0.0, 250.0, 1.0)), i
    ref = O.run(h.layer, act, reqs[2][0], reqs[2][1], "3D8", 0.0, 250.0, 1.0)
    assert rel_l2(batch[2], ref) <= TOL
    g.close()


def test_class_cells_above_dense_grid_limit(ctx):
    """Grids above 2^27 cells build the class cells through a hash map
    instead of the dense cell->slot array.  The same heart in a 40^3 grid and
    at the far corner of a 520^3 grid (electrodes moved with it) has the same
    activation times, the same classes and class cells, and the same ECG up
    to the FP32 rounding of the electrode offsets."""
    from ekgsim_amd import Heart
    n, N = 40, 520
    o = N - n - 1
    h = S.make_heart(n, seed=2)
    big = {}
    for name in ("layer", "seed", "delay"):
        a = np.zeros((N, N, N), np.uint8)
        a[o:o + n, o:o + n, o:o + n] = getattr(h, name)
        big[name] = a
    g1 = Heart(ctx, h.layer, h.seed)
    g1.excite_delay(h.delay)
    g2 = Heart(ctx, big["layer"], big["seed"])
    g2.excite_delay(big["delay"])
    a1 = g1.activation()
    a2 = g2.activation()[o:o + n, o:o + n, o:o + n]
    assert np.array_equal(a1, a2)
    i1, i2 = g1.layout_info("3D4"), g2.layout_info("3D4")
    assert i2["classes"] > 0
    for key in ("voxels", "class_voxels", "class_cells", "classes", "class_chunks"):
        assert i1[key] == i2[key], key
    k, pos = _req(n, 6, 99)
    e1 = g1.simulate(k, pos, "3D4", 0.0, 300.0, 1.0)
    e2 = g2.simulate(k, pos + o, "3D4", 0.0, 300.0, 1.0)
    err = rel_l2(e2, e1)
    print(f"40^3 heart in a 520^3 grid: rel-L2 {err:.3e} against the 40^3 grid")
    assert err <= 1e-6
    g1.close()
    g2.close()

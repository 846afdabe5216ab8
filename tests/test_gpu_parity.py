"""HIP path vs the FP64 oracle on the same seeded inputs.

Bars: activation times bit-exact (FP64 reference semantics and the integer
per-voxel-delay mode); every ECG within rel-L2 <= 1e-5 of the FP64 oracle
(the tolerance the simulation core is held to), plus max-error reporting.
"""
import glob
import os

import numpy as np
import pytest

from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import GOLDEN, load_golden, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _heart(ctx, h):
    from ekgsim_amd import Heart
    return Heart(ctx, h.layer, h.seed)


@pytest.mark.parametrize("n,seed", [(24, 3), (33, 7), (48, 1)])
def test_ca_fp64_bit_exact(ctx, n, seed):
    h = S.make_heart(n, seed=seed)
    T = S.layered_transfer(3, 12, 1.0, 1.5)
    ref = O.excite_fp64(h.layer, h.seed, T)
    g = _heart(ctx, h)
    g.excite_transfer(T)
    got = g.activation()
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_ca_fp64_irregular_weights_bit_exact(ctx):
    """Random conduction entries make path sums rounding-sensitive."""
    rng = np.random.default_rng(11)
    h = S.make_heart(40, seed=5)
    T = np.full((12, 12), -1.0)
    T[1:, 1:] = rng.uniform(0.1, 3.0, size=(11, 11))
    ref = O.excite_fp64(h.layer, h.seed, T)
    g = _heart(ctx, h)
    g.excite_transfer(T)
    assert np.array_equal(g.activation().view(np.uint64), ref.view(np.uint64))


def test_ca_2d_bit_exact(ctx):
    """A one-plane model floods over the 8-neighbourhood (simulator.cpp:251-254)."""
    h = S.make_heart(64, seed=2, dims=(1, 64, 64))
    if h.seed.sum() == 0:
        lay = h.layer.ravel()
        h.seed.ravel()[np.flatnonzero(lay)[0]] = 1
    T = S.layered_transfer(3, 12, 1.0, 2.0)
    ref = O.excite_fp64(h.layer, h.seed, T)
    g = _heart(ctx, h)
    g.excite_transfer(T)
    assert np.array_equal(g.activation(), ref)


@pytest.mark.parametrize("n", [48, 128])
def test_ca_int_bit_exact(ctx, n):
    h = S.make_heart(n, seed=1)
    ref = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    g = _heart(ctx, h)
    g.excite_delay(h.delay)
    assert np.array_equal(g.activation(), ref)


def _case(n, L, seed=1, pseed=1000):
    h = S.make_heart(n, seed=seed)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    x = S.random_params(3, L, pseed)
    k = S.layer_k_from_params(x, 3)
    pos = S.move_leads(S.lead_positions(n, L), S.displacements(x, 3, L))
    return h, act, k, pos


@pytest.mark.parametrize("nhood", ["3D4", "3D8", "2D8", "2D4"])
def test_ecg_parity_small(ctx, nhood):
    h, act, k, pos = _case(24, 2)
    ref = O.run(h.layer, act, k, pos, nhood, 0, 600, 1.0)
    g = _heart(ctx, h)
    g.set_activation(act)
    got = g.simulate(k, pos, nhood, 0, 600, 1.0)
    err = rel_l2(got, ref)
    print(f"24^3 {nhood}: rel-L2 {err:.3e} max|err| {np.abs(got - ref).max():.3e}")
    assert err <= TOL


@pytest.mark.parametrize("start,length,dt", [(100, 400, 1.0), (0, 600, 0.25), (0, 37, 1.0), (5, 64, 1.0)])
def test_ecg_parity_time_grids(ctx, start, length, dt):
    h, act, k, pos = _case(24, 3)
    ref = O.run(h.layer, act, k, pos, "3D4", start, length, dt)
    g = _heart(ctx, h)
    g.set_activation(act)
    got = g.simulate(k, pos, "3D4", start, length, dt)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) <= TOL


def test_ecg_parity_fp64_activation(ctx):
    """Non-integer activation times from the FP64 flood (sqrt(2), sqrt(3) lags)."""
    h = S.make_heart(32, seed=4)
    T = S.layered_transfer(3, 12, 1.0, 1.5)
    act = O.excite_fp64(h.layer, h.seed, T)
    x = S.random_params(3, 4, 77)
    k = S.layer_k_from_params(x, 3)
    pos = S.move_leads(S.lead_positions(32, 4), S.displacements(x, 3, 4))
    ref = O.run(h.layer, act, k, pos, "3D4", 0, 600, 1.0)
    g = _heart(ctx, h)
    g.excite_transfer(T)
    got = g.simulate(k, pos, "3D4", 0, 600, 1.0)
    assert rel_l2(got, ref) <= TOL


def test_ecg_parity_resting_offset(ctx):
    """k0 != 0 (a resting potential) enters through the per-chunk sum of G."""
    h, act, k, pos = _case(24, 2)
    k = k.copy()
    k[:, 0] = [-86.0, -84.0, -80.0]
    ref = O.run(h.layer, act, k, pos, "3D4", 0, 600, 1.0)
    g = _heart(ctx, h)
    g.set_activation(act)
    got = g.simulate(k, pos, "3D4", 0, 600, 1.0)
    assert rel_l2(got, ref) <= TOL


def test_ecg_parity_48_config(ctx):
    """48^3 heart, 2 leads, 600 ms at 1 ms, 3D4 (the CPU-reference case)."""
    h, act, k, pos = _case(48, 2)
    ref = O.run(h.layer, act, k, pos, "3D4", 0, 600, 1.0)
    g = _heart(ctx, h)
    g.excite_delay(h.delay)
    got = g.simulate(k, pos, "3D4", 0, 600, 1.0)
    err = rel_l2(got, ref)
    print(f"48^3: rel-L2 {err:.3e} max|err| {np.abs(got - ref).max():.3e}")
    assert err <= TOL


def test_ecg_parity_128_12_leads(ctx):
    """Full size: 128^3 heart, 12 leads, 600 ms at 1 ms."""
    h, act, k, pos = _case(128, 12)
    ref = O.run(h.layer, act, k, pos, "3D4", 0, 600, 1.0)
    g = _heart(ctx, h)
    g.excite_delay(h.delay)
    got = g.simulate(k, pos, "3D4", 0, 600, 1.0)
    errs = [rel_l2(got[i], ref[i]) for i in range(12)]
    print(f"128^3: rel-L2 {rel_l2(got, ref):.3e} worst lead {max(errs):.3e}")
    assert rel_l2(got, ref) <= TOL
    assert max(errs) <= TOL


def test_batch_equals_plain_bitwise(ctx):
    h = S.make_heart(40, seed=9)
    g = _heart(ctx, h)
    g.excite_delay(h.delay)
    ks, ps = [], []
    for b in range(5):
        x = S.random_params(3, 6, 500 + b)
        ks.append(S.layer_k_from_params(x, 3))
        ps.append(S.move_leads(S.lead_positions(40, 6), S.displacements(x, 3, 6)))
    batch = g.simulate_batch(np.stack(ks), np.stack(ps), "3D4", 0, 600, 1.0)
    for b in range(5):
        one = g.simulate(ks[b], ps[b], "3D4", 0, 600, 1.0)
        assert np.array_equal(batch[b], one)


def test_heart_without_seed_rejected(ctx):
    from ekgsim_amd import EkgError
    h = S.make_heart(16, seed=1)
    g = _heart(ctx, h)
    g2 = type(g)(ctx, h.layer, np.zeros_like(h.seed))
    with pytest.raises(EkgError):
        g2.excite_delay(h.delay)


def test_heart_destroyed_after_its_context(ctx):
    """A heart destroyed after its context frees its memory without touching
    the context (no stale device error reaches the next launch), and closing a
    Context closes its hearts first."""
    from ekgsim_amd import Context, Heart
    from ekgsim_amd._capi import lib
    h, act, k, pos = _case(16, 2)
    ref = O.run(h.layer, act, k, pos, "3D4", 0, 300, 1.0)
    c1 = Context(0)
    g1 = Heart(c1, h.layer, h.seed)
    g1.set_activation(act)
    g1.simulate(k, pos, "3D4", 0, 300, 1.0)
    raw, g1.h = g1.h, type(g1.h)()          # keep the handle past Context.close
    c1.close()
    lib().ekg_heart_destroy(raw)
    g2 = Heart(ctx, h.layer, h.seed)
    g2.set_activation(act)
    assert rel_l2(g2.simulate(k, pos, "3D4", 0, 300, 1.0), ref) <= TOL
    c3 = Context(0)
    g3 = Heart(c3, h.layer, h.seed)
    c3.close()
    assert not g3.h
    g2.close()


def test_simulate_before_excitation_rejected(ctx):
    from ekgsim_amd import EkgError
    h, act, k, pos = _case(16, 2)
    g = _heart(ctx, h)
    with pytest.raises(EkgError):
        g.simulate(k, pos)


# --- against the reference's own outputs (golden vectors) ---------------------


@pytest.mark.parametrize("name", ["h16", "h22_irregular", "h2d"])
def test_ca_matches_reference_golden_bitwise(ctx, name):
    g = load_golden(f"ref_ca_{name}.npz")
    from ekgsim_amd import Heart
    h = Heart(ctx, g["layer"], g["seed"])
    h.excite_transfer(g["transfer"])
    assert np.array_equal(h.activation().view(np.uint64), g["act"].view(np.uint64))


@pytest.mark.parametrize("name", sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "ref_run_*.npz"))))
def test_ecg_matches_reference_golden(ctx, name):
    g = load_golden(name)
    from ekgsim_amd import Heart
    h = Heart(ctx, g["layer"])
    h.set_activation(g["act"])
    got = h.simulate(g["layer_k"], g["pos"], str(g["nhood"]), float(g["start"]), float(g["length"]),
                     float(g["dt"]))
    assert got.shape == g["ecg"].shape
    assert rel_l2(got, g["ecg"]) <= TOL


def test_pipeline_matches_reference_golden(ctx):
    """Shape + conduction -> excitation on the GPU, file lead positions moved by
    the displacement genes through the C ABI, ECG on the GPU; against the
    reference EkgSim facade run on the same files."""
    import ekgsim_amd as E
    g = load_golden("ref_pipeline_16.npz")
    h = E.Heart(ctx, g["layer"], g["seed"])
    h.excite_transfer(g["transfer"])
    assert np.array_equal(h.activation(), g["act"])
    pos = E.move_leads(g["pos_file"], g["disp"])
    got = h.simulate(g["layer_k"], pos, "3D4", 0.0, 500.0, 1.0)
    assert rel_l2(got, g["ecg"]) <= TOL


@pytest.mark.parametrize("class_min", [0, -1])
def test_mixed_shapes_one_group_equal_alone(ctx, class_min):
    """Requests with different lead counts, durations, start times and time
    steps in one launch group equal the single-request calls bitwise and
    the FP64 oracle within 1e-5; with activation classes (default threshold)
    and without them (every voxel in a chunk, so every request's ECG comes
    from the partial rows of the reduction, with lead counts that are not a
    multiple of 4 and step counts that are not a multiple of the 8-step
    bound tile)."""
    h = S.make_heart(28, seed=3)
    g = _heart(ctx, h)
    g.excite_delay(h.delay)
    g.set_tune(class_min=class_min)
    act = O.excite_int(h.layer, h.seed, h.delay).astype(np.float64)
    rng = np.random.default_rng(5)
    reqs = []
    for r, (L, start, length, dt) in enumerate([(2, 0.0, 600.0, 1.0), (7, 50.0, 250.0, 0.5),
                                                (12, 0.0, 130.0, 0.25), (3, 10.0, 700.0, 1.0),
                                                (16, 0.0, 64.0, 1.0)]):
        x = S.random_params(3, L, 900 + r)
        reqs.append((S.layer_k_from_params(x, 3),
                     S.move_leads(S.lead_positions(28, L), S.displacements(x, 3, L)), start, length, dt))
    many = g.simulate_many(reqs)
    for (k, pos, start, length, dt), got in zip(reqs, many):
        alone = g.simulate(k, pos, "3D4", start, length, dt)
        assert np.array_equal(got, alone)
        ref = O.run(h.layer, act, k, pos, "3D4", start, length, dt)
        assert rel_l2(got, ref) <= TOL


# --- degenerate sizes ---------------------------------------------------------

@pytest.mark.parametrize("length,dt", [(0.0, 1.0), (0.5, 1.0), (1.0, 1.0), (3.0, 1.0)])
def test_ecg_parity_zero_one_and_few_samples(ctx, length, dt):
    """Durations of zero, one and a few samples: the same number of samples
    as the reference's rule and the oracle's values; one lead."""
    h, act, k, pos = _case(20, 4)
    pos = pos[:1]
    ref = O.run(h.layer, act, k, pos, "3D4", 0.0, length, dt)
    g = _heart(ctx, h)
    g.set_activation(act)
    got = g.simulate(k, pos, "3D4", 0.0, length, dt)
    assert got.shape == ref.shape == (1, O.num_steps(length, dt))
    if ref.size:
        assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("nhood", ["3D4", "3D8"])
def test_two_voxel_and_single_layer_hearts(ctx, nhood):
    """A heart of two voxels (one dipole; at class threshold 1 each voxel is
    a class of its own) and a one-layer heart, with activation classes on
    and off."""
    from ekgsim_amd import Heart
    lay = np.zeros((5, 5, 5), np.uint8)
    lay[2, 2, 2] = lay[3, 3, 3] = 1  # corner neighbours: a dipole in both 3-D neighbourhoods
    seed = np.zeros_like(lay)
    seed[2, 2, 2] = 1
    act = O.excite_int(lay, seed, np.ones_like(lay)).astype(np.float64)
    x = S.random_params(3, 2, 11)
    k = S.layer_k_from_params(x, 3)
    pos = np.array([[2.0, 9.5, 4.0], [-3.0, 1.5, 7.25]])
    ref = O.run(lay, act, k, pos, nhood, 0.0, 300.0, 1.0)
    assert np.abs(ref).max() > 0
    g = Heart(ctx, lay, seed)
    g.excite_delay(np.ones_like(lay))
    for cm in (0, 1, -1):
        g.set_tune(class_min=cm)
        got = g.simulate(k, pos, nhood, 0.0, 300.0, 1.0)
        assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max()), cm
    g.close()
    h1 = S.make_heart(20, seed=6, n_layers=1)
    act1 = O.excite_int(h1.layer, h1.seed, h1.delay).astype(np.float64)
    k1 = S.layer_k_from_params(S.random_params(1, 3, 12), 1)
    pos1 = S.lead_positions(20, 3)
    ref1 = O.run(h1.layer, act1, k1, pos1, nhood, 0.0, 300.0, 1.0)
    g1 = Heart(ctx, h1.layer, h1.seed)
    g1.excite_delay(h1.delay)
    assert rel_l2(g1.simulate(k1, pos1, nhood, 0.0, 300.0, 1.0), ref1) <= TOL
    g1.close()


@pytest.mark.parametrize("nhood", ["3D4", "3D8"])
@pytest.mark.parametrize("edge", ["slow_fast_noskip", "flat_sigmoid"])
def test_edge_layer_coefficients(ctx, nhood, edge):
    """Layer coefficients at the edges of the AP paths' shortcuts, classes on
    and off: a slow upstroke (k1 0.3: windows of ~170 steps, many transition
    tiles), k4 + k5 > k1/4 (no voxel is skipped before its upstroke) and a
    one-step upstroke (k1 40); or k1 = 0 (S = 1/2 for every voxel at every
    step: no tile is ever skipped or saturated)."""
    h, act, k, pos = _case(24, 3, seed=5, pseed=77)
    k = k.copy()
    if edge == "slow_fast_noskip":
        k[0, 1] = 0.3
        k[1, 4] = k[1, 1]
        k[2, 1] = 40.0
    else:
        k[:, 1] = 0.0
    ref = O.run(h.layer, act, k, pos, nhood, 0.0, 600.0, 1.0)
    g = _heart(ctx, h)
    g.set_activation(act)
    for cm in (0, -1):
        g.set_tune(class_min=cm)
        err = rel_l2(g.simulate(k, pos, nhood, 0.0, 600.0, 1.0), ref)
        print(f"{edge} {nhood} class_min {cm}: rel-L2 {err:.3e}")
        assert err <= TOL, cm
    g.close()

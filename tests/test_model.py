"""Host FP64 model functions of the library (no GPU): parameter vector ->
layer APs, ECG comparison, target resampling.

Pinned bit for bit against the reference's own application code
(tests/golden/ref_layer_aps.npz, ref_compare.npz, ref_resample.npz written by
make_golden.py from oracle/_ref/libekgrefapp.so), against the live reference
where it is built, and against closed forms (Pearson r by numpy, RMS,
cosine similarity, the 3-layer restatement in the oracle).
"""
import numpy as np
import pytest

from ekgsim_amd import _capi as A
from ekgsim_amd import synth as S
from oracle import oracle as O
from tests.conftest import load_golden

BASE = np.array(S.BASE_AP, float)
KMIN = np.array([0, 1.5, 100, 0.85, 0.05, 0.0003, 0.01, 0.01, 300.0])
KMAX = np.array([0, 3.5, 100, 0.95, 0.2, 0.0010, 0.10, 0.10, 450.0])
FREE = [5, 6, 7, 8]


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_layer_aps_match_reference_golden_bitwise():
    g = load_golden("ref_layer_aps.npz")
    row = 0
    for interp, nl, mid, genes in zip(g["interp"], g["n_layers"], g["mid"], g["genes"]):
        ng = 12 if interp == 2 else 8
        got, viol = A.layer_aps(np.stack([BASE] * 3), KMIN, KMAX, FREE, int(interp), float(mid), int(nl),
                                genes[:ng])
        ref = g["layer_k"][row:row + nl]
        row += nl
        assert np.array_equal(_bits(got), _bits(ref)), (interp, nl, mid)
    assert row == g["layer_k"].shape[0]


@pytest.mark.skipif(not O.have_refapp(), reason="reference application harness not built")
@pytest.mark.parametrize("interp", [1, 2])
def test_layer_aps_match_live_reference(interp):
    rng = np.random.default_rng(40 + interp)
    for _ in range(8):
        nl = int(rng.integers(3, 26))
        mid = float(rng.choice([0.6, 0.25, 0.8]))
        g = rng.uniform([3e-4, 0.01, 0.01, 250] * 3, [1e-3, 0.1, 0.1, 450] * 3)
        got, _ = A.layer_aps(np.stack([BASE] * 3), KMIN, KMAX, FREE, interp, mid, nl, g)
        fk, mk, bk = BASE.copy(), BASE.copy(), BASE.copy()
        fk[5:9], mk[5:9] = g[0:4], g[4:8]
        bk[5:9] = g[8:12] if interp == 2 else g[4:8]
        ref = O.ref_interp_layers(interp, nl, mid, fk, mk, bk)
        assert np.array_equal(_bits(got), _bits(ref))


def test_three_layers_equal_oracle_and_violation():
    """Three layers with mid at 0.6 use the vector directly (no matching):
    compare with the oracle's restatement, including the violation sum."""
    x = np.array([2e-4, 0.05, 0.05, 320, 6e-4, 0.2, 0.05, 500, 9e-4, 0.06, 0.07, 380])
    got, viol = A.layer_aps(np.stack([BASE] * 3), KMIN, KMAX, FREE, 2, 0.6, 3, x)
    ref, rviol = O.layer_aps_sim3(BASE, BASE, FREE, 0.6, KMIN, KMAX, x)
    assert np.array_equal(got, ref)
    # out of bounds: k5 = 2e-4 (< 3e-4), k6 = 0.2 (> 0.1), k8 = 500 (> 450)
    assert viol == pytest.approx((3e-4 - 2e-4) + (0.2 - 0.1) + (500 - 450), rel=1e-12)
    assert viol == rviol


def test_border_layers_and_blend_invariants():
    """Endo and epi rows carry the vector's values; with equal border APs
    every matched layer stays equal to them (zero misfit: no descent step)."""
    g = np.array([5e-4, 0.05, 0.06, 350] * 2)
    out, _ = A.layer_aps(np.stack([BASE] * 2), KMIN, KMAX, FREE, 1, 0.5, 9, g)
    assert np.all(out == out[0])
    assert np.array_equal(out[0, 5:9], g[:4]) and np.array_equal(out[0, :5], BASE[:5])


def test_num_genes_counts_distinct_free_coefficients():
    assert A.ap_num_genes([5, 6, 7, 8], "endo-mid-epi") == 12
    assert A.ap_num_genes([5, 6, 7, 8], "endo-epi") == 8
    assert A.ap_num_genes([5, 5, 6], "endo-epi") == 4


def test_short_vector_rejected():
    with pytest.raises(A.EkgError):
        A.layer_aps(np.stack([BASE] * 3), KMIN, KMAX, FREE, 2, 0.6, 3, np.zeros(11))


def test_criterion_matches_reference_golden_bitwise():
    g = load_golden("ref_compare.npz")
    for a, b, (mode, na, nb), off, val in zip(g["a"], g["b"], g["shape"], g["offset"], g["value"]):
        got = A.criterion(int(mode), a[:na], b[:nb], off)
        assert np.array_equal(_bits(got), _bits(val)) or (np.isnan(got) and np.isnan(val)), (mode, na, nb)


def test_criterion_closed_forms():
    rng = np.random.default_rng(3)
    a = rng.normal(size=500)
    b = 0.3 * a + rng.normal(size=500)
    n = 420
    r = np.corrcoef(a[:n], b[:n])[0, 1]
    assert A.criterion("correlation", a[:n], b) == pytest.approx(1 - r, rel=1e-12)
    assert A.criterion("rms", a, b) == pytest.approx(np.sqrt(np.mean((a - b) ** 2)), rel=1e-12)
    cos = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
    assert A.criterion("vector-correlation", a, b) == pytest.approx(1 - cos, rel=1e-12)
    # perfect correlation gives 0 (which the fitness maps to 2e10)
    assert abs(A.criterion("correlation", a, 2 * a + 1)) < 1e-15
    # deviation from linear: a curve equal to the target after range scaling
    # of a and the offset has a constant ratio, zero variance -> 1 / 1e-30
    t = np.linspace(0, 1, 50)
    assert A.criterion("dev-from-linear", t, t, 1.0) == 1 / 1e-30


def test_resample_matches_reference_golden_bitwise():
    g = load_golden("ref_resample.npz")
    for i in range(6):
        f, st = g[f"factor_{i}"]
        out = A.resample(g["src"], f, int(st))
        assert np.array_equal(_bits(out), _bits(g[f"out_{i}"]))


def test_resample_factor_one_ignores_start():
    src = np.arange(10.0)
    assert np.array_equal(A.resample(src, 1.0, 4), src)
    assert np.allclose(A.resample(src, 2.0, 0), np.r_[np.arange(0, 9.5, 0.5)])


def test_ap_value_matches_reference_golden():
    g = load_golden("ref_wohlfart.npz")
    for k, ts, vs in zip(g["k"], g["t"], g["v"]):
        got = np.array([A.ap_value(k, t) for t in ts])
        assert np.array_equal(_bits(got), _bits(vs))


@pytest.mark.skipif(not O.have_refapp(), reason="reference application harness not built")
def test_apd90_matches_live_reference_bitwise():
    """ekg_ap_apd90 (the facade's WohlfartPlus::apd90) against the compiled
    reference's WohlfartPlus::apd90 over the free-coefficient bounds."""
    rng = np.random.default_rng(90)
    for _ in range(40):
        k = BASE.copy()
        k[5:9] = rng.uniform([3e-4, 0.01, 0.01, 200], [1e-3, 0.1, 0.1, 450])
        assert np.array_equal(_bits([A.ap_apd90(k)]), _bits([O.ref_apd90(k)])), k


def test_apd90_special_cases():
    # a flat curve (k2 = 0) never falls through k0 + k2/10: -1
    k = BASE.copy()
    k[2] = 0.0
    assert A.ap_apd90(k) == -1.0
    # the crossing lies between the last sample above and the first at or
    # below k0 + k2/10, and later repolarisation (larger k8) moves it later
    k1, k2 = BASE.copy(), BASE.copy()
    k2[8] += 50.0
    a1, a2 = A.ap_apd90(k1), A.ap_apd90(k2)
    t = np.floor(a1)
    target = k1[0] + 0.1 * k1[2]
    assert A.ap_value(k1, t) > target >= A.ap_value(k1, t + 1)
    assert a2 > a1 + 40.0

"""The seeded synthetic workload generator."""
import numpy as np

from ekgsim_amd import synth as S


def test_myocardial_voxel_counts():
    # grid counts of the ellipsoid shells with centre (n-1)/2
    assert S.make_heart(48, seed=1).n_myo == 20136
    assert S.make_heart(128, seed=1).n_myo == 370656


def test_layers_seeds_delays():
    h = S.make_heart(48, seed=1)
    assert set(np.unique(h.layer)) == {0, 1, 2, 3}
    assert h.seed.sum() == 5 and np.all(h.layer[h.seed > 0] == 1)
    d = h.delay[h.layer > 0]
    assert d.min() == 1 and d.max() == 3 and not h.delay[h.layer == 0].any()
    # deterministic per seed
    assert np.array_equal(h.delay, S.make_heart(48, seed=1).delay)
    assert not np.array_equal(h.delay, S.make_heart(48, seed=2).delay)


def test_leads_on_sphere():
    pos = S.lead_positions(128, 12)
    c = np.full(3, 63.5)
    assert np.allclose(np.linalg.norm(pos - c, axis=1), 96.0)


def test_parameter_vector_layout():
    x = S.random_params(3, 12, 1000)
    assert x.size == 12 + 24
    k = S.layer_k_from_params(x, 3)
    for l in range(3):
        assert np.array_equal(k[l, 5:9], x[4 * l:4 * l + 4])
        assert np.all((k[l, 5:9] >= S.FREE_LO) & (k[l, 5:9] <= S.FREE_HI))
    d = S.displacements(x, 3, 12)
    assert d.shape == (12, 2) and np.all(np.abs(d) <= 50)


def test_matrix_writer_roundtrip(tmp_path):
    h = S.make_heart(12, seed=4)
    p = tmp_path / "h.matrix"
    S.write_matrix(p, h.layer, h.seed)
    lines = p.read_text().splitlines()
    assert lines[1].split() == ["3d", "12", "x", "12", "x", "12", "space"]
    vals = np.array([int(v) for ln in lines[2:] for v in ln.split()])
    assert vals.size == h.layer.size
    assert np.array_equal(np.abs(vals).reshape(h.layer.shape), h.layer)
    assert np.array_equal((vals < 0).reshape(h.layer.shape), h.seed > 0)

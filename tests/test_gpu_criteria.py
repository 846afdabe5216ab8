"""Criteria on the device, after the ECG reduction.

The per-lead part of the reference's fitness (sim.cpp:600-702): the
comparison value of each simulated lead against its target in the four
comparison modes (vectorMath.h:277-401), the peak distance when peaks are
criteria, and the 2e10 correction of degenerate values (sim.cpp:679-685).
The device runs the host restatement's loops one for one in FP64 without
contraction, so its values must equal the host's ekg_criterion bit for bit:

* on simulated batches (ekg_simulate_criteria, ECGs kept on the device),
  with targets longer and shorter than the ECG;
* on device ECGs given directly (ekg_criteria_device), including constant
  leads whose correlation is 0/0 (2e10, or NaN when raw output is asked for),
  on short and long rows, and with the cached target statistics reused
  across calls and recomputed when the targets or the ECG length change.
"""
import numpy as np
import pytest

from ekgsim_amd import synth as S

pytestmark = pytest.mark.gpu
MODES = ["rms", "correlation", "dev-from-linear", "vector-correlation"]


def _targets(n, length, seed):
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, length)
    out, offs = [], []
    for i in range(n):
        c = np.sin(2 * np.pi * (t * rng.uniform(1, 3) + rng.uniform())) + 0.3 * rng.standard_normal(length)
        lo, hi = c.min(), c.max()
        out.append(c / (hi - lo))
        offs.append(1.0 - lo / (hi - lo))
    return np.array(out), np.array(offs)


def _host(E, mode, ecg, tg, offs, peaks, raw):
    B, L, _ = ecg.shape
    n = min(L, tg.shape[0])
    out = np.zeros((B, n * (2 if peaks else 1)))
    for r in range(B):
        for i in range(n):
            v = E._capi.criterion(mode, ecg[r, i], tg[i], offs[i])
            if not raw and (v < 0 or not np.isfinite(v) or not abs(v) >= np.finfo(float).tiny):
                v = 2e10
            if peaks:
                out[r, 2 * i] = v
                out[r, 2 * i + 1] = abs(int(np.argmax(ecg[r, i])) - int(np.argmax(tg[i])))
            else:
                out[r, i] = v
    return out


@pytest.mark.parametrize("tlen", [300, 250, 340])
def test_simulate_criteria_equals_host(ctx, tlen):
    import ekgsim_amd as E
    h = S.make_heart(24, seed=3)
    g = E.Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    B, L = 5, 4
    ks, ps = [], []
    for r in range(B):
        x = S.random_params(3, L, 900 + r)
        ks.append(S.layer_k_from_params(x, 3))
        ps.append(S.move_leads(S.lead_positions(24, L), S.displacements(x, 3, L)))
    ks, ps = np.stack(ks), np.stack(ps)
    ecg = g.simulate_batch(ks, ps, "3D4", 0.0, 300.0, 1.0)
    tg, offs = _targets(3, tlen, tlen)
    for mode in MODES:
        for peaks in (False, True):
            for raw in (False, True):
                got = g.simulate_criteria(ks, ps, mode, tg, offs, "3D4", 0.0, 300.0, 1.0, peaks=peaks, raw=raw)
                ref = _host(E, mode, ecg, tg, offs, peaks, raw)
                assert got.shape == ref.shape
                assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (mode, peaks, raw)
    g.close()


@pytest.mark.parametrize("T", [200, 6000])
def test_device_criteria_degenerate_leads(ctx, T):
    import torch
    import ekgsim_amd as E
    rng = np.random.default_rng(5)
    B, L = 3, 3
    ecg = rng.standard_normal((B, L, T))
    ecg[1, 0, :] = 0.25            # constant: correlation 0/0, RMS still finite
    ecg[2, 2, :] = ecg[2, 2, 0]    # constant
    tg, offs = _targets(3, T, 7)
    tg[1, :] = 0.5                 # constant target
    dev = torch.device("cuda", 0)
    d_e = torch.from_numpy(ecg).to(dev)
    for mode in MODES:
        for peaks in (False, True):
            for raw in (False, True):
                per = L * (2 if peaks else 1)
                d_o = torch.full((B, per), -1.0, dtype=torch.float64, device=dev)
                torch.cuda.synchronize()   # torch's fill before the library's stream writes
                ctx.criteria_device(mode, tg, offs, d_e.data_ptr(), B, L, T, d_o.data_ptr(), peaks=peaks, raw=raw)
                ctx.synchronize()
                got = d_o.cpu().numpy()
                ref = _host(E, mode, ecg, tg, offs, peaks, raw)
                same = (got.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(got) & np.isnan(ref))
                assert same.all(), (mode, peaks, raw, got, ref)
                if not raw:
                    assert np.isfinite(got).all()

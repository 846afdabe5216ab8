"""The allocation-free batch entry point (ekg_simulate_batch_device_noalloc)
against ekg_simulate_batch_device: bitwise the same ECGs on two calls in a
row, the second with a smaller batch, on both contraction paths; a heart
whose layout is not built and a batch larger than the static workspace are
refused before any kernel runs."""
import numpy as np
import pytest

import ekgsim_amd as E
from ekgsim_amd import synth as S

pytestmark = pytest.mark.gpu


def _requests(n, L, B, s0):
    ks, ps = [], []
    for b in range(B):
        x = S.random_params(3, L, s0 + b)
        ks.append(S.layer_k_from_params(x, 3))
        ps.append(E.move_leads(S.lead_positions(n, L), S.displacements(x, 3, L)))
    return np.stack(ks), np.stack(ps)


@pytest.mark.parametrize("kernel,class_min", [("mfma", 0), ("mfma", -1), ("valu", 0)])
def test_noalloc_equals_device_batch(kernel, class_min):
    import torch
    dev = torch.device("cuda", 0)
    ctx = E.Context(0)
    ctx.set_kernel(kernel)
    n, L, T = 32, 6, 400
    h = S.make_heart(n, seed=8)
    g = E.Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    g.set_tune(class_min=class_min)
    g.layout_info("3D4")  # builds the layout the allocation-free call needs
    for call, B in enumerate((9, 4)):
        print(f"noalloc {kernel} class_min {class_min} call {call} batch {B}", flush=True)
        k, p = _requests(n, L, B, 500 + 17 * call)
        d_k = torch.from_numpy(k).to(dev)
        d_p = torch.from_numpy(p).to(dev)
        ref = torch.zeros((B, L, T), dtype=torch.float64, device=dev)
        got = torch.full((B, L, T), float("nan"), dtype=torch.float64, device=dev)
        g.simulate_batch_device(d_k.data_ptr(), d_p.data_ptr(), ref.data_ptr(), B, 3, L, "3D4", 0.0, float(T), 1.0)
        ctx.synchronize()
        g.simulate_batch_device_noalloc(d_k.data_ptr(), d_p.data_ptr(), got.data_ptr(), B, 3, L, "3D4", 0.0,
                                        float(T), 1.0)
        ctx.synchronize()
        a, b = got.cpu().numpy(), ref.cpu().numpy()
        assert np.all(np.isfinite(a))
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), call
    g.close()
    ctx.close()


def test_noalloc_refuses_unbuilt_layout_and_oversized_batch():
    import torch
    dev = torch.device("cuda", 0)
    ctx = E.Context(0)
    n, L, T = 24, 4, 300
    h = S.make_heart(n, seed=9)
    g = E.Heart(ctx, h.layer, h.seed)
    g.excite_delay(h.delay)
    k, p = _requests(n, L, 2, 700)
    d_k = torch.from_numpy(k).to(dev)
    d_p = torch.from_numpy(p).to(dev)
    out = torch.zeros((2, L, T), dtype=torch.float64, device=dev)
    with pytest.raises(E.EkgError) as ei:
        g.simulate_batch_device_noalloc(d_k.data_ptr(), d_p.data_ptr(), out.data_ptr(), 2, 3, L, "3D4", 0.0,
                                        float(T), 1.0)
    assert ei.value.code == 6  # EKG_ESTATE: layout not built
    g.layout_info("3D4")
    # far more requests than the static workspace holds: refused (the
    # pointers are never dereferenced because the buffers are carved first)
    big = 1 << 20
    with pytest.raises(E.EkgError) as ei:
        g.simulate_batch_device_noalloc(d_k.data_ptr(), d_p.data_ptr(), out.data_ptr(), big, 3, L, "3D4", 0.0,
                                        float(T), 1.0)
    assert ei.value.code == 2  # EKG_ENOMEM
    # the entry point still works after a refusal
    g.simulate_batch_device_noalloc(d_k.data_ptr(), d_p.data_ptr(), out.data_ptr(), 2, 3, L, "3D4", 0.0,
                                    float(T), 1.0)
    ctx.synchronize()
    assert np.all(np.isfinite(out.cpu().numpy()))
    g.close()
    ctx.close()

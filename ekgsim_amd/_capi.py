"""ctypes binding of include/ekg.h (lib/libekgsim.so).

The library is loaded from the package's lib/ folder.  A missing library is
an error: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libekgsim.so")

EKG_OK = 0
ERRORS = {
    1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "ENOSEED", 5: "ETRANSFER", 6: "ESTATE",
    7: "ENOTFOUND", 8: "EIO",
}
NHOOD = {"2D4": 0, "2D8": 1, "3D4": 2, "3D8": 3, "cube": 3}
MAX_LEADS = 16

# Every symbol include/ekg.h declares.
EXPORTS = (
    "ekg_last_error", "ekg_version", "ekg_ctx_create", "ekg_ctx_destroy", "ekg_ctx_set_profiling",
    "ekg_ctx_kernel_time", "ekg_ctx_reset_profiling", "ekg_heart_create", "ekg_heart_destroy",
    "ekg_heart_num_myocardial", "ekg_heart_max_layer", "ekg_heart_excite_transfer",
    "ekg_heart_excite_delay", "ekg_heart_set_activation", "ekg_heart_get_activation",
    "ekg_num_steps", "ekg_simulate", "ekg_simulate_batch", "ekg_lead_basis", "ekg_move_lead",
    "ekg_server_create", "ekg_server_destroy", "ekg_server_add_heart", "ekg_server_submit",
    "ekg_server_poll", "ekg_server_wait", "ekg_server_wait_view", "ekg_server_release", "ekg_server_stats",
    "ekg_simulate_batch_device",
    "ekg_ctx_synchronize", "ekg_ctx_set_kernel", "ekg_ap_value", "ekg_ap_num_genes", "ekg_layer_aps",
    "ekg_criterion", "ekg_resample", "ekg_heart_autotune", "ekg_heart_set_tune", "ekg_heart_get_tune",
    "ekg_ctx_set_graphs", "ekg_ctx_graph_stats", "ekg_heart_layout_info", "ekg_criteria_device",
    "ekg_simulate_criteria", "ekg_ap_apd90", "ekg_simulate_batch_device_noalloc",
)
KERNELS = {"auto": 0, "valu": 1, "mfma": 2}


class EkgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class SimDesc(C.Structure):
    _fields_ = [
        ("n_layers", C.c_int),
        ("layer_k", C.POINTER(C.c_double)),
        ("n_leads", C.c_int),
        ("lead_pos", C.POINTER(C.c_double)),
        ("nhood", C.c_int),
        ("start", C.c_double),
        ("length", C.c_double),
        ("dt", C.c_double),
    ]


class Tune(C.Structure):
    _fields_ = [
        ("chunk_voxels", C.c_int),
        ("mfma_waves", C.c_int),
        ("lead_phase", C.c_int),
        ("kernel", C.c_int),
        ("class_min", C.c_int),
        ("ms_per_batch", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LayoutInfo(C.Structure):
    _fields_ = [
        ("voxels", C.c_int64),
        ("class_voxels", C.c_int64),
        ("class_cells", C.c_int64),
        ("chunks", C.c_int),
        ("class_chunks", C.c_int),
        ("classes", C.c_int),
        ("class_pieces", C.c_int),
        ("class_entries", C.c_int64),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_inflight", C.c_int),
        ("max_batch", C.c_int),
        ("n_streams", C.c_int),
        ("batch_window_us", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m ekgsim_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    pd = C.POINTER(C.c_double)
    L.ekg_last_error.restype = C.c_char_p
    L.ekg_version.restype = C.c_char_p
    L.ekg_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.ekg_ctx_destroy.argtypes = [vp]
    L.ekg_ctx_destroy.restype = None
    L.ekg_ctx_set_profiling.argtypes = [vp, C.c_int]
    L.ekg_ctx_kernel_time.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64), pd]
    L.ekg_ctx_reset_profiling.argtypes = [vp]
    L.ekg_heart_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.POINTER(vp)]
    L.ekg_heart_destroy.argtypes = [vp]
    L.ekg_heart_destroy.restype = None
    L.ekg_heart_num_myocardial.argtypes = [vp, C.POINTER(C.c_int64)]
    L.ekg_heart_max_layer.argtypes = [vp, C.POINTER(C.c_int)]
    L.ekg_heart_excite_transfer.argtypes = [vp, vp, C.c_int, C.c_int]
    L.ekg_heart_excite_delay.argtypes = [vp, vp]
    L.ekg_heart_set_activation.argtypes = [vp, vp]
    L.ekg_heart_get_activation.argtypes = [vp, vp]
    L.ekg_num_steps.argtypes = [C.c_double, C.c_double]
    L.ekg_num_steps.restype = C.c_int64
    L.ekg_simulate.argtypes = [vp, C.POINTER(SimDesc), vp]
    L.ekg_simulate_batch.argtypes = [vp, C.c_int, C.POINTER(SimDesc), vp]
    L.ekg_simulate_batch_device.argtypes = [vp, C.POINTER(BatchDev)]
    L.ekg_simulate_batch_device_noalloc.argtypes = [vp, C.POINTER(BatchDev)]
    L.ekg_ctx_synchronize.argtypes = [vp]
    L.ekg_ctx_set_kernel.argtypes = [vp, C.c_int]
    L.ekg_lead_basis.argtypes = [vp, vp, vp]
    L.ekg_lead_basis.restype = None
    L.ekg_move_lead.argtypes = [vp, C.c_double, C.c_double, vp]
    L.ekg_move_lead.restype = None
    L.ekg_server_create.argtypes = [vp, C.POINTER(ServerOpts), C.POINTER(vp)]
    L.ekg_server_destroy.argtypes = [vp]
    L.ekg_server_destroy.restype = None
    L.ekg_server_add_heart.argtypes = [vp, vp, C.POINTER(C.c_int)]
    L.ekg_server_submit.argtypes = [vp, C.c_int, C.POINTER(SimDesc), C.POINTER(C.c_uint64)]
    L.ekg_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(C.c_int)]
    L.ekg_server_wait.argtypes = [vp, C.c_uint64, vp, C.c_size_t, pd]
    L.ekg_server_wait_view.argtypes = [vp, C.c_uint64, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_size_t), pd]
    L.ekg_server_release.argtypes = [vp, C.c_uint64]
    L.ekg_server_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
    L.ekg_ap_value.argtypes = [vp, C.c_double]
    L.ekg_ap_value.restype = C.c_double
    L.ekg_ap_apd90.argtypes = [vp]
    L.ekg_ap_apd90.restype = C.c_double
    L.ekg_ap_num_genes.argtypes = [C.POINTER(ApModel), C.POINTER(C.c_int)]
    L.ekg_layer_aps.argtypes = [C.POINTER(ApModel), C.c_int, vp, C.c_int, vp, pd]
    L.ekg_criterion.argtypes = [C.c_int, vp, C.c_int64, vp, C.c_int64, C.c_double, pd]
    L.ekg_heart_autotune.argtypes = [vp, C.POINTER(SimDesc), C.c_int, C.POINTER(Tune)]
    L.ekg_heart_set_tune.argtypes = [vp, C.POINTER(Tune)]
    L.ekg_heart_get_tune.argtypes = [vp, C.POINTER(Tune)]
    L.ekg_heart_layout_info.argtypes = [vp, C.c_int, C.POINTER(LayoutInfo)]
    L.ekg_criteria_device.argtypes = [vp, C.POINTER(CritDesc), vp, C.c_int, C.c_int, C.c_int64, vp]
    L.ekg_simulate_criteria.argtypes = [vp, C.c_int, C.POINTER(SimDesc), C.POINTER(CritDesc), vp]
    L.ekg_ctx_set_graphs.argtypes = [vp, C.c_int]
    L.ekg_ctx_graph_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.ekg_resample.argtypes = [vp, C.c_int64, C.c_double, C.c_int, vp, C.c_int64, C.POINTER(C.c_int64)]
    _lib = L
    return L


def check(rc):
    if rc != EKG_OK:
        raise EkgError(rc, lib().ekg_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def num_steps(length, dt):
    return int(lib().ekg_num_steps(float(length), float(dt)))


def lead_basis(p):
    p = np.ascontiguousarray(p, np.float64)
    u = np.zeros(3)
    v = np.zeros(3)
    lib().ekg_lead_basis(_p(p), _p(u), _p(v))
    return u, v


def move_leads(pos, disp):
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.zeros_like(pos)
    for i in range(pos.shape[0]):
        o = np.zeros(3)
        lib().ekg_move_lead(_p(np.ascontiguousarray(pos[i])), float(disp[i][0]), float(disp[i][1]), _p(o))
        out[i] = o
    return out


class ApModel(C.Structure):
    _fields_ = [
        ("n_base", C.c_int),
        ("base_k", C.c_void_p),
        ("k_min", C.c_void_p),
        ("k_max", C.c_void_p),
        ("n_free", C.c_int),
        ("free_k", C.c_void_p),
        ("interpolation", C.c_int),
        ("mid_position", C.c_double),
    ]


INTERP = {"endo-epi": 1, "endo-mid-epi": 2}
COMPARE = {"rms": 1, "correlation": 2, "dev-from-linear": 3, "vector-correlation": 4}


def ap_value(k, t):
    """WohlfartPlus W(t) in FP64 (host)."""
    k = np.ascontiguousarray(k, np.float64)
    return float(lib().ekg_ap_value(_p(k), float(t)))


def ap_apd90(k):
    """WohlfartPlus::apd90 of coefficients k (host, FP64; -1 if the curve never
    falls through k0 + k2/10 within 1000 ms)."""
    k = np.ascontiguousarray(k, np.float64)
    return float(lib().ekg_ap_apd90(_p(k)))


def _ap_model(base_k, k_min, k_max, free_k, interpolation, mid_position):
    keep = dict(base=np.ascontiguousarray(base_k, np.float64).reshape(-1, 9),
                kmin=np.ascontiguousarray(k_min, np.float64), kmax=np.ascontiguousarray(k_max, np.float64),
                free=np.ascontiguousarray(free_k, np.int32))
    m = ApModel(keep["base"].shape[0], keep["base"].ctypes.data, keep["kmin"].ctypes.data,
                keep["kmax"].ctypes.data, keep["free"].size, keep["free"].ctypes.data if keep["free"].size else None,
                INTERP.get(interpolation, interpolation), float(mid_position))
    return m, keep


def ap_num_genes(free_k, interpolation):
    m, keep = _ap_model(np.zeros(9), np.zeros(9), np.zeros(9), free_k, interpolation, 0.5)
    n = C.c_int()
    check(lib().ekg_ap_num_genes(C.byref(m), C.byref(n)))
    return n.value


def layer_aps(base_k, k_min, k_max, free_k, interpolation, mid_position, n_layers, genes):
    """Layer APs [n_layers][9] and violation from the leading parameter-vector
    entries (the `[wohlfart ap]` model of simulator.ini)."""
    m, keep = _ap_model(base_k, k_min, k_max, free_k, interpolation, mid_position)
    g = np.ascontiguousarray(genes, np.float64)
    out = np.zeros((n_layers, 9))
    viol = C.c_double()
    check(lib().ekg_layer_aps(C.byref(m), int(n_layers), _p(g) if g.size else None, g.size, _p(out),
                              C.byref(viol)))
    return out, viol.value


def criterion(mode, sim, target, target_offset=0.0):
    """Comparison value of one simulated lead against a target (offset 0)."""
    a = np.ascontiguousarray(sim, np.float64)
    b = np.ascontiguousarray(target, np.float64)
    v = C.c_double()
    check(lib().ekg_criterion(COMPARE.get(mode, mode), _p(a), a.size, _p(b), b.size, float(target_offset),
                              C.byref(v)))
    return v.value


def resample(src, factor, start):
    """Target resampling of the reference's target loader."""
    s = np.ascontiguousarray(src, np.float64)
    n = C.c_int64()
    check(lib().ekg_resample(_p(s), s.size, float(factor), int(start), None, 0, C.byref(n)))
    out = np.zeros(n.value)
    check(lib().ekg_resample(_p(s), s.size, float(factor), int(start), _p(out), out.size, C.byref(n)))
    return out


class CritDesc(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("peaks", C.c_int),
        ("raw", C.c_int),
        ("n_targets", C.c_int),
        ("target_len", C.c_int64),
        ("targets", C.c_void_p),
        ("target_offsets", C.c_void_p),
    ]


class _Crit:
    """Keeps the target arrays behind one CritDesc alive."""

    def __init__(self, mode, targets, offsets, peaks, raw):
        self.t = np.ascontiguousarray(np.atleast_2d(targets), np.float64)
        self.o = np.ascontiguousarray(offsets, np.float64).reshape(-1)
        if self.o.size != self.t.shape[0]:
            raise ValueError("one offset per target")
        self.s = CritDesc(COMPARE.get(mode, mode), int(bool(peaks)), int(bool(raw)), self.t.shape[0],
                          self.t.shape[1], self.t.ctypes.data, self.o.ctypes.data)


class BatchDev(C.Structure):
    _fields_ = [
        ("batch", C.c_int),
        ("n_layers", C.c_int),
        ("layer_k_dev", C.c_void_p),
        ("n_leads", C.c_int),
        ("lead_pos_dev", C.c_void_p),
        ("nhood", C.c_int),
        ("start", C.c_double),
        ("length", C.c_double),
        ("dt", C.c_double),
        ("ecg_dev", C.c_void_p),
    ]


class _Desc:
    """Keeps the arrays behind one SimDesc alive."""

    def __init__(self, layer_k, lead_pos, nhood, start, length, dt):
        self.k = np.ascontiguousarray(layer_k, np.float64).reshape(-1, 9)
        self.p = np.ascontiguousarray(lead_pos, np.float64).reshape(-1, 3)
        self.s = SimDesc(self.k.shape[0], _pd(self.k), self.p.shape[0], _pd(self.p),
                         NHOOD[nhood] if isinstance(nhood, str) else int(nhood),
                         float(start), float(length), float(dt))


class Context:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        self._servers = weakref.WeakSet()
        self._hearts = weakref.WeakSet()
        check(lib().ekg_ctx_create(int(device), C.byref(self.h)))

    def close(self):
        """Closes the context's servers and hearts, then the context."""
        if self.h:
            for s in list(self._servers):
                s.close()
            for h in list(self._hearts):
                h.close()
            lib().ekg_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, on: bool):
        check(lib().ekg_ctx_set_profiling(self.h, 1 if on else 0))

    def kernel_time(self, kernel="ecg_fused"):
        n = C.c_int64()
        ms = C.c_double()
        check(lib().ekg_ctx_kernel_time(self.h, kernel.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def reset_profiling(self):
        check(lib().ekg_ctx_reset_profiling(self.h))

    def synchronize(self):
        check(lib().ekg_ctx_synchronize(self.h))

    def set_kernel(self, kind: str):
        """'auto', 'valu' or 'mfma': lead-contraction path of the fused kernel."""
        check(lib().ekg_ctx_set_kernel(self.h, KERNELS[kind]))

    def heart(self, layer, seed=None):
        return Heart(self, layer, seed)

    def criteria_device(self, mode, targets, offsets, ecg_ptr, batch, n_leads, n_steps, out_ptr,
                        peaks=False, raw=False):
        """Criteria of device ECGs [batch][n_leads][n_steps] into a device
        buffer [batch][n_crit (x2 with peaks)], on the context's stream."""
        c = _Crit(mode, targets, offsets, peaks, raw)
        check(lib().ekg_criteria_device(self.h, C.byref(c.s), C.c_void_p(ecg_ptr), int(batch), int(n_leads),
                                        int(n_steps), C.c_void_p(out_ptr)))

    def set_graphs(self, enable: bool):
        """Replay repeated simulation groups from a captured hipGraph."""
        check(lib().ekg_ctx_set_graphs(self.h, 1 if enable else 0))

    def graph_stats(self):
        c = C.c_int64()
        n = C.c_int64()
        check(lib().ekg_ctx_graph_stats(self.h, C.byref(c), C.byref(n)))
        return {"captures": c.value, "launches": n.value}


class Heart:
    def __init__(self, ctx: Context, layer, seed=None):
        self.ctx = ctx
        self.layer = np.ascontiguousarray(layer, np.uint8)
        if self.layer.ndim != 3:
            raise ValueError("layer must be a 3-D (z, y, x) array")
        self.seed = None if seed is None else np.ascontiguousarray(seed, np.uint8)
        if self.seed is not None and self.seed.shape != self.layer.shape:
            raise ValueError("seed grid shape differs from the layer grid")
        self.h = C.c_void_p()
        n0, n1, n2 = self.layer.shape
        check(lib().ekg_heart_create(ctx.h, n0, n1, n2, _p(self.layer),
                                     None if self.seed is None else _p(self.seed), C.byref(self.h)))
        ctx._hearts.add(self)

    def close(self):
        if self.h:
            lib().ekg_heart_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_myo(self):
        n = C.c_int64()
        check(lib().ekg_heart_num_myocardial(self.h, C.byref(n)))
        return n.value

    @property
    def max_layer(self):
        n = C.c_int()
        check(lib().ekg_heart_max_layer(self.h, C.byref(n)))
        return n.value

    def excite_transfer(self, transfer):
        t = np.ascontiguousarray(transfer, np.float64)
        check(lib().ekg_heart_excite_transfer(self.h, _p(t), t.shape[0], t.shape[1]))

    def excite_delay(self, delay):
        d = np.ascontiguousarray(delay, np.uint8)
        if d.shape != self.layer.shape:
            raise ValueError("delay grid shape differs from the layer grid")
        check(lib().ekg_heart_excite_delay(self.h, _p(d)))

    def set_activation(self, act):
        a = np.ascontiguousarray(act, np.float64)
        if a.shape != self.layer.shape:
            raise ValueError("activation grid shape differs from the layer grid")
        check(lib().ekg_heart_set_activation(self.h, _p(a)))

    def activation(self):
        a = np.zeros(self.layer.shape, np.float64)
        check(lib().ekg_heart_get_activation(self.h, _p(a)))
        return a

    def simulate(self, layer_k, lead_pos, nhood="3D4", start=0.0, length=600.0, dt=1.0):
        d = _Desc(layer_k, lead_pos, nhood, start, length, dt)
        T = num_steps(length, dt)
        out = np.zeros((d.p.shape[0], T), np.float64)
        check(lib().ekg_simulate(self.h, C.byref(d.s), _p(out)))
        return out

    def simulate_batch(self, layer_k, lead_pos, nhood="3D4", start=0.0, length=600.0, dt=1.0):
        """layer_k: (B, n_layers, 9); lead_pos: (B, L, 3) -> (B, L, T)."""
        layer_k = np.asarray(layer_k, np.float64)
        lead_pos = np.asarray(lead_pos, np.float64)
        B = layer_k.shape[0]
        ds = [_Desc(layer_k[b], lead_pos[b], nhood, start, length, dt) for b in range(B)]
        arr = (SimDesc * B)(*[d.s for d in ds])
        T = num_steps(length, dt)
        out = np.zeros((B, lead_pos.shape[1], T), np.float64)
        check(lib().ekg_simulate_batch(self.h, B, arr, _p(out)))
        return out


    def simulate_criteria(self, layer_k, lead_pos, mode, targets, offsets, nhood="3D4", start=0.0,
                          length=600.0, dt=1.0, peaks=False, raw=False):
        """layer_k: (B, n_layers, 9); lead_pos: (B, L, 3) -> criteria (B, n_crit (x2 with peaks));
        the ECGs stay on the device."""
        layer_k = np.asarray(layer_k, np.float64)
        lead_pos = np.asarray(lead_pos, np.float64)
        B = layer_k.shape[0]
        ds = [_Desc(layer_k[b], lead_pos[b], nhood, start, length, dt) for b in range(B)]
        arr = (SimDesc * B)(*[d.s for d in ds])
        c = _Crit(mode, targets, offsets, peaks, raw)
        n_crit = min(lead_pos.shape[1], c.t.shape[0])
        out = np.zeros((B, n_crit * (2 if peaks else 1)), np.float64)
        check(lib().ekg_simulate_criteria(self.h, B, arr, C.byref(c.s), _p(out)))
        return out

    def autotune(self, layer_k, lead_pos, nhood="3D4", start=0.0, length=600.0, dt=1.0, batch=64):
        """Time the launch-shape knobs on `batch` copies of one request and
        keep the fastest for this heart; returns the choice."""
        d = _Desc(layer_k, lead_pos, nhood, start, length, dt)
        t = Tune()
        check(lib().ekg_heart_autotune(self.h, C.byref(d.s), int(batch), C.byref(t)))
        return t.as_dict()

    def set_tune(self, chunk_voxels=0, mfma_waves=0, lead_phase=0, kernel="auto", class_min=0):
        """class_min: activation classes of at least this many voxels take the
        class path (0 = default, -1 = none)."""
        t = Tune(int(chunk_voxels), int(mfma_waves), int(lead_phase),
                 KERNELS[kernel] if isinstance(kernel, str) else int(kernel), int(class_min), 0.0)
        check(lib().ekg_heart_set_tune(self.h, C.byref(t)))

    def get_tune(self):
        t = Tune()
        check(lib().ekg_heart_get_tune(self.h, C.byref(t)))
        return t.as_dict()

    def layout_info(self, nhood="3D4"):
        """Chunks, activation classes and class cells of the matrix-core
        layout as currently tuned."""
        o = LayoutInfo()
        check(lib().ekg_heart_layout_info(self.h, NHOOD[nhood], C.byref(o)))
        return {f: getattr(o, f) for f, _ in o._fields_ if f != "pad"}

    def simulate_many(self, requests, nhood="3D4"):
        """One launch group of requests that may differ in lead count, layer
        count, start, length and time step: requests = [(layer_k, lead_pos,
        start, length, dt), ...] -> list of (L, T) arrays."""
        ds = [_Desc(k, p, nhood, st, ln, dt) for (k, p, st, ln, dt) in requests]
        arr = (SimDesc * len(ds))(*[d.s for d in ds])
        shapes = [(d.p.shape[0], num_steps(r[3], r[4])) for d, r in zip(ds, requests)]
        out = np.zeros(sum(L * T for L, T in shapes), np.float64)
        check(lib().ekg_simulate_batch(self.h, len(ds), arr, _p(out)))
        res, o = [], 0
        for L, T in shapes:
            res.append(out[o:o + L * T].reshape(L, T).copy())
            o += L * T
        return res

    def simulate_batch_device(self, layer_k_ptr, lead_pos_ptr, ecg_ptr, batch, n_layers, n_leads,
                              nhood="3D4", start=0.0, length=600.0, dt=1.0):
        """Enqueue a batch whose arrays are device pointers (ints), e.g. the
        data_ptr() of torch tensors on the GPU; returns without waiting."""
        d = BatchDev(int(batch), int(n_layers), C.c_void_p(int(layer_k_ptr)), int(n_leads),
                     C.c_void_p(int(lead_pos_ptr)),
                     NHOOD[nhood] if isinstance(nhood, str) else int(nhood),
                     float(start), float(length), float(dt), C.c_void_p(int(ecg_ptr)))
        check(lib().ekg_simulate_batch_device(self.h, C.byref(d)))

    def simulate_batch_device_noalloc(self, layer_k_ptr, lead_pos_ptr, ecg_ptr, batch, n_layers, n_leads,
                                      nhood="3D4", start=0.0, length=600.0, dt=1.0):
        """simulate_batch_device with no device allocation: every group buffer
        lives in the GPU's static workspace (ekg_simulate_batch_device_noalloc);
        the heart's layout for `nhood` must be built (an earlier simulation, or
        layout_info on the matrix-core path)."""
        d = BatchDev(int(batch), int(n_layers), C.c_void_p(int(layer_k_ptr)), int(n_leads),
                     C.c_void_p(int(lead_pos_ptr)),
                     NHOOD[nhood] if isinstance(nhood, str) else int(nhood),
                     float(start), float(length), float(dt), C.c_void_p(int(ecg_ptr)))
        check(lib().ekg_simulate_batch_device_noalloc(self.h, C.byref(d)))


class Server:
    def __init__(self, ctx: Context, hbm_budget_bytes=0, max_inflight=0, max_batch=0, n_streams=0,
                 batch_window_us=0):
        self.ctx = ctx
        self.h = C.c_void_p()
        o = ServerOpts(int(hbm_budget_bytes), int(max_inflight), int(max_batch), int(n_streams),
                       int(batch_window_us))
        check(lib().ekg_server_create(ctx.h, C.byref(o), C.byref(self.h)))
        ctx._servers.add(self)
        self._hearts = []
        self._shapes = {}

    def close(self):
        if self.h:
            lib().ekg_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_heart(self, heart: Heart) -> int:
        i = C.c_int()
        check(lib().ekg_server_add_heart(self.h, heart.h, C.byref(i)))
        self._hearts.append(heart)
        return i.value

    def submit(self, heart_id, layer_k, lead_pos, nhood="3D4", start=0.0, length=600.0, dt=1.0):
        d = _Desc(layer_k, lead_pos, nhood, start, length, dt)
        t = C.c_uint64()
        check(lib().ekg_server_submit(self.h, int(heart_id), C.byref(d.s), C.byref(t)))
        self._shapes[t.value] = (d.p.shape[0], num_steps(length, dt))
        return t.value

    def poll(self, ticket) -> bool:
        d = C.c_int()
        check(lib().ekg_server_poll(self.h, C.c_uint64(ticket), C.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        L, T = self._shapes.pop(ticket)
        out = np.zeros((L, T), np.float64)
        lat = C.c_double()
        check(lib().ekg_server_wait(self.h, C.c_uint64(ticket), _p(out), out.size, C.byref(lat)))
        return out, lat.value

    def wait_view(self, ticket):
        """Waits for the ticket and returns (ECG, latency): the ECG is a
        read-only view of the server's pinned buffer, valid until
        release(ticket)."""
        L, T = self._shapes.pop(ticket)
        ptr = C.POINTER(C.c_double)()
        n = C.c_size_t()
        lat = C.c_double()
        check(lib().ekg_server_wait_view(self.h, C.c_uint64(ticket), C.byref(ptr), C.byref(n), C.byref(lat)))
        if n.value != L * T:
            raise EkgError(6, f"result holds {n.value} doubles, {L * T} expected")
        view = np.ctypeslib.as_array(ptr, shape=(L, T)) if n.value else np.zeros((L, T))
        view.flags.writeable = False
        return view, lat.value

    def release(self, ticket):
        check(lib().ekg_server_release(self.h, C.c_uint64(ticket)))

    def stats(self):
        c = C.c_uint64()
        g = C.c_uint64()
        p = C.c_int()
        b = C.c_size_t()
        check(lib().ekg_server_stats(self.h, C.byref(c), C.byref(g), C.byref(p), C.byref(b)))
        return {"completed": c.value, "groups": g.value, "peak_inflight": p.value,
                "peak_bytes": b.value}

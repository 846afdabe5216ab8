"""ctypes front-end of the FP64 oracle (liboracle.so) and, where it was built,
of the compiled reference harness (_ref/libekgref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libekgref.so")
_REFAPP = os.path.join(HERE, "_ref", "libekgrefapp.so")
REF_CLI = os.path.join(HERE, "_ref", "ekgSim")

NHOOD = {"2D4": 0, "2D8": 1, "3D4": 2, "3D8": 3, "cube": 3}

_d = np.ctypeslib.ndpointer(np.float64, flags="C")
_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
_i = C.c_int
_dbl = C.c_double
_i64 = C.c_int64


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and the reference harness when ref=True and the
    reference sources exist)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir(os.environ.get("REF", "/root/reference")):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_wohlfart.restype = _dbl
        L.orc_wohlfart.argtypes = [_d, _dbl]
        L.orc_cell_ap.restype = _dbl
        L.orc_cell_ap.argtypes = [_d, _dbl, _dbl]
        L.orc_wohlfart_series.argtypes = [_d, _d, _i64, _d]
        L.orc_neighbourhood.restype = _i
        L.orc_neighbourhood.argtypes = [_i, _i32]
        L.orc_excite_fp64.restype = _i
        L.orc_excite_fp64.argtypes = [_i, _i, _i, _u8, _u8, _d, _i, _i, _d]
        L.orc_excite_fp64_bf.restype = _i
        L.orc_excite_fp64_bf.argtypes = [_i, _i, _i, _u8, _u8, _d, _i, _i, _d]
        L.orc_excite_int.restype = _i
        L.orc_excite_int.argtypes = [_i, _i, _i, _u8, _u8, _u8, _i32]
        L.orc_lead_basis.argtypes = [_d, _d, _d]
        L.orc_move_point.argtypes = [_d, _d, _d, _dbl, _dbl, _d]
        L.orc_num_steps.restype = _i64
        L.orc_num_steps.argtypes = [_dbl, _dbl]
        for fn in (L.orc_run_literal, L.orc_run_regrouped):
            fn.restype = _i
            fn.argtypes = [_i, _i, _i, _u8, _d, _i, _d, _i, _dbl, _dbl, _dbl, _i, _d, _d]
        L.orc_lead_field.restype = _i64
        L.orc_lead_field.argtypes = [_i, _i, _i, _u8, _i, _i, _d, _d]
        L.orc_layer_aps_sim3.restype = _i
        L.orc_layer_aps_sim3.argtypes = [_d, _d, _i32, _i, _dbl, _d, _d, _d, _d, _d]
        L.orc_pearson_crit.restype = _dbl
        L.orc_pearson_crit.argtypes = [_d, _i64, _d, _i64]
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(_REF)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(_REF)
        R.ref_last_error.restype = C.c_char_p
        R.ref_wohlfart.restype = _dbl
        R.ref_wohlfart.argtypes = [_d, _dbl]
        R.ref_cell_ap.restype = _dbl
        R.ref_cell_ap.argtypes = [_d, _dbl, _dbl]
        R.ref_excite.restype = _i
        R.ref_excite.argtypes = [_i, _i, _i, _u8, _u8, _d, _i, _i, C.c_char_p, _d]
        R.ref_run.restype = _i
        R.ref_run.argtypes = [_i, _i, _i, _u8, _d, _i, _d, _i, _dbl, _dbl, _dbl, _i, _d, _d, _i64]
        R.ref_ekgsim_pipeline.restype = _i
        R.ref_ekgsim_pipeline.argtypes = [C.c_char_p, C.c_void_p, _i, _d, _i, _d, _i64, C.c_void_p, _i64]
        _ref = R
    return _ref


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class OracleError(RuntimeError):
    pass


def _chk(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with code {rc}")


# -- AP ---------------------------------------------------------------------
def wohlfart(k, t):
    k = _c(k, np.float64)
    t = np.atleast_1d(_c(t, np.float64))
    out = np.empty_like(t)
    lib().orc_wohlfart_series(k, t, t.size, out)
    return out


def cell_ap(layer_k, at, time):
    return lib().orc_cell_ap(_c(layer_k, np.float64), float(at), float(time))


def neighbourhood(name):
    buf = np.zeros((26, 3), np.int32)
    n = lib().orc_neighbourhood(NHOOD[name], buf)
    return buf[:n].copy()


def num_steps(length, dt):
    return int(lib().orc_num_steps(float(length), float(dt)))


# -- CA ---------------------------------------------------------------------
def excite_fp64(layer, seed, transfer, brute_force=False):
    layer = _c(layer, np.uint8)
    seed = _c(seed, np.uint8)
    t = _c(transfer, np.float64)
    act = np.zeros(layer.shape, np.float64)
    n0, n1, n2 = layer.shape
    fn = lib().orc_excite_fp64_bf if brute_force else lib().orc_excite_fp64
    _chk(fn(n0, n1, n2, layer, seed, t, t.shape[0], t.shape[1], act), "orc_excite_fp64")
    return act


def excite_int(layer, seed, delay):
    layer = _c(layer, np.uint8)
    act = np.zeros(layer.shape, np.int32)
    n0, n1, n2 = layer.shape
    _chk(lib().orc_excite_int(n0, n1, n2, layer, _c(seed, np.uint8), _c(delay, np.uint8), act),
         "orc_excite_int")
    return act


# -- ECG --------------------------------------------------------------------
def run(layer, act, layer_k, pos, nhood="3D4", start=0.0, length=600.0, dt=1.0, literal=False):
    layer = _c(layer, np.uint8)
    act = _c(act, np.float64)
    k = _c(layer_k, np.float64)
    pos = _c(pos, np.float64)
    T = num_steps(length, dt)
    L = pos.shape[0]
    ecg = np.zeros((L, T), np.float64)
    n0, n1, n2 = layer.shape
    fn = lib().orc_run_literal if literal else lib().orc_run_regrouped
    _chk(fn(n0, n1, n2, layer, act, k.shape[0], k, NHOOD[nhood], float(start), float(length),
            float(dt), L, pos, ecg), "orc_run")
    return ecg


def lead_field(layer, pos, nhood="3D4"):
    layer = _c(layer, np.uint8)
    pos = _c(pos, np.float64)
    V = int(np.count_nonzero(layer))
    G = np.zeros((V, pos.shape[0]), np.float64)
    n0, n1, n2 = layer.shape
    got = lib().orc_lead_field(n0, n1, n2, layer, NHOOD[nhood], pos.shape[0], pos, G)
    if got != V:
        raise OracleError("orc_lead_field failed")
    return G


def lead_basis(p):
    U = np.zeros(3)
    V = np.zeros(3)
    lib().orc_lead_basis(_c(p, np.float64), U, V)
    return U, V


def move_leads(pos, disp):
    out = np.zeros_like(_c(pos, np.float64))
    for i, p in enumerate(pos):
        U, V = lead_basis(p)
        o = np.zeros(3)
        lib().orc_move_point(_c(p, np.float64), U, V, float(disp[i][0]), float(disp[i][1]), o)
        out[i] = o
    return out


def layer_aps_sim3(base_first, base_last, free_k, mid_position, kmin, kmax, x):
    lk = np.zeros((3, 9))
    viol = np.zeros(1)
    fk = _c(free_k, np.int32)
    _chk(lib().orc_layer_aps_sim3(_c(base_first, np.float64), _c(base_last, np.float64), fk,
                                  len(fk), float(mid_position), _c(kmin, np.float64),
                                  _c(kmax, np.float64), _c(x, np.float64), lk, viol),
         "orc_layer_aps_sim3")
    return lk, float(viol[0])


def pearson_crit(a, b):
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    return lib().orc_pearson_crit(a, a.size, b, b.size)


# -- compiled reference (only where oracle/_ref was built) -------------------
def ref_wohlfart(k, t):
    return ref().ref_wohlfart(_c(k, np.float64), float(t))


def ref_cell_ap(layer_k, at, time):
    return ref().ref_cell_ap(_c(layer_k, np.float64), float(at), float(time))


def ref_excite(layer, seed, transfer, tmp_path):
    layer = _c(layer, np.uint8)
    t = _c(transfer, np.float64)
    act = np.zeros(layer.shape, np.float64)
    n0, n1, n2 = layer.shape
    rc = ref().ref_excite(n0, n1, n2, layer, _c(seed, np.uint8), t, t.shape[0], t.shape[1],
                          str(tmp_path).encode(), act)
    if rc:
        raise OracleError(ref().ref_last_error().decode())
    return act


def ref_run(layer, act, layer_k, pos, nhood="3D4", start=0.0, length=600.0, dt=1.0):
    layer = _c(layer, np.uint8)
    k = _c(layer_k, np.float64)
    pos = _c(pos, np.float64)
    T = num_steps(length, dt)
    ecg = np.zeros((pos.shape[0], T), np.float64)
    n0, n1, n2 = layer.shape
    rc = ref().ref_run(n0, n1, n2, layer, _c(act, np.float64), k.shape[0], k, NHOOD[nhood],
                       float(start), float(length), float(dt), pos.shape[0], pos, ecg, ecg.size)
    if rc:
        raise OracleError(ref().ref_last_error().decode())
    return ecg


def ref_pipeline(ini_path, disp, layer_k, n_leads, n_steps, act_size=0):
    k = _c(layer_k, np.float64)
    ecg = np.zeros((n_leads, n_steps), np.float64)
    d = None if disp is None else _c(disp, np.float64)
    act = np.zeros(max(act_size, 1), np.float64)
    rc = ref().ref_ekgsim_pipeline(str(ini_path).encode(),
                                   None if d is None else d.ctypes.data, n_leads, k, k.shape[0],
                                   ecg, ecg.size, act.ctypes.data if act_size else None, act_size)
    if rc:
        raise OracleError(ref().ref_last_error().decode())
    return ecg, (act if act_size else None)


# -- compiled reference application code (oracle/_ref/libekgrefapp.so) --------
_refapp = None


def have_refapp() -> bool:
    return os.path.exists(_REFAPP)


def refapp():
    global _refapp
    if _refapp is None:
        L = C.CDLL(_REFAPP)
        L.ref_interp_layers.restype = _i
        L.ref_interp_layers.argtypes = [_i, _i, _dbl, _d, _d, _d, _d]
        L.ref_apd90.restype = _dbl
        L.ref_apd90.argtypes = [_d]
        L.ref_compare.restype = _dbl
        L.ref_compare.argtypes = [_i, _d, _i64, _d, _i64, _dbl]
        L.ref_resample.restype = _i64
        L.ref_resample.argtypes = [_d, _i64, _dbl, _i, _d, _i64]
        _refapp = L
    return _refapp


def ref_interp_layers(interp, n_layers, mid_position, front_k, mid_k, back_k):
    out = np.zeros((n_layers, 9))
    refapp().ref_interp_layers(int(interp), int(n_layers), float(mid_position), _c(front_k, np.float64),
                               _c(mid_k, np.float64), _c(back_k, np.float64), out)
    return out


def ref_apd90(k):
    return refapp().ref_apd90(_c(k, np.float64))


def ref_compare(mode, a, b, bofs=0.0):
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    return refapp().ref_compare(int(mode), a, a.size, b, b.size, float(bofs))


def ref_resample(src, factor, start):
    src = _c(src, np.float64)
    n = refapp().ref_resample(src, src.size, float(factor), int(start), np.zeros(1), 0)
    out = np.zeros(max(n, 1))
    refapp().ref_resample(src, src.size, float(factor), int(start), out, n)
    return out[:n]

"""The ekgSim command line against the reference program's own outputs.

tests/golden/cli/<case>/ holds inputs and what the compiled reference ekgSim
wrote for them (make_cli_golden.py).  CPU tests pin those fixtures with the
FP64 oracle; GPU tests run this build's ekgSim on the same inputs and compare
every file it writes: ECG columns within rel-L2 1e-5 per lead (values are
printed with 6 significant digits, so the print rounding of both sides is
part of the error), host-computed files (layer and cell APs, resampled
targets, excitation export) byte for byte, criteria within the print
resolution.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT, rel_l2

CLI = os.path.join(ROOT, "ekgsim_amd", "bin", "ekgSim")
CASES = sorted(os.listdir(os.path.join(GOLDEN, "cli"))) if os.path.isdir(os.path.join(GOLDEN, "cli")) else []
TOL = 1e-5


def _case_dir(case):
    return os.path.join(GOLDEN, "cli", case)


def _args(case):
    with open(os.path.join(_case_dir(case), "args.txt")) as f:
        return [ln.rstrip("\n") for ln in f if ln.strip()]


def _expected_criteria(case):
    with open(os.path.join(_case_dir(case), "expected", "criteria.txt")) as f:
        crit, viol = f.read().split("\n")[:2]
    return np.array([float(v) for v in crit.split(",")]), float(viol)


def read_column(path):
    """(header lines, data rows) of a .column file."""
    head, rows = [], []
    with open(path) as f:
        for ln in f:
            if ln.startswith("#") or ln.startswith("%"):
                head.append(ln.rstrip("\n"))
            elif ln.strip():
                rows.append([float(v) for v in ln.split()])
    return head, np.array(rows)


def read_matrix(path):
    """A heart in this build's .matrix writer format (negative = seed)."""
    with open(path) as f:
        f.readline()
        hdr = f.readline().split()
        nx, ny, nz = int(hdr[1]), int(hdr[3]), int(hdr[5])
        vals = np.array(f.read().split(), np.int64)
    vals = vals.reshape(nz, ny, nx)
    return np.abs(vals).astype(np.uint8), (vals < 0).astype(np.uint8)


def read_ini(path):
    sec, out = "", {}
    with open(path) as f:
        for ln in f:
            ln = ln.strip()
            if not ln or ln.startswith(";"):
                continue
            if ln.startswith("["):
                sec = ln.strip("[]").strip()
            elif "=" in ln:
                k, v = ln.split("=", 1)
                out[(sec, k.strip())] = v.strip()
    return out


# ---- CPU: the fixtures against the FP64 oracle --------------------------------


def test_cli_fixtures_present():
    assert {"heart3", "heart6", "ring2d", "heart3_excitation_file", "heart3_approx_only"} <= set(CASES)


@pytest.mark.parametrize("case", ["heart3", "heart3_excitation_file"])
def test_reference_cli_output_matches_oracle(case):
    """Recompute the reference program's ECG with the oracle from the case's
    files: CA from the conduction matrix (or the excitation file), layer APs
    from the -sim vector, electrodes moved by the displacement genes."""
    from oracle import oracle as O
    d = _case_dir(case)
    ini = read_ini(os.path.join(d, "simulator.ini"))
    layer, seed = read_matrix(os.path.join(d, "heart.matrix"))
    with open(os.path.join(d, "conduction.matrix")) as f:
        f.readline()
        hdr = f.readline().split()
        T = np.array(f.read().split(), float).reshape(int(hdr[3]), int(hdr[1]))
    if ("model", "excitation sequence") in ini:
        with open(os.path.join(d, ini[("model", "excitation sequence")])) as f:
            f.readline()
            f.readline()
            act = np.array(f.read().split(), float).reshape(layer.shape)
    else:
        act = O.excite_fp64(layer, seed, T)
    x = np.array([float(v) for v in _args(case)[2].split(",")])
    base = np.array([0, 2.5, 100, 0.9, 0.1, 0.001, 0.1, 0.1, 250.0])
    k = np.tile(base, (3, 1))
    for l in range(3):
        k[l, 5:9] = x[4 * l:4 * l + 4]
    pts = []
    with open(os.path.join(d, "points.txt")) as f:
        for ln in f:
            xx, yy, zz = (float(v) for v in ln.split(","))
            pts.append([zz, yy, xx])
    pos = O.move_leads(np.array(pts), x[12:].reshape(-1, 2))
    start = float(ini[("simulation", "start")])
    length = float(ini[("simulation", "length")])
    dt = float(ini[("simulation", "time step")])
    ecg = O.run(layer, act, k, pos, ini[("simulation", "neighbourhood type")], start, length, dt)
    _, rows = read_column(os.path.join(d, "expected", "result.column"))
    assert rows.shape == (ecg.shape[1], 1 + ecg.shape[0])
    assert np.allclose(rows[:, 0], start + dt * np.arange(ecg.shape[1]))
    for i in range(ecg.shape[0]):
        assert rel_l2(rows[:, 1 + i], ecg[i]) <= TOL
    # criteria: 1 - Pearson (or deviation from linear) of the printed ECG
    # against the unshifted, range-normalised targets
    crit, viol = _expected_criteria(case)
    assert viol == 0.0
    _, tg = read_column(os.path.join(d, "targets.column"))
    if ini[("optimization", "comparison mode")] == "2":
        for i in range(ecg.shape[0]):
            assert abs(O.pearson_crit(ecg[i], tg[:, 1 + i]) - crit[i]) <= 2e-5 * abs(crit[i]) + 1e-6


def test_cli_help_runs_without_gpu():
    if not os.path.exists(CLI):
        pytest.skip("ekgSim not built")
    r = subprocess.run([CLI, "-?"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 0 and "-sim" in r.stdout


# ---- GPU: this build's ekgSim on the same inputs ------------------------------


def _run_cli(case, tmp_path, args=None):
    src = _case_dir(case)
    for fn in os.listdir(src):
        if fn != "expected":
            shutil.copy(os.path.join(src, fn), tmp_path / fn)
    print(f"ekgSim case {case}", flush=True)
    r = subprocess.run([CLI] + (args if args is not None else _args(case)[1:]), cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


def _criteria_line(stdout):
    m = re.search(r"criteria = <([^>]*)>, violation = (\S+)", stdout)
    assert m, stdout[-2000:]
    return np.array([float(v) for v in m.group(1).split(",")]), float(m.group(2))


def _close(a, b):
    return np.all(np.abs(a - b) <= 2e-5 * np.abs(b) + 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_cli_matches_reference_program(case, tmp_path):
    r = _run_cli(case, tmp_path)
    crit, viol = _criteria_line(r.stdout)
    exp_crit, exp_viol = _expected_criteria(case)
    assert crit.shape == exp_crit.shape
    assert _close(crit, exp_crit), (crit, exp_crit)
    assert viol == exp_viol
    exp_dir = os.path.join(_case_dir(case), "expected")
    for fn in sorted(os.listdir(exp_dir)):
        if fn == "criteria.txt":
            continue
        got_path = tmp_path / fn
        assert got_path.exists(), fn
        if fn == "result.column":
            h1, r1 = read_column(got_path)
            h0, r0 = read_column(os.path.join(exp_dir, fn))
            assert h1 == h0
            assert r1.shape == r0.shape
            if r0.size == 0:  # simulation skipped: header only
                continue
            assert np.array_equal(r1[:, 0], r0[:, 0])
            for i in range(1, r0.shape[1]):
                e = rel_l2(r1[:, i], r0[:, i])
                print(f"{case} lead {i}: rel-L2 {e:.2e}")
                assert e <= TOL
        else:
            with open(got_path) as f1, open(os.path.join(exp_dir, fn)) as f0:
                assert f1.read() == f0.read(), fn


@pytest.mark.gpu
def test_cli_batch_equals_single_runs(tmp_path):
    """-batch evaluates many vectors in one GPU launch set; each line equals
    the criteria of the corresponding -sim run."""
    case = "heart3"
    x0 = _args(case)[2]
    rng = np.random.default_rng(7)
    xs = [x0]
    for _ in range(3):
        g = list(rng.uniform([3e-4, 0.01, 0.01, 300] * 3, [1e-3, 0.1, 0.1, 450] * 3)) + list(rng.uniform(-50, 50, 4))
        xs.append(",".join(f"{v:.6g}" for v in g))
    (tmp_path / "in.txt").write_text("# vectors\n" + "\n".join(xs) + "\n")
    _run_cli(case, tmp_path, ["-batch", "in.txt", "-batch-out", "out.txt"])
    lines = (tmp_path / "out.txt").read_text().split("\n")
    for i, x in enumerate(xs):
        r = _run_cli(case, tmp_path, ["-sim", x])
        crit, viol = _criteria_line(r.stdout)
        vals = np.array([float(v) for v in lines[i].split()])
        assert _close(vals[:-1], crit) and vals[-1] == viol


@pytest.mark.gpu
def test_cli_external_evaluation_protocol(tmp_path):
    """input.txt / output.txt as AMS-DEMO's ExternalEvaluation writes and
    reads them: '#' comment, one value per line; '# violation v' then one
    criterion per line."""
    case = "heart3"
    x = [float(v) for v in _args(case)[2].split(",")]
    (tmp_path / "input.txt").write_text("# file generated by ExternalEvaluation class\n" +
                                        "".join(f"{v!r}\n" for v in x))
    _run_cli(case, tmp_path, ["-external", str(tmp_path)])
    out = (tmp_path / "output.txt").read_text().split("\n")
    assert out[0].startswith("# violation ")
    vals = np.array([float(v) for v in out[1:] if v.strip()])
    exp_crit, exp_viol = _expected_criteria(case)
    assert float(out[0].split()[2]) == exp_viol
    assert _close(vals, exp_crit)

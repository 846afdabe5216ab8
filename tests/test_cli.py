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
    assert {"heart3", "heart6", "ring2d", "heart3_excitation_file", "heart3_approx_only", "model24"} <= set(CASES)


def _model24():
    d = _case_dir("model24")
    z = np.load(os.path.join(d, "heart.npz"), allow_pickle=False)
    with open(os.path.join(d, "conduction.matrix")) as f:
        f.readline()
        hdr = f.readline().split()
        T = np.array(f.read().split(), float).reshape(int(hdr[3]), int(hdr[1]))
    with open(os.path.join(d, "expected", "activation.txt")) as f:
        sha, counts = f.read().split("\n")[:2]
    return z["layer"], z["seed"], T, sha, counts.split()


def test_model24_fixture_is_the_shipped_heart(tmp_path):
    """The stored heart is the reference's testRun/model_24.matrix (93 x 124 x
    124, 24 layers) and survives the round trip through the .matrix writer;
    the conduction matrix is the 26 x 26 FP64 conduction_24 (0.166667 on the
    diagonal, |i - j| off it, -1 for no tissue)."""
    layer, seed, T, sha, counts = _model24()
    assert layer.shape == (124, 124, 93)
    assert layer.max() == 24 and int(counts[0]) == int((layer > 0).sum())
    assert seed.sum() > 0 and np.all(layer[seed > 0] > 0)
    assert T.shape == (26, 26)
    i = np.arange(1, 26)
    assert np.allclose(np.diag(T)[1:], 0.166667)
    assert np.array_equal(T[1:, 1:][~np.eye(25, dtype=bool)], np.abs(i[:, None] - i[None, :])[~np.eye(25, dtype=bool)])
    _materialise("model24", tmp_path)
    back, bseed = read_matrix(tmp_path / "heart.matrix")
    assert np.array_equal(back, layer) and np.array_equal(bseed, seed)


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


def _materialise(case, tmp_path):
    """Copy a case's inputs; a heart stored as heart.npz (layer, seed) is
    written out as heart.matrix in the reference's text format."""
    src = _case_dir(case)
    for fn in os.listdir(src):
        if fn != "expected":
            shutil.copy(os.path.join(src, fn), tmp_path / fn)
    if os.path.exists(os.path.join(src, "heart.npz")):
        from ekgsim_amd import synth as S
        z = np.load(os.path.join(src, "heart.npz"), allow_pickle=False)
        S.write_matrix(str(tmp_path / "heart.matrix"), z["layer"], z["seed"])


def _run_cli(case, tmp_path, args=None, binary=CLI):
    _materialise(case, tmp_path)
    print(f"{os.path.basename(binary)} case {case}", flush=True)
    r = subprocess.run([binary] + (args if args is not None else _args(case)[1:]), cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


def _criteria_line(stdout):
    m = re.search(r"criteria = <([^>]*)>, violation = (\S+)", stdout)
    assert m, stdout[-2000:]
    return np.array([float(v) for v in m.group(1).split(",")]), float(m.group(2))


def _close(a, b):
    return np.all(np.abs(a - b) <= 2e-5 * np.abs(b) + 1e-6)


def _check_outputs(case, tmp_path, r):
    crit, viol = _criteria_line(r.stdout)
    exp_crit, exp_viol = _expected_criteria(case)
    assert crit.shape == exp_crit.shape
    assert _close(crit, exp_crit), (crit, exp_crit)
    assert viol == exp_viol
    exp_dir = os.path.join(_case_dir(case), "expected")
    for fn in sorted(os.listdir(exp_dir)):
        if fn in ("criteria.txt", "activation.txt"):  # the printed criteria; the activation hash (not a file the CLI writes)
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
@pytest.mark.parametrize("case", CASES)
def test_cli_matches_reference_program(case, tmp_path):
    _check_outputs(case, tmp_path, _run_cli(case, tmp_path))


# The reference's own main.cpp and sim.cpp, unmodified, compiled against this
# build's simlib/sim_lib.h and linked with libekgsim.so (oracle/Makefile,
# target _ref/ekgSim_facade; built where the reference sources are present).
FACADE_CLI = os.path.join(ROOT, "oracle", "_ref", "ekgSim_facade")
REF_SRC = "/root/reference"


def test_reference_driver_compiles_against_facade():
    """sim.cpp (SimImplementation: ActionPotential::wohl, WohlfartPlus,
    EkgSim::PositionVec, ColumnFile, saveVecElement) and main.cpp compile
    against the facade header and link with the library alone."""
    if not os.path.isdir(REF_SRC):
        pytest.skip("reference sources absent")
    if not os.path.exists(os.path.join(ROOT, "ekgsim_amd", "lib", "libekgsim.so")):
        pytest.skip("libekgsim.so not built")
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "_ref/ekgSim_facade"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    needed = subprocess.run(["readelf", "-d", FACADE_CLI], stdout=subprocess.PIPE, text=True).stdout
    assert "libekgsim.so" in needed
    # nothing of the reference's simulation core is linked in
    syms = subprocess.run(["nm", "-C", FACADE_CLI], stdout=subprocess.PIPE, text=True).stdout
    assert "SimLib::Simulation::run" not in syms
    assert "calculateExcitationSequence" not in syms


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_reference_driver_on_facade_matches_reference_program(case, tmp_path):
    """The reference's command-line program, running its own evaluation code
    (sim.cpp) over this build's EkgSim, writes the files the compiled
    reference wrote for the same case."""
    if not os.path.exists(FACADE_CLI):
        pytest.skip("ekgSim_facade not built (needs the reference sources)")
    _check_outputs(case, tmp_path, _run_cli(case, tmp_path, binary=FACADE_CLI))


@pytest.mark.gpu
def test_model24_activation_bit_exact():
    """FP64 excitation of the 24-layer shipped heart on the GPU against the
    SHA-256 of the compiled reference's activation times: every voxel's
    time, bit for bit (rounding-sensitive sums over 24 conduction layers)."""
    import hashlib

    import ekgsim_amd as E
    layer, seed, T, sha, counts = _model24()
    ctx = E.Context(0)
    h = E.Heart(ctx, layer, seed)
    h.excite_transfer(T)
    act = h.activation()
    myo = layer > 0
    print(f"model24: {int(myo.sum())} myocardial voxels, max activation {act.max()!r}", flush=True)
    assert int((act[myo] == 0).sum()) == int(counts[1])
    assert act.max() == float(counts[2])
    assert hashlib.sha256(np.ascontiguousarray(act, np.float64).tobytes()).hexdigest() == sha
    h.close()
    ctx.close()


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

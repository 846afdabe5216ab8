"""Write the command-line golden cases under tests/golden/cli/ by running the
reference's own ekgSim program (oracle/_ref/ekgSim, compiled from the
reference sources by `make -C oracle ref`) on synthetic inputs.

Each case directory holds the inputs (simulator.ini, heart, conduction
matrix, measuring points, targets), `args.txt` (the command-line arguments)
and `expected/`: the files the reference wrote and `criteria.txt` with the
criteria and violation it printed.  Nothing of the reference's source is
stored.  Regenerate with:

    python tests/golden/make_cli_golden.py           # every case
    python tests/golden/make_cli_golden.py model24   # the shipped 24-layer heart only
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "cli")
sys.path.insert(0, ROOT)

from ekgsim_amd import synth as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

BASE = "0, 2.5, 100, 0.90, 0.1, 0.0010, 0.10,  0.10, 250"
KMIN = "0, 1.5, 100, 0.85, 0.05, 0.0003, 0.01,  0.01, 300"
KMAX = "0, 3.5, 100, 0.95, 0.2, 0.0010, 0.10,  0.10, 450"

INI = """; synthetic case for the ekgSim command line
[model]
{shape_line}
transfer = conduction.matrix
points = points.txt
{excitation_line}

[output files]
results filename = result.column
{excitation_out}

[simulation]
start = {start}
length = {length}
neighbourhood type = {nhood}
time step = {dt}

[measuring points]
displacement min = -50
displacement max = 50

[wohlfart ap]
interpolation = {interp}
mid AP position = {mid}
base ap 1 = {base}
base ap 2 = {base}
base ap 3 = {base}
k min = {kmin}
k max = {kmax}
free k = 5, 6, 7, 8

[optimization]
mode = 1
targets filename = targets.column
comparison mode = {cmp}
endo-epi minimization criterion epi delay = {endo_epi}
peak position is criterion = {peak}
optimize measuring points = {move}
fast approximation is criterion = {fa}
fast approximation epi delay = {fa_delay}
fast approximation limit = {fa_limit}
"""


def targets(path, n_rows, n_leads, step, with_time, rng):
    t = np.arange(n_rows) * step
    cols = []
    for i in range(n_leads):
        c = 0.4 * np.exp(-((t - 250 - 20 * i) / 45.0) ** 2) - 0.15 * np.exp(-((t - 120) / 20.0) ** 2)
        c += 0.01 * rng.standard_normal(n_rows) + 0.05 * i
        cols.append(c * (300 + 100 * i))
    if not with_time:
        # a file of one column is a single target without time; with more
        # columns the first one is always read as time (sim.cpp:1003-1009)
        cols = cols[:1]
    with open(path, "w") as f:
        f.write("# synthetic target ECGs\n% second comment line\n")
        for r in range(n_rows):
            vals = ([f"{t[r]:g}"] if with_time else []) + [f"{c[r]:.6f}" for c in cols]
            f.write(" ".join(vals) + "\n")


def case(name, *, n=20, n_layers=3, heart_seed=3, n_leads=2, shape=True, start=0, length=400, nhood="3D4",
         dt=1, interp="endo-mid-epi", mid=0.6, cmp=2, endo_epi=-0.5, peak=0, move=1, fa=0, fa_delay=10.0,
         fa_limit=2, target_rows=600, target_step=1.0, target_time=True, conduction_size=12, out=("result",),
         excitation_in=None, excitation_out=None, x_seed=1000, extra_files=()):
    d = os.path.join(OUT, name)
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(os.path.join(d, "expected"))
    rng = np.random.default_rng(x_seed)
    if shape:
        h = S.make_heart(n, seed=heart_seed, n_layers=n_layers)
        S.write_matrix(os.path.join(d, "heart.matrix"), h.layer, h.seed)
        dims = h.layer.shape
    else:
        dims = (1, 160, 120)
        n_layers = 12
    S.write_transfer(os.path.join(d, "conduction.matrix"),
                     S.layered_transfer(n_layers, conduction_size, 1.0, 1.5))
    pos = S.lead_positions(max(dims), n_leads, dims=dims)
    S.write_points(os.path.join(d, "points.txt"), pos)
    targets(os.path.join(d, "targets.column"), target_rows, n_leads, target_step, target_time, rng)
    for src, dst in extra_files:
        shutil.copy(src, os.path.join(d, dst))
    with open(os.path.join(d, "simulator.ini"), "w") as f:
        f.write(INI.format(
            shape_line="shape = heart.matrix" if shape else "; no shape: the built-in 2-D ring heart",
            excitation_line=f"excitation sequence = {excitation_in}" if excitation_in else "",
            excitation_out=f"excitation sequence = {excitation_out}" if excitation_out else "",
            start=start, length=length, nhood=nhood, dt=dt, interp=interp, mid=mid, base=BASE, kmin=KMIN,
            kmax=KMAX, cmp=cmp, endo_epi=endo_epi, peak=peak, move=move, fa=fa, fa_delay=fa_delay,
            fa_limit=fa_limit))
    reps = 3 if interp == "endo-mid-epi" else 2
    genes = []
    for _ in range(reps):
        genes += list(rng.uniform([3e-4, 0.01, 0.01, 300], [1e-3, 0.1, 0.1, 450]))
    if move:
        genes += list(rng.uniform(-50, 50, size=2 * n_leads))
    x = ",".join(f"{v:.6g}" for v in genes)
    args = ["ignored", "-sim", x]
    for o in out:
        args += ["-out", o]
    with open(os.path.join(d, "args.txt"), "w") as f:
        f.write("\n".join(args) + "\n")
    before = set(os.listdir(d))
    r = subprocess.run([O.REF_CLI] + args, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=600)
    m = re.search(r"criteria = <([^>]*)>, violation = (\S+)", r.stdout)
    if not m:
        raise SystemExit(f"{name}: reference run failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    with open(os.path.join(d, "expected", "criteria.txt"), "w") as f:
        f.write(m.group(1) + "\n" + m.group(2) + "\n")
    for fn in sorted(set(os.listdir(d)) - before):
        shutil.move(os.path.join(d, fn), os.path.join(d, "expected", fn))
    size = sum(os.path.getsize(os.path.join(dp, f)) for dp, _, fs in os.walk(d) for f in fs)
    print(f"{name}: criteria <{m.group(1)}> violation {m.group(2)}; {size} bytes")


TESTRUN = "/root/reference/testRun"
# The parameter vector the reference's README gives for its shipped heart
# (README.md:108): endo, mid and epi k5..k8, then two displacements per lead.
TESTRUN_X = ("0.00035813,0.0890636,0.0632915,226.183,0.000369406,0.0965625,0.0523254,232.278,"
             "0.000710767,0.0720323,0.0187579,200.93,23,22,15,13")


def read_reference_matrix(path):
    """(layer, seed) of a reference .matrix heart written with the 'space'
    separator: '3d X x Y x Z space', then Z blocks of Y rows of X integers; a
    negative value marks an excitation start of layer -v."""
    with open(path) as f:
        f.readline()
        hdr = f.readline().split()
        assert hdr[0] in ("3d", "3D") and hdr[-1] == "space", hdr
        nx, ny, nz = int(hdr[1]), int(hdr[3]), int(hdr[5])
        vals = np.array(f.read().split(), np.int64)
    assert vals.size == nx * ny * nz
    vals = vals.reshape(nz, ny, nx)
    return np.abs(vals).astype(np.uint8), (vals < 0).astype(np.uint8)


def case_testrun(name="model24"):
    """The reference's shipped heart (testRun/: model_24.matrix, 24 layers,
    conduction_24.matrix, two measuring points, target_ecg_v2_v5.column) with
    the simulation settings of testRun/simulator.ini (start 100, length 400,
    3D4, 1 ms, endo-mid-epi at 0.6, 1 - Pearson, moved electrodes) and the
    README vector.  The heart is stored as heart.npz (layer, seed); the tests
    write heart.matrix from it.  expected/activation.txt holds the SHA-256 of
    the reference's FP64 activation times (Simulation::calculateExcitationSequence
    through oracle/_ref/libekgref.so) on that heart, in (z, y, x) order."""
    import hashlib
    d = os.path.join(OUT, name)
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(os.path.join(d, "expected"))
    layer, seed = read_reference_matrix(os.path.join(TESTRUN, "model_24.matrix"))
    np.savez_compressed(os.path.join(d, "heart.npz"), layer=layer, seed=seed)
    S.write_matrix(os.path.join(d, "heart.matrix"), layer, seed)
    shutil.copy(os.path.join(TESTRUN, "conduction_24.matrix"), os.path.join(d, "conduction.matrix"))
    shutil.copy(os.path.join(TESTRUN, "model_24_measuring_pos.txt"), os.path.join(d, "points.txt"))
    shutil.copy(os.path.join(TESTRUN, "target_ecg_v2_v5.column"), os.path.join(d, "targets.column"))
    for fn in ("conduction.matrix", "points.txt", "targets.column"):
        os.chmod(os.path.join(d, fn), 0o644)
    with open(os.path.join(d, "simulator.ini"), "w") as f:
        f.write(INI.format(
            shape_line="shape = heart.matrix", excitation_line="", excitation_out="", start=100, length=400,
            nhood="3D4", dt=1, interp="endo-mid-epi", mid=0.6, base=BASE, kmin=KMIN, kmax=KMAX, cmp=2,
            endo_epi=-0.5, peak=0, move=1, fa=0, fa_delay=10.0, fa_limit=2))
    args = ["ignored", "-sim", TESTRUN_X, "-out", "result", "-out", "layer_aps"]
    with open(os.path.join(d, "args.txt"), "w") as f:
        f.write("\n".join(args) + "\n")
    # activation times of the reference's Dijkstra on this heart
    with open(os.path.join(d, "conduction.matrix")) as f:
        f.readline()
        hdr = f.readline().split()
        T = np.array(f.read().split(), float).reshape(int(hdr[3]), int(hdr[1]))
    act = O.ref_excite(layer, seed, T, os.path.join(d, "tmp_transfer.matrix"))
    os.remove(os.path.join(d, "tmp_transfer.matrix"))
    myo = layer > 0
    with open(os.path.join(d, "expected", "activation.txt"), "w") as f:
        f.write(hashlib.sha256(np.ascontiguousarray(act, np.float64).tobytes()).hexdigest() + "\n")
        f.write(f"{int(myo.sum())} {int((act[myo] == 0).sum())} {float(act.max())!r}\n")
    before = set(os.listdir(d))
    r = subprocess.run([O.REF_CLI] + args, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=7200)
    m = re.search(r"criteria = <([^>]*)>, violation = (\S+)", r.stdout)
    if not m:
        raise SystemExit(f"{name}: reference run failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    with open(os.path.join(d, "expected", "criteria.txt"), "w") as f:
        f.write(m.group(1) + "\n" + m.group(2) + "\n")
    for fn in sorted(set(os.listdir(d)) - before):
        shutil.move(os.path.join(d, fn), os.path.join(d, "expected", fn))
    os.remove(os.path.join(d, "heart.matrix"))
    size = sum(os.path.getsize(os.path.join(dp, f)) for dp, _, fs in os.walk(d) for f in fs)
    print(f"{name}: criteria <{m.group(1)}> violation {m.group(2)}; {size} bytes")


def main():
    if not os.path.exists(O.REF_CLI):
        raise SystemExit("oracle/_ref/ekgSim is not built (make -C oracle ref)")
    if sys.argv[1:] == ["model24"]:
        case_testrun()
        return
    shutil.rmtree(OUT, ignore_errors=True)
    # 3 layers, endo-mid-epi, moved electrodes, 1 - Pearson, start 100 (the
    # targets are not shifted: resampling factor 1), excitation exported
    case("heart3", n=24, start=100, length=400, out=("result", "layer_aps"),
         excitation_out="excitation_out.matrix")
    # 6 layers, endo-epi with transmural shape matching, RMS with peak
    # positions, endo-epi criterion, fast-approximation criterion, 26
    # neighbours, 0.5 ms step (targets resampled), fixed electrodes, cell APs
    case("heart6", n=20, n_layers=6, heart_seed=4, n_leads=3, nhood="3D8", dt=0.5, start=20, length=300,
         interp="endo-epi", cmp=1, peak=1, endo_epi=5.0, fa=1, move=0,
         out=("result", "layer_aps", "cell_aps 0 5 17 40"), x_seed=1001)
    # built-in 2-D ring heart (12 layers, 10 of them matched), 8 in-plane
    # neighbours, vector correlation, one target without a time column for
    # two leads (one criterion)
    case("ring2d", shape=False, n_leads=2, nhood="2D8", length=350, cmp=4, conduction_size=13,
         target_time=False, out=("result", "layer_aps"), x_seed=1002)
    # excitation sequence read from a file (rounded to 3 decimals), the
    # deviation-from-linear comparison and 2-D corner neighbours on a flat heart
    exc = os.path.join(OUT, "heart3", "expected", "excitation_out.matrix")
    case("heart3_excitation_file", n=24, start=100, length=400, cmp=3, excitation_in="excitation_in.matrix",
         extra_files=[(exc, "excitation_in.matrix")], out=("result",), x_seed=1003)
    # fast approximation limit below the approximation's criterion: the
    # simulation is skipped and the approximation is reported
    case("heart3_approx_only", n=20, heart_seed=6, fa_limit=0.05, out=("result",), x_seed=1004)
    if os.path.isdir(TESTRUN):
        case_testrun()


if __name__ == "__main__":
    main()

"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs in the build container only (it needs oracle/_ref, built from the
reference sources by `make -C oracle ref`).  Every fixture holds inputs and
the reference's outputs for them; nothing of the reference's source is
stored.  Regenerate with:

    python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ekgsim_amd import synth as S  # noqa: E402
from oracle import oracle as O  # noqa: E402


def save(name, **arrs):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrs)
    print(f"{name}: {os.path.getsize(path)} bytes")


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    rng = np.random.default_rng(2024)
    tmp = tempfile.mkdtemp()

    # 1. WohlfartPlus values (Wohlfart.h:195-203) over a spread of coefficients and times
    ks, ts, vs = [], [], []
    for r in range(24):
        k = np.array(S.BASE_AP, float)
        k[[5, 6, 7, 8]] = rng.uniform(S.FREE_LO, S.FREE_HI)
        if r % 6 == 1:
            k[0] = -86.0
        if r % 6 == 2:
            k[3] = 1.0
        t = np.concatenate([np.linspace(-50, 700, 31), [k[8], 0.0, 0.5]])
        ks.append(k)
        ts.append(t)
        vs.append([O.ref_wohlfart(k, x) for x in t])
    save("ref_wohlfart.npz", k=np.array(ks), t=np.array(ts), v=np.array(vs))

    # 2. excitation sequence, FP64 reference semantics (simulator.cpp:212-286)
    cases = {}
    h = S.make_heart(16, seed=5)
    cases["h16"] = (h.layer, h.seed, S.layered_transfer(3, 12, 1.0, 1.5))
    h = S.make_heart(22, seed=8)
    T = np.full((12, 12), -1.0)
    T[1:, 1:] = rng.uniform(0.1, 3.0, size=(11, 11))
    cases["h22_irregular"] = (h.layer, h.seed, T)
    h = S.make_heart(48, seed=2, dims=(1, 48, 48))
    seed = h.seed.copy()
    if seed.sum() == 0:
        seed.ravel()[np.flatnonzero(h.layer.ravel())[0]] = 1
    cases["h2d"] = (h.layer, seed, S.layered_transfer(3, 12, 1.0, 2.0))
    for name, (lay, sd, tr) in cases.items():
        act = O.ref_excite(lay, sd, tr, os.path.join(tmp, "t.matrix"))
        save(f"ref_ca_{name}.npz", layer=lay, seed=sd, transfer=tr, act=act)

    # 3. Simulation::run (simulator.cpp:452-550), literal reference ECGs
    h = S.make_heart(16, seed=5)
    act16 = O.ref_excite(h.layer, h.seed, cases["h16"][2], os.path.join(tmp, "t.matrix"))
    runs = [
        ("r16_3d4", h.layer, act16, "3D4", 0.0, 600.0, 1.0, 2, 0),
        ("r16_cube", h.layer, act16, "3D8", 0.0, 300.0, 1.0, 3, 0),
        ("r16_start", h.layer, act16, "3D4", 100.0, 400.0, 1.0, 2, 0),
        ("r16_dt025", h.layer, act16, "3D4", 0.0, 120.0, 0.25, 2, 0),
        ("r16_k0", h.layer, act16, "3D4", 0.0, 600.0, 1.0, 2, 1),
    ]
    h2 = cases["h2d"]
    act2 = O.ref_excite(h2[0], h2[1], h2[2], os.path.join(tmp, "t.matrix"))
    runs += [("r2d_2d8", h2[0], act2, "2D8", 0.0, 400.0, 1.0, 2, 0),
             ("r2d_2d4", h2[0], act2, "2D4", 0.0, 400.0, 1.0, 2, 0)]
    for i, (name, lay, act, nh, st, ln, dt, L, k0) in enumerate(runs):
        x = S.random_params(3, L, 300 + i)
        k = S.layer_k_from_params(x, 3)
        if k0:
            k[:, 0] = [-86.0, -84.0, -80.0]
        n = lay.shape[1]
        pos = O.move_leads(S.lead_positions(n, L, dims=lay.shape), S.displacements(x, 3, L))
        ecg = O.ref_run(lay, act, k, pos, nh, st, ln, dt)
        save(f"ref_run_{name}.npz", layer=lay, act=act, layer_k=k, pos=pos,
             nhood=np.array(nh), start=st, length=ln, dt=dt, ecg=ecg)

    # 4. EkgSim facade pipeline with lead displacement (sim_lib.h:131-252)
    h = S.make_heart(16, seed=6)
    d = tempfile.mkdtemp()
    S.write_matrix(os.path.join(d, "heart.matrix"), h.layer, h.seed)
    S.write_transfer(os.path.join(d, "conduction.matrix"), S.layered_transfer(3, 12, 1.0, 1.5))
    pos0 = S.lead_positions(16, 2)
    S.write_points(os.path.join(d, "points.txt"), pos0)
    with open(os.path.join(d, "simulator.ini"), "w") as f:
        f.write("[model]\nshape = heart.matrix\ntransfer = conduction.matrix\npoints = points.txt\n"
                "[simulation]\nstart = 0\nlength = 500\nneighbourhood type = 3D4\ntime step = 1\n")
    x = S.random_params(3, 2, 77)
    k = S.layer_k_from_params(x, 3)
    disp = S.displacements(x, 3, 2)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        ecg, act = O.ref_pipeline("simulator.ini", disp.ravel(), k, 2, 500, act_size=h.layer.size)
    finally:
        os.chdir(cwd)
    pts_read = np.loadtxt(os.path.join(d, "points.txt"), delimiter=",")[:, ::-1].copy()
    save("ref_pipeline_16.npz", layer=h.layer, seed=h.seed, transfer=S.layered_transfer(3, 12, 1.0, 1.5),
         pos_file=pts_read, disp=disp, layer_k=k, act=act.reshape(h.layer.shape), ecg=ecg)

    # 5. application model code (sim.cpp:72-313, 751-916; vectorMath.h:179-404)
    if O.have_refapp():
        model_goldens(rng)


def model_goldens(rng):
    """Layer APs of the transmural interpolation, comparison values and
    target resampling, from the reference's own application code."""
    base = np.array(S.BASE_AP, float)
    rows = []
    for interp in (1, 2):
        for n_layers in (1, 2, 3, 4, 7, 24):
            for mid in (0.6, 0.0, 1.0, 0.35):
                if interp == 1 and mid != 0.6:
                    continue
                g = rng.uniform([3e-4, 0.01, 0.01, 300] * 3, [1e-3, 0.1, 0.1, 450] * 3)
                fk, mk, bk = base.copy(), base.copy(), base.copy()
                fk[5:9] = g[0:4]
                mk[5:9] = g[4:8]
                bk[5:9] = g[8:12] if interp == 2 else g[4:8]
                out = O.ref_interp_layers(interp, n_layers, mid, fk, mk, bk)
                rows.append((interp, n_layers, mid, g, out))
    save("ref_layer_aps.npz", interp=np.array([r[0] for r in rows]), n_layers=np.array([r[1] for r in rows]),
         mid=np.array([r[2] for r in rows]), genes=np.array([r[3] for r in rows]),
         layer_k=np.concatenate([r[4] for r in rows]))

    # comparisons of a simulated lead against a target, every mode, equal
    # and unequal lengths, a constant curve (zero variance)
    a_list, b_list, modes, offs, vals = [], [], [], [], []
    t = np.arange(400.0)
    for mode in (1, 2, 3, 4):
        for na, nb in ((400, 400), (400, 335), (250, 400)):
            a = 30 * np.exp(-((t[:na] - 200) / 40) ** 2) + rng.normal(0, 1, na)
            b = np.exp(-((t[:nb] - 210) / 45) ** 2) + 0.01 * rng.normal(0, 1, nb)
            off = 1.0 - b.min() / (b.max() - b.min())
            a_list.append(np.pad(a, (0, 400 - na), constant_values=np.nan))
            b_list.append(np.pad(b, (0, 400 - nb), constant_values=np.nan))
            modes.append((mode, na, nb))
            offs.append(off)
            vals.append(O.ref_compare(mode, a, b, off))
        c = np.full(300, 2.5)
        a_list.append(np.pad(c, (0, 100), constant_values=np.nan))
        b_list.append(np.pad(c * 0.5 + 1, (0, 100), constant_values=np.nan))
        modes.append((mode, 300, 300))
        offs.append(0.5)
        vals.append(O.ref_compare(mode, c, c * 0.5 + 1, 0.5))
    save("ref_compare.npz", a=np.array(a_list), b=np.array(b_list), shape=np.array(modes), offset=np.array(offs),
         value=np.array(vals))

    # target resampling (vectorMath.h:209-227)
    src = np.sin(np.arange(120) / 7.0) + 0.1 * rng.normal(size=120)
    res = {}
    for i, (factor, start) in enumerate(((1.0, 20), (2.0, 0), (2.0, 15), (0.5, 3), (4.0, 100), (1.0 / 3.0, 0))):
        res[f"factor_{i}"] = np.array([factor, start])
        res[f"out_{i}"] = O.ref_resample(src, factor, start)
    save("ref_resample.npz", src=src, **res)


if __name__ == "__main__":
    main()

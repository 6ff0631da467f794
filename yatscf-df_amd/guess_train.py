"""Training of the learned initial-guess network on densities the GPU SCF converges.

Data, family "alkane": seeded all-trans C_nH_(2n+2), n uniform in 4..20,
random torsions in +/-30 deg about the C-C bonds and 0.05 A jitter per
coordinate.  Family "flake": H-capped hexagonal graphene flakes C6H6, C24H12
and C54H18 with 0.05 A jitter.  Each geometry is converged by the TinySCF
engine on the GPU (cc-pVDZ / cc-pVDZ-RI, FP64, |dE| < 1e-10).
Samples: every atom pair within 6 A of every molecule, with the pair model's
features (host/guess_ml.c) and the projection of D_converged - D_SAD onto its
coefficients.  Network: 3 x 128 SiLU in FP32, trained on the GPU by the
library's forward / backward / Adam kernels (ygpu_mlp_trainer_*).  Evaluation:
SCF iterations from SAD against iterations from the learned guess on held-out
molecules (the C30H62 workload and fresh alkanes for "alkane", the C96H24
workload for "flake"), converged energies compared.

Usage:
  python guess_train.py [--family alkane|flake] [--geoms N] [--epochs 120] [--out FILE] [--report FILE]
"""
import argparse
import json
import os
import struct
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import geometry  # noqa: E402
import ygpu  # noqa: E402

NH = 128
XYZ = os.path.join(os.path.dirname(HERE), "tests", "golden", "xyz")


def weights_path(family):
    return os.path.join(HERE, "data", "guess_mlp_%s.bin" % family)


def log(*a):
    print(*a, flush=True)


def alkane_set(n_geoms, seed):
    """(n, seed) of each training geometry."""
    rng = np.random.RandomState(seed)
    return [(int(rng.randint(4, 21)), int(rng.randint(1, 2 ** 31 - 1))) for _ in range(n_geoms)]


def write_alkane(dirname, n, seed):
    syms, xyz = geometry.alkane(n, jitter=0.05, seed=seed, torsion=30.0)
    path = os.path.join(dirname, "c%dh%d_%d.xyz" % (n, 2 * n + 2, seed))
    geometry.write_xyz(path, syms, xyz)
    return path


def flake_set(n_geoms, seed):
    """(rings, seed): C6H6 : C24H12 : C54H18 in the ratio 3 : 3 : 1."""
    rng = np.random.RandomState(seed)
    rings = [0, 0, 0, 1, 1, 1, 2]
    return [(rings[k % len(rings)], int(rng.randint(1, 2 ** 31 - 1))) for k in range(n_geoms)]


def write_flake(dirname, rings, seed):
    syms, xyz = geometry.graphene(rings, jitter=0.05, seed=seed)
    path = os.path.join(dirname, "flake%d_%d.xyz" % (rings, seed))
    geometry.write_xyz(path, syms, xyz)
    return path


def training_set(family, n_geoms, seed, dirname):
    """(label, xyz path writer) of each training geometry."""
    if family == "alkane":
        return [("C%dH%d" % (n, 2 * n + 2), (lambda n=n, s=s: write_alkane(dirname, n, s)))
                for n, s in alkane_set(n_geoms, seed)]
    return [("flake%d" % r, (lambda r=r, s=s: write_flake(dirname, r, s))) for r, s in flake_set(n_geoms, seed)]


def held_out(family, seed, dirname):
    if family == "alkane":
        ev = [("c30h62", os.path.join(XYZ, "c30h62.xyz"), 1e-9)]
        for n, s in alkane_set(4, seed + 7):
            ev.append(("c%dh%d seed %d" % (n, 2 * n + 2, s), write_alkane(dirname, n, s), 1e-9))
        return ev
    return [("c96h24", os.path.join(XYZ, "c96h24.xyz"), 1e-8)]


def solve(xyz, weights=None, ene_tol=1e-10, niters=80):
    """Converge one molecule; returns the engine's results (the engine is closed)."""
    T = ygpu.TinySCF(xyz, niters=niters)
    try:
        T.L.tinyscf_set_guess_weights(T.T, (weights or "").encode())
        T.setd("ene_tol", ene_tol)
        T.seti("quirk_a2", 0)
        t0 = time.perf_counter()
        T.setup()
        D0 = T.mat("D_mat").copy()
        T.build_df_tensor()
        T.do_scf()
        dt = time.perf_counter() - t0
        out = dict(iters=T.geti("iter"), energy=T.getd("total_energy"), converged=T.geti("converged"),
                   guess_ml=T.geti("guess_ml"), seconds=dt, D0=D0, D=T.mat("D_mat").copy())
        out["model"] = ygpu.PairModel.of_engine(T)
        return out
    finally:
        T.close()


def samples(res):
    pm = res["model"]
    X = pm.features()
    Y, M = pm.targets(res["D"] - res["D0"])
    return X, Y, M, pm.types()


def norm_stats(X, Y, M, t):
    in_mean = X.mean(0)
    in_std = np.maximum(X.std(0), 1e-3)
    nt, no = ygpu.YSML_NTYPES, ygpu.YSML_NOUT
    out_mean = np.zeros((nt, no), dtype=np.float32)
    out_std = np.ones((nt, no), dtype=np.float32)
    for k in range(nt):
        sel = t == k
        if not np.any(sel):
            continue
        for q in range(no):
            m = M[sel, q] > 0
            if np.sum(m) > 1:
                v = Y[sel, q][m]
                out_mean[k, q] = v.mean()
                out_std[k, q] = max(float(v.std()), 1e-6)
    return in_mean.astype(np.float32), in_std.astype(np.float32), out_mean, out_std


def init_params(n_in, nh, n_out, seed):
    rng = np.random.RandomState(seed)
    parts = []
    for fi, fo in ((n_in, nh), (nh, nh), (nh, nh), (nh, n_out)):
        lim = np.sqrt(6.0 / (fi + fo))
        parts += [rng.uniform(-lim, lim, size=fi * fo), np.zeros(fo)]
    return np.concatenate(parts).astype(np.float32)


def write_weights(path, nh, in_mean, in_std, out_mean, out_std, params):
    with open(path, "wb") as f:
        f.write(b"YSML")
        f.write(struct.pack("<6i", 1, ygpu.YSML_NIN, nh, 3, ygpu.YSML_NOUT, ygpu.YSML_NTYPES))
        for a in (in_mean, in_std, out_mean, out_std, params):
            f.write(np.ascontiguousarray(a, dtype="<f4").tobytes())


def read_weights(path):
    with open(path, "rb") as f:
        assert f.read(4) == b"YSML"
        ver, n_in, nh, nl, n_out, nt = struct.unpack("<6i", f.read(24))
        rd = lambda n: np.frombuffer(f.read(4 * n), dtype="<f4").copy()  # noqa: E731
        in_mean, in_std = rd(n_in), rd(n_in)
        out_mean, out_std = rd(nt * n_out).reshape(nt, n_out), rd(nt * n_out).reshape(nt, n_out)
        params = rd(ygpu.mlp_param_count(n_in, nh, n_out))
    return dict(nh=nh, in_mean=in_mean, in_std=in_std, out_mean=out_mean, out_std=out_std, params=params)


def train(X, Y, M, t, epochs, batch, lr, seed, val_frac=0.1):
    in_mean, in_std, out_mean, out_std = norm_stats(X, Y, M, t)
    Xn = (X - in_mean) / in_std
    Yn = ((Y - out_mean[t]) / out_std[t]) * M
    n = len(X)
    rng = np.random.RandomState(seed)
    perm = rng.permutation(n)
    nval = int(n * val_frac)
    val, tr = perm[:nval], perm[nval:]
    params = init_params(ygpu.YSML_NIN, NH, ygpu.YSML_NOUT, seed)
    eng = ygpu.Engine(0)
    hist = []
    try:
        tr_ = ygpu.MLPTrainer(eng, NH, Xn[tr], Yn[tr], M[tr], params)
        va_ = ygpu.MLPTrainer(eng, NH, Xn[val], Yn[val], M[val], params) if nval else None
        try:
            for ep in range(epochs):
                lr_e = lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs))      # cosine decay
                l = tr_.epoch(batch, lr_e, seed * 1000 + ep)
                lv = None
                if va_ is not None and (ep % 10 == 9 or ep == epochs - 1):
                    # validation loss with the current parameters: a gradient call on the held-out set
                    p = tr_.params()
                    va_.close()
                    va_ = ygpu.MLPTrainer(eng, NH, Xn[val], Yn[val], M[val], p)
                    lv, _ = va_.grad(np.arange(nval))
                hist.append((ep, l, lv))
                if lv is not None:
                    log("epoch %3d  train loss %.5f  validation loss %.5f" % (ep, l, lv))
            params = tr_.params()
        finally:
            tr_.close()
            if va_ is not None:
                va_.close()
    finally:
        eng.close()
    return dict(nh=NH, in_mean=in_mean, in_std=in_std, out_mean=out_mean, out_std=out_std, params=params,
                history=hist)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--family", default="alkane", choices=["alkane", "flake"])
    ap.add_argument("--geoms", type=int, default=None, help="default 256 (alkane), 112 (flake)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None, help="default data/guess_mlp_<family>.bin")
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    if args.geoms is None:
        args.geoms = 256 if args.family == "alkane" else 112
    if args.out is None:
        args.out = weights_path(args.family)
    if ygpu.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    tmp = tempfile.mkdtemp(prefix="ysml_")
    Xs, Ys, Ms, Ts = [], [], [], []
    t0 = time.perf_counter()
    for k, (label, make) in enumerate(training_set(args.family, args.geoms, args.seed, tmp)):
        r = solve(make())
        if not r["converged"]:
            log("geometry %d (%s) did not converge; skipped" % (k, label))
            continue
        X, Y, M, t = samples(r)
        Xs.append(X); Ys.append(Y); Ms.append(M); Ts.append(t)
        if k % 16 == 15:
            log("converged %d / %d geometries (%.1f s)" % (k + 1, args.geoms, time.perf_counter() - t0))
    X, Y, M, t = np.vstack(Xs), np.vstack(Ys), np.vstack(Ms), np.concatenate(Ts)
    log("samples: %d atom pairs from %d molecules" % (len(X), len(Xs)))
    t1 = time.perf_counter()
    w = train(X, Y, M, t, args.epochs, args.batch, args.lr, args.seed)
    t2 = time.perf_counter()
    write_weights(args.out, w["nh"], w["in_mean"], w["in_std"], w["out_mean"], w["out_std"], w["params"])
    log("weights written to %s (training %.1f s)" % (args.out, t2 - t1))
    # held-out evaluation
    report = dict(family=args.family, samples=int(len(X)), molecules=len(Xs), data_seconds=t1 - t0,
                  train_seconds=t2 - t1, final_train_loss=w["history"][-1][1],
                  final_validation_loss=w["history"][-1][2], held_out=[])
    for name, path, tol in held_out(args.family, args.seed, tmp):
        if not os.path.exists(path):
            continue
        a = solve(path, None, ene_tol=tol)
        b = solve(path, args.out, ene_tol=tol)
        row = dict(molecule=name, sad_iters=a["iters"], learned_iters=b["iters"], iterations_saved=a["iters"] - b["iters"],
                   sad_seconds=a["seconds"], learned_seconds=b["seconds"], dE_Eh=b["energy"] - a["energy"],
                   learned_applied=bool(b["guess_ml"]), converged=bool(a["converged"] and b["converged"]))
        report["held_out"].append(row)
        log("held out %s: SAD %d iterations, learned %d (dE %.2e)" % (name, a["iters"], b["iters"],
                                                                       b["energy"] - a["energy"]))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    log(json.dumps(report))


if __name__ == "__main__":
    main()

"""Benchmark: DF-RHF time-to-converged energy and J/K build time per iteration.

Workload (default): linear alkane C10H22, all-trans zig-zag, C-C 1.54 A,
C-H 1.09 A, seeded +/-0.02 A jitter, cc-pVDZ / cc-pVDZ-RI (Cartesian: nbf 260,
naux 990, 41 occupied), Schwarz 1e-11, FP64 on the GPU.

One step = one complete SCF solve through the TinySCF host engine and the
device library, timed from after input parsing (basis shells and nuclei
already uploaded) to the passing convergence test (|dE| < 1e-11):
S/H/X, Schwarz screening, pair layout, SAD guess, DF tensor, SCF iterations.
Reported: value = seconds to the converged energy (lower is better), J/K
milliseconds per iteration (median over iterations >= 2), the converged
energy and its difference from the CPU oracle, a roofline entry for the
dominant per-iteration kernel (the X = C_occ^T B build on FP64 MFMA, timed
with HIP events around its single launch) and the CPU baseline (the oracle
restatement of the reference algorithm on the host cores, same workload).

With --precision mixed the timed solve starts in the mixed-precision mode
(K from BF16-split MFMA products with FP32 accumulation, incremental J from
the FP32 tensor) and switches to FP64 once the DIIS error is below 1e-5.
Unless --no-mixed-compare is given, one mixed solve of the same workload is
also run after the timed FP64 steps and reported under "mixed_vs_fp64"
(time, iterations, energy difference, max |dF| and max |d eps| at
convergence).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c10h22]
                       [--precision fp64|mixed] [--no-cpu-baseline] [--no-mixed-compare]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "yatscf-df_amd")
ORACLE = os.path.join(ROOT, "oracle")
XYZ = os.path.join(ROOT, "tests", "golden", "xyz")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

WORKLOADS = {
    "c10h22": "C10H22 linear alkane, seeded 0.02 A jitter, cc-pVDZ/cc-pVDZ-RI (cart. nbf 260, naux 990)",
    "water": "H2O, O-H 0.9572 A, 104.52 deg, cc-pVDZ/cc-pVDZ-RI (cart. nbf 25, naux 96)",
    "c54h18": "C54H18 circumcoronene, seeded 0.01 A jitter, cc-pVDZ/cc-pVDZ-RI (cart. nbf 900, naux 3834)",
    "c30h62": "C30H62 alkane, torsions +/-30 deg, 0.05 A jitter, cc-pVDZ/cc-pVDZ-RI (cart. nbf 760, naux 2910)",
    "c96h24": "C96H24 graphene flake, seeded 0.01 A jitter, cc-pVDZ/cc-pVDZ-RI (cart. nbf 1560, naux 6696)",
}
# reference defaults (DIIS rule of the reference, |dE| < 1e-11) for the small
# cases; the large ones converge to the tolerance their configuration names with
# the DIIS slot of the largest stored residual replaced
SETTINGS = {
    "water": dict(ene_tol=1e-11, quirk_a2=1, family=None),
    "c10h22": dict(ene_tol=1e-11, quirk_a2=1, family="alkane"),
    "c54h18": dict(ene_tol=1e-9, quirk_a2=0, family="flake"),
    "c30h62": dict(ene_tol=1e-9, quirk_a2=0, family="alkane"),
    "c96h24": dict(ene_tol=1e-8, quirk_a2=0, family="flake"),
}
FP64_MFMA_PEAK_TFLOPS = 78.6
FP64_VECTOR_PEAK_TFLOPS = 78.6                 # MI355X FP64 vector rate (same as the FP64 matrix rate)
CUS = 256
BF16_SPLIT_PEAK_TFLOPS = 2500.0 / 3.0   # dense BF16 MFMA peak, three products per FP32 product
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(workload, mixed):
    """HBM bytes per launch of the roofline kernel from its committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), or None."""
    name = "k_xbuild_%s_%s_traffic.md" % (workload, "mixed" if mixed else "fp64")
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1]), "profiles/" + name
    return None, None


def profiled_dominant(workload, mixed, nbf):
    """The longest-running kernel of one SCF of this workload from its committed rocprofv3
    kernel-trace summary (first row of the table, sorted by total time), or None.  For the
    one-workgroup tridiagonal reduction the achieved FP64 rate is given against one CU's
    FP64 vector peak (the kernel runs on one CU by design) and against the chip's."""
    name = "scf_%s_%s_kernel_trace.md" % (workload, "mixed_learned" if mixed else "fp64")
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| ") and not line.startswith("| kernel") and not line.startswith("|---"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            kern, calls, total_us, avg_us, pct = cells[0], int(cells[1]), float(cells[2]), float(cells[3]), float(cells[4])
            dom = {"kernel": kern, "share_of_kernel_time": pct / 100.0, "calls": calls, "avg_us": avg_us,
                   "source": "profiles/" + name}
            if kern.startswith("k_sytrd_reg"):
                flops = 4.0 * nbf ** 3 / 3.0                 # Householder tridiagonalisation
                ach = flops / (avg_us * 1e-6) / 1e12
                dom.update({"bound": "fp64 vector, one CU (latency of the n - 2 dependent steps)",
                            "flops_per_launch": flops, "achieved_TFLOPs": ach,
                            "one_cu_fp64_peak_TFLOPs": FP64_VECTOR_PEAK_TFLOPS / CUS,
                            "frac_of_one_cu": ach / (FP64_VECTOR_PEAK_TFLOPS / CUS),
                            "frac_of_chip": ach / FP64_VECTOR_PEAK_TFLOPS})
            return dom
    return None


def run_gpu_scf(xyz, niters, settings, mixed=False, fock_check=False, weights=None):
    import ygpu
    T = ygpu.TinySCF(xyz, niters=niters)        # parse + upload shells (untimed)
    T.L.tinyscf_set_guess_weights(T.T, (weights or "").encode())
    T.setd("ene_tol", settings["ene_tol"])
    T.seti("quirk_a2", settings["quirk_a2"])
    T.seti("mixed", int(mixed))
    t0 = time.perf_counter()
    T.setup()
    T.build_df_tensor()
    T.do_scf()
    dt = time.perf_counter() - t0
    n = T.geti("iter")
    hist = T.d("t_hist").reshape(-1, 6).copy()
    res = dict(seconds=dt, iters=n, energy=T.getd("total_energy"), converged=T.geti("converged"),
               hist=hist, t_df=T.d("t_df").copy(), nbf=T.geti("nbasfuncs"), naux=T.geti("df_nbf"),
               n_occ=T.geti("n_occ"), nnz=T.geti("num_bf_pair_scr"), switch=T.geti("mixed_switch_iter"),
               F=T.mat("F_mat").copy(), eigval=T.d("eigval").copy(), guess_ml=T.geti("guess_ml"))
    import numpy as np
    mask = T.i("bf_pair_mask").reshape(res["nbf"], res["nbf"])
    iu = np.triu_indices(res["nbf"])
    res["n_up"] = int(np.sum(mask[iu] >= 0))
    if fock_check:
        # one FP64 and one mixed Fock build on the converged density (untimed): the
        # arithmetic's own error in F and in the orbital energies of X^T F X
        X = T.mat("X_mat").copy()
        F64 = T.build_fock(ygpu.FOCK_FP64)
        Fmx = T.build_fock(ygpu.FOCK_MIXED)
        e64 = np.linalg.eigvalsh(X.T @ F64 @ X)
        emx = np.linalg.eigvalsh(X.T @ Fmx @ X)
        no = res["n_occ"]
        res["fock_check"] = {"max_abs_dF": float(np.max(np.abs(Fmx - F64))),
                             "max_abs_d_orbital_energy": float(np.max(np.abs(emx - e64))),
                             "max_abs_d_occupied_orbital_energy": float(np.max(np.abs(emx[:no] - e64[:no])))}
    T.close()
    return res


def run_cpu_baseline(xyz, niters, settings):
    import basis
    import oracle
    s = basis.System(os.path.join(PKG, "data", "cc-pvdz", "cc-pvdz.gbs"),
                     os.path.join(PKG, "data", "cc-pVDZ-RI", "cc-pVDZ-RI.gbs"), xyz)
    o = oracle.OracleSCF(s, niters=niters, ene_tol=settings["ene_tol"], quirk_a2=bool(settings["quirk_a2"]))
    t0 = time.perf_counter()
    o.setup()
    o.build_df_tensor()
    o.do_scf()
    dt = time.perf_counter() - t0
    return dict(seconds=dt, energy=o.getd("total_energy"), iters=o.geti("iter"), cores=oracle.get_threads(),
                mkl=oracle.is_mkl(), converged=o.geti("converged"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c10h22", choices=sorted(WORKLOADS))
    ap.add_argument("--niters", type=int, default=60)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--guess", default="sad", choices=["sad", "learned"],
                    help="initial guess of the timed solves: SAD, or SAD plus the learned pair correction "
                         "(network trained on the workload's family, yatscf-df_amd/data/guess_mlp_<family>.bin)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed-compare", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (torch copy kernel): the known-size "
                         "reference for scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                          "copies": 5}))
        return
    settings = SETTINGS[args.workload]
    mixed = args.precision == "mixed"
    weights = None
    if args.guess == "learned":
        if not settings["family"]:
            print("no learned-guess network for workload %s" % args.workload, file=sys.stderr)
            sys.exit(2)
        weights = os.path.join(PKG, "data", "guess_mlp_%s.bin" % settings["family"])
        if not os.path.exists(weights):
            print("learned-guess network %s missing (train it with yatscf-df_amd/guess_train.py)" % weights,
                  file=sys.stderr)
            sys.exit(2)
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the path is not "
              "sharded across devices (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import ygpu
    if ygpu.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    xyz = os.path.join(XYZ, args.workload + ".xyz")
    for w in range(args.warmup):
        r = run_gpu_scf(xyz, args.niters, settings, mixed, weights=weights)
        log("warmup %d: %.3f s, %d iterations, E = %.10f" % (w, r["seconds"], r["iters"], r["energy"]))
    runs = []
    for k in range(args.steps):
        # the FP64 Fock check runs after the timer of the last step stops
        check = (not mixed) and (not args.no_mixed_compare) and k == args.steps - 1
        r = run_gpu_scf(xyz, args.niters, settings, mixed, fock_check=check, weights=weights)
        runs.append(r)
        log("step %d: %.3f s, %d iterations, E = %.10f%s" % (k, r["seconds"], r["iters"], r["energy"],
            (", FP64 from iteration %d" % r["switch"]) if mixed else ""))
    secs = [r["seconds"] for r in runs]
    last = runs[-1]

    def rows(r, mixed_run, want_mixed):
        # per-iteration phase times; iterations >= 2 of the requested arithmetic
        h = r["hist"]
        sw = (r["switch"] if r["switch"] >= 0 else len(h)) if mixed_run else 0
        idx = [i for i in range(2, len(h)) if (i < sw) == want_mixed]
        return h[idx]

    hist = np.vstack([rows(r, mixed, mixed) for r in runs])
    if len(hist) == 0:
        hist = np.vstack([r["hist"][2:] for r in runs])
    # J and K run on two streams at once: the J/K figure is the wall time of the pair (start of
    # J to F assembled); j_ms and k_ms are each timed on their own stream beside the other
    jk_ms = float(np.median(hist[:, 5]) * 1e3)
    j_ms = float(np.median(hist[:, 0] + hist[:, 1]) * 1e3)
    k_ms = float(np.median(hist[:, 2] + hist[:, 3]) * 1e3)
    xb_ms = float(np.median(hist[:, 2]) * 1e3)
    syrk_ms = float(np.median(hist[:, 3]) * 1e3)
    nbf, naux, n_occ, nnz, n_up = last["nbf"], last["naux"], last["n_occ"], last["nnz"], last["n_up"]
    # X = C_occ^T B over both triangles of the pair list: 2 nnz n_occ naux flops per launch
    xb_flops = 2.0 * nnz * n_occ * naux
    achieved = xb_flops / (xb_ms * 1e-3) / 1e12
    peak = BF16_SPLIT_PEAK_TFLOPS if mixed else FP64_MFMA_PEAK_TFLOPS
    traffic, traffic_src = profiled_traffic(args.workload, mixed)
    # J reads the packed tensor twice (FP32 copy in the mixed phase) and D, J once
    j_bytes = 2.0 * n_up * naux * (4 if mixed else 8) + 16.0 * nbf * nbf
    out = {
        "metric": "SCF time-to-converged energy (s) + J/K build ms/iter; mixed vs FP64 at 1 uEh",
        "value": statistics.mean(secs),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": statistics.mean(secs) * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-split/f32-acc K, f32-tensor incremental J, then f64" if mixed else "f64",
        "data": "synthetic (seeded generated geometry; reference basis-set files)",
        "config": {"workload": WORKLOADS[args.workload],
                   "guess": ("SAD + learned pair correction" if last["guess_ml"] else "SAD"),
                   "precision": ("mixed until DIIS error < 1e-5, then FP64" if mixed else "FP64 (plain path)"),
                   "ene_tol": settings["ene_tol"],
                   "diis_replace": "largest residual ever (reference)" if settings["quirk_a2"] else
                                   "largest stored residual",
                   "basis": "cc-pVDZ", "aux_basis": "cc-pVDZ-RI", "nbf": nbf, "naux": naux, "n_occ": n_occ,
                   "pairs_both_triangles": nnz, "pairs_packed": n_up, "schwarz_tol": 1e-11},
        "jk_ms_per_iter": jk_ms,
        "j_ms_per_iter": j_ms,
        "k_ms_per_iter": k_ms,
        "iterations": last["iters"],
        "energy_Eh": last["energy"],
        "converged": bool(last["converged"]),
        "df_build_s": {"3c": last["t_df"][0], "2c": last["t_df"][1], "inv_sqrt": last["t_df"][2],
                       "df_gemm": last["t_df"][3]},
        "roofline": {"kernel": ("k_xbuild_bf3 (X_i = C_occ^T B rows of i, BF16-split MFMA 32x32x16)" if mixed
                               else "k_xbuild (X_i = C_occ^T B rows of i, FP64 MFMA 16x16x4)"),
                     "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                     "launch_ms": xb_ms, "flops_per_launch": xb_flops},
        "j_build": {"bound": "hbm", "ms": j_ms, "bytes_per_call": j_bytes,
                    "achieved_GBs": j_bytes / (j_ms * 1e-3) / 1e9, "peak_GBs": HBM_PEAK_GBS},
        "syrk_ms": syrk_ms,
        # the roofline entry above is the Fock build's MFMA kernel; at small nbf the longest-running
        # kernel of the step is the density eigensolver's one-workgroup reduction
        "dominant_kernel": profiled_dominant(args.workload, mixed, nbf),
    }
    if mixed:
        out["mixed_switch_iteration"] = last["switch"]
    if not args.no_mixed_compare:
        # the other arithmetic on the same workload, for the mixed-vs-FP64 comparison
        # mixed (with the timed run's guess) against the plain FP64 path (SAD guess)
        if mixed:
            fp, mx = run_gpu_scf(xyz, args.niters, settings, False, fock_check=True), last
        else:
            fp, mx = last, run_gpu_scf(xyz, args.niters, settings, True, weights=weights)
        if "fock_check" not in fp:
            fp = run_gpu_scf(xyz, args.niters, settings, False, fock_check=True)
        mh = rows(mx, True, True)
        out["mixed_vs_fp64"] = {
            "fp64_guess": "learned" if fp["guess_ml"] else "SAD",
            "mixed_guess": "learned" if mx["guess_ml"] else "SAD",
            "fp64_s": fp["seconds"], "mixed_s": mx["seconds"], "speedup": fp["seconds"] / mx["seconds"],
            "fp64_iters": fp["iters"], "mixed_iters": mx["iters"], "mixed_switch_iteration": mx["switch"],
            "dE_Eh": mx["energy"] - fp["energy"],
            "fock_at_converged_density": fp["fock_check"],
            "mixed_k_ms_per_iter": float(np.median(mh[:, 2] + mh[:, 3]) * 1e3) if len(mh) else None,
            "mixed_j_ms_per_iter": float(np.median(mh[:, 0] + mh[:, 1]) * 1e3) if len(mh) else None,
            "mixed_xbuild_tflops": (xb_flops / (float(np.median(mh[:, 2])) * 1e12)) if len(mh) else None,
            "converged": bool(mx["converged"] and fp["converged"]),
        }
        log("mixed vs FP64: %s" % json.dumps(out["mixed_vs_fp64"]))
    if not args.no_cpu_baseline:
        cb = run_cpu_baseline(xyz, args.niters, settings)
        out["cpu_baseline"] = {"value": cb["seconds"], "unit": "s", "cores": cb["cores"], "kind": "port",
                               "sample": "one full SCF of the same workload (setup, DF build, %d iterations) by the "
                                         "oracle (C/OpenMP + %s)" % (cb["iters"], "MKL" if cb["mkl"] else "built-in BLAS")}
        out["energy_vs_oracle_Eh"] = last["energy"] - cb["energy"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: DF-RHF time-to-converged energy, mixed precision plus learned guess
against the plain FP64 path, with the J/K build time per iteration.

Workload (default): the C96H24 graphene flake (hexagonal lattice, C-C 1.42 A,
C-H 1.09 A, planar, seeded +/-0.01 A jitter), cc-pVDZ / cc-pVDZ-RI in the
reference's Cartesian convention (nbf 1560, naux 6696, 300 occupied),
Schwarz 1e-11, converged to |dE| < 1e-8 Eh between FP64 energies.

One step = one complete SCF solve through the TinySCF host engine and the
device library, timed from after input parsing (basis shells and nuclei
uploaded) to the passing convergence test: S/H/X, Schwarz screening, pair
layout, initial guess, DF tensor, SCF iterations.  The timed solves start from
the learned guess (SAD plus the flake network's pair correction; C96H24 is not
in its training set) and run the mixed-precision Fock builds (K from
BF16-split MFMA products with FP32 accumulation, incremental J from the FP32
tensor) until the DIIS error is small, then FP64 to convergence.
value = mean seconds of the timed solves (lower is better).

In the same run, after the timed steps: the plain FP64 GPU path from SAD (the
reference's arithmetic and guess) -- speedup, energy difference, iterations --
and the FP64 path from the learned guess (iterations saved by the guess
alone); one mixed Fock build on the FP64 solution gives the reduced-precision
error in F and in the orbital energies.  J ms/iter with GB/s against 8 TB/s,
K ms/iter with TF/s against the matrix-core peak.  The roofline entry is the
kernel with the most device time in a timed solve, its duration taken live
with HIP events on the stream it runs on.  cpu_baseline: the oracle (the
C/OpenMP + MKL restatement of the reference algorithm) on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c96h24]
                       [--precision mixed|fp64] [--guess learned|sad]
                       [--no-cpu-baseline] [--no-compare]
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
    "c96h24": "C96H24 graphene flake, seeded 0.01 A jitter, cc-pVDZ/cc-pVDZ-RI",
    "c10h22": "C10H22 linear alkane, seeded 0.02 A jitter, cc-pVDZ/cc-pVDZ-RI",
    "water": "H2O, O-H 0.9572 A, 104.52 deg, cc-pVDZ/cc-pVDZ-RI",
    "c54h18": "C54H18 circumcoronene, seeded 0.01 A jitter, cc-pVDZ/cc-pVDZ-RI",
    "c30h62": "C30H62 alkane, torsions +/-30 deg, 0.05 A jitter, cc-pVDZ/cc-pVDZ-RI",
}
# convergence tolerance of each configuration; the small cases keep the reference's DIIS rule
# and |dE| < 1e-11, the large ones replace the DIIS slot of the largest stored residual
SETTINGS = {
    "water": dict(ene_tol=1e-11, quirk_a2=1, family=None),
    "c10h22": dict(ene_tol=1e-11, quirk_a2=1, family="alkane"),
    "c54h18": dict(ene_tol=1e-9, quirk_a2=0, family="flake"),
    "c30h62": dict(ene_tol=1e-9, quirk_a2=0, family="alkane"),
    "c96h24": dict(ene_tol=1e-8, quirk_a2=0, family="flake"),
}
# the CPU baseline's workload: the largest configuration whose complete oracle SCF stays near
# half a minute on the host (C54H18 needs several minutes, C96H24 ~286 GB of host memory in the
# reference layout)
CPU_WORKLOAD = {"c96h24": "c30h62", "c54h18": "c30h62", "c30h62": "c30h62", "c10h22": "c10h22", "water": "water"}
FP64_PEAK_TFLOPS = 78.6                  # MI355X FP64 matrix (= vector) peak
BF16_PEAK_TFLOPS = 2500.0                # dense BF16 MFMA peak
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    mem = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal:"):
                mem = round(int(line.split()[1]) / 1048576.0, 1)
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "ram_GiB": mem}


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summaries
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size): the line
    "traffic_bytes_per_launch <kernel> <workload> ... = <bytes>" of a profiles/*_traffic.md file."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    for name in sorted(os.listdir(d)):
        if not name.endswith("_traffic.md"):
            continue
        for line in open(os.path.join(d, name)):
            t = line.split()
            if len(t) >= 4 and t[0] == "traffic_bytes_per_launch" and t[1] == kernel and t[2] == workload:
                return float(line.split("=")[1]), "profiles/" + name
    return None, None


FOCK_FP64, FOCK_MIXED, FOCK_MIXED6, FOCK_INCR = 0, 1, 2, 3
KIND_NAMES = {FOCK_FP64: "FP64", FOCK_MIXED: "BF16x3", FOCK_MIXED6: "BF16x6", FOCK_INCR: "FP64 reference + difference"}


def kinds(r):
    """Fock arithmetic of each SCF iteration of a solve (FP64 throughout for a plain solve)."""
    import numpy as np
    n = len(r["hist"])
    if not r["mixed"]:
        return np.zeros(n, dtype=int)
    return np.asarray(r["phases"][:n], dtype=int)


def phase_rows(r, kind, first=2):
    """Per-iteration phase timers (aux J, J, aux K, K, symm, J/K wall) of iterations >= first
    whose Fock build ran the given arithmetic."""
    import numpy as np
    h, k = r["hist"], kinds(r)
    idx = [i for i in range(first, len(h)) if k[i] == kind]
    return h[idx] if idx else np.zeros((0, 6))


def kernel_table(runs, mixed):
    """Device time per solve of the matrix-core and bandwidth kernels of the timed solves, from
    their HIP-event timers (each around its launches on its own stream): launches, mean
    duration, algorithmic work per launch and the rate against the peak."""
    import numpy as np
    last = runs[-1]
    nbf, naux, n_occ, nnz, n_up = last["nbf"], last["naux"], last["n_occ"], last["nnz"], last["n_up"]
    xb_flops = 2.0 * nnz * n_occ * naux                   # X = C_occ^T B over both triangles of the pairs
    syrk_flops = float(nbf) * (nbf + 1) * n_occ * naux     # K = X X^T, upper triangle
    out = {}

    def add(name, kind, per_launch_ms, launches_per_solve, work, peak, unit, arith):
        if not per_launch_ms:
            return
        ms = float(np.mean(per_launch_ms))
        rate = work / (ms * 1e-3) / (1e12 if unit == "TFLOP/s" else 1e9)
        out[name] = {"bound": kind, "arith": arith, "launches_per_solve": launches_per_solve, "launch_ms": ms,
                     "total_ms_per_solve": ms * launches_per_solve, "work_per_launch": work,
                     "achieved": rate, "peak": peak, "unit": unit, "frac": rate / peak}

    df_ms = [r["t_df"][3] * 1e3 for r in runs]
    if os.environ.get("YATSCF_METRIC", "cholesky") == "invsqrt":
        add("k_gemm128_nt (B = pqA J^-1/2, DF tensor)", "mfma", df_ms, 1, 2.0 * n_up * naux * naux,
            FP64_PEAK_TFLOPS, "TFLOP/s", "fp64")
    else:
        # B = pqA L^-T with L^-1 lower triangular: column q sums q + 1 products
        add("k_gemm128_nt (B = pqA L^-T, triangular, DF tensor)", "mfma", df_ms, 1,
            float(n_up) * naux * (naux + 1), FP64_PEAK_TFLOPS, "TFLOP/s", "fp64")
    for kind in (FOCK_MIXED, FOCK_MIXED6, FOCK_INCR, FOCK_FP64):
        rows = [phase_rows(r, kind) for r in runs]
        n = int(round(np.mean([int(np.sum(kinds(r) == kind)) for r in runs])))
        if n == 0:
            continue
        xb = [v for x in rows for v in x[:, 2] * 1e3]
        sy = [v for x in rows for v in x[:, 3] * 1e3]
        jj = [v for x in rows for v in (x[:, 0] + x[:, 1]) * 1e3]
        j32 = 2.0 * n_up * naux * 4 + 16.0 * nbf * nbf
        if kind in (FOCK_MIXED, FOCK_MIXED6):
            np_ = 3 if kind == FOCK_MIXED else 6
            peak = BF16_PEAK_TFLOPS / np_
            tag = "BF16x%d split" % np_
            # two parts: from the k-contiguous operand blocks; three: the gathering X build
            xname = "k_xbuild_ps" if kind == FOCK_MIXED else "k_xbuild_bf"
            add("%s (X = C_occ^T B, %s)" % (xname, tag), "mfma", xb, n, xb_flops, peak, "TFLOP/s", tag + ", f32 acc")
            # two parts: the panel engine over the panel-major X; three: 128 x 128 tiles
            sname = "k_pgemm_ps" if kind == FOCK_MIXED else "k_syrk_bf"
            add("%s (K = X X^T, %s)" % (sname, tag), "mfma", sy, n, syrk_flops, peak, "TFLOP/s", tag + ", f32 acc")
            add("J build (%s iterations, FP32 tensor, two passes)" % tag, "hbm", jj, n, j32, HBM_PEAK_GBS, "GB/s",
                "f32 tensor, f64 sums")
        elif kind == FOCK_INCR:
            peak = BF16_PEAK_TFLOPS / 3.0
            # X^P and X^M together: the X build over 2 n_occ columns; G = X^M (X^P)^T: every tile
            add("k_xbuild_ps (X^[P|M], 2 n_occ columns, BF16x3 split)", "mfma", xb, n, 2.0 * xb_flops, peak, "TFLOP/s",
                "BF16x3 split, f32 acc")
            add("k_pgemm_ps (G = X^M X^P^T, difference K, BF16x3 split)", "mfma", sy, n,
                2.0 * nbf * nbf * n_occ * naux, peak, "TFLOP/s", "BF16x3 split, f32 acc")
            add("J build (difference iterations, FP32 tensor, two passes)", "hbm", jj, n, j32, HBM_PEAK_GBS, "GB/s",
                "f32 tensor, f64 sums")
        else:
            add("k_xbuild (X = C_occ^T B, FP64)", "mfma", xb, n, xb_flops, FP64_PEAK_TFLOPS, "TFLOP/s", "fp64")
            add("k_syrk128 (K = X X^T, FP64)", "mfma", sy, n, syrk_flops, FP64_PEAK_TFLOPS, "TFLOP/s", "fp64")
            add("J build (FP64 tensor, two passes)", "hbm", jj, n, 2.0 * n_up * naux * 8 + 16.0 * nbf * nbf,
                HBM_PEAK_GBS, "GB/s", "fp64")
    return out


def dominant_kernel(kernels):
    """The kernel function with the most device time per solve, its launches of every phase
    together (a function that runs in two arithmetic phases, such as the X build of the
    reduced-precision and the difference iterations, is one kernel): rate = total algorithmic
    work / total time, against the peak of its arithmetic."""
    groups = {}
    for name, k in kernels.items():
        groups.setdefault(name.split()[0], []).append(k)
    fn = max(groups, key=lambda f: sum(k["total_ms_per_solve"] for k in groups[f]))
    ks = groups[fn]
    launches = sum(k["launches_per_solve"] for k in ks)
    ms = sum(k["total_ms_per_solve"] for k in ks)
    work = sum(k["work_per_launch"] * k["launches_per_solve"] for k in ks)
    unit, peak = ks[0]["unit"], ks[0]["peak"]
    rate = work / (ms * 1e-3) / (1e12 if unit == "TFLOP/s" else 1e9)
    return fn, {"bound": ks[0]["bound"], "achieved": rate, "peak": peak, "unit": unit, "frac": rate / peak,
                "launch_ms": ms / launches, "work_per_launch": work / launches, "launches_per_solve": launches,
                "total_ms_per_solve": ms, "phases": [n for n in kernels if n.split()[0] == fn]}


def run_cpu_baseline(name):
    import basis
    import oracle
    settings = SETTINGS[name]
    s = basis.System(os.path.join(PKG, "data", "cc-pvdz", "cc-pvdz.gbs"),
                     os.path.join(PKG, "data", "cc-pVDZ-RI", "cc-pVDZ-RI.gbs"), os.path.join(XYZ, name + ".xyz"))
    o = oracle.OracleSCF(s, niters=80, ene_tol=settings["ene_tol"], quirk_a2=bool(settings["quirk_a2"]))
    t0 = time.perf_counter()
    o.setup()
    o.build_df_tensor()
    o.do_scf()
    dt = time.perf_counter() - t0
    return dict(seconds=dt, energy=o.getd("total_energy"), iters=o.geti("iter"), cores=oracle.get_threads(),
                mkl=oracle.is_mkl(), converged=o.geti("converged"), nbf=s.nbf, naux=s.naux)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c96h24", choices=sorted(WORKLOADS))
    ap.add_argument("--niters", type=int, default=80)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--basis-type", default="cartesian", choices=["cartesian", "spherical"],
                    help="Cartesian components (the reference's) or the spherical-harmonic basis (the sizes "
                         "quoted for cc-pVDZ / cc-pVDZ-RI: C96H24 nbf 1464, naux 5712) for every solve")
    ap.add_argument("--guess", default="learned", choices=["sad", "learned"],
                    help="initial guess of the timed solves: SAD, or SAD plus the learned pair correction "
                         "(network trained on the workload's family, yatscf-df_amd/data/guess_mlp_<family>.bin)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 comparison solves after the timed steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (torch copy kernel): the known-size "
                         "reference for scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the path is not "
              "sharded across devices (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
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
    if settings["family"]:
        weights = os.path.join(PKG, "data", "guess_mlp_%s.bin" % settings["family"])
    if args.guess == "learned":
        if not weights:
            print("no learned-guess network for workload %s" % args.workload, file=sys.stderr)
            sys.exit(2)
        if not os.path.exists(weights):
            print("learned-guess network %s missing (train it with yatscf-df_amd/guess_train.py)" % weights,
                  file=sys.stderr)
            sys.exit(2)
    import numpy as np
    import ygpu
    if ygpu.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    xyz = os.path.join(XYZ, args.workload + ".xyz")
    kw = dict(niters=args.niters, ene_tol=settings["ene_tol"], quirk_a2=settings["quirk_a2"],
              basis_type=args.basis_type)
    timed_weights = weights if args.guess == "learned" else None

    def solve(mx, w, **extra):
        r = ygpu.solve(xyz, mixed=mx, weights=w, **kw, **extra)
        r["mixed"] = mx
        return r

    for w in range(args.warmup):
        r = solve(mixed, timed_weights)
        log("warmup %d: %.3f s, %d iterations, E = %.10f" % (w, r["seconds"], r["iters"], r["energy"]))
    runs = []
    for k in range(args.steps):
        r = solve(mixed, timed_weights)
        runs.append(r)
        log("step %d: %.3f s (setup %.3f, DF %.3f, SCF %.3f), %d iterations%s, E = %.10f"
            % (k, r["seconds"], r["setup_s"], r["df_s"], r["scf_s"], r["iters"],
               (", FP64 from iteration %d" % r["switch"]) if mixed else "", r["energy"]))
    secs = [r["seconds"] for r in runs]
    last = runs[-1]
    # J/K per iteration: of the reduced-precision iterations of a mixed solve, else of all
    hist = np.vstack([phase_rows(r, FOCK_MIXED if mixed else FOCK_FP64) for r in runs])
    if len(hist) == 0:
        hist = np.vstack([r["hist"][2:] for r in runs])
    # the J/K figure is the wall time of the pair (start of J to F assembled); J runs first, each
    # kernel with the whole GPU (YGPU_J_BESIDE=1 puts it on a second stream beside the K build)
    jk_ms = float(np.median(hist[:, 5]) * 1e3)
    j_ms = float(np.median(hist[:, 0] + hist[:, 1]) * 1e3)
    k_ms = float(np.median(hist[:, 2] + hist[:, 3]) * 1e3)
    nbf, naux, n_occ, nnz, n_up = last["nbf"], last["naux"], last["n_occ"], last["nnz"], last["n_up"]
    kernels = kernel_table(runs, mixed)
    dom_name, dom = dominant_kernel(kernels)
    traffic, traffic_src = profiled_traffic(dom_name, args.workload)
    if mixed:
        dtype = ("bf16x3-split products with f32 accumulation (K), f32 tensor with f64 sums (J), then f64 "
                 "(full builds and difference builds against them)")
    else:
        dtype = "f64"
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
        "dtype": dtype,
        "data": "synthetic (seeded generated geometry; reference basis-set files)",
        "config": {"workload": WORKLOADS[args.workload],
                   "guess": ("SAD + learned pair correction" if last["guess_ml"] else "SAD"),
                   "precision": ("BF16x3-split K (FP32 accumulation) and FP32-tensor incremental J until the DIIS "
                                 "error reaches the split's floor; then one full FP64 reference build and "
                                 "difference-density builds K = K_ref + K(D - D_ref) (BF16x3 products of "
                                 "aligned orbitals), whose energies are exact to second order in D - D_ref; "
                                 "converged when two FP64 or difference-build energies agree; densities by "
                                 "SP2 purification (FP64), the last by the eigensolver" if mixed else
                                 "FP64 (plain path)"),
                   "ene_tol": settings["ene_tol"],
                   "diis_replace": "largest residual ever (reference)" if settings["quirk_a2"] else
                                   "largest stored residual",
                   "basis": "cc-pVDZ", "aux_basis": "cc-pVDZ-RI", "basis_type": args.basis_type,
                   "nbf": last["nbf_sph"], "nbf_cartesian_components": nbf, "naux": naux, "n_occ": n_occ,
                   "pairs_both_triangles": nnz, "pairs_packed": n_up, "schwarz_tol": 1e-11},
        "iterations": last["iters"],
        "iterations_by_arithmetic": {KIND_NAMES[k]: int(np.sum(kinds(last) == k)) for k in KIND_NAMES
                                     if np.sum(kinds(last) == k)},
        "mixed_switch_iteration": last["switch"] if mixed else None,
        "energy_Eh": last["energy"],
        "converged": bool(last["converged"]),
        "phases_s": {"setup_and_guess": statistics.mean(r["setup_s"] for r in runs),
                     "df_build": statistics.mean(r["df_s"] for r in runs),
                     "scf": statistics.mean(r["scf_s"] for r in runs)},
        "df_build_s": {"3c": last["t_df"][0], "2c": last["t_df"][1], "metric_factor": last["t_df"][2],
                       "df_gemm": last["t_df"][3]},
        "jk_ms_per_iter": jk_ms,
        "j_ms_per_iter": j_ms,
        "k_ms_per_iter": k_ms,
        "roofline": {"kernel": dom_name, "bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"],
                     "unit": dom["unit"], "frac": dom["frac"], "traffic": traffic, "traffic_source": traffic_src,
                     "launch_ms": dom["launch_ms"], "work_per_launch": dom["work_per_launch"],
                     "launches_per_solve": dom["launches_per_solve"], "total_ms_per_solve": dom["total_ms_per_solve"],
                     "phases": dom["phases"]},
        "kernels": kernels,
        "host": host_info(),
    }
    if not args.no_compare:
        # the plain FP64 path from SAD (the reference's arithmetic and guess), with one mixed Fock
        # build on its converged density; then FP64 from the learned guess (the guess alone)
        fp = solve(False, None, fock_check=True)
        log("FP64 from SAD: %.3f s, %d iterations, E = %.10f" % (fp["seconds"], fp["iters"], fp["energy"]))
        cmp = {"fp64_sad_s": fp["seconds"], "fp64_sad_iterations": fp["iters"], "fp64_sad_energy_Eh": fp["energy"],
               "speedup": fp["seconds"] / out["value"], "dE_Eh": last["energy"] - fp["energy"],
               "iterations_saved": fp["iters"] - last["iters"],
               "fock_at_converged_density": fp["fock_check"],
               "converged": bool(fp["converged"] and last["converged"])}
        fph = phase_rows(fp, FOCK_FP64)
        if len(fph):
            cmp["fp64_jk_ms_per_iter"] = float(np.median(fph[:, 5]) * 1e3)
            cmp["fp64_k_ms_per_iter"] = float(np.median(fph[:, 2] + fph[:, 3]) * 1e3)
            cmp["fp64_j_ms_per_iter"] = float(np.median(fph[:, 0] + fph[:, 1]) * 1e3)
        if weights and os.path.exists(weights):
            gl = solve(False, weights)
            log("FP64 from the learned guess: %.3f s, %d iterations, E = %.10f"
                % (gl["seconds"], gl["iters"], gl["energy"]))
            cmp["fp64_learned_s"] = gl["seconds"]
            cmp["fp64_learned_iterations"] = gl["iters"]
            cmp["iterations_saved_by_guess_alone"] = fp["iters"] - gl["iters"]
            cmp["fp64_learned_dE_Eh"] = gl["energy"] - fp["energy"]
        out["vs_fp64"] = cmp
        log("vs FP64: %s" % json.dumps(cmp))
    if not args.no_cpu_baseline:
        cw = CPU_WORKLOAD[args.workload]
        cb = run_cpu_baseline(cw)
        sample = ("one complete SCF of the %s workload (nbf %d, naux %d; setup, DF build, %d iterations to "
                  "|dE| < %g) by the oracle (C/OpenMP + %s)" % (cw.upper(), cb["nbf"], cb["naux"], cb["iters"],
                                                               SETTINGS[cw]["ene_tol"],
                                                               "MKL" if cb["mkl"] else "built-in BLAS"))
        if cw != args.workload:
            sample += ("; the largest configuration whose complete oracle SCF stays near half a minute here "
                       "(%s needs far more time and host memory in the reference layout)" % args.workload.upper())
        out["cpu_baseline"] = {"value": cb["seconds"], "unit": "s", "cores": cb["cores"], "kind": "port",
                               "sample": sample, "workload": cw, "energy_Eh": cb["energy"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

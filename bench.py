#!/usr/bin/env python3
"""Benchmark: seconds to FP64 relative residual 1e-10 for the 3-D Poisson solve
at 512^3 interior unknowns on one MI355X -- learned-preconditioned FCG with
FP32 inner vectors and FP64 outer refinement, against plain FP64 CG.

Workload: unit cube, 7-point FD, homogeneous Dirichlet, f = seeded sum of 64
Gaussian sources (centres U[0,1]^3, widths U[0.01,0.1], amplitudes N(0,1)),
x0 = 0, stop at relative residual 1e-10 (baseCG.hpp:188,224); every form's
true FP64 residual ||b - A x|| / ||b_adj|| is checked after the solve.

A step is one complete solve with the right-hand side resident in HBM: reset
x = 0, restore b, normalise, iterate, true FP64 residual, un-normalise.  Each
form is timed over K steps after W warm-up steps, bracketed by device
synchronisation.  Forms:
  - ir_nnfcg (headline): FP64 outer residual/update refinement around FCG in
    FP32 preconditioned by the learned network (BF16 MFMA convolutions), inner
    tolerance 1e-4, 2e-3 for the first outer step: the first FP32 correction
    leaves a true residual of about 2e-3 (kappa * eps32 at 512^3 is ~6e-3)
    however far the inner solve goes;
  - nn_fcg_fp64: the same network under FP64 FCG;
  - plain_fp64_cg: the reference method.
The network is the committed one (parallelpoissonsolver_amd/weights/nn_64.npy,
trained on the GPU by train_nn.py on 64^3 CG iterates); without it the bench
trains one first (about a minute, outside the timed region).

Roofline: every network layer and Krylov kernel is timed live with HIP events
on the context's stream; the dominant kernel of the headline solve (largest
time per iteration) is reported against 8 TB/s HBM.  CPU baseline: the
reference CPU solver (oracle/_ref, built from the reference sources) under MPI
on the host's cores, on a bounded sample of the same workload.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_INTERIOR = 512
SEED = 3
NSRC = 64
TOL = 1e-10
FINISH = 10  # true-residual finish of the CG/FCG forms: at most this many residual replacements
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
WEIGHTS = os.path.join(ROOT, "parallelpoissonsolver_amd", "weights", "nn_64.npy")
LAYER_NAMES = ["lift", "b0", "b1", "c0", "c1", "c2", "c3", "c4", "c5", "c6", "c7", "e0", "e1", "out"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def chains():
    import parallelpoissonsolver_amd._lib as L
    return {
        "plain_fp64_cg": [dict(kind=L.CG, tol=TOL, maxit=200000, true_finish=FINISH), dict(kind=L.NONE)],
        "ir_fp32cg_fp64": [dict(kind=L.IR, tol=TOL, maxit=100), dict(kind=L.CG, precision=32, tol=1e-5, maxit=200000),
                           dict(kind=L.NONE, precision=32)],
        "nn_fcg_fp64": [dict(kind=L.FCG, tol=TOL, maxit=20000, true_finish=FINISH),
                        dict(kind=L.NN, stall_window=50, stall_ratio=0.9)],
        "ir_nnfcg": [dict(kind=L.IR, tol=TOL, maxit=100, first_tol=2e-3, adapt=0.5),
                     dict(kind=L.FCG, precision=32, tol=1e-4, maxit=20000),
                     dict(kind=L.NN, precision=32, stall_window=50, stall_ratio=0.9)],
        # the reference's default GPU method (solverPoissonMPI_alpaka/include/inputParam.hpp:36,
        # solverSetup.hpp:58-73): BiCGSTAB + Chebyshev(24, global eigenvalues rescaled by
        # 100 and 1 - 1e-4, epsilon 0) in FP64
        "bicgstab_cheb24_fp64": [dict(kind=L.BICGSTAB, tol=TOL, maxit=20000, true_finish=FINISH),
                                 dict(kind=L.CHEBYSHEV, maxit=24, cheb_eps=0.0, cheb_rmin=100.0,
                                      cheb_rmax=1 - 1e-4)],
    }


def run_form(C, chain, steps, warmup, name):
    import parallelpoissonsolver_amd._lib as L
    res = None
    for w in range(warmup):
        C.fill(L.FIELD_X, 0.0)
        C.copy_field(L.FIELD_B, L.FIELD_T3)
        res = C.solve(chain)
        log(f"{name}: warm-up {w + 1}/{warmup}: {res['iterations']} it, {res['seconds']:.3f} s device")
    dev, wall = [], []
    for s in range(steps):
        C.fill(L.FIELD_X, 0.0)
        C.copy_field(L.FIELD_B, L.FIELD_T3)
        C.synchronize()
        t0 = time.perf_counter()
        res = C.solve(chain)
        C.synchronize()
        wall.append(time.perf_counter() - t0)
        dev.append(res["seconds"])
        log(f"{name}: step {s + 1}/{steps}: {res['iterations']} it (outer {res['outer_iterations']}), "
            f"true rel residual {res['error_compute_operator']:.3e}, fallback {res['fallback_switches']}, "
            f"{wall[-1]:.3f} s")
        # every form must reach a true FP64 relative residual below TOL: the
        # refinement forms stop on it; CG and FCG stop on their recursive
        # residual as the reference does (baseCG.hpp:224) and then run the
        # true-residual finish (r = b - A x replaces the drifted recursive
        # residual and the loop resumes until the true one is below TOL)
        if not res["error_compute_operator"] < TOL:
            raise RuntimeError(f"{name} did not reach {TOL}: {res['error_compute_operator']}")
    return {"s": sum(wall) / len(wall), "device_s": sum(dev) / len(dev), "iterations": res["iterations"],
            "outer_iterations": res["outer_iterations"], "true_rel_residual": res["error_compute_operator"],
            "fallback_switches": res["fallback_switches"]}


KERNELS = {"cg_stencil_f64": 0, "cg_update_f64": 1, "residual_f64": 2, "copy_f64": 3, "cheb_sweep_f64": 4,
           "cg_stencil_f32": 5, "cg_update_f32": 6, "fcg_stencil_f64": 7, "fcg_stencil_f32": 8,
           "copy_flat_f64": 9, "cheb_pair_f64": 10}  # PPS_K_* ids of include/pps_abi.h


def profiled_traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    counter summaries (scaled by a plain copy kernel), or None."""
    for name in (f"krylov_{n}_hbm_traffic.md", f"nn_{n}_hbm_traffic.md"):
        path = os.path.join(ROOT, "profiles", name)
        try:
            for line in open(path):
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                    return float(parts[2])
        except OSError:
            pass
    return None


def krylov_kernels(C, reps):
    out = {}
    for name, kid in KERNELS.items():
        ms, nbytes = C.time_kernel(kid, reps)
        out[name] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9}
        log(f"kernel {name}: {ms:.4f} ms/launch, {out[name]['gbs']:.0f} GB/s algorithmic")
    return out


def nn_layers(C, reps):
    """Per-layer timing on FP32 fields, as the headline (mixed-precision) solve runs the network."""
    out = {}
    for i, name in enumerate(LAYER_NAMES):
        ms, nbytes, flops = C.nn_time_layer(i, reps, prec=32)
        out[name] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                     "tflops": flops / (ms * 1e-3) / 1e12}
        log(f"layer {name}: {ms:.4f} ms/launch, {out[name]['gbs']:.0f} GB/s, {out[name]['tflops']:.1f} TFLOP/s "
            f"({100 * out[name]['tflops'] / PEAK_BF16_TFLOPS:.1f}% of BF16 MFMA peak)")
    return out


def host_cpu():
    """CPU model and physical core count of this host (from /proc/cpuinfo)."""
    model, cores = "", set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and not model:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    if phys is not None and core is not None:
        cores.add((phys, core))
    return model, len(cores) or None


def cpu_baseline(gpu_iterations):
    """Reference CPU solver (compiled from the reference sources into oracle/_ref)
    on the host cores: 100 CG iterations of the same 512^3 workload under MPI,
    extrapolated to the iteration count the solve needs.  Fails loudly when the
    reference build or mpirun is missing (pass --no-cpu to skip the baseline)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_solver_sample")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    if not (os.path.exists(exe) and os.path.exists(mpirun)):
        raise RuntimeError(f"CPU baseline: {exe} or mpirun missing (build with __graft_entry__.build() where the "
                           f"reference sources are present, or pass --no-cpu)")
    npg = N_INTERIOR + 2
    ds = 1.0 / (npg - 1)
    ranks = (2, 2, 2)  # 514 = 2 * 257: the reference needs ranks dividing npglobal (blockGrid.hpp:165)
    nproc = ranks[0] * ranks[1] * ranks[2]
    with tempfile.TemporaryDirectory() as tmp:
        prefix = os.path.join(tmp, "cpu")
        cmd = [mpirun, "-n", str(nproc), exe, "--ranks", *map(str, ranks), "--np", str(npg), str(npg), str(npg),
               "--ds", repr(ds), repr(ds), repr(ds), "--problem", "2", "--rhs", "gauss", "--seed", str(SEED),
               "--nsrc", str(NSRC), "--solver", "cg", "--precond", "none", "--tol", repr(TOL), "--out", prefix]
        log("CPU baseline: " + " ".join(cmd[:3]) + " ... (reference CPU CG, 100 iterations)")
        t0 = time.perf_counter()
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if p.returncode != 0:
            raise RuntimeError(f"CPU baseline failed: {p.stderr[-500:]}")
        meta = json.load(open(prefix + ".json"))
    it = meta["iterations"]
    per_it = meta["solver_in_function_time"] / it
    setup = meta["solver_time"] - meta["solver_in_function_time"]
    est = setup + per_it * gpu_iterations
    log(f"CPU baseline: {it} iterations in {meta['solver_in_function_time']:.2f} s ({per_it:.3f} s/it), "
        f"extrapolated {est:.1f} s for {gpu_iterations} iterations (run took {wall:.1f} s)")
    cpu_model, phys = host_cpu()
    return {"value": est, "unit": "s", "cores": nproc, "kind": "reference",
            "sample": f"reference solverPoissonMPI_CPU BaseCG+NoneSolver (plain CG), 512^3 Gaussian-source rhs, "
                      f"{nproc} MPI ranks (2x2x2, one thread each: 514 = 2 * 257 admits no other layout), "
                      f"{it} iterations timed ({per_it:.4f} s/iteration) + setup, extrapolated to the "
                      f"{gpu_iterations} iterations of the full plain-CG solve",
            "cpu_model": cpu_model, "host_physical_cores": phys, "host_logical_cpus": os.cpu_count(),
            "iterations_timed": it, "seconds_per_iteration": round(per_it, 5)}


def network():
    import parallelpoissonsolver_amd._lib as L
    from parallelpoissonsolver_amd import train_nn
    if os.path.exists(WEIGHTS):
        net = L.Network(seed=0)
        info = train_nn.load(net, WEIGHTS)
        log(f"network: {WEIGHTS} ({net.n_effective} parameters, trained loss {info.get('loss_last', float('nan')):.4f})")
        return net, dict(info, source="committed weights")
    log("network: no committed weights, training on 64^3 CG iterates first")
    net, info = train_nn.train()
    log(f"network: trained in {info['generate_seconds'] + info['train_seconds']:.1f} s, loss {info['loss_last']:.4f}")
    return net, dict(info, source="trained in this run")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=N_INTERIOR, help="interior points per direction")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="ir_nnfcg,nn_fcg_fp64,plain_fp64_cg,bicgstab_cheb24_fp64")
    ap.add_argument("--reps", type=int, default=20, help="launches per kernel timing")
    ap.add_argument("--kernels-only", action="store_true", help="only the per-kernel timings")
    ap.add_argument("--rhs", default="gauss", choices=("gauss", "multiscale"),
                    help="right-hand side: 64 Gaussian sources (the metric's), or those plus a k^-2 random spectrum")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {a.gpus})", file=sys.stderr)
        sys.exit(2)

    import parallelpoissonsolver_amd._lib as L

    npg = a.n + 2
    ds = 1.0 / (npg - 1)
    C = L.Context((npg,) * 3, (ds,) * 3, (0, 0, 0), (0,) * 6)
    C.fill(L.FIELD_X, 0.0)
    C.fill(L.FIELD_B, 0.0)
    C.rhs_gauss(SEED, NSRC, 0.01, 0.1)
    if a.rhs == "multiscale":
        C.rhs_spectrum(SEED, kmax=16, accumulate=True)
    C.copy_field(L.FIELD_T3, L.FIELD_B)
    log(f"workload: {a.n}^3 interior, Gaussian sources (seed {SEED}, {NSRC} sources)"
        f"{' plus a k^-2 spectrum' if a.rhs == 'multiscale' else ''}, tol {TOL}")
    forms = a.forms.split(",")
    needs_nn = a.kernels_only or any(f in ("ir_nnfcg", "nn_fcg_fp64") for f in forms)
    net, net_info = network() if needs_nn else (None, None)
    if net is not None:
        net.attach(C)

    if a.kernels_only:
        # the per-kernel timings alone (short enough to run under hardware counters)
        krylov_kernels(C, a.reps)
        nn_layers(C, a.reps)
        C.close()
        return

    results = {}
    all_forms = chains()
    for name in forms:
        results[name] = run_form(C, all_forms[name], a.steps, a.warmup, name)

    headline = forms[0]
    kern = krylov_kernels(C, a.reps)
    layers = nn_layers(C, a.reps) if net is not None else {}
    # dominant kernel of the headline solve: largest device time per iteration.
    # Per learned-FCG iteration: one network apply (each layer once; the output
    # layer also sums gamma and the lift forms the residual update) and the
    # fused FCG p-stencil pass; per plain-CG iteration: the fused stencil and
    # update.
    if headline in ("ir_nnfcg", "nn_fcg_fp64"):
        per_it = {f"nn_{k}": v for k, v in layers.items()}
        sfx = "f32" if headline == "ir_nnfcg" else "f64"
        per_it[f"fcg_stencil_{sfx}"] = kern[f"fcg_stencil_{sfx}"]
    elif headline == "ir_fp32cg_fp64":
        per_it = {k: kern[k] for k in ("cg_stencil_f32", "cg_update_f32")}
    else:
        per_it = {k: kern[k] for k in ("cg_stencil_f64", "cg_update_f64")}
    dom = max(per_it, key=lambda k: per_it[k]["ms"])
    k = per_it[dom]
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(k["gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(k["gbs"] / PEAK_HBM_GBS, 4), "traffic": profiled_traffic(dom, a.n),
            "bytes_per_launch": k["bytes"], "ms_per_launch": round(k["ms"], 4)}
    if "tflops" in k:
        roof["mfma_tflops"] = round(k["tflops"], 2)
        roof["mfma_frac"] = round(k["tflops"] / PEAK_BF16_TFLOPS, 4)
    # the spec peak next to a measured one: this box's plain copy kernel
    if "copy_flat_f64" in kern:
        roof["peak_measured_copy"] = round(kern["copy_flat_f64"]["gbs"], 1)

    cpu = None
    if not a.no_cpu and a.n == N_INTERIOR and a.rhs == "gauss":
        ref_it = results["plain_fp64_cg"]["iterations"] if "plain_fp64_cg" in results else 2406
        cpu = cpu_baseline(ref_it)
    C.close()

    hv = results[headline]
    out = {
        "metric": "s to 1e-10 rel. residual at 512³: NN-FCG+FP32/FP64 vs plain FP64 CG; HBM GB/s",
        "value": round(hv["s"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(hv["s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"ir_nnfcg": "f32+f64+bf16", "nn_fcg_fp64": "f64+bf16", "ir_fp32cg_fp64": "f32+f64"}.get(headline, "f64"),
        "data": "synthetic",
        "config": {"workload": f"poisson3d_{a.n}^3_dirichlet_" + (f"gauss{NSRC}" if a.rhs == "gauss" else
                                                                  f"gauss{NSRC}_spectrum16"),
                   "n_interior": a.n, "rhs_seed": SEED,
                   "solver": headline, "tol": TOL},
        "forms": results,
        "ratio_headline_vs_plain": (round(hv["s"] / results["plain_fp64_cg"]["s"], 4)
                                    if "plain_fp64_cg" in results else None),
        "network": ({k2: net_info[k2] for k2 in ("source", "n", "pairs", "steps", "loss_last", "params_effective")
                     if k2 in net_info} if net_info else None),
        "kernels": {k2: {"gbs": round(v["gbs"], 1), "ms": round(v["ms"], 4)} for k2, v in kern.items()},
        "nn_layers": {k2: {"gbs": round(v["gbs"], 1), "tflops": round(v["tflops"], 2), "ms": round(v["ms"], 4)}
                      for k2, v in layers.items()},
        "roofline": roof,
        # Chebyshev sweeps at this size: one per launch (4 fields) and two fused per launch (5 fields)
        "chebyshev": ({"single_sweep_ms": round(kern["cheb_sweep_f64"]["ms"], 4),
                       "single_sweep_gbs": round(kern["cheb_sweep_f64"]["gbs"], 1),
                       "pair_ms": round(kern["cheb_pair_f64"]["ms"], 4),
                       "pair_gbs": round(kern["cheb_pair_f64"]["gbs"], 1),
                       "per_sweep_ms_fused": round(kern["cheb_pair_f64"]["ms"] / 2, 4)}
                      if "cheb_pair_f64" in kern else None),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

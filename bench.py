"""Benchmark of the projection step on one MI355X.

Workload: the 64-bubble swarm on a 2048x2048 grid (1x1 m box, liquid depth
0.8 m, density ratio 1000:1), the configuration the solver metric is quoted
on.  One step = one projection step from device-resident inputs: operator and
RHS build from the predictor fields (ComputePossionMatrixCoefficient /
SetPoissonMatrix / SetPoissonVectors), PCG + multigrid solve from x0 = 0 to a
relative residual of 1e-8, pressure scatter + ghosts, velocity correction and
mass residual (UpdatePUV's correction).  Every step redoes all of it.

Prints one JSON line.  ``value`` is Mcell-steps/s of the headline path (the
mixed-precision solve); the plain-FP64 path is timed beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ni", type=int, default=2048)
    ap.add_argument("--nj", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def run_path(case, fields, precision, steps, warmup, tol):
    from simcoflow_amd.projection import Projection, params_for
    u, v, DpU, DpV = fields
    prm = params_for(case.fluids, tol=tol, precision=precision, max_iter=200)
    pj = Projection(case.grid, prm)
    pj.set_geometry(case.geo)
    pj.set_interface(case.itf.vof, case.itf.phi)
    pj.set_predictor(u, v, DpU, DpV)
    pj.upload("U", u)
    pj.upload("V", v)
    infos = []

    def step():
        pj.build(0.0)
        info = pj.solve()
        pj.velocity_correct(0.0)
        return info

    for _ in range(warmup):
        step()
    pj.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        infos.append(step())
    pj.synchronize()
    t1 = time.perf_counter()
    Pp = pj.download("PP")
    mres = pj.download("MRES")
    kern = kernel_roofline(pj, case, precision)
    pj.close()
    ms = (t1 - t0) * 1e3 / steps
    return {
        "ms_per_step": ms,
        "poisson_ms": float(np.mean([i.ms for i in infos])),
        "setup_ms": float(np.mean([i.setup_ms for i in infos])),
        "iterations": int(np.mean([i.iterations for i in infos])),
        "relres": float(max(i.relres for i in infos)),
        "max_row_residual": float(max(i.max_row_residual for i in infos)),
        "max_mres": float(np.abs(mres).max()),
        "Pp": Pp,
        "kernel": kern,
    }


# Algorithmic bytes per fine cell of the fine-level up-sweep k_up5 (the
# kernel with the largest share of a solve): reads b = PCG residual (FP64),
# x after pre-smoothing, diag/cE/cN, the 9 prolongation weights and the
# coarse correction (both at 1/4 resolution) and q (FP64, for the fused
# dot); writes z (FP64).  w = word size of the multigrid data.
def up_sweep_bytes_per_cell(w):
    return 8 + w + 3 * w + 9 * w / 4 + w / 4 + 8 + 8


def kernel_roofline(pj, case, precision):
    """Average duration of one kernel, timed with HIP events on the
    library's stream by sg_time_kernel, against its algorithmic bytes."""
    import ctypes
    from simcoflow_amd import _lib
    L = _lib.load()
    fn = L.sg_time_kernel
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    out = {}
    n = case.grid.ncells
    w = 4 if precision == 1 else 8
    for which, name, bpc in [(2, f"k_up5 (fine multigrid up-sweep + PCG dots, {'FP32' if w == 4 else 'FP64'})",
                              up_sweep_bytes_per_cell(w)),
                             (0, "k_apply5 (fine operator apply, FP64)", 5 * 8)]:
        ms = ctypes.c_double(0.0)
        if fn(pj.ctx, which, 100, ctypes.byref(ms)) != 0:
            continue
        b = bpc * n
        out[which] = {"name": name, "avg_ms": ms.value, "bytes": b, "GBs": b / (ms.value * 1e-3) / 1e9}
    # a plain copy of one padded fine-level FP64 vector: the calibration
    # kernel for the FETCH_SIZE / WRITE_SIZE counters
    ld = ((case.grid.ni + 1 + 31) // 32) * 32
    ncopy = (case.grid.nj + 3) * ld
    ms = ctypes.c_double(0.0)
    if fn(pj.ctx, 3, 20, ctypes.byref(ms)) == 0:
        out[3] = {"name": "k_copy", "avg_ms": ms.value, "bytes": 16 * ncopy,
                  "GBs": 16 * ncopy / (ms.value * 1e-3) / 1e9}
    return out


def measured_traffic(kernel_name):
    """HBM bytes per launch of the roofline kernel from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel),
    or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", "traffic_up_sweep_swarm2048.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("bytes_per_launch") if d.get("kernel", "") in kernel_name else None
    except (OSError, ValueError):
        return None


def cpu_baseline(case, fields, threads):
    """The oracle's FP64 PCG + multigrid on the host cores, on a bounded
    sample: solves of the same workload at 512x512 (a 1/16 sub-sample of the
    grid), scaled to Mcell-steps/s."""
    from oracle import oracle as O
    from simcoflow_amd import cases as C
    O.build()
    if threads > 0:
        O.lib().orc_set_num_threads(threads)
    nthreads = O.lib().orc_num_threads()
    small = C.case_swarm(512, 512, seed=0)
    u, v, DpU, DpV = C.predictor_fields(small)
    g = small.grid
    t0 = time.perf_counter()
    n = 0
    its = []
    while True:
        M = O.poisson_matrix(g, small.geo, O.poisson_coef(g, small.itf, DpU, DpV, small.fluids), small.bc)
        b = O.poisson_rhs(g, small.geo, u, v)
        x, it, rr = O.pcg_mg_solve(g, M, b, tol=1e-8, max_iter=2000)
        its.append(abs(it))
        n += 1
        if time.perf_counter() - t0 > 10.0 or n >= 50:
            break
    dt = (time.perf_counter() - t0) / n
    return {"value": g.ncells / dt / 1e6, "unit": "Mcell-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"{n} projection solves (build+PCG-MG to 1e-8, {int(np.mean(its))} iters) of the swarm "
                      f"at 512x512, {dt * 1e3:.1f} ms each"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the projection is "
              "not sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    from simcoflow_amd import cases as C
    case = C.case_swarm(a.ni, a.nj, seed=a.seed)
    fields = C.predictor_fields(case)
    tol = 1e-8
    fp64 = run_path(case, fields, 0, a.steps, a.warmup, tol)
    mixed = run_path(case, fields, 1, a.steps, a.warmup, tol)
    err = float(np.linalg.norm(mixed["Pp"] - fp64["Pp"]) / np.linalg.norm(fp64["Pp"]))
    ncell = case.grid.ncells
    value = ncell / (mixed["ms_per_step"] * 1e-3) / 1e6
    out = {
        "metric": "Poisson ms/step & Mcell-steps/s, learned+FP32 vs plain FP64 PCG at 1e-8 residual",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 multigrid",
        "data": "synthetic (seeded bubble swarm, seeded predictor fields)",
        "config": {"workload": f"bubble swarm {a.ni}x{a.nj}, 64 bubbles seed {a.seed}, projection step "
                               f"(build + PCG-MG to 1e-8 + correction)", "ni": a.ni, "nj": a.nj},
        "poisson_ms_per_step": round(mixed["poisson_ms"], 4),
        "fp64_poisson_ms_per_step": round(fp64["poisson_ms"], 4),
        "fp64_ms_per_step": round(fp64["ms_per_step"], 4),
        "fp64_value": round(ncell / (fp64["ms_per_step"] * 1e-3) / 1e6, 3),
        "speedup_mixed_vs_fp64": round(fp64["poisson_ms"] / mixed["poisson_ms"], 3),
        "iterations": {"fp64": fp64["iterations"], "mixed": mixed["iterations"]},
        "relres": {"fp64": fp64["relres"], "mixed": mixed["relres"]},
        "pressure_rel_l2_mixed_vs_fp64": err,
        "max_mass_residual": {"fp64": fp64["max_mres"], "mixed": mixed["max_mres"]},
    }
    k = mixed["kernel"].get(2)
    if k:
        out["roofline"] = {"bound": "hbm", "achieved": round(k["GBs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": round(k["GBs"] / PEAK_HBM_GBS, 4), "traffic": measured_traffic(k["name"]),
                           "kernel": k["name"], "avg_ms": round(k["avg_ms"], 5),
                           "bytes_alg": int(k["bytes"])}
        out["kernels"] = {
            "fp64_up_sweep": fp64["kernel"].get(2),
            "fp64_apply": fp64["kernel"].get(0),
        }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, fields, a.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

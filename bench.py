"""Benchmark of the hot path on one MI355X: full time steps and the
projection (Poisson) solve.

Workload: the 64-bubble swarm on a 2048x2048 grid (1x1 m box, liquid depth
0.8 m, density ratio 1000:1), the configuration the solver metric is quoted
on.  One step = one full time step of the hot path, device resident, in the
order of AdamBasforthCrankNicolson (Solver_Mod.f90:246-319): CFL time step,
interface transport (Coupled_LS_VOF, 6 sub-steps) and the interface on the
U/V cells, the box ghost layer, the momentum predictor (Predictor_UV, two
implicit solves), the projection (operator and RHS build, PCG + multigrid to
a relative residual of 1e-8, pressure scatter) and the velocity correction.
The swarm starts at rest; W steps of warm-up, then K timed steps.

Three solver paths run the same steps:
  * plain FP64 PCG-multigrid from x0 = 0 (the reference's start);
  * mixed precision (FP32 PCG vectors and multigrid cycle; FP64 x, dots and
    residual recomputation) from x0 = 0;
  * learned + mixed: the mixed solver started from the last step's pressure
    plus the learned correction (the net predicts it from the residual that
    pressure leaves, Galerkin-scaled; warm_start = 3).  The net is trained
    inside this run, with the library's own kernels, on (rhs, p) pairs the
    plain solver makes for seeded 512x512 swarms; it runs on the 256x256
    level of the hierarchy.
Two more full-step paths are reported beside them: the learned start alone
(from zero, warm_start = 2) and the mixed solver from the last step's
pressure without the net (warm_start = 1).
``value`` is the learned + mixed path's Mcell-steps/s over full time steps
(the headline metric); ``poisson_ms_per_step`` gives the Poisson solve time
per step of every path, inside those steps and, separately, on a fixed
projection step from seeded predictor fields (the solver alone, repeated).

Prints one JSON line.
"""
import argparse
import ctypes
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ni", type=int, default=2048)
    ap.add_argument("--nj", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip training and the learned path")
    ap.add_argument("--no-full-step", action="store_true", help="skip the per-part step timing")
    ap.add_argument("--no-particles", action="store_true", help="skip the particle and diagnostics timing")
    ap.add_argument("--only-particles", action="store_true",
                    help="time only the particle tracker and the diagnostics (for counter runs); prints their JSON")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--nn-width", type=int, default=256, help="widest multigrid level the net runs on")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# The transport's sweep: 8 FP64 words per cell and sweep (16 per sweep pair:
# read vof, phi, the two normal components, the plane constant, the face
# velocity, the two apertures and vofS; write vof, phi and the normal and
# plane constant the next reconstruction reads).
SWEEP_WORDS = 8

# Algorithmic bytes per fine cell and PCG iteration (compulsory traffic at
# the stored precision): operator apply 5 words (FP64 operator, the vector
# counted at FP64 too), PCG vector work 9 words of the vector precision (+ 2 FP32 words when the
# vectors are FP32: x stays FP64), one V-cycle 26 words of the multigrid
# precision; plus 15 FP64 words per solve for the coefficient and RHS build.
def solve_bytes(n, iters, w):
    x_extra = 8 if w == 4 else 0
    return n * (15 * 8 + iters * (5 * 8 + 9 * w + x_extra + 26 * w))


# Algorithmic bytes per fine cell of the fine-level up-sweep k_up5 (the
# kernel with the largest share of a solve): reads b = PCG residual, x after
# pre-smoothing, diag/cE/cN, the 9 prolongation weights and the coarse
# correction (both at 1/4 resolution) and q (for the fused dot); writes z.
# w = word size of the solve (FP64: 8; mixed precision: 4, vectors included).
def up_sweep_bytes_per_cell(w):
    return w + w + 3 * w + 9 * w / 4 + w / 4 + w + w


def kernel_roofline(pj, n, precision):
    """Average duration of the fine up-sweep (and, for calibration, a plain
    copy), timed with HIP events on the library's stream by sg_time_kernel."""
    from simcoflow_amd import _lib
    fn = _lib.load().sg_time_kernel
    w = 4 if precision == 1 else 8
    out = {}
    ms = ctypes.c_double(0.0)
    if fn(pj.ctx, 2, 100, ctypes.byref(ms)) == 0:
        b = up_sweep_bytes_per_cell(w) * n
        out["up"] = {"name": f"k_up5 (fine multigrid up-sweep + PCG dots, {'FP32' if w == 4 else 'FP64'})",
                     "avg_ms": ms.value, "bytes": b, "GBs": b / (ms.value * 1e-3) / 1e9}
    if fn(pj.ctx, 0, 100, ctypes.byref(ms)) == 0:
        b = 5 * 8 * n
        out["apply"] = {"name": "k_apply5 (fine operator apply, FP64)", "avg_ms": ms.value, "bytes": b,
                        "GBs": b / (ms.value * 1e-3) / 1e9}
    # a plain copy of one padded fine-level FP64 vector: the calibration
    # kernel of the FETCH_SIZE / WRITE_SIZE counters
    ld = ((pj.ni + 1 + 31) // 32) * 32
    ncopy = (pj.nj + 3) * ld
    if fn(pj.ctx, 3, 20, ctypes.byref(ms)) == 0:
        out["copy"] = {"name": "k_copy (FP64 vector copy)", "avg_ms": ms.value, "bytes": 16 * ncopy,
                       "GBs": 16 * ncopy / (ms.value * 1e-3) / 1e9}
    return out


def run_path(case, fields, precision, steps, warmup, tol, learned=None, width=256):
    from simcoflow_amd import learned as Lr
    from simcoflow_amd.projection import Projection, params_for
    u, v, DpU, DpV = fields
    prm = params_for(case.fluids, tol=tol, precision=precision, max_iter=200, warm_start=2 if learned else 0)
    pj = Projection(case.grid, prm)
    pj.set_geometry(case.geo)
    pj.set_interface(case.itf.vof, case.itf.phi)
    pj.set_predictor(u, v, DpU, DpV)
    pj.upload("U", u)
    pj.upload("V", v)
    if learned:
        w, K = learned
        # deflation needs FP64 directions; the mixed solve takes the warm start
        Lr.enable(pj, w, Lr.pick_level(pj, width), K, deflate=(precision == 0))
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
    kern = kernel_roofline(pj, case.grid.ncells, precision) if not learned else {}
    pj.close()
    return {
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "poisson_ms": float(np.mean([i.ms for i in infos])),
        "setup_ms": float(np.mean([i.setup_ms for i in infos])),
        "warm_ms": float(np.mean([i.warm_ms for i in infos])),
        "iterations": float(np.mean([i.iterations for i in infos])),
        "relres": float(max(i.relres for i in infos)),
        "warm_relres": float(np.mean([i.warm_relres for i in infos])),
        "max_row_residual": float(max(i.max_row_residual for i in infos)),
        "max_mres": float(np.abs(mres).max()),
        "Pp": Pp,
        "kernel": kern,
    }


def learned_bound(case, fields, width):
    """What any predictor on the net's level could gain: the exact pressure
    (zero start) and the exact correction of a previous pressure (warm_start
    3, a second system whose predictor velocities differ by 10 %), sampled
    at the level's points as the training targets are and handed to the
    solve in place of the net's output.  Iterations to 1e-8, FP64, on the
    net's level and on the first coarse level."""
    from simcoflow_amd import learned as Lr
    from simcoflow_amd.projection import Projection, params_for
    u, v, DpU, DpV = fields
    rng = np.random.default_rng(7)
    # a smooth second predictor field: the first plus 10 % of a shifted copy
    u2 = u + 0.1 * np.roll(u, 37, axis=1) * rng.uniform(0.5, 1.5)
    v2 = v + 0.1 * np.roll(v, 53, axis=0) * rng.uniform(0.5, 1.5)
    pj = Projection(case.grid, params_for(case.fluids, tol=1e-8, max_iter=300))
    try:
        pj.set_geometry(case.geo)
        pj.set_interface(case.itf.vof, case.itf.phi)

        def solve(uu, vv, tol=1e-8, ws=0):
            pj.set_predictor(uu, vv, DpU, DpV)
            pj.set_params(params_for(case.fluids, tol=tol, max_iter=300, warm_start=ws))
            pj.build(0.0)
            info = pj.solve()
            return info, pj.download("PP")[1:-1, 1:-1].copy()
        net_level = Lr.pick_level(pj, width)
        _, x2 = solve(u2, v2, 1e-12)
        _, x1 = solve(u, v, 1e-12)
        zero, _ = solve(u, v)
        solve(u, v, 1e-12)
        prev, _ = solve(u2, v2, ws=1)
        out = {"net_level": net_level, "net_level_grid": list(Lr.level_dims(pj, net_level)),
               "zero_start_iterations": zero.iterations, "zero_start_ms": round(zero.ms, 4),
               "previous_pressure_iterations": prev.iterations, "previous_pressure_ms": round(prev.ms, 4)}
        for lv in sorted({net_level, 1}):
            Lr.set_prediction(pj, lv, Lr.sample_at_level(pj, lv, x1))
            a, _ = solve(u, v, ws=2)
            Lr.set_prediction(pj, lv, Lr.sample_at_level(pj, lv, x2 - x1))
            solve(u, v, 1e-12)
            b, _ = solve(u2, v2, ws=3)
            Lr.set_prediction(pj, lv, None)
            out[f"level_{lv}"] = {"grid": list(Lr.level_dims(pj, lv)),
                                  "exact_pressure_iterations": a.iterations,
                                  "exact_correction_iterations": b.iterations}
    finally:
        pj.close()
    return out


def train_warm_start(seed, width):
    """The learned warm start, trained in this run on the GPU: pairs from the
    plain FP64 solver on four seeded 512x512 swarms, 48 right-hand sides
    each, then Adam with the library's kernels."""
    from simcoflow_amd import cases as C
    from simcoflow_amd import learned as Lr
    t0 = time.perf_counter()
    X, T, M, level = Lr.collect_pairs(lambda s: C.case_swarm(512, 512, seed=s), seeds=[seed + 101 + k for k in range(4)],
                                      per_case=48, max_width=width)
    t1 = time.perf_counter()
    w, K, hist = Lr.train(X, T, M, epochs=10, batch=16, lr=3e-3)
    t2 = time.perf_counter()
    return w, K, {"pairs": int(X.shape[0]), "level_grid": [int(X.shape[1] - 2), int(X.shape[2] - 2)],
                  "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
                  "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1]}


def stepping(case, stags, precision, steps, warmup, learned=None, width=256, previous=False, sweep_events=False):
    """K timed full time steps after W warm-up steps (TimeStepper), with the
    Poisson device time of each step from its solve report; with
    sweep_events, every transport sweep launch of the K steps is also timed
    with HIP events on the library's stream (sg_sweep_timing)."""
    from simcoflow_amd import _lib
    from simcoflow_amd.solver import TimeStepper
    lrn = None
    if learned:
        w, K = learned
        # deflation needs FP64 directions; the mixed solve takes the warm start
        lrn = (w, None, K, precision == 0)
    st = TimeStepper(case, precision=precision, tol=1e-8, max_iter=200, learned=lrn, stags=stags,
                     learned_width=width, previous_pressure=previous)
    try:
        for _ in range(warmup):
            st.step()
        st.pj.synchronize()
        if sweep_events:
            _lib.check(st.pj.lib.sg_sweep_timing(st.pj.ctx, 1, None, None))
        reps = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reps.append(st.step())
        st.pj.synchronize()
        t1 = time.perf_counter()
        sweeps = None
        if sweep_events:
            nl, tot = ctypes.c_longlong(0), ctypes.c_double(0.0)
            _lib.check(st.pj.lib.sg_sweep_timing(st.pj.ctx, 0, ctypes.byref(nl), ctypes.byref(tot)))
            sweeps = {"launches": int(nl.value), "avg_ms": tot.value / max(nl.value, 1),
                      "per_step_ms": tot.value / steps}
        vof = st.pj.download("VOF")
    finally:
        st.close()
    p = [r["poisson"] for r in reps]
    return {"sweeps": sweeps, "ms_per_step": (t1 - t0) * 1e3 / steps,
            "poisson_ms": float(np.mean([r["ms"] for r in p])),
            "warm_ms": float(np.mean([r["warm_ms"] for r in p])),
            "iterations": float(np.mean([r["iterations"] for r in p])),
            "relres": float(max(r["relres"] for r in p)),
            "dt_mean": float(np.mean([r["dt"] for r in reps])), "vof": vof}


def step_parts(case, stags, precision, steps=3, warmup=2):
    """The whole time step on the swarm (TimeStepper), with the device time
    of each part from host timers, the stream drained at each boundary."""
    from simcoflow_amd import _lib
    from simcoflow_amd.solver import TimeStepper
    st = TimeStepper(case, precision=precision, tol=1e-8, max_iter=200, stags=stags)
    pj, lib, ctx = st.pj, st.pj.lib, st.pj.ctx
    parts = {}
    try:
        for k in range(warmup + steps):
            if k < warmup:
                st.step()
                continue
            # the same sequence as TimeStepper.step, each part timed
            st.itt += 1
            t = time.perf_counter
            a = t()
            for dst, src in (("VOFO", "VOF"), ("PHIO", "PHI"), ("NXO", "NX"), ("NYO", "NY"), ("UVOFO", "UVOF"),
                             ("VVOFO", "VVOF")):
                _lib.check(lib.sg_field_copy(ctx, _lib.F[dst], _lib.F[src]))
            dt = ctypes.c_double()
            _lib.check(lib.sg_time_step_size(ctx, st.cfl, st.vint, st.uref, 0.0, 0.0, ctypes.byref(dt)))
            b = t()
            pj.clsvof_advect_resident(dt.value, st.time)
            pj.synchronize()
            c = t()
            _lib.check(lib.sg_interface_to_uv_resident(ctx))
            _lib.check(lib.sg_closed_box_bc(ctx))
            pj.synchronize()
            d = t()
            pj.predictor_resident(dt.value, 0.0, False)
            e = t()
            pj.build(0.0)
            info = pj.solve()
            pj.velocity_correct(0.0)
            pj.synchronize()
            f = t()
            st.time += dt.value
            for name, ms in (("copies_and_dt", b - a), ("transport", c - b), ("interface_to_uv_and_bc", d - c),
                             ("predictor", e - d), ("projection", f - e), ("step", f - a)):
                parts.setdefault(name, []).append(ms * 1e3)
    finally:
        st.close()
    return {k: round(float(np.mean(v)), 4) for k, v in parts.items()}


PARTICLE_STREAM_BYTES = 72  # x, y, u, v, d read + x, y, u, v written, FP64
PARTICLE_GATHER_BYTES = 48  # vof, vofS, two u faces, two v faces of the particle's cell
DIAG_BYTES_PER_CELL = 40  # vof, vofS, mres, u, v read once, FP64


def particles(case, fields, n=1 << 20, reps=20, seed=0):
    """The Lagrangian tracker (TrackingParticles, 10 sub-steps of the
    exponential drag integrator) on n particles spread uniformly over the
    swarm box, diameters log-normal about 50 um (sigma 0.5), density 2500
    kg/m^3, in the seeded predictor velocity field, and one row of the
    per-step diagnostics on the same fields; average device times over reps
    launches (HIP events on the context's stream)."""
    from simcoflow_amd import _lib
    from simcoflow_amd.projection import Projection, params_for
    g = case.grid
    rng = np.random.default_rng(seed)
    lx, ly = float(g.dx.sum()), float(g.dy.sum())
    x, y = rng.uniform(0.0, lx, n), rng.uniform(0.0, ly, n)
    u, v = rng.normal(0.0, 0.05, n), rng.normal(0.0, 0.05, n)
    d = np.exp(rng.normal(np.log(50e-6), 0.5, n))
    f = case.fluids
    prm = dict(rho_l=f.rho_l, rho_g=f.rho_g, nu_l=f.mu_l / f.rho_l, nu_g=f.mu_g / f.rho_g, rho_p=2500.0, gx=0.0,
               gy=-f.g, lref=1.0, uref=1.0, us=0.0, vs=0.0)
    pj = Projection(g, params_for(f))
    try:
        pj.set_geometry(case.geo)
        pj.upload("VOF", case.itf.vof)
        pj.upload("U", fields[0])
        pj.upload("V", fields[1])
        pj.particles_set(x, y, u, v, d)
        flagged = pj.particles_track(case.dt, prm)
        ms = ctypes.c_double()
        _lib.check(pj.lib.sg_time_kernel(pj.ctx, 4, reps, ctypes.byref(ms)))
        dms = ctypes.c_double()
        _lib.check(pj.lib.sg_time_kernel(pj.ctx, 5, reps, ctypes.byref(dms)))
    finally:
        pj.close()
    b = n * (PARTICLE_STREAM_BYTES + PARTICLE_GATHER_BYTES)
    gbs = b / (ms.value * 1e-3) / 1e9
    part = {"kernel": "k_particles (TrackingParticles, FP64, one thread per particle)", "particles": n,
            "avg_ms": round(ms.value, 5), "mparticle_steps_per_s": round(n / (ms.value * 1e-3) / 1e6, 1),
            "bytes_alg": int(b), "GBs": round(gbs, 1), "frac_hbm": round(gbs / PEAK_HBM_GBS, 4), "flagged": flagged}
    db = g.ni * g.nj * DIAG_BYTES_PER_CELL
    dgbs = db / (dms.value * 1e-3) / 1e9
    diag = {"kernel": "k_diag + k_diag_fin (one diagnostics row: volumes, gas centroid, max |mres|, |u|, |v|)",
            "avg_ms": round(dms.value, 5), "bytes_alg": int(db), "GBs": round(dgbs, 1),
            "frac_hbm": round(dgbs / PEAK_HBM_GBS, 4)}
    return part, diag


def geometry_rebuild(case, reps=10):
    """The cut-cell geometry of a moving cylinder (radius 0.1 m) rebuilt on
    the device on the swarm grid (sg_geometry_rebuild: P, U and V cells),
    host wall time per call (the call waits for the closing-rule passes)."""
    from simcoflow_amd.projection import Projection, params_for
    g = case.grid
    pj = Projection(g, params_for(case.fluids))
    try:
        times = []
        for k in range(reps + 2):
            t0 = time.perf_counter()
            pj.geometry_rebuild("circle", 0.5 + 1e-3 * k, 0.4 + 7e-4 * k, 0.1)
            if k >= 2:
                times.append(time.perf_counter() - t0)
    finally:
        pj.close()
    ms = float(np.median(times)) * 1e3
    return {"what": "sg_geometry_rebuild, circle r = 0.1 m moved between calls", "ms_per_call": round(ms, 4),
            "mcells_per_s": round(g.ni * g.nj / (ms * 1e-3) / 1e6, 1)}


def measured_traffic(kernel_name):
    """HBM bytes per launch of the roofline kernel from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel),
    or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", "traffic_sweep_swarm2048.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("bytes_per_launch") if d.get("kernel", "") in kernel_name else None
    except (OSError, ValueError):
        return None


def cpu_baseline(threads):
    """The oracle's full time step on the host cores, on a bounded sample:
    time steps of the same swarm on a 256x256 grid (a 1/64 sub-sample),
    scaled to Mcell-steps/s.  The oracle's PCG-multigrid uses every host
    thread (OpenMP); its interface transport and predictor assembly are
    single-threaded C, its momentum solves scipy's direct sparse solver."""
    from oracle import oracle as O
    from oracle.step import OracleStepper
    from simcoflow_amd import cases as C
    O.build()
    if threads > 0:
        O.lib().orc_set_num_threads(threads)
    nthreads = O.lib().orc_num_threads()
    small = C.case_swarm(256, 256, seed=0)
    st = OracleStepper(O, small, tol=1e-8)
    st.step()
    st.step()
    t0 = time.perf_counter()
    n = 0
    while True:
        st.step()
        n += 1
        if time.perf_counter() - t0 > 10.0 or n >= 20:
            break
    dt = (time.perf_counter() - t0) / n
    # the same steps on one host thread (a few of them)
    O.lib().orc_set_num_threads(1)
    t1 = time.perf_counter()
    n1 = 0
    while True:
        st.step()
        n1 += 1
        if time.perf_counter() - t1 > 5.0 or n1 >= 10:
            break
    dt1 = (time.perf_counter() - t1) / n1
    O.lib().orc_set_num_threads(nthreads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(small.grid.ncells / dt / 1e6, 4), "unit": "Mcell-steps/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{n} full time steps (transport, U/V interface, predictor with direct momentum solves, "
                      f"PCG-MG projection to 1e-8, correction) of the swarm at 256x256 after 2 warm-up steps, "
                      f"{dt * 1e3:.0f} ms each; PCG-MG on {nthreads} host threads, the rest single-threaded",
            "one_thread": {"value": round(small.grid.ncells / dt1 / 1e6, 4), "steps": n1},
            "host": {"cpu_model": model, "nproc": os.cpu_count()}}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the hot path is "
              "not sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    if not a.no_learned:
        # torch (the training data's device buffers) loads its HIP runtime
        # first; the library then binds to the same one
        import torch  # noqa: F401
    from simcoflow_amd import cases as C
    log("building the swarm case and its staggered geometry")
    case = C.case_swarm(a.ni, a.nj, seed=a.seed)
    fields = C.predictor_fields(case)
    if a.only_particles:
        part, diag = particles(case, fields)
        print(json.dumps({"particles": part, "diagnostics": diag, "geometry_rebuild": geometry_rebuild(case)}))
        return
    stags = C.staggered_geometry(case.grid, case.geo, case.solid)
    tol = 1e-8
    n = case.grid.ncells
    w = K = train = None
    if not a.no_learned:
        log("training the warm-start net")
        w, K, train = train_warm_start(a.seed, a.nn_width)
    # full time steps, one per solver path
    paths = [("fp64", 0, None), ("mixed", 1, None)] + ([("learned_mixed", 1, (w, K))] if w is not None else [])
    full = {}
    for name, prec, lrn in paths:
        log(f"full time steps, {name}")
        # the learned path starts from the last pressure plus the net's correction
        full[name] = stepping(case, stags, prec, a.steps, a.warmup, learned=lrn, width=a.nn_width,
                              previous=lrn is not None)
    # beside the metric's path (full steps only; on fixed fields every repeat
    # would start from the answer): the learned start from zero, and the last
    # step's pressure without the net
    if w is not None:
        log("full time steps, learned start from zero")
        full["learned_mixed_zero_start"] = stepping(case, stags, 1, a.steps, a.warmup, learned=(w, K),
                                                    width=a.nn_width)
    log("full time steps, mixed from the previous pressure")
    full["mixed_previous_pressure"] = stepping(case, stags, 1, a.steps, a.warmup, previous=True)
    # the projection alone on one fixed set of predictor fields
    proj = {}
    for name, prec, lrn in paths:
        log(f"projection step on fixed fields, {name}")
        proj[name] = run_path(case, fields, prec, a.steps, a.warmup, tol, learned=lrn, width=a.nn_width)
    head_name = paths[-1][0]
    head = full[head_name]
    value = n / (head["ms_per_step"] * 1e-3) / 1e6
    ref_p = proj["fp64"]["Pp"]
    ref_vof = full["fp64"]["vof"]
    cellvol = np.outer(case.grid.dy, case.grid.dx)
    vol_ref = float((ref_vof * cellvol).sum())
    out = {
        "metric": "Poisson ms/step & Mcell-steps/s, learned+FP32 vs plain FP64 PCG at 1e-8 residual",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 PCG vectors and multigrid, f64 x / dots / residual recomputation / operator in q = A p; "
                 "f64 transport and predictor",
        "data": "synthetic (seeded bubble swarm at rest; seeded predictor fields for the fixed projection)",
        "config": {"workload": f"bubble swarm {a.ni}x{a.nj} (1x1 m, liquid depth 0.8 m, 64 bubbles, seed {a.seed}), "
                               "full time step: CFL dt, CLSVOF transport (6 sub-steps), U/V interface, box ghost "
                               "layer, momentum predictor, projection (PCG-multigrid to 1e-8), correction",
                   "ni": a.ni, "nj": a.nj,
                   "path": ("previous pressure + learned correction (warm_start 3), mixed precision" if w is not None
                            else "mixed precision")},
        "poisson_ms_per_step": {k: round(v["poisson_ms"], 4) for k, v in full.items()},
        "ms_per_step_by_path": {k: round(v["ms_per_step"], 4) for k, v in full.items()},
        "mcell_steps_per_s": {k: round(n / (v["ms_per_step"] * 1e-3) / 1e6, 3) for k, v in full.items()},
        "iterations": {k: v["iterations"] for k, v in full.items()},
        "relres": {k: v["relres"] for k, v in full.items()},
        "poisson_speedup_vs_fp64": {k: round(full["fp64"]["poisson_ms"] / v["poisson_ms"], 3)
                                    for k, v in full.items() if k != "fp64"},
        "vof_max_abs_diff_vs_fp64": {k: float(np.abs(v["vof"] - ref_vof).max())
                                     for k, v in full.items() if k != "fp64"},
        "liquid_volume_rel_diff_vs_fp64": {k: abs(float((v["vof"] * cellvol).sum()) - vol_ref) / vol_ref
                                           for k, v in full.items() if k != "fp64"},
    }
    fixed = {"poisson_ms": {}, "setup_ms": {}, "iterations": {}, "relres": {}, "pressure_rel_l2_vs_fp64": {},
             "max_mass_residual": {}, "speedup_vs_fp64": {}, "solve_hbm": {}}
    for k, v in proj.items():
        fixed["poisson_ms"][k] = round(v["poisson_ms"], 4)
        fixed["setup_ms"][k] = round(v["setup_ms"], 4)
        fixed["iterations"][k] = v["iterations"]
        fixed["relres"][k] = v["relres"]
        fixed["max_mass_residual"][k] = v["max_mres"]
        if k != "fp64":
            fixed["pressure_rel_l2_vs_fp64"][k] = float(np.linalg.norm(v["Pp"] - ref_p) / np.linalg.norm(ref_p))
            fixed["speedup_vs_fp64"][k] = round(proj["fp64"]["poisson_ms"] / v["poisson_ms"], 3)
        if k != "learned_mixed":
            b = solve_bytes(n, v["iterations"], 8 if k == "fp64" else 4)
            gbs = b / (v["poisson_ms"] * 1e-3) / 1e9
            fixed["solve_hbm"][k] = {"bytes_alg": int(b), "GBs": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4)}
    out["projection_fixed_fields"] = fixed
    if w is not None:
        out["learned"] = dict(train, warm_start_ms=round(proj["learned_mixed"]["warm_ms"], 4),
                              start_relres=proj["learned_mixed"]["warm_relres"])
        log("upper bound of a learned start on the net's level")
        out["learned"]["upper_bound"] = learned_bound(case, fields, a.nn_width)
    # roofline: the transport sweeps, the largest kernel family of the
    # metric's time step, timed launch by launch with HIP events during
    # steps of the metric's own path (a second pass, so the events stay out
    # of the headline's timed steps); bytes: 8 FP64 words per cell per sweep
    log("transport sweeps timed with HIP events, metric path")
    lrn = (w, K) if w is not None else None
    ev = stepping(case, stags, 1, a.steps, a.warmup, learned=lrn, width=a.nn_width, previous=lrn is not None,
                  sweep_events=True)["sweeps"]
    b_sweep = SWEEP_WORDS * 8 * n
    gbs = b_sweep / (ev["avg_ms"] * 1e-3) / 1e9
    whole = {}
    for k in ("fp64", "mixed"):
        if k in fixed["solve_hbm"]:
            whole[k] = dict(fixed["solve_hbm"][k], ms=fixed["poisson_ms"][k], iterations=fixed["iterations"][k])
    out["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": measured_traffic("k_sweep"),
                       "kernel": "k_sweep (CLSVOF X/Y sweep, FP64: the step's largest kernel family)",
                       "avg_ms": round(ev["avg_ms"], 5), "launches": ev["launches"],
                       "sweeps_ms_per_step": round(ev["per_step_ms"], 4), "bytes_alg": int(b_sweep),
                       "whole_solve": whole}
    k = proj["mixed"]["kernel"].get("up")
    if k:
        out["kernels"] = {"mixed_up_sweep": k, "fp64_up_sweep": proj["fp64"]["kernel"].get("up"),
                          "fp64_apply": proj["fp64"]["kernel"].get("apply"),
                          "copy": proj["fp64"]["kernel"].get("copy")}
    if not a.no_particles:
        log("particle tracking and diagnostics")
        out["particles"], out["diagnostics"] = particles(case, fields)
        out["geometry_rebuild"] = geometry_rebuild(case)
    if not a.no_full_step:
        log("time step parts")
        out["step_parts_ms"] = {"mixed": step_parts(case, stags, 1)}
    if not a.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(a.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

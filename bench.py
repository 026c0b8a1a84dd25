"""Benchmark of the soft-body step on one MI355X: beam updates per second at the 2 kHz substep.

Workload (the single-large-body configuration): the 64x32x24 lattice (49 152 nodes, 420 848 beams, per-beam k
log-uniform in [1e5, 5e6] N/m, 2 % of the beams weakened to break strain 0.01) with its faces as collision
triangles and self contact on, standing on a concrete ground box, and 2 000 rigid balls (radius 0.1-0.3 m, 5 kg) resting on its top and around
it; inter-body contacts through the spatial hash (pair rebuild every 17 substeps, narrow phase and ordered response
every substep), ball-lattice contacts in every substep from the first. 4000 substeps of 0.5 ms in frames of 33
substeps (the app's 60 Hz frame). One step = one substep of every beam, node and contact. Inputs are resident in HBM
before the timed region; the rate is on the wall clock around the timed frames.

Prints one JSON line with the throughput, the dominant kernel's HBM roofline figures (achieved = algorithmic bytes
per launch / its average duration, from HIP events around each launch on the step's stream) and the compiled
reference's CPU step timed on a bounded sample of the same workload on this host's cores. The beam path alone (the
same lattice on the ground without contact geometry, where the step runs as one persistent kernel) is reported
beside it.

The metric's second half, implicit FEM solves/s, is the object "implicit_fem": the co-rotational shell plate of
256 x 256 cells (66 049 nodes, 65 025 free, 390 150 unknowns, 131 072 triangles; E 70 GPa, nu 0.33, 1.5 mm,
clamped edges), 200 implicit steps of 0.5 ms (the reference's step) under seeded random point loads (32 free nodes
a step, normal to the plate, N(0, 50 N), uploaded with the step), each step the element pass, the 6x6-block assembly, the supernodal Cholesky
(FP64 fronts on the matrix cores), the substitutions and the update, its residual measured on the device. Its
roofline is the factorization's: FP64 flops of the numeric factorization / its device time, against the FP64
matrix-core peak. Its CPU baseline is the compiled reference's FemFrame step on the same plate (one thread, one
step after a warm step), run in a child process beside the GPU legs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
FEM_CELLS = 256
FEM_H = 0.0005  # the reference's step (kDefaultDt, src/phys/softbody.h:25), the one its fembench solves with
METRIC = "beam updates/s at 2 kHz substep; implicit FEM solves/s at 1e-6 residual"
FRAME = 33
# the forms the contact step can take: its force kernel's workgroup size and beam ends per thread
CONTACT_CANDIDATES = [dict(mode=1, block=256, incidences_per_thread=p, substeps_per_launch=0, use_graph=g)
                      for g in (1, 0) for p in (1, 2, 4, 8)] + \
                     [dict(mode=1, block=512, incidences_per_thread=p, substeps_per_launch=0, use_graph=1) for p in (2, 4)]


def build_world(bodies):
    from beamlab_amd import scenes
    w = scenes.World()
    scenes.ground_box(w)
    for b in bodies:
        w.add_body(b)
    return w


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(bodies, sample_substeps, threads, runs=5):
    """The compiled reference's World::step_substeps (oracle/_ref) on the same scene; the C restatement when the
    reference library is absent. A warm-up of 2 frames, then `runs` timed samples of sample_substeps / runs substeps
    each, in frames of 33; the median sample's rate."""
    import statistics
    from beamlab_amd import abi, scenes
    from oracle import oracle
    w = scenes.World()
    scenes.ground_box(w)
    if oracle.have_ref():
        cw, kind = oracle.RefWorld(abi.default_settings(), w.statics.boxes, threads=threads), "reference"
    else:
        cw, kind, threads = oracle.OracleWorld(abi.default_settings(), w.statics.boxes), "port", 1
    for b in bodies:
        cw.add_body(b)
    cw.step_substeps(2 * FRAME)  # warm-up
    frames = max(1, sample_substeps // FRAME // runs)
    rates, total = [], 0.0
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(frames):
            cw.step_substeps(FRAME)
        dt = time.perf_counter() - t0
        total += dt
        rates.append(len(bodies[0].beams) * frames * FRAME / dt)
    n = frames * FRAME
    return {"value": statistics.median(rates), "unit": "beam updates/s", "cores": threads, "kind": kind,
            "cpu_model": cpu_model(), "runs": [round(r, 1) for r in rates],
            "sample": f"median of {runs} samples of {n} substeps each ({frames} frames of {FRAME}) of the same scene "
                      f"(lattice and 2000 balls), after a 2-frame warm-up; {total:.1f} s timed"}


def batched_leg(steps, retune, log):
    """The batched scene (512 hollow crates, lognormal sizes, mean degree 7, inter-body contacts on, 200 x 200 m):
    beam updates/s in frames of 32 substeps, with the autotuner's form for it."""
    from beamlab_amd import scenes
    t0 = time.perf_counter()
    tune = tuned(lambda: build_world(scenes.config4_scene()), None, retune, log)
    bodies = scenes.config4_scene()
    nb, nn = sum(len(b.beams) for b in bodies), sum(len(b.nodes) for b in bodies)
    w = build_world(bodies)
    w.set_tuning(**tune)
    log(f"batched scene ready ({time.perf_counter() - t0:.1f} s): {len(bodies)} bodies, {nn} nodes, {nb} beams")
    w.step_substeps(64)
    t_ms = 0.0
    done = 0
    st0 = w.stats()
    t0 = time.perf_counter()
    while done < steps:
        k = min(32, steps - done)
        t_ms += w.time_substeps(k)
        done += k
    wall = time.perf_counter() - t0
    st = w.stats()
    tun = w.get_tuning()
    w.close()
    contacts = (st["contacts"] - st0["contacts"]) / done
    if contacts <= 0:
        raise RuntimeError("the batched scene's timed substeps had no inter-body contacts")
    alg = 52 * nb + 56 * nn  # the beam-path substep's algorithmic bytes (beam records + CSR, node state)
    achieved = alg / (t_ms / 1e3 / done) / 1e9
    return {"value": nb * done / wall, "unit": "beam updates/s", "ms_per_step": wall / done * 1e3,
            "stream_ms_per_step": t_ms / done, "steps": done,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "the whole substep",
                         "alg_bytes_per_substep": alg},
            "config": {"workload": "512 hollow crates (random struts, mean degree 7), lognormal sizes (median 2000, "
                                   "500-8000 nodes), in stacks of two (the smaller resting on the larger) over a "
                                   "200 x 200 m arena, inter-body contacts", "bodies": len(bodies),
                       "contacts_per_substep": contacts,
                       "nodes": nn, "beams": nb, "frame_substeps": 32,
                       "form": {1: "stepped", 2: "persistent", 3: "resident", 4: "halo"}.get(tun.mode, "?"),
                       "block": tun.block, "incidences_per_thread": tun.incidences_per_thread,
                       "substeps_per_launch": tun.substeps_per_launch, "graph": tun.use_graph,
                       "contacts": st["contacts"], "pair_rebuilds": st["pair_rebuilds"]}}


LEARNED_K = 8  # PCG iterations before the Cholesky fallback, preconditioned by the last factor


def learned_run(net, steps, warm, precond, max_iter, log, name):
    """`steps` timed steps of the plate, each solved by PCG (warm start, preconditioner, K) with the Cholesky as the
    fallback, after 3 Cholesky steps (the first factor); the same seeded loads for every variant."""
    import numpy as np
    from beamlab_amd import fem, learned
    G = fem.ShellBody.plate(FEM_CELLS, FEM_CELLS)
    loads = fem_loads(G.nn, G.fixed, 3 + steps + 9, seed=21)
    for s in range(3):
        G.step(loads[s], FEM_H)
    reps = []
    t0 = time.perf_counter()
    for s in range(steps):
        reps.append(learned.pcg_step(G, net, loads[3 + s], h=FEM_H, max_iter=max_iter, tol=1e-6, warm=warm,
                                     precond=precond))
    wall = time.perf_counter() - t0
    mean = lambda k: float(np.mean([r[k] for r in reps]))
    out = {"value": steps / wall, "steps": steps, "ms_per_step": wall / steps * 1e3, "max_iter": max_iter,
           "converged_rate": mean("converged"), "fallback_rate": mean("fallback_used"),
           "iterations_mean": mean("iterations"), "residual_max": max(r["residual"] for r in reps),
           "phases_ms": {k: mean(k + "_ms") for k in ("assemble", "gnn", "pcg", "fallback", "update")}}
    log(f"learned [{name}]: {out['value']:.1f} solves/s, {out['iterations_mean']:.2f} iterations, "
        f"fallback {out['fallback_rate']:.2f}")
    return G, loads, out


def learned_leg(steps, plain_solves_per_s, log):
    """The learned warm start on the plate workload. The headline: each step solved by PCG from the trained network's
    prediction, preconditioned by the engine's last Cholesky factor (factor reuse between solves), at most LEARNED_K
    iterations to 1e-6, the supernodal Cholesky as the fallback (which refreshes the factor). Beside it, timed the
    same way: that PCG from the previous step's solution and from zero (what the network gains), and block-Jacobi PCG
    from the network's prediction with K = 50. Then, untimed, the network's error per solve against the FP64
    Cholesky and the iterations an uncapped block-Jacobi PCG needs."""
    from beamlab_amd import abi, learned
    net = learned.Network.load()
    G, loads, head = learned_run(net, steps, abi.WARM_NETWORK, abi.PRECOND_FACTOR, LEARNED_K, log, "network, factor")
    checks = [learned.pcg_step(G, net, loads[3 + steps + k], h=FEM_H, max_iter=LEARNED_K, tol=1e-6,
                               warm=abi.WARM_NETWORK, precond=abi.PRECOND_FACTOR, check_error=True) for k in range(8)]
    G.close()
    variants = {}
    for name, warm, prec, k, n in (("previous_warm_factor", abi.WARM_PREVIOUS, abi.PRECOND_FACTOR, LEARNED_K, steps),
                                   ("zero_warm_factor", abi.WARM_ZERO, abi.PRECOND_FACTOR, LEARNED_K, steps),
                                   ("network_warm_block_jacobi", abi.WARM_NETWORK, abi.PRECOND_BLOCK_JACOBI, 50,
                                    max(4, steps // 5))):
        Gv, lv, out = learned_run(net, n, warm, prec, k, log, name)
        if name == "network_warm_block_jacobi":
            out["per_solve_error_vs_fp64_cholesky"] = [
                {"warm_start": c["warm_error"], "pcg": c["pcg_error"], "fallback_used": c["fallback_used"]}
                for c in (learned.pcg_step(Gv, net, lv[3 + n + q], h=FEM_H, max_iter=50, tol=1e-6,
                                           warm=abi.WARM_NETWORK, precond=abi.PRECOND_BLOCK_JACOBI, check_error=True)
                          for q in range(4))]
            uncapped = {}
            for uname, warm_u in (("network", abi.WARM_NETWORK), ("zero", abi.WARM_ZERO)):
                state = Gv.state()
                r = learned.pcg_step(Gv, net, lv[3 + n + 4], h=FEM_H, max_iter=20000, tol=1e-6, warm=warm_u,
                                     precond=abi.PRECOND_BLOCK_JACOBI, fallback=False)
                Gv.set_state(*state)
                uncapped[uname] = {"iterations": r["iterations"], "converged": r["converged"], "pcg_ms": r["pcg_ms"]}
            out["uncapped_pcg_to_1e-6"] = uncapped
        Gv.close()
        out["vs_plain_cholesky"] = out["value"] / plain_solves_per_s if plain_solves_per_s else None
        variants[name] = out
    head.update({
        "unit": "implicit FEM solves/s", "dtype": "bf16 network (fp32 sums), f64 PCG and fallback",
        "config": {"workload": f"the {FEM_CELLS}x{FEM_CELLS} plate of implicit_fem", "network": "4 layers x 64, "
                   "trained on 20000 Cholesky-solved steps of a 64x64 plate (beamlab_amd/gnn_plate.npz)",
                   "pcg": "warm start from the network, preconditioned by the last Cholesky factor",
                   "max_iter": LEARNED_K, "tol": 1e-6, "fallback": "supernodal Cholesky (refreshes the factor)"},
        "vs_plain_cholesky": head["value"] / plain_solves_per_s if plain_solves_per_s else None,
        "per_solve_error_vs_fp64_cholesky": [{"warm_start": c["warm_error"], "pcg": c["pcg_error"],
                                              "iterations": c["iterations"], "fallback_used": c["fallback_used"]}
                                             for c in checks],
        "variants": variants})
    return head


def fem_loads(nn, fixed, steps, seed=7, per_step=32, sigma=50.0):
    """The seeded random point loads of the plate workload: per step `per_step` free nodes, a force normal to the
    plate (y) of N(0, sigma) N."""
    import numpy as np
    rng = np.random.default_rng(seed)
    free = np.nonzero(fixed == 0)[0]
    out = np.zeros((steps, nn, 3), np.float32)
    for s in range(steps):
        idx = rng.choice(free, per_step, replace=False)
        out[s, idx, 1] = rng.normal(0.0, sigma, per_step).astype(np.float32)
    return out


def fem_leg(steps, warmup, log):
    """Implicit FEM solves/s on the plate (see the module docstring)."""
    import numpy as np
    from beamlab_amd import fem
    G = fem.ShellBody.plate(FEM_CELLS, FEM_CELLS)
    loads = fem_loads(G.nn, G.fixed, warmup + steps)
    for s in range(warmup):
        G.step(loads[s], FEM_H)
    reps = []
    t0 = time.perf_counter()
    for s in range(steps):
        reps.append(G.step(loads[warmup + s], FEM_H))
        if (s + 1) % 50 == 0:
            log(f"fem step {s + 1}/{steps}")
    wall = time.perf_counter() - t0
    G.close()
    mean = lambda k: float(np.mean([r[k] for r in reps]))
    flops = reps[0]["factor_flops"]
    fac_s = mean("factor_ms") / 1e3
    achieved = flops / fac_s / 1e12
    return {
        "value": steps / wall, "unit": "implicit FEM solves/s", "ms_per_step": wall / steps * 1e3, "steps": steps,
        "warmup": warmup, "dtype": "f64",
        "config": {"workload": f"co-rotational shell plate {FEM_CELLS}x{FEM_CELLS} cells, clamped edges, seeded "
                               "random point loads", "nodes": int(G.nn), "triangles": int(G.nt),
                   "unknowns": 6 * int((G.fixed == 0).sum()), "h_s": FEM_H, "theta": 0.5, "dissipation": 0.02},
        "residual_max": max(r["residual"] for r in reps), "failures": sum(r["failures"] for r in reps),
        "phases_ms": {k: mean(k + "_ms") for k in ("assemble", "factor", "solve", "update")},
        "factorization": {"supernodes": reps[0]["supernodes"], "largest_front": reps[0]["fronts_max"],
                          "factor_nnz": reps[0]["factor_nnz"], "flops": flops},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": None, "kernel": "k_fstep (all panel steps)",
                     "alg_flops_per_factorization": flops, "factor_ms": fac_s * 1e3},
    }


def cpu_fem_worker(cells, steps):
    """The compiled reference's FemFrame step on the plate (child process of the bench): prints one JSON line."""
    from oracle import oracle
    from beamlab_amd import fem
    p, t, fx = fem.plate_mesh(cells, cells)
    loads = fem_loads(len(p), fx, steps + 1)
    t0 = time.perf_counter()
    P = oracle.RefPlate(cells, cells)
    build = time.perf_counter() - t0
    P.step(loads[0], FEM_H)  # warm step
    t1 = time.perf_counter()
    for s in range(steps):
        P.step(loads[1 + s], FEM_H)
    dt = time.perf_counter() - t1
    print(json.dumps({"value": steps / dt, "unit": "implicit FEM solves/s", "cores": 1, "kind": "reference",
                      "sample": f"{steps} step(s) of the same {cells}x{cells} plate after one warm step: "
                                f"FemFrame::compute_forces + FemFrame::solve, one thread ({dt:.1f} s; the "
                                f"reference's analysis and first factorization took {build:.1f} s)"}), flush=True)


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (the runtime's copy kernel), ten times: the known-size reference for
    scaling rocprof's FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0 and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(10):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a), hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 10}), flush=True)


def measured_traffic(tun):
    """HBM bytes per launch of the force kernel as rocprof's FETCH_SIZE / WRITE_SIZE counted them (scaled by a plain
    copy's counts), from the committed profile of the same kernel form; None for another form."""
    path = os.path.join(ROOT, "profiles", "k_forces_c2_scene_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t["traffic_bytes_per_launch"] if t.get("kernel") == f"k_forces<{tun.block}, {tun.incidences_per_thread}>" else None


def tuned(make, candidates, retune, log):
    """The autotuner's choice for this workload (cached by shape), found on a scratch world."""
    from beamlab_amd import autotune
    probe = make()
    tune = None if retune else autotune.cached(probe)
    if tune is None:
        tune, _ = autotune.autotune(probe, substeps=100, candidates=candidates, log=log)
    probe.close()
    return tune


def beam_path_only(steps, retune, log):
    """The lattice alone on the ground (no contact geometry): the persistent beam kernel's throughput and roofline."""
    from beamlab_amd import scenes
    tune = tuned(lambda: build_world([scenes.config2_body(seed=2)]), None, retune, log)
    body = scenes.config2_body(seed=2)
    w = build_world([body])
    w.set_tuning(**tune)
    w.step_substeps(200)
    t_ms = sum(w.time_substeps(400) for _ in range(steps // 400))
    tot_ms, ker_ms, launches = w.time_kernels(400)
    nb, nn = len(body.beams), len(body.nodes)
    alg = (52 * nb + 56 * nn) * (400 / launches)
    avg = ker_ms / launches / 1e3
    tun = w.get_tuning()
    w.close()
    return {"value": nb * (steps // 400) * 400 / (t_ms / 1e3), "unit": "beam updates/s",
            "kernel": {1: "k_substep", 2: "k_body_persistent", 3: "k_resident", 4: "k_halo"}.get(tun.mode, "?"),
            "achieved_GBs": alg / avg / 1e9, "frac": alg / avg / 1e9 / HBM_PEAK_GBS, "avg_launch_us": avg * 1e6,
            "alg_bytes_per_launch": alg}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-sample", type=int, default=3300, help="substeps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-beam-only", action="store_true")
    ap.add_argument("--retune", action="store_true", help="ignore the autotuner's cache")
    ap.add_argument("--copy-calib", type=int, default=0, help="only run a plain copy of this many MiB (profiling)")
    ap.add_argument("--fem-steps", type=int, default=200)
    ap.add_argument("--batched-steps", type=int, default=2000)
    ap.add_argument("--learned-steps", type=int, default=200)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-fem", action="store_true")
    ap.add_argument("--no-beams", action="store_true", help="the implicit-FEM leg only")
    ap.add_argument("--cpu-fem-worker", type=int, nargs=2, metavar=("CELLS", "STEPS"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_fem_worker:
        cpu_fem_worker(*args.cpu_fem_worker)
        return
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the step is single-GPU in this build",
              file=sys.stderr)
        sys.exit(2)

    from beamlab_amd import scenes
    log = lambda m: print(m, file=sys.stderr, flush=True)
    # the reference's FEM step on the CPU, beside the GPU legs (a child process, one core)
    fem_cpu = None
    if not args.no_cpu and not args.no_fem:
        import subprocess
        fem_cpu = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-fem-worker", str(FEM_CELLS), "1"],
                                   stdout=subprocess.PIPE, text=True)
    fem_out = None
    if not args.no_fem:
        fem_out = fem_leg(args.fem_steps, 3, log)
        log(f"fem: {fem_out['value']:.1f} solves/s")
        if not args.no_learned:
            fem_out["learned"] = learned_leg(args.learned_steps, fem_out["value"], log)
    if args.no_beams:
        out = {"metric": METRIC, "value": fem_out["value"], "unit": fem_out["unit"], "n_gpus": 1,
               "steps": args.fem_steps, "warmup": 3, "ms_per_step": fem_out["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": fem_out["config"],
               "roofline": fem_out["roofline"], "implicit_fem": fem_out}
        if fem_cpu is not None:
            fem_out["cpu_baseline"] = wait_cpu_fem(fem_cpu, log)
            out["cpu_baseline"] = fem_out["cpu_baseline"]
        print(json.dumps(out), flush=True)
        return
    tune = tuned(lambda: build_world(scenes.config2_scene()), CONTACT_CANDIDATES, args.retune, log)
    bodies = scenes.config2_scene()
    nb, nn = len(bodies[0].beams), sum(len(b.nodes) for b in bodies)
    w = build_world(bodies)
    w.set_tuning(**tune)
    done = 0
    while done < args.warmup:
        k = min(FRAME, args.warmup - done)
        w.step_substeps(k)
        done += k
    # timed region: frames of 33 substeps, inputs resident; the wall clock around all of it (the host's frame work
    # included), each frame's time on the stream beside it
    t_ms = 0.0
    done = 0
    st0 = w.stats()
    t0 = time.perf_counter()
    while done < args.steps:
        k = min(FRAME, args.steps - done)
        t_ms += w.time_substeps(k)
        done += k
    wall = time.perf_counter() - t0
    st = w.stats()
    contacts = (st["contacts"] - st0["contacts"]) / args.steps
    if contacts <= 0:
        raise RuntimeError("the contact scene's timed substeps had no contacts")
    # the dominant kernel: the force phase (beams, yield and break, gather into F), timed launch by launch
    tot_ms, ker_ms, launches = w.time_kernels(FRAME)
    avg_launch_s = ker_ms / launches / 1e3
    alg_bytes = 52 * nb + 44 * nn  # beam records + CSR, node p / v / inv_mass read, F written, CSR offset
    achieved = alg_bytes / avg_launch_s / 1e9
    tun = w.get_tuning()
    w.close()
    out = {
        "metric": METRIC,
        "value": nb * args.steps / wall,
        "unit": "beam updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64x32x24 beam lattice, faces as triangles, self contact, 2000 rigid balls, ground box",
                   "nodes": nn, "beams": nb, "bodies": len(bodies), "substep_s": 0.0005, "substeps": args.steps,
                   "frame_substeps": FRAME, "block": tun.block, "incidences_per_thread": tun.incidences_per_thread,
                   "wall_s": wall, "stream_ms_per_step": t_ms / args.steps, "contacts_per_substep": contacts,
                   "contacts": st["contacts"], "pair_rebuilds": st["pair_rebuilds"],
                   "narrow_tests": st["narrow_tests"], "broken_beams": st["broken_beams"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(tun), "kernel": "k_forces",
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_us": avg_launch_s * 1e6,
                     "share_of_step": ker_ms / tot_ms},
    }
    if not args.no_beam_only:
        out["beam_path_only"] = beam_path_only(4000, args.retune, log)
    if not args.no_batched:
        out["batched"] = batched_leg(args.batched_steps, args.retune, log)
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16) - (1 if fem_cpu is not None else 0)
        out["cpu_baseline"] = cpu_baseline(scenes.config2_scene(), args.cpu_sample, max(1, threads))
    if fem_out is not None:
        if fem_cpu is not None:
            fem_out["cpu_baseline"] = wait_cpu_fem(fem_cpu, log)
        out["implicit_fem"] = fem_out
    print(json.dumps(out), flush=True)


def wait_cpu_fem(proc, log, limit_s=900):
    """The reference FEM worker's line (it keeps running beside the GPU legs; progress every 30 s meanwhile)."""
    import subprocess
    t0 = time.perf_counter()
    while True:
        try:
            line = proc.communicate(timeout=30)[0]
            break
        except subprocess.TimeoutExpired:
            if time.perf_counter() - t0 > limit_s:
                proc.kill()
                return None
            log(f"waiting for the reference FEM step on the CPU ({time.perf_counter() - t0:.0f} s)")
    try:
        return json.loads(line.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return None


if __name__ == "__main__":
    main()

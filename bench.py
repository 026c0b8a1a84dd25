"""Benchmark of the soft-body step on one MI355X: beam updates per second at the 2 kHz substep.

Workload (the single-large-body configuration): the 64x32x24 lattice (49 152 nodes, 420 848 beams, per-beam k
log-uniform in [1e5, 5e6] N/m, 2 % of the beams weakened to break strain 0.01) with its faces as collision
triangles and self contact on, resting on a concrete ground box, and 2 000 rigid balls (radius 0.1-0.3 m, 5 kg) on
its top and around it; inter-body contacts through the spatial hash (pair rebuild every 17 substeps, narrow phase
and ordered response every substep). 4000 substeps of 0.5 ms in frames of 33 substeps (the app's 60 Hz frame). One
step = one substep of every beam, node and contact. Inputs are resident in HBM before the timed region.

Prints one JSON line with the throughput, the dominant kernel's HBM roofline figures (achieved = algorithmic bytes
per launch / its average duration, from HIP events around each launch on the step's stream) and the compiled
reference's CPU step timed on a bounded sample of the same workload on this host's cores. The beam path alone (the
same lattice on the ground without contact geometry, where the step runs as one persistent kernel) is reported
beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "beam updates/s at 2 kHz substep; implicit FEM solves/s at 1e-6 residual"
FRAME = 33
# the forms the contact step can take: its force kernel's workgroup size and beam ends per thread
CONTACT_CANDIDATES = [dict(mode=1, block=256, incidences_per_thread=p, substeps_per_launch=0, use_graph=0)
                      for p in (1, 2, 4, 8)] + \
                     [dict(mode=1, block=512, incidences_per_thread=p, substeps_per_launch=0, use_graph=0) for p in (2, 4)]


def build_world(bodies):
    from beamlab_amd import scenes
    w = scenes.World()
    scenes.ground_box(w)
    for b in bodies:
        w.add_body(b)
    return w


def cpu_baseline(bodies, sample_substeps, threads):
    """The compiled reference's World::step_substeps (oracle/_ref) on the same scene; the C restatement when the
    reference library is absent."""
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
    cw.step_substeps(FRAME)  # warm-up
    frames = sample_substeps // FRAME
    t0 = time.perf_counter()
    for _ in range(frames):
        cw.step_substeps(FRAME)
    dt = time.perf_counter() - t0
    n = frames * FRAME
    return {"value": len(bodies[0].beams) * n / dt, "unit": "beam updates/s", "cores": threads, "kind": kind,
            "sample": f"{n} substeps of the same scene (lattice and 2000 balls) in {frames} frames of {FRAME} "
                      f"({dt:.1f} s)"}


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
    args = ap.parse_args()
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the step is single-GPU in this build",
              file=sys.stderr)
        sys.exit(2)

    from beamlab_amd import scenes
    log = lambda m: print(m, file=sys.stderr, flush=True)
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
    # timed region: frames of 33 substeps, inputs resident (each frame's wall time on the stream)
    t_ms = 0.0
    done = 0
    t0 = time.perf_counter()
    while done < args.steps:
        k = min(FRAME, args.steps - done)
        t_ms += w.time_substeps(k)
        done += k
    wall = time.perf_counter() - t0
    st = w.stats()
    # the dominant kernel: the force phase (beams, yield and break, gather into F), timed launch by launch
    tot_ms, ker_ms, launches = w.time_kernels(FRAME)
    avg_launch_s = ker_ms / launches / 1e3
    alg_bytes = 52 * nb + 44 * nn  # beam records + CSR, node p / v / inv_mass read, F written, CSR offset
    achieved = alg_bytes / avg_launch_s / 1e9
    tun = w.get_tuning()
    w.close()
    out = {
        "metric": METRIC,
        "value": nb * args.steps / (t_ms / 1e3),
        "unit": "beam updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64x32x24 beam lattice, faces as triangles, self contact, 2000 rigid balls, ground box",
                   "nodes": nn, "beams": nb, "bodies": len(bodies), "substep_s": 0.0005, "substeps": args.steps,
                   "frame_substeps": FRAME, "block": tun.block, "incidences_per_thread": tun.incidences_per_thread,
                   "wall_s": wall, "contacts": st["contacts"], "pair_rebuilds": st["pair_rebuilds"],
                   "narrow_tests": st["narrow_tests"], "broken_beams": st["broken_beams"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_forces",
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_us": avg_launch_s * 1e6,
                     "share_of_step": ker_ms / tot_ms},
    }
    if not args.no_beam_only:
        out["beam_path_only"] = beam_path_only(4000, args.retune, log)
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(scenes.config2_scene(), args.cpu_sample, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

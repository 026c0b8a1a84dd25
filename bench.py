"""Benchmark of the soft-body step on one MI355X: beam updates per second at the 2 kHz substep.

Workload: the 64x32x24 lattice (49 152 nodes, 420 848 beams, per-beam k log-uniform in [1e5, 5e6] N/m, 2 % of the
beams weakened to break strain 0.01) dropped onto a concrete ground box, 4000 substeps of 0.5 ms. One step = one
substep of every beam and node (beam forces, yield and break, gather, ground contact, integration). Inputs are
resident in HBM before the timed region.

Prints one JSON line with the throughput, the dominant kernel's HBM roofline figures (achieved = algorithmic bytes
per substep launch / its average duration, from HIP events around each launch on the step's stream) and the
compiled reference's CPU step timed on a bounded sample of the same workload on this host's cores.
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


def build_world(body):
    from beamlab_amd import scenes
    w = scenes.World()
    scenes.ground_box(w)
    w.add_body(body)
    return w


def cpu_baseline(body, sample_substeps, threads):
    """The compiled reference's World::step_substeps (oracle/_ref) on the same body; the C restatement when the
    reference library is absent."""
    from beamlab_amd import abi, scenes
    from oracle import oracle
    w = scenes.World()
    scenes.ground_box(w)
    if oracle.have_ref():
        cw, kind = oracle.RefWorld(abi.default_settings(), w.statics.boxes, threads=threads), "reference"
    else:
        cw, kind, threads = oracle.OracleWorld(abi.default_settings(), w.statics.boxes), "port", 1
    cw.add_body(body)
    cw.step_substeps(50)  # warm-up
    frames = sample_substeps // 33
    t0 = time.perf_counter()
    for _ in range(frames):
        cw.step_substeps(33)  # 60 Hz frames of 33 substeps, as the app steps
    dt = time.perf_counter() - t0
    n = frames * 33
    return {"value": len(body.beams) * n / dt, "unit": "beam updates/s", "cores": threads, "kind": kind,
            "sample": f"{n} substeps of the same lattice in {frames} frames of 33 ({dt:.1f} s)"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-sample", type=int, default=3300, help="substeps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
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

    from beamlab_amd import autotune, scenes
    body = scenes.config2_body(seed=2)
    nb, nn = len(body.beams), len(body.nodes)
    # the execution form: the autotuner on a scratch world of the same workload, then a fresh world from the seed
    tune = None if args.retune else autotune.cached(build_world(body))
    if tune is None:
        scratch = build_world(scenes.config2_body(seed=2))
        tune, table = autotune.autotune(scratch, substeps=200, log=lambda m: print(m, file=sys.stderr, flush=True))
        scratch.close()
    w = build_world(body)
    w.set_tuning(**tune)
    w.step_substeps(args.warmup)
    # timed region: the steps in frames, inputs resident
    frame = 400
    t_ms = 0.0
    done = 0
    while done < args.steps:
        k = min(frame, args.steps - done)
        t_ms += w.time_substeps(k)
        done += k
    tun = w.get_tuning()
    # the dominant kernel's launches timed one by one (no graph): algorithmic bytes / its average duration
    tot_ms, ker_ms, launches = w.time_kernels(400)
    avg_launch_s = ker_ms / launches / 1e3
    substeps_per_launch = 400 / launches
    alg_bytes = (52 * nb + 56 * nn) * substeps_per_launch
    achieved = alg_bytes / avg_launch_s / 1e9
    value = nb * args.steps / (t_ms / 1e3)
    out = {
        "metric": METRIC,
        "value": value,
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
        "config": {"workload": "64x32x24 beam lattice on a ground box (no self or sphere contacts)",
                   "nodes": nn, "beams": nb, "substep_s": 0.0005, "substeps": args.steps,
                   "mode": {1: "stepped", 2: "persistent", 3: "resident", 4: "halo"}.get(tun.mode, str(tun.mode)),
                   "block": tun.block, "incidences_per_thread": tun.incidences_per_thread,
                   "substeps_per_launch": tun.substeps_per_launch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": {1: "k_substep", 2: "k_body_persistent", 3: "k_resident", 4: "k_halo"}.get(tun.mode, "?"),
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_us": avg_launch_s * 1e6},
    }
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(scenes.config2_body(seed=2), args.cpu_sample, threads)
    w.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

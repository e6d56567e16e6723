"""Benchmark of the particle step on one MI355X.

Workload (default): 2-D Weibel, 2048 x 2048 cells, 32 particles per cell per
species, 2 species (268,435,456 particles), Courant 0.45, c/omega_p = 10
cells, drift +-0.5c, thermal spread 0.01c.  --config weibel3d_256: 3-D
Weibel, 256^3 cells, 8 particles per cell per species (268,435,456
particles), drift +-0.3c, thermal spread 0.02c.  A step is one whole time step
(B half step, push + deposit + tile binning, B half step, current halo fold,
E update + AddCurrent, halos) with all inputs resident in HBM.

value = particles * steps / wall time of the timed steps (FP32 split form).
The FP64 form runs the same workload and is reported beside it.
roofline: the push/deposit kernel, timed with HIP events on its own stream;
algorithmic bytes per launch = particles * bytes per particle + cells * grid
bytes (BYTES_PER_PARTICLE, GRID_BYTES below).  cpu_baseline: the FP64 C
oracle (restated reference CPU path, OpenMP) on a bounded number of steps of
the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X

CONFIGS = {
    # name: (ndim, nx, ny, nz, epc per species, drift, thermal spread)
    "weibel2d_2048": (2, 2048, 2048, 1, 32, 0.5, 0.01),
    "weibel2d_512": (2, 512, 512, 1, 16, 0.5, 0.01),
    "weibel2d_256": (2, 256, 256, 1, 16, 0.5, 0.01),
    "weibel3d_256": (3, 256, 256, 256, 8, 0.3, 0.02),
    "weibel3d_128": (3, 128, 128, 128, 8, 0.3, 0.02),
}

# algorithmic bytes per particle update of the push kernel (read + write)
BYTES_PER_PARTICLE = {"fp32_split": 2 * (4 + 3 * 4 + 3 * 4), "fp64": 7 * 8 + 4 + 6 * 8}
# grid bytes per cell of the push kernel: read E,B (6 values), write J (3 FP64 values)
GRID_BYTES = {"fp32_split": 6 * 4 + 3 * 8, "fp64": 6 * 8 + 3 * 8}


def measured_traffic(form, workload, tile, ndim):
    # HBM bytes per launch of the push kernel from the committed rocprofv3 counter runs
    # (FETCH_SIZE x 2.0 + WRITE_SIZE, see the file named by "source"), or None
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(f"k_push{ndim}d_" + ("split" if form == "fp32_split" else "f64"))
    if e and e.get("workload") == workload and e.get("tile") == tile:
        return e["bytes_per_launch"]
    return None


def run_gpu(form, ndim, nx, ny, nz, epc, P, steps, warmup, calibrate=False, tile=0):
    from pictor_amd import Simulation

    sim = Simulation(nx, ny, nz, ndim=ndim, epc=epc, c=0.45, c_ompe=10.0, form=form, tile=tile)
    t0 = time.time()
    sim.upload_particles(**P)
    sim.synchronize()
    t_upload = time.time() - t0
    np_ = sim.particle_count()
    sim.step(warmup)
    sim.synchronize()
    sim.timer_reset()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    dt = time.perf_counter() - t0
    push_ms, npush = sim.timer("push")
    fld_ms, _ = sim.timer("fields")
    step_ms, _ = sim.timer("step")
    diag = sim.diagnostics()
    mem = sim.memory_bytes()
    if calibrate:
        # plain streaming kernels of known size for scaling HBM counters:
        # FP32 -> FP64 field conversion (read 4 B, write 8 B per cell) and
        # current scaling (read 8 B, write 8 B per cell), 9 arrays of (mx*my)
        sim.download_fields()
    tile = sim.tile
    sim.close()
    avg_push = push_ms / max(npush, 1) / 1e3
    ncell = nx * ny * (nz if ndim == 3 else 1)
    bytes_launch = np_ * BYTES_PER_PARTICLE[form] + ncell * GRID_BYTES[form]
    achieved = bytes_launch / avg_push
    return {
        "value": np_ * steps / dt,
        "ms_per_step": dt / steps * 1e3,
        "particles": np_,
        "push_ms": avg_push * 1e3,
        "field_ms": fld_ms / max(npush, 1),
        "event_step_ms": step_ms / max(npush, 1),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_launch": bytes_launch},
        "diag": diag,
        "upload_s": t_upload,
        "device_bytes": mem,
        "tile": tile,
    }


def run_cpu(ndim, nx, ny, nz, epc, P, steps, threads):
    from oracle.pyoracle import Oracle

    o = Oracle(ndim, nx, ny, nz, epc=epc, c=0.45, c_ompe=10.0, nthreads=threads, cap=int(len(P["x"]) * 1.05) + 4096)
    o.load_particles(**P)
    o.step(1)  # warm-up step (first touch of the work arrays)
    t0 = time.perf_counter()
    o.run(2, steps)
    dt = time.perf_counter() - t0
    return len(P["x"]) * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="weibel2d_2048", choices=sorted(CONFIGS))
    ap.add_argument("--forms", default="fp32_split,fp64")
    ap.add_argument("--cpu-steps", type=int, default=-1, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--calibrate", action="store_true", help="also run plain copy kernels (counter scaling)")
    ap.add_argument("--tile", type=int, default=0, help="tile edge in cells (0 = library default)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not code)", file=sys.stderr)
        sys.exit(2)

    import numpy as np  # noqa: F401
    from pictor_amd import setup_weibel

    ndim, nx, ny, nz, epc, drift, uth = CONFIGS[args.config]
    t0 = time.time()
    P = setup_weibel(nx, ny, nz, ndim=ndim, epc=epc, drift=drift, uth=uth, seed=12345)
    print(f"# setup {len(P['x'])} particles in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    forms = [f for f in args.forms.split(",") if f]
    res = {}
    for form in forms:
        res[form] = run_gpu(form, ndim, nx, ny, nz, epc, P, args.steps, args.warmup, args.calibrate, args.tile)
        r = res[form]
        r["roofline"]["traffic"] = measured_traffic(form, args.config, r["tile"], ndim)
        print(f"# {form}: {r['value']:.4e} particle-updates/s, {r['ms_per_step']:.3f} ms/step, "
              f"push {r['push_ms']:.3f} ms ({r['roofline']['frac'] * 100:.1f}% HBM)", file=sys.stderr, flush=True)

    cpu = None
    big = len(P["x"]) > 100_000_000
    cpu_steps = args.cpu_steps if args.cpu_steps >= 0 else (2 if ndim == 3 and big else (5 if big else 20))
    if cpu_steps > 0:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        v, dt = run_cpu(ndim, nx, ny, nz, epc, P, cpu_steps, threads)
        cpu = {"value": v, "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "sample": f"{cpu_steps} full steps of the same workload ({len(P['x'])} particles, FP64 oracle, "
                         f"OpenMP) after 1 warm-up step, {dt:.1f}s"}
        print(f"# cpu: {v:.4e} particle-updates/s on {threads} threads ({dt:.1f}s)", file=sys.stderr, flush=True)

    head = forms[0]
    r = res[head]
    line = {
        "metric": "particle-updates/s/GPU, 2D Weibel: FP32-split vs FP64 GPU; adjoint/forward cost",
        "value": r["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32_split" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {ndim}-D Weibel "
                               + ("x".join(str(v) for v in ((nx, ny, nz) if ndim == 3 else (nx, ny))))
                               + f", {epc} ppc/species, 2 species, {r['particles']} particles, drift +-{drift}c, "
                               f"thermal {uth}c, c=0.45, c/wp=10 cells",
                   "form": head, "tile": r["tile"], "particles": r["particles"]},
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
    }
    for form in forms[1:]:
        o = res[form]
        line[f"{form}_form"] = {"value": o["value"], "ms_per_step": o["ms_per_step"], "roofline": o["roofline"],
                                "dtype": "f64" if form == "fp64" else "f32"}
    if "fp64" in res and head != "fp64":
        line["split_over_fp64"] = r["value"] / res["fp64"]["value"]
    line["detail"] = {f: {k: res[f][k] for k in ("push_ms", "field_ms", "event_step_ms", "upload_s", "device_bytes",
                                                  "diag")} for f in forms}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

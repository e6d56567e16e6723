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
bytes (bytes_per_particle, GRID_BYTES below).  cpu_baseline: the FP64 C
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

# algorithmic bytes per particle update of the push kernel (read + write): the
# record a particle needs; FP64 the reference SoA (q,x,y,z,u,v,w + flv read, x..w
# written); split: cell word + in-plane FP32 offsets + FP32 momenta, species
# implicit in the bin (2-D: cell, ox, oy, u, v, w; 3-D: also oz)
def bytes_per_particle(form, ndim):
    if form == "fp64":
        return 7 * 8 + 4 + 6 * 8
    return 2 * (4 + ndim * 4 + 3 * 4)


# grid bytes per cell of the push kernel: read E,B (6 values), write J (3 values), at
# the field precision s (8 FP64, 4 split)
GRID_BYTES = {"fp32_split": 9 * 4, "fp64": 9 * 8}


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


def run_gpu(form, ndim, nx, ny, nz, epc, P, steps, warmup, calibrate=False, tile=0, curr_filter=0, late=None):
    """Upload, `warmup` untimed steps, `steps` timed steps.  late = (start, n, total):
    continue untimed to step `start`, time steps start+1 .. start+n, continue to step
    `total`, and report the late window and the mean over all `total` steps."""
    from pictor_amd import Simulation

    sim = Simulation(nx, ny, nz, ndim=ndim, epc=epc, c=0.45, c_ompe=10.0, form=form, tile=tile,
                     curr_filter=curr_filter)
    t0 = time.time()
    sim.upload_particles(**P)
    sim.synchronize()
    t_upload = time.time() - t0
    np_ = sim.particle_count()
    t_run0 = time.perf_counter()
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
    late_res = None
    if late:
        start, n, total = late
        done = warmup + steps
        sim.step(start - done)
        sim.synchronize()
        sim.timer_reset()
        t1 = time.perf_counter()
        sim.step(n)
        sim.synchronize()
        dt_late = time.perf_counter() - t1
        lpush, lnp = sim.timer("push")
        sim.step(total - start - n)
        sim.synchronize()
        t_all = time.perf_counter() - t_run0
        nre, nsp = sim.relayout_stats()
        late_res = {"window": f"steps {start + 1}-{start + n}", "ms_per_step": dt_late / n * 1e3,
                    "push_ms": lpush / max(lnp, 1), "value": np_ * n / dt_late,
                    "mean_ms_per_step": t_all / total * 1e3, "mean_over_steps": total,
                    "relayouts": nre, "spilled_particles": nsp,
                    "particles_end": sim.particle_count()}
    if calibrate:
        # plain streaming kernels of known size for scaling HBM counters:
        # FP32 -> FP64 field conversion (read 4 B, write 8 B per cell) and
        # current scaling (read 8 B, write 8 B per cell), 9 arrays of (mx*my)
        sim.download_fields()
    tile = sim.tile
    sim.close()
    avg_push = push_ms / max(npush, 1) / 1e3
    ncell = nx * ny * (nz if ndim == 3 else 1)
    bytes_launch = np_ * bytes_per_particle(form, ndim) + ncell * GRID_BYTES[form]
    achieved = bytes_launch / avg_push
    out = {
        "value": np_ * steps / dt,
        "ms_per_step": dt / steps * 1e3,
        "particles": np_,
        "push_ms": avg_push * 1e3,
        "field_ms": fld_ms / max(npush, 1),
        "event_step_ms": step_ms / max(npush, 1),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_launch": bytes_launch,
                     "bytes_per_particle": bytes_per_particle(form, ndim),
                     "grid_bytes_per_cell": GRID_BYTES[form]},
        "diag": diag,
        "upload_s": t_upload,
        "device_bytes": mem,
        "tile": tile,
    }
    if late_res:
        late_res["frac"] = bytes_launch / (late_res["push_ms"] / 1e3) / HBM_PEAK
        out["late"] = late_res
    return out


def run_cpu(ndim, nx, ny, nz, epc, P, steps, threads):
    from oracle.pyoracle import Oracle

    o = Oracle(ndim, nx, ny, nz, epc=epc, c=0.45, c_ompe=10.0, nthreads=threads, cap=int(len(P["x"]) * 1.05) + 4096)
    o.load_particles(**P)
    o.step(1)  # warm-up step (first touch of the work arrays)
    t0 = time.perf_counter()
    o.run(2, steps)
    dt = time.perf_counter() - t0
    return len(P["x"]) * steps / dt, dt


def _series_gpu(form, ndim, nx, ny, nz, epc, P, steps, every):
    from pictor_amd import Simulation

    rows = []
    with Simulation(nx, ny, nz, ndim=ndim, epc=epc, c=0.45, c_ompe=10.0, form=form) as sim:
        sim.upload_particles(**P)
        rows.append((0, sim.diagnostics()))
        for t in range(every, steps + 1, every):
            sim.step(every)
            rows.append((t, sim.diagnostics()))
    return rows


def _series_oracle(ndim, nx, ny, nz, epc, P, steps, every, threads):
    from oracle.pyoracle import Oracle

    o = Oracle(ndim, nx, ny, nz, epc=epc, c=0.45, c_ompe=10.0, nthreads=threads,
               cap=int(len(P["x"]) * 1.05) + 4096)
    o.load_particles(**P)

    def diag():
        e = o.energy()
        g = o.gauss()
        return {"B_energy": e[0], "E_energy": e[1], "prtl_energy_1": e[2], "prtl_energy_2": e[3],
                "gauss_residual": g[0], "rho_max": g[1]}

    rows = [(0, diag())]
    for t in range(1, steps + 1):
        o.step(t)
        if t % every == 0:
            rows.append((t, diag()))
    return rows


def _growth_window(ref):
    # linear phase of the instability: magnetic energy between 10x its smallest
    # value and 1/10 of its largest (the reference series decides the window)
    import numpy as np

    t = np.array([r[0] for r in ref], dtype=float)
    eb = np.array([r[1]["B_energy"] for r in ref])
    lo, hi = np.log(eb[eb > 0].min()) + np.log(10.0), np.log(eb.max()) - np.log(10.0)
    le = np.log(np.maximum(eb, 1e-300))
    i_hi = int(np.argmax(le >= hi)) if np.any(le >= hi) else len(le) - 1
    cand = np.nonzero((le >= lo) & (np.arange(len(le)) <= i_hi))[0]
    if len(cand) < 3:
        return None
    return t[cand[0]], t[i_hi]


def _growth_rate(rows, window, ompe):
    import numpy as np

    if window is None:
        return None
    t = np.array([r[0] for r in rows], dtype=float)
    eb = np.array([r[1]["B_energy"] for r in rows])
    m = (t >= window[0]) & (t <= window[1])
    slope = np.polyfit(t[m], np.log(eb[m]), 1)[0]
    return slope / 2.0 / ompe     # amplitude growth rate in units of omega_p


def accuracy_report(args, ndim, nx, ny, nz, epc, P):
    """FP32 split form vs the FP64 forms: field energy, growth rate and Gauss residual over a run."""
    import numpy as np

    steps, every = args.steps, args.every
    ompe = 0.45 / 10.0
    series = {}
    for form in ("fp64", "fp32_split"):
        t0 = time.time()
        series[form] = _series_gpu(form, ndim, nx, ny, nz, epc, P, steps, every)
        print(f"# {form}: {steps} steps in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    # the sensitivity floor: the FP64 form from momenta rounded to FP32 (a 2^-24 relative
    # change of the input); after saturation the Weibel filaments amplify any such change
    Pr = dict(P)
    for k in ("u", "v", "w"):
        Pr[k] = np.asarray(P[k]).astype(np.float32).astype(np.float64)
    t0 = time.time()
    series["fp64_f32_momenta"] = _series_gpu("fp64", ndim, nx, ny, nz, epc, Pr, steps, every)
    print(f"# fp64 from FP32-rounded momenta: {steps} steps in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    # an FP64 ensemble: the same run from momenta perturbed by 2^-24 relative Gaussian noise
    # (FP32's relative resolution), one seed per member; its spread around the reference is
    # what the chaotic late phase does to a perturbation of the size the split record rounds to
    for k in range(1, args.ensemble + 1):
        rng = np.random.default_rng(k)
        Pk = dict(P)
        for c in ("u", "v", "w"):
            a = np.asarray(P[c])
            Pk[c] = a * (1.0 + 2.0 ** -24 * rng.standard_normal(len(a)))
        t0 = time.time()
        series[f"fp64_ensemble_{k}"] = _series_gpu("fp64", ndim, nx, ny, nz, epc, Pk, steps, every)
        print(f"# fp64 ensemble member {k}: {steps} steps in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    if args.cpu_steps != 0:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        t0 = time.time()
        series["oracle"] = _series_oracle(ndim, nx, ny, nz, epc, P, steps, every, threads)
        print(f"# oracle: {steps} steps in {time.time() - t0:.1f}s on {threads} threads", file=sys.stderr, flush=True)
    ref_name = "oracle" if "oracle" in series else "fp64"
    ref = series[ref_name]
    window = _growth_window(ref)

    def fe(r):
        return r[1]["B_energy"] + r[1]["E_energy"]

    out = {"config": args.config, "steps": steps, "every": every, "reference": ref_name,
           "growth_window_steps": window, "forms": {}}
    for form, rows in series.items():
        if form == ref_name:
            continue
        err = np.array([abs(fe(a) - fe(b)) / fe(b) for a, b in zip(rows, ref) if fe(b) > 0])
        errb = np.array([abs(a[1]["B_energy"] - b[1]["B_energy"]) / b[1]["B_energy"]
                         for a, b in zip(rows, ref) if b[1]["B_energy"] > 0])
        lin = [i for i, r in enumerate(ref) if window and r[0] <= window[1]]
        out["forms"][form] = {
            "field_energy_rel_err_max": float(err.max()),
            "field_energy_rel_err_final": float(err[-1]),
            "field_energy_rel_err_max_linear_phase": float(err[lin].max()) if lin else None,
            "B_energy_rel_err_max": float(errb.max()),
            "growth_rate": _growth_rate(rows, window, ompe),
            # max over the run of max|div E - rho|, relative to the run's max|rho| (at t = 0 rho is
            # zero up to rounding: the two species start at the same positions)
            "gauss_residual_rel_max": float(max(r[1]["gauss_residual"] for r in rows)
                                            / max(max(r[1]["rho_max"] for r in rows), 1e-300)),
        }
    out["growth_rate_reference"] = _growth_rate(ref, window, ompe)
    out["gauss_residual_rel_max_reference"] = float(max(r[1]["gauss_residual"] for r in ref)
                                                    / max(max(r[1]["rho_max"] for r in ref), 1e-300))
    out["series"] = {f: [(t, d["B_energy"], d["E_energy"], d["gauss_residual"], d["rho_max"]) for t, d in rows]
                     for f, rows in series.items()}
    return out


def adjoint_report(args):
    """C5: gradient of the final magnetic energy of the 2-D Weibel run with
    respect to the drift speed and all initial momenta, FP32 split forward
    with checkpoints every --segment steps; checked against central finite
    differences of the FP64 engine (eps = 1e-4) and the FP64 adjoint."""
    import numpy as np
    import torch

    from pictor_amd import Simulation, setup_weibel
    from pictor_amd.adjoint import ParticleStepObjective, drift_boost, drift_signs

    nx = ny = args.adjoint_n
    epc, drift, uth, nsteps, seg = args.adjoint_epc, 0.5, 0.01, args.adjoint_steps, args.segment
    P0 = setup_weibel(nx, ny, epc=epc, drift=0.0, uth=uth, seed=12345)
    sgn = torch.from_numpy(drift_signs(len(P0["x"]), epc))
    th = [torch.from_numpy(P0[k]) for k in ("u", "v", "w")]
    out = {"workload": f"adjoint: 2-D Weibel {nx}x{ny}, {epc} ppc/species, {len(P0['x'])} particles, "
                       f"{nsteps} steps, checkpoints every {seg}", "forms": {}}
    # the autograd graph on GPU tensors: the momenta, particles and gradients stay in
    # device memory (pictor_adjoint_*_device)
    dev = "cuda"
    sgn_d, th_d = sgn.to(dev), [t.to(dev) for t in th]
    P0d = {k: torch.from_numpy(a).to(dev) for k, a in P0.items()}
    out["tensors"] = dev
    grads = {}
    for form in args.adjoint_forms.split(","):
        d = torch.tensor(drift, dtype=torch.float64, device=dev, requires_grad=True)
        tl = [t.clone().requires_grad_(True) for t in th_d]
        u, v, w = drift_boost(tl[0], tl[1], tl[2], sgn_d, d)
        for t in (u, v, w):
            t.retain_grad()
        with Simulation(nx, ny, epc=epc, form=form) as sim:
            sim.adjoint_engine(args.adjoint_engine)
            L = ParticleStepObjective.apply(u, v, w, sim, P0d, nsteps, seg)
            L.backward()
            st = sim.adjoint_stats()
            # plain forward of the same steps with the production kernels
            P = dict(P0)
            P["u"], P["v"], P["w"] = [t.detach().cpu().numpy() for t in (u, v, w)]
            sim.upload_particles(**P)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.step(nsteps)
            sim.synchronize()
            t_prod = (time.perf_counter() - t0) * 1e3
            L_prod = sim.diagnostics()["B_energy"]
        grads[form] = (float(d.grad), [t.grad.cpu().numpy().copy() for t in (u, v, w)], float(L))
        out["forms"][form] = {
            "objective": float(L), "objective_production_forward": L_prod, "d_drift": float(d.grad),
            "forward_ms": st["forward_ms"], "recompute_ms": st["recompute_ms"], "reverse_ms": st["reverse_ms"],
            "production_forward_ms": t_prod, "device_bytes": st["device_bytes"],
            "backward_over_forward": (st["recompute_ms"] + st["reverse_ms"]) / st["forward_ms"],
            "gradient_over_production_forward": (st["forward_ms"] + st["recompute_ms"] + st["reverse_ms"]) / t_prod,
        }
        print(f"# adjoint {form}: {out['forms'][form]}", file=sys.stderr, flush=True)
    # the FP64 adjoint from momenta perturbed at FP32's relative resolution: how far the
    # gradient itself moves when the trajectory changes by that much (the floor for the
    # split form, whose forward steps round to FP32)
    if "fp64" in grads and args.adjoint_ensemble > 0:
        out["fp64_ensemble"] = []
        g64 = grads["fp64"]
        for k in range(1, args.adjoint_ensemble + 1):
            gen = torch.Generator().manual_seed(k)
            d = torch.tensor(drift, dtype=torch.float64, device=dev, requires_grad=True)
            tl = [(t * (1.0 + 2.0 ** -24 * torch.randn(t.shape, generator=gen, dtype=torch.float64))).to(dev)
                  .requires_grad_(True) for t in th]
            u, v, w = drift_boost(tl[0], tl[1], tl[2], sgn_d, d)
            for t in (u, v, w):
                t.retain_grad()
            with Simulation(nx, ny, epc=epc, form="fp64") as sim:
                L = ParticleStepObjective.apply(u, v, w, sim, P0d, nsteps, seg)
                L.backward()
            gm = [t.grad.cpu().numpy() for t in (u, v, w)]
            num = np.sqrt(sum(np.sum((gm[c] - g64[1][c]) ** 2) for c in range(3)))
            den = np.sqrt(sum(np.sum(g64[1][c] ** 2) for c in range(3)))
            rec = {"member": k, "objective": float(L), "d_drift": float(d.grad),
                   "rel_diff_drift": abs(float(d.grad) - g64[0]) / abs(g64[0]), "rel_l2_momenta": float(num / den)}
            out["fp64_ensemble"].append(rec)
            print(f"# fp64 ensemble member {k}: {rec}", file=sys.stderr, flush=True)
    if "fp64" in grads and "fp32_split" in grads:
        a64, a32 = grads["fp64"], grads["fp32_split"]
        num = np.sqrt(sum(np.sum((a32[1][c] - a64[1][c]) ** 2) for c in range(3)))
        den = np.sqrt(sum(np.sum(a64[1][c] ** 2) for c in range(3)))
        out["rel_err_split_vs_fp64_drift"] = abs(a32[0] - a64[0]) / abs(a64[0])
        out["rel_l2_split_vs_fp64_momenta"] = float(num / den)
    if "fp64" not in grads or "fp32_split" not in grads or not args.adjoint_fd:
        return out
    # central finite differences of the FP64 objective in the drift speed; the zigzag
    # current jumps where a particle's crossing state changes, so a wide interval
    # measures those jumps as well (see tests/test_gpu_adjoint.py)
    g64, g32 = grads["fp64"], grads["fp32_split"]
    out["fd_drift_fp64"] = {}
    with Simulation(nx, ny, epc=epc, form="fp64") as sim:
        for eps in (1e-4, 1e-5, 1e-6, 1e-7, 1e-8):
            vals = []
            for dd in (drift + eps, drift - eps):
                u, v, w = drift_boost(th[0], th[1], th[2], sgn, torch.tensor(dd, dtype=torch.float64))
                vals.append(sim.adjoint_forward(P0["q"], P0["x"], P0["y"], P0["z"], u.numpy(), v.numpy(),
                                                w.numpy(), P0["flv"], nsteps, seg))
            fd = (vals[0] - vals[1]) / (2 * eps)
            out["fd_drift_fp64"][f"{eps:g}"] = {"fd": fd, "rel_err_fp64_adjoint": abs(g64[0] - fd) / abs(fd),
                                                "rel_err_split_adjoint": abs(g32[0] - fd) / abs(fd)}
            print(f"# fd eps={eps:g}: {out['fd_drift_fp64'][f'{eps:g}']}", file=sys.stderr, flush=True)
    # single initial momenta: a perturbation of one particle rarely moves a relay point across
    # a face, so these differences measure the derivative itself
    out["fd_single_fp64"] = []
    u0, v0, w0 = drift_boost(th[0], th[1], th[2], sgn, torch.tensor(drift, dtype=torch.float64))
    base = [u0.numpy().copy(), v0.numpy().copy(), w0.numpy().copy()]
    rng = np.random.default_rng(7)
    picks = []
    for comp in range(3):
        mag = np.abs(g64[1][comp])
        picks += [(comp, int(i)) for i in rng.choice(np.argsort(mag)[-1000:], 2, replace=False)]
    with Simulation(nx, ny, epc=epc, form="fp64") as sim:
        for comp, idx in picks:
            eps = 1e-6
            vals = []
            for sgn_e in (1.0, -1.0):
                mom = [b.copy() for b in base]
                mom[comp][idx] += sgn_e * eps
                vals.append(sim.adjoint_forward(P0["q"], P0["x"], P0["y"], P0["z"], mom[0], mom[1], mom[2],
                                                P0["flv"], nsteps, seg))
            fd = (vals[0] - vals[1]) / (2 * eps)
            adj = g64[1][comp][idx]
            rec = {"component": "uvw"[comp], "particle": idx, "eps": eps, "fd": fd, "adjoint_fp64": float(adj),
                   "adjoint_split": float(g32[1][comp][idx]), "rel_err_fp64": abs(adj - fd) / abs(fd)}
            out["fd_single_fp64"].append(rec)
            print(f"# fd single: {rec}", file=sys.stderr, flush=True)
    num = np.sqrt(sum(np.sum((g32[1][k] - g64[1][k]) ** 2) for k in range(3)))
    den = np.sqrt(sum(np.sum(g64[1][k] ** 2) for k in range(3)))
    out["rel_err_split_vs_fp64_drift"] = abs(g32[0] - g64[0]) / abs(g64[0])
    out["rel_l2_split_vs_fp64_momenta"] = float(num / den)
    return out


CYL_CONFIGS = {
    # name: (ndim, nr, ntheta, nz, particles per cell per species, c) -- particles uniform in index
    # space.  The cells next to the axis are 0.5 dtheta wide in theta, so the reference's scheme
    # needs c < 1/sqrt(1 + (ntheta/pi)^2 (+1)): 8 theta cells and c = 0.3.
    "cyl2d_4096": (2, 4096, 8, 1, 128, 0.3),
    "cyl3d_c4": (3, 512, 8, 4096, 4, 0.3),
}


def cyl_report(args):
    """The cylindrical geometry (the reference's -Dcyl build with the axis):
    FP64 steps of an electron-positron plasma filling the grid uniformly in
    index space (particles per cell constant in r), momenta Gaussian 0.1c.
    value = particles * steps / wall time; roofline of k_cyl_push with the
    FP64 record (108 B per particle); cpu_baseline: the cylindrical oracle
    (one thread) on a bounded sample."""
    import numpy as np

    from oracle.pyoracle import CylOracle
    from pictor_amd import Simulation

    ndim, nr, nth, nz, epc, c = CYL_CONFIGS[args.cyl]
    rng = np.random.default_rng(12345)
    ncell = nr * nth * (nz if ndim == 3 else 1)
    n = 2 * epc * ncell
    rmax = nr - 6.0
    half = n // 2
    x = 3.5 + rng.uniform(0.0, rmax, half)
    y = 3.0 + rng.uniform(0.0, nth, half)
    z = rng.uniform(1.0, 2.0, half) if ndim == 2 else rng.uniform(3.0, nz + 3.0, half)
    P = {"x": np.repeat(x, 2), "y": np.repeat(y, 2), "z": np.repeat(z, 2)}
    P["u"], P["v"], P["w"] = rng.normal(0.0, 0.1, (3, 2 * half))
    P["flv"] = np.tile(np.array([1, 2], dtype=np.int32), half)
    P["q"] = np.where(P["flv"] == 1, 1.0, -1.0)
    npart = len(P["x"])
    print(f"# cyl {args.cyl}: {npart} particles", file=sys.stderr, flush=True)
    with Simulation(nr, nth, nz, ndim=ndim, c=c, epc=epc, form="fp64", geometry="cyl") as sim:
        sim.upload_particles(P["q"], P["x"], P["y"], P["z"], P["u"], P["v"], P["w"], P["flv"])
        sim.step(args.warmup)
        sim.synchronize()
        sim.timer_reset()
        t0 = time.time()
        sim.step(args.steps)
        sim.synchronize()
        wall = time.time() - t0
        push_ms, npush = sim.timer("push")
    push_ms /= max(npush, 1)
    value = npart * args.steps / wall
    mx, my, mz = nr + 5, nth + 5, (nz + 5 if ndim == 3 else 1)
    algo = npart * bytes_per_particle("fp64", 3) + mx * my * mz * GRID_BYTES["fp64"]
    achieved = algo / (push_ms * 1e-3)
    cpu = None
    if args.cpu_steps != 0:
        # a bounded sample: the first particles of the same plasma on the same grid
        m = min(npart, 2_000_000)
        o = CylOracle(ndim, nr, nth, nz, c=c, epc=epc, cap=m + 16)
        o.load_particles(*(P[k][:m] for k in ("q", "x", "y", "z", "u", "v", "w", "flv")))
        o.step(0)
        steps = 10 if ndim == 2 else 4
        t0 = time.time()
        o.run(1, steps)
        dt = time.time() - t0
        cpu = {"value": m * steps / dt, "unit": "particle-updates/s", "cores": 1, "kind": "port",
               "sample": f"{steps} steps of the first {m} particles of the same plasma on the same grid "
                         f"(cylindrical FP64 oracle, one thread) after 1 warm-up step, {dt:.1f}s"}
    return {
        "metric": "particle-updates/s/GPU, cylindrical (r, theta, z) FP64 step",
        "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.cyl}: {ndim}-D cylindrical grid nr={nr} ntheta={nth}"
                               + (f" nz={nz}" if ndim == 3 else "") + f", {epc} ppc/species, {npart} particles, "
                               f"momenta N(0, 0.1) per component, c={c}, axis included, conducting outer wall",
                   "form": "fp64", "particles": npart},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "k_cyl_push", "push_ms": push_ms},
        "cpu_baseline": cpu,
    }


LWFA_CONFIGS = {
    # name: (nr, nz, dr in axial cells, c/omega_p in axial cells, radial plasma cells, particles per cell, c)
    # dz = 0.05 c/omega_p, dr = 0.2 c/omega_p; c dt = 0.8 dz sits below the m = 1 axis limit
    # 1/sqrt(1 + 5/dr^2) = 0.873
    "lwfa_4096x512": (512, 4096, 4.0, 20.0, 480, 16, 0.8),
    "lwfa_1024x128": (128, 1024, 4.0, 20.0, 112, 16, 0.8),
}


# algorithmic bytes per particle update of the modal push: the FP64 record (q, x, y, z,
# u, v, w, flavour read; x, y, z, u, v, w written) or the split record (cell word, 2 FP32
# offsets, 32-bit angle, 3 FP32 momenta, FP32 weight read; all but the weight written)
LWFA_BYTES_PER_PARTICLE = {"fp64": 7 * 8 + 4 + 6 * 8, "fp32_split": 8 * 4 + 7 * 4}


def _lwfa_run(args, form):
    from pictor_amd import Simulation

    nr, nz, dr, cps, nrp, ppc, c = LWFA_CONFIGS[args.lwfa]
    a0, omega0, tau, w0 = 2.0, (1.0 / 0.005) ** 0.5, 10.0, 15.0
    zc = nz - 3.0 * tau * cps            # the pulse centre three lengths behind the front
    with Simulation(nr, 2, nz, ndim=3, epc=ppc, c=c, c_ompe=cps, mi_me=1e30, form=form,
                    geometry="cyl_modal", dr=dr, fdtd=args.lwfa_fdtd) as sim:
        sim.cyl_plasma(nrp, 3, nz, ppc, seed=2024)
        sim.cyl_laser(a0, omega0, tau, w0, zc)
        sim.moving_window(True, nr_plasma=nrp, ppc=ppc, seed=2024)
        sim.cyl_sort_period(args.lwfa_sort)
        sim.step(args.warmup)
        sim.synchronize()
        sim.timer_reset()
        t0 = time.time()
        sim.step(args.steps)
        sim.synchronize()
        wall = time.time() - t0
        push_ms, npush = sim.timer("push")
        d = sim.diagnostics()
        mem = sim.memory_bytes()
    return wall, push_ms / max(npush, 1), d, mem


def lwfa_report(args):
    """Laser wakefield in the azimuthal-mode (m = 0, 1) cylinder with the Lehe
    B update (dispersion along z corrected): cold electrons at n/n_c = 0.005
    (omega_0 = 14.14 omega_p) over immobile ions, 16 particles per (r, z) cell,
    a linearly polarised Gaussian pulse a0 = 2, tau = 10/omega_p, w0 = 15
    c/omega_p, and the moving window (one axial cell per 1/c steps, new plasma
    injected at the front).  value = particles * steps / wall time of the timed
    steps of the split form (whole steps: fields, push + modal deposit, halos,
    window shifts and injection); the FP64 form runs beside it.  roofline:
    k_cylm_push_split (60 B per particle) or k_cylm_push (108 B) + 72 B per mode
    cell (E, B read as FP64, J written).  cpu_baseline: the modal oracle (one
    thread) on a 256-cell axial slab of the same plasma.  FDTD: Lehe (the
    default) or Yee."""
    from oracle.pyoracle import CylModalOracle

    nr, nz, dr, cps, nrp, ppc, c = LWFA_CONFIGS[args.lwfa]
    npart = nrp * nz * ppc
    ncell = (nr + 5) * 3 * (nz + 5)
    forms = [f for f in args.forms.split(",") if f]
    res = {}
    for form in forms:
        print(f"# lwfa {args.lwfa} {form}: {npart} particles", file=sys.stderr, flush=True)
        wall, push_ms, d, mem = _lwfa_run(args, form)
        algo = npart * LWFA_BYTES_PER_PARTICLE[form] + ncell * GRID_BYTES["fp64"]
        res[form] = {"value": npart * args.steps / wall, "ms_per_step": wall / args.steps * 1e3, "push_ms": push_ms,
                     "achieved": algo / (push_ms * 1e-3), "energies": {k: d[k] for k in ("E_energy", "B_energy")},
                     "device_bytes": mem}
        print(f"# {form}: {res[form]['ms_per_step']:.3f} ms/step, push {push_ms:.3f} ms", file=sys.stderr, flush=True)
    main_form = forms[0]
    r = res[main_form]
    cpu = None
    if args.cpu_steps != 0:
        nzs, steps = 256, 20
        cap = nrp * nzs * ppc + 16
        o = CylModalOracle(nr, nzs, dr=dr, c=c, c_ompe=cps, epc=ppc, cap=cap, fdtd=1 if args.lwfa_fdtd == "lehe" else 0)
        o.load_plasma(nrp, 3, nzs, ppc, seed=2024, cap=cap)
        o.step()
        t0 = time.time()
        o.run(steps)
        dt = time.time() - t0
        m = nrp * nzs * ppc
        cpu = {"value": m * steps / dt, "unit": "particle-updates/s", "cores": 1, "kind": "port",
               "sample": f"{steps} steps of the same plasma on a {nr} x {nzs} slab ({m} particles) "
                         f"with the modal oracle, one thread, after 1 warm-up step, {dt:.1f}s"}
    kernel = "k_cylm_push_split" if main_form == "fp32_split" else "k_cylm_push"
    traffic = None   # HBM bytes per launch from the committed counter runs (FETCH_SIZE x 2 + WRITE_SIZE)
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(kernel)
        if e and e.get("workload") == args.lwfa:
            traffic = e["bytes_per_launch"]
    except (OSError, ValueError):
        pass
    out = {
        "metric": "particle-updates/s/GPU, cylindrical laser wakefield (modes m = 0, 1)",
        "value": r["value"], "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if main_form == "fp32_split" else "f64", "data": "synthetic",
        "config": {"workload": f"{args.lwfa}: modal cylinder nz={nz} nr={nr}, dz=0.05, dr=0.2 c/omega_p, "
                               f"c dt = {c} dz, n/n_c=0.005, immobile ions, {ppc} ppc in {nrp} radial cells, "
                               f"a0=2.0, tau=10.0/omega_p, w0=15.0 c/omega_p, moving window, "
                               f"{'Lehe' if args.lwfa_fdtd == 'lehe' else 'Yee'} FDTD, particles sorted by cell "
                               f"every {args.lwfa_sort} steps",
                   "form": main_form, "particles": npart, "energies": r["energies"],
                   "device_bytes": r["device_bytes"]},
        "roofline": {"bound": "hbm", "achieved": r["achieved"] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": r["achieved"] / HBM_PEAK, "traffic": traffic, "kernel": kernel, "push_ms": r["push_ms"]},
        "cpu_baseline": cpu,
    }
    for form in forms[1:]:
        out[f"{form}_form"] = res[form]
        out[f"{main_form}_over_{form}"] = r["value"] / res[form]["value"]
    return out


def lwfa_accuracy_report(args):
    """The split form against the FP64 form on the laser wakefield (lwfa_1024x128,
    Lehe, no window, so that the Gauss law can be followed): field energies every
    `every` steps over `steps` steps, and per mode the change of the discrete Gauss
    residual (flux of E minus the node charge, the modal oracle's definition) from
    the start to the end, relative to its largest value."""
    import numpy as np

    from oracle.pyoracle import CylModalOracle
    from pictor_amd import Simulation

    name = "lwfa_1024x128"
    nr, nz, dr, cps, nrp, ppc, c = LWFA_CONFIGS[name]
    a0, omega0, tau, w0 = 2.0, (1.0 / 0.005) ** 0.5, 10.0, 15.0
    steps, every = args.steps, args.every
    series, gauss = {}, {}
    for form in ("fp64", "fp32_split"):
        with Simulation(nr, 2, nz, ndim=3, epc=ppc, c=c, c_ompe=cps, mi_me=1e30, form=form, geometry="cyl_modal",
                        dr=dr, fdtd=args.lwfa_fdtd) as sim:
            sim.cyl_plasma(nrp, 3, nz, ppc, seed=2024)
            P0 = sim.download_particles()
            sim.cyl_laser(a0, omega0, tau, w0, nz / 2.0)
            F0 = sim.download_fields()
            rows = []
            for t in range(0, steps + 1, every):
                if t:
                    sim.step(every)
                d = sim.diagnostics()
                rows.append((t, d["E_energy"], d["B_energy"], d["prtl_energy_2"]))
                print(f"# {form} step {t}: E {d['E_energy']:.9e} B {d['B_energy']:.9e}", file=sys.stderr, flush=True)
            F1 = sim.download_fields()
            P1 = sim.download_particles()
        series[form] = rows
        cap = len(P0["x"]) + 16
        o = CylModalOracle(nr, nz, dr=dr, c=c, c_ompe=cps, epc=ppc, cap=cap)
        res = []
        for F, P in ((F0, P0), (F1, P1)):
            o.set_fields(**{k: F[k] for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")})
            o.load_particles(P["q"], P["x"], P["y"], P["z"], P["u"], P["v"], P["w"], P["flv"], cap=cap)
            res.append(o.gauss())
        gauss[form] = [float(np.abs(res[1] - res[0])[:, s, :].max() / np.abs(res[0][:, s, :]).max()) for s in range(3)]
    a, b = series["fp64"], series["fp32_split"]
    rel = {k: max(abs(y[i] - x[i]) / abs(x[i]) for x, y in zip(a, b) if x[i] != 0) for k, i in
           (("E_energy", 1), ("B_energy", 2), ("prtl_energy_2", 3))}
    return {"report": "lwfa accuracy: split vs FP64", "config": name, "steps": steps, "every": every,
            "fdtd": args.lwfa_fdtd, "max_rel_diff": rel, "gauss_change_per_mode": gauss,
            "series": {f: [list(r) for r in series[f]] for f in series}}


def atomics_report(iters=4096):
    """Atomic throughput behind the accumulator choices: the split push's current
    window (LDS FP64), the adjoint's field-adjoint window (LDS FP64, FP32 only for
    the 32 x 32 tile), the tile flush (global FP64)."""
    import ctypes

    from pictor_amd._lib import check, lib

    out = (ctypes.c_double * 16)()
    check(lib().pictor_probe_atomics(iters, out))
    names = ("lds_f32_add", "lds_f64_add", "lds_u32_add", "lds_u64_add", "global_f64_add", "global_f32_add")
    return {"workload": f"atomic adds at pseudo-random addresses, 4 workgroups x 256 lanes per CU, {iters} each; "
                        "LDS window 32 KiB, global array 4 Mi elements",
            "cus": int(out[6]), "clock_hz": out[7],
            "lane_ops_per_s": {k: out[i] for i, k in enumerate(names)},
            "lane_ops_per_cu_clock": {k: out[8 + i] for i, k in enumerate(names)}}


def stream_report(n=268435456):
    """The stream ceiling of the 2-D split record's access pattern: six FP32
    arrays of n elements read and six written (48 B per element), with one,
    two (lane-contiguous) or four elements per lane (pictor_probe_stream)."""
    import ctypes

    from pictor_amd._lib import check, lib

    out = (ctypes.c_double * 8)()
    check(lib().pictor_probe_stream(n, out))
    names = ("one_4B_element_per_lane", "two_4B_elements_per_lane", "float2_per_lane", "float4_per_lane")
    return {"workload": f"6 float arrays of {n} elements read, 6 written (48 B per element)",
            "bytes_per_s": {k: out[i] for i, k in enumerate(names)},
            "ms": {k: out[4 + i] for i, k in enumerate(names)},
            "frac_of_8TBps": {k: out[i] / 8e12 for i, k in enumerate(names)}}


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    return {"nproc": os.cpu_count() or 1, "affinity": aff, "cpu_model": model, "omp_threads": omp}


def c1_full_cpu(threads):
    """The oracle on the 256^2 CPU workload in full: 16 ppc/species, 1000 FP64 steps."""
    from pictor_amd import setup_weibel

    ndim, nx, ny, nz, epc, drift, uth = CONFIGS["weibel2d_256"]
    P = setup_weibel(nx, ny, nz, ndim=ndim, epc=epc, drift=drift, uth=uth, seed=12345)
    from oracle.pyoracle import Oracle

    o = Oracle(ndim, nx, ny, nz, epc=epc, c=0.45, c_ompe=10.0, nthreads=threads, cap=int(len(P["x"]) * 1.05) + 4096)
    o.load_particles(**P)
    steps = 1000
    t0 = time.perf_counter()
    o.run(1, steps)
    dt = time.perf_counter() - t0
    return {"workload": "weibel2d_256: 256x256, 16 ppc/species, 2 species, 1000 steps, FP64",
            "particles": len(P["x"]), "steps": steps, "seconds": dt, "value": len(P["x"]) * steps / dt,
            "unit": "particle-updates/s", "threads": threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="weibel2d_2048", choices=sorted(CONFIGS))
    ap.add_argument("--forms", default="fp32_split,fp64")
    ap.add_argument("--cpu-steps", type=int, default=-1, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0 = OMP_NUM_THREADS, the host cores given to one GPU, else all cores)")
    ap.add_argument("--c1-full", type=int, default=1, help="also time the oracle on the 256^2 workload in full")
    ap.add_argument("--late", default="1000,50,2000",
                    help="START,N,TOTAL: for the first form also time steps START+1..START+N and the mean over "
                         "TOTAL steps (empty = skip)")
    ap.add_argument("--calibrate", action="store_true", help="also run plain copy kernels (counter scaling)")
    ap.add_argument("--tile", type=int, default=0, help="tile edge in cells (0 = library default)")
    ap.add_argument("--curr-filter", type=int, default=0, help="passes of the 1-2-1 current filter per step")
    ap.add_argument("--accuracy", action="store_true",
                    help="instead of timing: run --steps steps of both forms (and the oracle unless "
                         "--cpu-steps 0) and report field energy, growth rate and Gauss residual errors")
    ap.add_argument("--every", type=int, default=10, help="diagnostic interval of --accuracy")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="--accuracy: FP64 members from momenta perturbed by 2^-24 relative noise")
    ap.add_argument("--adjoint", action="store_true",
                    help="instead of timing the step: the adjoint workload (gradients, FD check, cost ratios) "
                         "over --steps steps with checkpoints every --segment")
    ap.add_argument("--segment", type=int, default=20)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-n", type=int, default=512)
    ap.add_argument("--adjoint-epc", type=int, default=16)
    ap.add_argument("--adjoint-engine", default="sorted", choices=("sorted", "unsorted"))
    ap.add_argument("--adjoint-forms", default="fp32_split,fp64")
    ap.add_argument("--adjoint-fd", type=int, default=1, help="finite-difference check of the drift gradient")
    ap.add_argument("--adjoint-ensemble", type=int, default=0,
                    help="--adjoint: FP64 gradients from momenta perturbed by 2^-24 relative noise")
    ap.add_argument("--cyl", default=None, choices=sorted(CYL_CONFIGS),
                    help="instead of the Weibel step: time the cylindrical geometry on this grid")
    ap.add_argument("--lwfa", default=None, choices=sorted(LWFA_CONFIGS),
                    help="instead of the Weibel step: time the laser wakefield in the modal cylinder")
    ap.add_argument("--lwfa-fdtd", default="lehe", choices=("lehe", "yee"))
    ap.add_argument("--lwfa-sort", type=int, default=20, help="steps between sorts of the modal particles by cell")
    ap.add_argument("--lwfa-accuracy", action="store_true",
                    help="split vs FP64 on lwfa_1024x128 (no window): energies every --every steps, Gauss change")
    ap.add_argument("--atomics", action="store_true",
                    help="instead of the Weibel step: LDS versus global atomic throughput (pictor_probe_atomics)")
    ap.add_argument("--stream", action="store_true",
                    help="instead of the Weibel step: the stream ceiling of the particle arrays (pictor_probe_stream)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not code)", file=sys.stderr)
        sys.exit(2)

    if args.adjoint:
        print(json.dumps(adjoint_report(args)), flush=True)
        return
    if args.cyl:
        print(json.dumps(cyl_report(args)), flush=True)
        return
    if args.atomics:
        print(json.dumps(atomics_report()), flush=True)
        return
    if args.stream:
        print(json.dumps(stream_report()), flush=True)
        return
    if args.lwfa_accuracy:
        print(json.dumps(lwfa_accuracy_report(args)), flush=True)
        return
    if args.lwfa:
        print(json.dumps(lwfa_report(args)), flush=True)
        return

    import numpy as np  # noqa: F401
    from pictor_amd import setup_weibel

    ndim, nx, ny, nz, epc, drift, uth = CONFIGS[args.config]
    t0 = time.time()
    P = setup_weibel(nx, ny, nz, ndim=ndim, epc=epc, drift=drift, uth=uth, seed=12345)
    print(f"# setup {len(P['x'])} particles in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    if args.accuracy:
        print(json.dumps(accuracy_report(args, ndim, nx, ny, nz, epc, P)), flush=True)
        return

    forms = [f for f in args.forms.split(",") if f]
    late = tuple(int(v) for v in args.late.split(",")) if args.late else None
    res = {}
    for i, form in enumerate(forms):
        res[form] = run_gpu(form, ndim, nx, ny, nz, epc, P, args.steps, args.warmup, args.calibrate, args.tile,
                            args.curr_filter, late=late if i == 0 else None)
        r = res[form]
        r["roofline"]["traffic"] = measured_traffic(form, args.config, r["tile"], ndim)
        print(f"# {form}: {r['value']:.4e} particle-updates/s, {r['ms_per_step']:.3f} ms/step, "
              f"push {r['push_ms']:.3f} ms ({r['roofline']['frac'] * 100:.1f}% HBM)", file=sys.stderr, flush=True)
        if "late" in r:
            print(f"# {form} late: {json.dumps(r['late'])}", file=sys.stderr, flush=True)

    cpu = None
    big = len(P["x"]) > 100_000_000
    cpu_steps = args.cpu_steps if args.cpu_steps >= 0 else (2 if ndim == 3 and big else (5 if big else 20))
    if cpu_steps > 0:
        host = host_info()
        threads = args.cpu_threads or host["omp_threads"] or host["nproc"]
        threads = max(1, min(threads, host["nproc"]))
        v, dt = run_cpu(ndim, nx, ny, nz, epc, P, cpu_steps, threads)
        cpu = {"value": v, "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "sample": f"{cpu_steps} full steps of the same workload ({len(P['x'])} particles, FP64 oracle, "
                         f"OpenMP) after 1 warm-up step, {dt:.1f}s",
               "nproc": host["nproc"], "affinity_cores": host["affinity"], "cpu_model": host["cpu_model"],
               "threads_note": "OpenMP threads = OMP_NUM_THREADS, the host cores given to one GPU"
                               if host["omp_threads"] and not args.cpu_threads else "--cpu-threads"}
        print(f"# cpu: {v:.4e} particle-updates/s on {threads} threads ({dt:.1f}s)", file=sys.stderr, flush=True)
        if args.c1_full:
            c1 = c1_full_cpu(threads)
            cpu["c1_full"] = c1
            print(f"# cpu C1 in full: {json.dumps(c1)}", file=sys.stderr, flush=True)

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
                   "form": head, "tile": r["tile"], "particles": r["particles"], "curr_filter": args.curr_filter},
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
    }
    for form in forms[1:]:
        o = res[form]
        line[f"{form}_form"] = {"value": o["value"], "ms_per_step": o["ms_per_step"], "roofline": o["roofline"],
                                "dtype": "f64" if form == "fp64" else "f32"}
    if "fp64" in res and head != "fp64":
        line["split_over_fp64"] = r["value"] / res["fp64"]["value"]
    if "late" in r:
        line["late"] = r["late"]
    line["detail"] = {f: {k: res[f][k] for k in ("push_ms", "field_ms", "event_step_ms", "upload_s", "device_bytes",
                                                  "diag")} for f in forms}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Coupled-step throughput of the LeachingSim hot path on one MI355X.

Workload (default): 512^3 uint16 image with pore + three mineral phases
(Rock 0.30, Sulphide 0.25, Precipitate 0.15, each from its own seeded
smoothed field, sigma 3), a 4-voxel planar fracture at a seeded orientation
through the centre, 2-plane Air slabs on both j faces (surface precipitation
active), constant-pressure inlet/outlet planes on the i (CLI z) faces,
dt = 0.5 s, pressure converged to 1e-8 relative residual.

A step is one full coupled step (snapshot, properties, pressure, flux,
saturation, surface effects, conc + acid transport, precipitation,
leaching) on device-resident state.  The timed region holds `--steps` steps
after `--warmup` untimed steps, bracketed by device synchronisation.

Runs (each on a fresh context, same workload and window):
  headline  learned+mixed: flexible CG with the U-Net coarse-level
            preconditioner, raced against the recorded curve of
            mixed-precision PCG (FP32 CG on the red-black Schur complement
            inside FP64 reliable updates), which takes over when the network
            falls behind; warm start = projection onto earlier solutions
            plus the network direction (trained weights from
            leachingsim_amd/weights/);
  baseline  plain FP64 Jacobi-PCG warm-started from the previous pressure
            (the reference's algorithm, src/Simulation.cpp:131-159, 808);
  ablations mixed-precision PCG + projected warm start (no network), and
            plain FP64 PCG + projected warm start.

Prints one JSON line: steps/s of the headline, the baseline and ablations,
the roofline of the dominant pressure kernel measured with HIP events
during the timed region (HBM traffic from the committed rocprofv3 counter
summary), the plain FP32 / FP64 stencil y = A x timed alone on the same
operator, and the CPU oracle baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

from leachingsim_amd import geometry as G  # noqa: E402
from leachingsim_amd import lsgpu as L  # noqa: E402

HBM_PEAK_GBS = 8000.0

SOLVER_NAMES = {L.SOLVER_PLAIN64: "plain_fp64_pcg", L.SOLVER_MIXED: "mixed_fp32_fp64ir",
                L.SOLVER_LEARNED: "learned_fcg_mixed"}
WARM_NAMES = {L.WARM_PREVIOUS: "previous", L.WARM_PROJECTED: "projected",
              L.WARM_LEARNED: "projected+network"}
TRAFFIC_FILE = ROOT / "profiles" / "pressure_stencil_traffic.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def workload(name: str, size: int | None):
    """(labels, label bytes, dt, description) of a named workload."""
    if name == "fractured512":
        n = size or 512
        cache = os.environ.get("LS_GEOM_CACHE")
        path = Path(cache) / f"fractured_{n}.npy" if cache else None
        if path is not None and path.exists():
            img = np.load(path)
        else:
            img = G.named_geometry(name, size=n)
            if path is not None:
                path.parent.mkdir(parents=True, exist_ok=True)
                np.save(path, img)
        return img, 2, 0.5, (f"{n}^3 uint16, pore + 3 minerals + 4-voxel fracture, Air j-slabs, "
                             "z-face inlet/outlet, dt 0.5")
    if name == "smooth338":
        return (G.named_geometry(name), 1, 1.0,
                "338x338x283 uint8 sigma 3 porosity 0.30, z-face inlet/outlet, dt 1")
    if name == "smooth64":
        return (G.named_geometry(name), 1, 1.0,
                "64^3 uint8 sigma 2 porosity 0.40, z-face inlet/outlet, dt 1")
    raise SystemExit(f"unknown workload {name}")


COMPARE_FIELDS = (("pressure", L.F_PRESSURE), ("conc", L.F_CONC), ("acid", L.F_ACID),
                  ("saturation", L.F_SATURATION), ("qx", L.F_QX), ("qy", L.F_QY),
                  ("qz", L.F_QZ), ("inventory", L.F_INVENTORY))


def field_errors(a, b):
    """Per-field relative L2 difference of run a against the FP64 run b, and
    whether the phase masks (labels) and the leaching fraction are equal."""
    out = {}
    for name, _ in COMPARE_FIELDS:
        x, y = a[name], b[name]
        nb = float(np.linalg.norm(y))
        out[name] = float(np.linalg.norm(x - y)) / nb if nb > 0 else float(np.linalg.norm(x))
    out["labels_bit_exact"] = bool(np.array_equal(a["img"], b["img"]))
    out["frac_bit_exact"] = bool(np.array_equal(a["frac"], b["frac"]))
    return out


def run_solver(img, label_bytes, dt, solver, steps, warmup, profile=False, warm=L.WARM_PREVIOUS,
               weights=None, keep_fields=False):
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=solver, warm_start=warm)
    sim = L.GpuSimulation(img.shape, P, label_bytes)
    sim.read_labels(img)
    sim.init_state()
    if weights is not None:
        sim.set_network(weights)
    name = f"{SOLVER_NAMES[solver]}/{WARM_NAMES[warm]}"
    t = 0.0
    reps = []
    for n in range(warmup):
        r = sim.step(t, dt)
        if r.status != 0:
            raise RuntimeError(f"warmup step {n} failed")
        t += dt
        log(f"  [{name}] warmup {n}: {r.t_total_ms:.1f} ms, press iters {r.press_iters}")
    sim.set_profiling(profile)
    device_sync()
    t0 = time.perf_counter()
    for n in range(steps):
        r = sim.step(t, dt)
        if r.status != 0:
            raise RuntimeError(f"timed step {n} failed: {r.as_dict()}")
        reps.append(r.as_dict())
        t += dt
    device_sync()
    t1 = time.perf_counter()
    for n, r in enumerate(reps):
        log(f"  [{name}] step {warmup + n}: {r['t_total_ms']:.1f} ms "
            f"(pressure {r['t_pressure_ms']:.1f}, transport {r['t_transport_ms']:.1f}), "
            f"press iters {r['press_iters']} outer {r['press_outer']} "
            f"fallback {r['press_fallback']} unet {r['press_unet_applies']} "
            f"warm relres {r['press_relres0']:.2e} basis {r['warm_basis']}+{r['warm_net']} "
            f"relres {r['press_relres']:.2e}, adr iters {r['adr_iters']}")
    copy_gbs = None
    if profile:
        nb = min(img.size * 8, 1 << 30)
        ms = sim.time_copy(nb, 20)
        copy_gbs = 2 * nb / ms / 1e6
    fields = None
    if keep_fields:
        fields = {name: sim.download(f) for name, f in COMPARE_FIELDS}
        fields["img"] = sim.download(L.F_IMG)
        fields["frac"] = sim.download(L.F_FRAC)
    sim.close()
    return (t1 - t0), reps, copy_gbs, fields


def cpu_baseline(img, label_bytes, dt, crop: int, nsteps: int):
    """The FP64 oracle (C++/OpenMP) on a bounded sample: a centred crop of
    the same geometry, `nsteps` coupled steps, scaled to the full grid by
    voxel count (the CPU's PCG iteration count grows with the grid, so the
    scaled figure flatters the CPU)."""
    from oracle.oracle import OracleSimulation, set_threads
    ncpu = len(os.sched_getaffinity(0))
    set_threads(ncpu)
    c = [max(0, (s - crop) // 2) for s in img.shape]
    sub = np.ascontiguousarray(img[c[0]:c[0] + crop, c[1]:c[1] + crop, c[2]:c[2] + crop])
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64)
    o = OracleSimulation(sub.shape, P, label_bytes)
    o.read_labels(sub)
    o.init_state()
    t = 0.0
    times = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        r = o.step(t, dt)
        times.append(time.perf_counter() - t0)
        t += dt
        if r.status != 0:
            break
    sec = float(np.mean(times))
    scale = sub.size / img.size
    return {"value": scale / sec, "unit": "steps/s", "cores": ncpu, "kind": "port",
            "sample": f"oracle FP64 Jacobi-PCG, {nsteps} coupled steps on the centred "
                      f"{'x'.join(map(str, sub.shape))} crop ({sec:.2f} s/step), scaled "
                      f"by voxel count to {'x'.join(map(str, img.shape))}"}


KERNELS = {"stencil64": (64, "k_stencil_vec<double>"), "stencil32": (32, "k_stencil_vec<float>"),
           "rb_gather": (1, "k_rb_vec<1>"), "rb_apply": (2, "k_rb_vec<2>"),
           "schur": (3, "k_schur_apply<16, 4>")}


def stencil_rooflines(img, label_bytes, reps=20, rounds=5):
    """The plain 7-point stencil y = A x alone, FP32 and FP64 (20 / 40 B per
    voxel), on the pressure operator of the initial state: HIP-event time
    per launch (the median over `rounds` rounds of `reps` back-to-back
    launches, each round's average listed), achieved GB/s against the 8 TB/s
    peak, and the HBM traffic of the committed counter runs."""
    P = L.default_params(bc_mode=L.BC_ZFACES, press_maxit=1)
    sim = L.GpuSimulation(img.shape, P, label_bytes)
    sim.read_labels(img)
    sim.init_state()
    sim.phase(L.PH_PROPERTIES)
    sim.phase(L.PH_PRESSURE, 1.0)      # assembles; the 1-iteration solve is irrelevant
    tr = load_traffic() or {}
    out = {}
    for kernel in ("stencil32", "stencil64"):
        precision, kname = KERNELS[kernel]
        per_round = []
        for _ in range(rounds):
            ms, by = sim.time_stencil(precision, reps)
            per_round.append(ms)
        ms = float(np.median(per_round))
        achieved = by / ms / 1e6
        t = tr.get(kernel, {})
        out[kname] = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": by, "avg_launch_ms": ms,
                      "launches": reps, "rounds_ms": per_round,
                      "traffic": t.get("bytes_per_launch") if t.get("grid") == list(img.shape)
                      else None}
    sim.close()
    return out


def stencil_only(img, label_bytes, kernel, reps):
    """Assemble the pressure operator of the initial state, then time `reps`
    launches of one pressure kernel and `reps` copies of 1 GiB (HIP events)."""
    precision, kname = KERNELS[kernel]
    P = L.default_params(bc_mode=L.BC_ZFACES, press_maxit=1)
    sim = L.GpuSimulation(img.shape, P, label_bytes)
    sim.read_labels(img)
    sim.init_state()
    sim.phase(L.PH_PROPERTIES)
    sim.phase(L.PH_PRESSURE, 1.0)      # assembles; the 1-iteration solve is irrelevant
    ms, by = sim.time_stencil(precision, reps)
    nb = min(img.size * 8, 1 << 30)
    cms = sim.time_copy(nb, reps)
    sim.close()
    print(json.dumps({"kernel": kname,
                      "grid": list(img.shape), "reps": reps, "avg_launch_ms": ms,
                      "bytes_per_launch": by, "achieved_GBs": by / ms / 1e6,
                      "copy_bytes": 2 * nb, "copy_ms": cms, "copy_GBs": 2 * nb / cms / 1e6}),
          flush=True)


def unet_only(size, reps):
    """Time the U-Net forward on the grid the learned preconditioner uses for a
    size^3 pressure grid (halved, padded to multiples of 16), with the tiled
    convolution kernels and with the general kernel everywhere (HIP events,
    trained weights, seeded input)."""
    from leachingsim_amd import unet_spec as S
    from leachingsim_amd.learn import WEIGHTS, load_weights
    from leachingsim_amd.unet import Net
    n = -(-((size + 1) // 2) // 16) * 16
    net = Net((n, n, n), batch=1, trainable=False, weights=load_weights(WEIGHTS))
    x = np.random.default_rng(0).standard_normal((1, n, n, n, S.CIN)).astype(np.float32)
    net.forward(x)
    del x
    flop = sum(2.0 * 27 * L["cin"] * L["cout"] * n ** 3 / 8 ** L["level"] for L in S.LAYERS)
    out = {"kernel": "unet_forward", "grid": [n, n, n], "reps": reps, "gflop": flop / 1e9}
    for tiled in (True, False):
        net.set_conv(tiled)
        ms = net.time_forward(reps)
        key = "tiled" if tiled else "general"
        out[f"ms_{key}"] = ms
        out[f"tflops_{key}"] = flop / ms / 1e9
    net.close()
    print(json.dumps(out), flush=True)


def summary(sec, reps, steps):
    return {"steps_per_s": steps / sec,
            "press_iters_per_step": [r["press_iters"] for r in reps],
            "warm_relres": [r["press_relres0"] for r in reps],
            "final_relres_max": max(r["press_relres"] for r in reps),
            "fallbacks": sum(r["press_fallback"] for r in reps),
            "unet_applies": sum(r["press_unet_applies"] for r in reps)}


def load_traffic():
    """HBM bytes per stencil launch from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs summarised in profiles/ (None if absent)."""
    try:
        return json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fractured512", choices=list(G.WORKLOADS))
    ap.add_argument("--size", type=int, default=None,
                    help="edge length of the fractured workload (default 512)")
    ap.add_argument("--solver", default="learned", choices=["plain", "mixed", "learned"])
    ap.add_argument("--warm", default=None, choices=["previous", "projected", "learned"],
                    help="warm start of the headline (default: learned for the learned "
                         "solver, projected otherwise)")
    ap.add_argument("--weights", default=None, help="U-Net weights (float32 file)")
    ap.add_argument("--no-baseline", action="store_true", help="skip the plain FP64 run")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip the mixed+projected and plain+projected runs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-crop", type=int, default=96)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--mode", default="steps", choices=["steps", "stencil", "unet"],
                    help="stencil: assemble the pressure operator once and time only the "
                         "stencil and a copy kernel (for counter collection); unet: time "
                         "the U-Net forward on the preconditioner's grid")
    ap.add_argument("--kernel", default="schur", choices=list(KERNELS),
                    help="kernel timed in --mode stencil")
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the i-slab decomposition across GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)

    if args.mode == "unet":
        unet_only(args.size or 512, args.reps)
        return
    t0 = time.perf_counter()
    img, lb, dt, desc = workload(args.workload, args.size)
    log(f"workload {desc}: generated in {time.perf_counter() - t0:.1f} s; labels "
        + ", ".join(f"{k}:{int((img == k).sum())}" for k in range(1, 6)))
    if args.mode == "stencil":
        stencil_only(img, lb, args.kernel, args.reps)
        return
    head = {"plain": L.SOLVER_PLAIN64, "mixed": L.SOLVER_MIXED,
            "learned": L.SOLVER_LEARNED}[args.solver]
    warm = {"previous": L.WARM_PREVIOUS, "projected": L.WARM_PROJECTED,
            "learned": L.WARM_LEARNED, None: None}[args.warm]
    if warm is None:
        warm = L.WARM_LEARNED if head == L.SOLVER_LEARNED else L.WARM_PROJECTED
    weights = None
    if head == L.SOLVER_LEARNED or warm == L.WARM_LEARNED:
        from leachingsim_amd.learn import WEIGHTS, load_weights
        weights = load_weights(args.weights or WEIGHTS)

    keep = not args.no_baseline
    sec, reps, copy_gbs, hfields = run_solver(img, lb, dt, head, args.steps, args.warmup,
                                              profile=True, warm=warm, weights=weights,
                                              keep_fields=keep)
    value = args.steps / sec
    base = None
    errors = None
    if not args.no_baseline:
        bsec, breps, _, bfields = run_solver(img, lb, dt, L.SOLVER_PLAIN64, args.steps,
                                             args.warmup, keep_fields=True)
        base = summary(bsec, breps, args.steps)
        errors = field_errors(hfields, bfields)
        del hfields, bfields
    ablations = {}
    if not args.no_ablation:
        for sv in (L.SOLVER_MIXED, L.SOLVER_PLAIN64):
            if (sv, L.WARM_PROJECTED) == (head, warm):
                continue
            asec, areps, _, _ = run_solver(img, lb, dt, sv, args.steps, args.warmup,
                                           warm=L.WARM_PROJECTED)
            ablations[f"{SOLVER_NAMES[sv]}/projected"] = summary(asec, areps, args.steps)

    # roofline of the sampled pressure kernel (HIP events in the timed
    # region): the fused Schur apply on compacted vectors when the mixed
    # solver runs on the Schur complement (3 faces + half-length p and q:
    # 16 B per voxel), its gather pass in the two-pass form, else the FP32
    # or FP64 stencil (x, 3 faces, y: 20 B / 40 B per voxel)
    if sum(r["prof_schur_n"] for r in reps) > 0:
        key = "prof_schur"
        bytes_per_launch = reps[-1]["prof_schur_bytes"]
        kernel = "schur" if bytes_per_launch == img.size * 16 else "rb_gather"
    elif head == L.SOLVER_PLAIN64 or sum(r["prof_stencil32_n"] for r in reps) == 0:
        key, kernel, bytes_per_launch = "prof_stencil64", "stencil64", img.size * 40
    else:
        key, kernel, bytes_per_launch = "prof_stencil32", "stencil32", img.size * 20
    n = sum(r[key + "_n"] for r in reps)
    ms = sum(r[key + "_ms"] for r in reps) / max(n, 1)
    kname = KERNELS[kernel][1]
    # the vectorised stencils need nk % 4 == 0 (FP32) / nk % 2 == 0 (FP64);
    # otherwise the solver ran the scalar stencil
    if kernel == "stencil32" and img.shape[2] % 4 != 0:
        kname = "k_stencil<float, 0>"
    if kernel == "stencil64" and img.shape[2] % 2 != 0:
        kname = "k_stencil<double, 0>"
    if kernel == "rb_gather" and img.shape[2] % 4 != 0:
        kname = "k_rb_scalar<1>"
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9 if n else 0.0
    traffic = None
    tr = load_traffic()
    if tr is not None and kernel in tr and tr[kernel].get("grid") == list(img.shape):
        traffic = tr[kernel]["bytes_per_launch"]

    stencils = stencil_rooflines(img, lb)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(img, lb, dt, min(args.cpu_crop, *img.shape), args.cpu_steps)

    out = {
        "metric": "steps/s at 512^3, 1e-8 rel residual: learned+mixed vs plain FP64 CG",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == L.SOLVER_PLAIN64 else "f32+f64+bf16",
        "data": "synthetic",
        "config": {"workload": desc, "grid": list(img.shape), "solver": SOLVER_NAMES[head],
                   "warm_start": WARM_NAMES[warm], "pressure_rtol": 1e-8,
                   "headline": summary(sec, reps, args.steps),
                   "plain_fp64_pcg_previous": base,
                   "speedup_vs_plain_fp64": (value / base["steps_per_s"]) if base else None,
                   "error_vs_plain_fp64_after_window": errors,
                   "ablations": ablations},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "kernel": kname,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": ms,
                     "sampled_launches": n, "measured_copy_GBs": copy_gbs, "traffic": traffic},
        "stencil_roofline": stencils,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

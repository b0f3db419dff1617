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

Each run is a fresh context making the workload's first 100 steps (the
timed window lies inside them, so the same steps are compared); the cold
solves of steps 0 and 1 are reported apart, with the whole 100-step time.
  headline  mixed precision: FP32 CG on the red-black Schur complement inside
            FP64 reliable updates, warm start = projection onto earlier
            solutions (the fastest configuration measured);
  baseline  plain FP64 Jacobi-PCG with the same warm start (like for like),
            and plain FP64 Jacobi-PCG from the previous pressure, the
            reference's algorithm (src/Simulation.cpp:131-159, 808), timed
            on its first steps (each is a full solve);
  ablation  the learned solver: flexible CG with the U-Net coarse-level
            preconditioner raced against mixed PCG, warm start with the
            network direction (trained weights from leachingsim_amd/weights/).

Prints one JSON line: steps/s of the headline over the window, every run's
window and 100-step figures, the speedups over plain FP64, the roofline of
the kernel with the largest share of the window (the transport's red Schur
pass) and of the pressure kernel of the cold solves, both timed live with
HIP events, the plain FP32 / FP64 stencil y = A x timed alone on the same
operator, and the CPU oracle baseline over the same window.
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
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_fractured512.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


_hip = None


def _hip_rt():
    global _hip
    import ctypes
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7")
    return _hip


def device_sync():
    """hipDeviceSynchronize through the HIP runtime the library links."""
    if _hip_rt().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def device_free_bytes():
    """(free, total) device memory (hipMemGetInfo)."""
    import ctypes
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    if _hip_rt().hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return f.value, t.value


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
    if name == "fractured1024x768":
        n = size or 1024
        img = G.named_geometry(name, size=n)
        return img, 1, 0.5, (f"{n}x{n}x{3 * n // 4} (x, y, z) uint8 sigma 3 porosity 0.25, two "
                             "orthogonal 6-voxel fractures, Air j-slabs, z-face inlet/outlet, "
                             "dt 0.5")
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


class Run:
    """One configuration on a fresh context: steps 0 .. total-1 of the
    workload; the timed window is steps warmup .. warmup+steps-1 (device
    synchronised on both sides); every step's wall time and report kept."""

    def __init__(self, solver, warm, **params):
        self.solver, self.warm, self.params = solver, warm, params
        self.name = f"{SOLVER_NAMES[solver]}/{WARM_NAMES[warm]}"

    def run(self, img, label_bytes, dt, warmup, steps, total, profile=False, weights=None,
            keep_fields=False):
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=self.solver, warm_start=self.warm,
                             **self.params)
        device_sync()
        free0, self.device_total = device_free_bytes()
        sim = L.GpuSimulation(img.shape, P, label_bytes)
        sim.read_labels(img)
        sim.init_state()
        if weights is not None:
            sim.set_network(weights)
        self.reports, self.wall = [], []
        self.fields = None
        self.window_s = None
        t = 0.0
        total = max(total, warmup + steps)
        # sampled kernel timing (one event pair per solver batch) for the
        # whole run: the cold solves' pressure kernels and the window's
        sim.set_profiling(profile)
        for n in range(total):
            if n == warmup:
                device_sync()
                w0 = time.perf_counter()
            t0 = time.perf_counter()
            r = sim.step(t, dt)
            t1 = time.perf_counter()
            if r.status != 0:
                raise RuntimeError(f"[{self.name}] step {n} failed: {sim_error(sim)}")
            t += dt
            self.reports.append(r.as_dict())
            self.wall.append(t1 - t0)
            if n == warmup + steps - 1:
                device_sync()
                self.window_s = time.perf_counter() - w0
                if keep_fields:
                    self.fields = {name: sim.download(f) for name, f in COMPARE_FIELDS}
                    self.fields["img"] = sim.download(L.F_IMG)
                    self.fields["frac"] = sim.download(L.F_FRAC)
            if n < 3 or warmup <= n < warmup + steps or n == total - 1 or n % 20 == 0:
                log_step(self.name, n, r.as_dict(), 1e3 * (t1 - t0))
        device_sync()
        self.device_bytes = free0 - device_free_bytes()[0]
        self.copy_gbs = None
        if profile:
            nb = min(img.size * 8, 1 << 30)
            ms = sim.time_copy(nb, 20)
            self.copy_gbs = 2 * nb / ms / 1e6
        sim.close()
        self.warmup, self.steps, self.total = warmup, steps, total
        self.voxels = img.size
        return self

    def window(self):
        return self.reports[self.warmup:self.warmup + self.steps]

    def summary(self):
        reps = self.window()
        out = {"steps_per_s": self.steps / self.window_s,
               "ms_per_step": 1e3 * self.window_s / self.steps,
               "press_iters_per_step": [r["press_iters"] for r in reps],
               "adr_iters_per_step": [r["adr_iters"] for r in reps],
               "adr_cycles_per_step": [r["adr_cycles"] for r in reps],
               "warm_relres": [r["press_relres0"] for r in reps],
               "final_relres_max": max(r["press_relres"] for r in reps),
               "fallbacks": sum(r["press_fallback"] for r in reps),
               "unet_applies": sum(r["press_unet_applies"] for r in reps),
               "device_bytes_per_voxel": self.device_bytes / self.voxels}
        # the two solves from a cold start (step 0: p = 0; step 1: the ramp has
        # grown the drive 1e6-fold, so no earlier solution is used)
        out["hard_solves"] = [{"step": n, "ms": 1e3 * self.wall[n],
                               "press_ms": self.reports[n]["t_pressure_ms"],
                               "press_iters": self.reports[n]["press_iters"]}
                              for n in range(min(2, len(self.reports)))]
        if self.total >= 100:
            out["first_100_steps"] = {"s": sum(self.wall[:100]),
                                      "steps_per_s": 100.0 / sum(self.wall[:100])}
        return out


def sim_error(sim):
    return sim.lib.ls_last_error(sim.h).decode()


def log_step(name, n, r, ms):
    log(f"  [{name}] step {n}: {ms:.1f} ms (pressure {r['t_pressure_ms']:.1f}, "
        f"transport {r['t_transport_ms']:.1f}), press iters {r['press_iters']} "
        f"outer {r['press_outer']} fallback {r['press_fallback']} unet {r['press_unet_applies']} "
        f"warm relres {r['press_relres0']:.2e} basis {r['warm_basis']}+{r['warm_net']} "
        f"relres {r['press_relres']:.2e}, adr iters {r['adr_iters']}")


def host_cpu():
    """Model name, sockets and logical CPUs of this host (lscpu, else
    /proc/cpuinfo)."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except (OSError, ValueError):
        pass
    if "model" not in info:
        try:
            for line in Path("/proc/cpuinfo").read_text().splitlines():
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return info


def cpu_run(sub, label_bytes, dt, warmup, nsteps, reference_ksp):
    """Seconds per step of the oracle over steps warmup..warmup+nsteps-1 and
    the total linear iterations of the timed steps."""
    from oracle.oracle import OracleSimulation
    if reference_ksp:
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64,
                             press_rtol=1e-6, adr_rtol=1e-6)
    else:
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64)
    o = OracleSimulation(sub.shape, P, label_bytes)
    o.set_ksp(1 if reference_ksp else 0)
    o.read_labels(sub)
    o.init_state()
    t = 0.0
    times, its = [], [0, 0]
    for n in range(warmup + nsteps):
        t0 = time.perf_counter()
        r = o.step(t, dt)
        if n >= warmup:
            times.append(time.perf_counter() - t0)
            its[0] += r.press_iters
            its[1] += r.adr_iters_conc + r.adr_iters_acid
        t += dt
        if r.status != 0:
            break
    o.close()
    return float(np.mean(times)), len(times), its


def cpu_baseline(img, label_bytes, dt, crop: int, warmup: int, nsteps: int):
    """The FP64 oracle (C++/OpenMP, the reference's algorithm) on a bounded
    sample of the workload: the centred crop^3 of the same geometry over the
    same step window as the GPU line (warmup untimed steps, then nsteps
    timed), scaled to the full grid by voxel count, with two sets of linear
    solvers:
      reference_ksp: the reference's KSP (GMRES(30) + left Jacobi, pressure on
        the full reference-form system, every solve issued twice, rtol 1e-6
        on the preconditioned residual; src/Simulation.cpp:146-157);
      pcg_1e-8: Jacobi-PCG from the previous pressure and BiCGStab
        transport, both to 1e-8 on the true residual (the GPU line's
        accuracy).
    `value` is the faster of the two.  The solvers' iteration counts grow
    with the grid, so the scaled figures flatter the CPU."""
    os.environ.setdefault("OMP_PROC_BIND", "close")     # read when the oracle loads
    from oracle.oracle import set_threads
    ncpu = len(os.sched_getaffinity(0))
    set_threads(ncpu)
    c = [max(0, (s - crop) // 2) for s in img.shape]
    sub = np.ascontiguousarray(img[c[0]:c[0] + crop, c[1]:c[1] + crop, c[2]:c[2] + crop])
    # keep the workload's class: face slabs of Air (the surface-effect and
    # transport sources) are laid on the crop's faces too
    for ax in range(3):
        for side in (0, -1):
            n = 0
            while n < img.shape[ax] // 2 and np.all(
                    np.take(img, n if side == 0 else img.shape[ax] - 1 - n, axis=ax) == G.AIR):
                n += 1
            if n:
                sl = [slice(None)] * 3
                sl[ax] = slice(0, n) if side == 0 else slice(sub.shape[ax] - n, None)
                sub[tuple(sl)] = G.AIR
    scale = sub.size / img.size
    grid = "x".join(map(str, img.shape))
    settings = {}
    for name, ref in (("reference_ksp", True), ("pcg_1e-8", False)):
        sec, n, its = cpu_run(sub, label_bytes, dt, warmup, nsteps, ref)
        settings[name] = {"value": scale / sec, "s_per_step_on_crop": sec, "steps": n,
                          "pressure_iters": its[0], "transport_iters": its[1]}
    best = max(settings, key=lambda k: settings[k]["value"])
    return {"value": settings[best]["value"], "unit": "steps/s", "cores": ncpu, "kind": "port",
            "host": host_cpu(), "setting": best, "settings": settings,
            "sample": f"oracle FP64, steps {warmup}..{warmup + nsteps - 1} (after {warmup} "
                      f"untimed) of the workload on its centred "
                      f"{'x'.join(map(str, sub.shape))} crop, scaled by voxel count to {grid}; "
                      f"value = the faster of reference_ksp (GMRES(30)+Jacobi, rtol 1e-6, "
                      f"each solve twice) and pcg_1e-8 (Jacobi-PCG + BiCGStab to 1e-8); "
                      f"OMP threads {ncpu}"}


KERNELS = {"stencil64": (64, "k_stencil_vec<double>"), "stencil32": (32, "k_stencil_vec<float>"),
           "rb_gather": (1, "k_rb_vec<1, W>"), "rb_apply": (2, "k_rb_vec<2, W>"),
           "schur": (3, "k_schur_apply<16, 4, T>"), "transport_red": (4, "k_tschur_red_v<1, 0>"),
           "transport_black": (5, "k_tschur_black_v<0, 1>"),
           "transport_red32": (6, "k_tmix_red<4, 1>"),
           "transport_black32": (7, "k_tmix_black<0, 4, 1>")}


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


def big_run(args, img, lb, dt, desc, head, warm, extra, weights):
    """The 1024x1024x768 workload on one GPU: the headline configuration
    alone, steps 0 .. warmup+steps-1 from the initial state (cold solves
    included when warmup is 0), with the device memory the context holds
    after the last step."""
    W, K = args.warmup, args.steps
    r = Run(head, warm, **extra).run(img, lb, dt, W, K, W + K, profile=False,
                                     weights=weights if head == L.SOLVER_LEARNED or
                                     warm == L.WARM_LEARNED else None)
    reps = r.window()
    summ = r.summary()
    out = {
        "metric": "steps/s at 1024x1024x768, 1e-8 rel residual",
        "value": K / r.window_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * r.window_s / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == L.SOLVER_PLAIN64 else (
            "f32+f64+bf16" if head == L.SOLVER_LEARNED else "f32+f64"),
        "data": "synthetic",
        "config": {"workload": desc, "grid": list(img.shape), "voxels": int(img.size),
                   "solver": SOLVER_NAMES[head], "warm_start": WARM_NAMES[warm],
                   "pressure_rtol": 1e-8, "transport_rtol": 1e-8,
                   "window": f"steps {W}..{W + K - 1}",
                   "press_relres_per_step": [x["press_relres"] for x in reps],
                   "adr_relres_per_step": [x["adr_relres"] for x in reps],
                   "press_iters_per_step": summ["press_iters_per_step"],
                   "step_ms": [1e3 * w for w in r.wall[W:W + K]],
                   "device_bytes": r.device_bytes,
                   "device_bytes_per_voxel": r.device_bytes / img.size,
                   "device_total_bytes": r.device_total,
                   "labels": {str(k): int((img == k).sum()) for k in range(1, 6)}},
    }
    print(json.dumps(out), flush=True)


def stencil_only(img, label_bytes, kernel, reps, dt=0.5):
    """Assemble the pressure operator of the initial state (for a transport
    pass: make three coupled steps of the headline configuration, which
    leave the conc system's rows in place), then time `reps` launches of the
    kernel and `reps` copies of 1 GiB (HIP events)."""
    precision, kname = KERNELS[kernel]
    nk = img.shape[2]
    kname = kname.replace("W", "4" if nk % 4 == 0 else "2") if nk % 2 == 0 else \
        kname.replace("k_rb_vec", "k_rb_scalar").replace(", W", "")
    kname = kname.replace("T>", "512>" if nk <= 512 else "1024>")
    if precision in (L.TIME_TRED, L.TIME_TBLACK, L.TIME_TRED32, L.TIME_TBLACK32):
        # the FP64 passes as the FP64 form runs them, the FP32 ones in the
        # mixed-precision form
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_MIXED,
                             warm_start=L.WARM_PROJECTED,
                             adr_mixed=int(precision in (L.TIME_TRED32, L.TIME_TBLACK32)))
        sim = L.GpuSimulation(img.shape, P, label_bytes)
        sim.read_labels(img)
        sim.init_state()
        for n in range(3):
            r = sim.step(n * dt, dt)
            if r.status != 0:
                raise RuntimeError(f"step {n} failed: {sim_error(sim)}")
    else:
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


def load_traffic():
    """HBM bytes per stencil launch from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs summarised in profiles/ (None if absent)."""
    try:
        return json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None


def sampled(reps, key):
    n = sum(r[key + "_n"] for r in reps)
    ms = sum(r[key + "_ms"] for r in reps) / max(n, 1)
    return n, ms


def kernel_roofline(name, bytes_per_launch, ms, n, traffic=None):
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9 if n else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": name,
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": ms, "sampled_launches": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--total", type=int, default=100,
                    help="steps each full run makes (the workload's 100 steps at dt 0.5); the "
                         "timed window is steps warmup .. warmup+steps-1 of them")
    ap.add_argument("--workload", default="fractured512", choices=list(G.WORKLOADS))
    ap.add_argument("--size", type=int, default=None,
                    help="edge length of the fractured workload (default 512)")
    ap.add_argument("--solver", default="mixed", choices=["plain", "mixed", "learned"],
                    help="headline solver (default: mixed, the fastest measured)")
    ap.add_argument("--warm", default=None, choices=["previous", "projected", "learned"],
                    help="warm start of the headline (default: learned for the learned "
                         "solver, projected otherwise)")
    ap.add_argument("--warm-history", type=int, default=None,
                    help="projection basis size of every projected run (default: the "
                         "library's)")
    ap.add_argument("--adr-mixed", type=int, default=None, choices=[0, 1],
                    help="transport of the headline and learned runs: 1 FP32 BiCGStab "
                         "cycles inside FP64 refinement, 0 FP64 BiCGStab (default: the "
                         "library's, 1); the plain FP64 runs always use 0")
    ap.add_argument("--adr-mixed-inner-rtol", type=float, default=None,
                    help="reduction of a mixed-precision transport cycle's FP32 solve "
                         "(default: the library's, 1e-6)")
    ap.add_argument("--weights", default=None, help="U-Net weights (float32 file)")
    ap.add_argument("--no-baseline", action="store_true",
                    help="skip the plain FP64 runs (projected and previous-pressure)")
    ap.add_argument("--previous-steps", type=int, default=3,
                    help="steps of the plain FP64 previous-pressure run (every one is a "
                         "full solve of ~4000 iterations; timed from step 0)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned ablation")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-crop", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=5)
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
        stencil_only(img, lb, args.kernel, args.reps, dt)
        return
    head = {"plain": L.SOLVER_PLAIN64, "mixed": L.SOLVER_MIXED,
            "learned": L.SOLVER_LEARNED}[args.solver]
    warm = {"previous": L.WARM_PREVIOUS, "projected": L.WARM_PROJECTED,
            "learned": L.WARM_LEARNED, None: None}[args.warm]
    if warm is None:
        warm = L.WARM_LEARNED if head == L.SOLVER_LEARNED else L.WARM_PROJECTED
    extra = {} if args.warm_history is None else {"warm_history": args.warm_history}
    if args.adr_mixed is not None:
        extra["adr_mixed"] = args.adr_mixed
    if args.adr_mixed_inner_rtol is not None:
        extra["adr_mixed_inner_rtol"] = args.adr_mixed_inner_rtol
    weights = None
    from leachingsim_amd.learn import WEIGHTS, load_weights
    if head == L.SOLVER_LEARNED or warm == L.WARM_LEARNED or not args.no_learned:
        weights = load_weights(args.weights or WEIGHTS)
    W, K, T = args.warmup, args.steps, args.total
    if args.workload == "fractured1024x768":
        big_run(args, img, lb, dt, desc, head, warm, extra, weights)
        return
    keep = not args.no_baseline

    hr = Run(head, warm, **extra).run(img, lb, dt, W, K, T, profile=True,
                                      weights=weights if head == L.SOLVER_LEARNED or
                                      warm == L.WARM_LEARNED else None, keep_fields=keep)
    value = K / hr.window_s
    runs = {hr.name: hr}
    base = prev = None
    errors = None
    if not args.no_baseline:
        # like for like: plain FP64 PCG with the headline's warm start
        bw = warm if warm != L.WARM_LEARNED else L.WARM_PROJECTED
        base = Run(L.SOLVER_PLAIN64, bw, **{**extra, "adr_mixed": 0}).run(
            img, lb, dt, W, K, T, keep_fields=True)
        runs[base.name] = base
        errors = field_errors(hr.fields, base.fields)
        hr.fields = base.fields = None
        # the reference's algorithm: plain FP64 PCG from the previous pressure;
        # every step is a full solve, so a short run from step 0 is timed
        npv = max(1, args.previous_steps)
        prev = Run(L.SOLVER_PLAIN64, L.WARM_PREVIOUS, adr_mixed=0).run(img, lb, dt, 0, npv, npv)
        runs[prev.name] = prev
    learned = None
    if not args.no_learned and head != L.SOLVER_LEARNED:
        learned = Run(L.SOLVER_LEARNED, L.WARM_LEARNED, **extra).run(img, lb, dt, W, K, T,
                                                                     weights=weights)
        runs[learned.name] = learned

    # roofline: the kernel with the largest share of the timed window, the
    # transport's red Schur pass (k_tschur_red, 32 B per voxel of the grid
    # for one system: the red rows' 6 coefficients and p, w at half length),
    # timed live with HIP events on the context's stream; the pressure
    # kernels that dominate the cold solves beside it
    reps = hr.window()
    rl = {}
    n, ms = sampled(reps, "prof_tred")
    if n:
        # bytes per voxel of the grid: 32 / 40 the FP64 pass for one / two
        # systems, 16 / 20 the FP32 pass of the mixed-precision form
        per = reps[-1]["prof_tred_bytes"] / img.size
        sysn = "1" if per in (32, 16) else "3"
        kname = (f"k_tmix_red<4, {sysn}>" if per in (16, 20) else
                 f"k_tschur_red_v<{sysn}, 0>" if img.shape[2] % 4 == 0 else
                 f"k_tschur_red<{sysn}, 0>")
        rl["transport_red_pass"] = kernel_roofline(kname, reps[-1]["prof_tred_bytes"], ms, n)
    hard = hr.reports[:min(2, len(hr.reports))]
    n, ms = sampled(hard, "prof_schur")
    if n:
        b = hard[-1]["prof_schur_bytes"]
        nt = 512 if img.shape[2] <= 512 else 1024     # threads per block of the fused apply
        kname = f"k_schur_apply<16, 4, {nt}>" if b == img.size * 16 else (
            "k_rb_vec<1, 4>" if img.shape[2] % 4 == 0 else
            "k_rb_vec<1, 2>" if img.shape[2] % 2 == 0 else "k_rb_scalar<1>")
        tr = load_traffic() or {}
        t = tr.get("schur" if b == img.size * 16 else "rb_gather", {})
        rl["pressure_schur_apply_cold_solves"] = kernel_roofline(
            kname, b, ms, n, t.get("bytes_per_launch") if t.get("grid") == list(img.shape)
            else None)
    tr = load_traffic() or {}
    t = tr.get("transport_red32" if rl.get("transport_red_pass", {}).get("kernel", "")
               .startswith("k_tmix") else "transport_red", {})
    if "transport_red_pass" in rl and t.get("grid") == list(img.shape) and \
            t.get("algorithmic_bytes") == rl["transport_red_pass"]["bytes_per_launch"]:
        rl["transport_red_pass"]["traffic"] = t.get("bytes_per_launch")
    roof = rl.get("transport_red_pass") or rl.get("pressure_schur_apply_cold_solves") or \
        kernel_roofline("none", 0.0, 0.0, 0)

    stencils = stencil_rooflines(img, lb)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(img, lb, dt, min(args.cpu_crop, *img.shape), W,
                           min(args.cpu_steps, K))

    summ = {name: r.summary() for name, r in runs.items()}
    if prev is not None:
        summ[prev.name]["note"] = (f"steps 0..{prev.steps - 1} from the initial state; every "
                                   "step is a full solve from the previous pressure")
    full = {name: summ[name].get("first_100_steps") for name in summ
            if summ[name].get("first_100_steps")}
    out = {
        "metric": "steps/s at 512^3, 1e-8 rel residual: learned+mixed vs plain FP64 CG",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * hr.window_s / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == L.SOLVER_PLAIN64 else (
            "f32+f64+bf16" if head == L.SOLVER_LEARNED else "f32+f64"),
        "data": "synthetic",
        "config": {"workload": desc, "grid": list(img.shape), "solver": SOLVER_NAMES[head],
                   "warm_start": WARM_NAMES[warm], "pressure_rtol": 1e-8,
                   "transport_rtol": 1e-8, "window": f"steps {W}..{W + K - 1}",
                   "runs": summ,
                   "speedup_vs_plain_fp64": (value / summ[base.name]["steps_per_s"])
                   if base else None,
                   "speedup_vs_plain_fp64_note": "plain FP64 Jacobi-PCG with the headline's "
                   "warm start, same window",
                   "speedup_vs_plain_fp64_previous_pressure":
                       (value / summ[prev.name]["steps_per_s"]) if prev else None,
                   "full_run_first_100_steps": full,
                   "full_run_speedup_vs_plain_fp64":
                       (full[hr.name]["steps_per_s"] / full[base.name]["steps_per_s"])
                       if base and hr.name in full and base.name in full else None,
                   "error_vs_plain_fp64_after_window": errors},
        "roofline": roof,
        "rooflines": rl,
        "stencil_roofline": stencils,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Training and evaluation of the learned coarse-level pressure preconditioner.

Data.  Seeded geometries (binary Pore / Rock, porosity U[0.2, 0.45],
smoothing sigma U[1, 4], z-face inlet/outlet; the configured run uses 200 of
them at 256^3).  For each, the library's own solves give:
  * iterates of capped plain FP64 PCG solves from p = 0 at a random ramp
    time, against the reference solution of the same system (rtol 1e-10);
  * the start of the pressure solve after a coupled step (projected warm
    start plus one PCG iteration), against the reference solution at the new
    time: successive pressure fields p_t, p_t+dt.
Each sample's error e = sqrt(d) (p* - p) lives in the scaled variables of the
reduced system, with r = A^ e.  Everything per sample is built on the device
(ls_debug_learned_sample: the network input of r, rc = P^T r and e.A^e) and
kept in the network's device store together with the Galerkin operator
A_c = P^T A^ P of its system (ls_debug_learned_galerkin).

Loss.  With u = P UNet(R r) (P copies the coarse value to the solved fine
children), 1 - (u.r)^2 / (u.A^u  e.A^e): the share of the A-norm error that
the best multiple of the coarse correction leaves (the solver applies that
best multiple, the Galerkin step).  Since u.r = out.rc and
u.A^u = out.A_c out, it is exact on the network grid; loss, output gradient,
backward pass and Adam run on the GPU (ls_net_energy_step).  The host
restatements below (apply_op, prolong, restrict_sum, galerkin_operator,
energy_loss_grad) are the checks of that device path.

Training streams: geometries are generated on the host by a process pool
started before the GPU is touched, each geometry's samples enter a ring of
slots in the device store, and a fixed number of Adam steps on batches drawn
from the ring follows every geometry.

    python -m leachingsim_amd.learn train --geometries 200 --size 256 --out W
    python -m leachingsim_amd.learn evaluate --seeds 20 --size 512 --weights W
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

from . import geometry as G
from . import unet_spec as S

WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_coarse.f32"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def training_geometry(seed: int, n: int) -> np.ndarray:
    """Seeded binary geometry of the training distribution: porosity
    U[0.2, 0.45], smoothing sigma U[1, 4]."""
    rng = np.random.default_rng(10_000 + seed)
    por = rng.uniform(0.2, 0.45)
    sig = rng.uniform(1.0, 4.0)
    return G.smoothed_threshold((n, n, n), sig, por, 20_000 + seed)


# --- host restatements (checks of the device path) -------------------------
def apply_op(u, faces):
    """A^ u = u - sum over the 6 faces of c^ u_n (scaled operator)."""
    cx, cy, cz = faces
    y = u.copy()
    y[:-1] -= cx[:-1] * u[1:]
    y[1:] -= cx[:-1] * u[:-1]
    y[:, :-1] -= cy[:, :-1] * u[:, 1:]
    y[:, 1:] -= cy[:, :-1] * u[:, :-1]
    y[:, :, :-1] -= cz[:, :, :-1] * u[:, :, 1:]
    y[:, :, 1:] -= cz[:, :, :-1] * u[:, :, :-1]
    return y


def prolong(out, mask):
    """P: the coarse value on every solved fine child."""
    ni, nj, nk = mask.shape
    u = np.repeat(np.repeat(np.repeat(out, 2, 0), 2, 1), 2, 2)[:ni, :nj, :nk]
    return u * mask


def restrict_sum(g, dims):
    """P^T: sum of the fine children onto the (padded) coarse grid."""
    out = np.zeros(dims)
    for a in range(2):
        for b in range(2):
            for c in range(2):
                sub = g[a::2, b::2, c::2]
                out[:sub.shape[0], :sub.shape[1], :sub.shape[2]] += sub
    return out


def galerkin_operator(faces, mask, dims):
    """A_c = P^T A^ P on the coarse grid `dims` (2x2x2 aggregation, P copying
    a coarse value to the solved fine children): [4][ci][cj][ck] = the
    diagonal (solved children minus twice the faces inside the aggregate)
    and the couplings to the +i, +j, +k neighbours (minus the faces that
    cross into them).  u.A^u = out.A_c out for u = P out."""
    ci, cj, ck = dims
    ni, nj, nk = mask.shape
    pads = [(0, 2 * ci - ni), (0, 2 * cj - nj), (0, 2 * ck - nk)]
    m = np.pad(np.asarray(mask, dtype=np.float64), pads)

    def pair_sum(a, ax):
        # sum over the two children along every axis but `ax`, whose planes
        # were already taken one per aggregate
        shape = []
        for d, c in enumerate(dims):
            shape += [c, 1] if d == ax else [c, 2]
        return a.reshape(shape).sum(axis=(1, 3, 5))

    ac = np.zeros((4, ci, cj, ck))
    ac[0] = m.reshape(ci, 2, cj, 2, ck, 2).sum(axis=(1, 3, 5))
    for ax, c in enumerate(faces):
        c = np.array(c, dtype=np.float64)
        last = [slice(None)] * 3
        last[ax] = -1
        c[tuple(last)] = 0.0                       # no coupling past the last plane
        c = np.pad(c, pads)
        inner = [slice(None)] * 3
        cross = [slice(None)] * 3
        inner[ax] = slice(0, None, 2)              # faces inside an aggregate
        cross[ax] = slice(1, None, 2)              # faces into the next aggregate
        ac[0] -= 2.0 * pair_sum(c[tuple(inner)], ax)
        ac[1 + ax] = -pair_sum(c[tuple(cross)], ax)
    return ac


def coarse_apply(ac, out):
    """A_c out for the [4][ci][cj][ck] operator of galerkin_operator."""
    y = ac[0] * out
    y[:-1] += ac[1][:-1] * out[1:]
    y[1:] += ac[1][:-1] * out[:-1]
    y[:, :-1] += ac[2][:, :-1] * out[:, 1:]
    y[:, 1:] += ac[2][:, :-1] * out[:, :-1]
    y[:, :, :-1] += ac[3][:, :, :-1] * out[:, :, 1:]
    y[:, :, 1:] += ac[3][:, :, :-1] * out[:, :, :-1]
    return y


def energy_loss_grad(out, r, e, faces, mask):
    """Loss 1 - rho, rho = (u.r)^2 / (|u|_A^2 |e|_A^2), u = P out: the share of
    the A-norm error that the best multiple of the coarse correction leaves.
    Returns (loss, d loss / d out), on the fine grid."""
    u = prolong(out.astype(np.float64), mask)
    Au = apply_op(u, faces)
    ur = float((u * r).sum())
    uAu = float((u * Au).sum())
    eAe = float((e * r).sum())
    if uAu <= 0 or eAe <= 0:
        return 1.0, np.zeros_like(out)
    rho = ur * ur / (uAu * eAe)
    drho = (2 * ur / (uAu * eAe)) * r - (2 * ur * ur / (uAu * uAu * eAe)) * Au
    g = -restrict_sum(drho * mask, out.shape)
    return 1.0 - rho, g.astype(np.float32)


def coarse_loss_grad(out, rc, eae, ac):
    """The same loss on the network grid from rc = P^T r, e.A^e and A_c."""
    out = out.astype(np.float64)
    Au = coarse_apply(ac, out)
    ur = float((out * rc).sum())
    uAu = float((out * Au).sum())
    if uAu <= 0 or eae <= 0:
        return 1.0, np.zeros_like(out)
    g = -(2 * ur / (uAu * eae)) * rc + (2 * ur * ur / (uAu * uAu * eae)) * Au
    return 1.0 - ur * ur / (uAu * eae), g


# --- data on the device -----------------------------------------------------
def _geometry_job(args):
    seed, n = args
    return seed, training_geometry(seed, n)


class GeometryStream:
    """Seeded training geometries (seed, image), in seed order, generated by a
    pool of processes.  The pool is forked when the stream is made, so make
    it before any GPU call: the workers only run numpy."""

    def __init__(self, seeds, n, workers):
        import multiprocessing as mp
        jobs = [(s, n) for s in seeds]
        self.pool = None
        if workers > 1:
            self.pool = mp.get_context("fork").Pool(workers)
            self.it = self.pool.imap(_geometry_job, jobs)
        else:
            self.it = map(_geometry_job, jobs)

    def __iter__(self):
        return self

    def __next__(self):
        return next(self.it)

    def close(self):
        if self.pool is not None:
            self.pool.terminate()
            self.pool.join()
            self.pool = None


def geometry_samples(img, seed, ks=(30, 100, 300), dt_pair=20.0):
    """Groups (A_c, [(x, rc, eAe), ...]) of one geometry: the capped PCG
    iterates (one operator), then the start of the solve after a coupled
    step (the operator at the new time)."""
    from . import lsgpu as L
    rng = np.random.default_rng(seed)
    t = float(rng.uniform(60.0, 3500.0))
    zeros = np.zeros(S.n_weights(), np.float32)      # the restriction needs a network
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_MIXED, press_rtol=1e-10,
                         warm_start=L.WARM_PROJECTED)
    groups = []
    sim = L.GpuSimulation(img.shape, P)
    try:
        sim.read_labels(img)
        sim.init_state()
        sim.set_network(zeros)
        sim.phase(L.PH_PROPERTIES)
        rep = sim.phase(L.PH_PRESSURE, t)
        if rep.status != 0:
            raise RuntimeError(f"seed {seed}: reference solve did not converge")
        pstar = sim.download(L.F_PRESSURE)
        ac = sim.learned_galerkin()
        samples = []
        for k in ks:
            sim.set_params(L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64,
                                            press_rtol=1e-10, press_maxit=k))
            sim.upload(L.F_PRESSURE, np.zeros(img.shape))
            sim.phase(L.PH_PRESSURE, t)
            x, rc, eae = sim.learned_sample(pstar)
            if eae > 0:
                samples.append((x, rc, eae))
        groups.append((ac, samples))
        # successive fields: a coupled step at t, then the new system's start
        sim.set_params(P)
        sim.upload(L.F_PRESSURE, pstar)
        sim.step(t, dt_pair)
        sim.set_params(L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64,
                                        press_rtol=1e-10, press_maxit=1,
                                        warm_start=L.WARM_PROJECTED))
        sim.phase(L.PH_PROPERTIES)
        sim.phase(L.PH_PRESSURE, t + dt_pair)
        pstart = sim.download(L.F_PRESSURE)
        sim.set_params(P)
        rep = sim.phase(L.PH_PRESSURE, t + dt_pair)
        if rep.status == 0:
            p1 = sim.download(L.F_PRESSURE)
            sim.upload(L.F_PRESSURE, pstart)
            x, rc, eae = sim.learned_sample(p1)
            if eae > 0:
                groups.append((sim.learned_galerkin(), [(x, rc, eae)]))
    finally:
        sim.close()
    return groups


class Store:
    """Ring of device-store slots: every geometry takes `ops_per_geom`
    operator slots and `samples_per_geom` sample slots, evicting the oldest."""

    def __init__(self, net, geoms, samples_per_geom=4, ops_per_geom=2):
        self.net, self.geoms = net, geoms
        self.spg, self.opg = samples_per_geom, ops_per_geom
        net.store_alloc(geoms * samples_per_geom, geoms * ops_per_geom)
        self.filled = {}          # geometry ring position -> sample slots
        self.pos = 0

    def add(self, groups):
        g = self.pos % self.geoms
        slots = []
        s0, o0 = g * self.spg, g * self.opg
        for oi, (ac, samples) in enumerate(groups[:self.opg]):
            self.net.store_operator(o0 + oi, ac)
            for x, rc, eae in samples:
                if len(slots) == self.spg:
                    break
                s = s0 + len(slots)
                self.net.store_sample(s, o0 + oi, x, rc, eae)
                slots.append(s)
        self.filled[g] = slots
        self.pos += 1

    def slots(self):
        return [s for v in self.filled.values() for s in v]


def mean_loss(net, slots):
    """Mean energy loss over stored samples (batches of the net's size; the
    last batch is padded with repeats, which are not counted)."""
    vals = []
    B = net.batch
    for i in range(0, len(slots), B):
        chunk = slots[i:i + B]
        pad = chunk + [chunk[-1]] * (B - len(chunk))
        losses, _ = net.energy_step(pad, 0.0)
        vals += list(losses[:len(chunk)])
    return float(np.mean(vals)) if vals else float("nan")


def train(n_geom=200, n=256, holdout=8, steps_per_geom=10, batch=4, lr=1e-3, lr_end=1e-4,
          ring=48, seed=0, workers=16, out=WEIGHTS, init=None, report=None):
    """Stream n_geom seeded n^3 geometries (then `holdout` unseen ones for the
    held-out loss) through the device store, with steps_per_geom Adam steps
    after each; returns the weights and a summary."""
    from .unet import Net
    t0 = time.time()
    stream = GeometryStream(range(n_geom + holdout), n, workers)
    cdims = tuple(-(-((d + 1) // 2) // 16) * 16 for d in (n, n, n))
    net = Net(cdims, batch=batch, trainable=True,
              weights=S.init_weights(seed) if init is None else init)
    store = Store(net, ring)
    hold = Net(cdims, batch=batch, trainable=True, weights=S.init_weights(seed))
    hold_store = Store(hold, holdout)
    rng = np.random.default_rng(seed)
    losses, hist = [], []
    total = n_geom * steps_per_geom
    it = 0
    t_data = t_train = 0.0
    for gi, (gseed, img) in enumerate(stream):
        if gi == n_geom + holdout - 1:
            stream.close()
        ta = time.time()
        groups = geometry_samples(img, gseed)
        t_data += time.time() - ta
        if gi >= n_geom:                  # unseen geometries for the held-out loss
            hold_store.add(groups)
            continue
        store.add(groups)
        tb = time.time()
        avail = store.slots()
        for _ in range(steps_per_geom):
            sl = list(rng.choice(avail, batch, replace=len(avail) < batch))
            lr_t = lr * (lr_end / lr) ** (it / max(total - 1, 1))
            l_, _ = net.energy_step(sl, lr_t)
            losses.append(float(np.mean(l_)))
            it += 1
        t_train += time.time() - tb
        if (gi + 1) % 10 == 0 or gi == n_geom - 1:
            recent = float(np.mean(losses[-10 * steps_per_geom:]))
            hist.append((gi + 1, it, recent))
            log(f"  geometry {gi + 1}/{n_geom}: {it} steps, loss {recent:.4f} "
                f"(data {t_data:.0f} s, training {t_train:.0f} s, {time.time() - t0:.0f} s)")
    w = net.get_weights()
    # held-out loss, trained and untrained weights, on the unseen geometries
    held_slots = hold_store.slots()
    held_untrained = mean_loss(hold, held_slots)
    hold.set_weights(w)
    held_trained = mean_loss(hold, held_slots)
    log(f"  held-out loss on {holdout} unseen geometries: trained {held_trained:.4f}, "
        f"untrained {held_untrained:.4f}")
    net.close()
    hold.close()
    summary = {"geometries": n_geom, "size": n, "holdout": holdout,
               "samples_per_geometry": store.spg, "steps": it, "batch": batch,
               "lr": [lr, lr_end], "ring_geometries": ring,
               "held_out_loss_trained": held_trained,
               "held_out_loss_untrained": held_untrained,
               "loss_history": hist, "seconds_data": t_data, "seconds_training": t_train,
               "seconds": time.time() - t0}
    if out is not None:
        Path(out).parent.mkdir(parents=True, exist_ok=True)
        w.astype("<f4").tofile(out)
        log(f"wrote {out}")
    if report is not None:
        Path(report).write_text(json.dumps(summary, indent=1))
    return w, summary


def evaluate(weights, seeds=20, n=512, first_seed=5000, workers=16, report=None):
    """On `seeds` unseen n^3 geometries of the training distribution, a cold
    pressure solve at a random ramp time (p = 0 start, rtol 1e-8):
      * mixed PCG (the method without the network) against the learned
        solver as shipped (flexible CG raced against mixed PCG's recorded
        curve, mixed PCG taking over when it falls behind): wall time and
        iterations;
      * the energy loss of the trained and the untrained network on the
        error of a 100-iteration PCG iterate (the network's own objective).
    """
    from . import lsgpu as L
    from .unet import Net
    rows = []
    untrained = S.init_weights(0)
    stream = GeometryStream(range(first_seed, first_seed + seeds), n, workers)
    for gseed, img in stream:
        rng = np.random.default_rng(gseed)
        t = float(rng.uniform(60.0, 3500.0))
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_MIXED)
        sim = L.GpuSimulation(img.shape, P)
        row = {"seed": gseed, "porosity": float((img == G.PORE).mean()), "t": t}
        try:
            sim.read_labels(img)
            sim.init_state()
            sim.set_network(weights)
            sim.phase(L.PH_PROPERTIES)
            for name, solver in (("mixed", L.SOLVER_MIXED), ("learned", L.SOLVER_LEARNED)):
                sim.set_params(L.default_params(bc_mode=L.BC_ZFACES, solver=solver))
                sim.upload(L.F_PRESSURE, np.zeros(img.shape))
                t0 = time.perf_counter()
                rep = sim.phase(L.PH_PRESSURE, t)
                row[name] = {"s": time.perf_counter() - t0, "iters": rep.press_iters,
                             "relres": rep.press_relres, "status": rep.status,
                             "fallback": rep.press_fallback, "unet_applies": rep.press_unet_applies}
            # the network's own objective on this geometry
            sim.set_params(L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_MIXED,
                                            press_rtol=1e-10))
            sim.phase(L.PH_PRESSURE, t)
            pstar = sim.download(L.F_PRESSURE)
            ac = sim.learned_galerkin()
            sim.set_params(L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64,
                                            press_rtol=1e-10, press_maxit=100))
            sim.upload(L.F_PRESSURE, np.zeros(img.shape))
            sim.phase(L.PH_PRESSURE, t)
            x, rc, eae = sim.learned_sample(pstar)
        finally:
            sim.close()
        net = Net(rc.shape, batch=1, trainable=True, weights=weights)
        try:
            net.store_alloc(1, 1)
            net.store_operator(0, ac)
            net.store_sample(0, 0, x, rc, eae)
            row["loss_trained"] = float(net.energy_step([0], 0.0)[0][0])
            net.set_weights(untrained)
            row["loss_untrained"] = float(net.energy_step([0], 0.0)[0][0])
        finally:
            net.close()
        rows.append(row)
        if len(rows) == seeds:
            stream.close()
        log(f"  seed {gseed}: mixed {row['mixed']['s']:.2f} s / {row['mixed']['iters']} it, "
            f"learned {row['learned']['s']:.2f} s / {row['learned']['iters']} it "
            f"(fallback {row['learned']['fallback']}, U-Net {row['learned']['unet_applies']}); "
            f"energy loss trained {row['loss_trained']:.4f} untrained {row['loss_untrained']:.4f}")
    tm = [r["mixed"]["s"] for r in rows]
    tl = [r["learned"]["s"] for r in rows]
    summary = {"seeds": seeds, "size": n, "first_seed": first_seed,
               "mixed_s_mean": float(np.mean(tm)), "learned_s_mean": float(np.mean(tl)),
               "learned_over_mixed_time": float(np.mean(tl) / np.mean(tm)),
               "learned_faster_seeds": int(sum(a < b for a, b in zip(tl, tm))),
               "fallbacks": int(sum(r["learned"]["fallback"] for r in rows)),
               "energy_removed_trained": float(1 - np.mean([r["loss_trained"] for r in rows])),
               "energy_removed_untrained": float(1 - np.mean([r["loss_untrained"] for r in rows])),
               "rows": rows}
    log(f"  {seeds} unseen {n}^3 seeds: mixed {summary['mixed_s_mean']:.2f} s, learned "
        f"{summary['learned_s_mean']:.2f} s; energy removed trained "
        f"{summary['energy_removed_trained']:.3f}, untrained "
        f"{summary['energy_removed_untrained']:.4f}")
    if report is not None:
        Path(report).write_text(json.dumps(summary, indent=1))
    return summary


def load_weights(path=WEIGHTS) -> np.ndarray:
    w = np.fromfile(path, dtype="<f4")
    if w.size != S.n_weights():
        raise ValueError(f"{path}: {w.size} weights, expected {S.n_weights()}")
    return w


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--geometries", type=int, default=200)
    t.add_argument("--size", type=int, default=256)
    t.add_argument("--holdout", type=int, default=8)
    t.add_argument("--steps-per-geometry", type=int, default=10)
    t.add_argument("--batch", type=int, default=4)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-end", type=float, default=1e-4)
    t.add_argument("--ring", type=int, default=48, help="geometries kept in the device store")
    t.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    t.add_argument("--init", default=None, help="continue from these weights")
    t.add_argument("--out", default=str(WEIGHTS))
    t.add_argument("--report", default=None)
    e = sub.add_parser("evaluate")
    e.add_argument("--seeds", type=int, default=20)
    e.add_argument("--size", type=int, default=512)
    e.add_argument("--first-seed", type=int, default=5000)
    e.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    e.add_argument("--weights", default=str(WEIGHTS))
    e.add_argument("--report", default=None)
    a = ap.parse_args()
    if a.cmd == "train":
        train(a.geometries, a.size, a.holdout, a.steps_per_geometry, a.batch, a.lr, a.lr_end,
              a.ring, workers=a.workers, out=a.out,
              init=load_weights(a.init) if a.init else None, report=a.report)
    else:
        evaluate(load_weights(a.weights), a.seeds, a.size, a.first_seed, a.workers, a.report)


if __name__ == "__main__":
    main()

"""Training of the learned coarse-level pressure preconditioner.

Data: for seeded geometries (porosity U[0.2, 0.45], smoothing sigma U[1, 4],
z-face inlet/outlet), the library's own FP64 PCG solves the pressure system
at a random ramp time to 1e-10; capped solves from p = 0 give iterates p_k.
Each pair is (scaled residual r_k, scaled error e_k = x* - x_k).  Successive
pressure fields give more pairs: after a coupled step at t, the solution at
t + dt against the energy-optimal multiple of the one at t (the projected
warm start), e = x*(t+dt) - c x*(t), r = A^ e.  The network input is the
library's restriction of r (ls_debug_learned_pack).

Loss: with u = P UNet(R r) (P copies the coarse value to the solved fine
children), 1 - (u.r)^2 / (|u|_A^2 |e|_A^2): the share of the A-norm error
that the best multiple of the coarse correction leaves (the solver applies
that best multiple, the Galerkin step).  Its gradient with respect to the
network output needs one application of the scaled operator (numpy, host);
the network's backward and Adam run on the GPU with the library's kernels.

    python -m leachingsim_amd.learn --out leachingsim_amd/weights/unet_coarse.f32
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import geometry as G
from . import lsgpu as L
from . import unet_spec as S

WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_coarse.f32"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def training_geometry(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(10_000 + seed)
    por = rng.uniform(0.2, 0.45)
    sig = rng.uniform(1.0, 4.0)
    return G.smoothed_threshold((n, n, n), sig, por, 20_000 + seed)


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


def energy_loss_grad(out, r, e, faces, mask):
    """Loss 1 - rho, rho = (u.r)^2 / (|u|_A^2 |e|_A^2), u = P out: the share of
    the A-norm error that the best multiple of the coarse correction leaves.
    Returns (loss, d loss / d out)."""
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


def collect(seed: int, n: int, ks=(30, 80, 200, 400, 800)):
    """Samples (x [cgrid,16], r, e, (faces, mask)) of one seeded geometry:
    scaled residual and error of capped solves, then of the projected warm
    start after a coupled step (each with the operator it belongs to)."""
    img = training_geometry(seed, n)
    rng = np.random.default_rng(seed)
    t = float(rng.uniform(60.0, 3500.0))
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64, press_rtol=1e-10)
    sim = L.GpuSimulation(img.shape, P)
    sim.read_labels(img)
    sim.init_state()
    sim.set_network(np.zeros(S.n_weights(), np.float32))   # the restriction needs a network
    sim.phase(L.PH_PROPERTIES)
    sim.phase(L.PH_PRESSURE, t)
    pstar = sim.download(L.F_PRESSURE)
    sd, bh, cl = sim.pressure_system()
    faces = sim.pressure_faces()
    R = cl == 1
    xs = sd * pstar
    geo = (faces.astype(np.float32), R.astype(np.float32))
    samples = []
    for k in ks:
        sim.set_params(L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64,
                                        press_rtol=1e-10, press_maxit=k))
        sim.upload(L.F_PRESSURE, np.zeros(img.shape))
        sim.phase(L.PH_PRESSURE, t)
        pk = sim.download(L.F_PRESSURE)
        e = np.where(R, xs - sd * pk, 0.0)
        r = np.where(R, apply_op(e, faces), 0.0)
        if not np.any(r):
            continue
        x = sim.learned_pack(r.astype(np.float32))
        samples.append((x, r.astype(np.float32), e.astype(np.float32), geo))
    sim.close()
    return samples + successive_pairs(img, t, rng)


def successive_pairs(img, t, rng, dts=(20.0, 60.0)):
    """(p_t, p_t+dt) pairs: the error of the projected warm start after a
    coupled step (the operator changes with the saturation)."""
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64, press_rtol=1e-10)
    out = []
    for dt in dts:
        sim = L.GpuSimulation(img.shape, P)
        sim.read_labels(img)
        sim.init_state()
        sim.set_network(np.zeros(S.n_weights(), np.float32))
        sim.step(t, dt)
        p0 = sim.download(L.F_PRESSURE)
        sim.phase(L.PH_PROPERTIES)
        sim.phase(L.PH_PRESSURE, t + dt)
        p1 = sim.download(L.F_PRESSURE)
        sd, bh, cl = sim.pressure_system()
        faces = sim.pressure_faces()
        R = cl == 1
        xs = np.where(R, sd * p1, 0.0)
        v = np.where(R, sd * p0, 0.0)
        av = apply_op(v, faces)
        b = apply_op(xs, faces)
        den = float((v * av).sum())
        c = float((v * b).sum()) / den if den > 0 else 0.0
        e = np.where(R, xs - c * v, 0.0)
        r = np.where(R, apply_op(e, faces), 0.0)
        if np.linalg.norm(r) > 1e-12 * max(np.linalg.norm(b), 1e-300):
            x = sim.learned_pack(r.astype(np.float32))
            out.append((x, r.astype(np.float32), e.astype(np.float32),
                        (faces.astype(np.float32), R.astype(np.float32))))
        sim.close()
    return out


def train(n_geom=32, n=64, steps=1500, batch=4, lr=1e-3, seed=0, log_every=100, out=WEIGHTS,
          holdout=4):
    from .unet import Net
    t0 = time.time()
    train_set, test_set = [], []
    for g in range(n_geom + holdout):
        (train_set if g < n_geom else test_set).extend(collect(g, n))
    log(f"collected {len(train_set)} training / {len(test_set)} held-out pairs from "
        f"{n_geom}+{holdout} geometries of {n}^3 in {time.time() - t0:.1f} s")
    cshape = train_set[0][0].shape[:3]
    net = Net(cshape, batch=batch, trainable=True, weights=S.init_weights(seed))
    rng = np.random.default_rng(seed)

    def evaluate(samples):
        vals = []
        for i in range(0, len(samples) - batch + 1, batch):
            chunk = samples[i:i + batch]
            outs = net.forward(np.stack([c[0] for c in chunk]))
            for o, (x, r, e, (faces, mask)) in zip(outs, chunk):
                vals.append(energy_loss_grad(o, r, e, faces, mask)[0])
        return float(np.mean(vals)) if vals else float("nan")

    losses = []
    log(f"  held-out loss before training: {evaluate(test_set):.4f}")
    t1 = time.time()
    for it in range(steps):
        idx = rng.choice(len(train_set), batch, replace=False)
        chunk = [train_set[i] for i in idx]
        outs = net.forward(np.stack([c[0] for c in chunk]))
        gs, ls = [], []
        for o, (x, r, e, (faces, mask)) in zip(outs, chunk):
            l_, g_ = energy_loss_grad(o, r, e, faces, mask)
            ls.append(l_)
            gs.append(g_ / batch)
        net.step(np.stack(gs), lr * (0.1 ** (it / steps)))
        losses.append(float(np.mean(ls)))
        if (it + 1) % log_every == 0:
            log(f"  step {it + 1}: loss {np.mean(losses[-log_every:]):.4f} ({time.time() - t1:.1f} s)")
    held = evaluate(test_set)
    log(f"  held-out loss after training: {held:.4f}")
    w = net.get_weights()
    net.close()
    if out is not None:
        Path(out).parent.mkdir(parents=True, exist_ok=True)
        w.astype("<f4").tofile(out)
        log(f"wrote {out}")
    return w, losses, held


def load_weights(path=WEIGHTS) -> np.ndarray:
    w = np.fromfile(path, dtype="<f4")
    if w.size != S.n_weights():
        raise ValueError(f"{path}: {w.size} weights, expected {S.n_weights()}")
    return w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--geometries", type=int, default=32)
    ap.add_argument("--size", type=int, default=64)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=str(WEIGHTS))
    a = ap.parse_args()
    train(a.geometries, a.size, a.steps, a.batch, a.lr, out=a.out)


if __name__ == "__main__":
    main()

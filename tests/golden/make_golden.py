"""Writes the known-answer vectors in this folder from closed forms, in
50-digit decimal arithmetic, independently of the oracle and the kernels.

The reference ships no golden vectors, tests or fixtures; these are the
mathematics of the formulas it implements (src/Simulation.cpp):
  vg.json        van Genuchten with n = 2, alpha = 1, s_res = 0.05 (:1596-1664):
                 krel = sqrt(Se) (1 - sqrt(1 - Se^2))^2, Se = clamp((s - 0.05) / 0.95, 0, 1);
                 Pc = sqrt(Se^-2 - 1), Se clamped to [1e-6, 1], Pc = 0 at Se = 1
  perm.json      Pore intrinsic permeability at precipitate inventory p (:187-248):
                 10^(p log10 1e-18 + (1 - p) log10 1e-12), times krel (1 at s = 1)
  channel.json   all-Pore (17, 2, 2) bar with the z-face planes at t = 1000 s: resistors
                 in series, end links harmonic(K, 1e-9), interior links K, over mu = 1e-3
                 (:662-798); pressure at every plane
  two_layer.json (12, 2, 2) bar, inventory 0.1 in the second half, t = 3600 s: the flux
                 through every link = ptop / sum of link resistances

    python tests/golden/make_golden.py
"""
import json
from decimal import Decimal as D, getcontext
from pathlib import Path

getcontext().prec = 50
HERE = Path(__file__).resolve().parent


def se_k(s):
    se = (D(s) - D("0.05")) / D("0.95")
    return min(D(1), max(D(0), se))


def krel(s):
    se = se_k(s)
    if se <= 0:
        return D(0)
    return se.sqrt() * (D(1) - (D(1) - se * se).sqrt()) ** 2


def pc(s):
    se = min(D(1), max(D("1e-6"), (D(s) - D("0.05")) / D("0.95")))
    if se >= 1:
        return D(0)
    return (D(1) / (se * se) - D(1)).sqrt()


def pow10(x):
    return (x * D(10).ln()).exp()


def hmean(a, b):
    return D(2) * a * b / (a + b)


def main():
    s_list = ["0.0", "0.03", "0.05", "0.1", "0.37", "0.9", "0.999999", "1.0"]
    json.dump({"s": [float(D(s)) for s in s_list],
               "krel": [str(krel(s)) for s in s_list],
               "pc": [str(pc(s)) for s in s_list]}, open(HERE / "vg.json", "w"), indent=1)

    p_list = ["0.0", "0.1", "0.25", "0.5", "0.75", "1.0"]
    perm = [pow10(D(p) * D(-18) + (D(1) - D(p)) * D(-12)) for p in p_list]
    json.dump({"inventory": [float(D(p)) for p in p_list], "perm_at_s1": [str(k) for k in perm]},
              open(HERE / "perm.json", "w"), indent=1)

    ni, t = 17, D(1000)
    ptop = D(10000) * t / D(3600) + D("1e-6")
    K = D("1e-12") * krel("0.1")
    mu = D("1e-3")
    c_mid = hmean(K, K) / mu
    c_end = hmean(K, D("1e-9")) / mu
    R = D(2) / c_end + D(ni - 1) / c_mid
    Q = ptop / R
    p = [ptop - Q / c_end - D(i) * Q / c_mid for i in range(ni)]
    json.dump({"shape": [ni, 2, 2], "t": float(t), "pressure": [str(x) for x in p]},
              open(HERE / "channel.json", "w"), indent=1)

    ni = 12
    kr = krel("0.1")
    Ks = [(D("1e-12") if i < ni // 2 else pow10(D("-12.6"))) * kr for i in range(ni)]
    links = [hmean(Ks[0], D("1e-9"))] + [hmean(Ks[i], Ks[i + 1]) for i in range(ni - 1)] + \
        [hmean(Ks[-1], D("1e-9"))]
    ptop = D(10000) + D("1e-6") * 0   # t = 3600: the ramp has ended
    Q = ptop / sum(mu / c for c in links)
    json.dump({"shape": [ni, 2, 2], "t": 3600.0, "inventory_second_half": 0.1,
               "link_conductance": [str(c / mu) for c in links], "flux": str(Q)},
              open(HERE / "two_layer.json", "w"), indent=1)


if __name__ == "__main__":
    main()

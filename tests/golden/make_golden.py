"""Generate the committed golden vectors under tests/golden/ (run once, offline).

    python tests/golden/make_golden.py

No reference code is imported or run: FLOWVPM (the package that holds the
hot path, Project.toml:25) is not in the reference tree.  Every vector here is
computed at 40 significant digits by mpmath along a route that shares no
formula with the oracle's pair algebra:

kernels.json
    g(rho) by quadrature 4 pi int_0^rho s^2 zeta(s) ds of the radial basis
    zeta of docs/src/theory/rvpm.md:84 (Gaussian filter rvpm.md:365,
    Winckelmans & Leonard 1993 rvpm.md:380), g'(rho) = 4 pi rho^2 zeta(rho)
    and the normalisation 4 pi int_0^inf s^2 zeta ds = 1.

uj_<kernel>_<case>.json
    U and J = grad U from the vector potential psi(x) = sum_j Gamma_j
    G_sigma_j(|x - x_j|), u = curl psi, with G the free-space solution of
    lap G = -zeta_sigma (checked here by a radial Laplacian at 40 digits) and
    every derivative taken by mpmath.diff (high-precision finite differences).
    The self pair is skipped, as the reference pair loop does for r == 0
    (src/FLOWUnsteady_processing_force.jl:895).  Estr of rvpm.md:251-262 in
    both stretching conventions is assembled from these J.
"""
from __future__ import annotations

import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))
PI = mp.pi


def zeta(kernel, rho):
    rho = mp.mpf(rho)
    if kernel == "gaussian":
        return 3 / (4 * PI) * mp.e ** (-rho ** 3)
    if kernel == "gaussianerf":
        return (2 * PI) ** mp.mpf(-1.5) * mp.e ** (-rho ** 2 / 2)
    if kernel == "winckelmans":
        return 15 / (8 * PI) * (rho ** 2 + 1) ** mp.mpf(-3.5)
    raise ValueError(kernel)


def g_quad(kernel, rho):
    return 4 * PI * mp.quad(lambda s: s ** 2 * zeta(kernel, s), [0, rho])


def green(kernel, r, sigma):
    """Free-space G with lap G = -sigma^-3 zeta(r/sigma)."""
    rho = r / sigma
    if kernel == "gaussianerf":
        return mp.erf(rho / mp.sqrt(2)) / (4 * PI * r)
    if kernel == "winckelmans":
        return (rho ** 2 + mp.mpf(1.5)) / (4 * PI * sigma * (rho ** 2 + 1) ** mp.mpf(1.5))
    if kernel == "singular":
        return 1 / (4 * PI * r)
    raise ValueError(kernel)


def check_green(kernel):
    """Radial Laplacian G'' + 2 G'/r must equal -zeta_sigma."""
    sigma = mp.mpf("0.37")
    for r in ("0.05", "0.4", "1.3", "3.0"):
        r = mp.mpf(r)
        G = lambda s: green(kernel, s, sigma)
        lap = mp.diff(G, r, 2) + 2 * mp.diff(G, r, 1) / r
        z = zeta(kernel, r / sigma) / sigma ** 3
        assert abs(lap + z) < mp.mpf(10) ** -25 * max(1, abs(z)), (kernel, r, lap, z)


EPS = {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (2, 1, 0): -1, (1, 0, 2): -1}


def uj_golden(kernel, X, G, S):
    n = len(X)
    Xm = [[mp.mpf(float(v)) for v in x] for x in X]
    Gm = [[mp.mpf(float(v)) for v in g] for g in G]
    Sm = [mp.mpf(float(v)) for v in S]
    U = np.zeros((n, 3))
    J = np.zeros((n, 9))
    for i in range(n):
        u = [mp.mpf(0)] * 3
        jm = [[mp.mpf(0)] * 3 for _ in range(3)]
        for j in range(n):
            if j == i:
                continue
            xj, sj = Xm[j], Sm[j]

            def Gf(a, b, c, xj=xj, sj=sj):
                r = mp.sqrt((a - xj[0]) ** 2 + (b - xj[1]) ** 2 + (c - xj[2]) ** 2)
                return green(kernel, r, sj)

            p = tuple(Xm[i])
            grad = [mp.diff(Gf, p, tuple(1 if q == a else 0 for q in range(3))) for a in range(3)]
            hess = [[None] * 3 for _ in range(3)]
            for a in range(3):
                for l in range(a, 3):
                    order = [0, 0, 0]
                    order[a] += 1
                    order[l] += 1
                    hess[a][l] = hess[l][a] = mp.diff(Gf, p, tuple(order))
            # u_k = eps_kab d_a G Gamma_b ;  J_kl = eps_kab d_l d_a G Gamma_b
            for (k, a, b), s in EPS.items():
                u[k] += s * grad[a] * Gm[j][b]
                for l in range(3):
                    jm[k][l] += s * hess[a][l] * Gm[j][b]
        U[i] = [float(v) for v in u]
        for k in range(3):
            for l in range(3):
                J[i, k + 3 * l] = float(jm[k][l])
    return U, J


def estr_golden(kernel, transposed, X, G, S, J):
    n = len(X)
    E = np.zeros((n, 3))
    for i in range(n):
        e = [mp.mpf(0)] * 3
        for j in range(n):
            r = mp.sqrt(sum((mp.mpf(float(X[i][k])) - mp.mpf(float(X[j][k]))) ** 2 for k in range(3)))
            sj = mp.mpf(float(S[j]))
            zs = zeta(kernel, r / sj) / sj ** 3
            dJ = [mp.mpf(float(J[i, m])) - mp.mpf(float(J[j, m])) for m in range(9)]
            for k in range(3):
                acc = mp.mpf(0)
                for l in range(3):
                    jkl = dJ[l + 3 * k] if transposed else dJ[k + 3 * l]
                    acc += jkl * mp.mpf(float(G[j][l]))
                e[k] += zs * acc
        E[i] = [float(v) for v in e]
    return E


def kernels_json():
    out = {}
    for k in ("gaussian", "gaussianerf", "winckelmans"):
        rows = []
        for rho in ("0.05", "0.3", "1", "2", "3.5", "6"):
            rows.append({"rho": float(rho), "zeta": float(zeta(k, rho)),
                         "g": float(g_quad(k, mp.mpf(rho))),
                         "dg": float(4 * PI * mp.mpf(rho) ** 2 * zeta(k, rho))})
        norm = 4 * PI * mp.quad(lambda s: s ** 2 * zeta(k, s), [0, 1, 5, mp.inf])
        out[k] = {"rows": rows, "normalisation": float(norm), "zeta0": float(zeta(k, 0))}
    return out


def case(seed, n, lo, hi, sig_lo, sig_hi, gstd):
    rng = np.random.default_rng(seed)
    X = rng.uniform(lo, hi, (n, 3))
    G = rng.normal(0.0, gstd, (n, 3))
    S = rng.uniform(sig_lo, sig_hi, n)
    return X, G, S


def main():
    for k in ("gaussianerf", "winckelmans"):
        check_green(k)
    with open(os.path.join(HERE, "kernels.json"), "w") as fh:
        json.dump(kernels_json(), fh, indent=1)
    cases = {
        # spread: rho from ~0.1 to ~10 across pairs
        "spread": dict(seed=11, n=24, lo=-1.0, hi=1.0, sig_lo=0.3, sig_hi=0.6, gstd=1.0),
        # clustered: many overlapping cores (rho < 1)
        "cluster": dict(seed=12, n=20, lo=-0.15, hi=0.15, sig_lo=0.1, sig_hi=0.25, gstd=1e-3),
    }
    for kern in ("gaussianerf", "winckelmans"):
        for cname, c in cases.items():
            X, G, S = case(**c)
            U, J = uj_golden(kern, X, G, S)
            Et = estr_golden(kern, True, X, G, S, J)
            Ec = estr_golden(kern, False, X, G, S, J)
            doc = {"kernel": kern, "case": cname, "params": c,
                   "X": X.tolist(), "Gamma": G.tolist(), "sigma": S.tolist(),
                   "U": U.tolist(), "J": J.tolist(),
                   "Estr_transposed": Et.tolist(), "Estr_classic": Ec.tolist()}
            with open(os.path.join(HERE, f"uj_{kern}_{cname}.json"), "w") as fh:
                json.dump(doc, fh)
            print("wrote", kern, cname, flush=True)


if __name__ == "__main__":
    main()

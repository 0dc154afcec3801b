"""Seeded synthetic particle sets for the parity tests and the benchmark.

Every generator returns (X (n,3), Gamma (n,3), sigma (n,), vol (n,)) as
float64 numpy arrays, deterministically from its seed (numpy PCG64).
"N(0, s)" below means standard deviation s.

  vortex_ring      thick ring R=1, core radius a=0.1: 40 azimuthal stations x
                   200 particles on a Vogel spiral over the core cross-section,
                   tangential Gamma with total circulation 1, sigma = 0.08,
                   positions jittered by N(0, 1e-3).                 (N = 8,000)
  uniform_cube     X ~ U[-1,1]^3, Gamma ~ N(0, 1e-3) iid, sigma ~ U[0.02, 0.05],
                   vol = sigma^3.                                      (N = 2^20)
  rotor_wake       3 helical tip filaments (radius 1, pitch 0.3, 40 turns, phases
                   0/120/240 deg) plus a root-vortex line on the z axis over
                   [0, 12], N/4 particles each; Gamma = 0.05 t_hat + N(0, 5e-4),
                   sigma = 0.03.                                       (N = 2^22)
  gaussian_blob    X ~ N(0, 0.5^2)^3, Gamma = 0.01 (-y, x, 0) + N(0, 1e-3),
                   sigma = exp(N(ln 0.03, 0.2)), vol = sigma^3.        (N = 2^21)
  batched_fields   64 fields with N_k = floor(exp(U[ln 5e3, ln 2.5e5])),
                   alternating rings and helices, sigma_k ~ U[0.02, 0.08].
"""
from __future__ import annotations

import numpy as np

GOLDEN_ANGLE_DEG = 137.508


def vortex_ring(n_stations=40, n_cross=200, R=1.0, a=0.1, sigma=0.08, jitter=1e-3,
                circulation=1.0, seed=1):
    rng = np.random.default_rng(seed)
    k = np.arange(1, n_cross + 1)
    rk = a * np.sqrt((k - 0.5) / n_cross)
    th = np.deg2rad(k * GOLDEN_ANGLE_DEG)
    phi = 2 * np.pi * np.arange(n_stations) / n_stations
    PHI, RK = np.meshgrid(phi, rk, indexing="ij")
    _, TH = np.meshgrid(phi, th, indexing="ij")
    PHI, RK, TH = PHI.ravel(), RK.ravel(), TH.ravel()
    rho = R + RK * np.cos(TH)                       # distance from the ring axis
    X = np.stack([rho * np.cos(PHI), rho * np.sin(PHI), RK * np.sin(TH)], axis=1)
    t_hat = np.stack([-np.sin(PHI), np.cos(PHI), np.zeros_like(PHI)], axis=1)
    dl = 2 * np.pi * rho / n_stations
    Gamma = (circulation / n_cross) * dl[:, None] * t_hat
    vol = (np.pi * a * a / n_cross) * dl
    X = X + rng.normal(0.0, jitter, X.shape)
    sig = np.full(X.shape[0], float(sigma))
    return X, Gamma, sig, vol


def uniform_cube(n=1 << 20, seed=2, gstd=1e-3, smin=0.02, smax=0.05):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1.0, 1.0, (n, 3))
    Gamma = rng.normal(0.0, gstd, (n, 3))
    sig = rng.uniform(smin, smax, n)
    return X, Gamma, sig, sig ** 3


def helix(n, radius=1.0, pitch=0.3, turns=40, phase=0.0):
    th = np.linspace(0.0, 2 * np.pi * turns, n, endpoint=False)
    X = np.stack([radius * np.cos(th + phase), radius * np.sin(th + phase),
                  pitch * th / (2 * np.pi)], axis=1)
    T = np.stack([-radius * np.sin(th + phase), radius * np.cos(th + phase),
                  np.full_like(th, pitch / (2 * np.pi))], axis=1)
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    return X, T


def rotor_wake(n=1 << 22, seed=3, gamma=0.05, gstd=5e-4, sigma=0.03, turns=40, pitch=0.3):
    rng = np.random.default_rng(seed)
    m = n // 4
    Xs, Ts = [], []
    for ph in (0.0, 2 * np.pi / 3, 4 * np.pi / 3):
        X, T = helix(m, 1.0, pitch, turns, ph)
        Xs.append(X)
        Ts.append(T)
    mr = n - 3 * m
    z = np.linspace(0.0, pitch * turns, mr, endpoint=False)
    Xs.append(np.stack([np.zeros(mr), np.zeros(mr), z], axis=1))
    Ts.append(np.tile([0.0, 0.0, 1.0], (mr, 1)))
    X = np.concatenate(Xs)
    Gamma = gamma * np.concatenate(Ts) + rng.normal(0.0, gstd, (n, 3))
    sig = np.full(n, float(sigma))
    return X, Gamma, sig, sig ** 3


def gaussian_blob(n=1 << 21, seed=4, std=0.5, swirl=0.01, gstd=1e-3, mu=np.log(0.03), s=0.2):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, std, (n, 3))
    Gamma = swirl * np.stack([-X[:, 1], X[:, 0], np.zeros(n)], axis=1) + rng.normal(0.0, gstd, (n, 3))
    sig = np.exp(rng.normal(mu, s, n))
    return X, Gamma, sig, sig ** 3


def batched_fields(nfields=64, seed=5, nmin=5e3, nmax=2.5e5):
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nfields))).astype(int)
    out = []
    for k, nk in enumerate(sizes):
        r = np.random.default_rng(seed + k + 1)
        sig = r.uniform(0.02, 0.08)
        if k % 2 == 0:
            nst = max(8, int(round(np.sqrt(nk / 5))))
            ncr = max(1, nk // nst)
            X, G, s, v = vortex_ring(n_stations=nst, n_cross=ncr, sigma=sig, seed=seed + k + 1)
        else:
            X, T = helix(nk, 1.0, 0.3, 10, r.uniform(0, 2 * np.pi))
            G = 0.05 * T + r.normal(0.0, 5e-4, (nk, 3))
            s = np.full(nk, sig)
            v = s ** 3
        out.append((X, G, s, v))
    return out


CONFIGS = {
    "ring8k": lambda: vortex_ring(),
    "cube1m": lambda: uniform_cube(),
    "rotor4m": lambda: rotor_wake(),
    "blob2m": lambda: gaussian_blob(),
}

"""Pins of the oracle's CPU FMM (oracle/fmm_oracle.c) against the direct sum
and closed forms -- nothing here restates the expansion algebra, except the
M2L sum written out by its definition to check the matrix-core M2L's unit
convention (lattice offsets) against direct separations."""
import math

import numpy as np
import pytest

from helpers import normwise
from oracle import oracle as O


def cloud(n=2500, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, 3))
    G = rng.normal(0, 1e-3, (n, 3))
    s = rng.uniform(0.02, 0.05, n)
    return X, G, s


def test_derivative_recurrence_matches_finite_differences():
    """b_k = d^k(1/r)/k!: compare with symbolic derivatives at a point."""
    r = np.array([0.7, -0.4, 1.1])
    b, mi = O.fmm_deriv(4, r)
    R = np.linalg.norm(r)
    idx = {tuple(m): i for i, m in enumerate(mi)}
    assert b[idx[(0, 0, 0)]] == pytest.approx(1 / R, rel=1e-15)
    assert b[idx[(1, 0, 0)]] == pytest.approx(-r[0] / R ** 3, rel=1e-14)
    # d2/dx dy (1/r) = 3 x y / r^5 ; /k! = 1
    assert b[idx[(1, 1, 0)]] == pytest.approx(3 * r[0] * r[1] / R ** 5, rel=1e-14)
    # d2/dx2 (1/r) = (3x^2 - r^2)/r^5 ; /2!
    assert b[idx[(2, 0, 0)]] == pytest.approx((3 * r[0] ** 2 - R ** 2) / R ** 5 / 2, rel=1e-14)
    # Laplace: sum of pure second derivatives vanishes -> b200 + b020 + b002 = 0
    assert abs(b[idx[(2, 0, 0)]] + b[idx[(0, 2, 0)]] + b[idx[(0, 0, 2)]]) < 1e-14
    # a fourth-order mixed term by central differences of the third-order ones
    h = 1e-4
    bp, _ = O.fmm_deriv(4, r + [0, 0, h])
    bm, _ = O.fmm_deriv(4, r - [0, 0, h])
    fd = (bp[idx[(2, 1, 0)]] - bm[idx[(2, 1, 0)]]) / (2 * h)   # d/dz of b_210 = b_211 * 1
    assert b[idx[(2, 1, 1)]] == pytest.approx(fd, rel=1e-7)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_error_falls_with_p(kernel):
    X, G, s = cloud()
    U0, J0 = O.uj_direct(kernel, X, X, G, s)
    errs = []
    for p in (2, 4, 6, 8, 10):
        U, J, st = O.uj_fmm(kernel, X, G, s, p=p, ncrit=50, theta=0.4,
                            sigcut=6.0 if kernel == "gaussianerf" else 25.0)
        assert st["nm2l"] > 0 and st["np2p"] > 0
        errs.append((normwise(U, U0)[0], normwise(J, J0)[0]))
    for (u1, j1), (u2, j2) in zip(errs, errs[1:]):
        assert u2 < 0.5 * u1 and j2 < 0.5 * j1, errs
    assert errs[-1][0] < 1e-6 and errs[-1][1] < 1e-6, errs


def test_theta_zero_is_direct():
    X, G, s = cloud(1500, seed=3)
    U0, J0 = O.uj_direct("gaussianerf", X, X, G, s)
    U, J, st = O.uj_fmm("gaussianerf", X, G, s, p=4, theta=1e-12)
    assert st["nm2l"] == 0
    assert st["p2p_pairs"] == len(X) ** 2
    assert normwise(U, U0)[0] < 1e-13 and normwise(J, J0)[0] < 1e-13


def test_far_cluster_taylor_remainder():
    """Two well-separated clusters: the far field of one on the other is the
    singular Biot-Savart sum up to the truncation error, which shrinks
    geometrically with p."""
    rng = np.random.default_rng(5)
    A = rng.uniform(-0.1, 0.1, (200, 3))
    B = rng.uniform(-0.1, 0.1, (200, 3)) + [3.0, 0.0, 0.0]
    X = np.concatenate([A, B])
    G = rng.normal(0, 1, (400, 3))
    s = np.full(400, 1e-3)
    U0, J0 = O.uj_direct("singular", X, X, G, s)
    e4 = normwise(O.uj_fmm("singular", X, G, s, p=4, ncrit=300, theta=0.4, sigcut=0)[0], U0)[0]
    e8 = normwise(O.uj_fmm("singular", X, G, s, p=8, ncrit=300, theta=0.4, sigcut=0)[0], U0)[0]
    assert e8 < 1e-3 * e4 or e8 < 1e-12


def test_sample_mode_matches_full():
    X, G, s = cloud(2000, seed=7)
    Uf, Jf, _ = O.uj_fmm("gaussianerf", X, G, s, p=6)
    idx = np.array([3, 100, 1999, 1500])
    Us, Js, st = O.uj_fmm("gaussianerf", X, G, s, p=6, sample=idx)
    np.testing.assert_array_equal(Us[idx], Uf[idx])
    np.testing.assert_array_equal(Js[idx], Jf[idx])


def test_m2l_lattice_offset_scaling():
    """The unit convention of the matrix-core M2L's operator table: with
    u = ext 2^-(lf+1) and a pair's separation u v (v integer, lf the finer
    level), l'_b = tau_A^-(|b|+1) sum_a [2^(dA(|a+b|+1)) E_{a+b}(v)]
    [m'_a tau_S^-|a| 2^((l_A - l_S)|a|)], tau = a cell's half width and
    dA = lf - l_A, equals the direct l'_b = sum_a E_{a+b}(c_A - c_S) m'_a."""
    P = 6
    rng = np.random.default_rng(3)
    ext, lo = 1.7, np.array([-0.3, 0.2, 0.1])
    _, mi = O.fmm_deriv(P, np.array([1.0, 0.0, 0.0]))
    mi = [tuple(m) for m in mi]
    idx = {m: i for i, m in enumerate(mi)}
    deg = np.array([sum(m) for m in mi])
    fact = np.array([float(math.prod(math.factorial(x) for x in m)) for m in mi])

    def E(r):
        b, _ = O.fmm_deriv(P, np.asarray(r, float))
        return b * fact

    def m2l(Ek, mp):
        out = np.zeros(len(mi))
        for j, b in enumerate(mi):
            for i, a in enumerate(mi):
                if sum(a) + sum(b) <= P:
                    out[j] += Ek[idx[(a[0] + b[0], a[1] + b[1], a[2] + b[2])]] * mp[i]
        return out

    for lA, lS in ((5, 5), (6, 5), (5, 7), (4, 6)):
        iA, iS = rng.integers(0, 2 ** lA, 3), rng.integers(0, 2 ** lS, 3)
        hA, hS = ext / 2 ** lA, ext / 2 ** lS
        cA, cS = lo + (iA + 0.5) * hA, lo + (iS + 0.5) * hS
        mp = rng.standard_normal(len(mi))
        ref = m2l(E(cA - cS), mp)
        lf = max(lA, lS)
        dA, dS = lf - lA, lf - lS
        v = (2 * iA + 1) * 2 ** dA - (2 * iS + 1) * 2 ** dS
        Ehat = E(v) * 2.0 ** (dA * (deg + 1))
        mpp = mp * (hS / 2) ** (-deg.astype(float)) * 2.0 ** ((lA - lS) * deg)
        out = m2l(Ehat, mpp) * (hA / 2) ** (-(deg + 1).astype(float))
        assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max(), (lA, lS)

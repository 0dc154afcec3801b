"""Pins of the CPU FP64 oracle (oracle/vpm_oracle.c).

No test here retypes the oracle's pair formula: each compares it with a route
that shares no algebra with it -- the 40-digit vector-potential golden vectors
(tests/golden/make_golden.py), finite differences of the oracle's own U,
closed-form flows of vortex rings and lines, the exact solution of y' = lambda y
and conservation identities of the rVPM equations.
"""
import json
import os

import numpy as np
import pytest

from helpers import normwise
from oracle import oracle as O

KERNELS = ["gaussianerf", "winckelmans"]
CASES = ["spread", "cluster"]


def load(golden_dir, kernel, case):
    with open(os.path.join(golden_dir, f"uj_{kernel}_{case}.json")) as fh:
        d = json.load(fh)
    return {k: np.array(v) if isinstance(v, list) else v for k, v in d.items()}


# ---- kernels -----------------------------------------------------------------------
def test_kernels_vs_quadrature(golden_dir):
    with open(os.path.join(golden_dir, "kernels.json")) as fh:
        gold = json.load(fh)
    for name, d in gold.items():
        assert abs(d["normalisation"] - 1.0) < 1e-14, name      # int zeta dV = 1
        assert abs(O.zeta0(name) - d["zeta0"]) < 1e-15 * d["zeta0"]
        for row in d["rows"]:
            z, g, dg = O.kernel(name, row["rho"])
            assert z == pytest.approx(row["zeta"], rel=1e-14)
            assert dg == pytest.approx(row["dg"], rel=1e-14)
            # closed-form g vs quadrature of 4 pi rho^2 zeta; the closed form
            # cancels to ~eps/rho^2 relative at small rho
            assert g == pytest.approx(row["g"], rel=1e-14 + 1e-15 / row["rho"] ** 2)


@pytest.mark.parametrize("name", ["gaussian", "gaussianerf", "winckelmans"])
def test_kernel_limits(name):
    assert O.kernel(name, 0.0)[1] == 0.0                   # g(0) = 0
    assert O.kernel(name, 1e3)[1] == pytest.approx(1.0, abs=1e-8)   # g(inf) = 1
    for rho in (0.2, 0.9, 2.7):                            # g' = 4 pi rho^2 zeta
        z, g, dg = O.kernel(name, rho)
        assert dg == pytest.approx(4 * np.pi * rho ** 2 * z, rel=1e-14)


# ---- U and J against the independent 40-digit route -------------------------------------
@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", CASES)
def test_uj_golden(golden_dir, kernel, case):
    d = load(golden_dir, kernel, case)
    U, J = O.uj_direct(kernel, d["X"], d["X"], d["Gamma"], d["sigma"])
    assert normwise(U, d["U"])[0] < 1e-13
    assert normwise(J, d["J"])[0] < 1e-13


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("transposed", [True, False])
def test_estr_golden(golden_dir, kernel, case, transposed):
    d = load(golden_dir, kernel, case)
    X, G, S = d["X"], d["Gamma"], d["sigma"]
    E = O.estr_direct(kernel, transposed, X, d["J"], X, G, S, d["J"])
    ref = d["Estr_transposed" if transposed else "Estr_classic"]
    assert normwise(E, ref)[0] < 1e-13


# ---- J = grad U by finite differences -----------------------------------------------------
def fd_jacobian(kernel, x, sX, sG, ssig, h=1e-5):
    """4th-order central differences of the oracle's U at points x (m,3)."""
    J = np.zeros((len(x), 9))
    for l in range(3):
        e = np.zeros(3)
        e[l] = h
        u = [O.uj_direct(kernel, x + s * e, sX, sG, ssig)[0] for s in (-2, -1, 1, 2)]
        d = (u[0] - 8 * u[1] + 8 * u[2] - u[3]) / (12 * h)
        J[:, 3 * l:3 * l + 3] = d
    return J


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
def test_J_is_grad_U(kernel):
    rng = np.random.default_rng(7)
    sX = rng.uniform(-1, 1, (60, 3))
    sG = rng.normal(0, 1, (60, 3))
    ssig = rng.uniform(0.2, 0.6, 60)
    x = rng.uniform(-1, 1, (12, 3))
    U, J = O.uj_direct(kernel, x, sX, sG, ssig)
    Jfd = fd_jacobian(kernel, x, sX, sG, ssig, h=1e-4 if kernel == "singular" else 1e-3)
    assert normwise(J, Jfd)[0] < 1e-8
    assert np.abs(J[:, 0] + J[:, 4] + J[:, 8]).max() < 1e-12 * np.abs(J).max()   # div U = 0


# ---- closed-form flows ------------------------------------------------------------------
def ring(n, R=1.0, circ=1.0):
    phi = 2 * np.pi * (np.arange(n) + 0.5) / n
    X = np.stack([R * np.cos(phi), R * np.sin(phi), np.zeros(n)], axis=1)
    T = np.stack([-np.sin(phi), np.cos(phi), np.zeros(n)], axis=1)
    return X, circ * (2 * np.pi * R / n) * T


@pytest.mark.parametrize("kernel", ["singular", "gaussianerf", "winckelmans"])
def test_ring_on_axis(kernel):
    """Circular ring (circulation 1, counter-clockwise about +z): on its axis
    u_z = R^2 / (2 (R^2 + z^2)^{3/2}) and u_x = u_y = 0 (Biot-Savart)."""
    R = 1.0
    X, G = ring(4000, R)
    sig = np.full(len(X), 0.01)
    z = np.array([0.0, 0.3, 1.0, 2.5])
    P = np.stack([np.zeros_like(z), np.zeros_like(z), z], axis=1)
    U, _ = O.uj_direct(kernel, P, X, G, sig)
    exact = R ** 2 / (2 * (R ** 2 + z ** 2) ** 1.5)
    tol = 1e-6 if kernel == "winckelmans" else 1e-9   # algebraic kernel: 1 - g ~ (sigma/r)^4
    np.testing.assert_allclose(U[:, 2], exact, rtol=tol)
    assert np.abs(U[:, :2]).max() < 1e-12


def test_straight_line():
    """Finite straight filament along +z (circulation 1, length L): at distance
    d from its midpoint u_y = (1/(4 pi d)) L / sqrt((L/2)^2 + d^2), u_x = u_z = 0."""
    n, L = 20000, 20.0
    z = -L / 2 + L * (np.arange(n) + 0.5) / n
    X = np.stack([np.zeros(n), np.zeros(n), z], axis=1)
    G = np.tile([0.0, 0.0, L / n], (n, 1))
    sig = np.full(n, 0.005)
    d = np.array([0.3, 1.0, 2.0])
    P = np.stack([d, np.zeros_like(d), np.zeros_like(d)], axis=1)
    U, _ = O.uj_direct("gaussianerf", P, X, G, sig)
    exact = L / (4 * np.pi * d * np.sqrt((L / 2) ** 2 + d ** 2))
    np.testing.assert_allclose(U[:, 1], exact, rtol=1e-7)
    assert np.abs(U[:, [0, 2]]).max() < 1e-12


def test_coincident_pair_skipped():
    X = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [0.5, 0.2, 0.3]])
    G = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    U, J = O.uj_direct("gaussianerf", X[:1], X[:2], G[:2], np.full(2, 0.1))
    assert np.all(U == 0) and np.all(J == 0)


# ---- stretching conventions --------------------------------------------------------------
def test_stretching_conventions_fd():
    """classic S = (Gamma . grad) U; transposed S_k = d(Gamma . U)/dx_k, both by
    central differences of the oracle's U."""
    rng = np.random.default_rng(3)
    sX = rng.uniform(-1, 1, (40, 3))
    sG = rng.normal(0, 1, (40, 3))
    ssig = rng.uniform(0.2, 0.5, 40)
    x = rng.uniform(-1, 1, 3)
    Gp = rng.normal(0, 1, 3)
    _, J = O.uj_direct("gaussianerf", x[None], sX, sG, ssig)

    def U_at(p):
        return O.uj_direct("gaussianerf", p[None], sX, sG, ssig)[0][0]

    def ddir(f, v, h=1e-3):   # 4th-order central difference of f along v
        return (f(x - 2 * h * v) - 8 * f(x - h * v) + 8 * f(x + h * v) - f(x + 2 * h * v)) / (12 * h)

    classic_fd = ddir(U_at, Gp)
    trans_fd = np.array([ddir(lambda p: Gp @ U_at(p), np.eye(3)[k]) for k in range(3)])
    Sc, St = O.stretching(False, J[0], Gp), O.stretching(True, J[0], Gp)
    assert np.linalg.norm(Sc - classic_fd) < 1e-8 * np.linalg.norm(classic_fd)
    assert np.linalg.norm(St - trans_fd) < 1e-8 * np.linalg.norm(trans_fd)
    assert np.linalg.norm(Sc - St) > 0.1 * np.linalg.norm(St)   # the two schemes differ


# ---- rVPM right-hand side ----------------------------------------------------------------
def test_rvpm_invariant():
    """rVPM (f=0, g=1/5), SFS off: d ln|Gamma|/dt + 2 d ln sigma/dt = 0, so
    |Gamma| sigma^2 is conserved; the part of dGamma/dt along Gamma is 2/5 of
    the stretching's and the normal part is the stretching's."""
    rng = np.random.default_rng(5)
    z0 = O.zeta0("gaussianerf")
    for _ in range(20):
        U, J, G = rng.normal(size=3), rng.normal(size=9), rng.normal(size=3)
        sig = rng.uniform(0.01, 0.2)
        for tr in (True, False):
            dX, dG, ds = O.rhs(tr, 0.0, 0.2, z0, [0.1, 0, 0], U, J, G, sig)
            gh = G / np.linalg.norm(G)
            S = O.stretching(tr, J, G)
            assert (gh @ dG) / np.linalg.norm(G) + 2 * ds / sig == pytest.approx(0, abs=1e-14)
            assert gh @ dG == pytest.approx(0.4 * (gh @ S), rel=1e-12)
            np.testing.assert_allclose(dG - (gh @ dG) * gh, S - (gh @ S) * gh, atol=1e-13)
            np.testing.assert_allclose(dX, U + [0.1, 0, 0], rtol=0, atol=0)
            # classic VPM: constant core, dGamma/dt = S
            dXc, dGc, dsc = O.rhs(tr, 0.0, 0.0, z0, [0, 0, 0], U, J, G, sig)
            assert dsc == 0.0
            np.testing.assert_allclose(dGc, S, rtol=1e-14)


def test_rhs_sfs_term():
    """With f = 0 the SFS enters only as -(C sigma^3 / zeta0) E in dGamma/dt."""
    rng = np.random.default_rng(9)
    z0 = O.zeta0("winckelmans")
    U, J, G, E = rng.normal(size=3), rng.normal(size=9), rng.normal(size=3), rng.normal(size=3)
    sig, C = 0.07, 0.3
    _, dG0, ds0 = O.rhs(True, 0.0, 0.2, z0, [0, 0, 0], U, J, G, sig)
    _, dG1, ds1 = O.rhs(True, 0.0, 0.2, z0, [0, 0, 0], U, J, G, sig, C=C, E=E)
    np.testing.assert_allclose(dG1 - dG0, -C * sig ** 3 / z0 * E, rtol=1e-12)
    assert ds1 == ds0


# ---- time integration ---------------------------------------------------------------------
@pytest.mark.parametrize("z", [-0.3 + 0j, -1 + 0.5j, 0.2 - 1.1j, -2.0 + 0j])
def test_lsrk3_linear_closed_form(z):
    y = O.lsrk3_linear(z, 1.0)
    assert abs(y - (1 + z + z ** 2 / 2 + z ** 3 / 6)) < 1e-15


def small_ring(seed=1):
    from flowunsteady_amd.workloads import vortex_ring
    return vortex_ring(n_stations=12, n_cross=10, sigma=0.12, jitter=1e-3, seed=seed)


def test_rk3_third_order():
    X, G, s, _ = small_ring()
    P = O.make_params("gaussianerf", True, "rVPM")
    T = 0.4

    def run(nsteps):
        F = O.Field(X, G, s)
        for _ in range(nsteps):
            O.rk3_step(P, F, T / nsteps)
        return np.concatenate([F.X.ravel(), F.G.ravel(), F.sigma])

    ref = run(64)
    e = [np.abs(run(n) - ref).max() for n in (4, 8, 16)]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 6.0 < r1 < 10.5 and 6.0 < r2 < 10.5, (e, r1, r2)


def test_rk3_freestream_translation_exact():
    F = O.Field([[0.1, 0.2, 0.3]], [[0.0, 0.0, 1.0]], [0.05])
    P = O.make_params("gaussianerf", True, "rVPM", Uinf=(1.0, -2.0, 0.5))
    O.rk3_step(P, F, 0.1)
    np.testing.assert_allclose(F.X[0], [0.2, 0.0, 0.35], rtol=0, atol=1e-15)
    np.testing.assert_array_equal(F.G[0], [0.0, 0.0, 1.0])
    assert F.sigma[0] == 0.05


def test_static_particles_do_not_move():
    X, G, s, _ = small_ring()
    st = np.zeros(len(X), np.uint8)
    st[::3] = 1
    F = O.Field(X, G, s, is_static=st)
    P = O.make_params("gaussianerf", True, "rVPM", Uinf=(0.3, 0, 0))
    O.rk3_step(P, F, 0.05)
    np.testing.assert_array_equal(F.X[st == 1], X[st == 1])
    np.testing.assert_array_equal(F.G[st == 1], G[st == 1])
    assert np.abs(F.X[st == 0] - X[st == 0]).max() > 1e-3


def test_ring_translates_along_axis():
    """Physics sanity (not a parity pin): a thick ring of circulation 1,
    R = 1, a = 0.1 moves along +z at roughly Gamma/(4 pi R) (ln(8R/a) - 1/4)."""
    X, G, s, _ = small_ring()
    U, _ = O.uj_direct("gaussianerf", X, X, G, s)
    uz = U[:, 2].mean()
    assert 0.15 < uz < 0.35
    assert np.abs(U[:, :2].mean(axis=0)).max() < 0.02 * uz


# ---- relaxation -----------------------------------------------------------------------------
def test_relaxation_identities():
    rng = np.random.default_rng(11)
    for _ in range(10):
        J, G = rng.normal(size=9), rng.normal(size=3)
        w = np.array([J[5] - J[7], J[6] - J[2], J[1] - J[3]])
        np.testing.assert_array_equal(O.relax(1, 0.0, J, G), G)
        np.testing.assert_allclose(O.relax(1, 1.0, J, G),
                                   np.linalg.norm(G) * w / np.linalg.norm(w), rtol=1e-14)
        for r in (0.1, 0.3, 0.7):
            Gc = O.relax(2, r, J, G)
            assert np.linalg.norm(Gc) == pytest.approx(np.linalg.norm(G), rel=1e-14)
            Gp = O.relax(1, r, J, G)
            np.testing.assert_allclose(Gc / np.linalg.norm(Gc), Gp / np.linalg.norm(Gp), rtol=1e-13)


def test_vorticity_mapping_is_curl_of_U():
    """The omega the relaxation and the enstrophy read from J is curl U
    (central differences of the oracle's U)."""
    rng = np.random.default_rng(2)
    sX = rng.uniform(-1, 1, (30, 3))
    sG = rng.normal(0, 1, (30, 3))
    ssig = rng.uniform(0.3, 0.6, 30)
    x = rng.uniform(-1, 1, (5, 3))
    _, J = O.uj_direct("gaussianerf", x, sX, sG, ssig)
    Jfd = fd_jacobian("gaussianerf", x, sX, sG, ssig)

    def d(Jm, k, l):   # dU_k/dx_l
        return Jm[:, k + 3 * l]

    curl = np.stack([d(Jfd, 2, 1) - d(Jfd, 1, 2), d(Jfd, 0, 2) - d(Jfd, 2, 0),
                     d(Jfd, 1, 0) - d(Jfd, 0, 1)], axis=1)
    w = np.stack([J[:, 5] - J[:, 7], J[:, 6] - J[:, 2], J[:, 1] - J[:, 3]], axis=1)
    assert normwise(w, curl)[0] < 1e-7
    ens = O.enstrophy(np.ones((5, 3)), J)
    assert ens == pytest.approx(0.5 * w.sum(), rel=1e-13)


def test_energy_definition():
    X = np.array([[1.0, 0, 0]])
    G = np.array([[0.0, 1.0, 0]])
    U = np.array([[0.0, 0.0, 2.0]])
    assert O.energy(X, G, U) == pytest.approx(2.0)   # U . (x x Gamma) = 2 * 1


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "singular"])
def test_ffv_direct_through_the_uj_oracle(kernel):
    """_Ffv_direct (src/FLOWUnsteady_processing_force.jl:879-929) by two other
    routes through the pinned UJ oracle: Ffv_b = sum_f U_b(x_f) x Gamma_f with
    U_b the velocity of bound particle b alone; Fbxf_b = the velocity at x_b of
    the free particles carrying strengths Gamma_b x Gamma_f and core sigma_b."""
    rng = np.random.default_rng(21)
    nb, nf = 7, 150
    bX = rng.uniform(-1, 1, (nb, 3)); bG = rng.normal(0, 1, (nb, 3))
    bs = rng.uniform(0.05, 0.3, nb)
    fX = rng.uniform(-1, 1, (nf, 3)); fG = rng.normal(0, 1, (nf, 3))
    fX[3] = bX[2]                     # a coincident pair is skipped
    out = O.ffv_direct(kernel, bX, bG, bs, fX, fG)
    for b in range(nb):
        Ub, _ = O.uj_direct(kernel, fX, bX[b:b + 1], bG[b:b + 1], bs[b:b + 1])
        np.testing.assert_allclose(out[b, :3], np.cross(Ub, fG).sum(axis=0), rtol=1e-12,
                                   atol=1e-12 * np.abs(out[b, :3]).max())
        Gp = np.cross(bG[b], fG)
        Uat, _ = O.uj_direct(kernel, bX[b:b + 1], fX, Gp, np.full(nf, bs[b]))
        np.testing.assert_allclose(out[b, 3:], Uat[0], rtol=1e-12,
                                   atol=1e-12 * np.abs(out[b, 3:]).max())

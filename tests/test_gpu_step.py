"""GPU parity of the time integrators (vpmx_rk3_step / vpmx_euler_step through
the C ABI) against the FP64 oracle's rungekutta3 / euler, including SFS,
relaxation, static particles, the classic formulation and the diagnostics."""
import numpy as np
import pytest

from flowunsteady_amd import vpm, workloads
from helpers import make_field, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def ring(n_stations=20, n_cross=100, sigma=0.08):
    return workloads.vortex_ring(n_stations=n_stations, n_cross=n_cross, sigma=sigma)


def state(pf):
    return vpm.get_X(pf).copy(), vpm.get_Gamma(pf).copy(), vpm.get_sigma(pf).copy()


def compare(pf, F, tol):
    X, G, s = state(pf)
    ex, eg, es = normwise(X, F.X), normwise(G, F.G), normwise(s[:, None], F.sigma[:, None])
    print("X", ex, "Gamma", eg, "sigma", es)
    assert ex[0] < tol and eg[0] < tol and es[0] < tol


CASES = {
    "rvpm_pedrizzetti": dict(formulation="rVPM", relax=("pedrizzetti", 1, 0.3), sfs=None),
    "rvpm_corrected_sfs": dict(formulation="rVPM", relax=("correctedpedrizzetti", 2, 0.3),
                               sfs=0.15),
    "cvpm_norelax": dict(formulation="cVPM", relax=("norelaxation", 0, 0.0), sfs=None),
}


def gpu_field(X, G, s, case, kernel="gaussianerf", precision="fp64", static=None,
              Uinf=(0.0, 0.0, 0.0)):
    c = CASES[case]
    rname = c["relax"][0]
    pf = make_field(X, G, s, kernel=kernel, precision=precision, static=static,
                    formulation=getattr(vpm, c["formulation"]),
                    relaxation=getattr(vpm, rname),
                    SFS=vpm.noSFS if c["sfs"] is None else vpm.ConstantSFS(c["sfs"]),
                    Uinf=lambda t: np.array(Uinf))
    return pf


def oracle_field(X, G, s, case, kernel="gaussianerf", static=None, Uinf=(0.0, 0.0, 0.0)):
    c = CASES[case]
    P = O.make_params(kernel, True, c["formulation"], Uinf=Uinf,
                      sfs=c["sfs"] is not None, Cs=c["sfs"] or 0.0, clip_backscatter=True,
                      relax_kind=c["relax"][1], rlxf=c["relax"][2])
    F = O.Field(X, G, s, is_static=None if static is None else static.astype(np.uint8))
    return P, F


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_rk3_vs_oracle(case, kernel):
    X, G, s, _ = ring()
    pf = gpu_field(X, G, s, case, kernel=kernel)
    P, F = oracle_field(X, G, s, case, kernel=kernel)
    dt = 0.01
    for i in range(1, 4):
        relax = vpm.relax_due(pf, i)
        vpm.nextstep(pf, dt, relax=relax)
        O.rk3_step(P, F, dt, relax)
    compare(pf, F, 1e-12)
    # U, J left in the matrix are those of the last UJ evaluation
    assert normwise(vpm.get_U(pf), F.U)[0] < 1e-12
    assert normwise(vpm.get_J(pf), F.J)[0] < 1e-12
    if CASES[case]["sfs"] is not None:
        assert normwise(vpm.get_SFS(pf), F.E)[0] < 1e-11
        np.testing.assert_array_equal(vpm.get_C(pf), F.C[:, 0])


@pytest.mark.parametrize("case", ["rvpm_pedrizzetti", "rvpm_corrected_sfs"])
def test_euler_vs_oracle(case):
    X, G, s, _ = ring()
    pf = gpu_field(X, G, s, case)
    pf.integration = vpm.euler_gpu
    P, F = oracle_field(X, G, s, case)
    for i in range(1, 4):
        vpm.nextstep(pf, 0.01, relax=True)
        O.euler_step(P, F, 0.01, True)
    compare(pf, F, 1e-12)


def test_static_particles_and_freestream():
    X, G, s, _ = ring()
    st = np.zeros(len(X))
    st[::4] = 1.0
    pf = gpu_field(X, G, s, "rvpm_pedrizzetti", static=st, Uinf=(0.5, 0.0, -0.2))
    P, F = oracle_field(X, G, s, "rvpm_pedrizzetti", static=st, Uinf=(0.5, 0.0, -0.2))
    vpm.nextstep(pf, 0.02, relax=True)
    O.rk3_step(P, F, 0.02, True)
    compare(pf, F, 1e-12)
    Xg, Gg, _ = state(pf)
    np.testing.assert_array_equal(Xg[st == 1], X[st == 1])
    np.testing.assert_array_equal(Gg[st == 1], G[st == 1])


def test_mixed_rk3_close_to_fp64():
    X, G, s, _ = workloads.vortex_ring()
    a = gpu_field(X, G, s, "rvpm_pedrizzetti", precision="fp64")
    b = gpu_field(X, G, s, "rvpm_pedrizzetti", precision="mixed")
    for i in range(1, 3):
        vpm.nextstep(a, 0.01, relax=True)
        vpm.nextstep(b, 0.01, relax=True)
    for get in (vpm.get_X, vpm.get_Gamma):
        assert normwise(get(b), get(a))[0] < 1e-6
    # dGamma/dt of the rVPM equations from each path's U, J on the same state
    z0 = O.zeta0("gaussianerf")
    Xa, Ga, sa = state(a)
    c = make_field(Xa, Ga, sa, precision="mixed")
    vpm.UJ_gpu(a)
    vpm.UJ_gpu(c)

    def dgdt(pf):
        J = vpm.get_J(pf)
        return np.array([O.rhs(True, 0.0, 0.2, z0, np.zeros(3), np.zeros(3), J[p], Ga[p],
                               sa[p])[1] for p in range(len(Ga))])

    e = normwise(dgdt(c), dgdt(a))
    print("dGamma/dt mixed vs fp64", e)
    # a ring's stretching S = J^T Gamma is small next to |J| |Gamma| (J is
    # mostly rotation), so the ~1e-6 relative J error shows up ~10x larger
    # relative to max |dGamma/dt|
    assert e[0] < 3e-5 and e[1] < 1e-5


def test_ring_translation_speed():
    """Physics sanity: the C1 ring keeps moving along +z and its |Gamma| sigma^2
    stays nearly constant under rVPM without relaxation."""
    X, G, s, _ = workloads.vortex_ring()
    pf = gpu_field(X, G, s, "cvpm_norelax")
    pf.formulation = vpm.rVPM
    z0 = X[:, 2].mean()
    inv0 = np.linalg.norm(G, axis=1) * s ** 2
    for i in range(1, 6):
        vpm.nextstep(pf, 0.01)
    Xn, Gn, sn = state(pf)
    speed = (Xn[:, 2].mean() - z0) / 0.05
    assert 0.15 < speed < 0.35
    inv = np.linalg.norm(Gn, axis=1) * sn ** 2
    assert np.abs(inv / inv0 - 1).max() < 1e-4


def test_resident_diagnostics():
    import ctypes
    from flowunsteady_amd import _native as N
    X, G, s, _ = ring()
    pf = gpu_field(X, G, s, "rvpm_pedrizzetti")
    ctx = vpm._context(0)
    L = N.lib()
    ctx.check(L.vpmx_upload(ctx.ptr, pf._ptr(), pf.np, ctypes.byref(pf.layout)))
    assert L.vpmx_resident_np(ctx.ptr) == pf.np
    ctx.check(L.vpmx_uj_resident(ctx.ptr, ctypes.byref(pf._uj_params())))
    ens, en = ctypes.c_double(), ctypes.c_double()
    ctx.check(L.vpmx_diagnostics(ctx.ptr, ctypes.byref(ens), ctypes.byref(en)))
    U, J = O.uj_direct("gaussianerf", X, X, G, s)
    assert ens.value == pytest.approx(O.enstrophy(G, J), rel=1e-12)
    assert en.value == pytest.approx(O.energy(X, G, U), rel=1e-12)
    ctx.check(L.vpmx_download(ctx.ptr, pf._ptr(), pf.np, ctypes.byref(pf.layout)))
    assert vpm.monitor_enstrophy(pf) == pytest.approx(ens.value, rel=1e-12)


def test_driver_loop_like_run_simulation():
    """The run_simulation pattern (src/FLOWUnsteady_simulation.jl:339-365):
    append static particles, nextstep, remove them again; then evaluate probes
    as Vvpm_on_Xs does (:494-570)."""
    X, G, s, _ = ring(n_stations=16, n_cross=50)
    pf = gpu_field(X, G, s, "rvpm_pedrizzetti")
    pf.particles = np.asfortranarray(np.pad(pf.particles, ((0, 0), (0, 200))))
    pf.maxparticles += 200
    P, F = oracle_field(X, G, s, "rvpm_pedrizzetti")
    rng = np.random.default_rng(0)
    for i in range(1, 3):
        org_np = vpm.get_np(pf)
        SX = rng.uniform(-1, 1, (40, 3))
        SG = rng.normal(0, 1e-3, (40, 3))
        for k in range(40):
            vpm.add_particle(pf, SX[k], SG[k], 0.05, static=True)
        vpm.nextstep(pf, 0.01, relax=vpm.relax_due(pf, i))
        for p in range(vpm.get_np(pf) - 1, org_np - 1, -1):
            vpm.remove_particle(pf, p)
        # oracle: same static particles appended for the step
        Fs = O.Field(np.concatenate([F.X, SX]), np.concatenate([F.G, SG]),
                     np.concatenate([F.sigma, np.full(40, 0.05)]),
                     is_static=np.concatenate([np.zeros(len(F.X)), np.ones(40)]).astype(np.uint8))
        O.rk3_step(P, Fs, 0.01, vpm.relax_due(pf, i))
        F = O.Field(Fs.X[:len(X)], Fs.G[:len(X)], Fs.sigma[:len(X)])
    compare(pf, F, 1e-12)
    # probes
    org = vpm.get_np(pf)
    probes = rng.uniform(-1.2, 1.2, (25, 3))
    for p in probes:
        vpm.add_particle(pf, p, np.zeros(3), 1e-6, vol=0.0)
    pf.UJ(pf)
    Up = vpm.get_U(pf)[org:].copy()
    for p in range(vpm.get_np(pf) - 1, org - 1, -1):
        vpm.remove_particle(pf, p)
    Ur, _ = O.uj_direct("gaussianerf", probes, F.X, F.G, F.sigma)
    assert normwise(Up, Ur)[0] < 1e-12


def test_c1_ring_twenty_steps_vs_oracle():
    """The 8,000-particle vortex ring (R = 1, a = 0.1, 40 stations x 200 per
    cross-section, sigma = 0.08, total circulation 1): gaussian-erf, rVPM,
    transposed stretching, Pedrizzetti relaxation every step, 20 RK3 steps of
    dt = 0.01 -- FP64 GPU against the oracle per particle.  Also the physics
    the ring must show: it translates along +z at about the thin-core speed
    Gamma/(4 pi R)(ln(8R/a) - 1/4) (Kelvin; the smoothing makes the core
    thicker, so slower) and |Gamma| sigma^2 changes only through relaxation."""
    X, G, s, v = workloads.vortex_ring()
    assert len(X) == 8000
    pf = make_field(X, G, s, vol=v, relaxation=vpm.pedrizzetti)
    P = O.make_params("gaussianerf", True, "rVPM", relax_kind=1, rlxf=0.3)
    F = O.Field(X, G, s)
    dt, nsteps = 0.01, 20
    z0 = X[:, 2].mean()
    inv0 = np.linalg.norm(G, axis=1) * s ** 2
    for i in range(nsteps):
        vpm.nextstep(pf, dt, relax=True)
        O.rk3_step(P, F, dt, True)
    compare(pf, F, 1e-12)
    Xg, Gg, sg = state(pf)
    speed = (Xg[:, 2].mean() - z0) / (nsteps * dt)
    kelvin = 1.0 / (4 * np.pi) * (np.log(8 / 0.1) - 0.25)
    drift = np.abs(np.linalg.norm(Gg, axis=1) * sg ** 2 / inv0 - 1.0).max()
    print(f"ring speed {speed:.5f} (thin-core {kelvin:.5f}), |Gamma| sigma^2 drift {drift:.2e}")
    assert 0.5 * kelvin < speed < 1.05 * kelvin     # oracle: 0.274 against 0.329
    assert drift < 1e-4                             # oracle: 3.7e-6

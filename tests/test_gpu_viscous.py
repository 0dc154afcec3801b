"""GPU parity of the viscous schemes (PSE and core spreading with its RBF
re-fit) against the oracle's restatement of the published methods."""
import ctypes

import numpy as np
import pytest

from flowunsteady_amd import _native as N
from flowunsteady_amd import vpm, workloads
from helpers import make_field, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def blob(n=1500, seed=21):
    X, G, s, v = workloads.gaussian_blob(n=n, seed=seed, std=0.25)
    return X, G, s * 2, (s * 2) ** 3


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian"])
def test_pse_rk3_vs_oracle(kernel):
    X, G, s, v = blob()
    visc = vpm.ParticleStrengthExchange(0.02)
    pf = make_field(X, G, s, vol=v, kernel=kernel, viscous=visc, relaxation=vpm.pedrizzetti)
    P = O.make_params(kernel, viscous="pse", nu=0.02, relax_kind=1, rlxf=0.3)
    F = O.Field(X, G, s, vol=v)
    for i in range(1, 3):
        vpm.nextstep(pf, 0.002, relax=True)
        O.rk3_step(P, F, 0.002, True)
    assert normwise(vpm.get_X(pf), F.X)[0] < 1e-12
    assert normwise(vpm.get_Gamma(pf), F.G)[0] < 1e-12


def test_pse_is_the_difference_of_viscous_and_inviscid_steps():
    X, G, s, v = blob(800, seed=3)
    a = make_field(X, G, s, vol=v, viscous=vpm.ParticleStrengthExchange(0.05))
    b = make_field(X, G, s, vol=v)
    a.integration = b.integration = vpm.euler_gpu
    vpm.nextstep(a, 1e-3)
    vpm.nextstep(b, 1e-3)
    d = vpm.get_Gamma(a) - vpm.get_Gamma(b)
    ref = 1e-3 * O.pse("gaussianerf", 0.05, X, G, s, v)
    assert normwise(d, ref)[0] < 1e-10
    # the PSE rate conserves the total circulation
    assert np.abs(d.sum(axis=0)).max() < 1e-12 * np.abs(d).max() * len(d)


def test_core_spreading_growth_vs_oracle():
    X, G, s, v = blob(1000, seed=5)
    cs = vpm.CoreSpreading(0.01, sgm0=0.05, beta=100.0)
    pf = make_field(X, G, s, vol=v, viscous=cs)
    P = O.make_params("gaussianerf", viscous="corespreading", nu=0.01, sgm0=0.05, beta=100.0)
    F = O.Field(X, G, s, vol=v)
    for _ in range(3):
        vpm.nextstep(pf, 0.003)
        O.rk3_step(P, F, 0.003)
    assert normwise(vpm.get_sigma(pf)[:, None], F.sigma[:, None])[0] < 1e-12
    assert normwise(vpm.get_Gamma(pf), F.G)[0] < 1e-12


def lattice_field(h=0.05, half=5, seed=7, nstatic=6):
    rng = np.random.default_rng(seed)
    a = np.arange(-half, half + 1) * h
    Z, Y, Xg = np.meshgrid(a, a, a, indexing="ij")
    X = np.stack([Xg.ravel(), Y.ravel(), Z.ravel()], axis=1) + rng.uniform(-0.1 * h, 0.1 * h,
                                                                           (len(a) ** 3, 3))
    G = np.exp(-np.sum(X ** 2, axis=1) / 0.02)[:, None] * np.array([0.2, -0.1, 1.0]) * h ** 3
    st = np.zeros(len(X))
    st[:nstatic] = 1
    return X, G, st


def test_core_spreading_reset_vs_oracle():
    """A step whose growth crosses beta * sgm0: the GPU re-fit agrees with the
    oracle's to the CG tolerance, static particles untouched."""
    X, G, st = lattice_field()
    s = np.full(len(X), 0.0555)
    cs = vpm.CoreSpreading(0.05, sgm0=0.04, beta=1.4, itmax=400, tol=1e-10)
    pf = make_field(X, G, s, static=st, viscous=cs)
    vpm.nextstep(pf, 0.01)
    P = O.make_params("gaussianerf", viscous="corespreading", nu=0.05, sgm0=0.04, beta=1.4,
                      itmax=400, tol=1e-10)
    F = O.Field(X, G, s, is_static=st.astype(np.uint8))
    O.rk3_step(P, F, 0.01)
    free = st == 0
    assert np.all(vpm.get_sigma(pf)[free] == 0.04)
    np.testing.assert_array_equal(vpm.get_sigma(pf)[~free], s[~free])
    np.testing.assert_array_equal(vpm.get_Gamma(pf)[~free], G[~free])
    eg = normwise(vpm.get_Gamma(pf), F.G)
    print("reset Gamma vs oracle", eg)
    assert eg[0] < 1e-7


def test_rbf_conjugategradient_keeps_the_vorticity():
    X, G, st = lattice_field(nstatic=0)
    s = np.full(len(X), 0.06)
    pf = make_field(X, G, s, viscous=vpm.CoreSpreading(0.0, sgm0=0.04, itmax=500, tol=1e-8))
    w0 = O.zeta_sum("gaussianerf", X, X, G, s)
    did, it, res = vpm.rbf_conjugategradient(pf)
    assert did and 0 < it < 500 and max(res) <= 1e-8
    w1 = O.zeta_sum("gaussianerf", X, X, vpm.get_Gamma(pf), s0=0.04)
    # the residual the solver reports is the one the field has
    for k in range(3):
        assert np.linalg.norm(w1[:, k] - w0[:, k]) / np.linalg.norm(w0[:, k]) <= 1.01 * res[k] + 1e-13
    np.testing.assert_array_equal(vpm.get_sigma(pf), 0.04)


def test_viscous_errors():
    X, G, s, v = blob(50)
    ctx = vpm._context(0)
    pf = make_field(X, G, s, kernel="winckelmans", viscous=vpm.CoreSpreading(0.01, 0.05))
    with pytest.raises(N.VPMXError) as e:
        vpm.nextstep(pf, 0.01)
    assert e.value.code == N.EINVAL and "gaussian-erf" in str(e.value)
    pf = make_field(X, G, s, viscous=vpm.CoreSpreading(0.01, 0.0))
    with pytest.raises(N.VPMXError):
        vpm.nextstep(pf, 0.01)
    a = make_field(X, G, s, viscous=vpm.ParticleStrengthExchange(0.01))
    b = make_field(X, G, s, viscous=vpm.ParticleStrengthExchange(0.01))
    with pytest.raises(N.VPMXError) as e:
        vpm.nextstep_batch([a, b], 0.01)
    assert e.value.code == N.EUNSUPPORTED


FMM_VISC = dict(p=8, ncrit=50, theta=0.4, nonzero_sigma=True)   # core-aware acceptance


@pytest.mark.parametrize("kernel,sigcut", [("gaussianerf", 7.0), ("gaussian", 3.0)])
def test_pse_fmm_rk3_vs_oracle(kernel, sigcut):
    """PSE in an FMM time step runs over the near lists of each stage's FMM
    evaluation: GPU FP64 against the oracle's FMM step with the same tree and
    lists (PSE over its near lists), per particle."""
    X, G, s, v = blob(4000, seed=23)
    pf = make_field(X, G, s, vol=v, kernel=kernel, viscous=vpm.ParticleStrengthExchange(0.02),
                    relaxation=vpm.pedrizzetti, fmm=vpm.FMM(**FMM_VISC))
    pf.UJ = vpm.UJ_fmm_gpu
    P = O.make_params(kernel, viscous="pse", nu=0.02, relax_kind=1, rlxf=0.3, fmm=True, p=8,
                      ncrit=50, theta=0.4, sigcut=sigcut)
    F = O.Field(X, G, s, vol=v)
    for i in range(2):
        vpm.nextstep(pf, 0.002, relax=True)
        O.rk3_step(P, F, 0.002, True)
    for get, ref in ((vpm.get_X, F.X), (vpm.get_Gamma, F.G)):
        e = normwise(get(pf), ref)
        print(kernel, get.__name__, e)
        assert e[0] < 1e-12


def test_core_spreading_fmm_reset_vs_oracle():
    """A step with the FMM whose core growth crosses beta * sgm0: the re-fit's
    vorticity sums run over the near lists of a tree built at the step's
    final positions (zeta_fmm); against the oracle's same procedure, static
    particles untouched."""
    X, G, st = lattice_field()
    s = np.full(len(X), 0.0555)
    cs = vpm.CoreSpreading(0.05, sgm0=0.04, zeta=vpm.zeta_fmm, beta=1.4, itmax=400, tol=1e-10)
    fmm = vpm.FMM(p=8, ncrit=20, theta=0.4, nonzero_sigma=True)
    pf = make_field(X, G, s, static=st, viscous=cs, fmm=fmm)
    pf.UJ = vpm.UJ_fmm_gpu
    vpm.nextstep(pf, 0.01)
    P = O.make_params("gaussianerf", viscous="corespreading", nu=0.05, sgm0=0.04, beta=1.4,
                      itmax=400, tol=1e-10, fmm=True, p=8, ncrit=20, theta=0.4, sigcut=7.0)
    F = O.Field(X, G, s, is_static=st.astype(np.uint8))
    O.rk3_step(P, F, 0.01)
    free = st == 0
    assert np.all(vpm.get_sigma(pf)[free] == 0.04)
    np.testing.assert_array_equal(vpm.get_Gamma(pf)[~free], G[~free])
    eg = normwise(vpm.get_Gamma(pf), F.G)
    print("fmm reset Gamma vs oracle", eg)
    assert eg[0] < 1e-7


def test_rbf_conjugategradient_zeta_fmm_matches_zeta_direct():
    """rbf_conjugategradient with zeta_fmm (near lists at the core-aware
    acceptance) and with zeta_direct fit the same strengths: the far pairs the
    FMM leaves out are beyond 7 cores."""
    X, G, st = lattice_field(nstatic=0)
    s = np.full(len(X), 0.06)
    out = {}
    for z in (vpm.zeta_direct, vpm.zeta_fmm):
        pf = make_field(X, G, s, fmm=vpm.FMM(p=4, ncrit=16, theta=0.4, nonzero_sigma=True),
                        viscous=vpm.CoreSpreading(0.0, sgm0=0.04, zeta=z, itmax=500, tol=1e-11))
        did, it, res = vpm.rbf_conjugategradient(pf)
        assert did and max(res) <= 1e-11
        out[z.name] = vpm.get_Gamma(pf).copy()
    e = normwise(out["zeta_fmm"], out["zeta_direct"])
    print("zeta_fmm vs zeta_direct Gamma", e)
    assert e[0] < 1e-8

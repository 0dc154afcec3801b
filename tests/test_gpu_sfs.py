"""GPU parity of the SFS paths: the dynamic procedure (direct UJ) against the
oracle's, and the SFS stretching over the FMM's near-field lists against the
direct oracle."""
import numpy as np
import pytest

from flowunsteady_amd import vpm, workloads
from helpers import make_field, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def blob(n=3000, seed=4):
    X, G, s, _ = workloads.gaussian_blob(n=n, seed=seed, std=0.3)
    return X, G, s * 3


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_dynamic_sfs_rk3_vs_oracle(kernel):
    """Direct FP64 RK3 with the dynamic procedure and Pedrizzetti relaxation
    against the oracle: state, coefficient and its averaged numerator and
    denominator per particle."""
    X, G, s = blob()
    sfs = vpm.DynamicSFS(alpha=0.999, rlxf=0.005, minC=0.0, maxC=1.0)
    pf = make_field(X, G, s, kernel=kernel, SFS=sfs, relaxation=vpm.pedrizzetti)
    P = O.make_params(kernel, True, "rVPM", sfs="dynamic", alpha=0.999, sfs_rlxf=0.005,
                      minC=0.0, maxC=1.0, relax_kind=1, rlxf=0.3)
    F = O.Field(X, G, s)
    for i in range(1, 3):
        vpm.nextstep(pf, 0.002, relax=True)
        O.rk3_step(P, F, 0.002, True)
    for get, ref in ((vpm.get_X, F.X), (vpm.get_Gamma, F.G), (vpm.get_sigma, F.sigma)):
        assert normwise(get(pf), ref)[0] < 1e-12
    C = pf.particles[vpm.C_INDEX, :pf.np].T
    assert np.abs(C[:, 0] - F.C[:, 0]).max() <= 1e-12
    for k in (1, 2):
        assert np.abs(C[:, k] - F.C[:, k]).max() <= 1e-12 * np.abs(F.C[:, k]).max(), k
    assert np.any(C[:, 0] > 0)


def _blob_c4(n, seed=4):
    """The C4 blob's distribution at n particles with the 2M blob's particle
    density (std scaled by (n / 2^21)^(1/3)): swirl Gamma, lognormal sigma."""
    std = 0.5 * (n / float(1 << 21)) ** (1.0 / 3.0)
    X, G, s, _ = workloads.gaussian_blob(n=n, seed=seed, std=std, swirl=0.01 * 0.5 / std)
    return X, G, s


C4_FMM = dict(p=8, ncrit=50, theta=0.4)   # FLOWUnsteady's default geometric acceptance


def test_fmm_dynamic_sfs_rk3_vs_oracle():
    """The full rVPM step of the rotor-hover high-fidelity setting on a 20k
    blob: FMM p = 8 (geometric acceptance), dynamic SFS, Pedrizzetti
    relaxation, RK3 -- FP64 GPU against the oracle's CPU FMM step (same tree,
    lists and near-field sums), per particle."""
    X, G, s = _blob_c4(20000)
    pf = make_field(X, G, s, SFS=vpm.DynamicSFS(), relaxation=vpm.pedrizzetti,
                    fmm=vpm.FMM(**C4_FMM))
    pf.UJ = vpm.UJ_fmm_gpu
    P = O.make_params("gaussianerf", True, "rVPM", sfs="dynamic", relax_kind=1, rlxf=0.3,
                      fmm=True, sigcut=0.0, **C4_FMM)
    F = O.Field(X, G, s)
    dt = 1e-3
    for i in range(2):
        vpm.nextstep(pf, dt, relax=True)
        O.rk3_step(P, F, dt, True)
        print("step", i, flush=True)
    st = vpm._context(0).fmm_stats()
    print("fmm stats", st)
    for get, ref in ((vpm.get_X, F.X), (vpm.get_Gamma, F.G), (vpm.get_sigma, F.sigma)):
        e = normwise(get(pf), ref)
        print(get.__name__, e)
        assert e[0] < 1e-12
    C = pf.particles[vpm.C_INDEX, :pf.np].T
    print("C range", F.C[:, 0].min(), F.C[:, 0].max(), "clipped", np.mean(F.C[:, 0] == 0))
    assert np.abs(C[:, 0] - F.C[:, 0]).max() <= 1e-11
    for k in (1, 2):
        assert np.abs(C[:, k] - F.C[:, k]).max() <= 1e-12 * np.abs(F.C[:, k]).max(), k
    assert np.any((C[:, 0] > 0) & (C[:, 0] < 1))


def test_estr_fmm_vs_direct():
    X, G, s, _ = workloads.rotor_wake(n=1 << 14, seed=3)
    pf = make_field(X, G, s, fmm=vpm.FMM(p=10, ncrit=50, theta=0.4, nonzero_sigma=True))
    import ctypes
    from flowunsteady_amd import _native as N
    P = pf._uj_params(sfs=True, fmm=True)
    ctx = vpm._context(0)
    ctx.check(N.lib().vpmx_uj(ctx.ptr, pf._ptr(), pf.np, ctypes.byref(pf.layout), ctypes.byref(P)))
    U, J = O.uj_direct("gaussianerf", X, X, G, s)
    E = O.estr_direct("gaussianerf", True, X, J, X, G, s, J)
    e = normwise(vpm.get_SFS(pf), E)
    print("Estr fmm vs direct", e)
    assert e[0] < 1e-5


def test_fmm_rk3_with_constant_sfs():
    X, G, s = blob(4000)
    pf = make_field(X, G, s, SFS=vpm.ConstantSFS(0.2), relaxation=vpm.pedrizzetti,
                    fmm=vpm.FMM(p=10, ncrit=50, theta=0.4, nonzero_sigma=True))
    pf.UJ = vpm.UJ_fmm_gpu
    P = O.make_params("gaussianerf", True, "rVPM", sfs=True, Cs=0.2, relax_kind=1, rlxf=0.3)
    F = O.Field(X, G, s)
    for i in range(1, 3):
        vpm.nextstep(pf, 0.002, relax=True)
        O.rk3_step(P, F, 0.002, True)
    ex, eg = normwise(vpm.get_X(pf), F.X), normwise(vpm.get_Gamma(pf), F.G)
    print("fmm rk3 + sfs vs direct oracle", ex, eg)
    assert ex[0] < 1e-7 and eg[0] < 1e-5


@pytest.mark.parametrize("fmm", [False, True])
def test_dynamic_sfs_mixed_per_particle(fmm):
    """Mixed against FP64 per particle after one RK3 step with the dynamic
    procedure: the averaged numerator and denominator (FP32 pair math in the
    core-size derivative sums, FP64 accumulation) normwise, and the
    coefficient of every particle that neither run clipped or bounded."""
    X, G, s = blob(3000) if not fmm else _blob_c4(20000)
    kw = dict(SFS=vpm.DynamicSFS(), relaxation=vpm.pedrizzetti)
    if fmm:
        kw["fmm"] = vpm.FMM(**C4_FMM)
    a = make_field(X, G, s, precision="fp64", **kw)
    b = make_field(X, G, s, precision="mixed", **kw)
    for pf in (a, b):
        if fmm:
            pf.UJ = vpm.UJ_fmm_gpu
        vpm.nextstep(pf, 1e-3, relax=True)
    Ca = a.particles[vpm.C_INDEX, :a.np].T
    Cb = b.particles[vpm.C_INDEX, :b.np].T
    assert np.all(np.isfinite(Cb)) and np.all((Cb[:, 0] >= 0) & (Cb[:, 0] <= 1))
    en = normwise(Cb[:, 1:2], Ca[:, 1:2])
    ed = normwise(Cb[:, 2:3], Ca[:, 2:3])
    free = (Ca[:, 0] > 0) & (Ca[:, 0] < 1) & (Cb[:, 0] > 0) & (Cb[:, 0] < 1)
    dC = np.abs(Cb[free, 0] - Ca[free, 0])
    rel = dC / np.abs(Ca[free, 0])
    flips = np.mean(((Ca[:, 0] > 0) & (Ca[:, 0] < 1)) != ((Cb[:, 0] > 0) & (Cb[:, 0] < 1)))
    print(f"dynamic mixed vs fp64 (fmm={fmm}): num {en}, den {ed}, free {free.mean():.3f}, "
          f"|dC| max {dC.max():.2e} p99 {np.quantile(dC, 0.99):.2e}, rel max {rel.max():.2e} "
          f"p99 {np.quantile(rel, 0.99):.2e}, regime flips {flips:.4f}; "
          f"Gamma {normwise(vpm.get_Gamma(b), vpm.get_Gamma(a))}")
    # measured: num 6.5e-7 / 1.0e-6, den 1.9e-6 / 7.6e-7, |dC| <= 9.2e-5, no flips
    assert en[0] < 1e-5 and ed[0] < 1e-5
    assert dC.max() < 5e-4 and flips < 1e-3


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian"])
@pytest.mark.parametrize("transposed", [True, False])
def test_estr_fmm_mixed_vs_oracle(kernel, transposed):
    """The mixed SFS stretching over the near-field lists (FP32 pair math,
    FP64 accumulation) against the oracle's FP64 all-pairs Estr on the same J
    (the J the mixed UJ produced), so only the Estr pair precision and the
    neglected far field (below e^-24.5 zeta(0) per pair for the gaussian
    kernels) separate them."""
    import ctypes
    from flowunsteady_amd import _native as N
    X, G, s, _ = workloads.rotor_wake(n=1 << 14, seed=3)
    pf = make_field(X, G, s, kernel=kernel, precision="mixed", transposed=transposed,
                    fmm=vpm.FMM(p=8, ncrit=50, theta=0.4, nonzero_sigma=True))
    P = pf._uj_params(sfs=True, fmm=True)
    ctx = vpm._context(0)
    ctx.check(N.lib().vpmx_uj(ctx.ptr, pf._ptr(), pf.np, ctypes.byref(pf.layout), ctypes.byref(P)))
    J = vpm.get_J(pf)
    E = O.estr_direct(kernel, transposed, X, J, X, G, s, J)
    e = normwise(vpm.get_SFS(pf), E)
    print(kernel, transposed, "mixed Estr fmm vs oracle", e)
    assert e[0] < 2e-5 and e[1] < 5e-6

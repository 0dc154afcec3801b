"""GPU parity of the fast multipole UJ (vpmx_uj with fmm = 1).

  FP64 GPU FMM vs the oracle's CPU FMM with the same p, ncrit, theta and
  sigma cut: same tree and lists (checked through the counts), <= 1e-12.
  GPU FMM vs the direct oracle: the truncation error, falling with p.
  Mixed near field vs FP64 near field: <= 1e-6 (U), 3e-6 (J).
"""
import numpy as np
import pytest

from flowunsteady_amd import _native as N
from flowunsteady_amd import vpm, workloads
from helpers import make_field, normwise, sample_targets
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEOMS = {
    "cube": lambda: workloads.uniform_cube(n=20000, seed=2),
    "rotor": lambda: workloads.rotor_wake(n=1 << 15, seed=3),
    "blob": lambda: workloads.gaussian_blob(n=20000, seed=4),
}


def fmm_field(X, G, s, precision="fp64", kernel="gaussianerf", **fmm):
    fmm.setdefault("nonzero_sigma", True)
    pf = make_field(X, G, s, kernel=kernel, precision=precision, fmm=vpm.FMM(**fmm))
    pf.UJ = vpm.UJ_fmm_gpu
    return pf


def run(pf):
    pf.UJ(pf)
    return vpm.get_U(pf).copy(), vpm.get_J(pf).copy()


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("p", [4, 8])
def test_fmm_fp64_vs_oracle_fmm(geom, p):
    X, G, s, _ = GEOMS[geom]()
    pf = fmm_field(X, G, s, p=p, ncrit=50, theta=0.4, sigma_cut=7.0)
    U, J = run(pf)
    st = vpm._context(0).fmm_stats()
    Uo, Jo, so = O.uj_fmm("gaussianerf", X, G, s, p=p, ncrit=50, theta=0.4, sigcut=7.0)
    # the device-built tree and lists are the oracle's, cell for cell
    assert int(st["cells"]) == so["ncell"] and int(st["leaves"]) == so["nleaf"]
    assert int(st["m2l_pairs"]) == so["nm2l"] and int(st["p2p_pairs"]) == so["p2p_pairs"]
    eu, ej = normwise(U, Uo), normwise(J, Jo)
    print(geom, p, "gpu fmm vs cpu fmm", eu, ej, st)
    assert eu[0] < 1e-12 and ej[0] < 1e-12


@pytest.mark.parametrize("geom", ["rotor", "blob"])
def test_fmm_error_vs_direct_falls_with_p(geom):
    X, G, s, _ = GEOMS[geom]()
    idx = sample_targets(len(X), 1500, seed=5)
    Ur, Jr = O.uj_direct("gaussianerf", X[idx], X, G, s)
    errs = []
    for p in (2, 4, 6, 8, 10):
        U, J = run(fmm_field(X, G, s, p=p))
        errs.append((normwise(U[idx], Ur)[0], normwise(J[idx], Jr)[0]))
    print(geom, errs)
    for (u1, j1), (u2, j2) in zip(errs, errs[1:]):
        assert u2 < 0.6 * u1 and j2 < 0.6 * j1
    assert errs[3][0] < 3e-5 and errs[4][0] < 3e-6


@pytest.mark.parametrize("geom", ["rotor", "cube"])
def test_fmm_mixed_near_field(geom):
    X, G, s, _ = GEOMS[geom]()
    U64, J64 = run(fmm_field(X, G, s, precision="fp64", p=8))
    Umx, Jmx = run(fmm_field(X, G, s, precision="mixed", p=8))
    eu, ej = normwise(Umx, U64), normwise(Jmx, J64)
    print(geom, "mixed vs fp64 fmm", eu, ej)
    assert eu[0] < 1e-6 and ej[0] < 3e-6


@pytest.mark.parametrize("ncrit,theta,p", [(1, 0.4, 3), (7, 0.6, 5), (300, 0.3, 4), (50, 1e-9, 2),
                                           (50, 0.4, 0), (64, 0.4, 12), (20, 0.5, 1), (30, 0.4, 7),
                                           (40, 0.4, 9)])
def test_fmm_parameters(ncrit, theta, p):
    X, G, s, _ = workloads.gaussian_blob(n=3000, seed=9)
    U, J = run(fmm_field(X, G, s, p=p, ncrit=ncrit, theta=theta, sigma_cut=7.0))
    Uo, Jo, _ = O.uj_fmm("gaussianerf", X, G, s, p=p, ncrit=ncrit, theta=theta, sigcut=7.0)
    assert normwise(U, Uo)[0] < 1e-12 and normwise(J, Jo)[0] < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 65, 1000])
def test_fmm_small_and_degenerate(n):
    X, G, s, _ = workloads.uniform_cube(n=n, seed=n, smin=0.001, smax=0.002)
    U, J = run(fmm_field(X, G, s, p=6, ncrit=2))
    Ur, Jr = O.uj_direct("gaussianerf", X, X, G, s)
    if n == 1:
        assert np.all(U == 0) and np.all(J == 0)
    else:
        assert normwise(U, Ur)[0] < 1e-4 and normwise(J, Jr)[0] < 1e-4


def test_fmm_coincident_cluster():
    """All particles at one point except a few: keys equal, the tree bottoms
    out at the finest level, coincident pairs are skipped."""
    X = np.zeros((200, 3))
    X[:5] = np.random.default_rng(0).uniform(-1, 1, (5, 3))
    G = np.random.default_rng(1).normal(0, 1, (200, 3))
    s = np.full(200, 0.05)
    U, J = run(fmm_field(X, G, s, p=4, ncrit=10))
    Ur, Jr = O.uj_direct("gaussianerf", X, X, G, s)
    assert np.all(np.isfinite(U)) and normwise(U, Ur)[0] < 1e-6


def test_fmm_rk3_step_close_to_direct():
    X, G, s, _ = workloads.vortex_ring()
    a = make_field(X, G, s, relaxation=vpm.pedrizzetti)
    b = make_field(X, G, s, relaxation=vpm.pedrizzetti,
                   fmm=vpm.FMM(p=10, ncrit=50, theta=0.4, nonzero_sigma=True))
    b.UJ = vpm.UJ_fmm_gpu
    ctx = vpm._context(0)
    for pf in (a, b):
        P = pf._uj_params()
        S = pf._step_params()
        import ctypes
        ctx.check(N.lib().vpmx_rk3_step(ctx.ptr, pf._ptr(), pf.np, ctypes.byref(pf.layout),
                                        ctypes.byref(P), ctypes.byref(S), 0.01, 1))
    assert normwise(vpm.get_X(b), vpm.get_X(a))[0] < 1e-6
    assert normwise(vpm.get_Gamma(b), vpm.get_Gamma(a))[0] < 1e-5


def test_geometric_criterion_matches_the_oracle():
    """nonzero_sigma = False (FLOWUnsteady's default FMM): the purely geometric
    acceptance, same tree and lists as the oracle's FMM with sigcut = 0."""
    X, G, s, _ = GEOMS["blob"]()
    pf = fmm_field(X, G, s, p=6, ncrit=50, theta=0.4, nonzero_sigma=False)
    U, J = run(pf)
    st = vpm._context(0).fmm_stats()
    Uo, Jo, so = O.uj_fmm("gaussianerf", X, G, s, p=6, ncrit=50, theta=0.4, sigcut=0.0)
    assert int(st["m2l_pairs"]) == so["nm2l"]
    assert normwise(U, Uo)[0] < 1e-12 and normwise(J, Jo)[0] < 1e-12
    # the core-aware criterion only refuses acceptances: never fewer near pairs
    pf7 = fmm_field(X, G, s, p=6, ncrit=50, theta=0.4, nonzero_sigma=True)
    run(pf7)
    assert vpm._context(0).fmm_stats()["p2p_pairs"] >= st["p2p_pairs"]



def clustered(n=6000, nout=40, seed=1):
    """A dense cube of particles beside sparse outliers: with ncrit = 4 the
    outliers' large leaves see thousands of the cluster's leaves in their near
    field (484 list entries per leaf on average)."""
    rng = np.random.default_rng(seed)
    X = np.vstack([rng.uniform(-0.05, 0.05, (n, 3)) + 0.3, rng.uniform(-1, 1, (nout, 3))])
    G = rng.normal(0, 1e-3, X.shape)
    return X, G, np.full(len(X), 0.004)


def test_fmm_long_near_lists():
    """Near-field lists longer than one work item (512 source leaves) are cut
    into several workgroups whose FP64 partial sums are added in order: FP64
    against the oracle's CPU FMM, mixed against FP64, and the SFS stretching
    over the same lists (FP64 and mixed) against the oracle's all-pairs Estr."""
    import ctypes
    from flowunsteady_amd import _native as N
    X, G, s = clustered()
    Uo, Jo, so = O.uj_fmm("gaussianerf", X, G, s, p=4, ncrit=4, theta=0.4, sigcut=7.0)
    assert so["np2p"] > 400 * so["nleaf"]
    out = {}
    for prec in ("fp64", "mixed"):
        pf = make_field(X, G, s, precision=prec,
                        fmm=vpm.FMM(p=4, ncrit=4, theta=0.4, nonzero_sigma=True))
        P = pf._uj_params(sfs=True, fmm=True)
        ctx = vpm._context(0)
        ctx.check(N.lib().vpmx_uj(ctx.ptr, pf._ptr(), pf.np, ctypes.byref(pf.layout),
                                  ctypes.byref(P)))
        st = ctx.fmm_stats()
        assert int(st["p2p_pairs"]) == so["p2p_pairs"]
        U, J, E = vpm.get_U(pf).copy(), vpm.get_J(pf).copy(), vpm.get_SFS(pf).copy()
        Eo = O.estr_direct("gaussianerf", True, X, J, X, G, s, J)
        out[prec] = (U, J)
        ee = normwise(E, Eo)
        print(prec, "estr vs oracle", ee)
        assert ee[0] < (1e-9 if prec == "fp64" else 2e-5)
    eu, ej = normwise(out["fp64"][0], Uo), normwise(out["fp64"][1], Jo)
    print("fp64 vs oracle fmm", eu, ej)
    assert eu[0] < 1e-12 and ej[0] < 1e-12
    mu, mj = normwise(out["mixed"][0], out["fp64"][0]), normwise(out["mixed"][1], out["fp64"][1])
    print("mixed vs fp64", mu, mj)
    assert mu[0] < 5e-6 and mj[0] < 1e-5


@pytest.mark.parametrize("p", [4, 8, 10])
def test_fmm_m2l_operator_table_and_recurrence(p):
    """A dense blob with a sparse, distant shell under the geometric criterion:
    the blob's pairs take the matrix-core M2L's operator table, the pairs
    between blob cells and the coarse shell cells (level gaps and offsets past
    the table) the in-kernel recurrence; both against the oracle's FMM."""
    X, G, s, _ = workloads.gaussian_blob(n=20000, seed=6)
    rng = np.random.default_rng(7)
    d = rng.standard_normal((600, 3))
    Xs = 40.0 * d / np.linalg.norm(d, axis=1, keepdims=True)
    X = np.vstack([X, Xs])
    G = np.vstack([G, 1e-3 * rng.standard_normal((600, 3))])
    s = np.concatenate([s, np.full(600, 0.05)])
    pf = fmm_field(X, G, s, p=p, ncrit=50, theta=0.4, sigma_cut=0.0, nonzero_sigma=False)
    U, J = run(pf)
    st = vpm._context(0).fmm_stats()
    Uo, Jo, so = O.uj_fmm("gaussianerf", X, G, s, p=p, ncrit=50, theta=0.4, sigcut=0.0)
    assert int(st["m2l_pairs"]) == so["nm2l"]
    eu, ej = normwise(U, Uo), normwise(J, Jo)
    print("blob + shell", p, eu, ej, st)
    assert eu[0] < 1e-12 and ej[0] < 1e-12

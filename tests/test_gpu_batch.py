"""GPU parity of the batched fields (many independent fields in one launch),
the schedules and autotuner of the batched pairwise pass, and graph replay of
time steps.

  FP64 batch vs each field alone (vpmx_uj per field): <= 1e-14 at any schedule
  (a field alone small enough to leave the device idle splits its source range
  across workgroups, which only regroups the FP64 sums).
  Mixed batch at block 256 vs each field alone: <= 1e-13 (same Morton order,
  blocks and FP32 tiles; the FP64 folds regroup as above).  Other blocks:
  within the mixed tolerance of FP64.
  Batched RK3 with a freestream per field vs the oracle's rungekutta3 per field:
  <= 1e-12.  Graph replay vs eager launches: bitwise, every replay.
"""
import ctypes

import numpy as np
import pytest

from flowunsteady_amd import _native as N
from flowunsteady_amd import vpm, workloads
from helpers import make_field, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCHEDULES = [(N.SCHED_FLAT, 64), (N.SCHED_QUEUE, 64), (N.SCHED_FLAT, 128),
             (N.SCHED_QUEUE, 128), (N.SCHED_FLAT, 256), (N.SCHED_QUEUE, 256)]


def sweep(nfields=6, seed=11, nmin=40, nmax=2500):
    """Rings and helices of uneven sizes, plus an empty field and a one-particle field."""
    cases = workloads.batched_fields(nfields=nfields, seed=seed, nmin=nmin, nmax=nmax)
    cases.insert(2, (np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0)))
    cases.append((np.array([[0.1, 0.2, 0.3]]), np.array([[0.0, 0.0, 1e-3]]), np.array([0.05]),
                  np.array([1e-4])))
    return cases


def fields(cases, precision="fp64", kernel="gaussianerf", **kw):
    return [make_field(X, G, s, precision=precision, kernel=kernel, **kw)
            for X, G, s, _ in cases]


def ctx():
    return vpm._context(0)


@pytest.fixture(autouse=True)
def _restore_schedule():
    yield
    ctx().set_schedule(N.SCHED_QUEUE, 256, 0)
    ctx().set_graph(False)


@pytest.mark.parametrize("sched,block", SCHEDULES)
@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_batch_fp64_matches_each_field_alone(sched, block, kernel):
    cases = sweep()
    alone = fields(cases, kernel=kernel)
    for pf in alone:
        vpm.UJ_gpu(pf)
    batch = fields(cases, kernel=kernel)
    ctx().set_schedule(sched, block)
    vpm.UJ_gpu_batch(batch)
    for a, b in zip(alone, batch):
        if a.np < 2:
            np.testing.assert_array_equal(vpm.get_U(b), vpm.get_U(a))
            continue
        assert normwise(vpm.get_U(b), vpm.get_U(a))[0] < 1e-14
        assert normwise(vpm.get_J(b), vpm.get_J(a))[0] < 1e-14
    # and the oracle on one small field
    X, G, s, _ = cases[1]
    Ur, Jr = O.uj_direct(kernel, X, X, G, s)
    assert normwise(vpm.get_U(batch[1]), Ur)[0] < 1e-12
    assert normwise(vpm.get_J(batch[1]), Jr)[0] < 1e-12


@pytest.mark.parametrize("sched", [N.SCHED_FLAT, N.SCHED_QUEUE])
def test_batch_mixed_block256_matches_each_field_alone(sched):
    cases = sweep()
    alone = fields(cases, precision="mixed")
    for pf in alone:
        vpm.UJ_gpu(pf)
    batch = fields(cases, precision="mixed")
    ctx().set_schedule(sched, 256)
    vpm.UJ_gpu_batch(batch)
    for a, b in zip(alone, batch):
        if a.np < 2:
            np.testing.assert_array_equal(vpm.get_U(b), vpm.get_U(a))
            continue
        assert normwise(vpm.get_U(b), vpm.get_U(a))[0] < 1e-13
        assert normwise(vpm.get_J(b), vpm.get_J(a))[0] < 1e-13


@pytest.mark.parametrize("sched,block", SCHEDULES)
def test_batch_mixed_vs_fp64(sched, block):
    cases = sweep(nfields=5, seed=12, nmin=500, nmax=6000)
    ref = fields(cases)
    vpm.UJ_gpu_batch(ref)
    mx = fields(cases, precision="mixed")
    ctx().set_schedule(sched, block)
    vpm.UJ_gpu_batch(mx)
    for a, b in zip(ref, mx):
        if a.np < 2:
            continue
        eu, ej = normwise(vpm.get_U(b), vpm.get_U(a)), normwise(vpm.get_J(b), vpm.get_J(a))
        print("mixed vs fp64", sched, block, a.np, eu, ej)
        # rings and helices: measured U <= 7.1e-7, J <= 1.33e-6
        assert eu[0] < 1e-6 and ej[0] < 2e-6, (a.np, eu, ej)


def test_batch_rk3_vs_oracle_with_freestream_per_field():
    cases = sweep(nfields=4, seed=13, nmin=100, nmax=1200)
    pfs = fields(cases, relaxation=vpm.pedrizzetti)
    uinfs = [np.array([0.1 * k, -0.05, 0.02 * k]) for k in range(len(pfs))]
    for pf, u in zip(pfs, uinfs):
        pf.Uinf = (lambda u: (lambda t: u))(u)
    refs = []
    for (X, G, s, _), u in zip(cases, uinfs):
        P = O.make_params("gaussianerf", True, "rVPM", Uinf=tuple(u), relax_kind=1, rlxf=0.3)
        refs.append((P, O.Field(X, G, s)))
    for i in range(1, 3):
        vpm.nextstep_batch(pfs, 0.01, relax=True)
        for P, F in refs:
            if len(F.X):
                O.rk3_step(P, F, 0.01, True)
    for pf, (P, F) in zip(pfs, refs):
        if pf.np == 0:
            continue
        assert normwise(vpm.get_X(pf), F.X)[0] < 1e-12
        assert normwise(vpm.get_Gamma(pf), F.G)[0] < 1e-12
        assert normwise(vpm.get_sigma(pf)[:, None], F.sigma[:, None])[0] < 1e-12
        assert pf.t == pytest.approx(0.02) and pf.nt == 2


def _resident_steps(pfs, graph, precision, nsteps=5):
    """Upload the fields, run nsteps resident RK3 steps with or without graph
    replay (relaxation from the third step on, so the graph is captured,
    replayed, re-captured and replayed), download after every step (eager work
    between replays)."""
    c = ctx()
    L = N.lib()
    c.set_graph(graph)
    a = pfs[0]
    ptrs = N.pointer_array([pf.particles for pf in pfs])
    nps = (ctypes.c_int64 * len(pfs))(*[pf.np for pf in pfs])
    c.check(L.vpmx_batch_upload(c.ptr, len(pfs), ptrs, nps, ctypes.byref(a.layout)))
    U = a._uj_params(fmm=False)
    S = a._step_params()
    states = []
    for i in range(nsteps):
        c.check(L.vpmx_batch_rk3_resident(c.ptr, ctypes.byref(U), ctypes.byref(S), None, 0.01,
                                          int(i >= 2)))
        ens = ctypes.c_double()
        c.check(L.vpmx_diagnostics(c.ptr, ctypes.byref(ens), None))
        c.check(L.vpmx_batch_download(c.ptr, len(pfs), ptrs, nps, ctypes.byref(a.layout)))
        states.append([pf.particles[:, :pf.np].copy() for pf in pfs] + [ens.value])
    c.set_graph(False)
    return states


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nf", [1, 5])
def test_graph_replay_bitwise_equals_eager(precision, nf):
    cases = sweep(nfields=nf, seed=14, nmin=300, nmax=3000)[:nf] if nf > 1 else \
        [workloads.vortex_ring(n_stations=10, n_cross=60)]
    eager = _resident_steps(fields(cases, precision=precision, relaxation=vpm.pedrizzetti),
                            False, precision)
    graph = _resident_steps(fields(cases, precision=precision, relaxation=vpm.pedrizzetti),
                            True, precision)
    for se, sg in zip(eager, graph):
        for a, b in zip(se[:-1], sg[:-1]):
            np.testing.assert_array_equal(a, b)
        assert se[-1] == sg[-1]


def test_graph_recaptures_when_dt_changes():
    X, G, s, _ = workloads.vortex_ring(n_stations=10, n_cross=60)
    runs = []
    for graph in (False, True):
        pf = make_field(X, G, s, relaxation=vpm.pedrizzetti)
        vpm.set_graphs(graph)
        for dt in (0.01, 0.01, 0.02, 0.01):
            vpm.nextstep(pf, dt, relax=True)
        vpm.set_graphs(False)
        runs.append(pf.particles[:, :pf.np].copy())
    np.testing.assert_array_equal(runs[0], runs[1])


def test_graph_timing_counts_replays():
    X, G, s, _ = workloads.vortex_ring(n_stations=10, n_cross=60)
    pf = make_field(X, G, s)
    c = ctx()
    c.set_timing(True)
    vpm.set_graphs(True)
    for _ in range(3):
        vpm.nextstep(pf, 0.01)
    vpm.set_graphs(False)
    ms, n = c.timing(N.T_GRAPH)
    c.set_timing(False)
    assert n == 3 and ms > 0.0


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_autotune_picks_a_candidate_and_results_hold(precision):
    cases = sweep(nfields=8, seed=15, nmin=200, nmax=20000)
    pfs = fields(cases, precision=precision)
    c = ctx()
    L = N.lib()
    a = pfs[0]
    ptrs = N.pointer_array([pf.particles for pf in pfs])
    nps = (ctypes.c_int64 * len(pfs))(*[pf.np for pf in pfs])
    c.check(L.vpmx_batch_upload(c.ptr, len(pfs), ptrs, nps, ctypes.byref(a.layout)))
    P = a._uj_params(fmm=False)
    res = c.autotune(P, reps=2)
    print(res)
    assert len(res) == (6 if precision == "mixed" else 3)
    assert all(ms > 0 for _, _, ms in res)
    best = min(res, key=lambda r: r[2])
    sched, block, grid = c.schedule()
    assert (sched, block) == best[:2] and grid > 0
    c.check(L.vpmx_uj_resident(c.ptr, ctypes.byref(P)))
    c.check(L.vpmx_batch_download(c.ptr, len(pfs), ptrs, nps, ctypes.byref(a.layout)))
    ref = fields(cases)
    for pf in ref:
        vpm.UJ_gpu(pf)
    for r, pf in zip(ref, pfs):
        if r.np > 1:
            assert normwise(vpm.get_U(pf), vpm.get_U(r))[0] < 1e-6


def test_batch_errors():
    cases = sweep(nfields=3, seed=16, nmin=50, nmax=200)
    pfs = fields(cases)
    c = ctx()
    L = N.lib()
    a = pfs[0]
    ptrs = N.pointer_array([pf.particles for pf in pfs])
    nps = (ctypes.c_int64 * len(pfs))(*[pf.np for pf in pfs])
    lay = ctypes.byref(a.layout)
    c.check(L.vpmx_batch_upload(c.ptr, len(pfs), ptrs, nps, lay))
    assert L.vpmx_batch_nfields(c.ptr) == len(pfs)
    P = a._uj_params(fmm=False)
    P.sfs = 1
    assert L.vpmx_uj_resident(c.ptr, ctypes.byref(P)) == N.EUNSUPPORTED
    P = a._uj_params(fmm=True)
    assert L.vpmx_uj_resident(c.ptr, ctypes.byref(P)) == N.EUNSUPPORTED
    # single-field download of a batch, wrong sizes, wrong count
    assert L.vpmx_download(c.ptr, a._ptr(), a.np, lay) == N.EINVAL
    bad = (ctypes.c_int64 * len(pfs))(*[pf.np + 1 for pf in pfs])
    assert L.vpmx_batch_download(c.ptr, len(pfs), ptrs, bad, lay) == N.EINVAL
    assert L.vpmx_batch_download(c.ptr, len(pfs) - 1, ptrs, nps, lay) == N.EINVAL
    assert L.vpmx_batch_set_schedule(c.ptr, 2, 256, 0) == N.EINVAL
    assert L.vpmx_batch_set_schedule(c.ptr, 0, 96, 0) == N.EINVAL
    # a bad sigma names its field and particle
    pfs[1].particles[vpm.SIGMA_INDEX, 3] = -1.0
    assert L.vpmx_batch_upload(c.ptr, len(pfs), ptrs, nps, lay) == N.EINVAL
    assert "field 2 particle 4" in N.last_error(c.ptr)
    # SFS in a batched step
    S = a._step_params()
    S.sfs = N.SFS_CONSTANT
    pfs[1].particles[vpm.SIGMA_INDEX, 3] = 0.05
    rc = L.vpmx_batch_rk3_step(c.ptr, len(pfs), ptrs, nps, lay, ctypes.byref(a._uj_params()),
                               ctypes.byref(S), None, 0.01, 0)
    assert rc == N.EUNSUPPORTED
    # a single resident field again after a batch
    c.check(L.vpmx_upload(c.ptr, a._ptr(), a.np, lay))
    assert L.vpmx_batch_nfields(c.ptr) == 1

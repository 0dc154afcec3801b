"""vpmx_uj_device_noalloc: the FP64 direct UJ on caller-owned device arrays
with the library's fixed scratch block, against vpmx_uj (the host-matrix
entry point running the same kernels) and the oracle, on two calls in a row
(the second smaller, so stale scratch from the first would show)."""
import ctypes

import numpy as np
import pytest

from helpers import make_field, normwise

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _cloud(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1.0, 1.0, (n, 3))
    G = rng.normal(0.0, 1e-3, (n, 3))
    s = rng.uniform(0.02, 0.05, n)
    return X, G, s


def _noalloc(ctx, X, G, s, kernel, reset=1, U0=None, J0=None, stream=None):
    """One call through the C ABI on torch device buffers (SoA, stride n)."""
    from flowunsteady_amd import _native as N
    from flowunsteady_amd import vpm
    n = len(X)
    dev = torch.device("cuda:0")
    dX = torch.tensor(np.ascontiguousarray(X.T), dtype=torch.float64, device=dev)
    dG = torch.tensor(np.ascontiguousarray(G.T), dtype=torch.float64, device=dev)
    ds = torch.tensor(s, dtype=torch.float64, device=dev)
    dU = (torch.full((3, n), float("nan"), dtype=torch.float64, device=dev) if U0 is None
          else torch.tensor(np.ascontiguousarray(U0.T), dtype=torch.float64, device=dev))
    dJ = (torch.full((9, n), float("nan"), dtype=torch.float64, device=dev) if J0 is None
          else torch.tensor(np.ascontiguousarray(J0.T), dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    P = _params(getattr(vpm, kernel).id, N.FP64, reset)
    rc = N.lib().vpmx_uj_device_noalloc(ctx.ptr, n, dX.data_ptr(), dG.data_ptr(),
                                        ds.data_ptr(), dU.data_ptr(), dJ.data_ptr(),
                                        ctypes.byref(P), stream)
    ctx.check(rc)
    torch.cuda.synchronize()
    return dU.cpu().numpy().T.copy(), dJ.cpu().numpy().T.copy()


def _params(kernel_id, precision, reset):
    from flowunsteady_amd import _native as N
    P = N.UJParams()
    P.kernel, P.precision, P.reset = kernel_id, precision, int(reset)
    P.transposed, P.sfs, P.fmm = 1, 0, 0
    return P


def _reference(X, G, s, kernel):
    from flowunsteady_amd import vpm
    pf = make_field(X, G, s, kernel=kernel)
    vpm.UJ_gpu(pf)
    return vpm.get_U(pf).copy(), vpm.get_J(pf).copy()


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_noalloc_matches_vpmx_uj_twice(kernel):
    from flowunsteady_amd import vpm
    from oracle import oracle as O
    ctx = vpm._context(0)
    stream = torch.cuda.current_stream().cuda_stream
    # first call: a split launch (fewer than 512 target blocks) on the
    # context's stream; second call: fewer particles on torch's stream
    for n, seed, st in ((40000, 11, None), (3000, 12, stream)):
        print(f"noalloc {kernel} n={n} stream={'ctx' if st is None else 'torch'}", flush=True)
        X, G, s = _cloud(n, seed)
        U, J = _noalloc(ctx, X, G, s, kernel, stream=st)
        Ur, Jr = _reference(X, G, s, kernel)
        eu, ej = normwise(U, Ur), normwise(J, Jr)
        print("  vs vpmx_uj", eu, ej, "bitwise", np.array_equal(U, Ur) and np.array_equal(J, Jr),
              flush=True)
        assert np.isfinite(U).all() and np.isfinite(J).all()
        assert eu[0] <= 1e-12 and ej[0] <= 1e-12
        idx = np.arange(0, n, max(1, n // 256))
        Uo, Jo = O.uj_direct(kernel, X[idx], X, G, s)
        assert normwise(U[idx], Uo)[0] <= 1e-12 and normwise(J[idx], Jo)[0] <= 1e-12


def test_noalloc_accumulates_and_rejects_oversize():
    from flowunsteady_amd import _native as N
    from flowunsteady_amd import vpm
    ctx = vpm._context(0)
    X, G, s = _cloud(2000, 13)
    U1, J1 = _noalloc(ctx, X, G, s, "gaussianerf")
    U2, J2 = _noalloc(ctx, X, G, s, "gaussianerf", reset=0, U0=U1, J0=J1)
    assert normwise(U2, 2.0 * U1)[0] <= 1e-15 and normwise(J2, 2.0 * J1)[0] <= 1e-15
    # above the fixed block's limit: a clear error before anything runs
    P = _params(vpm.gaussianerf.id, N.FP64, 1)
    rc = N.lib().vpmx_uj_device_noalloc(ctx.ptr, 131073, 1, 1, 1, 1, 1, ctypes.byref(P), None)
    assert rc != 0 and "no-allocation" in N.last_error(ctx.ptr)
    Pm = _params(vpm.gaussianerf.id, N.MIXED, 1)
    rc = N.lib().vpmx_uj_device_noalloc(ctx.ptr, 10, 1, 1, 1, 1, 1, ctypes.byref(Pm), None)
    assert rc != 0 and "FP64 direct" in N.last_error(ctx.ptr)

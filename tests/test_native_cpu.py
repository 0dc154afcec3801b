"""CPU checks of libvpmx.so: it loads, exports every symbol include/vpmx.h
declares, its structs have the header's layout, it reports a missing GPU as an
error code, and the FP32 radial functions of its mixed kernel meet their
accuracy bounds against the FP64 oracle.  No kernel is launched here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from flowunsteady_amd import _native as N
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vpmx.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vpmx_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    names = header_functions()
    assert len(names) >= 19
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(N.SIGNATURES), "ctypes binding and header disagree"


def test_version_string():
    assert N.lib().vpmx_version().decode().startswith("vpmx ")


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "vpmx.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(vpmx_layout), sizeof(vpmx_uj_params), sizeof(vpmx_step_params));
  printf("%zu %zu %zu %zu\n", offsetof(vpmx_uj_params, theta), offsetof(vpmx_step_params, Uinf),
         offsetof(vpmx_step_params, Cs), offsetof(vpmx_step_params, rlxf));
  printf("%zu %zu %zu\n", offsetof(vpmx_step_params, maxC), offsetof(vpmx_step_params, nu),
         offsetof(vpmx_step_params, tol));
  return 0;
}''')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    exp = [ctypes.sizeof(N.Layout), ctypes.sizeof(N.UJParams), ctypes.sizeof(N.StepParams),
           N.UJParams.theta.offset, N.StepParams.Uinf.offset, N.StepParams.Cs.offset,
           N.StepParams.rlxf.offset, N.StepParams.maxC.offset, N.StepParams.nu.offset,
           N.StepParams.tol.offset]
    assert got == exp


def _gpu_visible():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_gpu_visible(), reason="a GPU is visible")
def test_create_without_gpu_reports_enodev():
    p = ctypes.c_void_p()
    code = N.lib().vpmx_create(ctypes.byref(p), 0)
    assert code == N.ENODEV
    assert p.value is None
    assert "device" in N.last_error(None)


def test_null_arguments_rejected_without_device():
    L = N.lib()
    assert L.vpmx_create(None, 0) == N.EINVAL
    assert L.vpmx_uj(None, None, 0, None, None) == N.EINVAL
    assert "NULL" in N.last_error(None)
    assert L.vpmx_set_timing(None, 1) == N.EINVAL
    assert L.vpmx_resident_np(None) == -1
    L.vpmx_destroy(None)


def radial(kernel, r, sigma):
    A, B = ctypes.c_float(), ctypes.c_float()
    assert N.lib().vpmx_radial_f32(kernel, r, sigma, ctypes.byref(A), ctypes.byref(B)) == 0
    return A.value, B.value


def radial_ref(kernel, r, sigma):
    """A = g/r^3, B = g'/(sigma r^4) - 3 g/r^5 from the oracle's FP64 g, g'."""
    _, g, dg = O.kernel(kernel, r / sigma)
    return g / r ** 3, dg / (sigma * r ** 4) - 3 * g / r ** 5


@pytest.mark.parametrize("kernel,ident,rho_lo,tolA,tolB", [
    ("gaussianerf", N.GAUSSIANERF, 0.05, 5e-7, 1.5e-6),
    ("winckelmans", N.WINCKELMANS, 1e-3, 1e-6, 1.5e-6),
    ("gaussian", N.GAUSSIAN, 0.05, 6e-7, 2.5e-6),
    ("singular", N.SINGULAR, 1e-3, 3e-7, 5e-7),
])
def test_radial_f32_accuracy(kernel, ident, rho_lo, tolA, tolB):
    """Relative error of the mixed kernel's radial coefficients over rho in
    [rho_lo, 7] (the near-field branch; beyond rho = 7 the kernel uses g = 1).
    Below rho ~ 0.05 the FP64 closed form itself cancels (~eps/rho^2), so the
    gaussian kernels are compared there against their series."""
    sigma = 0.035
    worstA = worstB = 0.0
    for rho in np.geomspace(rho_lo, 7.0, 400):
        r = rho * sigma
        A, B = radial(ident, np.float32(r), np.float32(sigma))
        # reference evaluated at the FP32-rounded inputs
        rr, ss = float(np.float32(r)), float(np.float32(sigma))
        Ar, Br = radial_ref(kernel, rr, ss)
        assert Br < 0.0
        worstA = max(worstA, abs(A - Ar) / abs(Ar))
        worstB = max(worstB, abs(B - Br) / abs(Br))
    assert worstA < tolA, worstA
    assert worstB < tolB, worstB


def test_radial_f32_small_rho_series():
    """gaussianerf at rho -> 0: A -> sqrt(2/pi)/3 sigma^-3 and
    B -> -sqrt(2/pi)/5 sigma^-5 (limits of the series)."""
    sigma = 0.04
    A, B = radial(N.GAUSSIANERF, np.float32(1e-6 * sigma), np.float32(sigma))
    c = np.sqrt(2 / np.pi)
    assert A == pytest.approx(c / 3 / sigma ** 3, rel=1e-6)
    assert B == pytest.approx(-c / 5 / sigma ** 5, rel=1e-6)


def test_radial_far_field_is_singular():
    """Beyond rho = 7 the gaussian-erf g is 1 to FP32 rounding, so the
    far-field fast path A = r^-3, B = -3 r^-5 loses nothing."""
    sigma = 0.03
    for rho in (7.0, 9.0, 30.0):
        r = rho * sigma
        Ar, Br = radial_ref("gaussianerf", r, sigma)
        assert abs(Ar - r ** -3) / Ar < 6e-8
        assert abs(Br + 3 * r ** -5) / abs(Br) < 6e-8


@pytest.mark.parametrize("geom", ["cube", "rotor", "blob"])
@pytest.mark.parametrize("ncrit,theta,cut", [(50, 0.4, 7.0), (20, 0.5, 3.0), (1, 0.3, 0.0)])
def test_fmm_plan_matches_oracle(geom, ncrit, theta, cut):
    """The library's tree and dual-tree traversal (built on the host, as in
    vpmx_uj with fmm = 1) give the oracle FMM's cells, leaves and lists."""
    from flowunsteady_amd import workloads as W
    X, G, s, _ = {"cube": lambda: W.uniform_cube(n=6000, seed=2),
                  "rotor": lambda: W.rotor_wake(n=8192, seed=3),
                  "blob": lambda: W.gaussian_blob(n=6000, seed=4)}[geom]()
    a = N.fmm_plan_counts(X, s, ncrit, theta, cut)
    _, _, st = O.uj_fmm("gaussianerf", X, G, s, p=1, ncrit=ncrit, theta=theta, sigcut=cut)
    assert a == {k: st[k] for k in ("ncell", "nleaf", "nm2l", "np2p", "p2p_pairs")}


def test_fluid_domain_grid_and_preprocessing():
    """Host logic of the fluid-domain helpers (no GPU call)."""
    from flowunsteady_amd import processing, vpm
    P = processing.grid_nodes([0, 0, 0], [1, 2, 3], [1, 2, 3])
    assert P.shape == (2 * 3 * 4, 3)
    np.testing.assert_array_equal(P[0], [0, 0, 0])
    np.testing.assert_array_equal(P[1], [1, 0, 0])      # x fastest
    np.testing.assert_array_equal(P[2], [0, 1, 0])
    np.testing.assert_array_equal(P[-1], [1, 2, 3])
    R = np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1.0]])
    np.testing.assert_allclose(processing.grid_nodes([0, 0, 0], [1, 1, 1], [1, 1, 1], R)[1],
                               [0, 1, 0])
    pf = vpm.ParticleField(4)
    vpm.add_particles(pf, np.zeros((4, 3)), np.array([[np.nan, 0, 0], [3, 4, 0], [0, 0, 0],
                                                      [1e-3, 0, 0]]), [0.5, 0.05, 0.2, 0.01])
    counts = processing.preprocessing_fluiddomain_pfield(0.1, 1.0)(pf)
    assert counts == (2, 1, 1)
    np.testing.assert_array_equal(vpm.get_sigma(pf), [0.1, 0.05, 0.1, 0.01])
    G = vpm.get_Gamma(pf)
    np.testing.assert_array_equal(G[0], [1e-12, 1e-12, 1e-12])
    # |Gamma| = 5 > maxmagGamma is left as it is, as the reference's rebinding
    # `Gamma *= ...` leaves the particle
    np.testing.assert_array_equal(G[1], [3, 4, 0])
    assert np.all(G[2] > 0)            # zero strength lifted off zero
    np.testing.assert_array_equal(G[3], [1e-3, 0, 0])

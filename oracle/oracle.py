"""ctypes wrapper around the CPU FP64 oracle (oracle/vpm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.  Arrays are numpy
float64, AoS per particle (X shape (n, 3), J shape (n, 9) with
J[:, k + 3*l] = dU_k/dx_l).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")

SINGULAR, GAUSSIAN, GAUSSIANERF, WINCKELMANS = 0, 1, 2, 3
KERNELS = {"singular": SINGULAR, "gaussian": GAUSSIAN,
           "gaussianerf": GAUSSIANERF, "winckelmans": WINCKELMANS}

_dp = ctypes.POINTER(ctypes.c_double)


class _Field(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int64),
                ("X", _dp), ("G", _dp), ("sigma", _dp), ("U", _dp), ("J", _dp),
                ("E", _dp), ("M", _dp), ("C", _dp),
                ("is_static", ctypes.POINTER(ctypes.c_uint8)), ("vol", _dp)]


class _Params(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int), ("transposed", ctypes.c_int),
                ("f", ctypes.c_double), ("g", ctypes.c_double),
                ("Uinf", ctypes.c_double * 3),
                ("sfs", ctypes.c_int), ("Cs", ctypes.c_double),
                ("clip_backscatter", ctypes.c_int),
                ("relax_kind", ctypes.c_int), ("rlxf", ctypes.c_double),
                ("alpha", ctypes.c_double), ("sfs_rlxf", ctypes.c_double),
                ("minC", ctypes.c_double), ("maxC", ctypes.c_double),
                ("viscous", ctypes.c_int), ("nu", ctypes.c_double), ("sgm0", ctypes.c_double),
                ("beta", ctypes.c_double), ("itmax", ctypes.c_int), ("tol", ctypes.c_double),
                ("fmm", ctypes.c_int), ("p", ctypes.c_int), ("ncrit", ctypes.c_int),
                ("theta", ctypes.c_double), ("sigcut", ctypes.c_double)]


def build(force: bool = False) -> str:
    """Compile liborc.so with the committed Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liborc.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_kernel.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_zeta0.argtypes = [ctypes.c_int]
        L.orc_zeta0.restype = ctypes.c_double
        L.orc_uj_direct.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_int64,
                                    _dp, _dp, _dp, _dp, _dp]
        L.orc_estr_direct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _dp, _dp,
                                      ctypes.c_int64, _dp, _dp, _dp, _dp, _dp]
        L.orc_stretching.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_rhs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp,
                              _dp, _dp, _dp]
        L.orc_lsrk3_linear.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       _dp, _dp]
        L.orc_relax.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_rk3_step.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_Field),
                                   ctypes.c_double, ctypes.c_int]
        L.orc_euler_step.argtypes = L.orc_rk3_step.argtypes
        L.orc_ffv_direct.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int64,
                                     _dp, _dp, _dp]
        L.orc_enstrophy.argtypes = [ctypes.c_int64, _dp, _dp]
        L.orc_enstrophy.restype = ctypes.c_double
        L.orc_energy.argtypes = [ctypes.c_int64, _dp, _dp, _dp]
        L.orc_energy.restype = ctypes.c_double
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_dzeta.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_dzeta.restype = ctypes.c_double
        L.orc_kernel_dsig.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_dj_direct.argtypes = L.orc_uj_direct.argtypes
        L.orc_estr_dyn_direct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _dp, _dp,
                                          _dp, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                                          _dp]
        L.orc_dyn_terms.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                    _dp, _dp, _dp]
        L.orc_fmm_near.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int64, _dp, _dp,
                                   _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _kid(kernel):
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


def kernel(kernel, rho):
    z, g, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_kernel(_kid(kernel), float(rho), ctypes.byref(z), ctypes.byref(g), ctypes.byref(d))
    return z.value, g.value, d.value


def zeta0(kernel):
    return lib().orc_zeta0(_kid(kernel))


def uj_direct(kernel, tX, sX, sG, ssig):
    """U (nt,3) and J (nt,9) at targets tX induced by sources (sX, sG, ssig)."""
    tX = _c(tX, (-1, 3)); sX = _c(sX, (-1, 3)); sG = _c(sG, (-1, 3)); ssig = _c(ssig, (-1,))
    nt, ns = tX.shape[0], sX.shape[0]
    U = np.zeros((nt, 3)); J = np.zeros((nt, 9))
    lib().orc_uj_direct(_kid(kernel), nt, _p(tX), ns, _p(sX), _p(sG), _p(ssig), _p(U), _p(J))
    return U, J


def dzeta(kernel, rho):
    """zeta'(rho)."""
    return lib().orc_dzeta(_kid(kernel), float(rho))


def kernel_dsig(kernel, rho):
    """(h, h', zD) at rho: sigma d/dsigma of g(r/sigma) and of zeta(r/sigma)/sigma^3 (times
    sigma^3), with h' = dh/drho."""
    h, dh, zD = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_kernel_dsig(_kid(kernel), float(rho), ctypes.byref(h), ctypes.byref(dh),
                          ctypes.byref(zD))
    return h.value, dh.value, zD.value


def dj_direct(kernel, tX, sX, sG, ssig):
    """Core-size derivatives D[U] (nt,3), D[J] (nt,9), D = sigma d/dsigma of every core."""
    tX = _c(tX, (-1, 3)); sX = _c(sX, (-1, 3)); sG = _c(sG, (-1, 3)); ssig = _c(ssig, (-1,))
    nt, ns = tX.shape[0], sX.shape[0]
    DU = np.zeros((nt, 3)); DJ = np.zeros((nt, 9))
    lib().orc_dj_direct(_kid(kernel), nt, _p(tX), ns, _p(sX), _p(sG), _p(ssig), _p(DU), _p(DJ))
    return DU, DJ


def estr_dyn_direct(kernel, transposed, tX, tJ, tDJ, sX, sG, ssig, sJ, sDJ):
    """Estr (nt,3) and its core-size derivative D[E] (nt,3)."""
    tX = _c(tX, (-1, 3)); tJ = _c(tJ, (-1, 9)); tDJ = _c(tDJ, (-1, 9))
    sX = _c(sX, (-1, 3)); sG = _c(sG, (-1, 3)); ssig = _c(ssig, (-1,))
    sJ = _c(sJ, (-1, 9)); sDJ = _c(sDJ, (-1, 9))
    nt = tX.shape[0]
    E = np.zeros((nt, 3)); DE = np.zeros((nt, 3))
    lib().orc_estr_dyn_direct(_kid(kernel), int(transposed), nt, _p(tX), _p(tJ), _p(tDJ),
                              sX.shape[0], _p(sX), _p(sG), _p(ssig), _p(sJ), _p(sDJ), _p(E),
                              _p(DE))
    return E, DE


def dyn_terms(alpha, zeta0_, sigma, G, DJ, DE):
    """(num, den) of the dynamic coefficient of one particle."""
    G = _c(G, (3,)); DJ = _c(DJ, (9,)); DE = _c(DE, (3,))
    num, den = ctypes.c_double(), ctypes.c_double()
    lib().orc_dyn_terms(float(alpha), float(zeta0_), float(sigma), _p(G), _p(DJ), _p(DE),
                        ctypes.byref(num), ctypes.byref(den))
    return num.value, den.value


def fmm_near(kernel, transposed, what, X, G, sigma, J=None, DJ=None, ncrit=50, theta=0.4,
             sigcut=0.0):
    """Sums over the near-field lists of the oracle FMM's tree: what = 1 gives
    (D[U], D[J]); what = 2 gives (E, D[E]) from J (and DJ; D[E] is None without it)."""
    X = _c(X, (-1, 3)); G = _c(G, (-1, 3)); sigma = _c(sigma, (-1,))
    n = X.shape[0]
    Ja = _c(J, (-1, 9)) if J is not None else None
    DJa = _c(DJ, (-1, 9)) if DJ is not None else None
    o1 = np.zeros((n, 3))
    o2 = np.zeros((n, 9 if what == 1 else 3))
    rc = lib().orc_fmm_near(_kid(kernel), int(transposed), int(what), int(ncrit), float(theta),
                            float(sigcut), n, _p(X), _p(G), _p(sigma),
                            _p(Ja) if Ja is not None else None,
                            _p(DJa) if DJa is not None else None, _p(o1), _p(o2))
    if rc != 0:
        raise ValueError("orc_fmm_near: bad arguments")
    return o1, (o2 if (what == 1 or DJ is not None) else None)


def _viscous_ready(L):
    if not hasattr(L, "_viscous_ready"):
        L.orc_eta.restype = ctypes.c_double
        L.orc_eta.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_pse.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int64, _dp, _dp, _dp, _dp,
                              ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _dp]
        L.orc_zeta_sum.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_int64, _dp, _dp,
                                   _dp, ctypes.c_double, _dp]
        L.orc_rbf_cg.restype = ctypes.c_int
        L.orc_rbf_cg.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_double, _dp, _dp,
                                 ctypes.c_int, ctypes.c_double, _dp]
        L.orc_core_spreading_reset.restype = ctypes.c_int
        L.orc_core_spreading_reset.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_Field),
                                               ctypes.POINTER(ctypes.c_int)]
        L.orc_near_plan_build.restype = ctypes.c_void_p
        L.orc_near_plan_build.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int64, _dp, _dp]
        L.orc_near_plan_free.argtypes = [ctypes.c_void_p]
        L.orc_near_zeta.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_uint8), _dp]
        L.orc_near_pse.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp,
                                   _dp, _dp]
        L._viscous_ready = True
    return L


def _near_plan(L, X, sigma, ncrit, theta, sigcut):
    pl = L.orc_near_plan_build(int(ncrit), float(theta), float(sigcut), X.shape[0], _p(X),
                               _p(sigma))
    if not pl:
        raise ValueError("orc_near_plan_build: bad arguments")
    return pl


def near_zeta(kernel, X, G, sigma, s0=0.0, mask=None, ncrit=50, theta=0.4, sigcut=0.0):
    """The vorticity sums of zeta_sum (targets = sources = the particles) over
    the near-field lists of the oracle FMM's tree of (X, sigma)."""
    L = _viscous_ready(lib())
    X = _c(X, (-1, 3)); G = _c(G, (-1, 3)); sigma = _c(sigma, (-1,))
    m = None
    if mask is not None:
        mk = np.ascontiguousarray(mask, dtype=np.uint8)
        m = mk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    w = np.zeros((X.shape[0], 3))
    pl = _near_plan(L, X, sigma, ncrit, theta, sigcut)
    try:
        L.orc_near_zeta(pl, _kid(kernel), _p(X), _p(G), _p(sigma), float(s0), m, _p(w))
    finally:
        L.orc_near_plan_free(pl)
    return w


def near_pse(kernel, nu, X, G, sigma, vol, ncrit=50, theta=0.4, sigcut=0.0):
    """The PSE rate of pse() over the near-field lists of the oracle FMM's tree."""
    L = _viscous_ready(lib())
    X = _c(X, (-1, 3)); G = _c(G, (-1, 3)); sigma = _c(sigma, (-1,)); vol = _c(vol, (-1,))
    V = np.zeros((X.shape[0], 3))
    pl = _near_plan(L, X, sigma, ncrit, theta, sigcut)
    try:
        L.orc_near_pse(pl, _kid(kernel), float(nu), _p(X), _p(G), _p(sigma), _p(vol), _p(V))
    finally:
        L.orc_near_plan_free(pl)
    return V


def eta(kernel, rho):
    """eta(rho) = -zeta'(rho)/rho, the PSE kernel."""
    return _viscous_ready(lib()).orc_eta(_kid(kernel), float(rho))


def pse(kernel, nu, X, G, sigma, vol, targets=None):
    """PSE dGamma/dt (n or len(targets), 3) of the viscous term nu lap(omega)."""
    X = _c(X, (-1, 3)); G = _c(G, (-1, 3)); sigma = _c(sigma, (-1,)); vol = _c(vol, (-1,))
    n = X.shape[0]
    if targets is None:
        nt, tp = n, None
    else:
        t = np.ascontiguousarray(targets, dtype=np.int64)
        nt, tp = len(t), t.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    out = np.zeros((nt, 3))
    _viscous_ready(lib()).orc_pse(_kid(kernel), float(nu), n, _p(X), _p(G), _p(sigma), _p(vol),
                                  nt, tp, _p(out))
    return out


def zeta_sum(kernel, tX, sX, sG, ssig=None, s0=0.0):
    """Vorticity sum_j Gamma_j zeta_s(x - x_j) at tX (s = ssig_j, or s0 when s0 > 0)."""
    tX = _c(tX, (-1, 3)); sX = _c(sX, (-1, 3)); sG = _c(sG, (-1, 3))
    ssig = _c(np.zeros(len(sX)) if ssig is None else ssig, (-1,))
    w = np.zeros((tX.shape[0], 3))
    _viscous_ready(lib()).orc_zeta_sum(_kid(kernel), tX.shape[0], _p(tX), sX.shape[0], _p(sX),
                                       _p(sG), _p(ssig), float(s0), _p(w))
    return w


def rbf_cg(kernel, X, s0, w, G0, itmax=20, tol=1e-3):
    """CG on sum_j G_j zeta_s0(x_i - x_j) = w_i from G0: (G, iterations, residuals)."""
    X = _c(X, (-1, 3)); w = _c(w, (-1, 3))
    G = _c(G0, (-1, 3)).copy()
    res = np.zeros(3)
    it = _viscous_ready(lib()).orc_rbf_cg(_kid(kernel), X.shape[0], _p(X), float(s0), _p(w), _p(G),
                                          int(itmax), float(tol), _p(res))
    return G, it, res


def core_spreading_reset(params, field):
    """Spatial adaptation of core spreading on an oracle Field: (reset?, iterations)."""
    F = field._struct()
    it = ctypes.c_int()
    r = _viscous_ready(lib()).orc_core_spreading_reset(ctypes.byref(params), ctypes.byref(F),
                                                       ctypes.byref(it))
    return bool(r), it.value


def ffv_direct(kernel, bX, bG, bsig, fX, fG):
    """(nb, 6): Ffv/rho and Fbxf/rho of each bound particle from the free
    particles (src/FLOWUnsteady_processing_force.jl:879-929)."""
    bX = _c(bX, (-1, 3)); bG = _c(bG, (-1, 3)); bsig = _c(bsig, (-1,))
    fX = _c(fX, (-1, 3)); fG = _c(fG, (-1, 3))
    out = np.zeros((bX.shape[0], 6))
    lib().orc_ffv_direct(_kid(kernel), bX.shape[0], _p(bX), _p(bG), _p(bsig), fX.shape[0],
                         _p(fX), _p(fG), _p(out))
    return out


def estr_direct(kernel, transposed, tX, tJ, sX, sG, ssig, sJ):
    tX = _c(tX, (-1, 3)); tJ = _c(tJ, (-1, 9)); sX = _c(sX, (-1, 3)); sG = _c(sG, (-1, 3))
    ssig = _c(ssig, (-1,)); sJ = _c(sJ, (-1, 9))
    E = np.zeros((tX.shape[0], 3))
    lib().orc_estr_direct(_kid(kernel), int(transposed), tX.shape[0], _p(tX), _p(tJ),
                          sX.shape[0], _p(sX), _p(sG), _p(ssig), _p(sJ), _p(E))
    return E


def stretching(transposed, J, G):
    J = _c(J, (9,)); G = _c(G, (3,)); S = np.zeros(3)
    lib().orc_stretching(int(transposed), _p(J), _p(G), _p(S))
    return S


def rhs(transposed, f, g, zeta0_, Uinf, U, J, G, sigma, C=0.0, E=None):
    Uinf = _c(Uinf, (3,)); U = _c(U, (3,)); J = _c(J, (9,)); G = _c(G, (3,))
    Ep = None
    if E is not None:
        Ea = _c(E, (3,)); Ep = _p(Ea)
    dX = np.zeros(3); dG = np.zeros(3); ds = ctypes.c_double()
    lib().orc_rhs(int(transposed), f, g, zeta0_, _p(Uinf), _p(U), _p(J), _p(G), sigma, C,
                  Ep, _p(dX), _p(dG), ctypes.byref(ds))
    return dX, dG, ds.value


def lsrk3_linear(lam: complex, dt: float, y0: complex = 1.0 + 0j) -> complex:
    yr, yi = ctypes.c_double(y0.real), ctypes.c_double(y0.imag)
    lib().orc_lsrk3_linear(lam.real, lam.imag, dt, ctypes.byref(yr), ctypes.byref(yi))
    return complex(yr.value, yi.value)


def relax(kind, rlxf, J, G):
    J = _c(J, (9,)); G = _c(G, (3,)).copy()
    lib().orc_relax(int(kind), rlxf, _p(J), _p(G))
    return G


def enstrophy(G, J):
    G = _c(G, (-1, 3)); J = _c(J, (-1, 9))
    return lib().orc_enstrophy(G.shape[0], _p(G), _p(J))


def energy(X, G, U):
    X = _c(X, (-1, 3)); G = _c(G, (-1, 3)); U = _c(U, (-1, 3))
    return lib().orc_energy(X.shape[0], _p(X), _p(G), _p(U))


class Field:
    """A particle field for the oracle integrators (AoS float64 arrays)."""

    def __init__(self, X, G, sigma, is_static=None, vol=None):
        self.X = _c(X, (-1, 3)).copy()
        self.G = _c(G, (-1, 3)).copy()
        self.sigma = _c(sigma, (-1,)).copy()
        self.vol = (self.sigma ** 3 if vol is None else _c(vol, (-1,))).copy()
        n = self.X.shape[0]
        self.U = np.zeros((n, 3)); self.J = np.zeros((n, 9)); self.E = np.zeros((n, 3))
        self.C = np.zeros((n, 3)); self.M = np.zeros((n, 7))
        self.is_static = None if is_static is None else np.ascontiguousarray(is_static, np.uint8)

    def _struct(self):
        F = _Field()
        F.np = self.X.shape[0]
        F.X, F.G, F.sigma = _p(self.X), _p(self.G), _p(self.sigma)
        F.U, F.J, F.E, F.C, F.M = _p(self.U), _p(self.J), _p(self.E), _p(self.C), _p(self.M)
        F.is_static = (self.is_static.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
                       if self.is_static is not None else None)
        F.vol = _p(self.vol)
        return F


def make_params(kernel="gaussianerf", transposed=True, formulation="rVPM", Uinf=(0, 0, 0),
                sfs=False, Cs=0.0, clip_backscatter=True, relax_kind=0, rlxf=0.3,
                alpha=0.999, sfs_rlxf=0.005, minC=0.0, maxC=1.0, viscous=0, nu=0.0,
                sgm0=0.0, beta=1.5, itmax=20, tol=1e-3, fmm=False, p=8, ncrit=50, theta=0.4,
                sigcut=0.0):
    """sfs: False/0 none, True/1 constant Cs, "dynamic"/2 dynamic procedure.
    viscous: 0 inviscid, 1 / "corespreading", 2 / "pse".
    fmm: UJ by the oracle's CPU FMM (p, ncrit, theta, acceptance gap sigcut in
    core radii; 0 = purely geometric) with Estr and the core-size derivatives
    over its near-field lists."""
    P = _Params()
    P.kernel = _kid(kernel)
    P.transposed = int(transposed)
    if formulation == "rVPM":
        P.f, P.g = 0.0, 0.2
    elif formulation == "cVPM":
        P.f, P.g = 0.0, 0.0
    else:
        P.f, P.g = formulation
    for k in range(3):
        P.Uinf[k] = float(Uinf[k])
    P.sfs = 2 if sfs in ("dynamic", 2) else int(bool(sfs))
    P.Cs = float(Cs); P.clip_backscatter = int(clip_backscatter)
    P.relax_kind = int(relax_kind); P.rlxf = float(rlxf)
    P.alpha, P.sfs_rlxf, P.minC, P.maxC = float(alpha), float(sfs_rlxf), float(minC), float(maxC)
    P.viscous = {"corespreading": 1, "pse": 2}.get(viscous, viscous)
    P.nu, P.sgm0, P.beta = float(nu), float(sgm0), float(beta)
    P.itmax, P.tol = int(itmax), float(tol)
    P.fmm, P.p, P.ncrit = int(bool(fmm)), int(p), int(ncrit)
    P.theta, P.sigcut = float(theta), float(sigcut)
    return P


def rk3_step(params, field: Field, dt: float, relax: bool = False):
    _viscous_ready(lib())
    F = field._struct()
    lib().orc_rk3_step(ctypes.byref(params), ctypes.byref(F), dt, int(relax))


def euler_step(params, field: Field, dt: float, relax: bool = False):
    F = field._struct()
    lib().orc_euler_step(ctypes.byref(params), ctypes.byref(F), dt, int(relax))


def num_threads() -> int:
    return lib().orc_num_threads()


class FMMStats(ctypes.Structure):
    _fields_ = [("t_tree", ctypes.c_double), ("t_traverse", ctypes.c_double),
                ("t_p2m", ctypes.c_double), ("t_m2m", ctypes.c_double),
                ("t_m2l", ctypes.c_double), ("t_l2l", ctypes.c_double),
                ("t_l2p", ctypes.c_double), ("t_p2p", ctypes.c_double),
                ("ncell", ctypes.c_int64), ("nleaf", ctypes.c_int64),
                ("nm2l", ctypes.c_int64), ("np2p", ctypes.c_int64),
                ("p2p_pairs", ctypes.c_int64), ("nactive_leaf", ctypes.c_int64),
                ("p2p_pairs_all", ctypes.c_int64), ("nm2l_active", ctypes.c_int64),
                ("ncell_active", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def uj_fmm(kernel, X, G, sigma, p=4, ncrit=50, theta=0.4, sigcut=6.0, sample=None):
    """U, J of all particles by the oracle's CPU FMM (p, ncrit, theta, and the
    sigma-aware acceptance gap sigcut).  With `sample` (particle indices) only
    the leaves holding them are evaluated (others stay zero)."""
    L = lib()
    if not hasattr(L, "_fmm_ready"):
        L.orc_uj_fmm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp,
                                 ctypes.POINTER(ctypes.c_int64), ctypes.c_int64,
                                 ctypes.POINTER(FMMStats)]
        L.orc_fmm_deriv.argtypes = [ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
        L._fmm_ready = True
    X = _c(X, (-1, 3)); G = _c(G, (-1, 3)); sigma = _c(sigma, (-1,))
    n = X.shape[0]
    U = np.zeros((n, 3)); J = np.zeros((n, 9))
    st = FMMStats()
    sp, ns = None, 0
    if sample is not None:
        sa = np.ascontiguousarray(sample, dtype=np.int64)
        sp, ns = sa.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(sa)
    rc = L.orc_uj_fmm(_kid(kernel), int(p), int(ncrit), float(theta), float(sigcut), n, _p(X),
                      _p(G), _p(sigma), _p(U), _p(J), sp, ns, ctypes.byref(st))
    if rc != 0:
        raise ValueError("orc_uj_fmm: bad arguments")
    return U, J, st.as_dict()


def fmm_deriv(P, r):
    """b_k(r) = d^k(1/|r|)/k! for |k| <= P (graded order) and the multi-indices."""
    L = lib()
    if not hasattr(L, "_fmm_ready"):   # sets the FMM argtypes
        uj_fmm(0, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    K = (P + 1) * (P + 2) * (P + 3) // 6
    b = np.zeros(K)
    mi = np.zeros((K, 3), dtype=np.int32)
    r = _c(r, (3,))
    L.orc_fmm_deriv(int(P), _p(r), _p(b), mi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return b, mi

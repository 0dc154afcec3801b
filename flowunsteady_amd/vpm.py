"""Host-side mirror of the FLOWVPM particle-field interface that FLOWUnsteady
drives, backed by libvpmx.so.

FLOWUnsteady builds its field as

    vpm.ParticleField(max_particles, Float64; Uinf, formulation, viscous,
                      kernel, UJ, SFS, integration, transposed, relaxation, fmm)
                                              (src/FLOWUnsteady_simulation.jl:239-253)

and then calls `vpm.add_particle` (:486), `vpm.nextstep(pfield, dt; relax)`
(:358), `pfield.UJ(pfield)` (:544), `vpm.get_np` / `vpm.remove_particle`
(:350-365) and `vpm.monitor_enstrophy` (src/FLOWUnsteady_monitors.jl:614).
This module keeps those names, argument meanings and the column-per-particle
`particles` matrix, so the same driver logic runs on the GPU path:
`UJ=UJ_gpu` replaces `vpm.UJ_direct`/`vpm.UJ_fmm` and
`integration=rungekutta3_gpu` replaces `vpm.rungekutta3`.

Row layout of `particles` (0-based rows, one column per particle): the fields
FLOWVPM stores per particle, in its order (position, circulation, core size,
volume, circulation magnitude, velocity, vorticity, Jacobian, PSE,
integration memory, SFS coefficient, SFS term, static flag).  The library only
sees the row offsets through `vpmx_layout`, so a different row order (another
FLOWVPM version) is a different layout, not a different library.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _native as N

# ---- particle matrix rows -------------------------------------------------------
X_INDEX = slice(0, 3)
GAMMA_INDEX = slice(3, 6)
SIGMA_INDEX = 6
VOL_INDEX = 7
CIRCULATION_INDEX = 8
U_INDEX = slice(9, 12)
VORTICITY_INDEX = slice(12, 15)
J_INDEX = slice(15, 24)
PSE_INDEX = slice(24, 27)
M_INDEX = slice(27, 36)
C_INDEX = slice(36, 39)
SFS_INDEX = slice(39, 42)
STATIC_INDEX = 42
NFIELDS = 43


def default_layout() -> N.Layout:
    L = N.Layout()
    L.nrows = NFIELDS
    L.X, L.Gamma, L.sigma, L.vol = X_INDEX.start, GAMMA_INDEX.start, SIGMA_INDEX, VOL_INDEX
    L.U, L.J, L.SFS, L.C = U_INDEX.start, J_INDEX.start, SFS_INDEX.start, C_INDEX.start
    L.static_ = STATIC_INDEX
    return L


# ---- kernels (src/FLOWUnsteady_simulation.jl:37,128) ------------------------------
@dataclass(frozen=True)
class Kernel:
    """Radial basis zeta(rho), regularising function g(rho), g'(rho)
    (docs/src/theory/rvpm.md:84,89-99)."""
    name: str
    id: int
    zeta: Callable
    g: Callable
    dg: Callable

    def zeta0(self) -> float:
        return float(self.zeta(0.0))


def _erf(x):
    from scipy.special import erf
    return erf(x)


singular = Kernel("singular", N.SINGULAR, lambda r: np.zeros_like(np.asarray(r, float)),
                  lambda r: np.ones_like(np.asarray(r, float)),
                  lambda r: np.zeros_like(np.asarray(r, float)))
gaussian = Kernel("gaussian", N.GAUSSIAN, lambda r: 3 / (4 * np.pi) * np.exp(-np.asarray(r) ** 3),
                  lambda r: 1 - np.exp(-np.asarray(r) ** 3),
                  lambda r: 3 * np.asarray(r) ** 2 * np.exp(-np.asarray(r) ** 3))
gaussianerf = Kernel(
    "gaussianerf", N.GAUSSIANERF,
    lambda r: (2 * np.pi) ** -1.5 * np.exp(-np.asarray(r) ** 2 / 2),
    lambda r: _erf(np.asarray(r) / np.sqrt(2)) - np.sqrt(2 / np.pi) * np.asarray(r) * np.exp(-np.asarray(r) ** 2 / 2),
    lambda r: np.sqrt(2 / np.pi) * np.asarray(r) ** 2 * np.exp(-np.asarray(r) ** 2 / 2))
winckelmans = Kernel(
    "winckelmans", N.WINCKELMANS,
    lambda r: 15 / (8 * np.pi) * (np.asarray(r) ** 2 + 1) ** -3.5,
    lambda r: np.asarray(r) ** 3 * (np.asarray(r) ** 2 + 2.5) * (np.asarray(r) ** 2 + 1) ** -2.5,
    lambda r: 7.5 * np.asarray(r) ** 2 * (np.asarray(r) ** 2 + 1) ** -3.5)


# ---- formulations (src/FLOWUnsteady_simulation.jl:36,127) --------------------------
@dataclass(frozen=True)
class Formulation:
    name: str
    f: float
    g: float


rVPM = Formulation("rVPM", 0.0, 0.2)   # docs/src/theory/rvpm.md:196-236
cVPM = Formulation("cVPM", 0.0, 0.0)


# ---- relaxation (src/FLOWUnsteady_simulation.jl:44,135) ----------------------------
@dataclass(frozen=True)
class Relaxation:
    name: str
    kind: int
    nsteps_relax: int
    rlxf: float


norelaxation = Relaxation("norelaxation", N.RELAX_NONE, 0, 0.0)
pedrizzetti = Relaxation("pedrizzetti", N.RELAX_PEDRIZZETTI, 1, 0.3)
correctedpedrizzetti = Relaxation("correctedpedrizzetti", N.RELAX_CORRECTED, 1, 0.3)
relaxation_none = norelaxation


# ---- SFS (src/FLOWUnsteady_simulation.jl:39,130) -------------------------------------
@dataclass(frozen=True)
class SubFilterScale:
    name: str
    kind: int
    Cs: float = 0.0
    clip_backscatter: bool = True
    alpha: float = 0.999     # dynamic: test filter width alpha * sigma
    rlxf: float = 0.005      # dynamic: weight of the newest sample in the Lagrangian average
    minC: float = 0.0
    maxC: float = 1.0


noSFS = SubFilterScale("noSFS", N.SFS_NONE)
SFS_none = noSFS


def DynamicSFS(alpha: float = 0.999, rlxf: float = 0.005, minC: float = 0.0, maxC: float = 1.0,
               clip_backscatter: bool = True) -> SubFilterScale:
    """Dynamic model coefficient (vpm.DynamicSFS(Estr, pseudo3level_positive;
    alpha, maxC, clippings=[clipping_backscatter]), examples/rotorhover/rotorhover.jl:53-55):
    C = <Gamma . L> / <Gamma . m> (docs/src/theory/rvpm.md:264-296) with exact
    core-size derivatives D = sigma d/dsigma,
      Gamma . L = (3 alpha - 2)/sigma Gamma^T D[J] Gamma
        (L's unresolved u - ubar taken at the test filter alpha sigma, to first
         order in 1 - alpha: u - ubar ~ (alpha - 1) D[ubar]),
      Gamma . m = sigma^2/zeta0 Gamma . D[E],
    each with a per-particle Lagrangian average (weight rlxf of the newest
    sample), C bounded to [minC, maxC] (minC = 0: positive), then backscatter
    clipping."""
    return SubFilterScale("DynamicSFS", N.SFS_DYNAMIC, 0.0, bool(clip_backscatter),
                          float(alpha), float(rlxf), float(minC), float(maxC))


def ConstantSFS(Cs: float, clip_backscatter: bool = True) -> SubFilterScale:
    """Constant-coefficient model C = Cs with Estr_direct and, by default,
    backscatter clipping (docs/src/theory/rvpm.md:251-262,296)."""
    return SubFilterScale("ConstantSFS", N.SFS_CONSTANT, float(Cs), bool(clip_backscatter))


def SFS_Cs_nobackscatter(Cs: float = 0.15) -> SubFilterScale:
    return ConstantSFS(Cs, clip_backscatter=True)


# ---- FMM settings (src/FLOWUnsteady_simulation.jl:43) -------------------------------
@dataclass(frozen=True)
class FMM:
    """vpm.FMM(; p, ncrit, theta, nonzero_sigma, eps_tol) (the settings of
    src/FLOWUnsteady_simulation.jl:43, where eps_tol is written ε_tol):
    expansion order p, leaf size ncrit, acceptance (R_A + R_B) < theta d.
    nonzero_sigma = False (FLOWUnsteady's default) keeps the criterion purely
    geometric; True also requires the gap to exceed the kernel's default
    number of core radii of the source cell's largest sigma (gaussian-erf 7),
    so no source is expanded while its core still matters.  sigma_cut, when
    given, sets that number of core radii directly.

    eps_tol = None (FLOWUnsteady's default) evaluates at order p.  A number
    chooses the order instead: the smallest q <= MAX_P whose leading
    truncation term theta^(q+1) (the relative error of one accepted
    interaction of the Taylor expansion of 1/r to degree q when
    (R_A + R_B) = theta d) is at most eps_tol -- a per-interaction bound, so
    the field error it gives is smaller (measured at theta = 0.4: 1.7e-6 at
    p = 8 against 0.4^9 = 2.6e-4).  FLOWVPM's own use of ε_tol is not in the
    reference tree."""
    p: int = 4
    ncrit: int = 50
    theta: float = 0.4
    nonzero_sigma: bool = False
    sigma_cut: Optional[float] = None
    eps_tol: Optional[float] = None

    MAX_P = 12

    def mac_sigma_cut(self) -> float:
        if self.sigma_cut is not None:
            return float(self.sigma_cut)
        return -1.0 if self.nonzero_sigma else 0.0

    def order(self) -> int:
        """The expansion order an evaluation uses (p, or the one eps_tol picks)."""
        if self.eps_tol is None:
            return int(self.p)
        eps = float(self.eps_tol)
        if not (eps > 0.0) or not (0.0 < self.theta < 1.0):
            raise ValueError("eps_tol needs eps_tol > 0 and 0 < theta < 1")
        for q in range(0, self.MAX_P + 1):
            if self.theta ** (q + 1) <= eps:
                return q
        return self.MAX_P


# ---- viscous schemes (src/FLOWUnsteady_simulation.jl:42,133,229-234) ----------------------
@dataclass
class ViscousScheme:
    """vpm_viscous.  Mutable like FLOWVPM's: run_simulation sets `nu` (and a
    core spreading's `sgm0`) before building the field (:229-234)."""
    name: str
    kind: int
    nu: float = 0.0
    sgm0: float = 0.0
    beta: float = 1.5
    itmax: int = 20
    tol: float = 1e-3
    zeta: object = None


def Inviscid() -> ViscousScheme:
    return ViscousScheme("Inviscid", N.VISCOUS_NONE)


def ParticleStrengthExchange(nu: float) -> ViscousScheme:
    """Particle strength exchange: dGamma_i/dt += (2 nu / s^2) sum_j
    (vol_i Gamma_j - vol_j Gamma_i) eta_s(x_i - x_j), eta = -zeta'/rho."""
    return ViscousScheme("ParticleStrengthExchange", N.VISCOUS_PSE, float(nu))


@dataclass(frozen=True)
class ZetaScheme:
    """How core spreading evaluates the vorticity sum_j Gamma_j zeta_s(x - x_j)
    at the particles: all pairs (zeta_direct) or the near field of an FMM tree
    with the field's FMM settings (zeta_fmm; the far field of zeta is
    neglected, as the reference's zeta_fmm leaves it to the FMM's near field)."""
    name: str
    fmm: bool


zeta_direct = ZetaScheme("zeta_direct", False)
zeta_fmm = ZetaScheme("zeta_fmm", True)


def CoreSpreading(nu: float, sgm0: float, zeta=None, *, beta: float = 1.5, itmax: int = 20,
                  tol: float = 1e-3) -> ViscousScheme:
    """Core spreading (cores grow as d sigma/dt = nu/sigma) with radial-basis
    function re-fitting when a core exceeds beta * sgm0 (conjugate gradients,
    itmax iterations, relative residual tol), as
    vpm.CoreSpreading(nu, sgm0, zeta; beta, itmax, tol)
    (examples/rotorhover/rotorhover.jl:170).  `zeta` (zeta_direct or
    zeta_fmm) selects the vorticity sums of a standalone
    rbf_conjugategradient; inside a time step they follow the field's UJ
    scheme (UJ_fmm_gpu: the near lists of an FMM tree built at the step's
    final positions; UJ_gpu: all pairs).  The reference's examples pass
    beta=100.0, whose meaning in FLOWVPM is not available here; this trigger
    is beta * sgm0."""
    if zeta is not None and not isinstance(zeta, ZetaScheme):
        raise ValueError("zeta must be vpm.zeta_direct or vpm.zeta_fmm")
    return ViscousScheme("CoreSpreading", N.VISCOUS_CORESPREADING, float(nu), float(sgm0),
                         float(beta), int(itmax), float(tol), zeta)


def isinviscid(v: ViscousScheme) -> bool:
    return v.kind == N.VISCOUS_NONE


def iscorespreading(v: ViscousScheme) -> bool:
    return v.kind == N.VISCOUS_CORESPREADING


# ---- contexts: one per device, created on first use -----------------------------------
_contexts: dict = {}


def _context(device: int) -> N.Context:
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = _contexts[device] = N.Context(device)
    return ctx


# ---- the particle field -----------------------------------------------------------------
class ParticleField:
    """Column-per-particle particle matrix plus the solver hooks.

    `UJ` and `integration` are function-valued, as in FLOWVPM: the defaults
    are the GPU drop-ins `UJ_gpu` and `rungekutta3_gpu`.  `precision` selects
    the library's FP64 path or its mixed FP32/FP64 path.
    """

    def __init__(self, maxparticles: int, *, formulation: Formulation = rVPM,
                 kernel: Kernel = gaussianerf, UJ: Optional[Callable] = None,
                 SFS: SubFilterScale = noSFS, integration: Optional[Callable] = None,
                 transposed: bool = True, relaxation: Relaxation = pedrizzetti,
                 fmm: FMM = FMM(), Uinf: Callable = lambda t: np.zeros(3),
                 precision: str = "fp64", device: int = 0,
                 viscous: Optional[ViscousScheme] = None):
        if maxparticles < 0:
            raise ValueError("maxparticles must be non-negative")
        self.particles = np.zeros((NFIELDS, maxparticles), dtype=np.float64, order="F")
        self.maxparticles = maxparticles
        self.np = 0
        self.nt = 0
        self.t = 0.0
        self.formulation = formulation
        self.kernel = kernel
        self.UJ = UJ if UJ is not None else UJ_gpu
        self.SFS = SFS
        self.integration = integration if integration is not None else rungekutta3_gpu
        self.transposed = transposed
        self.relaxation = relaxation
        self.fmm = fmm
        self.viscous = viscous if viscous is not None else Inviscid()
        self.Uinf = Uinf
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.device = device
        self.layout = default_layout()

    # -- pointer handed to the C ABI: the whole column-major matrix
    def _ptr(self):
        return self.particles.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def _uj_params(self, *, reset=True, sfs=False, transposed=None, fmm=None) -> N.UJParams:
        P = N.UJParams()
        P.kernel = self.kernel.id
        P.precision = N.MIXED if self.precision == "mixed" else N.FP64
        P.transposed = int(self.transposed if transposed is None else transposed)
        P.reset = int(reset)
        P.sfs = int(bool(sfs))
        P.fmm = int(self.UJ is UJ_fmm_gpu) if fmm is None else int(fmm)
        P.p, P.ncrit, P.theta = self.fmm.order(), self.fmm.ncrit, self.fmm.theta
        P.sigma_cut = self.fmm.mac_sigma_cut()
        return P

    def _step_params(self) -> N.StepParams:
        S = N.StepParams()
        S.f, S.g = self.formulation.f, self.formulation.g
        u = np.asarray(self.Uinf(self.t), dtype=float)
        for k in range(3):
            S.Uinf[k] = float(u[k])
        S.sfs = self.SFS.kind
        S.Cs = self.SFS.Cs
        S.clip_backscatter = int(self.SFS.clip_backscatter)
        S.relax_kind = self.relaxation.kind
        S.rlxf = self.relaxation.rlxf
        S.alpha, S.sfs_rlxf = self.SFS.alpha, self.SFS.rlxf
        S.minC, S.maxC = self.SFS.minC, self.SFS.maxC
        v = self.viscous
        S.viscous, S.nu, S.sgm0, S.beta = v.kind, v.nu, v.sgm0, v.beta
        S.itmax, S.tol = v.itmax, v.tol
        return S


def add_particle(pfield: ParticleField, X, Gamma, sigma, *, vol=0.0, circulation=1.0,
                 C=0.0, static=False):
    """Append one particle (vpm.add_particle, src/FLOWUnsteady_simulation.jl:486)."""
    if pfield.np >= pfield.maxparticles:
        raise ValueError(f"PARTICLE OVERFLOW: maximum of {pfield.maxparticles} particles reached")
    col = pfield.particles[:, pfield.np]
    col[:] = 0.0
    col[X_INDEX] = X
    col[GAMMA_INDEX] = Gamma
    col[SIGMA_INDEX] = sigma
    col[VOL_INDEX] = vol
    col[CIRCULATION_INDEX] = circulation
    col[C_INDEX] = C
    col[STATIC_INDEX] = 1.0 if static else 0.0
    pfield.np += 1


def add_particles(pfield: ParticleField, X, Gamma, sigma, vol=None, static=None):
    """Vectorised add_particle for many particles at once."""
    X = np.asarray(X, float).reshape(-1, 3)
    n = X.shape[0]
    if pfield.np + n > pfield.maxparticles:
        raise ValueError(f"PARTICLE OVERFLOW: maximum of {pfield.maxparticles} particles reached")
    sl = slice(pfield.np, pfield.np + n)
    P = pfield.particles
    P[:, sl] = 0.0
    P[X_INDEX, sl] = X.T
    P[GAMMA_INDEX, sl] = np.asarray(Gamma, float).reshape(-1, 3).T
    P[SIGMA_INDEX, sl] = np.broadcast_to(np.asarray(sigma, float), (n,))
    if vol is not None:
        P[VOL_INDEX, sl] = np.broadcast_to(np.asarray(vol, float), (n,))
    P[CIRCULATION_INDEX, sl] = 1.0
    if static is not None:
        P[STATIC_INDEX, sl] = np.asarray(static, float)
    pfield.np += n


def remove_particle(pfield: ParticleField, i: int):
    """Remove particle i (0-based) by moving the last particle into its slot,
    as FLOWVPM's remove_particle does."""
    if not 0 <= i < pfield.np:
        raise IndexError(i)
    last = pfield.np - 1
    if i != last:
        pfield.particles[:, i] = pfield.particles[:, last]
    pfield.particles[:, last] = 0.0
    pfield.np -= 1


def get_np(pfield: ParticleField) -> int:
    return pfield.np


def get_X(pfield, i=None):
    return _rows(pfield, X_INDEX, i)


def get_Gamma(pfield, i=None):
    return _rows(pfield, GAMMA_INDEX, i)


def get_sigma(pfield, i=None):
    return pfield.particles[SIGMA_INDEX, :pfield.np] if i is None else pfield.particles[SIGMA_INDEX, i]


def get_U(pfield, i=None):
    return _rows(pfield, U_INDEX, i)


def get_J(pfield, i=None):
    """J[:, k + 3*l] = dU_k/dx_l (column-major 3x3 per particle)."""
    return _rows(pfield, J_INDEX, i)


def get_SFS(pfield, i=None):
    return _rows(pfield, SFS_INDEX, i)


def get_C(pfield, i=None):
    return pfield.particles[C_INDEX.start, :pfield.np] if i is None else pfield.particles[C_INDEX.start, i]


def _rows(pfield, sl, i):
    if i is None:
        return pfield.particles[sl, :pfield.np].T
    return pfield.particles[sl, i]


def _reset_particles(pfield: ParticleField):
    """Zero U and J (processing_force.jl:237)."""
    pfield.particles[U_INDEX, :pfield.np] = 0.0
    pfield.particles[J_INDEX, :pfield.np] = 0.0


# ---- the drop-in hooks --------------------------------------------------------------------
def UJ_gpu(pfield: ParticleField, *, reset: bool = True, reset_sfs: bool = False,
           sfs: bool = False, transposed: Optional[bool] = None, **kwargs):
    """Drop-in for vpm.UJ_direct / vpm.UJ_fmm (called as pfield.UJ(pfield) at
    src/FLOWUnsteady_simulation.jl:544).  Writes U and J of all particles
    (static and probes included) into the particle matrix; with `sfs`, also
    the Estr term into the SFS rows.  Extra keyword arguments are accepted and
    ignored, as FLOWVPM's integrators pass their own."""
    if reset_sfs and sfs:
        pfield.particles[SFS_INDEX, :pfield.np] = 0.0
    ctx = _context(pfield.device)
    P = pfield._uj_params(reset=reset, sfs=sfs, transposed=transposed, fmm=False)
    ctx.check(N.lib().vpmx_uj(ctx.ptr, pfield._ptr(), pfield.np, ctypes.byref(pfield.layout),
                              ctypes.byref(P)))


UJ_direct = UJ_gpu


def UJ_fmm_gpu(pfield: ParticleField, *, reset: bool = True, **kwargs):
    """Drop-in for vpm.UJ_fmm (the default vpm_UJ, src/FLOWUnsteady_simulation.jl:38,129):
    U and J by the fast multipole method with pfield.fmm's p, ncrit, theta."""
    ctx = _context(pfield.device)
    P = pfield._uj_params(reset=reset, fmm=True)
    ctx.check(N.lib().vpmx_uj(ctx.ptr, pfield._ptr(), pfield.np, ctypes.byref(pfield.layout),
                              ctypes.byref(P)))


UJ_fmm = UJ_fmm_gpu


def mirror_particles(pfield: ParticleField, mirror_X, mirror_normal):
    """Append the mirror image of every particle as a static particle, the
    method of images of Vvpm_on_Xs (src/FLOWUnsteady_simulation.jl:518-532):
    X_m = X - 2 ((X - X_0) . n) n and, as the reference writes it,
    Gamma_m = 2 (Gamma . n) Gamma / |Gamma| - Gamma, keeping sigma, vol,
    circulation and C.  A particle with Gamma = 0 gets Gamma_m = 0 (the
    reference's formula divides by |Gamma| there)."""
    n = pfield.np
    X0 = np.asarray(mirror_X, float).reshape(3)
    nv = np.asarray(mirror_normal, float).reshape(3)
    P = pfield.particles
    X = P[X_INDEX, :n].T.copy()
    G = P[GAMMA_INDEX, :n].T.copy()
    Xm = X - 2.0 * ((X - X0) @ nv)[:, None] * nv[None, :]
    gn = np.linalg.norm(G, axis=1)
    safe = np.where(gn > 0.0, gn, 1.0)
    Gm = np.where(gn[:, None] > 0.0, 2.0 * (G @ nv)[:, None] * G / safe[:, None] - G, 0.0)
    for i in range(n):
        add_particle(pfield, Xm[i], Gm[i], P[SIGMA_INDEX, i], vol=P[VOL_INDEX, i],
                     circulation=P[CIRCULATION_INDEX, i], C=P[C_INDEX, i].copy(), static=True)


def Vvpm_on_Xs(pfield: ParticleField, Xs, *, static_particles_fun=None, dt: float = 0.0,
               fsgm: float = 1.0, mirror: bool = False, mirror_X=None,
               mirror_normal=None) -> np.ndarray:
    """Velocity induced by the particle field at points Xs (m, 3), FLOWUnsteady's
    Vvpm_on_Xs (src/FLOWUnsteady_simulation.jl:494-570) with its arguments:
    static particles from static_particles_fun(pfield, t, dt) and, with mirror,
    mirror images (mirror_particles) are appended for the evaluation and
    removed afterwards; fsgm is applied and undone as the reference does it
    (one matrix entry -- the first particle's sigma -- scaled once per static
    particle already in the field).  The probes are not appended: the library
    evaluates U at them from the field (vpmx_uj_targets), so only the m x N
    target-source pairs are computed; the probes carry no circulation, so they
    add nothing to each other.  The evaluation is direct also for an FMM
    field."""
    Xs = np.ascontiguousarray(Xs, dtype=np.float64).reshape(-1, 3)
    m = Xs.shape[0]
    U = np.zeros((m, 3))
    J = np.zeros((m, 9))
    if m == 0:
        return U
    org_np = pfield.np
    nscaled = 0
    if abs(fsgm) != 1:
        for i in range(pfield.np):
            if pfield.particles[STATIC_INDEX, i] > 0:
                pfield.particles[SIGMA_INDEX, 0] *= fsgm
                nscaled += 1
    try:
        if static_particles_fun is not None:
            static_particles_fun(pfield, pfield.t, dt)
        if mirror:
            mirror_particles(pfield, mirror_X, mirror_normal)
        ctx = _context(pfield.device)
        L = N.lib()
        ctx.check(L.vpmx_upload(ctx.ptr, pfield._ptr(), pfield.np, ctypes.byref(pfield.layout)))
        P = pfield._uj_params(fmm=False)
        dp = ctypes.POINTER(ctypes.c_double)
        ctx.check(L.vpmx_uj_targets(ctx.ptr, Xs.ctypes.data_as(dp), m, U.ctypes.data_as(dp),
                                    J.ctypes.data_as(dp), ctypes.byref(P)))
    finally:
        for i in range(pfield.np - 1, org_np - 1, -1):
            remove_particle(pfield, i)
        for _ in range(nscaled):
            pfield.particles[SIGMA_INDEX, 0] /= fsgm
    return U


def rungekutta3_gpu(pfield: ParticleField, dt: float, *, relax: bool = False, **kwargs):
    """Drop-in for vpm.rungekutta3 (vpm_integration,
    src/FLOWUnsteady_simulation.jl:40,131): three low-storage RK stages with
    the particle state resident on the GPU, then relaxation if `relax`."""
    ctx = _context(pfield.device)
    U = pfield._uj_params(sfs=pfield.SFS.kind != N.SFS_NONE)
    S = pfield._step_params()
    ctx.check(N.lib().vpmx_rk3_step(ctx.ptr, pfield._ptr(), pfield.np,
                                    ctypes.byref(pfield.layout), ctypes.byref(U),
                                    ctypes.byref(S), float(dt), int(bool(relax))))


def euler_gpu(pfield: ParticleField, dt: float, *, relax: bool = False, **kwargs):
    """Drop-in for vpm.euler (src/FLOWUnsteady_simulation.jl:40)."""
    ctx = _context(pfield.device)
    U = pfield._uj_params(sfs=pfield.SFS.kind != N.SFS_NONE)
    S = pfield._step_params()
    ctx.check(N.lib().vpmx_euler_step(ctx.ptr, pfield._ptr(), pfield.np,
                                      ctypes.byref(pfield.layout), ctypes.byref(U),
                                      ctypes.byref(S), float(dt), int(bool(relax))))


rungekutta3 = rungekutta3_gpu
euler = euler_gpu


def nextstep(pfield: ParticleField, dt: float, *, relax: bool = False, custom_UJ=None):
    """vpm.nextstep (src/FLOWUnsteady_simulation.jl:358): one integration step,
    then advance t and nt."""
    pfield.integration(pfield, dt, relax=relax)
    pfield.t += dt
    pfield.nt += 1


# ---- sweeps: many independent fields advanced together ---------------------------------
def _batch_args(pfields, fmm_ok=False):
    """Pointer array, sizes and the shared settings of a list of fields.  The
    fields are independent cases (never interacting) that share the solver
    settings: kernel, precision, formulation, SFS model, relaxation, device."""
    pfields = list(pfields)
    if not pfields:
        raise ValueError("no particle fields")
    a = pfields[0]
    for pf in pfields[1:]:
        same = (pf.kernel.id == a.kernel.id and pf.precision == a.precision
                and pf.formulation == a.formulation and pf.SFS.kind == a.SFS.kind
                and pf.relaxation.kind == a.relaxation.kind
                and pf.relaxation.rlxf == a.relaxation.rlxf and pf.transposed == a.transposed
                and pf.device == a.device and pf.layout.nrows == a.layout.nrows)
        if not same:
            raise ValueError("batched fields must share kernel, precision, formulation, SFS, "
                             "relaxation, transposed and device")
    ptrs = N.pointer_array([pf.particles for pf in pfields])
    nps = (ctypes.c_int64 * len(pfields))(*[pf.np for pf in pfields])
    return pfields, a, ptrs, nps


def UJ_gpu_batch(pfields, *, reset: bool = True, **kwargs):
    """UJ_gpu of every field of a sweep in one launch (fields never interact):
    U and J of each field written into its own particle matrix."""
    pfields, a, ptrs, nps = _batch_args(pfields)
    ctx = _context(a.device)
    P = a._uj_params(reset=reset, fmm=False)
    ctx.check(N.lib().vpmx_batch_uj(ctx.ptr, len(pfields), ptrs, nps, ctypes.byref(a.layout),
                                    ctypes.byref(P)))


def rungekutta3_gpu_batch(pfields, dt: float, *, relax: bool = False, **kwargs):
    """rungekutta3_gpu of every field of a sweep as one batched step: the same
    dt and settings, each case with its own freestream Uinf(t)."""
    pfields, a, ptrs, nps = _batch_args(pfields)
    ctx = _context(a.device)
    U = a._uj_params(sfs=a.SFS.kind != N.SFS_NONE, fmm=False)
    S = a._step_params()
    uinf = np.concatenate([np.asarray(pf.Uinf(pf.t), dtype=float).reshape(3) for pf in pfields])
    ctx.check(N.lib().vpmx_batch_rk3_step(ctx.ptr, len(pfields), ptrs, nps,
                                          ctypes.byref(a.layout), ctypes.byref(U),
                                          ctypes.byref(S), uinf.ctypes.data_as(N._dp),
                                          float(dt), int(bool(relax))))


def nextstep_batch(pfields, dt: float, *, relax: bool = False):
    """vpm.nextstep of every field of a sweep (src/FLOWUnsteady_simulation.jl:358),
    as one batched step."""
    rungekutta3_gpu_batch(pfields, dt, relax=relax)
    for pf in pfields:
        pf.t += dt
        pf.nt += 1


def set_graphs(on: bool, device: int = 0):
    """Run time steps on the direct scheme as replayed hipGraphs (bitwise the
    same results as eager launches)."""
    _context(device).set_graph(on)


def rbf_conjugategradient(pfield: ParticleField, *, force: bool = True):
    """Core spreading's spatial adaptation on the field (vpm.rbf_conjugategradient):
    free cores reset to the scheme's sgm0 and strengths re-fitted to the
    vorticity at the particles.  Returns (reset?, CG iterations, residuals)."""
    if not iscorespreading(pfield.viscous):
        raise ValueError("the field's viscous scheme is not CoreSpreading")
    ctx = _context(pfield.device)
    L = N.lib()
    ctx.check(L.vpmx_upload(ctx.ptr, pfield._ptr(), pfield.np, ctypes.byref(pfield.layout)))
    did, it = ctypes.c_int32(), ctypes.c_int32()
    res = (ctypes.c_double * 3)()
    z = pfield.viscous.zeta
    fmm = z.fmm if z is not None else pfield.UJ is UJ_fmm_gpu
    ctx.check(L.vpmx_core_spreading_reset(ctx.ptr, ctypes.byref(pfield._uj_params(fmm=fmm)),
                                          ctypes.byref(pfield._step_params()), int(force),
                                          ctypes.byref(did), ctypes.byref(it), res))
    ctx.check(L.vpmx_download(ctx.ptr, pfield._ptr(), pfield.np, ctypes.byref(pfield.layout)))
    return bool(did.value), it.value, list(res)


def relax_due(pfield: ParticleField, i: int) -> bool:
    """Relaxation cadence of src/FLOWUnsteady_simulation.jl:346-348."""
    r = pfield.relaxation
    return r.kind != N.RELAX_NONE and r.nsteps_relax >= 1 and i > 0 and i % r.nsteps_relax == 0


def monitor_enstrophy(pfield: ParticleField) -> float:
    """0.5 sum_p Gamma_p . omega(x_p) with omega = curl U from J
    (src/FLOWUnsteady_monitors.jl:573-574)."""
    J = get_J(pfield)
    G = get_Gamma(pfield)
    w = np.stack([J[:, 5] - J[:, 7], J[:, 6] - J[:, 2], J[:, 1] - J[:, 3]], axis=1)
    return 0.5 * float(np.sum(G * w))


def save(pfield: ParticleField, file_name: str, *, path: str = "", add_num: bool = True,
         num=None, overwrite_time=None) -> str:
    """Write the field for visualisation and restart (FLOWVPM's `vpm.save`,
    called at src/FLOWUnsteady_simulation.jl:438): an HDF5 file of FLOWVPM's
    datasets and its XDMF descriptor (pfield_io module); returns the XDMF
    file name."""
    from . import pfield_io
    return pfield_io.save(pfield, file_name, path=path, add_num=add_num, num=num,
                          overwrite_time=overwrite_time)


def read(pfield: ParticleField, file_name: str, *, overwrite: bool = True,
         load_time: bool = False) -> ParticleField:
    """Restart from a file written by `save` (FLOWVPM's `vpm.read!`, called at
    src/FLOWUnsteady_simulation.jl:264 with overwrite=true, load_time=false)."""
    from . import pfield_io
    return pfield_io.read(pfield, file_name, overwrite=overwrite, load_time=load_time)

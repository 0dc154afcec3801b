"""Post-processing passes of FLOWUnsteady over a particle field, on libvpmx.so.

  Ffv_direct(source, target)     free-vortex force of the bound vortices in
                                 `source` from the free particles in `target`
                                 (src/FLOWUnsteady_processing_force.jl:860-929).
  fluid_domain(pfield, ...)      velocity and vorticity the field induces on
                                 the nodes of a Cartesian grid, the probing
                                 that computefluiddomain does
                                 (examples/rotorhover/rotorhover_fluiddomain.jl:93-104),
                                 without appending the nodes to the field.
  preprocessing_fluiddomain_pfield(maxsigma, maxmagGamma)
                                 the clipping of core sizes and strengths that
                                 uns.generate_preprocessing_fluiddomain_pfield
                                 applies before probing (same example, :83-84).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from . import vpm


def Ffv_direct(source: "vpm.ParticleField", target: "vpm.ParticleField"):
    """For every particle Pb of `source` (bound vortices, static included):
    Pb.M[1:3] = Ffv/rho = sum_f U_b(x_f) x Gamma_f and Pb.M[4:6] = Fbxf/rho,
    summed over every particle of `target` (static included); written into the
    first six M rows of `source.particles` as the reference does."""
    ctx = vpm._context(target.device)
    L = N.lib()
    ctx.check(L.vpmx_upload(ctx.ptr, target._ptr(), target.np, ctypes.byref(target.layout)))
    out = np.zeros((source.np, 6))
    ctx.check(L.vpmx_ffv(ctx.ptr, source._ptr(), source.np, ctypes.byref(source.layout),
                         source.kernel.id, out.ctypes.data_as(N._dp)))
    source.particles[vpm.M_INDEX.start:vpm.M_INDEX.start + 6, :source.np] = out.T
    return out


def grid_nodes(Pmin, Pmax, NDIVS, Oaxis=None, O=None):
    """Nodes (n, 3) of the Cartesian grid [Pmin, Pmax] with NDIVS cells per
    direction, x fastest, rotated by Oaxis (rows = axes) about O."""
    Pmin = np.asarray(Pmin, float)
    Pmax = np.asarray(Pmax, float)
    NDIVS = np.asarray(NDIVS, int)
    axes = [np.linspace(Pmin[k], Pmax[k], NDIVS[k] + 1) for k in range(3)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if Oaxis is not None:
        P = P @ np.asarray(Oaxis, float)
    if O is not None:
        P = P + np.asarray(O, float)
    return P


def fluid_domain(pfield: "vpm.ParticleField", Pmin, Pmax, NDIVS, Oaxis=None, O=None,
                 include_freestream=False):
    """U (n, 3), vorticity W (n, 3) and the nodes (n, 3) of the grid, induced
    by the particle field (the freestream excluded, as in the reference's
    fluid domain, unless include_freestream)."""
    P = grid_nodes(Pmin, Pmax, NDIVS, Oaxis, O)
    m = len(P)
    U = np.zeros((m, 3))
    J = np.zeros((m, 9))
    if m:
        ctx = vpm._context(pfield.device)
        L = N.lib()
        ctx.check(L.vpmx_upload(ctx.ptr, pfield._ptr(), pfield.np, ctypes.byref(pfield.layout)))
        prm = pfield._uj_params(fmm=False)
        ctx.check(L.vpmx_uj_targets(ctx.ptr, np.ascontiguousarray(P).ctypes.data_as(N._dp), m,
                                    U.ctypes.data_as(N._dp), J.ctypes.data_as(N._dp),
                                    ctypes.byref(prm)))
    # omega = curl U from J[k + 3 l] = dU_k / dx_l
    W = np.stack([J[:, 5] - J[:, 7], J[:, 6] - J[:, 2], J[:, 1] - J[:, 3]], axis=1)
    if include_freestream:
        U = U + np.asarray(pfield.Uinf(pfield.t), float)[None, :]
    return U, W, P


def preprocessing_fluiddomain_pfield(maxsigma: float, maxmagGamma: float):
    """A function of a particle field that prepares it for probing, in place,
    as generate_preprocessing_fluiddomain_pfield does
    (src/FLOWUnsteady_postprocessing.jl:187-240), static particles included:
    core sizes above maxsigma are set to maxsigma, NaN strengths to 1e-12 per
    component, and strengths below 1e-14 in magnitude get 1e-14 added.
    Strengths above maxmagGamma are counted but left as they are: the
    reference's `Gamma *= maxmagGamma / magGamma` rebinds its local name to a
    scaled copy, so the particle keeps its strength, and that behaviour is
    kept.  Returns (shrunk sigmas, NaN strengths, strengths above maxmagGamma),
    the counts the reference prints."""
    def fn(pfield: "vpm.ParticleField", *args, **kwargs):
        n = pfield.np
        s = pfield.particles[vpm.SIGMA_INDEX, :n]
        nsig = int(np.count_nonzero(s > maxsigma))
        np.minimum(s, maxsigma, out=s)
        G = pfield.particles[vpm.GAMMA_INDEX, :n]
        nan = np.any(np.isnan(G), axis=0)
        G[:, nan] = 1e-12
        mag = np.linalg.norm(G, axis=0)
        nbig = int(np.count_nonzero(mag > maxmagGamma))
        G[:, mag < 1e-14] += 1e-14
        return nsig, int(np.count_nonzero(nan)), nbig
    return fn

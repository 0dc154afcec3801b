"""Shared test helpers."""
import numpy as np


def normwise(a, ref):
    """(max_i |a_i - ref_i| / max_j |ref_j|, rms(a - ref) / rms(ref)) over
    per-particle vectors (rows)."""
    a = np.asarray(a, float).reshape(len(a), -1)
    ref = np.asarray(ref, float).reshape(len(ref), -1)
    d = np.linalg.norm(a - ref, axis=1)
    n = np.linalg.norm(ref, axis=1)
    mx = d.max() / max(n.max(), 1e-300)
    rms = np.sqrt(np.mean(d ** 2)) / max(np.sqrt(np.mean(n ** 2)), 1e-300)
    return float(mx), float(rms)


def make_field(X, G, sigma, *, kernel="gaussianerf", precision="fp64", vol=None, static=None,
               extra=0, **kw):
    """A ParticleField of the host mirror holding the given particles."""
    from flowunsteady_amd import vpm
    X = np.asarray(X, float).reshape(-1, 3)
    pf = vpm.ParticleField(len(X) + extra, kernel=getattr(vpm, kernel), precision=precision, **kw)
    vpm.add_particles(pf, X, G, sigma, vol=vol, static=static)
    return pf


def sample_targets(n, m, seed=0):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=min(m, n), replace=False))

"""GPU parity of the post-processing passes: the free-vortex force
(_Ffv_direct, src/FLOWUnsteady_processing_force.jl:860-929) and the fluid
domain probing, against the oracle."""
import numpy as np
import pytest

from flowunsteady_amd import processing, vpm, workloads
from helpers import make_field, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian"])
@pytest.mark.parametrize("nb", [1, 13, 700])
def test_ffv_vs_oracle(kernel, nb):
    X, G, s, _ = workloads.gaussian_blob(n=3000, seed=31, std=0.4)
    rng = np.random.default_rng(nb)
    bX = rng.uniform(-0.5, 0.5, (nb, 3))
    bG = rng.normal(0, 0.01, (nb, 3))
    bs = rng.uniform(0.02, 0.1, nb)
    bX[0] = X[5]                     # coincident with a free particle: skipped
    free = make_field(X, G, s, kernel=kernel)
    bound = make_field(bX, bG, bs, kernel=kernel, static=np.ones(nb))
    out = processing.Ffv_direct(bound, free)
    ref = O.ffv_direct(kernel, bX, bG, bs, X, G)
    assert normwise(out[:, :3], ref[:, :3])[0] < 1e-12
    assert normwise(out[:, 3:], ref[:, 3:])[0] < 1e-12
    Mrows = bound.particles[vpm.M_INDEX.start:vpm.M_INDEX.start + 6, :nb].T
    np.testing.assert_array_equal(Mrows, out)
    # bitwise repeatable
    np.testing.assert_array_equal(processing.Ffv_direct(bound, free), out)


def test_ffv_empty_fields():
    X, G, s, _ = workloads.gaussian_blob(n=10, seed=1)
    free = make_field(X, G, s)
    empty = vpm.ParticleField(4)
    assert processing.Ffv_direct(empty, free).shape == (0, 6)
    bound = make_field(X[:3], G[:3], s[:3])
    out = processing.Ffv_direct(bound, vpm.ParticleField(2))
    assert np.all(out == 0.0)


def test_fluid_domain_vs_oracle():
    X, G, s, _ = workloads.vortex_ring(n_stations=12, n_cross=40)
    pf = make_field(X, G, s)
    U, W, P = processing.fluid_domain(pf, [-1.5, -1.5, -0.5], [1.5, 1.5, 0.5], [12, 10, 4])
    assert P.shape == (13 * 11 * 5, 3)
    Ur, Jr = O.uj_direct("gaussianerf", P, X, G, s)
    assert normwise(U, Ur)[0] < 1e-12
    Wr = np.stack([Jr[:, 5] - Jr[:, 7], Jr[:, 6] - Jr[:, 2], Jr[:, 1] - Jr[:, 3]], axis=1)
    assert normwise(W, Wr)[0] < 1e-12

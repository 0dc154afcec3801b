"""Particle-field save / restart (host logic only, no GPU): round trips,
append and time semantics, and the XDMF descriptor pointing at the right
bytes."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from flowunsteady_amd import vpm


def field(n=37, seed=3, cap=None):
    rng = np.random.default_rng(seed)
    pf = vpm.ParticleField(cap or n)
    for i in range(n):
        vpm.add_particle(pf, rng.normal(size=3), rng.normal(size=3), rng.uniform(0.01, 0.1),
                         vol=rng.uniform(1e-6, 1e-5), circulation=rng.uniform(),
                         static=bool(i % 5 == 0))
    pf.particles[vpm.C_INDEX, :n] = rng.normal(size=(3, n))
    pf.particles[vpm.U_INDEX, :n] = rng.normal(size=(3, n))
    pf.t, pf.nt = 0.125, 42
    return pf


KEPT = [vpm.X_INDEX, vpm.GAMMA_INDEX, vpm.SIGMA_INDEX, vpm.VOL_INDEX, vpm.CIRCULATION_INDEX,
        vpm.STATIC_INDEX, vpm.C_INDEX, vpm.U_INDEX]


def test_round_trip_is_bitwise(tmp_path):
    a = field()
    xmf = vpm.save(a, "run_pfield", path=str(tmp_path))
    assert xmf == os.path.join(str(tmp_path), "run_pfield.42.xmf")
    b = vpm.ParticleField(a.np)
    vpm.read(b, xmf, load_time=True)
    assert b.np == a.np and b.t == 0.125 and b.nt == 42
    for rows in KEPT:
        np.testing.assert_array_equal(b.particles[rows, :b.np], a.particles[rows, :a.np])


def test_overwrite_append_and_time(tmp_path):
    a = field(10, seed=1)
    f = vpm.save(a, "x", path=str(tmp_path), add_num=False, overwrite_time=7.5)
    b = field(5, seed=2, cap=20)
    old = b.particles[:, :5].copy()
    vpm.read(b, f, overwrite=False)           # append, keep time
    assert b.np == 15 and b.t == 0.125 and b.nt == 42
    np.testing.assert_array_equal(b.particles[:, :5], old)
    np.testing.assert_array_equal(b.particles[vpm.X_INDEX, 5:15], a.particles[vpm.X_INDEX, :10])
    vpm.read(b, f, load_time=True)            # replace, take the stamped time
    assert b.np == 10 and b.t == 7.5
    assert np.all(b.particles[:, 10:15] == 0.0)
    with pytest.raises(ValueError):
        vpm.read(vpm.ParticleField(3), f)      # does not fit


def test_xdmf_points_at_the_arrays(tmp_path):
    a = field(9)
    xmf = vpm.save(a, "v", path=str(tmp_path), num=3)
    root = ET.parse(xmf).getroot()
    grid = root.find("Domain/Grid")
    assert grid.find("Topology").get("NumberOfElements") == "9"
    assert float(grid.find("Time").get("Value")) == 0.125

    def data(item):
        raw = np.fromfile(os.path.join(str(tmp_path), item.text), dtype="<f8")
        dims = [int(d) for d in item.get("Dimensions").split()]
        k = int(item.get("Seek")) // 8
        return raw[k:k + int(np.prod(dims))].reshape(dims)

    np.testing.assert_array_equal(data(grid.find("Geometry/DataItem")), vpm.get_X(a))
    attrs = {at.get("Name"): at for at in grid.findall("Attribute")}
    np.testing.assert_array_equal(data(attrs["Gamma"].find("DataItem")), vpm.get_Gamma(a))
    np.testing.assert_array_equal(data(attrs["sigma"].find("DataItem")), vpm.get_sigma(a))
    assert attrs["Gamma"].get("AttributeType") == "Vector"

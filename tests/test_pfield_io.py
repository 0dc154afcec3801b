"""Particle-field save / restart (host logic only, no GPU): round trips,
append and time semantics, and the XDMF descriptor pointing at the right
bytes."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from flowunsteady_amd import h5, vpm


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
        assert item.get("Format") == "HDF"
        fname, dset = item.text.split(":")
        d = h5.read(os.path.join(str(tmp_path), fname))[dset.lstrip("/")]
        assert list(d.shape) == [int(x) for x in item.get("Dimensions").split()]
        return d

    np.testing.assert_array_equal(data(grid.find("Geometry/DataItem")), vpm.get_X(a))
    attrs = {at.get("Name"): at for at in grid.findall("Attribute")}
    np.testing.assert_array_equal(data(attrs["Gamma"].find("DataItem")), vpm.get_Gamma(a))
    np.testing.assert_array_equal(data(attrs["sigma"].find("DataItem")), vpm.get_sigma(a))
    assert attrs["Gamma"].get("AttributeType") == "Vector"
    conn = data(grid.find("Topology/DataItem"))
    np.testing.assert_array_equal(conn.reshape(-1, 3), [[1, 1, k] for k in range(9)])


def test_file_holds_the_flowvpm_datasets(tmp_path):
    """The HDF5 file holds the datasets FLOWVPM's save writes, with their
    shapes and types (particle-major np x 3 for vectors, 1-based index)."""
    a = field(6)
    xmf = vpm.save(a, "f", path=str(tmp_path), add_num=False)
    d = h5.read(xmf[:-4] + ".h5")
    assert d["np"].shape == () and d["np"].dtype == np.int64 and int(d["np"]) == 6
    assert int(d["nt"]) == 42 and float(d["t"]) == 0.125
    for k in ("X", "Gamma"):
        assert d[k].shape == (6, 3) and d[k].dtype == np.float64
    for k in ("sigma", "circulation", "vol", "static"):
        assert d[k].shape == (6,) and d[k].dtype == np.float64
    np.testing.assert_array_equal(d["i"], np.arange(1, 7))
    np.testing.assert_array_equal(d["static"], a.particles[vpm.STATIC_INDEX, :6])


H5DUMP = "/opt/conda/bin/h5dump"


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="the HDF5 tools are not installed")
def test_the_hdf5_library_reads_the_file(tmp_path):
    """h5dump (the HDF5 library's own reader) parses the file: every dataset
    with its type and shape, and the values to full precision."""
    import subprocess
    a = field(7)
    xmf = vpm.save(a, "lib", path=str(tmp_path), add_num=False)
    f = xmf[:-4] + ".h5"
    head = subprocess.run([H5DUMP, "-H", f], capture_output=True, text=True, check=True).stdout
    for k in ("X", "Gamma", "sigma", "circulation", "vol", "static", "C", "U"):
        assert f'DATASET "{k}"' in head
    assert head.count("H5T_IEEE_F64LE") == 9 and head.count("H5T_STD_I64LE") == 4
    assert head.count("SCALAR") == 3 and "( 7, 3 )" in head and "( 21 )" in head

    def values(name):
        out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", "/" + name, "-y", "-w", "0", f],
                             capture_output=True, text=True, check=True).stdout
        body = out[out.index("DATA {") + 6: out.rindex("}", 0, out.rindex("}"))]
        return np.array([float(v) for v in body.replace(",", " ").split() if v != "}"])

    np.testing.assert_array_equal(values("X"), vpm.get_X(a).reshape(-1))
    np.testing.assert_array_equal(values("sigma"), vpm.get_sigma(a))
    assert values("t")[0] == 0.125


@pytest.mark.parametrize("kind", ["contiguous", "chunked", "latest"])
def test_reads_files_written_by_the_hdf5_library(golden_dir, kind):
    """Fixtures written by the HDF5 library (tests/golden/make_h5_fixtures.c):
    the default file format with contiguous and with chunked datasets, and
    the latest format (superblock 3, version-2 object headers, links in a
    fractal heap).  Every value is known in closed form."""
    pf = vpm.ParticleField(8)
    vpm.read(pf, os.path.join(golden_dir, f"pfield_{kind}.h5"), load_time=True)
    k = np.arange(5)
    assert pf.np == 5 and pf.nt == 12 and pf.t == 0.375
    X = np.array([[kk + c / 10.0 for c in range(3)] for kk in k])
    p = np.array([0.1, 0.1 / 10.0, 0.1 / 10.0 / 10.0])
    np.testing.assert_array_equal(vpm.get_X(pf), X)
    np.testing.assert_array_equal(vpm.get_Gamma(pf), -(k[:, None] + 1) * p[None, :])
    np.testing.assert_array_equal(vpm.get_sigma(pf), 0.01 * (k + 1))
    np.testing.assert_array_equal(pf.particles[vpm.VOL_INDEX, :5], 1e-3 / (k + 1))
    np.testing.assert_array_equal(pf.particles[vpm.CIRCULATION_INDEX, :5], 0.5 * k)
    np.testing.assert_array_equal(pf.particles[vpm.STATIC_INDEX, :5], k % 2)
    d = h5.read(os.path.join(golden_dir, f"pfield_{kind}.h5"))
    np.testing.assert_array_equal(d["connectivity"].reshape(-1, 3), [[1, 1, kk] for kk in k])
    assert d["i"].dtype == np.int64 and d["np"].shape == ()


def test_h5_module_round_trip_of_types_and_shapes(tmp_path):
    f = str(tmp_path / "t.h5")
    data = {"s64": np.float64(2.5), "i32": np.int32([1, -2, 3]), "u8": np.uint8([[1, 2], [3, 4]]),
            "f32": np.float32([0.5, 1.5]), "b": np.array([True, False]),
            "empty": np.zeros((0, 3)), "big": np.arange(10000, dtype=np.float64)}
    h5.write(f, data)
    r = h5.read(f)
    for k, v in data.items():
        v = np.asarray(v)
        if v.dtype == bool:
            v = v.astype(np.uint8)
        assert r[k].dtype == v.dtype and r[k].shape == v.shape, k
        np.testing.assert_array_equal(r[k], v)
    bad = tmp_path / "not.h5"
    bad.write_bytes(b"x" * 64)
    with pytest.raises(h5.FormatError):
        h5.read(str(bad))


def test_reads_the_earlier_raw_binary_form(tmp_path):
    """Files of the earlier XDMF + raw binary + JSON sidecar form still load."""
    import json
    a = field(4)
    n = a.np
    blobs, arrays, off = [], {}, 0
    for name, rows, nc in (("X", vpm.X_INDEX, 3), ("Gamma", vpm.GAMMA_INDEX, 3),
                           ("sigma", vpm.SIGMA_INDEX, 1)):
        v = np.ascontiguousarray(a.particles[rows, :n].T if nc > 1 else a.particles[rows, :n])
        arrays[name] = [off, nc]
        blobs.append(v.astype("<f8").tobytes())
        off += v.nbytes
    (tmp_path / "old.bin").write_bytes(b"".join(blobs))
    (tmp_path / "old.json").write_text(json.dumps({"format": "vpmx-pfield-1", "np": n, "nt": 5,
                                                   "t": 1.5, "arrays": arrays,
                                                   "bin": "old.bin"}))
    b = vpm.ParticleField(n)
    vpm.read(b, str(tmp_path / "old.json"), load_time=True)
    assert b.np == n and b.nt == 5 and b.t == 1.5
    np.testing.assert_array_equal(vpm.get_X(b), vpm.get_X(a))
    np.testing.assert_array_equal(vpm.get_sigma(b), vpm.get_sigma(a))

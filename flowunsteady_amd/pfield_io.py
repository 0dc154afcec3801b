"""Particle-field save and restart: the `vpm.save` / `vpm.read!` pair that
FLOWUnsteady calls to write the field every `nsteps_save` steps and to restart
from a file (src/FLOWUnsteady_simulation.jl:264,438;
src/FLOWUnsteady_processing_force.jl:796).

Files, as FLOWVPM writes them (`.{num}` only with add_num=True):

    {path}/{name}.{num}.h5    HDF5, flat datasets in the root group:
                              np, nt (int64 scalars), t (float64 scalar),
                              X, Gamma (np x 3), sigma, circulation, vol,
                              static (np, float64), i (np, int64, 1-based),
                              connectivity (3 np, int64: 1, 1, k per particle)
                              and, for a restart of this library's state,
                              C (np x 3: SFS coefficient, its averaged
                              numerator and denominator) and U (np x 3)
    {path}/{name}.{num}.xmf   XDMF 3 descriptor of the HDF5 datasets
                              (Format="HDF", file.h5:/dataset), which ParaView
                              opens (the `_pfield...xmf` files of
                              docs/src/examples/*.md)

Arrays are written particle-major: a 3 x np column-per-particle block of the
particle matrix becomes an np x 3 dataset, which is how a Julia 3 x np matrix
lands in HDF5 (column-major storage = row-major np x 3).  The HDF5 container is
written and read by `flowunsteady_amd.h5`; `read` also accepts files written
by the HDF5 library (contiguous or chunked datasets, old or new file format)
and, for files of this library's earlier XDMF + raw binary + JSON form, their
JSON sidecar.
"""
from __future__ import annotations

import json
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import h5

LEGACY_FORMAT = "vpmx-pfield-1"


def _arrays():
    """(dataset name, row slice or index of `particles`, components)."""
    from . import vpm
    return [("X", vpm.X_INDEX, 3), ("Gamma", vpm.GAMMA_INDEX, 3),
            ("sigma", vpm.SIGMA_INDEX, 1), ("circulation", vpm.CIRCULATION_INDEX, 1),
            ("vol", vpm.VOL_INDEX, 1), ("static", vpm.STATIC_INDEX, 1),
            ("C", vpm.C_INDEX, 3), ("U", vpm.U_INDEX, 3)]


def _base(file_name, path, add_num, num, nt):
    name = file_name
    if add_num:
        name += f".{nt if num is None else num}"
    return os.path.join(path, name) if path else name


def save(pfield, file_name: str, *, path: str = "", add_num: bool = True, num=None,
         overwrite_time=None) -> str:
    """Write the particles of `pfield` (FLOWVPM's `save(pfield, file_name; path,
    add_num, num, overwrite_time)`): the step number `num` (default pfield.nt)
    is appended to the name when add_num, and the time stamp is overwrite_time
    when given, else pfield.t.  Returns the XDMF file name."""
    if path:
        os.makedirs(path, exist_ok=True)
    base = _base(file_name, path, add_num, num, pfield.nt)
    n = pfield.np
    P = pfield.particles
    t = float(pfield.t if overwrite_time is None else overwrite_time)
    data = {"np": np.int64(n), "nt": np.int64(pfield.nt), "t": np.float64(t),
            "i": np.arange(1, n + 1, dtype=np.int64),
            "connectivity": np.stack([np.ones(n, np.int64), np.ones(n, np.int64),
                                      np.arange(n, dtype=np.int64)], axis=1).reshape(-1)}
    for name, rows, nc in _arrays():
        data[name] = np.array(P[rows, :n].T if nc > 1 else P[rows, :n], dtype="<f8", order="C")
    h5.write(base + ".h5", data)
    _write_xdmf(base, n, t, data)
    return base + ".xmf"


def _write_xdmf(base, n, t, data):
    h5f = os.path.basename(base + ".h5")
    X = ET.Element("Xdmf", Version="3.0")
    dom = ET.SubElement(X, "Domain")
    grid = ET.SubElement(dom, "Grid", Name="particles", GridType="Uniform")
    ET.SubElement(grid, "Time", Value=repr(t))
    top = ET.SubElement(grid, "Topology", TopologyType="Mixed", NumberOfElements=str(n))

    def item(parent, name, number_type="Float"):
        a = data[name]
        d = ET.SubElement(parent, "DataItem", Format="HDF", NumberType=number_type,
                          Precision=str(a.dtype.itemsize), Dimensions=" ".join(map(str, a.shape)))
        d.text = f"{h5f}:/{name}"

    item(top, "connectivity", "Int")
    geo = ET.SubElement(grid, "Geometry", GeometryType="XYZ")
    item(geo, "X")
    for name, _, nc in _arrays():
        if name == "X":
            continue
        at = ET.SubElement(grid, "Attribute", Name=name, Center="Node",
                           AttributeType="Vector" if nc == 3 else "Scalar")
        item(at, name)
    ET.ElementTree(X).write(base + ".xmf", xml_declaration=True, encoding="utf-8")


def _stem(file_name):
    stem = file_name
    for ext in (".xmf", ".h5", ".json", ".bin"):
        if stem.endswith(ext):
            return stem[: -len(ext)]
    return stem


def _load(file_name):
    """{name: array} of the particle datasets and (np, nt, t) of a file."""
    stem = _stem(file_name)
    if os.path.exists(stem + ".h5"):
        d = h5.read(stem + ".h5")
        for k in ("np", "X", "Gamma", "sigma"):
            if k not in d:
                raise ValueError(f"{stem}.h5: no dataset '{k}' (not a particle-field file)")
        n = int(np.asarray(d["np"]).reshape(-1)[0])
        nt = int(np.asarray(d["nt"]).reshape(-1)[0]) if "nt" in d else 0
        t = float(np.asarray(d["t"]).reshape(-1)[0]) if "t" in d else 0.0
        return d, n, nt, t
    if os.path.exists(stem + ".json"):                   # earlier raw-binary form
        with open(stem + ".json") as f:
            meta = json.load(f)
        if meta.get("format") != LEGACY_FORMAT:
            raise ValueError(f"{stem}.json: not a {LEGACY_FORMAT} file")
        n = int(meta["np"])
        raw = np.fromfile(os.path.join(os.path.dirname(stem), meta["bin"]), dtype="<f8")
        d = {}
        for name, (off, nc) in meta["arrays"].items():
            a = raw[off // 8: off // 8 + n * nc]
            d[name] = a.reshape(n, nc) if nc > 1 else a
        return d, n, int(meta["nt"]), float(meta["t"])
    raise FileNotFoundError(f"no {stem}.h5 (or .json) particle-field file")


def read(pfield, file_name: str, *, overwrite: bool = True, load_time: bool = False):
    """Load a particle-field file (FLOWVPM's `read!(pfield, file; overwrite,
    load_time)`): `file_name` is the .h5 or .xmf name or their common stem.
    overwrite=True replaces the particles, False appends them; load_time=True
    also takes t and nt from the file.  Raises ValueError when the particles
    do not fit in `pfield`."""
    d, n, nt, t = _load(file_name)
    first = 0 if overwrite else pfield.np
    if first + n > pfield.maxparticles:
        raise ValueError(f"{n} particles do not fit: {first} held, capacity {pfield.maxparticles}")
    P = pfield.particles
    if overwrite:
        P[:, :pfield.np] = 0.0
    sl = slice(first, first + n)
    for name, rows, nc in _arrays():
        if name not in d:
            continue
        a = np.asarray(d[name], dtype=np.float64)
        if nc > 1:
            if a.shape != (n, nc):
                raise ValueError(f"{name}: shape {a.shape} in the file, ({n}, {nc}) expected")
            P[rows, sl] = a.T
        else:
            if a.shape != (n,):
                raise ValueError(f"{name}: shape {a.shape} in the file, ({n},) expected")
            P[rows, sl] = a
    pfield.np = first + n
    if load_time:
        pfield.t = t
        pfield.nt = nt
    return pfield

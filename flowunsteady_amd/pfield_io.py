"""Particle-field save and restart: the `vpm.save` / `vpm.read!` pair that
FLOWUnsteady calls to write the field every `nsteps_save` steps and to restart
from a file (src/FLOWUnsteady_simulation.jl:264,438;
src/FLOWUnsteady_processing_force.jl:796).

FLOWVPM writes HDF5 with an XDMF descriptor.  This build has no HDF5 library,
so the same content goes to a raw little-endian binary file described by an
XDMF 3 file (`Format="Binary"` data items, which ParaView reads like the HDF5
ones) and a JSON sidecar holding the field state a restart needs.  Files:

    {path}/{name}.{num}.xmf   XDMF: Polyvertex topology, XYZ geometry, attributes
    {path}/{name}.{num}.bin   the arrays, one after another, float64, particle-major
    {path}/{name}.{num}.json  format tag, np, nt, t, array offsets

(`.{num}` only with add_num=True.)
"""
from __future__ import annotations

import json
import os
import xml.etree.ElementTree as ET

import numpy as np

FORMAT = "vpmx-pfield-1"

# (name in the files, row slice or index of `particles`, components)
_ARRAYS = None


def _arrays():
    global _ARRAYS
    if _ARRAYS is None:
        from . import vpm
        _ARRAYS = [("X", vpm.X_INDEX, 3), ("Gamma", vpm.GAMMA_INDEX, 3),
                   ("sigma", vpm.SIGMA_INDEX, 1), ("vol", vpm.VOL_INDEX, 1),
                   ("circulation", vpm.CIRCULATION_INDEX, 1), ("static", vpm.STATIC_INDEX, 1),
                   ("C", vpm.C_INDEX, 3), ("U", vpm.U_INDEX, 3)]
    return _ARRAYS


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
    offsets, blobs, off = {}, [], 0
    for name, rows, nc in _arrays():
        a = np.ascontiguousarray(P[rows, :n].T if nc > 1 else P[rows, :n], dtype="<f8")
        offsets[name] = [off, nc]
        blobs.append(a.tobytes())
        off += a.nbytes
    with open(base + ".bin", "wb") as f:
        for b in blobs:
            f.write(b)
    t = pfield.t if overwrite_time is None else overwrite_time
    meta = {"format": FORMAT, "np": n, "nt": int(pfield.nt), "t": float(t),
            "arrays": offsets, "bin": os.path.basename(base + ".bin")}
    with open(base + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    _write_xdmf(base, meta)
    return base + ".xmf"


def _write_xdmf(base, meta):
    binf, n = meta["bin"], meta["np"]
    X = ET.Element("Xdmf", Version="3.0")
    dom = ET.SubElement(X, "Domain")
    grid = ET.SubElement(dom, "Grid", Name="particles", GridType="Uniform")
    ET.SubElement(grid, "Time", Value=repr(meta["t"]))
    ET.SubElement(grid, "Topology", TopologyType="Polyvertex", NumberOfElements=str(n))

    def item(parent, name):
        off, nc = meta["arrays"][name]
        dims = f"{n} {nc}" if nc > 1 else f"{n}"
        d = ET.SubElement(parent, "DataItem", Format="Binary", Endian="Little",
                          NumberType="Float", Precision="8", Seek=str(off), Dimensions=dims)
        d.text = binf

    geo = ET.SubElement(grid, "Geometry", GeometryType="XYZ")
    item(geo, "X")
    for name, _, nc in _arrays():
        if name == "X":
            continue
        at = ET.SubElement(grid, "Attribute", Name=name, Center="Node",
                           AttributeType="Vector" if nc == 3 else "Scalar")
        item(at, name)
    ET.ElementTree(X).write(base + ".xmf", xml_declaration=True, encoding="utf-8")


def read(pfield, file_name: str, *, overwrite: bool = True, load_time: bool = False):
    """Load a field written by `save` (FLOWVPM's `read!(pfield, file;
    overwrite, load_time)`): `file_name` is the .xmf, .json or .bin name or
    their common stem.  overwrite=True replaces the particles, False appends
    them; load_time=True also takes t and nt from the file.  Raises ValueError
    when the particles do not fit in `pfield`."""
    stem = file_name
    for ext in (".xmf", ".json", ".bin"):
        if stem.endswith(ext):
            stem = stem[: -len(ext)]
    with open(stem + ".json") as f:
        meta = json.load(f)
    if meta.get("format") != FORMAT:
        raise ValueError(f"{stem}.json: not a {FORMAT} file")
    n = int(meta["np"])
    first = 0 if overwrite else pfield.np
    if first + n > pfield.maxparticles:
        raise ValueError(f"{n} particles do not fit: {first} held, capacity {pfield.maxparticles}")
    raw = np.fromfile(os.path.join(os.path.dirname(stem), meta["bin"]), dtype="<f8")
    P = pfield.particles
    if overwrite:
        P[:, :pfield.np] = 0.0
    sl = slice(first, first + n)
    for name, rows, nc in _arrays():
        if name not in meta["arrays"]:
            continue
        off, nc_file = meta["arrays"][name]
        if nc_file != nc:
            raise ValueError(f"{name}: {nc_file} components in the file, {nc} expected")
        a = raw[off // 8: off // 8 + n * nc]
        if nc > 1:
            P[rows, sl] = a.reshape(n, nc).T
        else:
            P[rows, sl] = a
    pfield.np = first + n
    if load_time:
        pfield.t = float(meta["t"])
        pfield.nt = int(meta["nt"])
    return pfield

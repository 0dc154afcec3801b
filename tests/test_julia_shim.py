"""The Julia binding (julia/VPMX.jl) against the C ABI (include/vpmx.h).

Julia is not installed here, so the shim is checked as text: every struct it
passes by reference has the fields of its C typedef in the same order with
the same types (Julia then lays it out as the C compiler does; the layout is
also rebuilt with ctypes and compared with the header's offsets), and every
ccall names an exported entry point with the header's number of arguments.
"""
import ctypes
import os
import re

import pytest

from flowunsteady_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "julia", "VPMX.jl")
HEADER = os.path.join(ROOT, "include", "vpmx.h")

JULIA_TO_C = {"Int32": "int32_t", "Float64": "double", "NTuple{3,Float64}": "double[3]"}
C_TO_CTYPES = {"int32_t": ctypes.c_int32, "double": ctypes.c_double,
               "double[3]": ctypes.c_double * 3}


def header_text():
    return re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)


def header_structs():
    """{typedef name: [(field, ctype)]} of the header's typedef structs."""
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", header_text(), re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"(int32_t|double)\s+(.*)", decl)
            assert m, decl
            ctype, names = m.groups()
            for nm in names.split(","):
                nm = nm.strip()
                a = re.match(r"(\w+)\[(\d+)\]", nm)
                fields.append((a.group(1), f"{ctype}[{a.group(2)}]") if a else (nm, ctype))
        out[name] = fields
    return out


def shim_structs():
    """{C typedef the struct mirrors: (julia name, [(field, ctype)])}."""
    txt = open(SHIM).read()
    out = {}
    for jname, cname, body in re.findall(r"^struct (\w+)\s*#\s*(vpmx_\w+)\n(.*?)^end", txt,
                                         re.S | re.M):
        fields = []
        for line in body.splitlines():
            line = line.split("#")[0].strip()
            if not line:
                continue
            f, t = [x.strip() for x in line.split("::")]
            fields.append((f, JULIA_TO_C[t.replace(" ", "")]))
        out[cname] = (jname, fields)
    return out


def header_arity():
    """{entry point: number of parameters} from the header's prototypes."""
    out = {}
    for name, args in re.findall(r"\b(vpmx_\w+)\s*\(([^;{]*?)\)\s*;", header_text(), re.S):
        args = args.strip()
        out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def _matching(txt, i):
    """Index of the parenthesis closing the one at txt[i]."""
    depth = 0
    for j in range(i, len(txt)):
        if txt[j] in "({":
            depth += 1
        elif txt[j] in ")}":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced")


def _top_level_count(s):
    s = s.strip()
    if not s:
        return 0
    depth, n = 0, 1
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        elif ch == "," and depth == 0:
            n += 1
    return n - (1 if s.rstrip().endswith(",") else 0)


def shim_ccalls():
    """[(entry point, number of argument types)] of every ccall in the shim."""
    txt = open(SHIM).read()
    out = []
    for m in re.finditer(r"ccall\(\(:(vpmx_\w+),\s*lib\),\s*\w+,\s*\(", txt):
        i = m.end() - 1
        j = _matching(txt, i)
        out.append((m.group(1), _top_level_count(txt[i + 1:j])))
    return out


def test_shim_mirrors_every_struct_it_passes():
    hs, js = header_structs(), shim_structs()
    assert set(hs) == {"vpmx_layout", "vpmx_uj_params", "vpmx_step_params"}
    assert set(js) == set(hs)
    for cname, (jname, fields) in js.items():
        assert fields == hs[cname], (cname, jname)


def test_shim_struct_layout_equals_the_header_layout():
    """The shim's fields rebuilt as a C struct have the sizes and offsets of
    the ctypes mirror, which test_native_cpu pins to gcc's layout of the header."""
    mirrors = {"vpmx_layout": N.Layout, "vpmx_uj_params": N.UJParams,
               "vpmx_step_params": N.StepParams}
    for cname, (jname, fields) in shim_structs().items():
        S = type(jname, (ctypes.Structure,),
                 {"_fields_": [(f, C_TO_CTYPES[t]) for f, t in fields]})
        ref = mirrors[cname]
        assert ctypes.sizeof(S) == ctypes.sizeof(ref), cname
        for f, _ in fields:
            assert getattr(S, f).offset == getattr(ref, f).offset, (cname, f)


def test_shim_ccalls_match_header_prototypes():
    arity = header_arity()
    calls = shim_ccalls()
    assert len(calls) >= 10
    L = N.lib()
    for name, n in calls:
        assert name in arity, name
        assert hasattr(L, name), name
        assert n == arity[name], (name, n, arity[name])
    names = {c[0] for c in calls}
    for need in ("vpmx_uj", "vpmx_rk3_step", "vpmx_euler_step", "vpmx_uj_targets",
                 "vpmx_upload", "vpmx_batch_rk3_step", "vpmx_create", "vpmx_last_error"):
        assert need in names, need


def test_shim_fills_every_step_setting():
    """The StepParams constructor call passes one value per field (viscous
    settings included), in a single call site."""
    txt = open(SHIM).read()
    i = txt.index("return StepParams(") + len("return StepParams")
    j = _matching(txt, i)
    nfields = len(shim_structs()["vpmx_step_params"][1])
    assert _top_level_count(txt[i + 1:j]) == nfields


@pytest.mark.parametrize("needle", ["sigma_cut(fmm) = prop(fmm, :nonzero_sigma, false) ? -1.0 : 0.0"])
def test_shim_sigma_cut_follows_the_header(needle):
    """sigma_cut 0 = geometric (FLOWVPM's nonzero_sigma = false), < 0 = the
    kernel's default, as vpmx_uj_params documents."""
    assert needle in open(SHIM).read()
    assert "0: the purely" in open(HEADER).read() and "< 0: the" in open(HEADER).read()

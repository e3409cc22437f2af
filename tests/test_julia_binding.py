"""The Julia binding (trixilw.jl_amd/julia/TrixiLWHip.jl) against the C header
and the ctypes mirror, without Julia: the structs' field order, types,
offsets and sizes follow C layout rules and must equal the header's typedefs
and the ctypes structures; every function the header declares has a ccall
with the header's argument count, argument types and return type."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "trixilw_hip.h"
JULIA = ROOT / "trixilw.jl_amd" / "julia" / "TrixiLWHip.jl"

# size and alignment of the scalar types on x86-64 / gfx950 hosts
C_SCALAR = {"int32_t": (4, 4), "int": (4, 4), "double": (8, 8), "int64_t": (8, 8), "size_t": (8, 8)}
JL_SCALAR = {"Int32": (4, 4), "Cint": (4, 4), "Float64": (8, 8), "Int64": (8, 8), "Csize_t": (8, 8)}

STRUCTS = {"tlw_mesh_desc": "MeshDesc", "tlw_basis_desc": "BasisDesc", "tlw_solver_desc": "SolverDesc",
           "tlw_parabolic_desc": "ParabolicDesc", "tlw_curved_desc": "CurvedDesc", "tlw_mortar_desc": "MortarDesc",
           "tlw_p4est_desc": "P4estDesc"}


def _strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", "", s, flags=re.S)


def _layout(fields):
    """[(name, size, align, count)] -> ([(name, offset, size·count)], sizeof) with C rules."""
    off, amax, out = 0, 1, []
    for name, size, align, count in fields:
        off = (off + align - 1) // align * align
        out.append((name, off, size * count))
        off += size * count
        amax = max(amax, align)
    return out, (off + amax - 1) // amax * amax


def header_structs():
    src = _strip_c_comments(HEADER.read_text())
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(\w+);", src, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(?:const )?(\w+)\s*(.+)$", decl)
            assert m, decl
            ty, declarators = m.groups()
            for d in declarators.split(","):  # "double alpha_max, alpha_min"
                dm = re.match(r"\s*(\*?)\s*(\w+)(?:\[(\d+)\])?\s*$", d)
                assert dm, decl
                ptr, fname, cnt = dm.groups()
                size, align = (8, 8) if ptr else C_SCALAR[ty]
                fields.append((fname, size, align, int(cnt or 1)))
        out[name] = fields
    return out


def julia_structs():
    src = JULIA.read_text()
    out = {}
    for name, body in re.findall(r"^struct (\w+)\n(.*?)^end", src, flags=re.S | re.M):
        fields = []
        for line in body.splitlines():
            line = line.split("#")[0].strip()
            if not line:
                continue
            m = re.match(r"(\w+)::(.+)$", line)
            assert m, line
            fname, ty = m.groups()
            if ty.startswith("Ptr{"):
                fields.append((fname, 8, 8, 1))
                continue
            t = re.match(r"NTuple\{(\d+),\s*(\w+)\}$", ty)
            if t:
                size, align = JL_SCALAR[t.group(2)]
                fields.append((fname, size, align, int(t.group(1))))
                continue
            size, align = JL_SCALAR[ty]
            fields.append((fname, size, align, 1))
        out[name] = fields
    return out


def ctypes_struct(cls):
    return [(f, getattr(cls, f).offset, getattr(cls, f).size) for f, _ in cls._fields_], C.sizeof(cls)


def test_struct_layouts_agree():
    import trixilw_hip._lib as L

    hs, js = header_structs(), julia_structs()
    mirror = {"MeshDesc": L.MeshDesc, "BasisDesc": L.BasisDesc, "SolverDesc": L.SolverDesc,
              "ParabolicDesc": L.ParabolicDesc, "CurvedDesc": L.CurvedDesc, "MortarDesc": L.MortarDesc, "P4estDesc": L.P4estDesc}
    assert set(hs) == set(STRUCTS)
    for cname, jname in STRUCTS.items():
        h_lay, h_size = _layout(hs[cname])
        j_lay, j_size = _layout(js[jname])
        c_lay, c_size = ctypes_struct(mirror[jname])
        assert [f for f, _, _ in h_lay] == [f for f, _, _ in j_lay] == [f for f, _, _ in c_lay], cname
        assert h_lay == j_lay == c_lay, cname
        assert h_size == j_size == c_size, cname


C_ARG = {
    "tlw_ctx*": {"Ptr{Cvoid}"}, "tlw_ctx**": {"Ref{Ptr{Cvoid}}"}, "double": {"Float64"},
    "tlw_ensemble*": {"Ptr{Cvoid}"}, "tlw_ensemble**": {"Ref{Ptr{Cvoid}}"},
    "int32_t": {"Int32", "Cint"}, "int64_t": {"Int64"}, "size_t": {"Csize_t"}, "void*": {"Ptr{Cvoid}"},
    "double*": {"Ptr{Float64}", "Ref{Float64}"}, "double**": {"Ref{Ptr{Float64}}"},
    "int64_t*": {"Ref{Int64}", "Ptr{Int64}"}, "int32_t*": {"Ref{Int32}", "Ptr{Int32}"},
}
C_RET = {"int": "Cint", "void": "Cvoid", "const char*": "Cstring", "size_t": "Csize_t"}


def header_functions():
    src = _strip_c_comments(HEADER.read_text())
    out = {}
    for ret, name, args in re.findall(r"^(int|void|const char\*|size_t)\s+(tlw_\w+)\((.*?)\);", src,
                                      flags=re.S | re.M):
        types = []
        for a in " ".join(args.split()).split(","):
            a = a.strip().replace("const ", "")
            m = re.match(r"(\w+)\s*(\**)\s*\w+$", a)
            assert m, (name, a)
            ty, stars = m.groups()
            if ty in STRUCTS:
                types.append(("struct", ty, stars))
            else:
                types.append(ty + stars)
        out[name] = (ret, types)
    return out


def _split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def julia_ccalls():
    src = JULIA.read_text()
    out = {}
    for name, ret, args in re.findall(r"ccall\(\(:(tlw_\w+),\s*lib\),\s*(\w+),\s*\((.*?)\),", src, flags=re.S):
        out.setdefault(name, []).append((ret, _split_top(" ".join(args.split()).rstrip(","))))
    return out


def test_every_entry_point_has_a_matching_ccall():
    hf, jc = header_functions(), julia_ccalls()
    import trixilw_hip._lib as L

    assert set(hf) == set(L.EXPORTED), sorted(set(hf) ^ set(L.EXPORTED))
    missing = sorted(set(hf) - set(jc))
    assert not missing, missing
    for name, (ret, ctypes_) in hf.items():
        for jret, jargs in jc[name]:
            assert jret == C_RET[ret], (name, jret)
            assert len(jargs) == len(ctypes_), (name, jargs)
            for ct, jt in zip(ctypes_, jargs):
                if isinstance(ct, tuple):  # a description struct passed by pointer
                    allowed = {f"Ref{{{STRUCTS[ct[1]]}}}", f"Ptr{{{STRUCTS[ct[1]]}}}"}
                else:
                    allowed = C_ARG[ct]
                assert jt in allowed, (name, ct, jt)


@pytest.mark.parametrize("field", ["reconstruction"])
def test_solver_desc_carries_every_header_field(field):
    assert field in [f for f, *_ in julia_structs()["SolverDesc"]]

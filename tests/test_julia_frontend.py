"""The Julia front end (gradpower.jl_amd/julia/GradPowerHIP.jl) against the C
ABI, on the CPU (no Julia in this image, so the module is checked as text):
every struct it passes through ccall has the fields of the header's struct in
the same order with the same sizes and offsets (C layout rules, which Julia's
isbits structs follow), the ctypes mirrors used by the Python front end agree
too, every symbol it calls is declared in the header and exported by the
library, and its method and policy constants equal the header's enums."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gradpower_hip.h")
JL = os.path.join(ROOT, "gradpower.jl_amd", "julia", "GradPowerHIP.jl")
PAIRS = {"gp_system_desc": "GpSystemDesc", "gp_event": "GpEvent", "gp_line_trip": "GpLineTrip",
         "gp_load_drop": "GpLoadDrop", "gp_device_drop": "GpDeviceDrop", "gp_options": "GpOptions",
         "gp_stats": "GpStats",
         "gp_objective": "GpObjective", "gp_adjoint_stats": "GpAdjointStats",
         "gp_symbolic_info": "GpSymbolicInfo"}
C_SCALARS = {"int32_t": 4, "int64_t": 8, "double": 8, "uint8_t": 1}
JL_SCALARS = {"Int32": 4, "Int64": 8, "Float64": 8, "UInt8": 1, "Cint": 4}


def _strip_c_comments(t):
    t = re.sub(r"/\*.*?\*/", "", t, flags=re.S)
    return re.sub(r"//[^\n]*", "", t)


def _c_structs():
    t = _strip_c_comments(open(HEADER).read())
    consts = {"GP_NDEVTYPES": 8}
    out = {}
    for name, body in ((m.group(2), m.group(1)) for m in
                       re.finditer(r"typedef struct \w+ \{(.*?)\}\s*(\w+);", t, flags=re.S)):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            decl = decl.replace("const ", "")
            base, rest = decl.split(None, 1)
            for item in rest.split(","):
                item = item.strip()
                ptr = item.startswith("*")
                item = item.lstrip("*").strip()
                arr = 1
                mm = re.match(r"(\w+)\[(\w+)\]", item)
                if mm:
                    item, dim = mm.group(1), mm.group(2)
                    arr = int(consts.get(dim, dim))
                size = 8 if ptr else C_SCALARS[base]
                fields.append((item, size, arr))
        out[name] = fields
    return out


def _jl_structs():
    t = open(JL).read()
    out = {}
    for m in re.finditer(r"^struct (Gp\w+)\n(.*?)^end", t, flags=re.S | re.M):
        fields = []
        for line in m.group(2).splitlines():
            line = line.split("#")[0].strip()
            if not line:
                continue
            name, ty = [x.strip() for x in line.split("::")]
            mt = re.match(r"NTuple\{(\d+),\s*(.+)\}$", ty)
            arr = 1
            if mt:
                arr, ty = int(mt.group(1)), mt.group(2)
            size = 8 if ty.startswith("Ptr{") else JL_SCALARS[ty]
            fields.append((name, size, arr))
        out[m.group(1)] = fields
    return out


def _layout(fields):
    """(name, offset, size) per field and the struct size, C rules (natural
    alignment, trailing padding to the largest member)."""
    off, amax, lay = 0, 1, []
    for name, size, arr in fields:
        off = (off + size - 1) // size * size
        lay.append((name, off, size * arr))
        off += size * arr
        amax = max(amax, size)
    return lay, (off + amax - 1) // amax * amax


def test_julia_structs_match_header():
    cs, js = _c_structs(), _jl_structs()
    assert set(PAIRS) <= set(cs), sorted(set(PAIRS) - set(cs))
    for cname, jname in PAIRS.items():
        assert jname in js, jname
        assert _layout(js[jname]) == _layout(cs[cname]), cname


def test_python_ctypes_mirrors_match_header():
    import gradpower_hip.desc as d
    import gradpower_hip.lib as lib
    mirrors = {"gp_system_desc": d.CSystemDesc, "gp_event": d.CEvent, "gp_line_trip": d.CLineTrip,
               "gp_load_drop": d.CLoadDrop, "gp_device_drop": d.CDeviceDrop, "gp_options": lib.COptions,
               "gp_stats": lib.CStats,
               "gp_objective": lib.CObjective, "gp_adjoint_stats": lib.CAdjStats,
               "gp_symbolic_info": lib.CSymInfo}
    cs = _c_structs()
    for cname, cls in mirrors.items():
        lay, size = _layout(cs[cname])
        assert C.sizeof(cls) == size, cname
        for name, off, _ in lay:
            assert getattr(cls, name).offset == off, (cname, name)


def test_julia_ccalls_are_declared_and_exported():
    from gradpower_hip import lib
    t = open(JL).read()
    called = set(re.findall(r"ccall\(\(:(\w+), LIB\)", t))
    assert {"gp_create", "gp_destroy", "gp_set_state", "gp_set_events", "gp_integrate", "gp_get_trajectory",
            "gp_adjoint"} <= called
    hdr = _strip_c_comments(open(HEADER).read())
    declared = set(re.findall(r"\b(gp_\w+)\s*\(", hdr))
    assert called <= declared, sorted(called - declared)
    so = lib.load()
    for s in called:
        assert hasattr(so, s), s


def test_julia_constants_match_header_enums():
    hdr = _strip_c_comments(open(HEADER).read())
    enum = dict((k, int(v)) for k, v in re.findall(r"\b(GP_(?:METHOD|REFACTOR|ERR)_\w+)\s*=\s*(-?\d+)", hdr))
    t = open(JL).read()
    m = re.search(r"const METHOD_MONO_FP64, METHOD_SCHUR_FP64, METHOD_SCHUR_FP32, METHOD_SCHUR_SHARED = "
                  r"Int32\((\d)\), Int32\((\d)\), Int32\((\d)\), Int32\((\d)\)", t)
    assert m and [int(x) for x in m.groups()] == [enum["GP_METHOD_MONO_FP64"], enum["GP_METHOD_SCHUR_FP64"],
                                                  enum["GP_METHOD_SCHUR_FP32"], enum["GP_METHOD_SCHUR_SHARED"]]
    assert "REFACTOR_EVERY_ITER, REFACTOR_FIRST_ITER, REFACTOR_ADAPTIVE, REFACTOR_ASYNC = Int32.(0:3)" in t
    assert [enum[k] for k in ("GP_REFACTOR_EVERY_ITER", "GP_REFACTOR_FIRST_ITER", "GP_REFACTOR_ADAPTIVE",
                              "GP_REFACTOR_ASYNC")] == [0, 1, 2, 3]
    assert re.search(r"GP_ERR_NEWTON_NOT_CONVERGED = Cint\((-?\d+)\)", t).group(1) == \
        str(enum["GP_ERR_NEWTON_NOT_CONVERGED"])
    assert "GP_GENROU, GP_IEESGO, GP_TGOV1, GP_SEXS, GP_ESDC1A, GP_IEEEST, GP_ZIPLOAD, GP_STATICGEN = 0:7" in t
    assert "const COLS = (6, 4, 4, 4, 4, 4, 2, 4)" in t

"""The Julia binding (celly.jl_amd/julia/CorePottsMI355XExt.jl) checked against the C header.

Julia is not installed here, so the binding cannot run; these checks are mechanical:
* CpmRelation, CpmMechanics and CpmModelDesc repeat cpm_relation_desc, cpm_mechanics_desc and
  cpm_model_desc field for field: same names, same order, the Julia type of each C type, so the
  isbits layouts (C rules) agree; the layout computed from the Julia field types has the size
  and offsets ctypes gives the C struct;
* every ccall names an exported symbol of the header with its return type and argument types;
* the dispatch: the workspace is the ninth type parameter (Q) of
  ScientificPottsIntegrator{S, C, R, A, G, P, W, M, Q, L, B} (lib/CorePotts/src/algorithms/
  sequential.jl:255-273), the MCS and report methods fix A to the algorithm, and the workspace
  hooks take the execution plan positionally (checkerboard.jl:159, lottery.jl:46).
"""
import ctypes as C
import os
import re

import pytest

from corepotts import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cpm_mi355x.h")
JULIA = os.path.join(ROOT, "celly.jl_amd", "julia", "CorePottsMI355XExt.jl")
CONSTS = {"CPM_MAX_DIRS": 128, "CPM_MAX_TYPES": 16, "CPM_MAX_MEDIA": 16, "CPM_MAX_MECHANICS": 2,
          "MAX_DIRS": 128, "MAX_TYPES": 16, "MAX_MEDIA": 16}
SCALARS = {"int32_t": "Int32", "uint32_t": "UInt32", "int64_t": "Int64", "uint64_t": "UInt64",
           "double": "Float64", "float": "Float32", "uint8_t": "UInt8", "void": "Cvoid"}
STRUCTS = {"cpm_relation_desc": "CpmRelation", "cpm_mechanics_desc": "CpmMechanics",
           "cpm_model_desc": "CpmModelDesc"}
SIZES = {"Int32": 4, "UInt32": 4, "Int64": 8, "UInt64": 8, "Float64": 8, "Float32": 4, "UInt8": 1}


def _strip_comments(text):
    return re.sub(r"/\*.*?\*/", " ", re.sub(r"//[^\n]*", " ", text), flags=re.S)


def _eval_dim(expr):
    for k, v in CONSTS.items():
        expr = re.sub(rf"\b{k}\b", str(v), expr)
    return int(eval(expr, {"__builtins__": {}}))


def c_structs():
    text = _strip_comments(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct \w+ \{(.*?)\} (\w+);", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(const )?(\w+) (\*?)(\w+)((?:\[[^\]]+\])*)$", decl)
            assert m, decl
            ctype, ptr, fname, dims = m.group(2), m.group(3), m.group(4), m.group(5)
            jt = STRUCTS.get(ctype) or SCALARS[ctype]
            if ptr:
                jt = f"Ptr{{{jt}}}"
            for d in reversed(re.findall(r"\[([^\]]+)\]", dims)):
                jt = f"NTuple{{{_eval_dim(d)}, {jt}}}"
            fields.append((fname, jt))
        out[STRUCTS.get(name, name)] = fields
    return out


def julia_structs():
    text = re.sub(r"#[^\n]*", "", open(JULIA).read())
    out = {}
    for name, body in re.findall(r"\nstruct (Cpm\w+)\n(.*?)\nend", text, flags=re.S):
        fields = []
        for line in body.splitlines():
            line = line.strip()
            if not line:
                continue
            fname, ftype = line.split("::", 1)
            ftype = ftype.replace(" ", "")
            ftype = re.sub(r"NTuple\{([^,{}]+),", lambda m: "NTuple{%d," % _eval_dim(m.group(1)), ftype)
            fields.append((fname.strip(), ftype.replace(",", ", ")))
        out[name] = fields
    return out


def _norm(t):
    return t.replace(" ", "")


def test_structs_match_the_header_field_for_field():
    cs, js = c_structs(), julia_structs()
    for name in ("CpmRelation", "CpmMechanics", "CpmModelDesc"):
        assert [f for f, _ in js[name]] == [f for f, _ in cs[name]], name
        for (fname, ct), (_, jt) in zip(cs[name], js[name]):
            assert _norm(ct) == _norm(jt), f"{name}.{fname}: C {ct} vs Julia {jt}"


def _layout(t, structs):
    """(size, alignment) of a Julia isbits type under C layout rules."""
    t = _norm(t)
    if t in SIZES:
        return SIZES[t], SIZES[t]
    if t.startswith("Ptr{"):
        return 8, 8
    m = re.match(r"NTuple\{(\d+),(.*)\}$", t)
    if m:
        size, align = _layout(m.group(2), structs)
        return int(m.group(1)) * size, align
    return _struct_layout(structs[t], structs)[:2]


def _struct_layout(fields, structs):
    off, align, offsets = 0, 1, []
    for _, t in fields:
        s, a = _layout(t, structs)
        off = (off + a - 1) // a * a
        offsets.append(off)
        off += s
        align = max(align, a)
    return (off + align - 1) // align * align, align, offsets


def test_julia_layout_equals_the_c_layout():
    js = julia_structs()
    size, _, offsets = _struct_layout(js["CpmModelDesc"], js)
    assert size == C.sizeof(native.CpmModelDesc)
    c_offsets = [getattr(native.CpmModelDesc, f).offset for f, _ in native.CpmModelDesc._fields_]
    assert offsets == c_offsets
    assert _struct_layout(js["CpmRelation"], js)[0] == C.sizeof(native.CpmRelation)
    assert _struct_layout(js["CpmMechanics"], js)[0] == C.sizeof(native.CpmMechanicsDesc)


def c_prototypes():
    text = _strip_comments(open(HEADER).read())
    out = {}
    for ret, name, args in re.findall(r"\n(?:const )?(\w+ \*?)(cpm_\w+)\(([^)]*)\);", text):
        out[name] = (ret.strip(), [a.strip() for a in args.split(",")] if args.strip() != "void" else [])
    return out


def _c_to_julia(ctype):
    ctype = " ".join(ctype.split())
    m = re.match(r"(const )?(\w+) ?(\**)\s*\w*(\[\d*\])?$", ctype)
    assert m, ctype
    base, stars, arr = m.group(2), m.group(3), m.group(4)
    if base in ("cpm_ctx", "cpm_batch", "cpm_snapshot"):
        return "Ref{Ptr{Cvoid}}" if len(stars) == 2 else "Ptr{Cvoid}"
    if base == "cpm_model_desc":
        return "Ref{CpmModelDesc}"
    if base == "char":
        return "Cstring"
    jt = SCALARS[base]
    return f"Ptr{{{jt}}}" if stars or arr else jt


def test_every_ccall_matches_a_header_prototype():
    protos = c_prototypes()
    text = open(JULIA).read()
    calls = re.findall(r"ccall\(\(:(\w+), LIB\), (\w+),\s*\(((?:[^()]|\([^()]*\))*)\)", text, flags=re.S)
    assert len(calls) >= 10
    for name, ret, args in calls:
        assert name in protos, name
        cret, cargs = protos[name]
        assert _c_to_julia(cret) == ret, f"{name}: return {cret} vs {ret}"
        jargs = [a.strip() for a in re.split(r",(?![^{]*\})", " ".join(args.split())) if a.strip()]
        assert [_norm(_c_to_julia(a)) for a in cargs] == [_norm(a) for a in jargs], name
        assert name in native.EXPORTS or name == "cpm_abi_version"


def test_dispatch_shape():
    text = open(JULIA).read()
    # the ninth parameter of ScientificPottsIntegrator{S, C, R, A, G, P, W, M, Q, L, B} is Q
    m = re.search(r"const MI355XIntegrator\{A\} = CorePotts\.ScientificPottsIntegrator\{\s*([^}]*)\}", text)
    params = [p.strip() for p in m.group(1).split(",")]
    assert params == ["S", "C", "R", "A", "G", "P", "W", "M", "MI355XWorkspace"]
    for alg in ("CheckerboardSweepCPM", "LotteryCPM"):
        assert re.search(rf"perform_scientific_mcs!\(integrator::MI355XIntegrator\{{A\}}, ::{alg}\) where \{{A <: {alg}\}}", text)
        assert re.search(rf"_current_mcs_report\(integrator::MI355XIntegrator\{{A\}}, ::{alg}\) where \{{A <: {alg}\}}", text)
    assert re.search(r"_allocate_checkerboard_workspace\(state, proposal_relation, accesses, plan::MI355XPlan\)", text)
    assert re.search(r"_allocate_lottery_workspace\(state, proposal_relation, accesses, plan::MI355XPlan\)", text)
    assert "function logical_state(integrator::MI355XIntegrator)" in text
    assert "function model_descriptor(" in text

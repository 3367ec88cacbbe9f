"""The Julia binding (julia/TitusHIP.jl) against the C header.

Julia is not installed here, so the binding cannot run; these tests read it
as text and check what a mismatch would break at a `ccall`: the field order
and types of its `titus_problem_v1` mirror, and, for every `ccall`, that the
symbol is declared in include/titus_hip.h with the same return type, the
same number of parameters and matching parameter types.
"""
import os
import re

import pytest

from test_abi import HEADER, ROOT, header_struct_fields

JL = os.path.join(ROOT, "julia", "TitusHIP.jl")

# Julia type of a struct field or ccall argument -> C type of the header
JULIA_TO_C = {
    "Int32": "int32_t", "Float64": "double", "Cint": "int", "Cvoid": "void",
    "Cstring": "char*",
    "Ptr{Float64}": "double*", "Ptr{Int32}": "int32_t*",
    "Ptr{Cvoid}": "titus_ctx*", "Ptr{Ptr{Cvoid}}": "titus_ctx**",
    "Ref{TitusProblemV1}": "titus_problem_v1*", "Ref{TitusBeamV1}": "titus_beam_v1*",
    "Ptr{Ptr{Float64}}": "double**", "Ptr{Int64}": "int64_t*",
}


def c_norm(t):
    """Normalise a C parameter type: drop const, `struct`, names."""
    t = t.replace("const", " ").replace("struct", " ")
    t = re.sub(r"\s+", " ", t).strip()
    t = re.sub(r"\s*\*\s*", "*", t)
    return t


def header_prototypes():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(titus_\w+)\s*\(([^;{]*?)\)\s*;", txt, re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        if ret.strip().startswith("typedef"):
            continue
        ptypes = []
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                # strip the parameter name (last identifier)
                p = re.sub(r"\b\w+$", "", p).strip()
                ptypes.append(c_norm(p))
        out[name] = (c_norm(ret), ptypes)
    return out


def julia_struct_fields():
    txt = open(JL).read()
    body = re.search(r"struct TitusProblemV1\n(.*?)\nend", txt, re.S).group(1)
    fields = []
    for line in body.splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        name, typ = line.split("::")
        fields.append((name.strip(), typ.strip()))
    return fields


def julia_ccalls():
    txt = open(JL).read()
    calls = []
    pat = re.compile(r"ccall\(\(:(titus_\w+),\s*libtitus\),\s*([\w{}]+),\s*\(([^()]*)\)")
    for m in pat.finditer(txt):
        args = [a.strip() for a in m.group(3).split(",") if a.strip()]
        calls.append((m.group(1), m.group(2), args))
    return calls


def test_julia_struct_mirrors_the_header():
    jl = julia_struct_fields()
    hd = header_struct_fields()
    assert [n for n, _ in jl] == [n for n, _ in hd]
    for (n, jt), (_, ct) in zip(jl, hd):
        assert JULIA_TO_C[jt] == ct, (n, jt, ct)


def test_every_ccall_matches_its_prototype():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 11
    seen = set()
    for name, ret, args in calls:
        assert name in protos, name
        cret, cparams = protos[name]
        assert JULIA_TO_C[ret] == cret, (name, ret, cret)
        assert len(args) == len(cparams), (name, args, cparams)
        for a, c in zip(args, cparams):
            assert JULIA_TO_C[a] == c, (name, a, c)
        seen.add(name)
    # the binding drives the whole forward, batch and gradient path
    for need in ("titus_create", "titus_solve", "titus_get_state", "titus_progress",
                 "titus_get_history", "titus_destroy", "titus_solve_batch",
                 "titus_solve_batch_lockstep",
                 "titus_set_checkpoints", "titus_adjoint_batch", "titus_adjoint_batch_lockstep",
                 "titus_last_error",
                 "titus_get_cfl", "titus_set_rejection", "titus_get_rejections"):
        assert need in seen, need


def test_header_parser_sees_every_declared_function():
    from test_abi import header_functions
    assert set(header_prototypes()) == set(header_functions())


@pytest.mark.parametrize("name", ["titus_problem", "SolveFirstCollisionSourceDLR",
                                  "SolveFirstCollisionSourceDLRBUGRejection",
                                  "SolveFirstCollisionSourceDLRParallel",
                                  "CSolveFirstCollisionSourceDLR", "plan_objective",
                                  "CudaSolveFirstCollisionSourceDLR4thOrder"])
def test_binding_defines_the_reference_entry_points(name):
    txt = open(JL).read()
    assert re.search(r"\nfunction " + name + r"\(", txt), name

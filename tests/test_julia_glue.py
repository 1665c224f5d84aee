"""CPU check of the Julia extension source (itensors.jl_amd/julia/
NDTensorsMI355XExt.jl) against the C ABI: Julia cannot run in this image, so
every `ccall` in the glue is parsed and its declared argument types and its
argument count are compared with the prototype of the same function in
include/itg.h (return type, arity, and a Julia type compatible with each C
parameter type). A mismatch would be a crash or silent corruption at the
boundary."""
import os
import re

import pytest

from helpers import ROOT

GLUE = os.path.join(ROOT, "itensors.jl_amd", "julia", "NDTensorsMI355XExt.jl")
HEADER = os.path.join(ROOT, "include", "itg.h")

# Julia types a C parameter type may be passed as
COMPAT = {
    "int": {"Cint"},
    "itg_dtype": {"Cint"},
    "itg_prec": {"Cint"},
    "unsigned": {"Cuint"},
    "int64_t": {"Int64"},
    "size_t": {"Csize_t"},
    "const int64_t*": {"Ptr{Int64}", "Ref{Int64}"},
    "int64_t*": {"Ptr{Int64}", "Ref{Int64}"},
    "const int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "const uint8_t*": {"Ptr{UInt8}"},
    "const double*": {"Ref{ComplexF64}", "Ptr{ComplexF64}", "Ptr{Cvoid}", "Ptr{Float64}"},
    "void*": "pointer",
    "const void*": "pointer",
    "void**": {"Ref{Ptr{Cvoid}}"},
}
RET = {"itg_status": "Cint", "const char*": "Cstring", "int64_t": "Int64"}


def _split_top(s):
    """Split on commas outside (), {}, []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(s, i):
    """The text inside the parenthesis opening at s[i]."""
    assert s[i] == "("
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return s[i + 1:j], j
    raise ValueError("unbalanced")


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"(itg_status|const char\*|int64_t)\s+(itg_\w+)\s*\(", text):
        inner, _ = _balanced(text, m.end() - 1)
        params = []
        for p in _split_top(inner):
            p = " ".join(p.split())
            if p == "void":
                continue
            # drop the parameter name
            t = re.sub(r"\s*\b\w+$", "", p).replace(" *", "*").strip()
            params.append(t)
        protos[m.group(2)] = (m.group(1), params)
    return protos


def glue_ccalls():
    text = open(GLUE).read()
    text = "\n".join(l.split("#", 1)[0] for l in text.splitlines())
    calls = []
    for m in re.finditer(r"ccall\(", text):
        inner, _ = _balanced(text, m.end() - 1)
        parts = _split_top(inner)
        name = re.match(r"\(:(\w+),\s*libitg\)", parts[0]).group(1)
        ret = parts[1]
        types = _split_top(parts[2].strip()[1:-1]) if parts[2].strip().startswith("(") else []
        calls.append((name, ret, types, parts[3:]))
    return calls


def _norm(t):
    t = t.replace(" ", "")
    t = re.sub(r"Ptr\{T\}|Ptr\{S\}", "Ptr{Cvoid}", t)
    return t


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = glue_ccalls()
    assert len(calls) >= 15
    seen = set()
    for name, ret, types, args in calls:
        assert name in protos, f"{name} is not declared in include/itg.h"
        cret, params = protos[name]
        assert ret == RET[cret], (name, ret, cret)
        assert len(types) == len(params), (name, types, params)
        assert len(args) == len(params), (name, len(args), len(params))
        for k, (jt, ct) in enumerate(zip(types, params)):
            jt = _norm(jt)
            if ct.endswith("**") or ct in ("itg_bs_plan**", "itg_handle**", "itg_graph**"):
                ok = jt in ("Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}")
            elif re.match(r"(const )?itg_\w+\*$", ct):
                ok = jt == "Ptr{Cvoid}"
            else:
                allowed = COMPAT[ct]
                ok = (jt.startswith(("Ptr{", "Ref{")) if allowed == "pointer" else jt in allowed)
            assert ok, f"{name} argument {k + 1}: Julia {jt} for C {ct}"
        seen.add(name)
    # the seams the glue must bind
    for need in ("itg_contract_dense", "itg_bs_plan_create", "itg_bs_execute",
                 "itg_bs_diag_plan_create", "itg_bs_diag_execute", "itg_convert",
                 "itg_memcpy_htod", "itg_memcpy_dtoh", "itg_memcpy_dtod", "itg_bs_copy"):
        assert need in seen, need


def test_unsupported_falls_back_in_every_contraction_method():
    """Each contraction method of the glue handles ITG_ERR_UNSUPPORTED by the
    stock path (NDTensorscuTENSORExt/contract.jl:45-53), not by throwing."""
    text = open(GLUE).read()
    methods = re.split(r"\nfunction NDTensors\.contract!?\(", text)[1:]
    assert len(methods) >= 3
    for body in methods:
        body = body.split("\nend\n", 1)[0]
        if "ccall((:itg_" not in body:
            continue  # a forwarding method
        assert "ITG_ERR_UNSUPPORTED" in body
        assert "fallback_contract" in body or "NDTensors.contract!(Rh" in body


@pytest.mark.parametrize("fn", ["setindex!", "copyto!"])
def test_device_vector_has_writers(fn):
    text = open(GLUE).read()
    assert re.search(r"function Base\." + re.escape(fn) + r"\(\w+::ItgVector", text)

"""The Julia bindings (particleincell.jl_amd/julia/*.jl) against the C header,
without Julia (absent here): every ccall names a function the header
declares, with as many argument types as the C parameters, each of the same
kind (pointer, int, int64_t, uint64_t, double), and an int (Cint) status, or
the string of pic_last_error."""
import glob
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pic_mi355x.h")
JL = sorted(glob.glob(os.path.join(ROOT, "particleincell.jl_amd", "julia", "**", "*.jl"),
                      recursive=True))


def c_decls():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, params in re.findall(r"\b(int|const char \*)\s*(pic_[a-z0-9_]+)\s*\(([^)]*)\)",
                                        text):
        params = params.strip()
        args = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        out[name] = (ret.strip(), [c_kind(a) for a in args])
    return out


def c_kind(param):
    if "*" in param or "[" in param:
        return "ptr"
    t = param.rsplit(" ", 1)[0]
    return {"int": "int", "int64_t": "i64", "uint64_t": "u64", "double": "f64"}[t]


def jl_kind(t):
    t = t.strip()
    if t.startswith(("Ptr{", "Ref{")) or t in ("Cstring", "CPtr"):
        return "ptr"
    return {"Cint": "int", "Int32": "int", "Int64": "i64", "UInt64": "u64", "Float64": "f64"}[t]


def split_types(tuple_text):
    """Top-level comma split of a Julia tuple's contents."""
    parts, depth, cur = [], 0, ""
    for ch in tuple_text:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p for p in (s.strip() for s in parts) if p]


def ccalls(path):
    """(name, return type, [argument types]) of every ccall in a file."""
    src = open(path).read()
    out = []
    for m in re.finditer(r"ccall\(\s*\(:(pic_[a-z0-9_]+),\s*\w+\)\s*,\s*(\w+)\s*,\s*\(", src):
        i, depth = m.end(), 1
        j = i
        while depth:
            depth += {"(": 1, ")": -1}.get(src[j], 0)
            j += 1
        out.append((m.group(1), m.group(2), split_types(src[i:j - 1])))
    return out


def test_julia_files_exist():
    names = [os.path.basename(p) for p in JL]
    assert "pic_mi355x.jl" in names and "ParticleInCellChainRulesExt.jl" in names


@pytest.mark.parametrize("path", JL, ids=[os.path.basename(p) for p in JL])
def test_ccalls_match_header(path):
    decls = c_decls()
    calls = ccalls(path)
    assert calls, path
    for name, ret, args in calls:
        assert name in decls, name
        cret, cargs = decls[name]
        assert ret == ("Cstring" if cret.startswith("const char") else "Cint"), name
        assert [jl_kind(a) for a in args] == cargs, (name, args, cargs)


def test_bindings_cover_the_reference_calls():
    """The drop-ins of src/pic.jl and the device methods of the hot path are
    all bound."""
    bound = {n for p in JL for n, _, _ in ccalls(p)}
    for name in ("pic_f90_push_v", "pic_f90_push_x", "pic_f90_deposition", "pic_compute_rho",
                 "pic_push_v", "pic_push_x", "pic_solve_e", "pic_run", "pic_adjoint",
                 "pic_run_maxwell", "pic_ua_run", "pic_sample_landau", "pic_read_particles"):
        assert name in bound, name

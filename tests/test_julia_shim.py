"""The Julia binding (nextla.jl_amd/julia/NextLAHIP.jl) against the C ABI.

Julia is not installed here, so the module cannot be run; what can be checked
without it is that every `ccall` in it names a function of include/nla.h with
the prototype's return type and the same number and kinds of arguments
(Cchar <-> char, Int32 <-> int32_t, Int64 <-> int64_t, Float64 <-> double,
Ptr/Cstring <-> pointers and opaque handles), and that the byte buffer it
passes as nla_mixed_info is as large as the C struct.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "nextla.jl_amd", "julia", "NextLAHIP.jl")
HDR = os.path.join(ROOT, "include", "nla.h")


def _split_top(s):
    """Split at commas outside (), {} and []."""
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


def _matching(s, i):
    """Index just past the parenthesis group that opens at s[i]."""
    assert s[i] == "("
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j + 1
    raise ValueError("unbalanced parentheses")


def c_prototypes():
    txt = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for mt in re.finditer(r"^([A-Za-z_][\w \*]*?)\b(nla_\w+)\s*\(", txt, flags=re.M):
        ret, name = mt.group(1).strip(), mt.group(2)
        i = mt.end() - 1
        j = _matching(txt, i)
        if txt[j:].lstrip()[:1] != ";":
            continue  # not a declaration
        params = _split_top(txt[i + 1:j - 1])
        if params == ["void"]:
            params = []
        protos[name] = (ret, params)
    return protos


def c_kind(decl):
    d = decl.strip()
    if "*" in d or d.split()[0] in ("nla_batch_plan", "nla_batch_mixed_plan", "nla_server"):
        return "ptr"
    t = d.split()[0] if not d.startswith("const") else d.split()[1]
    return {"char": "char", "int32_t": "i32", "int64_t": "i64", "double": "f64"}[t]


def jl_kind(t):
    t = t.strip()
    if t.startswith("Ptr") or t in ("Cstring",):
        return "ptr"
    return {"Cchar": "char", "Int32": "i32", "Int64": "i64", "Float64": "f64"}[t]


def julia_ccalls():
    txt = open(JL).read()
    calls = []
    # ccall((:nla_x, libnlahip), ...) and the trsm/trmm loop's
    # ccall(($(QuoteNode(sym)), libnlahip), ...) over (:nla_trsm, :nla_trmm)
    pat = r"ccall\(\((?::(nla_\w+)|\$\(QuoteNode\(sym\)\)),\s*libnlahip\),"
    for mt in re.finditer(pat, txt):
        names = [mt.group(1)] if mt.group(1) else ["nla_trsm", "nla_trmm"]
        i = mt.end()
        rest = txt[i:]
        k = rest.index(",")
        ret = rest[:k].strip()
        j = i + k + 1
        while txt[j].isspace():
            j += 1
        end = _matching(txt, j)
        types = _split_top(txt[j + 1:end - 1])
        calls.extend((name, ret, types) for name in names)
    return calls


def test_every_ccall_matches_the_header():
    protos = c_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 23, "the ccall parser found too few calls"
    for name, ret, types in calls:
        assert name in protos, f"{name} is not declared in include/nla.h"
        cret, params = protos[name]
        assert [jl_kind(t) for t in types] == [c_kind(p) for p in params], name
        want_ret = {"int32_t": "Int32", "int64_t": "Int64", "const char *": "Cstring"}[cret]
        assert ret == want_ret, (name, ret, cret)


def test_header_parser_sees_every_exported_function():
    """The prototype parser itself: it finds every function the ctypes
    mirror declares (which tests/test_abi.py checks against the library)."""
    sys_path = os.path.join(ROOT, "nextla.jl_amd")
    import sys
    if sys_path not in sys.path:
        sys.path.insert(0, sys_path)
    from nextla_amd import _lib
    assert set(c_prototypes()) == set(_lib.SIGNATURES)


def test_mixed_info_buffer_size():
    from nextla_amd import _lib
    txt = open(JL).read()
    mt = re.search(r"info = zeros\(UInt8, ([^)]*)\)", txt)
    assert mt, "rectrxm_mixed! info buffer not found"
    size = eval(mt.group(1), {"__builtins__": {}})
    assert size == ctypes.sizeof(_lib.MixedInfo)


@pytest.mark.parametrize("needle", ["nla_rectrxm", "nla_rectrxm_mixed", "nla_residual", "nla_rectrxm_noalloc",
                                    "nla_getrf2", "nla_lauum", "nla_larfb", "nla_batch_plan_create",
                                    "nla_trsm", "nla_trmm", "nla_gemm_add", "nla_gemm_sub", "nla_pamm",
                                    "nla_parfb", "nla_serve_create", "nla_serve_submit", "nla_serve_step",
                                    "nla_serve_drain", "nla_serve_destroy"])
def test_entry_points_bound(needle):
    assert needle in {c[0] for c in julia_ccalls()}

"""The Julia ccall shim (julia/QED2dHIP.jl) against the C ABI (include/qed2d.h).

Julia is not installed here, so the shim cannot run; this checks statically
that every ccall in it names an entry point of include/qed2d.h, passes as many
arguments as the prototype has, declares a Julia type that matches each C
parameter type and the return type, and that the shim defines every
reference function of the drop-in list (src/QED2d.jl:27-54)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qed2d.h")
SHIM = os.path.join(ROOT, "julia", "QED2dHIP.jl")

# The reference's exported hot-path functions (src/QED2d.jl:27-54) that the shim
# must define, plus the RNG seeding that replaces Random.seed!(CUDA.default_rng()).
REQUIRED = [
    "gamm5Dw", "gamm5", "Dw", "gamm5Dw_sqr", "gamm5Dw_sqr_msq", "gamm5Dw_sqr_musq",
    "gamm5Dw_sqr_sqr", "gamm5Dw_sqr_sqr_musq", "tr_dQwdU", "krnl_force!", "updt_fld!",
    "plaquette!", "qtop!", "Action", "Plaquette", "Qtop", "Hamiltonian", "CG", "MultiCG", "R",
    "generate_pseudofermion!", "power_method", "get_rhmc_params", "LuscherZ", "LuscherZp",
    "reweighting_factor", "HMC!", "OMF4!", "leapfrog!", "update_momenta!", "unfold_fld!",
    "unfold_reflect!", "save_gauge", "load_gauge", "seed!",
]

# C parameter type (const dropped, spaces normalised) -> Julia types that ccall
# converts to it.
C2J = {
    "int": {"Cint"},
    "long": {"Clong"},
    "double": {"Cdouble", "Float64"},
    "float": {"Cfloat", "Float32"},
    "uint64_t": {"UInt64"},
    "uint32_t": {"UInt32"},
    "void*": {"Ptr{Cvoid}"},
    "qed2d_ctx*": {"Ptr{Cvoid}"},
    "qed2d_batch*": {"Ptr{Cvoid}"},
    "qed2d_ctx**": {"Ptr{Ptr{Cvoid}}", "Ref{Ptr{Cvoid}}"},
    "qed2d_batch**": {"Ptr{Ptr{Cvoid}}", "Ref{Ptr{Cvoid}}"},
    "char*": {"Cstring"},
    # device fields travel as Ptr{Cvoid} (dptr); host arrays and out-parameters typed
    "double*": {"Ptr{Cvoid}", "Ptr{Cdouble}", "Ptr{Float64}", "Ref{Cdouble}"},
    "float*": {"Ptr{Cvoid}", "Ptr{Cfloat}"},
    "int*": {"Ptr{Cint}", "Ref{Cint}"},
    "long*": {"Ptr{Clong}", "Ref{Clong}"},
    "uint64_t*": {"Ptr{UInt64}"},
    "double**": {"Ptr{Ptr{Cvoid}}", "Ptr{Ptr{Cdouble}}", "Ref{Ptr{Cvoid}}"},
    "double*const*": {"Ptr{Ptr{Cvoid}}", "Ptr{Ptr{Cdouble}}"},
}
RET = {"int": "Cint", "void": "Cvoid", "char*": "Cstring", "double": "Cdouble"}


def _norm_ctype(t):
    t = re.sub(r"\bconst\b", "", t)
    return re.sub(r"\s+", "", t)


def parse_header():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", " ", src)
    src = re.sub(r"#[^\n]*", " ", src)
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(qed2d_\w+)\s*\(([^;{}]*?)\)\s*;", src):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        ret = _norm_ctype(ret.replace("typedef", "").replace("struct", ""))
        plist = []
        if params.strip() not in ("", "void"):
            for p in params.split(","):
                p = p.strip()
                arr = re.match(r"(.*?)(\w+)\s*\[\s*\d*\s*\]$", p)
                if arr:  # `int tiles[4]` is a pointer parameter
                    plist.append(_norm_ctype(arr.group(1)) + "*")
                    continue
                pm = re.match(r"(.*?)(\w+)$", p)
                plist.append(_norm_ctype(pm.group(1)))
        protos[name] = (ret, plist)
    return protos


def _split_top(s):
    """Split at top-level commas (outside (), {}, [])."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    tail = "".join(cur).strip()
    if tail:
        out.append(tail)
    return out


def parse_ccalls():
    src = open(SHIM).read()
    src = "\n".join(l.split("#", 1)[0] if not l.lstrip().startswith("#") else ""
                    for l in src.splitlines())
    calls = []
    for m in re.finditer(r"\bccall\(", src):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        parts = _split_top(src[m.end():i - 1])
        target = parts[0]
        nm = re.match(r"\(\s*:(\w+)\s*,\s*lib\s*\)", target)
        assert nm, f"ccall target not a literal (:name, lib): {target}"
        types = parts[2].strip()
        assert types.startswith("(") and types.endswith(")"), types
        tlist = _split_top(types[1:-1])
        calls.append((nm.group(1), parts[1].strip(), tlist, parts[3:]))
    return calls


def test_every_ccall_matches_the_header():
    protos = parse_header()
    calls = parse_ccalls()
    assert len(calls) >= 35
    for name, ret, types, args in calls:
        assert name in protos, f"{name}: not declared in include/qed2d.h"
        cret, cparams = protos[name]
        assert RET[cret] == ret, f"{name}: returns {cret}, shim declares {ret}"
        assert len(types) == len(cparams), \
            f"{name}: {len(cparams)} C parameters, shim declares {len(types)}"
        assert len(args) == len(types), \
            f"{name}: shim declares {len(types)} types but passes {len(args)} arguments"
        for k, (jt, ct) in enumerate(zip(types, cparams)):
            assert ct in C2J, f"{name}: C type {ct} has no Julia mapping"
            assert jt in C2J[ct], f"{name} parameter {k}: C {ct} vs Julia {jt}"


def test_shim_defines_the_reference_exports():
    src = open(SHIM).read()
    defined = set(re.findall(r"^function\s+([\w!]+)\s*\(", src, flags=re.M))
    # short-form methods `name(args) = ...` start at column 0 (calls are indented)
    defined |= set(re.findall(r"^([\w!]+)\(", src, flags=re.M))
    missing = [f for f in REQUIRED if f not in defined]
    assert not missing, f"shim lacks {missing}"
    exported = " ".join(re.findall(r"^export\s+(.*?)(?=^\S)", src, flags=re.M | re.S))
    for f in REQUIRED:
        assert re.search(r"(?<![\w!])" + re.escape(f) + r"(?![\w!])", exported), \
            f"{f} not exported"


def test_hmc_methods_dispatch_on_integrators():
    """HMC! takes integrator::Integrators (src/QED2dHMC.jl:19,64,116), never an Int."""
    src = open(SHIM).read()
    heads = re.findall(r"^function HMC!\((.*?)\)\n", src, flags=re.M | re.S)
    assert len(heads) == 3
    for h in heads:
        assert "integrator::Integrators" in h
    assert "integrator_id(::Leapfrog)" in src and "integrator_id(::OMF4)" in src


@pytest.mark.parametrize("name", ["qed2d_hmc_nf2", "qed2d_hmc_rhmc", "qed2d_cg",
                                  "qed2d_multicg", "qed2d_power_method"])
def test_key_entry_points_are_bound(name):
    assert name in {c[0] for c in parse_ccalls()}

"""The Julia binding (genericlinearalgebra.jl_amd/julia/GenericLinearAlgebraHIP)
against the C ABI it calls: every `ccall` goes through the module's
SIGNATURES table, and each entry must match include/gla_hip.h argument by
argument (C type -> the Julia types that pass it).  Julia is not installed
here, so the file is checked as text: the table, the reference functions it
extends, and the entry points of the header it leaves out.  CPU only."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gla_hip.h")
JL_DIR = os.path.join(ROOT, "genericlinearalgebra.jl_amd", "julia", "GenericLinearAlgebraHIP")
JL = os.path.join(JL_DIR, "src", "GenericLinearAlgebraHIP.jl")

# C parameter type -> Julia types that ccall accepts for it
C2JL = {
    "gla_handle_t": {"Ptr{Cvoid}"},
    "gla_batch_plan_t": {"Ptr{Cvoid}"},
    "gla_handle_t *": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "gla_batch_plan_t *": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "char": {"Cchar"},
    "int": {"Cint"},
    "int32_t": {"Int32"},
    "int64_t": {"Int64"},
    "double": {"Float64"},
    "float": {"Float32"},
    "void *": {"Ptr{Cvoid}"},
    "const char *": {"Cstring", "Ptr{Cchar}"},
    "char *": {"Ref{Cchar}", "Ptr{Cchar}"},
    "double *": {"Ptr{Float64}", "Ref{Float64}"},
    "const double *": {"Ptr{Float64}"},
    "float *": {"Ptr{Float32}", "Ref{Float32}"},
    "const float *": {"Ptr{Float32}"},
    "int64_t *": {"Ref{Int64}", "Ptr{Int64}"},
    "int32_t *": {"Ptr{Int32}", "Ref{Int32}"},
    "const int32_t *": {"Ptr{Int32}"},
    "double *const *": {"Ptr{Ptr{Float64}}"},
    "gla_cdouble *": {"Ptr{ComplexF64}"},
    "const gla_cdouble *": {"Ptr{ComplexF64}"},
    "gla_cfloat *": {"Ptr{ComplexF32}"},
    "const gla_cfloat *": {"Ptr{ComplexF32}"},
    "float *const *": {"Ptr{Ptr{Float32}}"},
}
RET2JL = {"int": {"Cint"}, "const char *": {"Cstring"}}


def _norm(t: str) -> str:
    t = re.sub(r"\s+", " ", t.strip())
    t = re.sub(r"\s*\*\s*", " *", t)          # "double* x" / "double *x" -> "double *"
    t = t.replace("* const", "*const").replace("*const *", "*const *")
    return t.strip()


def header_decls():
    """name -> (return type, [parameter types]) for every declared function."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^(int|const char \*)\s*(gla_[a-z0-9_]+)\s*\(([^;]*?)\)\s*;",
                         text, re.M | re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        types = []
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                pm = re.match(r"(.*?)([A-Za-z_][A-Za-z0-9_]*)$", p)
                types.append(_norm(pm.group(1)))
        out[name] = (_norm(ret), types)
    return out


def julia_signatures():
    """name -> (return type, [argument types]) from the SIGNATURES table."""
    text = open(JL).read()
    block = re.search(r"const SIGNATURES = Dict\{.*?\}\((.*?)\n\)\n", text, re.S).group(1)
    out = {}
    for m in re.finditer(r":(gla_[a-z0-9_]+) => \((\w+), \((.*?)\)\),?\s*$", block, re.M):
        args = [a.strip() for a in m.group(3).split(",") if a.strip()]
        out[m.group(1)] = (m.group(2), args)
    return out


def test_signatures_table_matches_the_header():
    decls = header_decls()
    sigs = julia_signatures()
    assert len(sigs) >= 40
    problems = []
    for name, (ret, args) in sigs.items():
        if name not in decls:
            problems.append(f"{name}: not declared in gla_hip.h")
            continue
        cret, cargs = decls[name]
        if ret not in RET2JL.get(cret, set()):
            problems.append(f"{name}: return {ret} for C {cret}")
        if len(args) != len(cargs):
            problems.append(f"{name}: {len(args)} arguments, header has {len(cargs)}")
            continue
        for i, (jt, ct) in enumerate(zip(args, cargs)):
            if jt not in C2JL.get(ct, set()):
                problems.append(f"{name} argument {i + 1}: Julia {jt} for C {ct}")
    assert not problems, "\n".join(problems)


def test_binding_covers_every_entry_point():
    """Every header entry point is bound, except the ones that exist for
    Python tooling only (listed here with the reason)."""
    not_bound = {
        "gla_chol_recursive_noalloc_d",   # for callers that cannot allocate; C/graph users
    }
    missing = sorted(set(header_decls()) - set(julia_signatures()) - not_bound)
    assert not missing, missing


def test_every_ccall_goes_through_the_table():
    text = open(JL).read()
    # raw ccalls are allowed into the HIP runtime only
    for m in re.finditer(r"ccall\(\((:\w+), (\w+)\)", text):
        sym, lib = m.groups()
        assert lib == "libhip" or (lib == "libgla" and sym == ":gla_batch_plan_destroy"), m.group(0)
    used = set(re.findall(r"@gla\(?\s*(gla_[a-z0-9_]+)", text))
    used |= set(re.findall(r"Symbol\(\"(gla_[a-z0-9_]+)\", sfx\)", text))
    assert used, "no @gla calls found"
    sigs = julia_signatures()
    for u in used:
        if u.endswith("_"):
            assert u + "d" in sigs and u + "s" in sigs, u
        else:
            assert u in sigs, u


def test_reference_methods_and_defaults():
    """The methods a user of the reference calls exist with the reference's
    defaults: cutoff = 1 (src/cholesky.jl:37), blocksize = 12 and a work
    matrix (src/qr.jl:113-118), a device copy for cholRecursive!(copy(A), ...)
    (README.md:36), and the ChainRulesCore imports the rrules use."""
    text = open(JL).read()
    assert re.search(r"function cholRecursive!\(A::HIPMatrix\{\$T\}, uplo::Type\{Val\{U\}\},\s*"
                     r"cutoff::Integer = 1\)", text)
    assert re.search(r"function qrBlocked!\(A::HIPMatrix\{\$T\}, blocksize::Integer = 12,", text)
    assert "work::HIPMatrix{$T} = hipzeros($T, blocksize, size(A, 2))" in text
    assert "Base.copy(A::HIPArray{T}) where {T}" in text
    for name in ("NoTangent", "unthunk"):
        assert re.search(r"using ChainRulesCore: .*\b%s\b" % name, text)
    for name in ("HIPVector", "hipzeros", "hipsimilar"):
        assert re.search(r"\b(const|function)\s+%s\b|^%s\(" % (name, name), text, re.M), name
    for fn in ("cholUnblocked!", "cholBlocked!", "cholRecursive!", "qrUnblocked!", "qrBlocked!",
               "rankUpdate!"):
        assert re.search(r"import GenericLinearAlgebra: [^\n]*(\n\s+[^\n]*)*" + re.escape(fn),
                         text), fn


def test_project_toml_declares_dependencies():
    toml = open(os.path.join(JL_DIR, "Project.toml")).read()
    for dep in ("ChainRulesCore", "GenericLinearAlgebra", "LinearAlgebra", "Libdl"):
        assert re.search(r"^%s = \"[0-9a-f-]{36}\"$" % dep, toml, re.M), dep
    # the reference's own package UUID (reference Project.toml)
    assert "14197337-ba66-59df-a3e3-ca00e7dcff7a" in toml


def test_julia_source_blocks_are_balanced():
    """A cheap structural check in place of a parse (no Julia here): block
    openers and `end` balance once strings and comments are removed."""
    text = open(JL).read()
    text = re.sub(r'"""[\s\S]*?"""', '""', text)
    text = re.sub(r'"(\\.|[^"\\])*"', '""', text)
    text = re.sub(r"#.*", "", text)
    text = re.sub(r"\[[^\[\]]*\bend\b[^\[\]]*\]", "[]", text)   # `end` as an index
    opens = len(re.findall(r"^\s*(module|function|macro|struct|mutable struct|for|while|if|"
                           r"let|begin|try)\b", text, re.M))
    opens += len(re.findall(r"@eval\s+(begin|function)\b", text))
    opens += len(re.findall(r"\bdo\b(\s+\w+)?\s*$", text, re.M))
    opens += len(re.findall(r"=\s*(begin|let)\b", text))
    ends = len(re.findall(r"\bend\b", text))
    assert opens == ends, (opens, ends)

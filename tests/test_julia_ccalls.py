"""Every `ccall((:name, lib), R, (T1, ..., Tn), ...)` of the Julia binding in
INTEGRATION.md matches the prototype include/mcx_hip.h declares for `name`:
the return type, the number of arguments and each argument's C type, in
order; and every declared entry point is bound (no GPU needed)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_HANDLES = {"mcx_ctx", "mcx_batch", "mcx_pt", "mcx_gen", "mcx_server"}
_STRUCTS = {"BatchDesc": "mcx_batch_desc", "PTDesc": "mcx_pt_desc", "GenDesc": "mcx_gen_desc",
            "GenSchedule": "mcx_gen_schedule", "GenProgress": "mcx_gen_progress",
            "LaunchShape": "mcx_launch_shape", "ServerDesc": "mcx_server_desc",
            "JobDesc": "mcx_job_desc", "JobResult": "mcx_job_result",
            "ServerStats": "mcx_server_stats_t", "SLOReport": "mcx_slo_report"}
_SCALARS = {"Int8": "int8_t", "Int32": "int32_t", "Int64": "int64_t", "UInt32": "uint32_t",
            "UInt64": "uint64_t", "Float64": "double", "Csize_t": "size_t"}


def _prototypes():
    txt = open(os.path.join(ROOT, "include", "mcx_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"(mcx_status|const char\*)\s+(mcx_\w+)\s*\(([^)]*)\)\s*;",
                                      txt):
        args = " ".join(args.split())
        params = [] if args in ("", "void") else [a.strip() for a in args.split(",")]
        # drop the parameter name: the C type is everything before it
        types = [re.sub(r"\s*\b\w+$", "", p).replace(" *", "*") for p in params]
        out[name] = (ret, types)
    return out


def _accepts(jl: str, c: str) -> bool:
    """Whether Julia argument type jl passes a C argument of type c."""
    c = c.replace("const ", "")
    if jl in _SCALARS:
        return c == _SCALARS[jl]
    m = re.fullmatch(r"(Ptr|Ref)\{(.+)\}", jl)
    if not m or not c.endswith("*"):
        return False
    inner, base = m.group(2), c[:-1]
    if inner == "Cvoid":
        return m.group(1) == "Ptr" and (base in _HANDLES or base == "void")
    if inner == "Ptr{Cvoid}":
        return base.endswith("*") and base[:-1] in _HANDLES
    if inner == "UInt8":
        return base in ("void", "uint8_t")
    if inner in _SCALARS:
        return base == _SCALARS[inner]
    if inner in _STRUCTS:
        return base == _STRUCTS[inner]
    return False


def _ccalls():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    pat = re.compile(r"ccall\(\(:(\w+),\s*lib\),\s*(\w+),\s*\(([^()]*)\)", re.S)
    for m in pat.finditer(txt):
        types = [t.strip() for t in m.group(3).replace("\n", " ").split(",") if t.strip()]
        yield m.group(1), m.group(2), types


def test_every_ccall_matches_the_header():
    protos = _prototypes()
    assert len(protos) > 70
    calls = list(_ccalls())
    assert len(calls) > 70
    bad = []
    for name, ret, types in calls:
        if name not in protos:
            bad.append(f"{name}: not declared in the header")
            continue
        cret, ctypes = protos[name]
        if (ret, cret) not in (("Int32", "mcx_status"), ("Cstring", "const char*")):
            bad.append(f"{name}: returns {ret}, header {cret}")
        if len(types) != len(ctypes):
            bad.append(f"{name}: {len(types)} arguments, header {len(ctypes)} ({ctypes})")
            continue
        for i, (jl, c) in enumerate(zip(types, ctypes)):
            if not _accepts(jl, c):
                bad.append(f"{name}: argument {i + 1} is {jl}, header {c}")
    assert not bad, "\n".join(bad)


def test_every_entry_point_is_bound():
    bound = {name for name, _, _ in _ccalls()}
    missing = sorted(set(_prototypes()) - bound)
    assert not missing, missing


def test_the_checker_rejects_mismatches():
    assert not _accepts("Int64", "int32_t")
    assert not _accepts("Ptr{Float64}", "int32_t*")
    assert not _accepts("Ptr{Cvoid}", "mcx_ctx**")
    assert _accepts("Ref{Ptr{Cvoid}}", "mcx_ctx**")
    assert _accepts("Ref{BatchDesc}", "const mcx_batch_desc*")
    assert not _accepts("Ref{JobDesc}", "const mcx_batch_desc*")
    assert _accepts("Ptr{UInt8}", "const void*")

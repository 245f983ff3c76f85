"""The ctypes mirrors of the C ABI's descriptor and result structs have the
layout that include/mcx_hip.h declares: the same size and the same offset for
every field, read from a C program compiled against the header with gcc.  The
program is generated from the ctypes field names, so a field that the header
lacks or names differently fails to compile."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mirrors():
    from montecarlox import _lib, muca, pt, serving
    return [
        ("mcx_batch_desc", _lib.BatchDesc),
        ("mcx_pt_desc", pt.PTDesc),
        ("mcx_gen_desc", muca.GenDesc),
        ("mcx_server_desc", _lib.ServerDesc),
        ("mcx_job_desc", _lib.JobDesc),
        ("mcx_job_result", _lib.JobResult),
        ("mcx_server_stats_t", _lib.ServerStats),
        ("mcx_launch_shape", _lib.LaunchShape),
        ("mcx_gen_schedule", muca.GenSchedule),
        ("mcx_gen_progress", muca.GenProgress),
        ("mcx_slo_report", serving.SLOReport),
    ]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_the_header(tmp_path):
    pairs = _mirrors()
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "mcx_hip.h"',
             "int main(void) {"]
    for cname, cls in pairs:
        lines.append(f'    printf("%zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'    printf("%zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True, capture_output=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    k = 0
    for cname, cls in pairs:
        assert got[k] == C.sizeof(cls), (cname, cls.__qualname__, got[k], C.sizeof(cls))
        k += 1
        for field, _ in cls._fields_:
            assert got[k] == getattr(cls, field).offset, (cname, field, got[k])
            k += 1


# Julia field type -> bytes
_JL_SIZE = {"Int32": 4, "UInt32": 4, "Int64": 8, "UInt64": 8, "Float64": 8}


def _julia_structs(text):
    """struct Name ... end blocks of INTEGRATION.md: name -> [(field, bytes)]."""
    import re
    out = {}
    for m in re.finditer(r"^struct (\w+)(.*?)\bend\b", text, re.M | re.S):
        fields = []
        for f, t in re.findall(r"(\w+)::([\w{},]+)", m.group(2)):
            if t.startswith("Ptr"):
                n = 8
            elif t.startswith("NTuple"):
                k, e = re.match(r"NTuple\{(\d+),(\w+)\}", t).groups()
                n = int(k) * _JL_SIZE[e]
            elif t in _JL_SIZE:
                n = _JL_SIZE[t]
            else:
                n = None  # a host-side Julia struct (not passed to C)
            fields.append((f, n))
        out[m.group(1)] = fields
    return out


def test_julia_bindings_match_the_ctypes_mirrors():
    """The Julia structs INTEGRATION.md gives a MonteCarloX.jl user have the
    fields, in order and of the sizes, of the mirrors checked above."""
    from montecarlox import _lib, muca, pt, serving
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    jl = _julia_structs(text)
    for name, cls in [("BatchDesc", _lib.BatchDesc), ("PTDesc", pt.PTDesc),
                      ("ServerDesc", _lib.ServerDesc), ("JobDesc", _lib.JobDesc),
                      ("JobResult", _lib.JobResult), ("LaunchShape", _lib.LaunchShape),
                      ("GenDesc", muca.GenDesc), ("GenSchedule", muca.GenSchedule),
                      ("GenProgress", muca.GenProgress), ("ServerStats", _lib.ServerStats),
                      ("SLOReport", serving.SLOReport)]:
        assert name in jl, name
        want = [(f, C.sizeof(t)) for f, t in cls._fields_]
        assert jl[name] == want, (name, jl[name], want)

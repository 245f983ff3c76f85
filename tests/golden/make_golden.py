"""Regenerates the committed fixtures in tests/golden/.

  philox_kat.json    Philox4x32-10 known-answer vectors.  The first three are
                     the published Random123 vectors (Salmon et al., SC'11,
                     kat_vectors "philox4x32 10"); every vector's output is
                     recomputed here by rocRAND's host engine
                     (philox_rocrand_check.cpp, built with hipcc) when hipcc is
                     available, and the run fails if the two disagree.
  bc_4x4_dos.json    exact density of states g(pair_sum, M, Q) of the 4x4
                     periodic Blume-Capel lattice, by enumerating all 3^16
                     configurations (about 90 s).
  ising2D_8x8.csv    (not generated) the reference's own exact Beale DOS data
                     file, SpinSystems/data/exact_solutions/ising2D_8x8.csv.

Run from the repository root:  python tests/golden/make_golden.py [--skip-bc]
"""
import json
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

RANDOM123 = [
    # ctr0..3, key0..1 -> out0..3
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]
EXTRA_INPUTS = [([1, 2, 3, 4], [7, 0]), ([17, 123456, (6 << 24) | (1 << 16), 0], [42, 4095])]


def rocrand_outputs():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        return None
    exe = "/tmp/philox_rocrand_check"
    subprocess.run([hipcc, "-O1", "-o", exe, os.path.join(HERE, "philox_rocrand_check.cpp")],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    res = []
    for line in out:
        if "->" not in line:
            continue
        lhs, rhs = line.split("->")
        v = [int(x, 16) for x in lhs.split()]
        res.append((v[:4], v[4:6], [int(x, 16) for x in rhs.split()]))
    return res


def make_philox():
    vecs = [{"ctr": c, "key": k, "out": o, "source": "Random123 kat_vectors"} for c, k, o in RANDOM123]
    rr = rocrand_outputs()
    if rr is not None:
        for (c, k, o), v in zip(rr, vecs):
            assert (c, k, o) == (v["ctr"], v["key"], v["out"]), "rocRAND disagrees with Random123"
            v["rocrand_checked"] = True
        for c, k, o in rr[len(vecs):]:
            vecs.append({"ctr": c, "key": k, "out": o, "source": "rocRAND host engine"})
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(vecs, f, indent=1)


def make_bc():
    from oracle import exact
    d = exact.dos_blume_capel_bruteforce(4)
    rows = sorted([k[0], k[1], k[2], v] for k, v in d.items())
    with open(os.path.join(HERE, "bc_4x4_dos.json"), "w") as f:
        json.dump({"L": 4, "columns": ["pair_sum", "M", "Q", "count"], "rows": rows}, f)


if __name__ == "__main__":
    make_philox()
    if "--skip-bc" not in sys.argv:
        make_bc()

"""Writes tests/golden/fixtures.json from the compiled reference.

Run in the container where /root/reference exists (oracle/Makefile builds the
reference's libqdt sources into oracle/_ref/libqdt_ref.so).  Every input here
is generated with exact IEEE operations only (splitmix64 uniforms, integer
lattices, sums of uniforms), so the same points come out on any machine and
the fixtures hold just the generator spec and the reference's answer:
triangle count, FNV-1a-64 of the canonical triangle list (each index as 8
little-endian bytes), and for small inputs the canonical list itself.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

def FIXTURE_INPUTS():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _fixture_inputs

    d = _fixture_inputs.inputs()
    d.update(_fixture_inputs.large_inputs())
    return d


def main():
    import pyoracle

    assert pyoracle.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    out = {"generator": "libqdt_amd.workloads + tests/_fixture_inputs.py", "cases": {}}
    for name, pts in FIXTURE_INPUTS().items():
        tris, used = pyoracle.ref_triangulate(pts, parallel=len(pts) > 2_000_000)
        canon = pyoracle.canonical(tris)
        case = {"n": int(len(pts)), "triangles": int(len(tris)), "fnv": "%016x" % pyoracle.canon_hash(tris)}
        if len(pts) <= 400:
            case["canonical"] = canon.tolist()
            case["used"] = used.tolist() if used is not None else None
        out["cases"][name] = case
        print(name, case["n"], case["triangles"], case["fnv"], flush=True)
    # the reference README's published counts (README.md:505-512, 518-528)
    out["published_counts"] = {
        "bench_uniform": {"100": 185, "1000": 1976, "10000": 19968, "100000": 199958, "1000000": 1999963},
        "dist_1M": {"uniform random": 1999953, "regular grid (shuffled)": 1996002,
                    "jittered grid (shuffled)": 1999961, "flat band (aspect 1e6:1)": 1999962,
                    "50% duplicate points": 999964, "uniform, presorted by x": 1999959},
    }
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

"""Extracts the numbers the reference's own committed logs hold for the
basis / geometry path into reference_logs.json (data only: inputs and
expected outputs).  Reads the reference tree; run here, not on the GPU box.

Sources (paths relative to the morganaPublic tree):
  tests/morganaFields/logs/fePr3dTest2.log      P1 on the reference tet at (1/4,1/4,1/4)
  tests/morganaFields/logs/fePr3dTest3.log      P2 on {(1,0,0),(0,0,0),(0,0,1),(0,1,0)}
                                                at (1/4,1/4,1/4)  (inputs: fePr3dTest3.cpp)
  tests/morganaGeometry/logs/geoMapInterfaceTest5.log   linearTetra map of the unit tet
                                                translated by (1,1,1), at Y=(0,.5,.5)
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_logs.json")


def blocks(path):
    lines = [l.strip() for l in open(path).read().split("\n")]
    out, key = {}, None
    for l in lines:
        if l in ("LocEval", "LocGradX", "LocGradY", "LocGradZ", "GlobalEval", "GlobalGradX",
                 "GlobalGradY", "GlobalGradZ"):
            key = l
            out[key] = []
        elif key and l:
            try:
                out[key].append(float(l))
            except ValueError:
                key = None
        elif key and not l and out[key]:
            key = None
    return out


def main():
    d = {}
    d["fePr3dTest2"] = {
        "source": "tests/morganaFields/logs/fePr3dTest2.log",
        "nodes": [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
        "Y": [0.25, 0.25, 0.25],
        **blocks(os.path.join(REF, "tests/morganaFields/logs/fePr3dTest2.log")),
    }
    d["fePr3dTest3"] = {
        "source": "tests/morganaFields/logs/fePr3dTest3.log",
        "nodes": [[1, 0, 0], [0, 0, 0], [0, 0, 1], [0, 1, 0]],
        "Y": [0.25, 0.25, 0.25],
        **blocks(os.path.join(REF, "tests/morganaFields/logs/fePr3dTest3.log")),
    }
    geo = open(os.path.join(REF, "tests/morganaGeometry/logs/geoMapInterfaceTest5.log")).read()
    ev = [float(v) for v in geo.split("eval:")[1].split("\n")[0].split()]
    g = geo.split("Tensor components")[1].strip().split("\n")[:3]
    grad = [[float(v) for v in row.split()] for row in g]
    d["geoMapInterfaceTest5"] = {
        "source": "tests/morganaGeometry/logs/geoMapInterfaceTest5.log",
        "nodes": [[1, 1, 1], [2, 1, 1], [1, 2, 1], [1, 1, 2]],
        "Y": [0.0, 0.5, 0.5],
        "eval": ev,
        "gradient": grad,
    }
    d["fePr3dTest2"]["inputs_from"] = "tests/morganaFields/fePr3dTest2.cpp"
    d["fePr3dTest3"]["inputs_from"] = "tests/morganaFields/fePr3dTest3.cpp"
    d["geoMapInterfaceTest5"]["inputs_from"] = "tests/morganaGeometry/geoMapInterfaceTest5.cpp"
    json.dump(d, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

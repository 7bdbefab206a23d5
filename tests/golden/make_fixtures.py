"""Regenerate the golden fixtures from the reference tree (run in a container
where the reference is mounted; the fixtures themselves are committed).

hilbert_kat.json   <- UnitTest/test_hilbert.cpp (known answers of hilbert::index)
lattice_tables.json <- src/lbm/constants.h (LBMethod<D2Q9>/<D3Q19> velocity,
                       weight and opposite tables, parsed as data)
reference_configs/  <- test/{couette,poiseuille,poisson}/*.json, test/sphere, test/step
                       (the reference's own integration-test inputs)
"""
import json
import os
import re
import shutil
import sys
from fractions import Fraction

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def hilbert():
    src = open(os.path.join(REF, "UnitTest/test_hilbert.cpp")).read()
    cases, cur = [], {}
    for line in src.splitlines():
        m = re.search(r"VectorD<(\d)>\s+(\w+)\s*=\s*\{([^}]*)\}", line)
        if m:
            cur[m.group(2)] = [float(v) for v in m.group(3).split(",")]
            continue
        m = re.search(r"hilbert::index<(\d)>\((\w+),\s*(\d+)\)\s*,\s*(\d+)\)", line)
        if m:
            cases.append({"dim": int(m.group(1)), "x": cur[m.group(2)], "level": int(m.group(3)),
                          "index": int(m.group(4))})
    json.dump({"source": "UnitTest/test_hilbert.cpp", "cases": cases},
              open(os.path.join(OUT, "hilbert_kat.json"), "w"), indent=1)


def tables():
    src = open(os.path.join(REF, "src/lbm/constants.h")).read()
    out = {}
    for name in ("D2Q9", "D3Q19"):
        m = re.search(r"class LBMethod<LBMethodType::%s>(.*?)m_name" % name, src, re.S)
        body = m.group(1)
        dirs = re.search(r"m_dirs = \{(.*?)\}\};", body, re.S).group(1)
        vecs = [[float(v) for v in t.split(",")] for t in re.findall(r"\{\{([-\d., ]+)\}\}", dirs)]
        opp = [int(v) for v in re.search(r"m_oppositeDist = \{([^}]*)\}", body).group(1).split(",")]
        wtxt = re.search(r"m_weights = \{([^}]*)\}", body).group(1)
        ws = [str(Fraction(int(a), int(b))) for a, b in re.findall(r"(\d+)\.0\s*/\s*(\d+)\.0", wtxt)]
        out[name] = {"c": vecs, "w": ws, "opp": opp}
    json.dump({"source": "src/lbm/constants.h", "tables": out},
              open(os.path.join(OUT, "lattice_tables.json"), "w"), indent=1)


def configs():
    dst = os.path.join(OUT, "reference_configs")
    os.makedirs(dst, exist_ok=True)
    for sub in ("couette", "poiseuille", "poisson", "sphere", "step"):
        for f in sorted(os.listdir(os.path.join(REF, "test", sub))):
            if f.endswith(".json"):
                shutil.copy(os.path.join(REF, "test", sub, f), os.path.join(dst, f"{sub}__{f}"))


if __name__ == "__main__":
    hilbert()
    tables()
    configs()

"""Writes the golden fixtures under tests/golden/ from the reference's data files.

Run in a container where the reference tree is mounted (default
/root/reference).  It reads only data:

* dictionaries/{3,4,5,6}.csv -> dict_k.json with the quick -> cgLocal table,
  the cgLocal -> cgGlobal list, the per-class quick lists and the SHA-256 of
  the source file (format parsed as QuickMapping::QuickMapping does,
  src/main/QuickMapping.cpp:24-39);
* datasets/test.edgelist and datasets/citeseer.edgelist, copied verbatim to
  tests/golden/datasets/.

The committed fixtures are what the tests use, so a machine without the
reference tree runs from them.
"""
import hashlib
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_dictionary(path):
    with open(path) as f:
        lines = [ln.strip() for ln in f if ln.strip()]
    i = 0
    assert lines[i].startswith("#quick"), lines[i]
    i += 1
    nq = int(lines[i])
    i += 1
    q2c = [None] * nq
    for _ in range(nq):
        q, c = lines[i].split(",")
        q2c[int(q)] = int(c)
        i += 1
    assert lines[i].startswith("#cg_local,cg_global"), lines[i]
    i += 1
    nc = int(lines[i])
    i += 1
    c2g = [None] * nc
    for _ in range(nc):
        c, g = lines[i].split(",")
        c2g[int(c)] = int(g)
        i += 1
    assert lines[i].startswith("#cg_local,amount"), lines[i]
    i += 1
    nc2 = int(lines[i])
    i += 1
    members = [None] * nc2
    for _ in range(nc2):
        parts = [int(x) for x in lines[i].split(",")]
        members[parts[0]] = parts[2:2 + parts[1]]
        assert len(members[parts[0]]) == parts[1]
        i += 1
    assert None not in q2c and None not in c2g and None not in members
    return q2c, c2g, members


def main(ref="/root/reference"):
    os.makedirs(os.path.join(HERE, "datasets"), exist_ok=True)
    for k in (3, 4, 5, 6):
        src = os.path.join(ref, "dictionaries", "%d.csv" % k)
        with open(src, "rb") as f:
            digest = hashlib.sha256(f.read()).hexdigest()
        q2c, c2g, members = parse_dictionary(src)
        out = {
            "k": k,
            "source": "dictionaries/%d.csv" % k,
            "sha256": digest,
            "quick_to_cg_local": q2c,
            "cg_local_to_global": c2g,
            "class_quicks": members,
        }
        with open(os.path.join(HERE, "dict_%d.json" % k), "w") as f:
            json.dump(out, f, separators=(",", ":"))
        print("dict_%d.json: %d quicks, %d classes" % (k, len(q2c), len(c2g)))
    for name in ("test.edgelist", "citeseer.edgelist"):
        shutil.copyfile(os.path.join(ref, "datasets", name),
                        os.path.join(HERE, "datasets", name))
        print("datasets/%s" % name)


if __name__ == "__main__":
    main(*sys.argv[1:])

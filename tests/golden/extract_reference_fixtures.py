"""Writes tests/golden/{nodes_tp,operators_tp}.json from the reference's unit-test data files.

Reads (data only, no reference code is run):
  input/testing/unit/nodes/nodes_tp.data          (used by src/testing/unit/test_unit_nodes.c)
  input/testing/unit/operators/operators_tp.data  (used by src/testing/unit/test_unit_operators_tp.c:88-170)
under the DPGSolver checkout given as argv[1] (default /root/reference).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _read_matrix(lines, i):
    """lines[i] is 'container Matrix_d'; returns (matrix rows, next index) or (None, next) for NULL."""
    assert lines[i].strip().startswith("container Matrix_d"), lines[i]
    hdr = lines[i + 1].split()
    if hdr[0] == "NULL":
        return None, i + 2
    r, c = int(hdr[1]), int(hdr[2])
    rows = [[float(v) for v in lines[i + 2 + k].split()] for k in range(r)]
    assert all(len(x) == c for x in rows)
    return rows, i + 2 + r


def _read_multiarray(lines, i):
    assert lines[i].strip().startswith("container Multiarray_d"), lines[i]
    hdr = lines[i + 1].split()
    order = int(hdr[1])
    ext = [int(v) for v in hdr[2:2 + order]]
    r = ext[0]
    rows = [[float(v) for v in lines[i + 2 + k].split()] for k in range(r)]
    return rows, i + 2 + r


def nodes(path):
    lines = open(path).read().splitlines()
    out = {}
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if s.startswith("Nodes_d"):
            name = s
            has_w = int(lines[i + 2].split(":")[1])
            assert lines[i + 3].strip() == "rst:"
            rst, j = _read_matrix(lines, i + 4)
            w = None
            if has_w:
                assert lines[j].strip() == "w:"
                w, j = _read_matrix(lines, j + 1)
                w = [x[0] for x in w]
            out[name] = {"rst": rst, "w": w}
            i = j
        else:
            i += 1
    return out


def operators(path):
    lines = open(path).read().splitlines()
    out = {}
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if s.endswith("_ops_tp:"):
            key = s[:-1]
            n = int(lines[i + 3].split()[1])
            j = i + 4
            mats = []
            for _ in range(n):
                m, j = _read_matrix(lines, j)
                mats.append(m)
            out[key] = mats
            i = j
        elif s.endswith("_op_std:"):
            m, i = _read_matrix(lines, i + 1)
            out[s[:-1]] = m
        elif s.endswith("_b:") or s.endswith("_c:"):
            m, i = _read_multiarray(lines, i + 1)
            out[s[:-1]] = m
        else:
            i += 1
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    unit = os.path.join(ref, "input", "testing", "unit")
    n = nodes(os.path.join(unit, "nodes", "nodes_tp.data"))
    o = operators(os.path.join(unit, "operators", "operators_tp.data"))
    with open(os.path.join(HERE, "nodes_tp.json"), "w") as f:
        json.dump({"source": "input/testing/unit/nodes/nodes_tp.data", "nodes": n}, f, indent=0)
    with open(os.path.join(HERE, "operators_tp.json"), "w") as f:
        json.dump({"source": "input/testing/unit/operators/operators_tp.data", "operators": o}, f, indent=0)
    print("nodes:", sorted(n), "operators:", sorted(o))


if __name__ == "__main__":
    main()

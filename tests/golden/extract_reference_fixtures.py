"""Writes tests/golden/*.json from the reference's unit- and integration-test data files.

Reads (data only, no reference code is run):
  input/testing/unit/nodes/nodes_tp.data          (used by src/testing/unit/test_unit_nodes.c)
  input/testing/unit/operators/operators_tp.data  (used by src/testing/unit/test_unit_operators_tp.c:88-170)
  input/testing/unit/nodes/nodes_correspondence.data (face node permutations per ind_ord,
                                                  src/element/nodes_correspondence.c)
  input/testing/integration/fe/*.fe.data          (Face::neigh_info{ind_v, ind_lf, ind_ord} of the shipped meshes,
                                                  compared by src/testing/support/test_support_fe_init.c:222)
  input/testing/unit/nodes/nodes_si.data          (2-D simplex nodes, used by src/testing/unit/test_unit_nodes.c)
  input/nodes/tri/WSH{1..5}.txt, WV{2..10}.txt    (the symmetric TRI node and cubature tables the reference reads
                                                  at run time, src/element/nodes.c:628-852)
  input/testing/integration/mesh/*.msh(.data)     (the 2-D test meshes and the connectivity the reference's mesh
                                                  integration test compares, copied as data into tests/golden/mesh)
under the DPGSolver checkout given as argv[1] (default /root/reference).
"""
import glob
import json
import os
import re
import shutil
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


def _read_vector_i_list(lines, i):
    """lines[i] is 'container Multiarray_Vector_i'; returns (list of int rows, next index)."""
    assert lines[i].strip() == "container Multiarray_Vector_i", lines[i]
    n = int(lines[i + 2].split()[1])
    assert lines[i + 3].strip() == "ext_0/data"
    rows = []
    for k in range(n):
        vals = [int(v) for v in lines[i + 4 + k].split()]
        assert vals[0] == len(vals) - 1
        rows.append(vals[1:])
    return rows, i + 4 + n


def correspondence(path):
    lines = open(path).read().splitlines()
    out = {}
    for i, s in enumerate(lines):
        s = s.strip()
        if re.fullmatch(r"(point|line|quad|tri)_\d+:", s):
            out[s[:-1]], _ = _read_vector_i_list(lines, i + 1)
    return out


def fe_faces(path):
    """Per face: [boundary, v0, lf0, ord0, v1, lf1, ord1] from a .fe.data file."""
    lines = [x.strip() for x in open(path).read().splitlines()]
    out = []
    i = 0
    while i < len(lines):
        if lines[i].startswith("Face "):
            bnd = int(lines[i + 1].split()[1])
            rec = [bnd]
            j = i + 1
            for side in (0, 1):
                while lines[j] != f"Neigh_Info {side}":
                    j += 1
                kv = {}
                for k in range(1, 6):
                    key, val = lines[j + k].split()
                    kv[key] = int(val)
                rec += [kv["ind_v"], kv["ind_lf"], kv["ind_ord"]]
                j += 6
            out.append(rec)
            i = j
        else:
            i += 1
    return out


def simplex_nodes(path):
    """Nodes_d2_* blocks of nodes_si.data: {name: {"rst": rows, "w": weights or None}}."""
    lines = open(path).read().splitlines()
    out = {}
    i = 0
    while i < len(lines):
        st = lines[i].strip()
        if st.startswith("Nodes_d2_"):
            has_w = int(lines[i + 2].split(":")[1])
            assert lines[i + 3].strip() == "rst:"
            rst, j = _read_matrix(lines, i + 4)
            w = None
            if has_w:
                assert lines[j].strip() == "w:"
                w, j = _read_matrix(lines, j + 1)
                w = [x[0] for x in w]
            out[st] = {"rst": rst, "w": w}
            i = j
        else:
            i += 1
    return out


def tri_table(path):
    """One input/nodes/tri/*.txt file: symmetry orbits as the reference stores them."""
    lines = [x for x in open(path).read().splitlines()]
    ngroups, nsymms = (int(v) for v in lines[0].split())
    symms = [[int(v) for v in lines[1 + k].split()] for k in range(nsymms)]
    k = 1 + nsymms
    while not lines[k].strip().startswith("Barycentric"):
        k += 1
    bary = [[float(v) for v in lines[k + 1 + g].split()] for g in range(ngroups)]
    k += 1 + ngroups
    while not lines[k].strip().startswith("Weights"):
        k += 1
    w = [float(lines[k + 1 + g].split()[0]) for g in range(ngroups)]
    return {"symms": symms, "bary": bary, "w": w}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    unit = os.path.join(ref, "input", "testing", "unit")
    n = nodes(os.path.join(unit, "nodes", "nodes_tp.data"))
    o = operators(os.path.join(unit, "operators", "operators_tp.data"))
    with open(os.path.join(HERE, "nodes_tp.json"), "w") as f:
        json.dump({"source": "input/testing/unit/nodes/nodes_tp.data", "nodes": n}, f, indent=0)
    with open(os.path.join(HERE, "operators_tp.json"), "w") as f:
        json.dump({"source": "input/testing/unit/operators/operators_tp.data", "operators": o}, f, indent=0)
    c = correspondence(os.path.join(unit, "nodes", "nodes_correspondence.data"))
    with open(os.path.join(HERE, "nodes_correspondence.json"), "w") as f:
        json.dump({"source": "input/testing/unit/nodes/nodes_correspondence.data", "corr": c}, f, indent=0)
    fe = {}
    for path in sorted(glob.glob(os.path.join(ref, "input", "testing", "integration", "fe", "*.fe.data"))):
        fe[os.path.basename(path)[:-len(".fe.data")]] = fe_faces(path)
    with open(os.path.join(HERE, "fe_faces.json"), "w") as f:
        json.dump({"source": "input/testing/integration/fe/*.fe.data",
                   "columns": ["boundary", "ind_v0", "ind_lf0", "ind_ord0", "ind_v1", "ind_lf1", "ind_ord1"],
                   "faces": fe}, f, indent=0)
    si = simplex_nodes(os.path.join(unit, "nodes", "nodes_si.data"))
    tables = {}
    for kind, orders in (("WSH", range(1, 6)), ("WV", range(2, 11, 2))):
        for q in orders:
            tables[f"{kind}{q}"] = tri_table(os.path.join(ref, "input", "nodes", "tri", f"{kind}{q}.txt"))
    with open(os.path.join(HERE, "tri_nodes.json"), "w") as f:
        json.dump({"source": "input/testing/unit/nodes/nodes_si.data (golden), input/nodes/tri/*.txt (tables)",
                   "golden": si, "tables": tables}, f, indent=0)
    meshes = []
    for path in sorted(glob.glob(os.path.join(ref, "input", "testing", "integration", "mesh", "*2d*.msh*"))):
        shutil.copyfile(path, os.path.join(HERE, "mesh", os.path.basename(path)))
        meshes.append(os.path.basename(path))
    print("nodes:", sorted(n), "operators:", sorted(o), "corr:", sorted(c), "fe:", sorted(fe), "si:", sorted(si),
          "tri tables:", sorted(tables), "meshes:", meshes)


if __name__ == "__main__":
    main()

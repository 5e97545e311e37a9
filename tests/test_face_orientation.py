"""Face orientation codes at the boundary: the reference's ind_ord, pinned by its own fixtures.

* input/testing/unit/nodes/nodes_correspondence.data (tests/golden/nodes_correspondence.json): the node
  permutation of every orientation case for LINE faces of 3 and 4 nodes and QUAD faces of 9 and 16 nodes
  (constructor_face_corr, src/element/nodes_correspondence.c).  The oracle, the mesh generator and the
  library (dgx_face_permutation, host only) must all reproduce it.
* input/testing/integration/fe/*.fe.data (tests/golden/fe_faces.json): Face::neigh_info[*].ind_ord of the
  shipped QUAD meshes; dgx_mesh_read_gmsh must write the same code for the same face.
* In 3-D the code written into the face table must be the one whose node table maps each left face node to
  the coincident right face node.  The reference derives ind_ord from a vertex match (face.c:498-540); that is
  restated here and compared on meshes with every orientation.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as O
from dpgsolver_amd import mesh as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CORR = json.load(open(os.path.join(GOLD, "nodes_correspondence.json")))["corr"]
FE = json.load(open(os.path.join(GOLD, "fe_faces.json")))["faces"]
CASES = [("line_3", 2, 2), ("line_4", 2, 3), ("quad_9", 3, 2), ("quad_16", 3, 3)]


def _oracle_perm(d, p, code):
    mesh = M.structured_tp_mesh(d, p, (2,) * d, (0,) * d, (1,) * d, (True,) * d)
    orc = O.Oracle(mesh, M.PDE_EULER, M.FLUX_LF)
    perm = np.zeros((p + 1) ** (d - 1), dtype=np.int32)
    O.lib().or_face_perm(orc.ctx, code, perm.ctypes.data_as(O.ip))
    return perm


@pytest.mark.parametrize("name,d,p", CASES)
def test_oracle_face_perm_matches_reference_table(name, d, p):
    for code, want in enumerate(CORR[name]):
        np.testing.assert_array_equal(_oracle_perm(d, p, code), want, err_msg=f"{name} case {code}")


@pytest.mark.parametrize("name,d,p", CASES)
def test_mesh_generator_face_perm_matches_reference_table(name, d, p):
    assert len(CORR[name]) == M.n_orientations(d)
    for code, want in enumerate(CORR[name]):
        np.testing.assert_array_equal(M.face_perm(d, p + 1, code), want, err_msg=f"{name} case {code}")


@pytest.mark.parametrize("name,d,p", CASES)
def test_library_face_perm_matches_reference_table(name, d, p):
    from dpgsolver_amd import _lib

    for code, want in enumerate(CORR[name]):
        np.testing.assert_array_equal(_lib.face_permutation(d, p, code), want, err_msg=f"{name} case {code}")


def test_library_rejects_codes_without_meaning():
    from dpgsolver_amd import _lib

    for d, bad in ((2, 2), (2, 4), (2, -1), (3, 8), (3, -1)):
        with pytest.raises(_lib.DgxError):
            _lib.face_permutation(d, 3, bad)


def test_bit_coded_orientation_fails_the_table():
    """The planted mistake this pin exists for: reading a LINE face's ind_ord = 1 as orientation bits (only
    bit 1 reverses) treats a reversed face as aligned."""
    bits_reading = np.arange(3)[::-1] if (1 & 2) else np.arange(3)
    assert not np.array_equal(bits_reading, CORR["line_3"][1])
    # and a transposed reading of QUAD case 5 (its inverse, case 6) differs from the table
    assert not np.array_equal(np.argsort(CORR["quad_9"][5]), CORR["quad_9"][5])


def _fixture_codes(name):
    """{frozenset({(v0, lf0), (v1, lf1)}): (ind_ord0, ind_ord1)} for the interior faces of a .fe.data file."""
    out = {}
    for bnd, v0, lf0, o0, v1, lf1, o1 in FE[name]:
        if bnd:
            assert o0 == -1 and o1 == -1
            continue
        out[frozenset({(v0, lf0), (v1, lf1)})] = (o0, o1)
    return out


@pytest.mark.parametrize("name", ["straight_2d_quad_periodic", "parametric_2d_quad_periodic_reflected"])
def test_gmsh_reader_writes_the_reference_ind_ord(name):
    from dpgsolver_amd import _lib

    m = _lib.read_gmsh(os.path.join(GOLD, "mesh", name + ".msh"), 2)
    want = _fixture_codes("TEST_" + name)
    got = {}
    for eL, fL, eR, fR, ord_, bc in m["faces"]:
        if bc == M.BC_INTERIOR:
            got[frozenset({(int(eL), int(fL)), (int(eR), int(fR))})] = int(ord_)
    assert set(got) == set(want)
    for k, code in got.items():
        assert (code, code) == want[k], (sorted(k), code, want[k])
    if name.startswith("parametric"):
        assert sum(v == 1 for v in got.values()) == 2       # the y-reflected periodic pair (ind_ord 1 on both sides)


# ------------------------------------------------------------------ 3-D: which neigh_info's code to pass
# face.c:516-528 vertex patterns of a QUAD face (vertex (a, b) = a + 2 b, a the first face coordinate)
_QUAD_MATCHES = [[0, 1, 2, 3], [1, 0, 3, 2], [2, 3, 0, 1], [3, 2, 1, 0], [0, 2, 1, 3], [2, 0, 3, 1], [1, 3, 0, 2],
                 [3, 1, 2, 0]]


def _face_corners(mesh, e, lf):
    N = mesh.p + 1
    idx = M.face_node_volume_index(mesh.d, N, lf)
    corners = [0, N - 1, N * (N - 1), N * N - 1]            # face nodes (a, b) at the vertices, a fastest
    return mesh.xyz[e][:, idx[corners]].T


def _matches(xm, xs, period):
    """constructor_matches_Vector_i_Matrix_d (face.c:636-690): for each master vertex, the slave vertex at the
    same place (here modulo the periodic box)."""
    out = []
    for a in xm:
        dd = a[None, :] - xs
        dd -= period * np.round(dd / period)
        out.append(int(np.argmin(np.abs(dd).sum(1))))
    return out


def test_3d_code_is_the_vertex_pattern_of_the_left_to_right_match():
    rot = np.random.default_rng(21).integers(0, 8, 27)
    mesh = M.structured_tp_mesh(3, 2, (3, 3, 3), (0,) * 3, (1,) * 3, (True,) * 3, rotate=rot)
    assert set(np.unique(mesh.faces[:, 4])) == set(range(8))
    inverse_pairs = 0
    for eL, fL, eR, fR, ord_, bc in mesh.faces:
        xL, xR = _face_corners(mesh, eL, fL), _face_corners(mesh, eR, fR)
        r_to_l = _matches(xL, xR, 1.0)      # neigh_info[1].ind_ord comes from this pattern
        l_to_r = _matches(xR, xL, 1.0)      # neigh_info[0].ind_ord from this one
        ind_ord1, ind_ord0 = _QUAD_MATCHES.index(r_to_l), _QUAD_MATCHES.index(l_to_r)
        assert ord_ == ind_ord0
        if ind_ord0 != ind_ord1:
            assert {ind_ord0, ind_ord1} == {5, 6}
            inverse_pairs += 1
    assert inverse_pairs > 0

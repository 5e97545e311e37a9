"""gmsh mesh ingest (dgx_mesh_read_gmsh) against the reference's own connectivity fixtures.

The `.msh.data` files are the reference's expected Mesh_Connectivity (v_to_v, v_to_lf, v_to_lf_wp) for the
shipped test meshes (input/testing/integration/mesh/, copied to tests/golden/mesh/).  Covered: a periodic
QUAD mesh, and a QUAD mesh with Riemann/slip-wall boundaries and a periodic boundary reflected in y; and, through
dgx_mesh_read_gmsh_mixed, the mixed TRI/QUAD mesh blended_2d_mixed (the reference's v_to_v and v_to_lf for its 8 TRI
and 4 QUAD volumes), every 2-D test mesh the reference ships, and the geometry of every interior face seen from
its two sides.  The QUAD reader refuses a TRI mesh.  The readers are host code: these tests call no GPU entry point.
"""
import os

import numpy as np
import pytest

from dpgsolver_amd import mesh as M

HERE = os.path.join(os.path.dirname(__file__), "golden", "mesh")


def _golden(path, name):
    """Parse one Multiarray_Vector_i block of a reference .msh.data file."""
    lines = open(path).read().splitlines()
    i = lines.index(name + ":")
    ext = int(lines[i + 3].split()[1])
    rows = []
    for k in range(ext):
        vals = [int(x) for x in lines[i + 5 + k].split()]
        assert vals[0] == len(vals) - 1
        rows.append(vals[1:])
    return np.array(rows)


@pytest.mark.parametrize("name", ["straight_2d_quad_periodic", "parametric_2d_quad_periodic_reflected"])
def test_connectivity_matches_reference_fixture(name):
    from dpgsolver_amd import _lib

    m = _lib.read_gmsh(os.path.join(HERE, name + ".msh"), 2)
    data = os.path.join(HERE, name + ".msh.data")
    for key in ("v_to_v", "v_to_lf", "v_to_lf_wp"):
        np.testing.assert_array_equal(m[key], _golden(data, key), err_msg=key)
    assert m["n_inverted"] == 0


def test_faces_and_geometry_of_periodic_mesh():
    from dpgsolver_amd import _lib

    m = _lib.read_gmsh(os.path.join(HERE, "straight_2d_quad_periodic.msh"), 3)
    assert (m["E"], m["F"]) == (4, 8)
    assert np.all(m["faces"][:, 2] >= 0)                       # fully periodic: no boundary faces
    # every element is a 1x1 square of [-1,1]^2 (the gmsh midpoints sit at -2.75e-12)
    x = m["xyz"]
    for e in range(4):
        span = x[e].max(axis=1) - x[e].min(axis=1)
        np.testing.assert_allclose(span, 1.0, atol=1e-11)
    # the GLL nodes of degree 3 along an element edge
    r = M.gauss_lobatto(4)[0]
    e0 = x[0][0, :4]
    np.testing.assert_allclose((e0 - e0[0]) * 2 - 1, r, atol=1e-11)


def test_boundary_faces_carry_the_reference_bc():
    from dpgsolver_amd import _lib

    m = _lib.read_gmsh(os.path.join(HERE, "parametric_2d_quad_periodic_reflected.msh"), 2)
    f = m["faces"]
    b = f[f[:, 2] < 0]
    # x101 (Riemann) and x102 (slip wall) tags map onto the library's codes
    tags = {(int(e), int(lf)): int(m["v_to_lf"][e, lf]) for e, lf in b[:, :2]}
    for (e, lf), row in zip(tags, b):
        want = M.BC_RIEMANN if tags[(e, lf)] % 10000 == 101 else M.BC_SLIPWALL
        assert row[5] == want
    assert m["n_unsupported_bc"] == 0


def test_tri_mesh_is_refused():
    from dpgsolver_amd import _lib

    with pytest.raises(_lib.DgxError):
        _lib.read_gmsh(os.path.join(HERE, "straight__2d__tri_ml0.msh"), 2)
    with pytest.raises(_lib.DgxError):
        _lib.read_gmsh(os.path.join(HERE, "no_such_file.msh"), 2)


# ---------------------------------------------------------------------------------------------- TRI and mixed meshes
MESHES_2D = ["blended_2d_mixed", "parametric_2d_quad_periodic_reflected", "parametric_vortex__2d__mixed_ml0",
             "parametric_vortex__2d__unaligned_mixed_ml1", "straight_2d_quad_periodic", "straight__2d__mixed_ml0",
             "straight__2d__quad_ml0", "straight__2d__tri_ml-1", "straight__2d__tri_ml0"]


def _golden_rows(path, name):
    """A Multiarray_Vector_i block with rows of different lengths (TRI 3, QUAD 4)."""
    lines = open(path).read().splitlines()
    i = lines.index(name + ":")
    ext = int(lines[i + 3].split()[1])
    rows = []
    for k in range(ext):
        vals = [int(x) for x in lines[i + 5 + k].split()]
        assert vals[0] == len(vals) - 1
        rows.append(vals[1:])
    return rows


def test_mixed_connectivity_matches_reference_fixture():
    """blended_2d_mixed: 8 TRI and 4 QUAD volumes with Riemann / slip-wall boundaries; the reference's
    v_to_v and v_to_lf (its Mesh_Connectivity) for every volume, TRI faces opposite their vertices."""
    from dpgsolver_amd import _lib

    m = _lib.read_gmsh_mixed(os.path.join(HERE, "blended_2d_mixed.msh"), 2)
    data = os.path.join(HERE, "blended_2d_mixed.msh.data")
    assert m["type"].tolist() == [1] * 8 + [0] * 4
    for key in ("v_to_v", "v_to_lf"):
        gold = _golden_rows(data, key)
        assert len(gold) == m["E"]
        for e, row in enumerate(gold):
            assert m[key][e][:len(row)].tolist() == row, (key, e)
            assert (m[key][e][len(row):] == -1).all()
    assert m["n_inverted"] == 0 and m["n_unsupported_bc"] == 0


@pytest.mark.parametrize("name", MESHES_2D)
def test_every_reference_2d_mesh_loads(name):
    from dpgsolver_amd import _lib

    for p in (1, 3):
        m = _lib.read_gmsh_mixed(os.path.join(HERE, name + ".msh"), p)
        assert m["E"] > 0 and m["n_inverted"] == 0
        if not m["type"].any():   # QUAD only: the same mesh as the QUAD reader's
            q = _lib.read_gmsh(os.path.join(HERE, name + ".msh"), p)
            np.testing.assert_array_equal(np.stack(m["xyz"]), q["xyz"])
            for key in ("faces", "v_to_v", "v_to_lf", "v_to_lf_wp"):
                np.testing.assert_array_equal(m[key], q[key], err_msg=key)


def _msh_volumes(path):
    """Vertex coordinates of the 2-D volumes of a gmsh 2.2 file, in file order (TRI 3, QUAD 4 vertices)."""
    lines = open(path).read().splitlines()
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    xy = {int(t[0]): (float(t[1]), float(t[2])) for t in (lines[i + 2 + k].split() for k in range(n))}
    j = lines.index("$Elements")
    vols = []
    for k in range(int(lines[j + 1])):
        t = [int(v) for v in lines[j + 2 + k].split()]
        if t[1] in (2, 3):
            vols.append(np.array([xy[v] for v in t[3 + t[2]:]]))
    return vols


@pytest.mark.parametrize("name", [n for n in MESHES_2D if "quad" not in n or "mixed" in n])
def test_mixed_faces_meet_at_the_same_points(name):
    """Every interior face seen from both sides: the face's GL points mapped through each side's own geometry and
    the orientation code coincide (pins the TRI face numbering, the face vertex order and ind_ord); each TRI's
    WSH nodes averaged with the WSH weights are the centroid of its three vertices (the affine map of a symmetric
    rule)."""
    import mixed_oracle as MO
    from dpgsolver_amd import _lib

    vols = _msh_volumes(os.path.join(HERE, name + ".msh"))
    for p in (1, 2, 4):
        mesh, order = M.mixed_mesh_from_gmsh(os.path.join(HERE, name + ".msh"), p)
        assert (np.diff(mesh.kind) >= 0).all()
        orc = MO.MixedOracle(mesh, M.FLUX_LF)
        u = np.zeros(orc.off[-1])
        n_int = 0
        for row in mesh.faces:
            if row[5] != 0:
                continue
            fs = orc._face_setup(u, row)
            xL = fs["IL"] @ mesh.xyz[row[0]].T
            xR = fs["IR"] @ mesh.xyz[row[2]].T
            np.testing.assert_allclose(xL, xR, atol=1e-12, err_msg=f"{name} p={p} face {row}")
            n_int += 1
        assert n_int > 0
        _, w = _lib.reference_nodes(1, p)
        for e in np.nonzero(mesh.kind == 1)[0]:
            X = vols[order[e]]
            assert X.shape == (3, 2)
            np.testing.assert_allclose(mesh.xyz[e] @ w / w.sum(), X.mean(axis=0), atol=1e-13)

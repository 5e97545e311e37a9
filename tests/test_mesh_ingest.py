"""gmsh mesh ingest (dgx_mesh_read_gmsh) against the reference's own connectivity fixtures.

The `.msh.data` files are the reference's expected Mesh_Connectivity (v_to_v, v_to_lf, v_to_lf_wp) for the
shipped test meshes (input/testing/integration/mesh/, copied to tests/golden/mesh/).  Covered: a periodic
QUAD mesh, and a QUAD mesh with Riemann/slip-wall boundaries and a periodic boundary reflected in y.  A TRI
mesh is refused.  The reader is host code: these tests call no GPU entry point.
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

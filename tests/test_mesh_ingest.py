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


# ----------------------------------------------------------------------------- 3-D HEX meshes
def _cube_rotations():
    """The 24 orientation-preserving symmetries of the unit cube as maps of corner coordinates (i0, i1, i2)."""
    import itertools

    out = []
    for perm in itertools.permutations(range(3)):
        sgn = np.linalg.det(np.eye(3)[list(perm)])
        for flips in itertools.product((0, 1), repeat=3):
            if sgn * (-1) ** sum(flips) > 0:
                out.append((perm, flips))
    return out


def _write_hex_gmsh(path, n, periodic, seed):
    """A gmsh 2.2 file of the n^3 unit-cube HEX mesh: each element's vertices listed after a random cube rotation,
    boundary QUADs tagged 51..56 (periodic master/slave per direction) or 10101 (Riemann)."""
    from dpgsolver_amd import mesh as M

    bc = None if periodic else (M.BC_RIEMANN,) * 3
    m1 = M.structured_tp_mesh(3, 1, (n, n, n), (0, 0, 0), (1, 1, 1), (periodic,) * 3, bc_low=bc, bc_high=bc)
    corners = np.transpose(m1.xyz, (0, 2, 1))              # (E, 8, 3) in tensor-product order
    ids, pts = {}, []
    def nid(x):
        k = tuple(np.round(x, 12))
        if k not in ids:
            ids[k] = len(pts) + 1
            pts.append(x)
        return ids[k]
    rots = _cube_rotations()
    rng = np.random.default_rng(seed)
    gm = (0, 1, 3, 2, 4, 5, 7, 6)                          # gmsh HEX vertex j = tensor-product corner gm[j]
    hexes = []
    for e in range(m1.E):
        perm, flips = rots[rng.integers(len(rots))]
        vs = []
        for j in range(8):
            c = [(gm[j] >> k) & 1 for k in range(3)]
            c = [c[perm[k]] ^ flips[k] for k in range(3)]
            vs.append(nid(corners[e, c[0] + 2 * c[1] + 4 * c[2]]))
        hexes.append(vs)
    quads = []
    for dirn in range(3):
        for side in (0, 1):
            for e in range(m1.E):
                sel = [k for k in range(8) if ((k >> dirn) & 1) == side]
                X = corners[e, sel]
                if np.allclose(X[:, dirn], float(side)):
                    tag = (51 + 2 * dirn + side) if periodic else 10101
                    quads.append((tag, [nid(x) for x in X[[0, 1, 3, 2]]]))
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(pts))
        for i, x in enumerate(pts):
            f.write("%d %.17g %.17g %.17g\n" % (i + 1, *x))
        f.write("$EndNodes\n$Elements\n%d\n" % (len(quads) + len(hexes)))
        k = 1
        for tag, vs in quads:
            f.write("%d 3 2 %d %d %s\n" % (k, tag, tag, " ".join(map(str, vs))))
            k += 1
        for vs in hexes:
            f.write("%d 5 2 1 1 %s\n" % (k, " ".join(map(str, vs))))
            k += 1
        f.write("$EndElements\n")


@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "riemann"])
def test_hex_gmsh_reader_matches_the_generated_mesh(tmp_path, periodic):
    """dgx_mesh_read_gmsh on a HEX mesh whose elements list their vertices after random cube rotations (every
    face orientation code occurs): the same elements (node sets), and the oracle's residual of a smooth state on the
    read mesh equals the one on the generated mesh at every node to 1e-12 -- a wrong orientation code or
    neighbour changes the face fluxes."""
    import pyoracle as O

    from dpgsolver_amd import _lib, mesh as M

    n, p = 3, 2
    path = str(tmp_path / "hex.msh")
    _write_hex_gmsh(path, n, periodic, seed=11)
    m = _lib.read_gmsh(path, p)
    assert m["d"] == 3 and m["E"] == n ** 3 and m["n_inverted"] == 0 and m["n_unsupported_bc"] == 0
    ref = M.structured_tp_mesh(3, p, (n, n, n), (0, 0, 0), (1, 1, 1), (periodic,) * 3,
                               bc_low=None if periodic else (M.BC_RIEMANN,) * 3,
                               bc_high=None if periodic else (M.BC_RIEMANN,) * 3)
    got = M.Mesh(3, p, m["xyz"], m["faces"].astype(np.int32))
    interior = m["faces"][:, 5] == M.BC_INTERIOR
    assert len(set(m["faces"][interior, 4])) > 4              # many orientation codes occur
    prim = M.freestream_prim(3)
    u_of = lambda mesh: M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(mesh)), 3))
    kw = {} if periodic else {"bc_prim": prim}
    r_got = O.Oracle(got, M.PDE_EULER, M.FLUX_ROE, **kw).rhs(u_of(got)).reshape(n ** 3, 5, -1)
    r_ref = O.Oracle(ref, M.PDE_EULER, M.FLUX_ROE, **kw).rhs(u_of(ref)).reshape(n ** 3, 5, -1)
    scale = np.abs(r_ref).max()
    for e in range(n ** 3):
        xg, xr = m["xyz"][e].T, ref.xyz[e].T                   # (Np, 3)
        d2 = ((xg[:, None, :] - xr[None, :, :]) ** 2).sum(-1)
        j = d2.argmin(1)
        assert d2[np.arange(len(j)), j].max() < 1e-24 and len(set(j)) == len(j), e
        assert np.abs(r_got[e] - r_ref[e][:, j]).max() <= 1e-12 * scale, e

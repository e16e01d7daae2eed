"""OpenFOAM polyMesh reader (rans_sst_py.Mesh.load_from_file on a polyMesh
directory; reference: core/src/Mesh.cpp:21-265).

Meshes are written here as ASCII polyMesh files: 2D cells (jittered quads,
and quads split into triangles) extruded one layer in z, faces in OpenFOAM's
order (internal faces by owner then neighbour, then the boundary patches).
The reader's result, written back as MSH1, must equal byte for byte what the
reference's own reader writes for the same files (compiled from its
sources); without that library, the geometry is checked against closed forms.
"""
import ctypes as C
import os

import numpy as np
import pytest

HEADER = """FoamFile
{{
    version     2.0;
    format      ascii;
    class       {cls};
    location    "constant/polyMesh";
    object      {obj};
}}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

"""


def polygons(nx, ny, jitter, split, seed):
    """2D nodes on [0, 2] x [0, 1] and ccw cells (quads, or two triangles per quad
    in a checkerboard when split)."""
    rng = np.random.default_rng(seed)
    xs, ys = np.linspace(0.0, 2.0, nx + 1), np.linspace(0.0, 1.0, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    hx, hy = 2.0 / nx, 1.0 / ny
    X[1:-1, 1:-1] += rng.uniform(-jitter, jitter, (nx - 1, ny - 1)) * hx
    Y[1:-1, 1:-1] += rng.uniform(-jitter, jitter, (nx - 1, ny - 1)) * hy
    nid = lambda i, j: i * (ny + 1) + j
    cells = []
    for i in range(nx):
        for j in range(ny):
            q = [nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)]
            if split and (i + j) % 2 == 1:
                cells += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
            else:
                cells.append(q)
    return np.stack([X.ravel(), Y.ravel()], axis=1), cells


def write_polymesh(d, xy, cells, dz=0.1):
    """Extrude the 2D cells by dz and write points/faces/owner/neighbour/boundary."""
    os.makedirs(d, exist_ok=True)
    n2 = len(xy)
    pts = np.concatenate([np.c_[xy, np.zeros(n2)], np.c_[xy, np.full(n2, dz)]])
    edges = {}
    for c, poly in enumerate(cells):
        for k in range(len(poly)):
            a, b = poly[k], poly[(k + 1) % len(poly)]
            edges.setdefault((min(a, b), max(a, b)), []).append((c, a, b))
    internal, bnd = [], {"bottom": [], "top": [], "inlet": [], "outlet": []}
    for (a0, b0), uses in edges.items():
        if len(uses) == 2:
            (c1, a1, b1), (c2, a2, b2) = uses
            own, nbr, a, b = (c1, c2, a1, b1) if c1 < c2 else (c2, c1, a2, b2)
            internal.append((own, nbr, [a, b, b + n2, a + n2]))
        else:
            c, a, b = uses[0]
            mid = 0.5 * (xy[a] + xy[b])
            key = ("bottom" if abs(mid[1]) < 1e-12 else "top" if abs(mid[1] - 1.0) < 1e-12
                   else "inlet" if abs(mid[0]) < 1e-12 else "outlet")
            bnd[key].append((c, [a, b, b + n2, a + n2]))
    internal.sort(key=lambda t: (t[0], t[1]))
    faces, owner, neigh = [], [], []
    for own, nbr, f in internal:
        faces.append(f); owner.append(own); neigh.append(nbr)
    patches = []
    for name in ("bottom", "top", "inlet", "outlet"):
        start = len(faces)
        for c, f in sorted(bnd[name]):
            faces.append(f); owner.append(c)
        patches.append((name, "wall" if name in ("bottom", "top") else "patch", start, len(faces) - start))
    start = len(faces)
    for c, poly in enumerate(cells):
        faces.append(list(reversed(poly))); owner.append(c)                 # z = 0, normal -z
        faces.append([v + n2 for v in poly]); owner.append(c)               # z = dz, normal +z
    patches.append(("frontAndBack", "empty", start, len(faces) - start))

    def put(obj, cls, n, lines, note=""):
        with open(os.path.join(d, obj), "w") as fh:
            fh.write(HEADER.format(cls=cls, obj=obj))
            fh.write("%d\n(\n" % n)
            fh.write("\n".join(lines))
            fh.write("\n)\n\n// ************************************************************************* //\n")
    put("points", "vectorField", len(pts), ["(%.17g %.17g %.17g)" % tuple(p) for p in pts])
    put("faces", "faceList", len(faces), ["%d(%s)" % (len(f), " ".join(map(str, f))) for f in faces])
    put("owner", "labelList", len(owner), [str(o) for o in owner])
    put("neighbour", "labelList", len(neigh), [str(o) for o in neigh])
    with open(os.path.join(d, "boundary"), "w") as fh:
        fh.write(HEADER.format(cls="polyBoundaryMesh", obj="boundary"))
        fh.write("%d\n(\n" % len(patches))
        for name, typ, st, nf in patches:
            fh.write("    %s\n    {\n        type            %s;\n        inGroups        List<word> 1(%s);\n"
                     "        nFaces          %d;\n        startFace       %d;\n    }\n" % (name, typ, typ, nf, st))
        fh.write(")\n")
    return len(cells), len(faces), len(internal)


CASES = {"quads": dict(nx=9, ny=6, jitter=0.2, split=False, seed=1),
         "tri_quad": dict(nx=8, ny=7, jitter=0.25, split=True, seed=2)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_reader_matches_reference(rs, ref, tmp_path, name):
    xy, cells = polygons(**CASES[name])
    poly = str(tmp_path / "case" / "constant" / "polyMesh")
    write_polymesh(poly, xy, cells)
    ours, theirs = str(tmp_path / "ours.msh"), str(tmp_path / "ref.msh")
    rs.Mesh.load_from_file(str(tmp_path / "case")).save_binary(ours)
    assert ref.ref_convert_mesh(poly.encode(), theirs.encode()) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()


@pytest.mark.parametrize("name", sorted(CASES))
def test_reader_geometry(rs, tmp_path, name):
    xy, cells = polygons(**CASES[name])
    nc, nf, nif = write_polymesh(str(tmp_path), xy, cells, dz=0.1)
    m = rs.Mesh.load_from_file(str(tmp_path))
    assert (m.n_cells(), m.n_faces(), m.n_internal_faces()) == (nc, nf, nif)
    path = str(tmp_path / "m.msh")
    m.save_binary(path)
    import rsg_abi as A
    msh = A.Msh(path)
    # extruded polygons: volume = 2D shoelace area x dz; the cells tile [0,2]x[0,1]x[0,0.1]
    for c, poly in enumerate(cells):
        p = xy[poly]
        area = 0.5 * abs(np.dot(p[:, 0], np.roll(p[:, 1], -1)) - np.dot(np.roll(p[:, 0], -1), p[:, 1]))
        assert msh.vol[c] == pytest.approx(0.1 * area, rel=1e-12)
    assert msh.vol.sum() == pytest.approx(0.2, rel=1e-12)
    # unit normals; internal normals point from owner to neighbour
    assert np.allclose(np.linalg.norm(msh.fn, axis=1), 1.0)
    d = msh.cc[msh.neighbor] - msh.cc[msh.owner[:nif]]
    assert np.all(np.sum(d * msh.fn[:nif], axis=1) > 0)
    assert msh.pnames == ["bottom", "top", "inlet", "outlet", "frontAndBack"]
    assert msh.ptypes[:2] == ["wall", "wall"] and msh.ptypes[4] == "empty"
    assert np.all((msh.weight[:nif] > 0) & (msh.weight[:nif] < 1))


def test_reader_rejects_missing_files(rs, tmp_path):
    with pytest.raises(RuntimeError):
        rs.Mesh.load_from_file(str(tmp_path))
    with pytest.raises(RuntimeError):
        rs.Mesh.load_from_file(str(tmp_path / "x.vtk"))

"""Mesh files (no GPU): the library's OpenFOAM and Gmsh readers against the
reference's own readers on the reference's fixture meshes (and two written
for these tests: Gmsh 4.1 with $Entities, tagged prisms; an OpenFOAM polyMesh
with headers and comments), bit for bit; the geometry of the read meshes
against the oracle; the VTU writer against the reference writer's arrays;
and the readers' refusals of malformed input."""
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import oracle_mesh_from_arrays, write_gmsh_tet_channel

N = pytest.importorskip("chuanflow_amd.native")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = {
    "singleHex_22": ("gmsh", "meshes/singleHex_22.msh", {}),
    "cubeHex_41": ("gmsh", "meshes/cubeHex_41.msh", {"inlet": "inlet", "outlet": "outlet"}),
    "cubeTets_22": ("gmsh", "meshes/cubeTets_22.msh", {}),
    "cubeTets_41": ("gmsh", "meshes/cubeTets_41.msh", {}),
    "prismChannel_22": ("gmsh", "meshes/prismChannel_22.msh", {"inlet": "inlet", "outlet": "outlet"}),
    "prismTagged_41": ("gmsh", "meshes/prismTagged_41.msh", {"inlet": "inlet"}),
    "openfoamBox": ("foam", "openfoamBox/polyMesh", {}),
    "jitterBox_foam": ("foam", "jitterBox_foam", {"front": "symmetry"}),
}


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLD, "mesh_readers.npz"))


def read(case):
    kind, rel, types = CASES[case]
    path = os.path.join(GOLD, rel)
    return N.HostMesh.read_gmsh(path, types) if kind == "gmsh" else N.HostMesh.read_openfoam(path, types)


@pytest.mark.parametrize("case", list(CASES))
def test_reader_matches_reference_reader(golden, case):
    hm = read(case)
    a = hm.arrays()
    for k in ("points", "faceStart", "faceVerts", "owner", "neighbour", "patchStart", "patchSize"):
        assert np.array_equal(np.asarray(a[k]).ravel(), golden[f"{case}__{k}"]), k
    assert a["nCells"] == int(golden[f"{case}__nCells"])
    spec = "".join(f"{n}={t};" for n, t in hm.patches())
    assert spec == str(golden[f"{case}__patches"])


@pytest.mark.parametrize("case", list(CASES))
def test_read_mesh_geometry_is_the_oracle_geometry_and_closed(case):
    hm = read(case)
    a, g = hm.arrays(), hm.geometry()
    om = oracle_mesh_from_arrays(a)
    for k in ("Sf", "Cf", "magSf", "C", "V", "dvec", "invDist"):
        assert np.array_equal(g[k], om.geom[k]), k
    # every cell is closed: the outward area vectors of its faces sum to zero
    own, nei = np.asarray(a["owner"]), np.asarray(a["neighbour"])
    Sf = np.asarray(g["Sf"]).reshape(-1, 3)
    close = np.zeros((a["nCells"], 3))
    np.add.at(close, own, Sf)
    np.add.at(close, nei, -Sf[:len(nei)])
    assert np.abs(close).max() < 1e-12 * np.abs(Sf).max()
    assert np.all(np.asarray(g["V"]) > 0)
    pts = np.asarray(a["points"]).reshape(-1, 3)
    box = np.prod(pts.max(axis=0) - pts.min(axis=0))
    if case in ("singleHex_22", "cubeHex_41", "cubeTets_22", "cubeTets_41", "openfoamBox", "prismTagged_41"):
        assert abs(np.sum(g["V"]) - box) < 1e-12 * box


def test_read_mesh_renumbers_and_keeps_volume():
    hm = read("cubeTets_41")
    v0 = np.sum(hm.geometry()["V"])
    bw0 = hm.bandwidth()
    hm.renumber_rcm()
    assert hm.bandwidth() <= bw0
    assert abs(np.sum(hm.geometry()["V"]) - v0) < 1e-14


def _vtu_arrays(path):
    text = open(path).read()
    out = {}
    for m in re.finditer(r'<DataArray([^>]*)>(.*?)</DataArray>', text, re.S):
        name = re.search(r'Name="([^"]*)"', m.group(1))
        out[name.group(1) if name else "Points"] = np.array(m.group(2).split(), dtype=np.float64)
    return out


def test_vtu_writer_matches_reference_writer(golden, tmp_path):
    hm = read("cubeTets_22")
    path = tmp_path / "t.vtu"
    hm.write_vtu(path, {"p": golden["vtu_in__p"], "U": golden["vtu_in__U"]})
    mine = _vtu_arrays(path)
    for k in ("Points", "connectivity", "offsets", "types", "faces", "p", "U"):
        assert np.array_equal(mine[k], golden[f"vtu__{k}"]), k
    # face offsets (VTK's index of the end of each cell's face stream) are consistent with the stream
    faces, fo = mine["faces"].astype(np.int64), mine["faceoffsets"].astype(np.int64)
    pos = 0
    for end in fo:
        nf = faces[pos]
        pos += 1
        for _ in range(nf):
            pos += 1 + faces[pos]
        assert pos == end
    assert pos == len(faces)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_openfoam_refusals(tmp_path):
    src = os.path.join(GOLD, "openfoamBox", "polyMesh")
    files = {n: open(os.path.join(src, n)).read() for n in ("points", "faces", "owner", "neighbour", "boundary")}

    def case(sub, **over):
        d = tmp_path / sub
        d.mkdir()
        for n, t in files.items():
            (d / n).write_text(over.get(n, t))
        return str(d)

    with pytest.raises(RuntimeError, match="binary"):
        N.HostMesh.read_openfoam(case("bin", points=files["points"].replace("ascii", "binary")))
    with pytest.raises(RuntimeError, match="fewer than 3 vertices|expected|truncated"):
        N.HostMesh.read_openfoam(case("short", faces=re.sub(r"\n4\(", "\n2(", files["faces"], count=1)))
    with pytest.raises(RuntimeError, match="outside"):
        N.HostMesh.read_openfoam(case("range", faces=re.sub(r"\((\d+) ", "(999 ", files["faces"], count=1)))
    with pytest.raises(RuntimeError, match="no 'points'"):
        N.HostMesh.read_openfoam(str(tmp_path / "nowhere"))
    # the case root is accepted as well as the polyMesh directory
    root = tmp_path / "case" / "constant"
    root.mkdir(parents=True)
    os.symlink(src, root / "polyMesh")
    assert N.HostMesh.read_openfoam(str(tmp_path / "case")).sizes()["nCells"] == 2


def test_gmsh_refusals(tmp_path):
    base = open(os.path.join(GOLD, "meshes", "singleHex_22.msh")).read()
    with pytest.raises(RuntimeError, match="binary"):
        N.HostMesh.read_gmsh(_write(tmp_path, "b.msh", base.replace("2.2 0 8", "2.2 1 8")))
    with pytest.raises(RuntimeError, match="pyramid"):
        pyr = base.replace("1 5 0 1 2 3 4 5 6 7 8", "1 7 0 1 2 3 4 5")
        N.HostMesh.read_gmsh(_write(tmp_path, "p.msh", pyr))
    with pytest.raises(RuntimeError, match="first-order"):
        N.HostMesh.read_gmsh(_write(tmp_path, "q.msh", base.replace("1 5 0 1 2 3 4 5 6 7 8", "1 11 0 1 2 3 4 5 6 7 8 1 2")))
    with pytest.raises(RuntimeError, match="no \\$Elements|has no"):
        N.HostMesh.read_gmsh(_write(tmp_path, "e.msh", base.split("$Elements")[0]))
    # a boundary face with no physical group, off the bounding-box planes
    tet = ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n5\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n5 1 1 1\n"
           "$EndNodes\n$Elements\n2\n1 4 0 1 2 3 4\n2 4 0 2 3 4 5\n$EndElements\n")
    with pytest.raises(RuntimeError, match="no physical group"):
        N.HostMesh.read_gmsh(_write(tmp_path, "t.msh", tet))


@pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not built")
def test_generated_tet_channel_reads_like_the_reference_reader(tmp_path):
    """A Gmsh 2.2 tet channel written by the tests (6,144 Kuhn tets, jittered,
    physical surfaces inlet/outlet/walls): the library's reader gives the
    reference reader's arrays and patches bit for bit, and closed cells of
    the box's volume."""
    import make_mesh_golden as MG
    path = str(tmp_path / "channel.msh")
    write_gmsh_tet_channel(path, 16, 8, 8, jitter=0.1, seed=3)
    hm = N.HostMesh.read_gmsh(path, {"inlet": "inlet", "outlet": "outlet"})
    lib = MG.ref_lib()
    h, ref = MG.ref_read(lib, 1, path, "inlet=inlet;outlet=outlet")
    lib.cfr_mesh_free(h)
    a = hm.arrays()
    for k in ("points", "faceStart", "faceVerts", "owner", "neighbour", "patchStart", "patchSize"):
        assert np.array_equal(np.asarray(a[k]).ravel(), ref[k]), k
    assert a["nCells"] == int(ref["nCells"]) == 16 * 8 * 8 * 6
    assert "".join(f"{n}={t};" for n, t in hm.patches()) == str(ref["patches"])
    g = hm.geometry()
    assert np.all(np.asarray(g["V"]) > 0) and abs(np.sum(g["V"]) - 0.0625) < 1e-14

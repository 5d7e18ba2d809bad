"""Mesh-file fixtures and the reference readers' outputs on them.

Run where the reference tree is present and ``make -C oracle ref`` has built
``oracle/_ref/libcfref.so``:

    python tests/golden/make_mesh_golden.py

Inputs:
  meshes/*.msh, openfoamBox/polyMesh   the mesh files the reference's own tests hold (copied as data)
  meshes/prismTagged_41.msh            written here: Gmsh 4.1 with $Entities, physical surface
                                       names, non-contiguous node tags and element tags, prisms
  jitterBox_foam/                      written here: an OpenFOAM polyMesh (FoamFile headers,
                                       comments, inGroups entries) of a jittered 4x3x2 BlockMesh
Outputs (mesh_readers.npz): per case the reference reader's points, faces (CSR), owner,
neighbour, patch ranges and "name=type;" patch string, and for one mesh the arrays of the
reference VtkWriter's file.
"""
import ctypes as C
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, REPO)
import oracle as O  # noqa: E402

# case name -> (kind 0 OpenFOAM / 1 Gmsh, path relative to tests/golden, patch types spec)
CASES = {
    "singleHex_22": (1, "meshes/singleHex_22.msh", ""),
    "cubeHex_41": (1, "meshes/cubeHex_41.msh", "inlet=inlet;outlet=outlet"),
    "cubeTets_22": (1, "meshes/cubeTets_22.msh", ""),
    "cubeTets_41": (1, "meshes/cubeTets_41.msh", ""),
    "prismChannel_22": (1, "meshes/prismChannel_22.msh", "inlet=inlet;outlet=outlet"),
    "prismTagged_41": (1, "meshes/prismTagged_41.msh", "inlet=inlet"),
    "openfoamBox": (0, "openfoamBox/polyMesh", ""),
    "jitterBox_foam": (0, "jitterBox_foam", "front=symmetry"),
}


def write_prism_41(path):
    """3x3 nodes per layer, 3 layers; each quad of the layer grid split into two triangles and
    extruded: 16 prisms.  Surfaces: bottom/top triangles, x/y sides as quads."""
    nx = ny = 3
    nz = 3
    pts = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                pts.append((0.5 * i, 0.25 * j, 0.2 * k))
    tag = lambda i, j, k: 10 + 2 * ((k * ny + j) * nx + i)  # non-contiguous node tags
    tris = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            a, b, c, d = (i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)
            tris += [(a, b, c), (a, c, d)]
    prisms = []
    for k in range(nz - 1):
        for t in tris:
            prisms.append([tag(*p, k) for p in t] + [tag(*p, k + 1) for p in t])
    bottom = [[tag(*p, 0) for p in t] for t in tris]
    top = [[tag(*p, nz - 1) for p in t] for t in tris]

    def quads(fixed_axis, value):
        out = []
        for k in range(nz - 1):
            for s in range(2):
                if fixed_axis == "x":
                    p0, p1 = (value, s), (value, s + 1)
                else:
                    p0, p1 = (s, value), (s + 1, value)
                out.append([tag(*p0, k), tag(*p1, k), tag(*p1, k + 1), tag(*p0, k + 1)])
        return out

    surfaces = [(1, "inlet", quads("x", 0)), (2, "outlet", quads("x", nx - 1)), (3, "walls", quads("y", 0)),
                (4, "walls", quads("y", ny - 1)), (5, "bottom", bottom), (6, "top", top)]
    phys = {"inlet": 11, "outlet": 12, "walls": 13, "bottom": 14, "top": 15}
    L = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(phys) + 1)]
    L += [f'2 {t} "{n}"' for n, t in phys.items()] + ['3 20 "fluid"', "$EndPhysicalNames"]
    L += ["$Entities", f"0 0 {len(surfaces)} 1"]
    for ent, name, _ in surfaces:
        L.append(f"{ent} 0 0 0 1 1 1 1 {phys[name]} 0")
    L.append("1 0 0 0 1 1 1 1 20 0")
    L += ["$EndEntities", "$Nodes", f"1 {len(pts)} {tag(0, 0, 0)} {tag(nx - 1, ny - 1, nz - 1)}",
          f"3 1 0 {len(pts)}"]
    L += [str(10 + 2 * q) for q in range(len(pts))]
    L += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in pts]
    L.append("$EndNodes")
    nel = sum(len(s) for _, _, s in surfaces) + len(prisms)
    L += ["$Elements", f"{len(surfaces) + 1} {nel} 1 {nel}"]
    et = 1
    for ent, _, elems in surfaces:
        L.append(f"2 {ent} {3 if len(elems[0]) == 4 else 2} {len(elems)}")
        for e in elems:
            L.append(" ".join(str(v) for v in [et] + e))
            et += 1
    L.append(f"3 1 6 {len(prisms)}")
    for e in prisms:
        L.append(" ".join(str(v) for v in [et] + e))
        et += 1
    L.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_foam_box(directory):
    """The jittered 4x3x2 BlockMesh as an OpenFOAM ASCII polyMesh with headers, comments and
    extra boundary entries."""
    from chuanflow_amd import native as N
    hm = N.HostMesh.box(4, 3, 2, (0, 0, 0), (1, 0.75, 0.5), 0.15, 5, False)
    a = hm.arrays()
    names = ["left", "right", "bottom", "top", "front", "back"]
    types = ["patch", "patch", "wall", "wall", "symmetryPlane", "empty"]
    os.makedirs(directory, exist_ok=True)

    def header(cls, obj):
        return ("/*--------------------------------*- C++ -*----------------------------------*\\\n"
                "  written for the mesh-reader fixtures\n"
                "\\*---------------------------------------------------------------------------*/\n"
                "FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
                f"    class       {cls};\n    location    \"constant/polyMesh\";\n    object      {obj};\n}}\n"
                "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //\n\n")

    pts = np.asarray(a["points"]).reshape(-1, 3)
    with open(os.path.join(directory, "points"), "w") as f:
        f.write(header("vectorField", "points") + f"{len(pts)}\n(\n")
        f.write("".join(f"({x:.17g} {y:.17g} {z:.17g})\n" for x, y, z in pts) + ")\n")
    fs, fv = np.asarray(a["faceStart"]), np.asarray(a["faceVerts"])
    with open(os.path.join(directory, "faces"), "w") as f:
        f.write(header("faceList", "faces") + f"{len(fs) - 1}\n(\n")
        for i in range(len(fs) - 1):
            v = fv[fs[i]:fs[i + 1]]
            f.write(f"{len(v)}(" + " ".join(str(x) for x in v) + ")\n")
        f.write(")\n")
    for name, arr in (("owner", a["owner"]), ("neighbour", a["neighbour"])):
        with open(os.path.join(directory, name), "w") as f:
            f.write(header("labelList", name) + f"{len(arr)}\n(\n" + "\n".join(str(x) for x in arr)
                    + "\n)\n/* end of list */\n")
    with open(os.path.join(directory, "boundary"), "w") as f:
        f.write(header("polyBoundaryMesh", "boundary") + f"{len(names)}\n(\n")
        for n, t, s, z in zip(names, types, a["patchStart"], a["patchSize"]):
            f.write(f"    {n}\n    {{\n        type            {t};\n        inGroups        1({t});\n"
                    f"        nFaces          {z};\n        startFace       {s};\n    }}\n")
        f.write(")\n")


def ref_read(lib, kind, path, spec):
    h = lib.cfr_read_mesh(kind, path.encode(), spec.encode())
    if not h:
        raise RuntimeError(lib.cfr_read_error().decode())
    s = np.zeros(6, np.int64)
    lib.cfr_mesh_sizes(C.c_void_p(h), s.ctypes.data_as(C.c_void_p))
    nP, nF, nI, nC, nPa, nFV = (int(x) for x in s)
    out = dict(points=np.zeros(3 * nP), faceStart=np.zeros(nF + 1, np.int64), faceVerts=np.zeros(nFV, np.int64),
               owner=np.zeros(nF, np.int64), neighbour=np.zeros(nI, np.int64), patchStart=np.zeros(nPa, np.int64),
               patchSize=np.zeros(nPa, np.int64))
    buf = C.create_string_buffer(1 << 16)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib.cfr_mesh_arrays(C.c_void_p(h), p(out["points"]), p(out["faceStart"]), p(out["faceVerts"]),
                             p(out["owner"]), p(out["neighbour"]), p(out["patchStart"]), p(out["patchSize"]), buf,
                             C.c_int64(len(buf)))
    assert rc == 0
    out["patches"] = np.array(buf.value.decode())
    out["nCells"] = np.array(nC)
    return h, out


def parse_vtu(path):
    """DataArray contents of a .vtu (name -> flat float array)."""
    text = open(path).read()
    arrays = {}
    for m in re.finditer(r'<DataArray([^>]*)>(.*?)</DataArray>', text, re.S):
        name = re.search(r'Name="([^"]*)"', m.group(1))
        arrays[name.group(1) if name else "Points"] = np.array(m.group(2).split(), dtype=np.float64)
    return arrays


def ref_lib():
    """The reference build with the mesh-reader entry points' signatures set."""
    lib = O.ref()
    lib.cfr_read_mesh.restype = C.c_void_p
    lib.cfr_read_mesh.argtypes = [C.c_int, C.c_char_p, C.c_char_p]
    lib.cfr_read_error.restype = C.c_char_p
    lib.cfr_mesh_sizes.argtypes = [C.c_void_p, C.c_void_p]
    lib.cfr_mesh_arrays.argtypes = [C.c_void_p] + [C.c_void_p] * 7 + [C.c_char_p, C.c_int64]
    lib.cfr_write_vtu.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_char_p, C.c_void_p]
    lib.cfr_mesh_free.argtypes = [C.c_void_p]
    return lib


def main():
    if not O.have_ref():
        sys.exit("reference build missing: run `make -C oracle ref` first")
    write_prism_41(os.path.join(HERE, "meshes", "prismTagged_41.msh"))
    write_foam_box(os.path.join(HERE, "jitterBox_foam"))
    lib = ref_lib()
    golden = {}
    for name, (kind, rel, spec) in CASES.items():
        h, out = ref_read(lib, kind, os.path.join(HERE, rel), spec)
        for k, v in out.items():
            golden[f"{name}__{k}"] = v
        if name == "cubeTets_22":
            n = int(out["nCells"])
            s = np.linspace(-1.0, 2.0, n) ** 3 / 7.0
            v = np.column_stack([np.sin(np.arange(n)), np.cos(np.arange(n)) / 3, np.arange(n) * 1e-3])
            tmp = os.path.join(HERE, "_vtu_tmp.vtu")
            assert lib.cfr_write_vtu(C.c_void_p(h), tmp.encode(), b"p", s.ctypes.data_as(C.c_void_p), b"U",
                                     np.ascontiguousarray(v).ctypes.data_as(C.c_void_p)) == 0
            for k, arr in parse_vtu(tmp).items():
                golden[f"vtu__{k}"] = arr
            golden["vtu_in__p"], golden["vtu_in__U"] = s, v
            os.remove(tmp)
        lib.cfr_mesh_free(C.c_void_p(h))
        print("read", name, int(out["nCells"]), "cells,", str(out["patches"]))
    np.savez_compressed(os.path.join(HERE, "mesh_readers.npz"), **golden)
    print("wrote mesh_readers.npz")


if __name__ == "__main__":
    main()

"""Wavefront OBJ import through the scene loader (ObjFileReader,
objreader.cpp:93-590; SceneJsonLoader::readObjAsSceneObjects,
sceneloader.cpp:405-478) on a host-only context, against an independent
parse of the same fixture.  CPU only.

The fixture (tests/golden/obj) covers quads and a pentagon (fan
triangulation), v/vt/vn and v//vn faces, negative (relative) indices, a face
without normals (face normal of Mesh::calcNormals), `o` and `g` objects with
`g default` ignored, usemtl resolved by a `_materials` override, by the .mtl
(Kd, Tf, map_Kd) and by the loading object's `mat`, an object with
inconsistent faces (dropped) and an empty object (dropped)."""
import os

import numpy as np

import rgapi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "obj")


def parse_obj(path):
    """Independent reader: list of (name, usemtl, triangles (n,3,3), normals
    (n,3,3) or None, ok)."""
    vs, ns = [], []
    objs = []
    cur = None

    def start(name):
        nonlocal cur
        cur = {"name": name, "mtl": "", "v": [], "n": [], "hasn": None, "hast": None, "ok": True}
        objs.append(cur)

    start("")
    for raw in open(path):
        line = raw.rstrip("\r\n")
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        if tok[0] == "v":
            vs.append([float(x) for x in tok[1:4]])
        elif tok[0] == "vn":
            ns.append([float(x) for x in tok[1:4]])
        elif tok[0] == "o":
            start(line[2:])
        elif tok[0] == "g":
            if line[2:] != "default":
                start("_".join(tok[1:]))
        elif tok[0] == "usemtl":
            cur["mtl"] = line[7:]
        elif tok[0] == "f":
            if not cur["ok"]:
                continue
            corners = [c.split("/") for c in tok[1:]]
            hast = any(len(c) > 1 and c[1] != "" for c in corners)
            hasn = any(len(c) > 2 and c[2] != "" for c in corners)
            if cur["hasn"] is not None and (hasn != cur["hasn"] or hast != cur["hast"]):
                cur["ok"] = False
                continue
            cur["hasn"], cur["hast"] = hasn, hast

            def idx(s, n):
                i = int(s)
                return i - 1 if i > 0 else n + i
            vi = [idx(c[0], len(vs)) for c in corners]
            ni = [idx(c[2], len(ns)) for c in corners] if hasn else None
            for i in range(1, len(vi) - 1):
                cur["v"].append([vs[vi[0]], vs[vi[i]], vs[vi[i + 1]]])
                if hasn:
                    cur["n"].append([ns[ni[0]], ns[ni[i]], ns[ni[i + 1]]])
    out = []
    for o in objs:
        if not o["ok"] or not o["v"]:
            continue
        v = np.array(o["v"], np.float64)
        if o["hasn"]:
            n = np.array(o["n"], np.float64)
        else:
            n = np.cross(v[:, 0] - v[:, 2], v[:, 0] - v[:, 1])
            n = np.repeat((n / np.linalg.norm(n, axis=1, keepdims=True))[:, None], 3, 1)
        out.append((o["name"], o["mtl"], v, n))
    return out


def test_obj_objects_geometry_and_materials():
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(GOLD, "obj_scene.json"))
    parts = parse_obj(os.path.join(GOLD, "box.obj"))
    assert [p[0] for p in parts] == ["cube", "penta_part", "flat", "named"]
    assert [len(p[2]) for p in parts] == [12, 3, 1, 1]
    # holder: its 4 OBJ objects, then its JSON child "after" (a plane); then
    # "second": its 4 OBJ objects
    sizes = [int((d["obj"] == k).sum()) for k in range(len(d["objects"]))]
    assert sizes == [12, 3, 1, 1, 2, 12, 3, 1, 1]
    view = np.array([0, 0, -10.0])
    for base, offset in ((0, np.array([1, 2, 3.0])), (5, np.zeros(3))):
        for k, (_, _, v, n) in enumerate(parts):
            got = d["v9"][d["obj"] == base + k].reshape(-1, 3, 3)
            np.testing.assert_allclose(got, v + offset + view, atol=1e-5)
            gn = d["n9"][d["obj"] == base + k].reshape(-1, 3, 3)
            np.testing.assert_allclose(gn, n / np.linalg.norm(n, axis=2, keepdims=True), atol=1e-6)
    objs = d["objects"]
    # .mtl: Kd -> colour, Tf -> transparency 1 - mean(Tf); blue keeps red's
    # transparency (ObjMaterial::reset does not clear it)
    np.testing.assert_allclose(objs[0]["color"], [0.8, 0.1, 0.1], rtol=1e-6)
    assert np.isclose(objs[0]["transparency"], 0.5) and np.isclose(objs[0]["roughness"], 0.5)
    np.testing.assert_allclose(objs[1]["color"], [0.1, 0.2, 0.9], rtol=1e-6)
    assert np.isclose(objs[1]["transparency"], 0.5)
    # usemtl with no library entry: the loading object's material as it
    # stood when its "mesh" key was read (holder: the default, since its
    # "mat" comes later; second: its own colour)
    np.testing.assert_allclose(objs[2]["color"], [0.8, 0.8, 0.8], rtol=1e-6)
    assert objs[2]["emission"] == 0
    np.testing.assert_allclose(objs[7]["color"], [0.2, 0.4, 0.6], rtol=1e-6)
    # _materials override by name
    np.testing.assert_allclose(objs[3]["color"], [0, 1, 0])
    assert np.isclose(objs[3]["glossy"], 0.5)
    w = "\n".join(d["warnings"])
    assert "blue.png" in w and "broken" in w
    assert "empty" not in w


def test_obj_millimetre_export_is_scaled(tmp_path):
    (tmp_path / "mm.obj").write_text("# File exported by 3ds Max\nv 0 0 0\nv 100 0 0\nv 0 100 0\nf 1 2 3\n")
    (tmp_path / "s.json").write_text('{"mainCamera": {}, "o": {"mesh": "mm.obj"}}')
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, str(tmp_path / "s.json"))
    v = d["v9"].reshape(-1, 3)
    assert np.isclose(v[:, 0].max(), 1.0) and np.isclose(v[:, 1].max(), 1.0)


def test_obj_missing_file_is_an_io_error(tmp_path):
    (tmp_path / "s.json").write_text('{"o": {"mesh": "nosuch.obj"}}')
    ctx = rgapi.Context(-1)
    try:
        rgapi.load_json(ctx, str(tmp_path / "s.json"))
    except rgapi.RGError as e:
        assert e.code == -8 and "nosuch.obj" in str(e)
    else:
        raise AssertionError("missing OBJ accepted")

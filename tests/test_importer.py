"""Scene import (XML camera/lights + OBJ/MTL, the reference's Importer.cpp
rules), EXR output, and the command line."""
import math
import struct

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import scenes
from craystal_amd.imageio import read_exr, write_exr
from craystal_amd.importer import ImportError_, Importer, parse_mtl, parse_obj, parse_scene_xml

XML = """<?xml version="1.0"?>
<camera type="perspective" width="64" height="48" fovy="39.3">
    <eye x="0" y="1" z="3.9"/>
    <lookat x="0" y="1" z="0"/>
    <up x="0" y="1" z="0"/>
</camera>
<light mtlname="Light" radiance="17.0, 12.0, 4.0"/>
"""


def write_obj(data: cr.SceneData, names, d, stem="scene", xml=XML):
    """Exports SceneData as OBJ + MTL (test helper): one `o`/`usemtl` per mesh."""
    mtl, obj = [], [f"mtllib {stem}.mtl"]
    for name, m in zip(names, data.materials):
        mtl += [f"newmtl {name}", "Kd %r %r %r" % tuple(m.diffuseRefl),
                "Ks %r %r %r" % tuple(m.specularRefl), f"Ns {m.shininess!r}"]
    base = 1
    for k, m in enumerate(data.meshes):
        obj.append(f"o mesh{k}")
        obj.append(f"usemtl {names[m.materialID]}")
        for p in m.position:
            obj.append("v %r %r %r" % tuple(float(x) for x in p))
        for n in m.normal:
            obj.append("vn %r %r %r" % tuple(float(x) for x in n))
        for t in range(0, len(m.index), 3):
            a, b, c = (int(i) + base for i in m.index[t:t + 3])
            obj.append(f"f {a}//{a} {b}//{b} {c}//{c}")
        base += len(m.position)
    (d / f"{stem}.obj").write_text("\n".join(obj) + "\n")
    (d / f"{stem}.mtl").write_text("\n".join(mtl) + "\n")
    (d / f"{stem}.xml").write_text(xml)
    return d / f"{stem}.xml"


def test_xml_camera_and_lights():
    camd, lights = parse_scene_xml(XML)
    assert camd["size"] == (64, 48) and math.isclose(camd["fov"], 39.3)
    assert camd["pos"] == (0.0, 1.0, 3.9) and camd["target"] == (0.0, 1.0, 0.0)
    assert lights == {"Light": (17.0, 12.0, 4.0)}
    with pytest.raises(ImportError_):
        parse_scene_xml("<light mtlname='a' radiance='1,1,1'/>")


def test_obj_fans_groups_flat_normals_and_emissive_ids(tmp_path):
    (tmp_path / "m.mtl").write_text(
        "newmtl White\nKd 0.7 0.7 0.7\nKs 0 0 0\nNs 10\nNi 1.5\nd 0.25\n"
        "newmtl Light\nKd 0 0 0\n")
    obj = """mtllib m.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
usemtl White
f 1 2 3 4
o second
f 1 2 5
usemtl Light
f 2 3 5
f 3 4 5
"""
    (tmp_path / "m.obj").write_text(obj)
    (tmp_path / "m.xml").write_text(XML.replace("Light", "Light"))
    data, camd = Importer().load_data(tmp_path / "m.xml")
    # meshes: quad (2 tris, fan), group "second" (1 tri), material change (2 tris)
    assert [len(m.index) // 3 for m in data.meshes] == [2, 1, 2]
    assert list(data.meshes[0].index) == [0, 1, 2, 0, 2, 3]
    # flat normal of the quad in the z = 0 plane, counter-clockwise -> +z
    assert np.allclose(data.meshes[0].normal, [0, 0, 1])
    white, light = data.materials
    assert white.type == cr.MaterialType.Phong and light.type == cr.MaterialType.Phong
    assert white.shininess == 10.0 and white.IoR == 1.5
    assert np.allclose(white.transmittance, 0.75)
    assert light.isEmissive() and not white.isEmissive()
    assert np.allclose(light.emission, (17, 12, 4))
    assert data.emissiveIndex == [3, 4]          # primitive IDs in mesh and face order


def test_texture_is_rejected(tmp_path):
    (tmp_path / "t.mtl").write_text("newmtl Tex\nKd 1 1 1\nmap_Kd wood.png\n")
    (tmp_path / "t.obj").write_text("mtllib t.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl Tex\nf 1 2 3\n")
    (tmp_path / "t.xml").write_text(XML)
    with pytest.raises(ImportError_, match="textures"):
        Importer().load_data(tmp_path / "t.xml")


def test_round_trip_of_the_cornell_box_matches_its_bvh(tmp_path):
    """The procedural Cornell box written as OBJ/MTL imports back to the same
    triangles (positions, normals, indices, emissive list), so the host BVH
    is bit-identical."""
    data, _ = scenes.cornell_box()
    names = ["White", "Red", "Green", "Light"]
    xml = write_obj(data, names, tmp_path)
    got, camd = Importer().load_data(xml)
    assert len(got.meshes) == len(data.meshes)
    for a, b in zip(got.meshes, data.meshes):
        assert np.array_equal(a.position, b.position) and np.array_equal(a.index, b.index)
        assert np.array_equal(a.normal, b.normal) and a.materialID == b.materialID
    assert got.emissiveIndex == data.emissiveIndex
    n0, p0 = cr.BVH(*data.concatenated()).export()
    n1, p1 = cr.BVH(*got.concatenated()).export()
    assert n0.tobytes() == n1.tobytes() and np.array_equal(p0, p1)


def test_exr_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(3)
    img = (rng.random((5, 7, 3)) * 10).astype(np.float32)
    write_exr(tmp_path / "a.exr", img)
    raw = (tmp_path / "a.exr").read_bytes()
    assert struct.unpack_from("<ii", raw, 0) == (20000630, 2)
    back, hdr = read_exr(tmp_path / "a.exr")
    assert [c for c, _ in hdr["channels"]] == ["B", "G", "R"]
    assert all(t == 1 for _, t in hdr["channels"])             # HALF
    assert hdr["dataWindow"] == (0, 0, 6, 4) and hdr["compression"] == 0
    assert np.array_equal(back[..., ::-1], img.astype(np.float16).astype(np.float32))
    with pytest.raises(ValueError):
        write_exr(tmp_path / "a.png", img)


@pytest.mark.gpu
def test_cli_renders_an_imported_scene(gpu, tmp_path):
    from craystal_amd import cli
    data, _ = scenes.cornell_box()
    xml = write_obj(data, ["White", "Red", "Green", "Light"], tmp_path)
    out = tmp_path / "out.exr"
    assert cli.main([str(xml), "--spp", "4", "-o", str(out)]) == 0
    img, hdr = read_exr(out)
    assert img.shape == (48, 64, 3) and np.isfinite(img).all() and img.mean() > 0.01


def _cpp_import(xml, exr):
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent / "cpp" / "_build" / "import_test"
    r = subprocess.run([str(exe), str(xml), str(exr)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def test_cpp_importer_matches_python(tmp_path):
    """include/craystal/Importer.h imports the same SceneData as the Python
    importer (meshes, indices, float positions, emissive IDs), including
    polygons and faces without normals."""
    data, _ = scenes.cornell_box()
    xml = write_obj(data, ["White", "Red", "Green", "Light"], tmp_path)
    # add a quad without normals in its own group
    with open(tmp_path / "scene.obj", "a") as f:
        n = sum(len(m.position) for m in data.meshes)
        f.write("g extra\nusemtl White\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
                f"f {n + 1} {n + 2} {n + 3} {n + 4}\n")
    got, camd = Importer().load_data(xml)
    lines = _cpp_import(xml, tmp_path / "c.exr")
    head = lines[0].split()
    assert int(head[1]) == len(got.meshes) and int(head[3]) == len(got.materials)
    assert int(head[5]) == got.triangle_count and int(head[7]) == len(got.emissiveIndex)
    assert (int(head[9]), int(head[10])) == camd["size"]
    meshes = [ln for ln in lines if ln.startswith("mesh ")]
    for ln, m in zip(meshes, got.meshes):
        t = ln.split()
        assert int(t[2]) == m.materialID and int(t[4]) == len(m.position)
        k = t.index("pos")
        assert [int(v) for v in t[6:k]] == [int(v) for v in m.index]
        assert np.array_equal(np.array(t[k + 1:], np.float32).reshape(-1, 3), m.position)
    assert [int(ln.split()[1]) for ln in lines if ln.startswith("em ")] == got.emissiveIndex


def test_cpp_exr_writer_matches_python_bytes(tmp_path):
    _cpp_import(write_obj(scenes.cornell_box()[0], ["W", "R", "G", "L"], tmp_path),
                tmp_path / "c.exr")
    w, h = 7, 5
    i = np.arange(w * h * 3, dtype=np.float32)
    img = (i * np.float32(0.37) - np.float32(3.0) + np.float32(1e-6) * (i * i)).astype(np.float32)
    img[4], img[5], img[6] = 70000.0, 1e-6, -2.5e-8
    write_exr(tmp_path / "p.exr", img.reshape(h, w, 3))
    assert (tmp_path / "c.exr").read_bytes() == (tmp_path / "p.exr").read_bytes()


@pytest.mark.gpu
def test_cpp_cli_matches_python_cli(gpu, tmp_path):
    """crayli (C++ over the C-ABI) and the Python CLI render the same EXR."""
    import subprocess
    from pathlib import Path
    from craystal_amd import cli
    data, _ = scenes.cornell_box()
    xml = write_obj(data, ["White", "Red", "Green", "Light"], tmp_path)
    exe = Path(cr.__file__).resolve().parent / "crayli"
    r = subprocess.run([str(exe), str(xml), "--spp", "3", "-o", str(tmp_path / "c.exr")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert cli.main([str(xml), "--spp", "3", "-o", str(tmp_path / "p.exr")]) == 0
    assert (tmp_path / "c.exr").read_bytes() == (tmp_path / "p.exr").read_bytes()

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


def _png_bytes(pix: np.ndarray, filters) -> bytes:
    """An 8-bit PNG of pix (H, W, C uint8, row 0 first in the file), encoded
    here from the PNG specification with the given filter type per row."""
    import zlib
    h, w, ch = pix.shape
    rows = pix.reshape(h, w * ch).astype(np.int64)
    raw = b""
    for y in range(h):
        f = filters[y % len(filters)]
        prev = rows[y - 1] if y > 0 else np.zeros(w * ch, np.int64)
        cur = rows[y]
        a = np.concatenate([np.zeros(ch, np.int64), cur[:-ch]])
        c = np.concatenate([np.zeros(ch, np.int64), prev[:-ch]])
        if f == 0:
            pred = np.zeros_like(cur)
        elif f == 1:
            pred = a
        elif f == 2:
            pred = prev
        elif f == 3:
            pred = (a + prev) // 2
        else:
            pa, pb, pc = np.abs(prev - c), np.abs(a - c), np.abs(a + prev - 2 * c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
        raw += bytes([f]) + ((cur - pred) & 0xFF).astype(np.uint8).tobytes()

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    ct = {1: 0, 3: 2, 4: 6}[ch]
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ct, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def _exr_zip_bytes(chans, h, w, comp=3) -> bytes:
    """A scanline OpenEXR with ZIP (3, 16 lines per block) or ZIPS (2)
    compression, encoded here from the specification: per block the lines'
    channel samples, bytes split into even / odd halves, then the delta
    predictor (+128), then zlib.  chans: list of (name, type, (H, W) array)."""
    import zlib
    chans = sorted(chans, key=lambda c: c[0])
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", t, 0, 1, 1)
                      for n, t, _ in chans) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)

    def attr(n, t, v):
        return n.encode() + b"\0" + t.encode() + b"\0" + struct.pack("<i", len(v)) + v
    head = struct.pack("<ii", 20000630, 2) + attr("channels", "chlist", chlist) + \
        attr("compression", "compression", bytes([comp])) + attr("dataWindow", "box2i", box) + \
        attr("displayWindow", "box2i", box) + attr("lineOrder", "lineOrder", b"\0") + b"\0"
    lines = 16 if comp == 3 else 1
    blocks = []
    dt = {0: "<u4", 1: "<f2", 2: "<f4"}
    for y0 in range(0, h, lines):
        raw = b"".join(np.asarray(a[y], dt[t]).tobytes()
                       for y in range(y0, min(h, y0 + lines)) for _, t, a in chans)
        b = np.frombuffer(raw, np.uint8)
        split = np.concatenate([b[0::2], b[1::2]]).astype(np.int64)
        delta = np.concatenate([split[:1], (split[1:] - split[:-1] + 128) & 0xFF])
        z = zlib.compress(delta.astype(np.uint8).tobytes())
        blocks.append(struct.pack("<ii", y0, len(z)) + z)
    table_at = len(head)
    offs, pos = [], table_at + 8 * len(blocks)
    for blk in blocks:
        offs.append(pos)
        pos += len(blk)
    return head + struct.pack(f"<{len(blocks)}Q", *offs) + b"".join(blocks)


def test_png_reader_known_answers(tmp_path):
    """Every PNG filter type, grey / RGB / RGBA: values /255 (and ^2.2 to
    linear), rows bottom-up as the reference's stb load flips them."""
    from craystal_amd.imageio import read_png, write_png
    rng = np.random.default_rng(1)
    for ch in (1, 3, 4):
        pix = rng.integers(0, 256, (9, 6, ch)).astype(np.uint8)
        (tmp_path / "t.png").write_bytes(_png_bytes(pix, [0, 1, 2, 3, 4]))
        got = read_png(tmp_path / "t.png", to_linear=False)
        assert np.array_equal(got, pix[::-1].astype(np.float32) / np.float32(255))
        lin = read_png(tmp_path / "t.png", to_linear=True)
        assert np.allclose(lin, (pix[::-1] / 255.0) ** 2.2, rtol=1e-6, atol=1e-7)
    img = rng.random((5, 4, 3)).astype(np.float32)
    write_png(tmp_path / "w.png", img, gamma_correct=True)
    back = read_png(tmp_path / "w.png", to_linear=False)[::-1]
    assert np.array_equal(back * 255, np.floor(np.clip(img ** (1 / 2.2), 0, 1) * 255))
    (tmp_path / "ga.png").write_bytes(_png_bytes(pix[..., :1], [0]).replace(
        b"IHDR" + struct.pack(">IIBB", 6, 9, 8, 0), b"IHDR" + struct.pack(">IIBB", 6, 9, 8, 4)))
    with pytest.raises(ValueError):
        read_png(tmp_path / "ga.png")


def test_exr_reader_zip_and_channel_mapping(tmp_path):
    """ZIP and ZIPS compressed EXR with FLOAT, HALF and UINT channels and an
    extra alpha: load_exr_image puts R, G, B in channels 0-2 and leaves the
    fourth channel zero (Image::readEXR copies R, G, B only)."""
    from craystal_amd.imageio import load_exr_image, read_exr
    rng = np.random.default_rng(2)
    h, w = 37, 11                                        # blocks of 16 with a ragged tail
    R = rng.random((h, w)).astype(np.float32) * 5
    G = rng.random((h, w)).astype(np.float16)
    B = rng.integers(0, 1000, (h, w)).astype(np.uint32)
    A = np.ones((h, w), np.float32)
    for comp in (3, 2):
        (tmp_path / "z.exr").write_bytes(_exr_zip_bytes(
            [("R", 2, R), ("G", 1, G), ("B", 0, B), ("A", 2, A)], h, w, comp))
        data, hdr = read_exr(tmp_path / "z.exr")
        assert [n for n, _ in hdr["channels"]] == ["A", "B", "G", "R"]
        img = load_exr_image(tmp_path / "z.exr")
        assert img.shape == (h, w, 4)
        assert np.array_equal(img[..., 0], R) and np.array_equal(img[..., 1], G.astype(np.float32))
        assert np.array_equal(img[..., 2], B.astype(np.float32)) and not img[..., 3].any()


def _cpp_image_dump(path):
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent / "cpp" / "_build" / "import_test"
    d = path.parent
    if not (d / "dump.xml").exists():
        write_obj(scenes.cornell_box()[0], ["W", "R", "G", "L"], d, stem="dump")
    r = subprocess.run([str(exe), str(d / "dump.xml"), str(d / "dump_out.exr"), str(path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    k = [i for i, ln in enumerate(lines) if ln.startswith("image ")][0]
    w, h, c = (int(v) for v in lines[k].split()[1:])
    vals = np.array([float.fromhex(v) for v in lines[k + 1:]], np.float32)
    return vals.reshape(c, h, w).transpose(1, 2, 0)


def test_cpp_image_readers_match_python(tmp_path):
    """Image::load (include/craystal/Image.h) reads PNG and ZIP-compressed EXR
    to the same floats as the Python readers."""
    from craystal_amd.imageio import load_image
    rng = np.random.default_rng(4)
    pix = rng.integers(0, 256, (7, 5, 3)).astype(np.uint8)
    (tmp_path / "a.png").write_bytes(_png_bytes(pix, [4, 3, 1, 2, 0]))
    assert np.array_equal(_cpp_image_dump(tmp_path / "a.png"), load_image(tmp_path / "a.png"))
    R = rng.random((20, 3)).astype(np.float32)
    (tmp_path / "b.exr").write_bytes(_exr_zip_bytes(
        [("R", 2, R), ("G", 1, R.astype(np.float16)), ("B", 2, 2 * R)], 20, 3))
    assert np.array_equal(_cpp_image_dump(tmp_path / "b.exr"), load_image(tmp_path / "b.exr"))


def test_map_kd_texture_is_loaded(tmp_path):
    """A map_Kd texture becomes a texture image of its material (flag and
    handle set, Utils/Importer.cpp:113-120), in both importers."""
    from craystal_amd.imageio import load_image
    pix = np.arange(4 * 3 * 3, dtype=np.uint8).reshape(4, 3, 3) * 7
    (tmp_path / "wood.png").write_bytes(_png_bytes(pix, [1]))
    (tmp_path / "t.mtl").write_text("newmtl Tex\nKd 1 1 1\nmap_Kd -s 1 1 1 wood.png\n"
                                    "newmtl Light\nKd 0 0 0\n")
    (tmp_path / "t.obj").write_text(
        "mtllib t.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\n"
        "usemtl Tex\nf 1/1 2/2 3/3\nusemtl Light\nf 1 3 2\n")
    (tmp_path / "t.xml").write_text(XML)
    data, _ = Importer().load_data(tmp_path / "t.xml")
    tex = data.materials[0]
    assert tex.flags & int(cr.MaterialFlags.hasDiffuseTexture) and tex.diffuseHandle == 0
    assert not data.materials[1].flags & int(cr.MaterialFlags.hasDiffuseTexture)
    assert len(data.textureImages) == 1
    assert np.array_equal(data.textureImages[0], load_image(tmp_path / "wood.png"))
    lines = _cpp_import(tmp_path / "t.xml", tmp_path / "c.exr")
    assert "texture 0 3x4x3" in lines


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

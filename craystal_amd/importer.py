"""Scene import: the reference's XML camera/light description plus a
Wavefront OBJ/MTL model (Utils/Importer.cpp:23-236), without assimp.

File format (the reference's): `scene.xml` holds top-level elements

    <camera width="W" height="H" fovy="deg">
        <eye x=".." y=".." z=".."/> <lookat .../> <up .../>
    </camera>
    <light mtlname="Name" radiance="r, g, b"/>   (any number)

and the model is `scene.obj` next to it (same stem), with its `mtllib`.

Rules kept from the reference:
* every material is Phong (Importer.cpp:145); Kd -> diffuse, Ks -> specular,
  Ns -> shininess, Ni -> IoR, opacity d -> transmittance 1 - d; a material
  named by a <light> gets that radiance as emission; then finalize();
* primitive IDs count triangles in mesh order and face order; a triangle is
  emissive iff its mesh's material is (Importer.cpp:185-197);
* polygons are triangulated as fans from their first vertex and faces
  without normals get their flat face normal (assimp's Triangulate and
  GenNormals steps); identical (position, uv, normal) corners are joined.
Meshes: a new mesh starts at every `o`/`g` statement and every material
change, in file order (assimp splits an OBJ per object/group and material).
A map_Kd diffuse texture (its last token, relative to the scene file) is
read by imageio.load_image (PNG or EXR, as Image::load) into
SceneData.textureImages, and its material gets hasDiffuseTexture and the
image's index (Importer.cpp:113-120).
"""
from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

from .core import (Camera, MaterialData, MaterialFlags, MaterialType, MeshData, Scene, SceneData,
                   Sensor)
from .imageio import load_image


class ImportError_(ValueError):
    """Malformed or unsupported scene files (the reference's CRAYSTAL_CHECK)."""


def _f3(node) -> tuple[float, float, float]:
    return (float(node.get("x", 0)), float(node.get("y", 0)), float(node.get("z", 0)))


def parse_scene_xml(text: str) -> tuple[dict, dict[str, tuple[float, float, float]]]:
    """Returns (camera dict, {material name: radiance rgb}).  The reference's
    file has several top-level elements, so the text is parsed inside a
    synthetic root."""
    body = re.sub(r"<\?xml[^>]*\?>", "", text)
    try:
        root = ET.fromstring("<root>" + body + "</root>")
    except ET.ParseError as e:
        raise ImportError_(f"Invalid XML: {e}") from None
    cam = root.find("camera")
    if cam is None:
        raise ImportError_("No camera node found.")
    camd = {"fov": float(cam.get("fovy", 0)), "size": (int(cam.get("width", 0)),
                                                       int(cam.get("height", 0)))}
    for key, tag in (("pos", "eye"), ("target", "lookat"), ("up", "up")):
        node = cam.find(tag)
        if node is not None:
            camd[key] = _f3(node)
    lights = {}
    for node in root.findall("light"):
        vals = [float(t) for t in node.get("radiance", "").split(",")[:3] if t.strip()]
        vals += [0.0] * (3 - len(vals))
        lights[node.get("mtlname", "")] = tuple(vals)
    return camd, lights


def parse_mtl(text: str) -> dict[str, dict]:
    mats: dict[str, dict] = {}
    cur = None
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        key, *rest = line.split()
        if key == "newmtl":
            cur = mats.setdefault(" ".join(rest), {})
        elif cur is None:
            continue
        elif key in ("Kd", "Ks", "Ke"):
            cur[key] = tuple(float(v) for v in rest[:3])
        elif key in ("Ns", "Ni", "d"):
            cur[key] = float(rest[0])
        elif key == "Tr":
            cur["d"] = 1.0 - float(rest[0])
        elif key == "map_Kd" and rest:
            cur["map_Kd"] = rest[-1]
    return mats


def parse_obj(text: str, mtl_loader=None):
    """Returns (meshes, materials): meshes as lists of triangles of corners
    (v, vt, vn) indices (0-based, -1 = absent) with a material name, and the
    MTL materials in first-use order."""
    V, VT, VN = [], [], []
    meshes: list[dict] = []
    mtl: dict[str, dict] = {}
    cur_mat = None
    start_new = True

    def idx(tok: str, n: int) -> int:
        if not tok:
            return -1
        i = int(tok)
        return i - 1 if i > 0 else n + i

    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        key, *rest = line.split()
        if key == "v":
            V.append([float(x) for x in rest[:3]])
        elif key == "vt":
            VT.append([float(x) for x in (rest + ["0"])[:2]])
        elif key == "vn":
            VN.append([float(x) for x in rest[:3]])
        elif key in ("o", "g"):
            start_new = True
        elif key == "usemtl":
            name = " ".join(rest)
            if name != cur_mat:
                cur_mat = name
                start_new = True
        elif key == "mtllib" and mtl_loader is not None:
            mtl.update(mtl_loader(" ".join(rest)))
        elif key == "f":
            corners = []
            for tok in rest:
                parts = (tok.split("/") + ["", ""])[:3]
                corners.append((idx(parts[0], len(V)), idx(parts[1], len(VT)),
                                idx(parts[2], len(VN))))
            if len(corners) < 3:
                continue
            if start_new or not meshes:
                meshes.append({"material": cur_mat, "tris": []})
                start_new = False
            for k in range(1, len(corners) - 1):
                meshes[-1]["tris"].append((corners[0], corners[k], corners[k + 1]))
    meshes = [m for m in meshes if m["tris"]]
    return (np.asarray(V, np.float64).reshape(-1, 3), np.asarray(VT, np.float64).reshape(-1, 2),
            np.asarray(VN, np.float64).reshape(-1, 3), meshes, mtl)


def build_scene_data(obj_text: str, lights: dict, mtl_loader=None,
                     image_loader=None) -> SceneData:
    """image_loader(name) -> (H, W, C) texture image of a map_Kd name."""
    V, VT, VN, meshes, mtl = parse_obj(obj_text, mtl_loader)
    names: list[str] = []
    for m in meshes:
        if m["material"] not in names:
            names.append(m["material"])
    materials, textures = [], []
    for name in names:
        d = mtl.get(name, {}) if name is not None else {}
        md = MaterialData(diffuseRefl=d.get("Kd", (0.0, 0.0, 0.0)),
                          specularRefl=d.get("Ks", (0.0, 0.0, 0.0)),
                          transmittance=((1.0 - d["d"],) * 3) if "d" in d else (0.0, 0.0, 0.0),
                          shininess=d.get("Ns", 0.0), IoR=d.get("Ni", 1.0),
                          type=MaterialType.Phong,
                          emission=lights.get(name, (0.0, 0.0, 0.0)) if name is not None
                          else (0.0, 0.0, 0.0))
        if "map_Kd" in d:      # Utils/Importer.cpp:113-120
            if image_loader is None:
                raise ImportError_(f"material {name!r}: no image loader for its texture")
            md.flags |= int(MaterialFlags.hasDiffuseTexture)
            md.diffuseHandle = len(textures)
            textures.append(image_loader(d["map_Kd"]))
        materials.append(md.finalize())
    out_meshes, emissive, prim = [], [], 0
    for m in meshes:
        mid = names.index(m["material"])
        corner_id: dict = {}
        P, N, T, I = [], [], [], []
        for tri in m["tris"]:
            if any(c[0] < 0 or c[0] >= len(V) for c in tri):
                raise ImportError_("face references a missing vertex")
            flat = None
            if any(c[2] < 0 for c in tri):
                a, b, c = (V[t[0]] for t in tri)
                n = np.cross(b - a, c - a)
                ln = np.linalg.norm(n)
                flat = tuple(n / ln) if ln > 0 else (0.0, 0.0, 1.0)
            for c in tri:
                if c[2] >= 0:
                    key = c
                    nrm = VN[c[2]]
                else:
                    key = (c[0], c[1], flat)
                    nrm = flat
                if key not in corner_id:
                    corner_id[key] = len(P)
                    P.append(V[c[0]])
                    N.append(nrm)
                    T.append(VT[c[1]] if 0 <= c[1] < len(VT) else (0.0, 0.0))
                I.append(corner_id[key])
        out_meshes.append(MeshData(np.asarray(P, np.float32), np.asarray(N, np.float32),
                                   np.asarray(I, np.uint32), mid, np.asarray(T, np.float32)))
        nt = len(I) // 3
        if materials[mid].isEmissive():
            emissive += list(range(prim, prim + nt))
        prim += nt
    data = SceneData(meshes=out_meshes, materials=materials, textureImages=textures)
    data.emissiveIndex = emissive
    return data


class Importer:
    """Importer::import (Utils/Importer.cpp:200-236): XML + OBJ -> Scene
    with its camera and sensor set."""

    def __init__(self):
        self.base_path = Path(".")

    def resolveResourcePath(self, rel) -> Path:
        return self.base_path / rel

    def load_data(self, path) -> tuple[SceneData, dict]:
        """The host part of the import: (SceneData, camera dict)."""
        path = Path(path)
        if path.suffix != ".xml":
            raise ImportError_("Imported file must be xml file.")
        if not path.is_file():
            raise ImportError_(f"Invalid XML file {path}.")
        self.base_path = path.parent
        camd, lights = parse_scene_xml(path.read_text())
        obj = path.with_suffix(".obj")
        if not obj.is_file():
            raise ImportError_(f"model file {obj} not found")

        def mtl_loader(rel):
            p = self.resolveResourcePath(rel)
            return parse_mtl(p.read_text()) if p.is_file() else {}

        def image_loader(rel):
            return load_image(self.resolveResourcePath(rel))

        return build_scene_data(obj.read_text(), lights, mtl_loader, image_loader), camd

    def load(self, path, device: int = 0) -> Scene:
        data, camd = self.load_data(path)
        scene = Scene(data, device)
        cam = Camera()
        cam.setFovY(camd["fov"])
        for key, setter in (("pos", cam.setWorldPosition), ("target", cam.setTarget),
                            ("up", cam.setUp)):
            if key in camd:
                setter(camd[key])
        cam.setSensor(Sensor(camd["size"], device=device))
        cam.calculateCameraData()
        scene.setCamera(cam)
        self.scene_data = data
        return scene

    import_ = load

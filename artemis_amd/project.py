"""Kobra project and scene files -> krt scene uploads.

Reads what the reference's editor writes and loads:

* ``project.den``: ``@scenes N`` followed by N scene file names
  (Project::save / Project::load, source/project.cpp).
* ``<scene>.kobra`` in the project dialect (s_load_scene, source/project.cpp):
  ``@entity <name>`` / ``@materials`` elements with ``.transform`` (position,
  rotation in degrees, scale: 9 floats), ``.mesh N`` followed by N lines
  ``<submesh file>, <material file>``, ``.renderable``, ``.list N`` and
  ``.camera <fov> <aspect>`` fields; ``#`` comments.
* ``<scene>.kobra`` in the older scene dialect (Scene::load,
  source/scene.cpp:349-388): ``[PROPERTIES]`` then ``[ENTITY]`` blocks with
  ``[TRANSFORM]`` (``position:`` / ``rotation:`` / ``scale:`` lines),
  ``[MESH]`` (``source: 0`` and raw ``submesh { v ... f ... }`` blocks),
  ``[RENDERABLE]``, ``[CAMERA]`` (``fov:`` / ``aspect:``) and ``[LIGHT]``.
* ``.cache/*.submesh``: int32 vertex count, int32 index count, the 56-byte
  vertices {position, normal, tex_coords, tangent, bitangent}
  (include/vertex.hpp:15-20), int32 indices (load_mesh, source/project.cpp).
* ``assets/*.mat``: the text of load_material (source/daemons/material.cpp:
  32-97); materials are looked up by name, the first file of a name wins.

Every submesh of an entity becomes one instance of its mesh with the
entity's model matrix, Transform::matrix() = T * mat4_cast(quat(radians(r)))
* S (include/transform.hpp), so a loaded project is the two-level scene the
reference's Accelerator builds (one GAS per submesh, one instance per
entity submesh); ``.flatten()`` gives the world-space upload.  The camera
entity gives eye = position, lookat = eye + forward(), up = up(), where
forward() / up() rotate by quat(rotation) with the rotation's degrees taken
as radians, as the reference does (transform.hpp:98-111).
"""
from __future__ import annotations

import math
import re
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .scenes import E_DIFFUSE, InstancedSceneData

VERTEX_DTYPE = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("tex_coords", "<f4", 2),
                         ("tangent", "<f4", 3), ("bitangent", "<f4", 3)])
assert VERTEX_DTYPE.itemsize == 56

# Material defaults (include/material.hpp): diffuse (1, 0, 1), specular 0,
# emission 0, roughness 0.001, refraction 1, type eDiffuse.
DEFAULT_MATERIAL = [1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.001, 1.0, E_DIFFUSE]


class ProjectError(ValueError):
    pass


# ---------------------------------------------------------------- transforms

def euler_quat(angles) -> np.ndarray:
    """glm::quat(vec3 euler) for angles in radians: (w, x, y, z) of
    Rz(z) Ry(y) Rx(x), in float32 as glm computes it."""
    a = np.asarray(angles, np.float32) * np.float32(0.5)
    c = np.cos(a).astype(np.float32)
    s = np.sin(a).astype(np.float32)
    w = c[0] * c[1] * c[2] + s[0] * s[1] * s[2]
    x = s[0] * c[1] * c[2] - c[0] * s[1] * s[2]
    y = c[0] * s[1] * c[2] + s[0] * c[1] * s[2]
    z = c[0] * c[1] * s[2] - s[0] * s[1] * c[2]
    return np.array([w, x, y, z], np.float32)


def quat_matrix(q) -> np.ndarray:
    """glm::mat3_cast of a unit quaternion (w, x, y, z), row-major 3x3."""
    w, x, y, z = (np.float32(v) for v in q)
    one, two = np.float32(1), np.float32(2)
    return np.array([
        [one - two * (y * y + z * z), two * (x * y - w * z), two * (x * z + w * y)],
        [two * (x * y + w * z), one - two * (x * x + z * z), two * (y * z - w * x)],
        [two * (x * z - w * y), two * (y * z + w * x), one - two * (x * x + y * y)],
    ], np.float32)


def transform_matrix(position, rotation_deg, scale) -> np.ndarray:
    """Transform::matrix() as a row-major 3x4: T * R(quat(radians(r))) * S."""
    R = quat_matrix(euler_quat(np.radians(np.asarray(rotation_deg, np.float64)).astype(np.float32)))
    M = np.zeros((3, 4), np.float32)
    M[:, :3] = R * np.asarray(scale, np.float32)[None, :]
    M[:, 3] = np.asarray(position, np.float32)
    return M


def transform_axes(rotation_deg):
    """forward() and up() of Transform: the rotation's numbers are passed to
    glm::quat unconverted (degrees read as radians), as the reference does."""
    R = quat_matrix(euler_quat(np.asarray(rotation_deg, np.float32)))
    fwd = R @ np.array([0.0, 0.0, -1.0], np.float32)
    up = R @ np.array([0.0, 1.0, 0.0], np.float32)
    return fwd / np.linalg.norm(fwd), up / np.linalg.norm(up)


# ------------------------------------------------------------------ elements

@dataclass
class Entity:
    name: str
    position: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0)
    scale: tuple = (1.0, 1.0, 1.0)
    # [(positions, normals, indices, material row, tex_coords, {tangent,
    # bitangent}, (albedo, normal, roughness) texture files or None)]
    submeshes: list = field(default_factory=list)
    renderable: bool = False
    camera: tuple | None = None                    # (fov, aspect)
    light: tuple | None = None                     # (color, power, type) of the scene dialect


@dataclass
class KobraScene:
    name: str
    entities: list
    environment_map: str = ""

    def instanced(self, renderable_only: bool = False) -> InstancedSceneData:
        """The two-level scene: one base mesh per distinct submesh, one
        instance per entity submesh (entity model matrix), materials
        deduplicated by value in first-use order."""
        meshes, mesh_of, inst_mesh, xforms, tri_mat, mats, mat_of = [], {}, [], [], [], [], {}
        uvs, tans, mtex, images, image_of = [], [], [], [], {}
        for e in self.entities:
            if not e.submeshes or (renderable_only and not e.renderable):
                continue
            M = transform_matrix(e.position, e.rotation, e.scale)
            for pos, nrm, idx, mrow, uv, tan, tex in e.submeshes:
                key = id(pos)
                if key not in mesh_of:
                    mesh_of[key] = len(meshes)
                    meshes.append((pos, nrm, idx))
                    uvs.append(np.zeros((len(pos), 2), np.float32) if uv is None else uv)
                    tans.append(np.zeros((len(pos), 6), np.float32) if tan is None else tan)
                slots = [-1] * 5  # diffuse, normal, specular, emission, roughness
                for slot, f in zip((0, 1, 4), tex or (None, None, None)):
                    if f is not None:
                        if f not in image_of:
                            image_of[f] = len(images)
                            images.append(f)
                        slots[slot] = image_of[f]
                mk = tuple(np.asarray(mrow, np.float32).tolist()) + tuple(slots)
                if mk not in mat_of:
                    mat_of[mk] = len(mats)
                    mats.append(mk[:12])
                    mtex.append(slots)
                inst_mesh.append(mesh_of[key])
                xforms.append(M)
                tri_mat.append(np.full(len(idx), mat_of[mk], np.uint32))
        if not inst_mesh:
            raise ProjectError(f"scene {self.name!r} has no meshes")
        d = InstancedSceneData(meshes, np.array(inst_mesh, np.int32), np.stack(xforms).astype(np.float32),
                               np.concatenate(tri_mat), np.array(mats, np.float32), camera=self.camera(),
                               name=self.name)
        if images:
            d.mesh_uvs, d.mesh_tangents = uvs, tans
            d.textures = [load_texture(f) for f in images]
            d.material_textures = np.array(mtex, np.int32)
        return d

    def camera(self) -> dict:
        """look_at arguments of the first camera entity (eye, lookat, up,
        fov); {} when the scene has none.  The aspect is camera_aspect()."""
        for e in self.entities:
            if e.camera is not None:
                fwd, up = transform_axes(e.rotation)
                eye = np.asarray(e.position, np.float32)
                return {"eye": tuple(eye.tolist()), "lookat": tuple((eye + fwd).tolist()),
                        "up": tuple(up.tolist()), "fov": float(e.camera[0])}
        return {}

    def camera_aspect(self) -> float | None:
        for e in self.entities:
            if e.camera is not None:
                return float(e.camera[1])
        return None


# --------------------------------------------------------------- file readers

def read_submesh(path, attributes: bool = False) -> tuple:
    """A ``.submesh`` cache file -> (positions, normals, indices), with
    attributes=True also tex_coords (nv, 2) and {tangent, bitangent}
    (nv, 6)."""
    data = Path(path).read_bytes()
    if len(data) < 8:
        raise ProjectError(f"{path}: truncated submesh header")
    nv, ni = struct.unpack_from("<ii", data, 0)
    need = 8 + 56 * nv + 4 * ni
    if nv < 0 or ni < 0 or ni % 3 or len(data) < need:
        raise ProjectError(f"{path}: {nv} vertices / {ni} indices do not fit {len(data)} bytes")
    v = np.frombuffer(data, VERTEX_DTYPE, nv, 8)
    idx = np.frombuffer(data, "<i4", ni, 8 + 56 * nv).astype(np.int64)
    if ni and (idx.min() < 0 or idx.max() >= nv):
        raise ProjectError(f"{path}: index out of range")
    out = (np.ascontiguousarray(v["position"]), np.ascontiguousarray(v["normal"]),
           idx.reshape(-1, 3).astype(np.uint32))
    if attributes:
        out += (np.ascontiguousarray(v["tex_coords"]),
                np.ascontiguousarray(np.concatenate([v["tangent"], v["bitangent"]], axis=1)))
    return out


def write_submesh(path, positions, normals, indices, uvs=None, tangents=None) -> None:
    """The ``.submesh`` layout that read_submesh takes (transcribe_submesh,
    source/project.cpp): tex_coords and {tangent, bitangent} as given, else
    zero."""
    positions = np.asarray(positions, np.float32)
    v = np.zeros(len(positions), VERTEX_DTYPE)
    v["position"] = positions
    v["normal"] = np.asarray(normals, np.float32)
    if uvs is not None:
        v["tex_coords"] = np.asarray(uvs, np.float32)
    if tangents is not None:
        t = np.asarray(tangents, np.float32)
        v["tangent"], v["bitangent"] = t[:, 0:3], t[:, 3:6]
    idx = np.asarray(indices, np.int32).reshape(-1)
    Path(path).write_bytes(struct.pack("<ii", len(v), len(idx)) + v.tobytes() + idx.tobytes())


_FLOAT = r"([-+0-9.eEinfaINFA]+)"


def _fields(line: str, key: str, n: int, sep: str):
    """sscanf-like: the first n numbers after `key` separated by `sep`, or
    as many as parse (the rest keep their defaults)."""
    if not line.startswith(key):
        return []
    parts = line[len(key):].strip().split(sep)
    out = []
    for p in parts[:n]:
        try:
            out.append(float(p.strip()))
        except ValueError:
            break
    return out


def read_material(path, textures: bool = False) -> tuple:
    """An ``assets/*.mat`` file -> (name, 12-float material row), with
    textures=True also the albedo / normal / roughness texture names
    (``""`` when absent: Material::has_albedo etc. treat "" and "0" as
    none)."""
    lines = Path(path).read_text().split("\n")
    lines += [""] * (10 - len(lines))
    row = list(DEFAULT_MATERIAL)
    name = lines[0][len("name: "):].split()[0] if lines[0].startswith("name: ") and lines[0][6:].split() else ""
    for k, (key, n, off) in enumerate([("diffuse: ", 3, 0), ("specular: ", 3, 3), ("emission: ", 3, 6),
                                       ("roughness: ", 1, 9), ("refraction: ", 1, 10)]):
        vals = _fields(lines[1 + k], key, n, ",")
        row[off:off + len(vals)] = vals
    t = _fields(lines[6], "type: ", 1, ",")
    if t:
        row[11] = float(int(t[0]))
    if not textures:
        return name, row
    names = []
    for k, key in enumerate(("albedo_texture: ", "normal_texture: ", "roughness_texture: ")):
        ln = lines[7 + k] if 7 + k < len(lines) else ""
        tok = ln[len(key):].split() if ln.startswith(key) else []
        names.append("" if not tok or tok[0] == "0" else tok[0])
    return name, row, tuple(names)


def load_texture(path) -> np.ndarray:
    """A texture file as (H, W, 4) uint8 in texture memory order: rows
    flipped, as TextureLoader::load_texture loads them (flip = true,
    include/backend.hpp:89, source/image.cpp:96).  PNG only."""
    from . import read_png

    return np.ascontiguousarray(read_png(path)[::-1])


def resolve_texture(name: str, directory) -> Path | None:
    """The file of a material's texture name: as given when absolute, else
    under the project directory or its assets/; None (a warning) when it is
    missing or not a PNG file, the reference's log-and-continue."""
    import warnings

    if not name:
        return None
    cands = [Path(name)] if Path(name).is_absolute() else [Path(directory) / name, Path(directory) / "assets" / name]
    for c in cands:
        if c.exists():
            if c.suffix.lower() != ".png":
                warnings.warn(f"texture {c}: only PNG textures are decoded; the material keeps its constants")
                return None
            return c.resolve()
    warnings.warn(f"texture {name!r} not found under {directory}")
    return None


def write_material(path, name: str, row, textures=("", "", "")) -> None:
    """The text of transcribe_material (source/project.cpp) for a 12-float
    row and its albedo / normal / roughness texture names."""
    r = [float(x) for x in row]
    a, n, ro = textures
    Path(path).write_text(
        f"name: {name}\ndiffuse: {r[0]:f},{r[1]:f},{r[2]:f}\nspecular: {r[3]:f},{r[4]:f},{r[5]:f}\n"
        f"emission: {r[6]:f},{r[7]:f},{r[8]:f}\nroughness: {r[9]:f}\nrefraction: {r[10]:f}\ntype: {int(r[11])}\n"
        f"albedo_texture: {a}\nnormal_texture: {n}\nroughness_texture: {ro}")


# ------------------------------------------------------ the project dialect

def parse_project_scene(text: str, directory, name: str = "") -> KobraScene:
    """s_load_scene (source/project.cpp): elements and their fields."""
    directory = Path(directory)
    lines = text.split("\n")
    elements = []
    cur = None
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if not line or line[0] == "#":
            continue
        if line[0] == "@":
            kind = line[1:].split()[0] if line[1:].split() else ""
            if kind == "entity":
                elements.append({"name": line[7:].strip(" ")})
            elif kind == "materials":
                elements.append({})
            else:
                continue  # unknown element: its fields go to the previous one, as in the reference
            cur = elements[-1]
        if line[0] == ".":
            if cur is None:
                continue
            fld = line[1:]
            if fld.startswith("transform"):
                v = [float(x) for x in fld[10:].split()[:9]]
                if len(v) != 9:
                    raise ProjectError(f"bad transform line: {line!r}")
                cur["transform"] = v
            elif fld.startswith("mesh"):
                count = int(fld[5:])
                meshes, mats = [], []
                for _ in range(count):
                    if i >= len(lines):
                        raise ProjectError("mesh list runs past the end of the file")
                    entry = lines[i].lstrip()
                    i += 1
                    mesh, _, rest = entry.partition(",")
                    meshes.append(mesh)
                    mats.append(rest.split()[0] if rest.split() else "")
                cur["mesh"], cur["material"] = meshes, mats
            elif fld.startswith("renderable"):
                cur["renderable"] = True
            elif fld.startswith("list"):
                count = int(fld[5:])
                cur["list"] = [lines[i + j].split()[0] for j in range(count)]
                i += count
            elif fld.startswith("camera"):
                fov, aspect = (float(x) for x in fld[7:].split()[:2])
                cur["camera"] = (fov, aspect)
    materials = {}  # name -> (row, textures): the daemon's name lookup, first load wins
    cache = {}
    entities = []
    for el in elements:
        if "name" not in el:
            continue
        e = Entity(el["name"])
        if "transform" in el:
            t = el["transform"]
            e.position, e.rotation, e.scale = tuple(t[0:3]), tuple(t[3:6]), tuple(t[6:9])
        for mesh, mat in zip(el.get("mesh", []), el.get("material", [])):
            mpath = directory / "assets" / mat
            row, tex = list(DEFAULT_MATERIAL), ("", "", "")
            if mpath.exists():
                mname, row, tex = read_material(mpath, textures=True)
                row, tex = materials.setdefault(mname, (row, tex))
            spath = directory / ".cache" / mesh
            if not spath.exists():
                continue  # the reference logs and skips a missing mesh file
            if spath not in cache:
                cache[spath] = read_submesh(spath, attributes=True)
            pos, nrm, idx, uv, tan = cache[spath]
            files = tuple(resolve_texture(t, directory) for t in tex)
            e.submeshes.append((pos, nrm, idx, row, uv, tan, files if any(files) else None))
        e.renderable = bool(el.get("renderable", False))
        e.camera = el.get("camera")
        entities.append(e)
    return KobraScene(name, entities)


# --------------------------------------------------------- the scene dialect

def _vec(line: str, key: str):
    if not line.startswith(key):
        raise ProjectError(f"expected {key!r}, got {line!r}")
    return tuple(float(x) for x in line[len(key):].split()[:3])


def parse_scene_file(text: str, name: str = "") -> KobraScene:
    """Scene::load (source/scene.cpp): [PROPERTIES], then [ENTITY] blocks."""
    lines = text.split("\n")
    pos = 0

    def header():
        nonlocal pos
        while pos < len(lines) and not lines[pos]:
            pos += 1
        if pos >= len(lines):
            return None
        pos += 1
        return lines[pos - 1]

    def line():
        nonlocal pos
        pos += 1
        return lines[pos - 1] if pos - 1 < len(lines) else ""

    if header() != "[PROPERTIES]":
        raise ProjectError("scene file does not start with [PROPERTIES]")
    env = line()[len("environment_map: "):]
    entities = []
    h = header()
    while h is not None:
        if h != "[ENTITY]":
            raise ProjectError(f"invalid header {h!r}")
        e = Entity(line()[len("name: "):])
        entities.append(e)
        while True:
            h = header()
            if h == "[TRANSFORM]":
                e.position = _vec(line(), "position: ")
                e.rotation = _vec(line(), "rotation: ")
                e.scale = _vec(line(), "scale: ")
            elif h == "[MESH]":
                src = line()
                m = re.match(r"source: (\S+)", src)
                if not m or m.group(1) != "0":
                    raise ProjectError("only raw meshes (source: 0) can be loaded here: external mesh files "
                                       "go through assimp in the reference")
                while True:
                    s = line()
                    if not s:
                        break
                    if s != "submesh {":
                        raise ProjectError(f"expected 'submesh {{', got {s!r}")
                    verts, faces = [], []
                    while True:
                        s = line()
                        if not s:
                            break
                        verts.append([float(x) for x in s.split()[1:4]])
                    while True:
                        s = line()
                        f = s.split()
                        if len(f) < 4 or f[0] != "f":
                            break
                        faces.append([int(x) for x in f[1:4]])
                    if s != "}":
                        raise ProjectError(f"expected '}}', got {s!r}")
                    p = np.array(verts, np.float32).reshape(-1, 3)
                    idx = np.array(faces, np.uint32).reshape(-1, 3)
                    # raw vertices carry no normals: the face normals, as the reference's
                    # Vertex(position) constructor leaves them unset
                    n = _vertex_normals(p, idx)
                    e.submeshes.append((p, n, idx, list(DEFAULT_MATERIAL), None, None, None))
            elif h == "[RENDERABLE]":
                e.renderable = True
            elif h == "[CAMERA]":
                fov = float(line()[len("fov: "):])
                aspect = float(line()[len("aspect: "):])
                e.camera = (fov, aspect)
            elif h == "[LIGHT]":
                color = _vec(line(), "color: ")
                power = float(line()[len("power: "):])
                kind = line()[len("type: "):]
                e.light = (color, power, kind)
            else:
                break
    return KobraScene(name, entities, env)


def _vertex_normals(p, idx):
    n = np.zeros_like(p)
    if len(idx):
        a, b, c = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
        fn = np.cross(b - a, c - a)
        for k in range(3):
            np.add.at(n, idx[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    return np.where(ln > 0, n / np.where(ln > 0, ln, 1), np.array([0, 1, 0], np.float32)).astype(np.float32)


# ----------------------------------------------------------------- entry points

def load_scene(path) -> KobraScene:
    """One ``.kobra`` file of either dialect (the project dialect reads its
    meshes and materials from the file's directory)."""
    path = Path(path)
    text = path.read_text()
    first = next((l for l in text.split("\n") if l.strip()), "")
    if first == "[PROPERTIES]":
        return parse_scene_file(text, path.stem)
    return parse_project_scene(text, path.parent, path.stem)


def load_project(directory) -> list:
    """Every scene of ``project.den`` (Project::load), in file order."""
    directory = Path(directory)
    lines = [l for l in (directory / "project.den").read_text().split("\n")]
    m = re.match(r"@scenes (\d+)", lines[0] if lines else "")
    if not m:
        raise ProjectError("project.den does not start with '@scenes N'")
    n = int(m.group(1))
    names = [l.strip() for l in lines[1:1 + n]]
    return [load_scene(directory / s) for s in names]


def save_project(directory, scenes: dict) -> None:
    """Writes the Project::save layout (source/project.cpp) for
    {scene name: InstancedSceneData + entity transforms}: every base mesh to
    ``.cache/submesh-<i>.submesh``, every material to ``assets/``, one
    ``.kobra`` per scene and ``project.den``.  `scenes` maps a name to
    (InstancedSceneData, [(entity name, position, rotation, scale, instance
    ids)], camera entity or None) — the transforms are the entities', so the
    instance transforms of the data are not used.  Textured data (mesh_uvs,
    mesh_tangents, textures) writes its tex_coords and tangent frames into
    the submeshes and every texture a material's albedo / normal /
    roughness slot uses as ``assets/<scene>-texture-<i>.png`` (rows flipped
    back to file order); the .mat format has no specular or emission
    texture, so those slots must be empty."""
    directory = Path(directory)
    (directory / ".cache").mkdir(parents=True, exist_ok=True)
    (directory / "assets").mkdir(parents=True, exist_ok=True)
    for sname, (data, ents, cam) in scenes.items():
        textured = getattr(data, "textures", None) is not None
        for m, (pos, nrm, idx) in enumerate(data.meshes):
            write_submesh(directory / ".cache" / f"{sname}-submesh-{m}.submesh", pos, nrm, idx,
                          data.mesh_uvs[m] if textured else None, data.mesh_tangents[m] if textured else None)
        if textured:
            from . import write_png

            for t, img in enumerate(data.textures):
                rgba = np.ascontiguousarray(np.asarray(img, np.uint8)[::-1])
                write_png(directory / "assets" / f"{sname}-texture-{t}.png",
                          rgba.view(np.uint32).reshape(rgba.shape[0], rgba.shape[1]))
        for k, row in enumerate(data.materials):
            names = ("", "", "")
            if textured:
                slots = np.asarray(data.material_textures[k])
                if slots[2] >= 0 or slots[3] >= 0:
                    raise ProjectError("the .mat format holds no specular or emission texture")
                names = tuple(f"{sname}-texture-{int(slots[j])}.png" if slots[j] >= 0 else "" for j in (0, 1, 4))
            write_material(directory / "assets" / f"{sname}-material-{k}.mat", f"{sname}-material-{k}", row, names)
        starts = np.concatenate([[0], np.cumsum([len(data.meshes[m][2]) for m in data.inst_mesh])])
        out = []
        for ename, p, r, s, insts in ents:
            out.append(f"\n@entity {ename}\n")
            out.append(".transform " + " ".join(f"{x:g}" for x in (*p, *r, *s)) + "\n")
            out.append(f".mesh {len(insts)}\n")
            for i in insts:
                mat = int(data.tri_material[starts[i]])
                out.append(f"\t{sname}-submesh-{int(data.inst_mesh[i])}.submesh, {sname}-material-{mat}.mat\n")
            out.append(".renderable\n")
        if cam is not None:
            ename, p, r, fov, aspect = cam
            out.append(f"\n@entity {ename}\n")
            out.append(".transform " + " ".join(f"{x:g}" for x in (*p, *r, 1, 1, 1)) + "\n")
            out.append(f".camera {fov:g} {aspect:g}\n")
        (directory / f"{sname}.kobra").write_text("".join(out))
    (directory / "project.den").write_text(f"@scenes {len(scenes)}\n" + "".join(f"{n}.kobra\n" for n in scenes))


def euler_from_matrix(R) -> tuple:
    """Rotation in degrees whose Transform::matrix() rotation is R (row-major
    3x3, Rz Ry Rx order) — the decomposition of Transform(const mat4 &),
    include/transform.hpp; for fixtures."""
    R = np.asarray(R, np.float64)
    x = math.atan2(R[2, 1], R[2, 2])
    y = math.atan2(-R[2, 0], math.hypot(R[2, 1], R[2, 2]))
    z = math.atan2(R[1, 0], R[0, 0])
    return tuple(math.degrees(v) for v in (x, y, z))

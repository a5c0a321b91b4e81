"""Kobra project / scene files (artemis_amd.project): the transform and
camera-axis conventions pinned against an independent rotation library and
closed forms, round trips through the reference's file layouts, the older
scene dialect, malformed inputs, and (GPU) a loaded project rendered against
the oracle."""
import math
import struct

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from artemis_amd import project as P
from artemis_amd import scenes


@pytest.mark.parametrize("rot", [(0, 0, 0), (90, 0, 0), (0, 90, 0), (0, 0, 90), (30, -45, 60), (170, 10, -100)])
def test_transform_matrix_is_t_r_s(rot):
    # glm::quat(euler) composes Rz Ry Rx: scipy's extrinsic 'xyz'
    pos, scale = (1.5, -2.0, 3.25), (2.0, 0.5, 1.0)
    M = P.transform_matrix(pos, rot, scale)
    R = Rotation.from_euler("xyz", rot, degrees=True).as_matrix()
    assert np.allclose(M[:, :3], R * np.asarray(scale)[None, :], atol=1e-6)
    assert np.array_equal(M[:, 3], np.asarray(pos, np.float32))


def test_transform_closed_forms():
    # 90 degrees about y takes +x to -z; about z takes +x to +y
    Mx = P.transform_matrix((0, 0, 0), (0, 90, 0), (1, 1, 1))
    assert np.allclose(Mx[:, :3] @ [1, 0, 0], [0, 0, -1], atol=1e-6)
    Mz = P.transform_matrix((0, 0, 0), (0, 0, 90), (1, 1, 1))
    assert np.allclose(Mz[:, :3] @ [1, 0, 0], [0, 1, 0], atol=1e-6)
    assert np.array_equal(P.transform_matrix((0, 0, 0), (0, 0, 0), (1, 1, 1)), np.eye(3, 4, dtype=np.float32))


def test_camera_axes_read_degrees_as_radians():
    fwd, up = P.transform_axes((0, 0, 0))
    assert np.allclose(fwd, [0, 0, -1]) and np.allclose(up, [0, 1, 0])
    # a rotation of pi/2 "degrees" is a quarter turn, as glm::quat(rotation) reads it
    fwd, up = P.transform_axes((0, math.pi / 2, 0))
    assert np.allclose(fwd, [-1, 0, 0], atol=1e-6) and np.allclose(up, [0, 1, 0], atol=1e-6)


def test_submesh_file_layout(tmp_path):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    nrm = np.array([[0, 0, 1]] * 4, np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    f = tmp_path / "a.submesh"
    P.write_submesh(f, pos, nrm, idx)
    raw = f.read_bytes()
    assert struct.unpack_from("<ii", raw) == (4, 6)
    assert len(raw) == 8 + 56 * 4 + 4 * 6  # 56-byte Vertex
    p, n, i = P.read_submesh(f)
    assert np.array_equal(p, pos) and np.array_equal(n, nrm) and np.array_equal(i, idx)


def test_material_file(tmp_path):
    f = tmp_path / "m.mat"
    f.write_text("name: brick\ndiffuse: 0.5,0.25,0.125\nspecular: 0,0,0\nemission: 3,2,1\nroughness: 0.75\n"
                 "refraction: 1.5\ntype: 4\nalbedo_texture: \nnormal_texture: \nroughness_texture: ")
    name, row = P.read_material(f)
    assert name == "brick"
    assert row == [0.5, 0.25, 0.125, 0, 0, 0, 3, 2, 1, 0.75, 1.5, 4.0]
    # lines that do not parse keep the defaults of include/material.hpp
    f.write_text("name: x\ndiffuse: junk\n")
    assert P.read_material(f)[1] == P.DEFAULT_MATERIAL


def _fixture(tmp_path):
    data = scenes.spheres(instances=6, slices=8, stacks=5, emissive=4, base_meshes=3, instanced=True)
    rng = np.random.default_rng(3)
    ents = []
    for i in range(len(data.inst_mesh)):
        p = tuple(np.round(rng.uniform(-4, 4, 3), 3))
        r = tuple(np.round(rng.uniform(-180, 180, 3), 2))
        s = tuple(np.round(rng.uniform(0.5, 1.5, 3), 3))
        ents.append((f"ball {i}", p, r, s, [i]))
    cam = ("camera", (0.0, 1.0, 12.0), (0.0, 0.0, 0.0), 45.0, 1.25)
    P.save_project(tmp_path, {"balls": (data, ents, cam)})
    return data, ents, cam


def test_project_round_trip(tmp_path):
    data, ents, cam = _fixture(tmp_path)
    assert (tmp_path / "project.den").read_text().startswith("@scenes 1\nballs.kobra")
    [ks] = P.load_project(tmp_path)
    inst = ks.instanced()
    assert len(inst.meshes) == len(data.meshes)
    for a, b in zip(inst.meshes, data.meshes):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y, x.dtype))
    assert np.array_equal(inst.inst_mesh, data.inst_mesh)
    for k, (_, p, r, s, _) in enumerate(ents):
        want = P.transform_matrix([float(f"{v:g}") for v in p], [float(f"{v:g}") for v in r],
                                  [float(f"{v:g}") for v in s])
        assert np.array_equal(inst.transforms[k], want)
    # materials round-trip through %f text and deduplicate by value
    written = [[float(f"{v:f}") for v in row] for row in data.materials]
    got = [list(m) for m in inst.materials]
    assert all(any(np.allclose(g, w) for w in written) for g in got)
    assert inst.triangles == data.triangles
    c = ks.camera()
    assert np.allclose(c["eye"], cam[1]) and np.allclose(c["lookat"], (0, 1, 11)) and c["fov"] == 45.0
    assert ks.camera_aspect() == 1.25
    flat = inst.flatten()
    assert flat.triangles == data.triangles and np.isfinite(flat.positions).all()


def test_project_comments_unknown_elements_and_missing_meshes(tmp_path):
    _fixture(tmp_path)
    text = (tmp_path / "balls.kobra").read_text()
    text = "# a comment\n" + text.replace("\n@entity ball 1\n", "\n@frobnicate\n@entity ball 1\n")
    text += "\n@entity ghost\n.transform 0 0 0 0 0 0 1 1 1\n.mesh 1\n\tnothere.submesh, nothere.mat\n"
    (tmp_path / "balls.kobra").write_text(text)
    ks = P.load_scene(tmp_path / "balls.kobra")
    names = [e.name for e in ks.entities]
    assert "ghost" in names and [e for e in ks.entities if e.name == "ghost"][0].submeshes == []
    assert len(ks.instanced().inst_mesh) == 6


def test_scene_dialect(tmp_path):
    f = tmp_path / "old.kobra"
    f.write_text("[PROPERTIES]\nenvironment_map: sky.hdr\n\n[ENTITY]\nname: floor\n[TRANSFORM]\n"
                 "position: 0 -1 0\nrotation: 0 0 0\nscale: 2 1 2\n[MESH]\nsource: 0\nsubmesh {\n"
                 "\tv -1 0 -1\n\tv 1 0 -1\n\tv 1 0 1\n\tv -1 0 1\n\n\tf 0 1 2\n\tf 0 2 3\n}\n\n[RENDERABLE]\n\n"
                 "[ENTITY]\nname: cam\n[TRANSFORM]\nposition: 0 1 5\nrotation: 0 0 0\nscale: 1 1 1\n"
                 "[CAMERA]\nfov: 40\naspect: 1.5\n\n[ENTITY]\nname: sun\n[LIGHT]\ncolor: 1 1 1\npower: 5\n"
                 "type: Area\n")
    ks = P.load_scene(f)
    assert ks.environment_map == "sky.hdr" and [e.name for e in ks.entities] == ["floor", "cam", "sun"]
    inst = ks.instanced()
    flat = inst.flatten()
    assert flat.triangles == 2
    assert np.allclose(flat.positions[:, 1], -1.0) and np.allclose(np.abs(flat.positions[:, 0]), 2.0)
    # (v1 - v0) x (v2 - v0) of the faces' winding
    assert np.allclose(flat.normals, [0, -1, 0])
    assert ks.camera()["fov"] == 40.0 and ks.camera_aspect() == 1.5
    assert ks.entities[2].light == ((1.0, 1.0, 1.0), 5.0, "Area")


def test_malformed_inputs(tmp_path):
    (tmp_path / "project.den").write_text("scenes 1\n")
    with pytest.raises(P.ProjectError):
        P.load_project(tmp_path)
    f = tmp_path / "bad.submesh"
    f.write_bytes(struct.pack("<ii", 5, 3) + b"\0" * 20)
    with pytest.raises(P.ProjectError):
        P.read_submesh(f)
    P.write_submesh(f, np.zeros((3, 3)), np.zeros((3, 3)), np.array([[0, 1, 7]]))
    with pytest.raises(P.ProjectError):
        P.read_submesh(f)
    g = tmp_path / "x.kobra"
    g.write_text("[PROPERTIES]\nenvironment_map: 0\n[ENTITY]\nname: a\n[MESH]\nsource: model.obj\n")
    with pytest.raises(P.ProjectError):
        P.load_scene(g)


@pytest.mark.gpu
def test_loaded_project_renders_like_oracle(tmp_path, krt_ctx):
    import oracle as O
    from artemis_amd import ArmadaRTX, PathTracer, look_at

    _fixture(tmp_path)
    [ks] = P.load_project(tmp_path)
    flat = ks.instanced().flatten()
    sc = krt_ctx.upload(flat)
    osc = O.OracleScene(flat, sc.serialized())
    W, H = 40, 32
    cam = look_at(**ks.camera(), aspect=ks.camera_aspect())
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("path_tracer", PathTracer(krt_ctx))
    arm.set_depth(3)
    print(f"[project] {flat.triangles} triangles {W}x{H}", flush=True)
    for _ in range(2):
        arm.render(sc, cam, accumulate=True)
    ref, _ = osc.render_pt(cam, W, H, frames=2, max_depth=3)
    g = arm.color_buffer()
    ex = float(np.mean(np.all(g.view(np.uint32) == ref[0].view(np.uint32), axis=-1)))
    print(f"  exact {ex:.4f}", flush=True)
    assert ex >= 0.99
    sc.free()


def _textured_fixture(tmp_path):
    """textured_room as a one-entity project (identity transform), one
    submesh per material, its albedo / normal / roughness textures as PNG
    files.  Returns the scene as the loader's flattening must give it back
    (triangles grouped by material)."""
    d = scenes.textured_room()
    mt = d.material_textures.copy()
    mt[:, 2] = mt[:, 3] = -1  # the .mat format has no specular / emission texture
    meshes, uvs, tans, tri_mat = [], [], [], []
    for m in range(len(d.materials)):
        tris = d.indices[d.tri_material == m]
        used, local = np.unique(tris, return_inverse=True)
        meshes.append((d.positions[used], d.normals[used], local.reshape(-1, 3).astype(np.uint32)))
        uvs.append(d.uvs[used])
        tans.append(d.tangents[used])
        tri_mat.append(np.full(len(tris), m, np.uint32))
    n = len(meshes)
    inst = scenes.InstancedSceneData(meshes, np.arange(n, dtype=np.int32),
                                     np.repeat(np.eye(3, 4, dtype=np.float32)[None], n, 0),
                                     np.concatenate(tri_mat), d.materials, camera=d.camera, name="room")
    inst.mesh_uvs, inst.mesh_tangents, inst.textures, inst.material_textures = uvs, tans, d.textures, mt
    P.save_project(tmp_path, {"room": (inst, [("room", (0, 0, 0), (0, 0, 0), (1, 1, 1), list(range(n)))], None)})
    return inst.flatten(), mt


def test_textured_project_round_trip(tmp_path):
    d, mt = _textured_fixture(tmp_path)
    [ks] = P.load_project(tmp_path)
    flat = ks.instanced().flatten()
    assert np.array_equal(flat.positions, d.positions) and np.array_equal(flat.uvs, d.uvs)
    assert np.array_equal(flat.tangents, d.tangents)  # identity transform, applied twice
    # the slots refer to the loader's own texture order: compare the images slot by slot
    for m in range(len(d.materials)):
        for j in (0, 1, 4):
            a, b = mt[m, j], flat.material_textures[m, j]
            assert (a < 0) == (b < 0)
            if a >= 0:
                assert np.array_equal(flat.textures[b], d.textures[a]), (m, j)
    assert np.all(flat.material_textures[:, 2:4] == -1)
    # the .mat files name PNG files whose rows run opposite to texture memory
    img = P.read_material(tmp_path / "assets" / "room-material-0.mat", textures=True)[2]
    assert img[0].endswith(".png") and img[1] == ""
    from artemis_amd import read_png
    first = read_png(tmp_path / "assets" / img[0])
    assert np.array_equal(first[::-1], d.textures[mt[0, 0]])


def test_textured_project_rotated_entity_tangents(tmp_path):
    """An entity's tangent frame goes through its model matrix twice, as
    calculate_normal applies it (amadeus_common.cuh:125-136)."""
    d = scenes.textured_room()
    inst = scenes.InstancedSceneData([(d.positions, d.normals, d.indices)], np.array([0], np.int32),
                                     P.transform_matrix((0.5, 0, 0), (0, 30, 0), (1, 1, 1))[None],
                                     d.tri_material, d.materials, camera=d.camera)
    inst.mesh_uvs, inst.mesh_tangents, inst.textures = [d.uvs], [d.tangents], d.textures
    inst.material_textures = d.material_textures
    flat = inst.flatten()
    R = Rotation.from_euler("xyz", (0, 60, 0), degrees=True).as_matrix()  # twice 30 degrees
    assert np.allclose(flat.tangents[:, 0:3], d.tangents[:, 0:3] @ R.T, atol=1e-5)
    assert np.allclose(flat.tangents[:, 3:6], d.tangents[:, 3:6] @ R.T, atol=1e-5)


def test_missing_or_unsupported_textures_warn(tmp_path):
    assert P.resolve_texture("", tmp_path) is None
    with pytest.warns(UserWarning):
        assert P.resolve_texture("nope.png", tmp_path) is None
    (tmp_path / "a.jpg").write_bytes(b"\xff\xd8")
    with pytest.warns(UserWarning):
        assert P.resolve_texture("a.jpg", tmp_path) is None


@pytest.mark.gpu
def test_textured_project_renders_like_oracle(tmp_path, krt_ctx):
    import oracle as O
    from artemis_amd import ArmadaRTX, PathTracer, look_at

    _textured_fixture(tmp_path)
    [ks] = P.load_project(tmp_path)
    flat = ks.instanced().flatten()
    sc = krt_ctx.upload(flat)
    osc = O.OracleScene(flat, sc.serialized())
    W, H = 40, 32
    cam = look_at(**scenes.textured_room().camera, aspect=W / H)  # the project has no camera entity
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("path_tracer", PathTracer(krt_ctx))
    arm.set_depth(3)
    print(f"[project] textured {flat.triangles} triangles {W}x{H}", flush=True)
    for _ in range(2):
        arm.render(sc, cam, accumulate=True)
    ref, _ = osc.render_pt(cam, W, H, frames=2, max_depth=3)
    for k in range(4):
        assert np.array_equal(arm._read(k).view(np.uint32), ref[k].view(np.uint32)), k
    sc.free()

"""Scene-JSON input: the C loader (lcsx_solver_load_scene_json) against the Python mirror and the
oracle, on the reference's own scene files and meshes (tests/golden/scenes).

CPU only: the host setup is built before the device is looked for, so its arrays are readable
on a host without a GPU (init then reports LCSX_ERR_NO_GPU).
"""
import os

import numpy as np
import pytest

import oracle
from luisacomputesimulator_amd import LcsxError, NewtonSolver, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
SCENE_DIR = os.path.join(HERE, "golden", "scenes")
SCENES = ["cloth_unit_test_square2", "cloth_unit_test_triangle1", "cloth_moving_boundary",
          "cloth_test_groud_ccd", "cloth_rotation_cylinder_7K"]
STATIC = ["rest_x", "faces", "edges", "hinges", "csr", "stretch_faces", "bending_hinges", "fixed_verts", "is_fixed",
          "face_area", "edge_area", "vert_area", "vert_mass", "offset", "dhat", "sf_Dm", "sf_area", "sf_mu",
          "sf_lambda", "bh_angle", "bh_hbar", "bh_k"]


def load_product(path):
    s = NewtonSolver()
    s.load_scene_json(path, SCENE_DIR)
    try:
        s.init_solver()
    except LcsxError as e:
        if "error 7" not in str(e):  # LCSX_ERR_NO_GPU on a CPU-only host
            raise
    return s


@pytest.mark.parametrize("name", SCENES)
def test_scene_json_setup_matches_oracle(name):
    path = os.path.join(SCENE_DIR, name + ".json")
    sc = scenes.from_json(path, SCENE_DIR)
    s = load_product(path)
    o = oracle.Oracle(sc)
    assert s.num_verts == sc.num_verts
    for key in STATIC:
        got, ref = s.read(key), o.get(key)
        assert got.shape == ref.shape, key
        assert np.array_equal(got, ref.astype(got.dtype)), key


def test_rotation_cylinder_7k_contents():
    """The reference's 7K rotating-cylinder scene: one closed cylinder, both end rings pinned and
    turned in opposite directions, gravity off, 50 PCG / 1 Newton iteration minimum."""
    path = os.path.join(SCENE_DIR, "cloth_rotation_cylinder_7K.json")
    sc = scenes.from_json(path, SCENE_DIR)
    assert len(sc.meshes) == 1
    m = sc.meshes[0]
    assert sc.params.gravity == (0.0, 0.0, 0.0) and not sc.params.use_floor
    assert sc.params.pcg_iter_count == 50 and sc.params.nonlinear_iter_count == 1
    assert [fp.method for fp in m.fixed] == ["Left", "Right"]
    assert [fp.anim.rot_deg_per_s for fp in m.fixed] == [-72.0, 72.0]
    s = load_product(path)
    fixed = s.read("fixed_verts")
    x = m.verts
    lo, hi = x[:, 0].min(), x[:, 0].max()
    # Left / Right select the vertices at the two ends of the cylinder axis (x)
    ends = np.isclose(x[fixed, 0], lo) | np.isclose(x[fixed, 0], hi)
    assert len(fixed) > 0 and ends.all()
    # each edge is shared by at most two triangles (closed tube with two boundary loops)
    e = s.read("edges").reshape(-1, 2)
    h = s.read("hinges").reshape(-1, 4)
    assert len(e) == len(h) + (len(e) - len(h))  # hinges = interior edges
    assert len(h) > 0.9 * len(e)


PER_ROW = {"rest_x": 3, "is_fixed": 1, "vert_area": 1, "vert_mass": 1, "offset": 1, "dhat": 1}


@pytest.mark.parametrize("name,bodies", [("rigid_test_solid", ["rigid_shell", "rigid_shell"]),
                                         ("cloth_rigid_coupling_drop", ["cloth", "rigid_shell"])])
def test_rigid_scene_json(name, bodies):
    """Scenes with "shell_type": "Rigid" (app_simulation_demo_config.cpp:362-375, 540-575): each
    rigid shell becomes an affine body (a rigid shell by default, RigidMaterial::is_shell) whose
    four rows (p, a0, a1, a2) follow the scene's vertices, starting at (0, I); the setup's vertex
    rows equal the oracle's."""
    path = os.path.join(SCENE_DIR, name + ".json")
    sc = scenes.from_json(path, SCENE_DIR)
    assert [m.material.body for m in sc.meshes] == bodies
    s = load_product(path)
    o = oracle.Oracle(sc)
    nb = sum(b != "cloth" for b in bodies)
    nv = sum(len(m.verts) for m in sc.meshes)
    assert s.num_verts == nv + 4 * nb
    for key in STATIC:
        got, ref = s.read(key), o.get(key)
        if key in PER_ROW:
            w = PER_ROW[key]
            assert len(got) == w * (nv + 4 * nb) and len(ref) == w * nv, key
            got = got[: w * nv]
        assert got.shape == ref.shape, key
        assert np.array_equal(got, ref.astype(got.dtype)), key
    q0 = s.read("rest_x").reshape(-1, 3)[nv:].reshape(nb, 4, 3)
    assert np.array_equal(q0, np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32), (nb, 1, 1)))


def test_rigid_shell_must_be_closed():
    """The reference rejects a rigid shell on an open mesh (init_mesh_data.cpp:404-412); so does
    the loader, before any device is needed."""
    path = os.path.join(SCENE_DIR, "rigid_test_shell.json")
    s = NewtonSolver()
    s.load_scene_json(path, SCENE_DIR)
    with pytest.raises(LcsxError, match="closed"):
        s.init_solver()


def test_rigid_material_fields_from_json(tmp_path):
    """A Rigid material's stiffness and is_shell (false: a rigid solid) reach the Python mirror."""
    import json
    js = {"shells": [{"model_name": "cube.obj", "shell_type": "Rigid", "scale": 0.1,
                      "material": {"is_shell": False, "stiffness": 1e8, "density": 500.0}}]}
    p = tmp_path / "solid.json"
    p.write_text(json.dumps(js))
    sc = scenes.from_json(str(p), SCENE_DIR)
    m = sc.meshes[0].material
    assert (m.body, m.rigid_stiffness, m.density) == ("rigid_solid", 1e8, 500.0)
    s = load_product(str(p))
    assert s.num_verts == 8 + 4
    assert np.all(s.read("offset")[:8] == 0.0)  # a rigid solid has no contact thickness


def test_solver_options_from_json(tmp_path):
    """contact_energy_type (0 quadratic, 1 log barrier) and use_energy_linesearch are read from the
    scene keys by both loaders; another contact_energy_type is rejected by both."""
    import json
    js = {"contact_energy_type": 0, "use_energy_linesearch": True,
          "shells": [{"model_name": "square2.obj", "shell_type": "Cloth"}]}
    p = tmp_path / "opts.json"
    p.write_text(json.dumps(js))
    prm = scenes.from_json(str(p), SCENE_DIR).params
    assert (prm.contact_energy_type, prm.use_energy_linesearch) == (0, True)
    load_product(str(p))
    js["contact_energy_type"] = 3
    p.write_text(json.dumps(js))
    with pytest.raises(ValueError):
        scenes.from_json(str(p), SCENE_DIR)
    with pytest.raises(LcsxError, match="contact_energy_type"):
        NewtonSolver().load_scene_json(str(p), SCENE_DIR)

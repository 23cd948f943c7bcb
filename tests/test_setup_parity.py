"""Host setup of the product (topology, masses, offsets, culled elements, CSR) against the oracle.

Runs without a GPU: lcsx_solver_init builds the host setup before it looks for a device, and
the setup buffers stay readable (it then reports LCSX_ERR_NO_GPU on a CPU-only host).
"""
import numpy as np
import pytest

import oracle
from luisacomputesimulator_amd import LcsxError, NewtonSolver, scenes

STATIC = ["rest_x", "faces", "edges", "hinges", "csr", "stretch_faces", "bending_hinges", "fixed_verts", "is_fixed",
          "face_area", "edge_area", "vert_area", "vert_mass", "offset", "dhat", "sf_Dm", "sf_area", "sf_mu",
          "sf_lambda", "bh_angle", "bh_hbar", "bh_k"]


def product_setup(sc):
    s = NewtonSolver(sc.params)
    for m in sc.meshes:
        mid = s.register_mesh(m.name, m.verts, m.tris, m.material)
        for fp in m.fixed:
            s.add_fixed_points(mid, fp.method, fp.range, fp.anim)
    try:
        s.init_solver()
    except LcsxError as e:
        if "error 7" not in str(e):  # LCSX_ERR_NO_GPU on a CPU-only host
            raise
    return s


@pytest.mark.parametrize("name", ["square", "two_sheets", "tube_small"])
def test_setup_bit_exact(name):
    if name == "square":
        sc = scenes.square_cloth(33)
    elif name == "two_sheets":
        sc = scenes.small_two_sheets(9)
    else:
        sc = scenes.cylinder_tube(rings=21, per_ring=24, cyl_segments=16, cyl_rings=6)
    o = oracle.Oracle(sc)
    s = product_setup(sc)
    for key in STATIC:
        got = s.read(key)
        ref = o.get(key if key != "rest_x" else "rest_x")
        assert got.shape == ref.shape, key
        assert np.array_equal(got.view(np.uint8) if got.dtype == np.uint8 else got, ref.astype(got.dtype)), key


def test_reference_sizes_square_cloth():
    # closed-form counts for the 33 x 33 grid: 3,136 edges, 3,008 hinges, 10,368 material blocks
    sc = scenes.square_cloth(33)
    s = product_setup(sc)
    assert len(s.read("edges")) // 2 == 3136
    assert len(s.read("hinges")) // 4 == 3008
    assert len(s.read("csr")) - (1089 + 1) == 10368
    assert s.read("fixed_verts").tolist() == [0, 32]  # LeftBack, RightBack corners
    assert s.read("vert_mass").sum() == pytest.approx(0.2, rel=1e-5)  # 0.2 kg/m^2 over 1 m^2


def test_cylinder_tube_counts():
    sc = scenes.cylinder_tube()
    assert sc.meshes[0].verts.shape[0] == 88209
    assert sc.meshes[0].tris.shape[0] == 175824
    s = product_setup(sc)
    fixed = s.read("fixed_verts")
    assert len(fixed) == 2 * 297 + 128 * 45  # both tube end rings and the whole cylinder

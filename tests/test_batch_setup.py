"""Batched form, host side: scene ranges and registration order (no GPU needed; the solver's
host setup runs before device initialisation)."""
import numpy as np

from luisacomputesimulator_amd import LcsxError, NewtonSolver, scenes


def _register(scs):
    g = NewtonSolver(scs[0].params)
    for k, sc in enumerate(scs):
        sid = g.begin_scene() if k else 0
        assert sid == k
        for m in sc.meshes:
            mid = g.register_mesh(m.name, m.verts, m.tris, m.material)
            for fp in m.fixed:
                g.add_fixed_points(mid, fp.method, fp.range, fp.anim)
    try:
        g.init_solver()
    except LcsxError:
        pass  # no device here: the host setup is still built
    return g


def test_batch_scene_ranges():
    scs = [scenes.square_cloth(9), scenes.cloth_on_sphere(17, 0.05, -0.02, sphere_subdiv=1), scenes.square_cloth(5)]
    g = _register(scs)
    assert g.num_scenes == 3
    first = 0
    for k, sc in enumerate(scs):
        a, n = g.scene_range(k)
        assert (a, n) == (first, sc.num_verts)
        first += n
    assert g.num_verts == first


def test_begin_scene_without_meshes_does_not_open_empty_scenes():
    g = NewtonSolver(scenes.Params())
    assert g.begin_scene() == 0  # nothing registered yet: still scene 0
    sc = scenes.square_cloth(5)
    g.register_mesh("a", sc.meshes[0].verts, sc.meshes[0].tris)
    assert g.begin_scene() == 1
    assert g.begin_scene() == 1  # scene 1 has no mesh yet


def test_sphere_batch_is_seeded():
    a = scenes.sphere_batch(8, seed=3)
    b = scenes.sphere_batch(8, seed=3)
    assert [s.num_verts for s in a] == [s.num_verts for s in b]
    assert all(np.array_equal(x.meshes[1].verts, y.meshes[1].verts) for x, y in zip(a, b))
    n = [int(round(np.sqrt(s.meshes[0].verts.shape[0]))) for s in a]
    assert all(17 <= k <= 161 for k in n)
    sv, st = scenes.icosphere(0.3, 2)
    assert len(sv) == 162 and len(st) == 320
    assert np.allclose(np.linalg.norm(sv, axis=1), 0.3, atol=1e-6)

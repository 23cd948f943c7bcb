"""PCG iterations replayed from captured hipGraphs against the same iterations launched directly.

Both forms run the same kernels with the same arguments in the same order, so without contacts
(whose slot order comes from atomics) every frame is bit-identical; each captured graph is replayed
once per Newton iteration, with the solver's own work and synchronisations in between, so the
frames check every replay.
"""
import numpy as np
import pytest

from luisacomputesimulator_amd import NewtonSolver, scenes

pytestmark = pytest.mark.gpu


def pair(make):
    a, b = make(), make()
    a.set_graphs(True)
    b.set_graphs(False)
    return a, b


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_graph_replay_bit_identical_without_contact(require_gpu, precision):
    def make():
        sc = scenes.square_cloth(33)
        sc.params.precision = precision
        return NewtonSolver.from_scene(sc)

    g, d = pair(make)
    for f in range(4):
        g.physics_step_gpu(1)
        d.physics_step_gpu(1)
        sg, sd = g.stats(), d.stats()
        print(f"[gpu-graphs] {precision} frame {f}: newton {sg['newton_iters']}/{sd['newton_iters']} "
              f"pcg {sg['pcg_iters_total']}/{sd['pcg_iters_total']}", flush=True)
        assert sg["pcg_iters_total"] == sd["pcg_iters_total"]
        assert np.array_equal(g.get_simulation_results(), d.get_simulation_results())


def test_graph_replay_batch_bit_identical(require_gpu):
    def make():
        return NewtonSolver.from_scenes([scenes.square_cloth(17), scenes.square_cloth(33), scenes.square_cloth(24)])

    g, d = pair(make)
    for f in range(3):
        g.physics_step_gpu(1)
        d.physics_step_gpu(1)
        for s in range(3):
            assert g.scene_stats(s)["pcg_iters_total"] == d.scene_stats(s)["pcg_iters_total"]
        assert np.array_equal(g.get_simulation_results(), d.get_simulation_results())
    print(f"[gpu-graphs] batch of 3: {g.stats()}", flush=True)


def test_graph_replay_with_contacts_tracks_direct(require_gpu):
    """Cloth on a sphere with boxes (affine bodies): contacts from the first frames; slot order
    differs between runs, so the bar is 1e-4 relative per frame."""
    def make():
        sc = scenes.cloth_rigid_coupling(n=17, boxes=2, seed=3, box_region=((-0.25, 0.25), (0.32, 0.5), (-0.25, 0.25)))
        for m in sc.meshes[1:]:
            # each box's lowest vertex 1.5 mm above the cloth (at y = 0.2)
            m.verts[:, 1] -= np.float32(m.verts[:, 1].min() - (0.2 + 1.5e-3))
        return NewtonSolver.from_scene(sc)

    g, d = pair(make)
    contacts = 0
    for f in range(6):
        g.physics_step_gpu(1)
        d.physics_step_gpu(1)
        xg, xd = g.get_simulation_results(), d.get_simulation_results()
        err = np.abs(xg - xd).max() / np.abs(xd).max()
        contacts += g.stats()["num_contacts"]
        print(f"[gpu-graphs] contact frame {f}: rel diff {err:.3e}, {g.stats()['num_contacts']} contacts", flush=True)
        assert np.isfinite(xg).all()
        assert err < 1e-4
    assert contacts > 0

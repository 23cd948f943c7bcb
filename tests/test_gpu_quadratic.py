"""The quadratic contact energy (contact_energy_type = 0, the reference's alternative to the IPC
log barrier) on the GPU against the oracle: DCD contact terms, ground terms, total energy, and
whole frames in FP32 and in the mixed form."""
import numpy as np
import pytest

import oracle
from luisacomputesimulator_amd import NewtonSolver, scenes

pytestmark = pytest.mark.gpu


def contact_key(idx, typ):
    idx = idx.reshape(-1, 4)
    k = np.concatenate([typ.reshape(-1, 1), idx], 1)
    order = np.lexsort(k.T[::-1])
    return k[order], order


def test_quadratic_dcd_and_ground_terms(require_gpu):
    """Two sheets pressed 0.4 mm into the contact band, the lower one 1.2 mm above the floor: the
    active set is the oracle's and the contact terms, gradient and diagonal agree as in the
    barrier form's parity test. Here k2 = kappa * area does not depend on the distance at all."""
    sc = scenes.small_two_sheets(9, gap=1.6e-3)
    sc.params.contact_energy_type = 0
    sc.params.use_floor = True
    sc.params.floor_y = -1.2e-3
    print(f"[gpu-quadratic] dcd scene={sc.name} verts={sc.num_verts}", flush=True)
    g, o = NewtonSolver.from_scene(sc), oracle.Oracle(sc, 32)
    for s in ("frame_begin", "predict", "iter_begin"):
        g.stage(s)
        o.stage(s)
    x = o.get("x").reshape(-1, 3).copy()
    x[81:] += np.float32(-4e-4) * np.array([0, 1, 0], dtype=np.float32)
    o.set("x", x)
    g.write("x", x)
    for s in ("material", "update_contacts", "contact_assemble"):
        g.stage(s)
        o.stage(s)
    kg, pg = contact_key(g.read("contact_idx"), g.read("contact_type"))
    ko, po = contact_key(o.get("contact_idx"), o.get("contact_type"))
    assert len(ko) > 50 and np.array_equal(kg, ko)
    for name, width in (("contact_n", 3), ("contact_k", 3), ("contact_w", 4)):
        a = g.read(name).reshape(-1, width)[pg]
        b = o.get(name).reshape(-1, width)[po]
        assert np.allclose(a, b, rtol=2e-5, atol=1e-6 * np.abs(b).max()), name
    for name in ("cgB", "diag"):
        a, b = g.read(name), o.get(name)
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(b).max()), name
    eg, eo = float(g.read("energy")[0]), o.energy()[0]
    print(f"[gpu-quadratic] energy gpu {eg:.9e} oracle {eo:.9e}", flush=True)
    assert abs(eg - eo) <= 1e-5 * abs(eo)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_quadratic_frames(require_gpu, precision):
    """Square cloth dropped from 12 mm onto the floor with the quadratic ground penalty, eight
    frames: it enters the contact band (offset 0.5 mm + d_hat 1 mm) on the third frame and is pushed
    back to its edge. FP32 tracks the FP32 oracle to 1e-4 relative; the mixed form stays within
    three times the FP32 oracle's error against the FP64 oracle (plus 1e-5)."""
    sc = scenes.square_cloth(33)
    sc.meshes[0].verts[:, 1] += np.float32(1.2e-2 - 0.5)
    sc.params.contact_energy_type = 0
    sc.params.precision = precision
    g, o, o64 = NewtonSolver.from_scene(sc), oracle.Oracle(sc, 32), oracle.Oracle(sc, 64)
    for f in range(8):
        g.physics_step_gpu(1)
        o.step()
        o64.step()
        xg, xo, x64 = g.get_simulation_results(), o.get("x").reshape(-1, 3), o64.get("x").reshape(-1, 3)
        scale = np.abs(x64).max()
        err_o, err_64, err_o32 = (np.abs(xg - xo).max() / scale, np.abs(xg - x64).max() / scale,
                                  np.abs(xo - x64).max() / scale)
        print(f"[gpu-quadratic] {precision} frame {f}: vs FP32 oracle {err_o:.3e}, vs FP64 {err_64:.3e}, "
              f"FP32 oracle vs FP64 {err_o32:.3e}, {g.stats()}", flush=True)
        assert np.isfinite(xg).all()
        if precision == "fp32":
            assert err_o < 1e-4
        else:
            assert err_64 <= 3.0 * err_o32 + 1e-5

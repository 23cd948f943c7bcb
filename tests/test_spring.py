"""The spring stretch model (ConstitutiveStretchModelCloth::Spring): a spring on every edge of the
cloth instead of the Baraff-Witkin membrane. CPU: the host setup against the oracle's, and the
oracle's spring terms pinned by a closed form (the energy 1/2 k (l - L)^2 with k = thickness * E /
(2 (1 + nu)) computed here from the material), by finite differences of the energy (gradient) and
of the gradient (Hessian times a direction, stretched springs, where the Hessian is exact).
GPU: the assembled system and frames against the oracle."""
import numpy as np
import pytest

import oracle
from luisacomputesimulator_amd import LcsxError, NewtonSolver, scenes


def spring_cloth(n=9, bending=1e-4, floor=True, gravity=(0.0, -9.8, 0.0), pinned=True):
    sc = scenes.square_cloth(n)
    if not pinned:
        sc.meshes[0].fixed = []
    sc.meshes[0].material.stretch_model = "Spring"
    sc.meshes[0].material.area_bending_stiffness = bending
    sc.params.use_floor = floor
    sc.params.gravity = gravity
    return sc


def host_setup(sc):
    s = NewtonSolver(sc.params)
    for m in sc.meshes:
        mid = s.register_mesh(m.name, m.verts, m.tris, m.material)
        for fp in m.fixed:
            s.add_fixed_points(mid, fp.method, fp.range, fp.anim)
    try:
        s.init_solver()
    except LcsxError as e:
        if "error 7" not in str(e):  # no GPU here: the host setup is still readable
            raise
    return s


def test_spring_setup_matches_oracle():
    """Springs on every edge with a free end, in edge order; no stretch faces; the CSR holds the
    spring neighbours; rest lengths and stiffness equal the oracle's bit for bit."""
    sc = spring_cloth()
    s, o = host_setup(sc), oracle.Oracle(sc, 32)
    sp = s.read("stretch_springs").reshape(-1, 2)
    edges, fixed = o.get("edges").reshape(-1, 2), o.get("is_fixed")
    free_edges = edges[(fixed[edges[:, 0]] == 0) | (fixed[edges[:, 1]] == 0)]
    assert np.array_equal(sp, free_edges.astype(sp.dtype))
    assert np.array_equal(sp, o.get("stretch_springs").reshape(-1, 2).astype(sp.dtype))
    assert len(s.read("stretch_faces")) == 0 and len(o.get("stretch_faces")) == 0
    for name in ("sp_rest", "sp_k", "csr"):
        assert np.array_equal(s.read(name), o.get(name).astype(s.read(name).dtype)), name
    m = sc.meshes[0].material
    k = np.float32(np.float64(m.youngs_modulus) / (2.0 * (1.0 + np.float64(np.float32(m.poisson_ratio))))
                   * np.float64(np.float32(m.thickness)))
    assert np.all(s.read("sp_k") == k)


def staged_oracle(sc, stretch=1.1, seed=0):
    o = oracle.Oracle(sc, 64)
    for st in ("frame_begin", "predict", "iter_begin"):
        o.stage(st)
    rest = o.get("rest_x").reshape(-1, 3)
    rng = np.random.default_rng(seed)
    x = rest * stretch + 1e-3 * rng.standard_normal(rest.shape)
    x[o.get("is_fixed") != 0] = rest[o.get("is_fixed") != 0]
    o.set("x", x)
    o.stage("material")
    return o, x


def test_spring_energy_closed_form():
    """No gravity, no floor, no bending, at rest velocity: E(x) minus the inertia term, computed
    here from the vertex masses, is sum 1/2 k (|x_a - x_b| - L)^2 over the springs."""
    sc = spring_cloth(bending=0.0, floor=False, gravity=(0.0, 0.0, 0.0))
    o, x = staged_oracle(sc)
    rest = o.get("rest_x").reshape(-1, 3).astype(np.float64)
    m = o.get("vert_mass").astype(np.float64)
    h = sc.params.implicit_dt
    fixed = o.get("is_fixed") != 0
    w = np.where(fixed, sc.params.stiffness_dirichlet, 1.0)
    inertia = np.sum(w * 0.5 * m / h ** 2 * np.sum((x - rest) ** 2, 1))
    mat = sc.meshes[0].material
    k = mat.thickness * mat.youngs_modulus / (2.0 * (1.0 + mat.poisson_ratio))
    sp = o.get("stretch_springs").reshape(-1, 2)
    L = np.linalg.norm(rest[sp[:, 0]] - rest[sp[:, 1]], axis=1)
    l = np.linalg.norm(x[sp[:, 0]] - x[sp[:, 1]], axis=1)
    expect = np.sum(0.5 * k * (l - L) ** 2)
    got = o.energy()[0] - inertia
    assert expect > 0
    assert abs(got - expect) <= 1e-5 * expect


def test_spring_gradient_by_finite_differences():
    sc = spring_cloth(floor=False)
    o, x = staged_oracle(sc, stretch=1.05, seed=1)
    g = -o.get("cgB").reshape(-1, 3)
    free = np.nonzero(o.get("is_fixed") == 0)[0]
    scale = np.abs(g[free]).max()
    for i in free[:: max(1, len(free) // 10)]:
        for d in range(3):
            xp = x.copy(); xp[i, d] += 1e-6; o.set("x", xp); ep = o.energy()[0]
            xm = x.copy(); xm[i, d] -= 1e-6; o.set("x", xm); em = o.energy()[0]
            assert abs((ep - em) / 2e-6 - g[i, d]) <= 2e-4 * scale, (i, d)


def test_spring_hessian_by_finite_differences():
    """Stretched springs (l > L everywhere: no pins, every position scaled by 1.2) and no bending:
    the assembled operator (diagonal and off-diagonal blocks) times v equals the central
    difference of the gradient along v. (With pins, springs next to a pin can be compressed,
    where the Hessian's lateral term is clamped at 0 and differs from the exact one.)"""
    sc = spring_cloth(bending=0.0, floor=False, pinned=False)
    o, x = staged_oracle(sc, stretch=1.2, seed=2)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(x.shape)
    v[o.get("is_fixed") != 0] = 0.0
    o.set("dx", v)
    o.stage("spmv_dx")
    Av = o.get("dx").reshape(-1, 3).copy()
    eps = 1e-6
    gs = []
    for sgn in (1, -1):
        o.set("x", x + sgn * eps * v)
        o.stage("material")
        gs.append(-o.get("cgB").reshape(-1, 3))
    fd = (gs[0] - gs[1]) / (2 * eps)
    free = o.get("is_fixed") == 0
    assert np.allclose(Av[free], fd[free], rtol=1e-4, atol=1e-6 * np.abs(fd[free]).max())


def test_stretch_model_from_json(tmp_path):
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    scene_dir = os.path.join(here, "golden", "scenes")
    for sm, spring in (("Spring", True), (0, True), ("FEM_BW98", False), (1, False)):
        js = {"shells": [{"model_name": "square2.obj", "shell_type": "Cloth", "material": {"stretch_model": sm}}]}
        p = tmp_path / "s.json"
        p.write_text(json.dumps(js))
        assert (scenes.from_json(str(p), scene_dir).meshes[0].material.stretch_model == "Spring") == spring
        s = NewtonSolver()
        s.load_scene_json(str(p), scene_dir)
        try:
            s.init_solver()
        except LcsxError as e:
            if "error 7" not in str(e):
                raise
        assert (len(s.read("stretch_springs")) > 0) == spring
        assert (len(s.read("stretch_faces")) > 0) == (not spring)


@pytest.mark.gpu
def test_spring_assembly_on_gpu(require_gpu):
    """The GPU's gradient and diagonal and off-diagonal blocks equal the FP32 oracle's (blocks
    bit-exact: only +, -, *, /, sqrt; gradient within the bending angle's few ulps)."""
    sc = spring_cloth(n=17)
    print(f"[gpu-spring] assembly verts={sc.num_verts}", flush=True)
    g, o = NewtonSolver.from_scene(sc), oracle.Oracle(sc, 32)
    for st in ("frame_begin", "predict", "iter_begin"):
        g.stage(st)
        o.stage(st)
    rest = o.get("rest_x").reshape(-1, 3)
    x = (rest * np.float32(1.05)).astype(np.float32)
    x[o.get("is_fixed") != 0] = rest[o.get("is_fixed") != 0]
    g.write("x", x)
    o.set("x", x)
    g.stage("material")
    o.stage("material")
    assert np.array_equal(g.read("diag"), o.get("diag"))
    assert np.array_equal(g.read("offdiag"), o.get("offdiag"))
    a, b = g.read("cgB"), o.get("cgB")
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(b).max())
    eg, eo = float(g.read("energy")[0]), o.energy()[0]
    assert abs(eg - eo) <= 1e-5 * abs(eo)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "mixed", "mixed_bf16"])
def test_spring_frames_on_gpu(require_gpu, precision):
    """Spring cloth falling under gravity with two pinned corners, six frames. FP32 tracks the
    FP32 oracle to 1e-4 relative (the bar of the membrane model's frames); every form is within
    three times the FP32 oracle's error against the FP64 oracle, plus 1e-5 (mixed) or the Newton
    step tolerance per frame (BF16, whose blocks carry 8 bits, as in the membrane model's test)."""
    sc = spring_cloth(n=33)
    sc.params.precision = precision
    g, o, o64 = NewtonSolver.from_scene(sc), oracle.Oracle(sc, 32), oracle.Oracle(sc, 64)
    for f in range(6):
        g.physics_step_gpu(1)
        o.step()
        o64.step()
        xg, xo, x64 = g.get_simulation_results(), o.get("x").reshape(-1, 3), o64.get("x").reshape(-1, 3)
        scale = np.abs(x64).max()
        err, err_g, err_o = (np.abs(xg - xo).max() / np.abs(xo).max(), np.abs(xg - x64).max() / scale,
                             np.abs(xo - x64).max() / scale)
        print(f"[gpu-spring] {precision} frame {f}: vs FP32 oracle {err:.3e}, vs FP64 {err_g:.3e}, "
              f"FP32 oracle vs FP64 {err_o:.3e}, {g.stats()}", flush=True)
        if precision == "fp32":
            assert err < 1e-4
        step_tol = (f + 1) * sc.params.newton_tol * sc.params.implicit_dt / scale
        assert err_g <= 3.0 * err_o + (step_tol if precision == "mixed_bf16" else 1e-5)

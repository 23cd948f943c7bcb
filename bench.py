#!/usr/bin/env python3
"""Benchmark: Newton steps/s of the cloth Newton step on one MI355X.

Workload: the 297 x 297-vertex cloth tube (88,209 vertices, 175,824 triangles, r = 0.11 m,
length 1 m) around a kinematic cylinder (r = 0.1 m, 5,760 vertices) spinning at 2*pi rad/s,
tube ends pinned to the spin, both started in the steady spin; dt = 1/60 s, gravity, self
contact with d_hat = 1e-3 m and kappa = 1e9, FP32. A bench "step" is one simulated frame (each
frame runs >= 3 Newton iterations, each with <= 100 PCG iterations); `value` is Newton
iterations per second over the timed frames, with every input resident on the GPU.

The JSON line carries:
  value         Newton steps/s in FP32 (the reference's precision).
  mixed         the same frames in mixed precision (FP16 off-diagonal blocks with a power-of-two
                scale per block inside the PCG, FP32 vectors with residual replacement every 25
                iterations, FP64 CCD and barrier terms): Newton steps/s and its SpMV figures.
  roofline      the PCG block SpMV k_spmv_pcg of the FP32 run (material blocks and contact
                slots): algorithmic bytes per launch = 40 B per off-diagonal block (36 B values +
                4 B column) + 84 B per row (36 B diagonal block, z and p_old of the row 24 B,
                p_new and q writes 24 B) + 4 B per 64-row slice + 24 B per contact slot,
                divided by the launch's average duration from hipEvents attached to the
                launches (hipExtLaunchKernelGGL) on the solver stream; peak = 8 TB/s HBM3E. The
                FP16 launch moves 23 B per off-diagonal block (18 B values, 1 B exponent, 4 B
                column) and the same per row.
  batched       the batched form on 256 cloth-on-sphere scenes (grid side n ~ U{17..161},
                sphere r = 0.3 m offset U(-0.1, 0.1) m in x and z, seed 0; 2.81M vertices): all
                scenes in one solver (one launch per stage) against the same scenes stepped one by
                one, each from the batch's state after the warm-up frames (the cloths reach the
                spheres at about frame 12). Scene-steps/s = scenes x frames / seconds.
  coupling      cloth-rigid coupling (129 x 129 cloth, 24 affine-body boxes): Newton steps/s in FP32
                and mixed precision over 60 frames, and both trajectories' relative position error
                against the FP64 oracle over the first 24 frames (the first box-cloth contacts
                come at about frame 20).
  cpu_baseline  the CPU oracle (C++ restatement of the reference path, 1 thread) timed on
                frame 0 of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newton steps/s @88K-vert cloth, mixed vs FP32; SpMV HBM GB/s; scene-steps/s"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(n_rows: int, n_blocks: int, slots: float = 0.0, block_bytes: int = 40) -> float:
    """Algorithmic bytes of one k_spmv_pcg launch: block_bytes per off-diagonal block (40 in FP32:
    36 B values + 4 B column; 23 with FP16 blocks: 18 B values + 1 B exponent + 4 B column) +
    84 B per row + 4 B per 64-row slice offset, and 24 B per contact slot (16 B coefficient, 4 B
    row, the contact's 4 B t_c)."""
    return block_bytes * n_blocks + 84 * n_rows + 4 * ((n_rows + 63) // 64) + 24 * slots


def traffic_from_profiles():
    """HBM bytes per SpMV launch and the algorithmic bytes of that same launch, from the committed
    rocprofv3 counter summary (lines `hbm_bytes_per_launch:` / `algorithmic_bytes_per_launch:` of
    profiles/k_spmv_pcg_tube88k_counters.md), or (None, None)."""
    path = os.path.join(ROOT, "profiles", "k_spmv_pcg_tube88k_counters.md")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                key, _, val = line.strip().partition(":")
                if key in ("hbm_bytes_per_launch", "algorithmic_bytes_per_launch") and val.strip():
                    out[key] = float(val.strip())
    return out.get("hbm_bytes_per_launch"), out.get("algorithmic_bytes_per_launch")


def profile_pcg(g, copy_mib: int = 256) -> None:
    """Counter-run workload (rocprofv3 --pmc): one device-to-device copy of `copy_mib` MiB (the
    calibration kernel of known size) followed by one PCG solve (up to 100 SpMV launches) on the
    system of the last Newton iteration."""
    import numpy as np

    from luisacomputesimulator_amd import hiprt

    src = hiprt.DeviceArray(np.zeros(copy_mib << 18, dtype=np.float32))
    dst = hiprt.DeviceArray(np.zeros(copy_mib << 18, dtype=np.float32))
    hiprt.copy_device(dst, src)
    hiprt.synchronize()
    g.stage("pcg")
    hiprt.synchronize()
    st = g.stats()
    print(f"[bench] profile-pcg: copied {copy_mib} MiB, PCG {st['pcg_iters']} iterations", file=sys.stderr, flush=True)


def batched_leg(args, precision, sequential):
    """Batched vs one-by-one scene-steps/s on the cloth-on-sphere batch."""
    from luisacomputesimulator_amd import NewtonSolver, hiprt, scenes

    scs = scenes.sphere_batch(args.batch_scenes, seed=0)
    for sc in scs:
        sc.params.precision = precision
    t = time.perf_counter()
    b = NewtonSolver.from_scenes(scs)
    print(f"[bench] batch {precision}: {len(scs)} scenes, {b.num_verts} vertices, setup {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    for f in range(args.batch_warmup):
        tw = time.perf_counter()
        b.physics_step_gpu(1)
        st = b.stats()
        print(f"[bench] batch {precision} warmup frame {f}: {st['newton_iters']} scene-Newton, {st['num_contacts']} contacts, "
              f"{time.perf_counter() - tw:.2f} s", file=sys.stderr, flush=True)
    x0, v0 = b.get_simulation_results(), b.get_velocities()
    ranges = [b.scene_range(g) for g in range(len(scs))]
    hiprt.synchronize()
    newton = 0
    t0 = time.perf_counter()
    for f in range(args.batch_steps):
        b.physics_step_gpu(1)
        newton += b.stats()["newton_iters"]
    hiprt.synchronize()
    el = time.perf_counter() - t0
    out = {"scene_steps_per_s": len(scs) * args.batch_steps / el, "ms_per_frame": 1e3 * el / args.batch_steps,
           "scene_newton_iters": newton, "contacts_last_iter": b.stats()["num_contacts"]}
    print(f"[bench] batch {precision} timed: {out}", file=sys.stderr, flush=True)
    del b
    if sequential:
        total, newton_s = 0.0, 0
        for g, sc in enumerate(scs):
            a, n = ranges[g]
            one = NewtonSolver.from_scene(sc)
            one.set_state(x0[a:a + n], v0[a:a + n])
            one.physics_step_gpu(1)  # warm-up frame (buffer sizes), then back to the batch's state
            one.set_state(x0[a:a + n], v0[a:a + n])
            hiprt.synchronize()
            t0 = time.perf_counter()
            for f in range(args.batch_steps):
                one.physics_step_gpu(1)
                newton_s += one.stats()["newton_iters"]
            hiprt.synchronize()
            total += time.perf_counter() - t0
            del one
            if g % 64 == 63:
                print(f"[bench] one-by-one: {g + 1} scenes, {total:.2f} s", file=sys.stderr, flush=True)
        out["one_by_one_scene_steps_per_s"] = len(scs) * args.batch_steps / total
        out["one_by_one_scene_newton_iters"] = newton_s
        out["speedup_vs_one_by_one"] = out["scene_steps_per_s"] / out["one_by_one_scene_steps_per_s"]
    return out


def coupling_leg(args):
    """Cloth-rigid coupling (129 x 129 cloth pinned at its corners, 24 affine-body boxes): Newton
    steps/s of all frames in FP32 and in mixed precision with FP16 and with BF16 block storage, and
    the position error of each against the FP64 oracle over the first frames (a bounded CPU
    sample)."""
    import numpy as np

    import oracle
    from luisacomputesimulator_amd import NewtonSolver, hiprt, scenes

    out = {"workload": "129x129 cloth pinned at 4 corners + 24 affine-body boxes (0.1 m, solid, density 500, "
                       "stiffness 1e8), seed 0", "frames": args.coupling_frames}
    traj = {}
    for prec in ("fp32", "mixed", "mixed_bf16"):
        sc = scenes.cloth_rigid_coupling(129, 24, seed=0)
        sc.params.precision = prec
        g = NewtonSolver.from_scene(sc)
        nv = sc.num_verts
        hiprt.synchronize()
        newton, xs = 0, []
        t0 = time.perf_counter()
        for f in range(args.coupling_frames):
            g.physics_step_gpu(1)
            newton += g.stats()["newton_iters"]
            if f < args.coupling_oracle_frames:
                xs.append(g.get_simulation_results()[:nv].copy())
        hiprt.synchronize()
        el = time.perf_counter() - t0
        out[prec] = {"newton_steps_per_s": newton / el, "frames_per_s": args.coupling_frames / el,
                     "newton_iters": newton, "contacts_last_iter": g.stats()["num_contacts"]}
        traj[prec] = xs
        print(f"[bench] coupling {prec}: {out[prec]}", file=sys.stderr, flush=True)
    out["mixed_speedup_vs_fp32"] = out["mixed"]["newton_steps_per_s"] / out["fp32"]["newton_steps_per_s"]
    out["mixed_bf16_speedup_vs_fp32"] = out["mixed_bf16"]["newton_steps_per_s"] / out["fp32"]["newton_steps_per_s"]
    if args.coupling_oracle_frames > 0:
        sc = scenes.cloth_rigid_coupling(129, 24, seed=0)
        o64 = oracle.Oracle(sc, 64)
        err = {"fp32": [], "mixed": [], "mixed_bf16": []}
        t0 = time.perf_counter()
        for f in range(args.coupling_oracle_frames):
            o64.step()
            x64 = o64.get("x").reshape(-1, 3)[:sc.num_verts]
            scale = np.abs(x64).max()
            for prec in err:
                err[prec].append(float(np.abs(traj[prec][f] - x64).max() / scale))
            print(f"[bench] coupling FP64 oracle frame {f}: fp32 {err['fp32'][-1]:.2e} mixed {err['mixed'][-1]:.2e} "
                  f"mixed_bf16 {err['mixed_bf16'][-1]:.2e}", file=sys.stderr, flush=True)
        out["rel_error_vs_fp64_per_frame"] = err
        out["oracle_seconds"] = time.perf_counter() - t0
    return out


def twist_error_leg(args):
    """Twisting sheet at a size the FP64 oracle finishes in minutes (n x n, edges turned at -/+ pi
    rad/s, 120 frames): relative position error of the GPU's FP32, mixed (FP16) and mixed (BF16)
    forms, and of the FP32 oracle, against the FP64 oracle at frames 30, 60 and 120."""
    import numpy as np

    import oracle
    from luisacomputesimulator_amd import NewtonSolver, scenes

    n, marks = args.twist_error_n, (30, 60, 120)
    out = {"workload": f"twisting sheet {n}x{n}, edges turned at -/+ pi rad/s, 120 frames", "frames": list(marks)}
    xs = {}
    for prec in ("fp32", "mixed", "mixed_bf16"):
        sc = scenes.twisting_sheet(n)
        sc.params.precision = prec
        g = NewtonSolver.from_scene(sc)
        t0, newton, xs[prec] = time.perf_counter(), 0, {}
        for f in range(1, marks[-1] + 1):
            g.physics_step_gpu(1)
            newton += g.stats()["newton_iters"]
            if f in marks:
                xs[prec][f] = g.get_simulation_results().copy()
        out[prec] = {"newton_iters": newton, "seconds": time.perf_counter() - t0}
        print(f"[bench] twist {n} {prec}: {newton} Newton iterations", file=sys.stderr, flush=True)
    sc = scenes.twisting_sheet(n)
    o64, o32 = oracle.Oracle(sc, 64), oracle.Oracle(sc, 32)
    err = {k: [] for k in ("fp32", "mixed", "mixed_bf16", "oracle_fp32")}
    t0 = time.perf_counter()
    for f in range(1, marks[-1] + 1):
        o64.step()
        o32.step()
        if f in marks:
            x64 = o64.get("x").reshape(-1, 3)
            scale = np.abs(x64).max()
            for prec in ("fp32", "mixed", "mixed_bf16"):
                err[prec].append(float(np.abs(xs[prec][f] - x64).max() / scale))
            err["oracle_fp32"].append(float(np.abs(o32.get("x").reshape(-1, 3) - x64).max() / scale))
            print(f"[bench] twist {n} frame {f} vs FP64: " + ", ".join(f"{k} {v[-1]:.2e}" for k, v in err.items()),
                  file=sys.stderr, flush=True)
    out["rel_error_vs_fp64"] = err
    out["oracle_seconds"] = time.perf_counter() - t0
    return out


def cpu_baseline(scene):
    """The CPU oracle (FP32, one thread) on frame 0 of the same scene: Newton iterations per second."""
    import oracle

    o = oracle.Oracle(scene, 32)
    t0 = time.perf_counter()
    o.step()
    dt = time.perf_counter() - t0
    its = int(o.scalars()["newton_iters"])
    return {"value": its / dt, "unit": "Newton steps/s", "cores": 1, "kind": "port",
            "sample": f"frame 0 of the same {scene.num_verts}-vertex scene ({its} Newton iterations, "
                      f"{dt:.1f} s; C++ oracle, FP32, single thread)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("tube", "twist"), default="tube",
                    help="tube: the 297 x 297 cloth tube on the spinning cylinder (the headline workload); "
                         "twist: the 513 x 513 sheet twisted by its two pinned edges")
    ap.add_argument("--rings", type=int, default=297, help="tube: rings (= vertices per ring)")
    ap.add_argument("--n", type=int, default=513, help="twist: vertices per side")
    ap.add_argument("--precision", choices=("fp32", "mixed", "both"), default="both",
                    help="fp32: the reference's precision (the headline value); mixed: FP16 blocks in the PCG "
                         "with FP32 residual replacement, FP64 CCD/barrier; both: FP32 headline plus a mixed "
                         "run of the same frames")
    ap.add_argument("--batch-scenes", type=int, default=256, help="batched leg: number of scenes (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3, help="batched leg: timed frames")
    ap.add_argument("--batch-warmup", type=int, default=16, help="batched leg: frames before timing")
    ap.add_argument("--coupling-frames", type=int, default=60, help="cloth-rigid leg: GPU frames (0: skip)")
    ap.add_argument("--legs", default="tube,batch,coupling",
                    help="comma list of legs: tube (the headline), batch, coupling, twist_error (not run by "
                         "default: its FP64 oracle takes minutes)")
    ap.add_argument("--twist-error-n", type=int, default=65, help="twist_error leg: vertices per side")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default runtime mappings")
    ap.add_argument("--coupling-oracle-frames", type=int, default=24,
                    help="cloth-rigid leg: frames compared with the FP64 oracle")
    ap.add_argument("--profile-pcg", action="store_true",
                    help="after the warm-up frames run one calibration copy and one PCG solve, then exit "
                         "(workload for rocprofv3 counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the workload is a single scene on one MI355X and the solver "
              "has no multi-GPU path. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from luisacomputesimulator_amd import NewtonSolver, hiprt, scenes

    def make_scene():
        return scenes.cylinder_tube(rings=args.rings, per_ring=args.rings) if args.workload == "tube" \
            else scenes.twisting_sheet(args.n)

    sc = make_scene()
    if args.workload == "tube":
        label = (f"cloth tube {args.rings}x{args.rings} ({sc.meshes[0].verts.shape[0]} verts, "
                 f"{sc.meshes[0].tris.shape[0]} tris) around a spinning kinematic cylinder, dt 1/60 s, kappa 1e9")
    else:
        label = (f"twisting sheet {args.n}x{args.n} ({sc.num_verts} verts, {sc.num_faces} tris), edges turned "
                 "at -/+ pi rad/s, d_hat 5e-4 m, dt 1/60 s")

    def run(precision):
        """Warm-up and timed frames of a fresh solver in one precision."""
        s = make_scene()
        s.params.precision = precision
        g = NewtonSolver.from_scene(s)
        g.set_timing(2)
        tag = precision

        def progress(kind, f, st, t):
            print(f"[bench] {tag} {kind} frame {f}: {st['newton_iters']} Newton, {st['pcg_iters_total']} PCG "
                  f"iterations, {st['num_contacts']} contacts, {t:.2f} s", file=sys.stderr, flush=True)

        for f in range(args.warmup):
            tw = time.perf_counter()
            g.physics_step_gpu(1)
            progress("warmup", f, g.stats(), time.perf_counter() - tw)
        tuned = None
        if not args.no_autotune:
            # runtime mappings timed on the warm state (results do not depend on them)
            tuned = g.autotune()["chosen"]
            print(f"[bench] {tag} autotune: {tuned}", file=sys.stderr, flush=True)
        if args.profile_pcg:
            profile_pcg(g)
            return None
        g.kernel_times(reset=True)
        hiprt.synchronize()
        newton = pcg = 0
        t0 = time.perf_counter()
        for f in range(args.steps):
            tf = time.perf_counter()
            g.physics_step_gpu(1)
            st = g.stats()
            newton += st["newton_iters"]
            pcg += st["pcg_iters_total"]
            progress("timed", f, st, time.perf_counter() - tf)
        hiprt.synchronize()
        elapsed = time.perf_counter() - t0
        kt = g.kernel_times()
        n_rows = g.num_verts
        n_blocks = len(g.read("csr")) - (n_rows + 1)
        spmv_ms, spmv_n = kt.get("spmv", (float("nan"), 0))
        slots = kt.get("spmv_slots", (0.0, 0))[0]
        nbytes = spmv_bytes(n_rows, n_blocks, slots, 23 if precision == "mixed" else 40)
        return {"elapsed": elapsed, "newton": newton, "pcg": pcg, "kt": kt, "st": g.stats(), "n_rows": n_rows, "tuned": tuned,
                "n_blocks": n_blocks, "spmv_ms": spmv_ms, "spmv_n": spmv_n, "slots": slots, "nbytes": nbytes,
                "achieved": nbytes / (spmv_ms * 1e-3) / 1e9 if spmv_n else None}

    precisions = {"fp32": ["fp32"], "mixed": ["mixed"], "both": ["fp32", "mixed"]}[args.precision]
    legs = set(args.legs.split(","))
    if "tube" not in legs:
        # a profiling run of the other legs: their JSON objects only
        line = {"metric": METRIC, "value": None, "unit": "Newton steps/s", "n_gpus": 1, "legs": sorted(legs)}
        if "batch" in legs and args.batch_scenes > 0:
            line["batched"] = {"fp32": batched_leg(args, precisions[0], sequential=False)}
        if "coupling" in legs and args.coupling_frames > 0:
            line["coupling"] = coupling_leg(args)
        if "twist_error" in legs:
            line["twist_error"] = twist_error_leg(args)
        print(json.dumps(line), flush=True)
        return
    res = {}
    for prec in precisions:
        res[prec] = run(prec)
        if res[prec] is None:
            return
    head = res[precisions[0]]
    n_rows, n_blocks, st, kt = head["n_rows"], head["n_blocks"], head["st"], head["kt"]
    traffic = traffic_from_profiles()
    line = {
        "metric": METRIC,
        "value": head["newton"] / head["elapsed"],
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if precisions[0] == "fp32" else "f16/f32/f64 mixed",
        "data": "synthetic",
        "config": {"workload": label + (", FP32" if precisions[0] == "fp32" else ", mixed precision"),
                   "verts_total": n_rows, "material_blocks": n_blocks, "frames": args.steps,
                   "newton_iters": head["newton"], "pcg_iters": head["pcg"],
                   "frames_per_s": args.steps / head["elapsed"],
                   "contacts_last_iter": st["num_contacts"],
                   "pairs_last_iter": st["num_pairs_vf"] + st["num_pairs_ee"]},
        "roofline": {"bound": "hbm", "achieved": head["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (head["achieved"] / HBM_PEAK_GBS) if head["achieved"] else None,
                     "traffic": traffic[0] if precisions[0] == "fp32" else None,
                     "traffic_algorithmic_bytes": traffic[1] if precisions[0] == "fp32" else None,
                     "kernel": "k_spmv_pcg", "bytes_per_launch": head["nbytes"], "avg_launch_ms": head["spmv_ms"],
                     "contact_slots_per_launch": head["slots"],
                     "launches": head["spmv_n"]},
        "kernel_ms": {k: round(v[0], 4) for k, v in kt.items() if k != "spmv_slots"},
        "autotuned": head["tuned"],
    }
    if "mixed" in res and precisions[0] != "mixed":
        m = res["mixed"]
        line["mixed"] = {"value": m["newton"] / m["elapsed"], "unit": "Newton steps/s",
                         "speedup_vs_fp32": (m["newton"] / m["elapsed"]) / line["value"],
                         "ms_per_step": 1e3 * m["elapsed"] / args.steps,
                         "newton_iters": m["newton"], "pcg_iters": m["pcg"],
                         "spmv_avg_launch_ms": m["spmv_ms"], "spmv_bytes_per_launch": m["nbytes"],
                         "spmv_achieved_gbs": m["achieved"],
                         "pcg_ms": round(m["kt"].get("pcg", (float("nan"), 0))[0], 4)}
    if "batch" in legs and args.batch_scenes > 0:
        line["batched"] = {"workload": f"{args.batch_scenes} cloth-on-sphere scenes, n ~ U{{17..161}}, seed 0",
                           "unit": "scene-steps/s", "frames": args.batch_steps, "warmup_frames": args.batch_warmup,
                           "fp32": batched_leg(args, "fp32", sequential=True)}
        if "mixed" in precisions:
            line["batched"]["mixed"] = batched_leg(args, "mixed", sequential=False)
    if "coupling" in legs and args.coupling_frames > 0:
        line["coupling"] = coupling_leg(args)
    if "twist_error" in legs:
        line["twist_error"] = twist_error_leg(args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

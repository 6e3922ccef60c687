"""Benchmark of the MI355X DGSEM rhs! (one JSON line on stdout).

Workload (default): 3D compressible Euler, polydeg 3, flux differencing with
flux_ranocha, LLF surface flux, periodic TreeMesh [-2,2]^3 at refinement
level 5 (32768 elements x 64 nodes x 5 variables), seeded smooth periodic
state.  A step is one rhs! evaluation (u -> du) with u resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64]

`value` is DOF-updates/s of the mixed-precision rhs! (FP32 arithmetic, FP64
state / log-means); the FP64 rhs! and the FP32-state rhs! (the reference's
RealT = Float32 mode: FP32 state, du and logarithms) are reported beside it,
each with its own roofline; each mode also times one CarpenterKennedy2N54
step (five fused rhs! + register-update stages, hipGraph) with its HBM
roofline.  The CPU baseline is the FP64 oracle restatement (OpenMP) on a
bounded sample of the same workload and state.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "trixicuda.jl_amd")
# rocprofv3 FETCH_SIZE/WRITE_SIZE of the rhs kernel on this workload, copy-calibrated
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "rhs_kernel_euler3d_fd_ranocha_llf_level5_traffic.txt")


def load(name, path, pkg_dir=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=pkg_dir)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def make_semi(th, level, precision):
    eq = th.CompressibleEulerEquations3D(1.4)
    solver = th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs,
                         volume_integral=th.VolumeIntegralFluxDifferencing(th.flux_ranocha), precision=precision)
    mesh = th.TreeMesh((-2.0,) * 3, (2.0,) * 3, initial_refinement_level=level, n_cells_max=10**8)
    return th.SemidiscretizationHyperbolicHIP(mesh, eq, th.initial_condition_smooth_periodic(20261202), solver)


def time_rhs(torch, th, semi, u, du, steps, warmup, stream):
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            th.rhs_hip_(du, u, semi, 0.0)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record(stream)
        for _ in range(steps):
            th.rhs_hip_(du, u, semi, 0.0)
        stop.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / steps, wall / steps * 1e3


def time_vjp(torch, th, semi, u, steps, warmup, stream):
    vbar = torch.ones_like(u)
    ubar = torch.empty_like(u)
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            th.rhs_vjp(ubar, u, vbar, semi)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            th.rhs_vjp(ubar, u, vbar, semi)
        stop.record(stream)
        torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def _events_ms(torch, stream, fn, reps, warmup=3):
    """Average ms per call of fn() on `stream` (HIP events on that stream)."""
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _graph_ms(torch, fn, reps, check=None):
    """Average ms per call of fn() replayed from a HIP graph of `reps` calls
    (no host launch overhead: the device time of back-to-back rhs! calls).
    `check()` returns the output tensor; the graph's result must equal the
    eager one bit for bit."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    eager = check().clone() if check is not None else None
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    if check is not None and not torch.equal(check(), eager):
        raise RuntimeError("graph replay of rhs! differs from the eager result")
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(3):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / (3 * reps)


def roofline_of(fm, key, dofs, ms, precision):
    """HBM roof (algorithmic bytes: read u, write du) with the flop fraction
    beside it, for one rhs! of `dofs` DOF taking `ms` milliseconds."""
    secs = ms * 1e-3
    gbps = fm.BYTES_PER_DOF[key] * dofs / secs / 1e9
    fpeak = fm.PEAK["fp64_TFLOPs"] if precision == "fp64" else fm.PEAK["fp32_TFLOPs"]
    tf = fm.FLOPS_PER_DOF[key] * dofs / secs / 1e12
    return {"bound": "hbm", "achieved": gbps, "peak": fm.PEAK["hbm_GBps"], "unit": "GB/s",
            "frac": gbps / fm.PEAK["hbm_GBps"], "bytes_per_dof": fm.BYTES_PER_DOF[key],
            "flop_per_dof": fm.FLOPS_PER_DOF[key], "flop_TFLOPs": tf, "flop_peak_TFLOPs": fpeak,
            "flop_frac": tf / fpeak, "flop_model": key}


def measure_configs(torch, th, fm, stream, reps):
    """The other workloads of the benchmark set, each on one GPU (rhs_ms:
    device time per rhs! from a HIP graph of back-to-back calls; rhs_ms_eager:
    the same calls launched one by one from Python, which bounds the small
    meshes by the host's launch rate):
    C2 2D Euler FD (flux_ranocha + LLF), [-2,2]^2 level 5, seeded random state;
    C4 3D GLM-MHD Alfven wave (flux_hindenlang_gassner + Powell, LLF), [-1,1]^3 level 4;
    C5 2D Euler HG shock capturing (alpha_max 0.5), [-2,2]^2 level 6, seeded weak blast:
    FP64 and mixed rhs! times, the VJP and its ratio to the FP64 rhs!, and for
    C5 a 100-step mixed forward with the checkpointed FP64 adjoint."""
    out = {}
    specs = {
        "C2": dict(key="euler2d_fd_ranocha_llf", eq=th.CompressibleEulerEquations2D(1.4), nd=2, lo=-2.0, hi=2.0,
                   level=5,
                   vi=th.VolumeIntegralFluxDifferencing(th.flux_ranocha), sf=th.flux_lax_friedrichs,
                   ic=th.initial_condition_seeded_random(20261202), vjp=True),
        "C4": dict(key="mhd3d_fd_hindenlang_gassner_llf", eq=th.IdealGlmMhdEquations3D(5.0 / 3.0, 1.0), nd=3,
                   lo=-1.0, hi=1.0, level=4,
                   vi=th.VolumeIntegralFluxDifferencing((th.flux_hindenlang_gassner, th.flux_nonconservative_powell)),
                   sf=(th.flux_lax_friedrichs, th.flux_nonconservative_powell),
                   ic=th.initial_condition_convergence_test, vjp=True),
        "C5": dict(key="euler2d_hg_ranocha_llf", eq=th.CompressibleEulerEquations2D(1.4), nd=2, lo=-2.0, hi=2.0,
                   level=6,
                   vi=th.VolumeIntegralShockCapturingHG(
                       th.IndicatorHennemannGassner(th.CompressibleEulerEquations2D(1.4), th.LobattoLegendreBasisHIP(3),
                                                    alpha_max=0.5, alpha_min=0.001, alpha_smooth=True,
                                                    variable=th.density_pressure),
                       volume_flux_dg=th.flux_ranocha, volume_flux_fv=th.flux_lax_friedrichs),
                   sf=th.flux_lax_friedrichs, ic=th.initial_condition_weak_blast_wave_seeded(20261202), vjp=True),
    }
    for name, c in specs.items():
        mesh = th.TreeMesh((c["lo"],) * c["nd"], (c["hi"],) * c["nd"], initial_refinement_level=c["level"],
                           n_cells_max=10**8)
        r, du = {}, {}
        semis = {}
        for prec in ("fp64", "mixed"):
            solver = th.DGSEMHIP(polydeg=3, surface_flux=c["sf"], volume_integral=c["vi"], precision=prec)
            semi = th.SemidiscretizationHyperbolicHIP(mesh, c["eq"], c["ic"], solver)
            semis[prec] = semi
            u = th.compute_coefficients(0.0, semi)
            d = torch.empty_like(u)
            eager_ms = _events_ms(torch, stream, lambda: th.rhs_hip_(d, u, semi, 0.0), reps)
            ms = _graph_ms(torch, lambda: th.rhs_hip_(d, u, semi, 0.0), reps, check=lambda: d)
            torch.cuda.synchronize()
            du[prec] = d
            r[prec] = {"rhs_ms": ms, "rhs_ms_eager": eager_ms, "dof_per_s": semi.ndofs / (ms * 1e-3),
                       "roofline": roofline_of(fm, c["key"], semi.ndofs, ms, prec)}
        r["ndofs"] = semis["fp64"].ndofs
        a, b = du["mixed"], du["fp64"]
        r["mixed_vs_fp64_rel_l2"] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        r["mixed_vs_fp64_rel_max"] = float((a - b).abs().max() / b.abs().max())
        r["mixed_over_fp64"] = r["fp64"]["rhs_ms"] / r["mixed"]["rhs_ms"]
        if c["vjp"]:
            s64 = semis["fp64"]
            u = th.compute_coefficients(0.0, s64)
            vb = torch.randn_like(u)
            ub = torch.empty_like(u)
            vms = _graph_ms(torch, lambda: th.rhs_vjp(ub, u, vb, s64), max(reps // 4, 5), check=lambda: ub)
            r["vjp_ms"] = vms
            r["adjoint_over_forward"] = vms / r["fp64"]["rhs_ms"]
        if name == "C5":
            u0 = th.compute_coefficients(0.0, semis["fp64"])
            dt = 0.5 * th.max_dt(u0, semis["fp64"])
            for prec in ("mixed", "fp64"):
                semi = semis[prec]
                with torch.cuda.stream(stream):
                    # untimed, same length: loads the kernels and allocates the
                    # checkpoint and adjoint buffers (the caching allocator keeps them)
                    warm = th.forward_with_checkpoints(semi, u0, dt, 100)
                    th.backward(warm, th.l2_norm_and_gradient(warm.u_final, semi)[1])
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    traj = th.forward_with_checkpoints(semi, u0, dt, 100)
                    torch.cuda.synchronize()
                    t1 = time.perf_counter()
                    _, g = th.l2_norm_and_gradient(traj.u_final, semi)
                    grad = th.backward(traj, g)
                    torch.cuda.synchronize()
                    t2 = time.perf_counter()
                r[f"trajectory_100_steps_{prec}_forward"] = {"forward_ms": (t1 - t0) * 1e3,
                                                             "checkpointed_adjoint_ms": (t2 - t1) * 1e3}
                r.setdefault("_grads", {})[prec] = grad
            ga, gb = r["_grads"]["mixed"], r["_grads"]["fp64"]
            r["gradient_mixed_vs_fp64_rel_l2"] = float(torch.linalg.norm(ga - gb) / torch.linalg.norm(gb))
            del r["_grads"]
        for semi in semis.values():
            semi.close()
        out[name] = r
    return out


def time_rk(torch, th, semi, u, steps, stream):
    k = torch.zeros_like(u)
    uu = u.clone()
    with torch.cuda.stream(stream):
        for i in range(3):
            th.rk_step(uu, k, 0.0, 1e-4, semi, use_graph=True, stream=stream)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for i in range(steps):
            th.rk_step(uu, k, 0.0, 1e-4, semi, use_graph=True, stream=stream)
        stop.record(stream)
        torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def trajectory(torch, th, semi, u0, dt, steps, stream):
    """`steps` CK2N54 steps of size dt from u0 (hipGraph per step); the
    final state in FP64."""
    k = torch.zeros_like(u0)
    uu = u0.clone()
    with torch.cuda.stream(stream):
        for i in range(steps):
            th.rk_step(uu, k, i * dt, dt, semi, use_graph=True, stream=stream)
        torch.cuda.synchronize()
    return uu.double()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed",
                    help="mode reported as `value` (both are measured)")
    ap.add_argument("--only", choices=("mixed", "fp64", "fp32"), default=None,
                    help="measure one mode only (profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rk-steps", type=int, default=20)
    ap.add_argument("--traj-steps", type=int, default=100,
                    help="RK steps of the precision-comparison trajectory (0: none)")
    ap.add_argument("--vjp-steps", type=int, default=50)
    ap.add_argument("--no-extra", action="store_true", help="skip the C2/C4/C5 measurements")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU path is not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                                               "launches": args.steps}}))
        return

    th = load("trixi_hip", os.path.join(PKG_DIR, "__init__.py"), [PKG_DIR])
    fm = load("trixi_hip_flop_model", os.path.join(PKG_DIR, "flop_model.py"))
    prof = load("trixi_hip_profiling", os.path.join(PKG_DIR, "profiling.py"))
    traffic = prof.read_traffic(TRAFFIC_SUMMARY) if (args.level == 5 and os.path.exists(TRAFFIC_SUMMARY)) else {}
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    key = "euler3d_fd_ranocha_llf"
    flops_dof, bytes_dof = fm.FLOPS_PER_DOF[key], fm.BYTES_PER_DOF[key]

    stream = torch.cuda.Stream()
    modes = [args.only] if args.only else ["fp64", "mixed", "fp32"]
    res, du_out = {}, {}
    traj, traj_dt = {}, None
    u = None
    semi_keep = None
    for mode in modes:
        semi = make_semi(th, args.level, mode)
        if u is None:
            u = th.compute_coefficients(0.0, semi)
            if semi.fp32_state:
                u = u.double()
            torch.cuda.synchronize()
        # the FP32-state mode reads the same state rounded to FP32 and writes
        # FP32 du: half the bytes per DOF
        um = u.float() if semi.fp32_state else u
        bytes_dof = fm.BYTES_PER_DOF[key] // (2 if semi.fp32_state else 1)
        du = torch.empty_like(um)
        ms, wall_ms = time_rhs(torch, th, semi, um, du, args.steps, args.warmup, stream)
        dofs = semi.ndofs
        secs = ms * 1e-3
        fpeak = fm.PEAK["fp64_TFLOPs"] if mode == "fp64" else fm.PEAK["fp32_TFLOPs"]
        achieved_tf = flops_dof * dofs / secs / 1e12
        gbps = bytes_dof * dofs / secs / 1e9
        # Roofline: 405 flop / 80 B = 5 flop/B lies below the FP64 (9.8) and FP32
        # (19.7) ridge points, so the HBM roof is the one that bounds the kernel.
        res[mode] = {
            "value": dofs / secs, "ms_per_step": ms, "wall_ms_per_step": wall_ms,
            "roofline": {"bound": "hbm", "achieved": gbps, "peak": fm.PEAK["hbm_GBps"], "unit": "GB/s",
                         "frac": gbps / fm.PEAK["hbm_GBps"], "traffic": traffic.get(mode),
                         "kernel": "rhs_kernel<3, 4, Euler<3>, FD, flux_ranocha, LLF, "
                                   + {"fp64": "double, M=4>", "mixed": "float, M=16, logs double>",
                                      "fp32": "float, M=16, logs float, state float>"}[mode],
                         "algorithmic_bytes": bytes_dof * dofs, "bytes_per_dof": bytes_dof,
                         "flop_per_dof": flops_dof, "flop_TFLOPs": achieved_tf, "flop_peak_TFLOPs": fpeak,
                         "flop_frac": achieved_tf / fpeak},
        }
        if args.rk_steps > 0:
            rk_ms = time_rk(torch, th, semi, um, args.rk_steps, stream)
            res[mode]["rk_step_ms"] = rk_ms
            # One CarpenterKennedy2N54 step: five fused stages; the first reads u
            # and writes k and the stage output, the other four also read k.
            vb = 5 * (4 if semi.fp32_state else 8)  # bytes of one node's state
            rk_bytes = (3 + 4 * 4) * vb * dofs
            rk_gbps = rk_bytes / (rk_ms * 1e-3) / 1e9
            res[mode]["rk_step_roofline"] = {"bound": "hbm", "achieved": rk_gbps, "peak": fm.PEAK["hbm_GBps"],
                                             "unit": "GB/s", "frac": rk_gbps / fm.PEAK["hbm_GBps"],
                                             "algorithmic_bytes": rk_bytes, "bytes_per_dof": rk_bytes // dofs}
        if not args.only and args.traj_steps > 0:
            # every mode advances the same initial state with the same steps
            # (CFL 1 from the FP64 state's max_dt), for the solution-level
            # difference between the precisions
            if traj_dt is None:
                traj_dt = th.max_dt(u, semi)
            traj[mode] = trajectory(torch, th, semi, um, traj_dt, args.traj_steps, stream)
        if mode == "fp64" and args.vjp_steps > 0:
            vms = time_vjp(torch, th, semi, u, args.vjp_steps, 3, stream)
            res[mode]["vjp_ms"] = vms
            res[mode]["adjoint_over_forward"] = vms / ms
        du_out[mode] = du.double() if semi.fp32_state else du.clone()
        if semi_keep is None and not semi.fp32_state:
            semi_keep = semi  # FP64 vectors: the CPU baseline's unpacking
        else:
            semi.close()
    torch.cuda.synchronize()

    extra = None
    if not args.no_extra and not args.only:
        extra = measure_configs(torch, th, fm, stream, 100)

    errors = {}
    for m in ("mixed", "fp32"):
        if "fp64" in du_out and m in du_out:
            a, b = du_out[m], du_out["fp64"]
            errors[f"{m}_vs_fp64_rel_max"] = float((a - b).abs().max() / b.abs().max())
            errors[f"{m}_vs_fp64_rel_l2"] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    trajectory_errors = None
    if "fp64" in traj:
        b = traj["fp64"]
        trajectory_errors = {"steps": args.traj_steps, "dt": traj_dt, "t_end": traj_dt * args.traj_steps,
                             "cfl": 1.0}
        for m in ("mixed", "fp32"):
            if m in traj:
                a = traj[m]
                trajectory_errors[f"{m}_vs_fp64_rel_max"] = float((a - b).abs().max() / b.abs().max())
                trajectory_errors[f"{m}_vs_fp64_rel_l2"] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
                # relative to the state's change over the trajectory
                db = b - u
                trajectory_errors[f"{m}_vs_fp64_increment_rel_l2"] = float(torch.linalg.norm(a - b)
                                                                           / torch.linalg.norm(db))
                trajectory_errors[f"{m}_vs_fp64_increment_rel_max"] = float((a - b).abs().max() / db.abs().max())

    if semi_keep is None:
        semi_keep = make_semi(th, args.level, "fp64")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        # the host's own cores: what `nproc` reports (the process's CPU affinity)
        threads = len(os.sched_getaffinity(0))
        O.set_threads(threads)
        cfg = O.make_cfg(3, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_LLF, ic_length=4.0)
        mesh = O.Mesh(3, args.level, -2.0, 2.0, 3)
        u_np = th.wrap_array_native(u, semi_keep).cpu().numpy()
        n, t0 = 0, time.perf_counter()
        du_ref = None
        while True:
            du_ref = O.rhs(cfg, mesh, u_np)
            n += 1
            if time.perf_counter() - t0 >= args.cpu_seconds and n >= 2:
                break
        el = time.perf_counter() - t0
        dofs = mesh.nelements * 64
        cpu = {"value": dofs * n / el, "unit": "DOF-updates/s", "cores": threads, "nproc": threads,
               "os_cpu_count": os.cpu_count(), "kind": "port",
               "sample": f"{n} x rhs! of the full workload state ({dofs} DOF), FP64 C++ oracle (C++ restatement, not Trixi), OpenMP"}
        for mode, d in du_out.items():
            g = th.wrap_array_native(d, semi_keep).cpu().numpy()
            errors[f"{mode}_vs_oracle_rel_max"] = float(np.abs(g - du_ref).max() / np.abs(du_ref).max())
            errors[f"{mode}_vs_oracle_rel_l2"] = float(np.linalg.norm(g - du_ref) / np.linalg.norm(du_ref))

    head = args.precision if args.precision in res else modes[0]
    r = res[head]
    line = {
        "metric": "rhs! DOF-updates/s", "value": r["value"], "unit": "DOF-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if head == "mixed" else "f64",
        "data": "synthetic (seeded smooth periodic state)",
        "config": {"workload": f"3D Euler FD flux_ranocha + LLF, polydeg 3, TreeMesh [-2,2]^3 level {args.level} "
                               f"({8 ** args.level} elements, {8 ** args.level * 64} DOF, 5 vars), one rhs!",
                   "precision": head, "parallelism": "single GPU"},
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
        "modes": res,
        "errors": errors,
        "trajectory_errors": trajectory_errors,
        "other_configs": extra,
    }
    for m in ("mixed", "fp32"):
        if "fp64" in res and m in res:
            line[f"{m}_over_fp64"] = res[m]["value"] / res["fp64"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()

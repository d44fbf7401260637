"""Residual-path benchmark on one MI355X.

A step is one FP64 residual evaluation R(W) over the whole mesh (boundary
ghosts, gradients + limiters, face fluxes, ordered cell sums with the SA
source), on a synthetic state resident in HBM.  Prints one JSON line:
whole-job cell-residual evaluations per second; the roofline of the dominant
pass of the two-pass residual model from HIP events recorded live around each
kernel, with the whole residual's fraction beside it; implicit step times of
both Newton modes on the workload; the Newton race to a 1e-10 residual drop
(laminar cylinder 256x128, FD/FP64 against JVP/FP32, loads against the CPU
restatement); and the CPU baseline (the CPU restatement on all host cores, on
a bounded sample: residual evaluations and implicit steps).  The default
workload is the SA-RANS flat plate (6 equations per cell).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config plate_sa]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the four timed stages of fcfd_residual (fcfd_profile) and their kernels
STAGES = ["ghost", "grad_lim", "face_flux", "cell_sum"]
KERNEL_NAMES = {"ghost": "ghost_kernel", "grad_lim": "grad_lim_kernel", "face_flux": "face_flux_kernel",
                "cell_sum": "cell_sum_kernel"}
FACE_PASS_KERNELS = ["face_flux", "cell_sum"]


def make_case(name, scale=1):
    from freecfd_amd import cases
    if name == "plate_sa":
        return cases.plate_case(512 // scale, 128 // scale, 64 // scale, noise=0.01, sa=True)
    if name == "plate":
        return cases.plate_case(512 // scale, 128 // scale, 64 // scale, noise=0.01)
    if name == "bump":
        return cases.bump_case(256 // scale, 96 // scale, 64 // scale, noise=0.01)
    if name == "tetbox":
        return cases.tetbox_case(64 // scale)
    if name == "cylinder":
        return cases.cylinder_case(256 // scale, 128 // scale, noise=0.01)
    if name == "wedge":
        return cases.wedge_case(96 // scale, 48 // scale)
    raise SystemExit("unknown config " + name)


WORKLOADS = {
    "plate_sa": "SA-RANS flat plate 512x128x64 hex, top 20% layers split into prisms, Mach 0.3, Re_L 5e6, nu~ = 3 nu, "
                "viscous Roe + Spalart-Allmaras (coupled), 2nd-order curvilinear gradients + Venkatakrishnan limiter, "
                "6 equations/cell",
    "plate": "flat plate 512x128x64 hex, top 20% layers split into prisms, Mach 0.3, laminar viscous Roe, "
             "2nd-order curvilinear gradients + Venkatakrishnan limiter, 5 equations/cell",
    "bump": "3-D bump channel 256x96x64 hex, Mach 0.85, inviscid Roe, curvilinear gradients + Venkatakrishnan",
    "tetbox": "unit box 64^3 hexes x 6 tets, Mach 0.5, viscous Roe, lsqr gradients + Venkatakrishnan",
    "cylinder": "cylinder O-grid 256x128x1, Mach 0.2 Re 40, AUSM+up, Sutherland, lsqr gradients",
    "wedge": "Mach 2 compression ramp 96x48x1, Roe, lsqr gradients + minmod",
}


def model_bytes(info, n_grad, has_phi, sw=False):
    """Algorithmic (compulsory) HBM bytes per residual evaluation by the
    two-pass model of the residual path: every array element a pass needs,
    read or written once, from the mesh's actual sizes (nv = 5, or 6 with the
    SA model; 8 B per FP64, 4 B per int32).

    pass 1 (cells: gradients + limiters): W (N+B) nv 8, centroids (N+B) 24,
      CSR offsets 2 (N+1) 4, gradient-map entries n_g (4 + 24), cell->face
      n_f 4, face centroid + parent/neighbour F 32, write grad W N 3nv 8,
      write phi N nv 8
    pass 2 (faces): face arrays F 104 (topology 8, n/area/t1 56, centroid 24,
      face-average weights 16), gather W (N+B) nv 8, grad W N 3nv 8, phi
      N nv 8, centroids (N+B) 24, write R N nv 8

    The per-face D - C buffer the face pass writes and sums back is not
    compulsory (the model's face pass writes R), so it counts as overhead, in
    the measured traffic only.  The SA source adds V and d read and V S
    written by pass 1 and read by pass 2 (N 3 8 and N 8).  The boundary-ghost
    update is outside the two-pass model; its own bytes are listed apart."""
    N, B, F, H, nv = info.n_cells, info.n_ghosts, info.n_faces, info.n_halffaces, info.nv
    phi = N * nv * 8 if has_phi else 0
    pass1 = ((N + B) * nv * 8 + (N + B) * 24 + 2 * (N + 1) * 4 + n_grad * (4 + 24) + H * 4 + F * 32
             + N * 3 * nv * 8 + phi)
    pass2 = F * (104 + (16 if sw else 0)) + (N + B) * nv * 8 + N * 3 * nv * 8 + phi + (N + B) * 24 + N * nv * 8
    if nv == 6:
        pass1 += N * 3 * 8
        pass2 += N * 8
    ghost = B * (4 + 4 + 4 + 24 + 24 + 8 + nv * 8 + nv * 8)
    return {"pass1_gradients_limiters": pass1, "pass2_face_pass": pass2, "ghost": ghost}


def measured_traffic(config, cells, kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE runs (scaled by a 1 GiB copy, see
    profiles/residual_<config>_hbm_traffic.md), or None if not measured."""
    path = os.path.join(ROOT, "profiles", "residual_%s_traffic.json" % config)
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    if d.get("workload") != config or d.get("cells") != cells or kernel not in d.get("kernels", {}):
        return None
    return d["kernels"][kernel]["traffic_bytes"]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name, target_s=12.0, scale=2, newton_steps=2, cfl=5.0):
    """The CPU restatement on the host cores (OpenMP over its face and cell
    loops, every sum in the serial order) on a bounded sample of the same
    workload: the configuration at 1/scale of the mesh per direction.
    Residual evaluations for about target_s seconds, then implicit steps
    (FD Jacobian, time terms, block-Jacobi GMRES(30) capped at 30
    iterations, W += dW) from the same state."""
    from oracle.oracle import Oracle, num_threads
    cs = make_case(name, scale=scale)
    t0 = time.perf_counter()
    o = Oracle(cs.raw, cs.opt)
    setup = time.perf_counter() - t0
    o.residual(cs.W0)   # warm
    n = 0
    t0 = time.perf_counter()
    while True:
        o.residual(cs.W0)
        n += 1
        if time.perf_counter() - t0 > target_s:
            break
    dt = (time.perf_counter() - t0) / n
    o.set_state(cs.W0)
    steps = []
    for _ in range(newton_steps):
        t1 = time.perf_counter()
        st = o.newton_step(cfl=cfl, restart=30, max_iters=30, rtol=1e-6)
        st["ms_per_step"] = 1e3 * (time.perf_counter() - t1)
        steps.append(st)
    o.close()
    cores = num_threads()
    out = {"value": cs.n_cells / dt, "unit": "cell-residual evals/s", "cores": cores, "kind": "port",
           "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
           "sample": "%s at 1/%d scale per direction (%d cells), %d residual evaluations (update_boundaries + "
                     "calc_cell_grads + calc_limiter + residual face loop) in %.1f s on %d OpenMP threads; "
                     "set-up %.1f s" % (name, scale, cs.n_cells, n, n * dt, cores, setup),
           "newton_step": {"cells": cs.n_cells, "cfl": cfl, "krylov": "block-Jacobi GMRES(30), at most 30 iterations",
                           "ms_per_step": [round(x["ms_per_step"], 1) for x in steps],
                           "ms_residual_jacobian": [round(x["ms_residual_jacobian"], 1) for x in steps],
                           "ms_precond": [round(x["ms_precond"], 1) for x in steps],
                           "ms_krylov": [round(x["ms_krylov"], 1) for x in steps],
                           "krylov_iters": [x["krylov_iters"] for x in steps]}}
    print("bench: cpu baseline %.3e cell-residual evals/s on %d threads, Newton step %s ms (%d cells)"
          % (out["value"], cores, out["newton_step"]["ms_per_step"], cs.n_cells), file=sys.stderr, flush=True)
    return out


def newton_compare(ns, cs, steps, cfl=5.0, max_rel_update=0.2):
    """Implicit steps from the same initial state with the reference-style FD
    Jacobian and FP64 GMRES(30), and with the exact tangent (matrix-free) and
    FP32 GMRES(30); both precondition with the FD diagonal blocks (block
    Jacobi), do at most 30 Krylov iterations per step and limit the step to a
    20% change of p or T (the impulsive start of the plate otherwise drives
    the cell next to its velocity inlet supersonic, which ends the run as in
    the reference).  Wall time per step and its split, Krylov iterations and
    ||R||_2 before each step; a mode whose step fails reports the error."""
    import torch
    from freecfd_amd import _abi as A
    out = {}
    for label, jac, fp32 in (("fd_fp64", A.JAC_FD, False), ("jvp_fp32", A.JAC_JVP_FULL, True)):
        opts = A.newton_opts(cfl=cfl, jacobian=jac, fp32=fp32, restart=30, max_iters=30, max_rel_update=max_rel_update)
        ns.set_state(cs.W0)
        ns.newton_step(opts, check=False)   # allocation of the Krylov workspace, untimed
        ns.set_state(cs.W0)
        ms, parts, its, hist, err = [], [], [], [], None
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc, st = ns.newton_step(opts, check=False)
            torch.cuda.synchronize()
            if rc != 0:
                from freecfd_amd._lib import lib
                err = "libfcfd error %d: %s" % (rc, lib().fcfd_last_error(ns.h).decode())
                break
            ms.append(1e3 * (time.perf_counter() - t0))
            parts.append([st.ms_residual, st.ms_jacobian, st.ms_krylov, st.ms_update])
            its.append(st.krylov_iters)
            hist.append(st.res_norm)
        if not ms:
            out[label] = {"error": err}
            print("bench: newton %s failed: %s" % (label, err), file=sys.stderr, flush=True)
            continue
        p = np.median(np.array(parts), axis=0)
        out[label] = {"ms_per_step": float(np.median(ms)), "ms_residual": float(p[0]), "ms_jacobian_or_precond": float(p[1]),
                      "ms_krylov": float(p[2]), "ms_update": float(p[3]), "krylov_iters": its,
                      "res_norm_before_step": [float(h) for h in hist]}
        if err:
            out[label]["error"] = err
        print("bench: newton %s %.1f ms/step, ||R|| %s" % (label, np.median(ms), " ".join("%.3e" % h for h in hist)),
              file=sys.stderr, flush=True)
    return {"cfl": cfl, "steps": steps, "restart": 30, "max_krylov_iters": 30, "max_rel_update": max_rel_update, "modes": out}


RACES = {
    "cylinder64": ("laminar cylinder O-grid 64x32x1 (Mach 0.2, Re_D 40, AUSM+up, Sutherland, lsqr)",
                   lambda: __import__("freecfd_amd.cases", fromlist=["x"]).cylinder_case(64, 32)),
    "cylinder128": ("laminar cylinder O-grid 128x64x1 (Mach 0.2, Re_D 40, AUSM+up, Sutherland, lsqr)",
                    lambda: __import__("freecfd_amd.cases", fromlist=["x"]).cylinder_case(128, 64)),
    "cylinder256": ("laminar cylinder O-grid 256x128x1 (Mach 0.2, Re_D 40, AUSM+up, Sutherland, lsqr)",
                    lambda: __import__("freecfd_amd.cases", fromlist=["x"]).cylinder_case(256, 128)),
}


def race(name, cfl=1000.0, drop=1e-10, max_steps=200):
    """Implicit solve to a `drop` residual reduction from the freestream with
    (a) the reference-style FD Jacobian and FP64 GMRES and (b) the exact
    tangent, matrix-free, with FP32 GMRES; both preconditioned by one V-cycle
    of aggregation multigrid on the FD Jacobian blocks (one block-SGS sweep
    per level and side), GMRES(30) capped at 60 iterations per step, constant
    local CFL.  Reports wall time to the drop and the surface loads of both
    converged states against the CPU restatement evaluated on the JVP/FP32
    state."""
    import torch
    from freecfd_amd import _abi as A
    from freecfd_amd.ns import NavierStokes
    from oracle.oracle import Oracle
    label, make = RACES[name]
    cs = make()
    ns = NavierStokes(cs.raw, cs.opt)
    modes = {}
    states = {}
    for mode, jac, fp32 in (("fd_fp64", A.JAC_FD, False), ("jvp_fp32", A.JAC_JVP_FULL, True)):
        ns.set_state(cs.W0)
        opts = A.newton_opts(cfl=cfl, jacobian=jac, fp32=fp32, restart=30, max_iters=60, rtol=1e-4,
                             precond=A.PC_MULTIGRID, sweeps=1)
        r = ns.solve(opts, max_steps=max_steps, drop=drop)
        loads = ns.loads()
        states[mode] = ns.state_host(include_ghosts=False)
        modes[mode] = {"status": r["status"], "steps": r["steps"], "seconds": r["total_ms"] / 1e3, "drop": r["drop"],
                       "krylov_iters": int(np.sum(r["krylov_iters"])), "loads": loads.tolist()}
        print("bench: race %s %s: %s in %d steps, %.3f s, drop %.2e" % (name, mode, r["status"], r["steps"],
                                                                          r["total_ms"] / 1e3, r["drop"]),
              file=sys.stderr, flush=True)
    # the reference loads: the CPU restatement's loads of the converged
    # FD/FP64 state; each mode's GPU loads of its own converged state are
    # compared with them (so the JVP/FP32 figure holds both the solution's
    # and the evaluation's error)
    o = Oracle(cs.raw, cs.opt)
    o.residual(states["fd_fp64"])
    ref = o.loads()
    o.close()
    wall = int(np.argmax(np.abs(ref).sum(axis=1)))
    for mode in modes:
        got = np.array(modes[mode]["loads"])
        modes[mode]["load_err_vs_oracle_fd_fp64"] = float(np.abs(got[wall, :3] - ref[wall, :3]).max() /
                                                          np.abs(ref[wall, :3]).max())
    fd, jv = modes["fd_fp64"], modes["jvp_fp32"]
    both = fd["status"] == "converged" and jv["status"] == "converged"
    return {"metric": "wall time to a 1e-10 residual drop", "value": jv["seconds"], "unit": "s", "n_gpus": 1,
            "higher_is_better": False, "vs_baseline": None, "dtype": "f64 residual/update, f32 Krylov (jvp_fp32)",
            "data": "synthetic", "config": {"workload": label, "cells": cs.n_cells, "cfl": cfl, "modes": modes,
                                            "speedup_fd_over_jvp": (fd["seconds"] / jv["seconds"]) if both else None,
                                            "drag_lift_fd_vs_jvp": float(np.abs(np.array(fd["loads"])[wall, :3] -
                                                                                np.array(jv["loads"])[wall, :3]).max() /
                                                                         np.abs(np.array(jv["loads"])[wall, :3]).max())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="plate_sa", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (a known-size kernel for HBM counter calibration)")
    ap.add_argument("--race", default=None, choices=sorted(RACES),
                    help="instead of the residual benchmark: Newton to a 1e-10 residual drop with FD/FP64 and "
                         "JVP/FP32 on this case, one JSON line")
    ap.add_argument("--newton-steps", type=int, default=3,
                    help="implicit steps timed per Jacobian mode (FD/FP64 and JVP/FP32) on the workload after the residual "
                         "(step times only: both modes are capped at 30 Krylov iterations); 0 skips")
    ap.add_argument("--no-race", action="store_true",
                    help="skip the Newton race to a 1e-10 drop (laminar cylinder 256x128, FD/FP64 vs JVP/FP32)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the residual path is measured on one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from freecfd_amd import _abi as A
    from freecfd_amd._lib import lib
    from freecfd_amd.ns import NavierStokes

    if args.race:
        print(json.dumps(race(args.race)), flush=True)
        return

    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print("bench: copied 1 GiB device to device (counter calibration)", file=sys.stderr, flush=True)
        del a, b
    t0 = time.perf_counter()
    cs = make_case(args.config)
    print("bench: mesh %s, %d cells (%.1f s)" % (args.config, cs.n_cells, time.perf_counter() - t0), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ns = NavierStokes(cs.raw, cs.opt)
    print("bench: set-up %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    ns.set_state(cs.W0)
    info = ns.info
    for _ in range(args.warmup):
        ns.assemble_residual()
    torch.cuda.synchronize()
    ns.status()

    L = lib()
    L.fcfd_profile(ns.h, 1, None, None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ns.assemble_residual()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stage = (C.c_double * 4)()
    calls = C.c_longlong()
    L.fcfd_profile(ns.h, 0, stage, C.byref(calls))
    ns.status()
    R = ns.R[:, :info.n_cells]
    if not bool(torch.isfinite(R).all()):
        raise SystemExit("non-finite residual")

    ms_step = 1e3 * elapsed / args.steps
    evals_s = args.steps / elapsed
    avg_ms = {STAGES[k]: stage[k] / max(calls.value, 1) for k in range(4)}
    has_phi = cs.opt.limiter != A.LIMITER_NONE and cs.opt.order == A.ORDER_SECOND
    mb = model_bytes(info, info.n_gradmap, has_phi, sw=cs.opt.convective_flux == A.FLUX_SW)
    # the two passes of the model and the kernels (fcfd_profile stages) that carry them
    passes = {
        "pass2_face_pass": (FACE_PASS_KERNELS, mb["pass2_face_pass"]),
        "pass1_gradients_limiters": (["grad_lim"], mb["pass1_gradients_limiters"]),
    }
    pass_ms = {k: sum(avg_ms[x] for x in ks) for k, (ks, _) in passes.items()}
    dom = max(pass_ms, key=lambda k: pass_ms[k])
    achieved = passes[dom][1] / (pass_ms[dom] * 1e-3) / 1e9
    model_total = mb["pass1_gradients_limiters"] + mb["pass2_face_pass"]
    traffic = [measured_traffic(args.config, info.n_cells, k) for k in passes[dom][0]]
    line = {
        "metric": "cell-residual evals/s",
        "value": cs.n_cells * evals_s,
        "unit": "cell-residual evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "cells": info.n_cells, "faces": info.n_faces,
                   "boundary_faces": info.n_ghosts, "half_faces": info.n_halffaces,
                   "gradmap_entries": info.n_gradmap,
                   "residual_evals_per_s": evals_s,
                   "stage_ms": {k: round(v, 5) for k, v in avg_ms.items()},
                   "model_bytes": mb,
                   "whole_residual": {"model_bytes_per_eval": model_total,
                                      "achieved_gbs": model_total / (ms_step * 1e-3) / 1e9,
                                      "frac": model_total / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
                   "pass_roofline": {k: {"kernels": passes[k][0], "ms": pass_ms[k], "model_bytes": passes[k][1],
                                         "achieved_gbs": passes[k][1] / (pass_ms[k] * 1e-3) / 1e9,
                                         "frac": passes[k][1] / (pass_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS}
                                     for k in passes}},
        "roofline": {"bound": "hbm", "kernel": dom + " (" + " + ".join(KERNEL_NAMES[k] for k in passes[dom][0]) + ")",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": sum(traffic) if all(t is not None for t in traffic) else None,
                     "algorithmic_bytes_per_launch": passes[dom][1], "avg_launch_ms": pass_ms[dom]},
    }
    if args.newton_steps > 0:
        line["config"]["newton_step_times"] = newton_compare(ns, cs, args.newton_steps)
    if not args.no_race:
        r = race("cylinder256")
        line["config"]["newton_race"] = {"workload": r["config"]["workload"], "cells": r["config"]["cells"],
                                         "cfl": r["config"]["cfl"], "modes": r["config"]["modes"],
                                         "speedup_fd_over_jvp": r["config"]["speedup_fd_over_jvp"],
                                         "drag_lift_fd_vs_jvp": r["config"]["drag_lift_fd_vs_jvp"]}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.config)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

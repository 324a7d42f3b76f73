#!/usr/bin/env python
"""Benchmark of the linear-tetmesh hot path on one MI355X.

Workload (one "step"): the Neo-Hookean Hessian of a 64x64x256-vertex beam (6,072,570 tets,
138,917,358 nonzeros) with the random smooth bend of the at-scale configuration, including the
simple_psd_fix(1e-3) element projection and the deterministic CSR assembly -- i.e. one call of
bh_linear_tetmesh_dq2_device with inputs already resident in HBM.  ``value`` is tets per second
(Mtets/s) over exactly K timed steps.

Also reported (not part of ``value``): the same call without the PSD projection, the gradient and
the energy kernels, the roofline of the dominant kernel (the assembly gather), measured live with
HIP events on the stream the kernels run on, and the FP64 CPU oracle (the reference algorithm
restated in C++: dense B^T H B, 144 triplets per tet, 12x12 Jacobi PSD, sort-and-fold) timed on a
bounded slab of the same mesh.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published


def algorithmic_bytes(nV, nT, nnz):
    """per-tet algorithmic bytes: E 16 + dphidX 96 + volume 8 + params 16, q once per vertex,
    and for the Hessian every CSR value written once (metadata excluded)."""
    b_in = 16 + 96 + 8 + 16
    b_q = b_in + 24.0 * nV / nT
    return dict(q=b_q, dq=b_q + 24.0 * nV / nT, dq2=b_q + 8.0 * nnz / nT)


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 counters (profiles/*.json), if committed"""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(name)
    except Exception:
        return None


def cpu_baseline(p, dX, vol, budget_s):
    """the FP64 oracle (reference algorithm, 1 thread) on leading slabs of the same mesh and deformation"""
    from oracle import oracle as O

    O.load()
    nV = p.V.shape[0]
    slab = 40000
    done, t0 = 0, time.perf_counter()
    while True:
        sl = slice(done, done + slab)
        O.tetmesh_dq2(0, nV, p.E[sl], p.q, dX[sl], vol[sl], p.params[sl], psd=True, psd_tol=1e-3)
        done += slab
        el = time.perf_counter() - t0
        if el >= budget_s or done >= p.E.shape[0]:
            break
    return dict(value=done / el / 1e6, unit="Mtets/s", cores=1, kind="port",
                sample=f"first {done} tets of the 64x64x256 bend, neohookean dq2 + simple_psd_fix(1e-3), "
                       f"{el:.1f} s on one host core")


def newton_steps(p, mesh, steps, pcg_max):
    """Newton-step wall time on the same beam (bent state, gravity, x = 0 clamped, PSD 1e-3, PCG to a
    1e-6 relative residual): FP64 elements with block-Jacobi PCG (the plain solver) against the FP32
    precision form with block Jacobi, block IC(0), block symmetric Gauss-Seidel (Eisenstat form),
    two-level block Jacobi (quadratic coarse space on vertex aggregates, also in FP64), block IC(0)
    with the same coarse correction (also in FP64), the learned factor with the coarse correction
    (untrained network) and the learned Eisenstat diagonal (when trained weights are present)."""
    import torch

    from bartels_amd import _lib, learned

    lib = _lib.load()
    nV = p.V.shape[0]
    fixed = np.zeros(3 * nV, np.uint8)
    fixed[(3 * p.fixed[:, None] + np.arange(3)).ravel()] = 1
    q0 = p.q.copy()
    q0[fixed == 1] = p.V.ravel()[fixed == 1]
    dev = torch.device("cuda:0")
    dparams = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device=dev)
    dfixed = torch.tensor(fixed, device=dev)
    f = torch.zeros(3 * nV, dtype=torch.float64, device=dev)
    lib.bh_linear_tetmesh_const_ext_force_device(mesh.handle, (ctypes.c_double * 3)(*(p.density * p.gravity)),
                                                 ctypes.c_void_p(f.data_ptr()), None)
    P = ctypes.c_void_p()
    _lib.check(lib.bh_precond_create(mesh.handle, ctypes.byref(P)), "bh_precond_create")
    net = learned.Network(weights=learned.DEFAULT_WEIGHTS) if os.path.exists(learned.DEFAULT_WEIGHTS) else None
    # precision 1: FP32 Hessian / CG operator, FP64 element gradient; 2: FP32 element gradient too
    variants = [("fp64_jacobi", 0, 0), ("fp32_jacobi", 1, 0), ("fp32_bic0", 1, 2), ("fp32_sgs", 1, 5),
                ("fp32full_sgs", 2, 5), ("fp64_jacobi2l", 0, 7), ("fp32_jacobi2l", 1, 7),
                ("fp64_bic0_2l", 0, 2, 1), ("fp32_bic0_2l", 1, 2, 1)]
    if net is not None:
        variants.append(("fp32_learned_sgs", 1, net.precond))
    # the learned factor form with the coarse correction, untrained: its decoders are zero, so the
    # network's factor is block Jacobi; the timing includes the network's inference per Newton step
    net0 = learned.Network(form="factor")
    variants.append(("fp32_learned_factor_untrained_2l", 1, net0.precond, 1))
    res, qs = {}, {}
    for name, prec, precond, *coarse in variants:
        o = _lib.BhNewtonOpts()
        lib.bh_newton_default_opts(ctypes.byref(o))
        o.model = 0
        o.psd = _lib.BhPsd(1, 1e-3)
        o.gradient_tol = 0.0
        o.max_iterations = steps - 1
        o.pcg_tol = 1e-6
        o.pcg_max_iterations = pcg_max
        o.precond = precond
        o.coarse_correction = coarse[0] if coarse else 0
        o.precond_handle = P
        o.element_precision = prec
        if precond == net0.precond:
            o.mgn = net0.h
        elif net is not None:
            o.mgn = net.h
        dq = torch.tensor(q0, device=dev)
        rep = _lib.BhNewtonReport()
        tr = np.zeros(4 * (steps + 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = lib.bh_newton_static_device(mesh.handle, ctypes.byref(o), ctypes.c_void_p(dparams.data_ptr()),
                                         ctypes.c_void_p(dfixed.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                                         ctypes.c_void_p(dq.data_ptr()), ctypes.byref(rep),
                                         tr.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        qs[name] = dq.cpu().numpy()
        res[name] = dict(status=st, newton_steps=rep.iterations, s_per_step=wall / max(1, rep.iterations),
                         pcg_iterations=[int(x) for x in tr.reshape(-1, 4)[:rep.iterations, 3]],
                         pcg_failures=rep.pcg_failures, pcg_fallbacks=rep.pcg_fallbacks,
                         objective=rep.energy)
    lib.bh_precond_destroy(P)
    if net is not None:
        net.close()
    net0.close()
    u = np.linalg.norm(qs["fp64_jacobi"] - q0)
    for name in qs:
        res[name]["rel_displacement_vs_fp64_jacobi"] = float(np.linalg.norm(qs[name] - qs["fp64_jacobi"]) / u)
    base = res["fp64_jacobi"]["s_per_step"]
    for name in res:
        res[name]["speedup_vs_fp64_jacobi"] = base / res[name]["s_per_step"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="256,64,64", help="vertices along x,y,z")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--gather-lanes", type=int, default=0, help="assembly kernel shape: 1, 2 or 4 lanes per vertex "
                    "(0: the library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--newton-steps", type=int, default=2,
                    help="Newton steps per solver variant in the Newton-step comparison (0: skip)")
    ap.add_argument("--pcg-max", type=int, default=20000)
    ap.add_argument("--quick", action="store_true", help="timed steps only (for counter collection)")
    ap.add_argument("--newton-only", action="store_true",
                    help="only the Newton-step comparison (for the kernel profile of the solvers)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (reference point for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding is a design, not code); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from bartels_amd import _lib, meshgen
    from bartels_amd import tetmesh as tm

    grid = tuple(int(x) for x in args.grid.split(","))
    t_setup = time.perf_counter()
    p = meshgen.config2(grid=grid)
    dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
    mesh = tm.TetMesh(p.V.shape[0], p.E, dX, vol, V=p.V)
    nV, nT, nnz = mesh.nV, mesh.nT, mesh.nnz
    dev = torch.device("cuda:0")
    q = torch.tensor(p.q, device=dev)
    params = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device=dev)
    H = torch.empty(nnz, dtype=torch.float64, device=dev)
    g = torch.empty(3 * nV, dtype=torch.float64, device=dev)
    e = torch.empty(1, dtype=torch.float64, device=dev)
    psd = tm.simple_psd_fix(1e-3)
    stream = torch.cuda.current_stream()
    lib0 = _lib.load()
    if args.gather_lanes:
        assert lib0.bh_mesh_set_gather_lanes(mesh.handle, args.gather_lanes) == 0
    setup_s = time.perf_counter() - t_setup
    if args.newton_only:
        print(json.dumps(newton_steps(p, mesh, max(1, args.newton_steps), args.pcg_max)), flush=True)
        return

    def step():
        mesh.dq2_device(tm.NEOHOOKEAN, q, params, H, callback=psd, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib = _lib.load()
    lib.bh_mesh_set_profiling(mesh.handle, 1)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    ms = np.zeros(len(_lib.KERNEL_CLASSES))
    cnt = np.zeros(len(_lib.KERNEL_CLASSES), dtype=np.int64)
    lib.bh_mesh_kernel_times(mesh.handle, ms.ctypes.data, cnt.ctypes.data)
    lib.bh_mesh_set_profiling(mesh.handle, 0)
    kt = {name: (ms[i] / cnt[i] if cnt[i] else None) for i, name in enumerate(_lib.KERNEL_CLASSES)}

    flagged = ctypes.c_int64()
    lib.bh_mesh_psd_count(mesh.handle, ctypes.byref(flagged))

    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.quick:
        print(json.dumps({"ms_per_step": ms_step, "kernels_ms": {k: v for k, v in kt.items() if v is not None}}))
        return

    # secondary kernels, each timed on its own (not part of value)
    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def timed_kernels(fn, reps=10):
        """(ms per call, {kernel class: ms per launch}) with the library's own event pairs"""
        fn()
        torch.cuda.synchronize()
        lib.bh_mesh_set_profiling(mesh.handle, 1)
        t = timed(fn, reps)
        lib.bh_mesh_kernel_times(mesh.handle, ms.ctypes.data, cnt.ctypes.data)
        lib.bh_mesh_set_profiling(mesh.handle, 0)
        return t, {name: round(ms[i] / cnt[i], 4) for i, name in enumerate(_lib.KERNEL_CLASSES) if cnt[i]}

    def tune(lanes=0, xcd=-1):
        assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(_lib.BhTuning(lanes, xcd))) == 0

    ms_nopsd, k_nopsd = timed_kernels(lambda: mesh.dq2_device(tm.NEOHOOKEAN, q, params, H, stream=stream))
    gather_nopsd = k_nopsd["hessian_gather"]
    ms_dq = timed(lambda: mesh.dq_device(tm.NEOHOOKEAN, q, params, g, stream=stream))
    ms_q = timed(lambda: mesh.q_device(tm.NEOHOOKEAN, q, params, e, stream=stream))
    # the other kernel schedules (gather lanes x XCD grouping, PSD two-pass / fused) and the FP32
    # precision form, same workload; none changes a result bit
    variants = {}
    par32 = params.to(torch.float32)
    H32 = torch.empty(nnz, dtype=torch.float32, device=dev)
    for lanes in (1, 2, 4):
        for xcd in (0, 1):
            tune(lanes, xcd)
            key = f"lanes{lanes}_xcd{xcd}"
            variants[key + "_psd"] = timed_kernels(step)
            variants[key + "_nopsd"] = timed_kernels(lambda: mesh.dq2_device(tm.NEOHOOKEAN, q, params, H, stream=stream))
            variants[key + "_fp32_psd"] = timed_kernels(
                lambda: mesh.dq2_f32_device(tm.NEOHOOKEAN, q, par32, H32, callback=psd, stream=stream))
            variants[key + "_fp32_nopsd"] = timed_kernels(
                lambda: mesh.dq2_f32_device(tm.NEOHOOKEAN, q, par32, H32, stream=stream))
    tune(args.gather_lanes or 4, 0)  # back to the library default (or the requested shape)
    del H32

    B = algorithmic_bytes(nV, nT, nnz)
    gather_ms = kt["hessian_gather"]
    achieved = B["dq2"] * nT / (gather_ms * 1e-3) / 1e9
    traffic = load_traffic("hessian_gather_psd")
    out = {
        "metric": "Mtets/s for neohookean_dq2 assembly",
        "value": nT / (ms_step * 1e-3) / 1e6,
        "unit": "Mtets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"neohookean_dq2 + simple_psd_fix(1e-3) + CSR assembly, {grid[1]}x{grid[2]}x{grid[0]} beam "
                        f"with random smooth bend, {nT} tets, {nnz} nnz, FP64",
            "grid": list(grid), "tets": nT, "vertices": nV, "nnz": nnz,
        },
        "roofline": {
            "kernel": "k_hessian_rec (assembly gather from the PSD records)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_tet": B["dq2"],
        },
        "kernels_ms": {k: v for k, v in kt.items() if v is not None},
        "extra": {
            "dq2_no_psd_ms": ms_nopsd,
            "dq2_no_psd_mtets_s": nT / (ms_nopsd * 1e-3) / 1e6,
            "gather_no_psd_ms": gather_nopsd,
            "kernels_no_psd_ms": k_nopsd,
            "gather_no_psd_hbm_frac": B["dq2"] * nT / (gather_nopsd * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "dq_ms": ms_dq,
            "dq_hbm_frac": B["dq"] * nT / (ms_dq * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "q_ms": ms_q,
            "q_hbm_frac": B["q"] * nT / (ms_q * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "variants_ms": variants,
            "psd_clamped_tets": int(flagged.value),
            "wall_s_timed": wall,
            "setup_s": setup_s,
        },
    }
    if args.newton_steps > 0:
        del H, g
        nres = newton_steps(p, mesh, args.newton_steps, args.pcg_max)
        out["extra"]["newton"] = nres
        # the fastest variant that stays within 1e-5 relative displacement of the FP64 block-Jacobi run
        ok = {k: v for k, v in nres.items() if v["rel_displacement_vs_fp64_jacobi"] < 1e-5 and v["status"] in (0, 4)}
        if ok:
            best = max(ok, key=lambda k: ok[k]["speedup_vs_fp64_jacobi"])
            out["extra"]["newton_best"] = dict(solver=best, s_per_step=ok[best]["s_per_step"],
                                               fp64_jacobi_s_per_step=nres["fp64_jacobi"]["s_per_step"],
                                               speedup_vs_fp64_jacobi=ok[best]["speedup_vs_fp64_jacobi"],
                                               rel_displacement=ok[best]["rel_displacement_vs_fp64_jacobi"],
                                               pcg_tol=1e-6)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, dX, vol, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python
"""Benchmark of the linear-tetmesh hot path on one MI355X.

Workload (one "step"): the Neo-Hookean Hessian of a 64x64x256-vertex beam (6,072,570 tets,
138,917,358 nonzeros) with the random smooth bend of the at-scale configuration, including the
simple_psd_fix(1e-3) element projection and the deterministic CSR assembly -- i.e. one call of
bh_linear_tetmesh_dq2_device with inputs already resident in HBM.  ``value`` is tets per second
(Mtets/s) over exactly K timed steps.

Also reported (not part of ``value``): the same call without the PSD projection, the gradient and
the energy kernels, the roofline of the dominant kernel (the assembly gather), measured live with
HIP events on the stream the kernels run on, the Newton-step and time-to-solution comparisons of
the PCG solvers against an FP64 judge, and the FP64 CPU oracle's timing build (the reference
algorithm restated in C++: dense B^T H B, 144 triplets per tet, 12x12 PSD by tridiagonalisation +
QR, sort-and-fold; one thread and all host threads) timed on bounded slabs of the same mesh.
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
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix), AMD published


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
    """The FP64 oracle's timing build (the reference algorithm restated in C++: index-functor B, dense
    B^T d2psi B, simple_psd_fix by tridiagonalisation + implicit QR, 144 triplets per tet, one
    setFromTriplets fold; g++ -O3 -march=native) on consecutive slabs of the same mesh and deformation,
    each slab renumbered to its own vertices so the fold's cost per tet is that of a whole mesh.  Two
    legs: one thread (what the reference runs) and every usable host thread (OpenMP over elements,
    triplets written in element order).  value = the all-threads leg."""
    from oracle import oracle as O

    O.load(fast=True)
    model, ncpu, _ = O.host_cpu()
    threads = max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu))))
    slab = 50000

    def leg(nthreads, budget):
        done, el, el_elem = 0, 0.0, 0.0
        while el < budget and done < p.E.shape[0]:
            E = p.E[done:done + slab]
            verts, Eloc = np.unique(E, return_inverse=True)
            Eloc = Eloc.reshape(E.shape).astype(np.int32)
            q = p.q.reshape(-1, 3)[verts].ravel()
            t0 = time.perf_counter()
            _, tt = O.tetmesh_dq2_timed(0, verts.size, Eloc, q, dX[done:done + slab], vol[done:done + slab],
                                        p.params[done:done + slab], psd=True, psd_tol=1e-3, threads=nthreads)
            el += time.perf_counter() - t0
            el_elem += tt[0]
            done += E.shape[0]
        return dict(mtets_s=done / el / 1e6, tets=done, seconds=el, element_seconds=el_elem)

    one = leg(1, 0.5 * budget_s)
    many = leg(threads, 0.5 * budget_s) if threads > 1 else one
    return dict(value=many["mtets_s"], unit="Mtets/s", cores=threads, kind="port",
                sample=f"consecutive 50,000-tet slabs of the 64x64x256 bend (neohookean dq2 + simple_psd_fix(1e-3) + "
                       f"setFromTriplets): {many['tets']} tets on {threads} threads in {many['seconds']:.1f} s; "
                       f"{one['tets']} tets on 1 thread in {one['seconds']:.1f} s",
                one_core=one, all_cores=many, host_cpu=model, host_threads_usable=ncpu,
                compiler_flags="g++ -O3 -march=native -fopenmp")


class NewtonBench:
    """Newton solves on the bench beam (bent state, gravity, x = 0 clamped, PSD 1e-3) with one solver
    variant per call: (precision, preconditioner[, coarse correction]).  precision 1 = FP32 Hessian and
    CG operator with the FP64 element gradient; 2 = FP32 element gradient too."""

    def __init__(self, p, mesh, pcg_max):
        import torch

        from bartels_amd import _lib, learned

        self.lib = lib = _lib.load()
        self._lib, self.torch, self.mesh, self.pcg_max = _lib, torch, mesh, pcg_max
        nV = p.V.shape[0]
        fixed = np.zeros(3 * nV, np.uint8)
        fixed[(3 * p.fixed[:, None] + np.arange(3)).ravel()] = 1
        self.q0 = p.q.copy()
        self.q0[fixed == 1] = p.V.ravel()[fixed == 1]
        dev = torch.device("cuda:0")
        self.dparams = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device=dev)
        self.dfixed = torch.tensor(fixed, device=dev)
        self.f = torch.zeros(3 * nV, dtype=torch.float64, device=dev)
        lib.bh_linear_tetmesh_const_ext_force_device(mesh.handle, (ctypes.c_double * 3)(*(p.density * p.gravity)),
                                                     ctypes.c_void_p(self.f.data_ptr()), None)
        self.P = ctypes.c_void_p()
        _lib.check(lib.bh_precond_create(mesh.handle, ctypes.byref(self.P)), "bh_precond_create")
        # the shipped weights (the tied factor form) and their untrained twin: at zero decoders the tied
        # factor is block symmetric Gauss-Seidel written as a factor
        self.net = (learned.Network(weights=learned.DEFAULT_WEIGHTS, form=learned.DEFAULT_FORM)
                    if os.path.exists(learned.DEFAULT_WEIGHTS) else None)
        self.net0 = learned.Network(form=learned.DEFAULT_FORM)

    def close(self):
        self.lib.bh_precond_destroy(self.P)
        if self.net is not None:
            self.net.close()
        self.net0.close()

    def run(self, prec, precond, coarse=0, pcg_tol=1e-6, max_it=1, gtol=0.0, mgn=None):
        _lib, torch = self._lib, self.torch
        o = _lib.BhNewtonOpts()
        self.lib.bh_newton_default_opts(ctypes.byref(o))
        o.model = 0
        o.psd = _lib.BhPsd(1, 1e-3)
        o.gradient_tol = gtol
        o.relative_tol = 1
        o.max_iterations = max_it - 1
        o.pcg_tol = pcg_tol
        o.pcg_max_iterations = self.pcg_max
        o.precond = precond
        o.coarse_correction = coarse
        o.precond_handle = self.P
        o.element_precision = prec
        if mgn is not None:
            o.mgn = mgn
        dq = torch.tensor(self.q0, device="cuda:0")
        rep = _lib.BhNewtonReport()
        tr = np.zeros(4 * (max_it + 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = self.lib.bh_newton_static_device(self.mesh.handle, ctypes.byref(o),
                                              ctypes.c_void_p(self.dparams.data_ptr()),
                                              ctypes.c_void_p(self.dfixed.data_ptr()),
                                              ctypes.c_void_p(self.f.data_ptr()), ctypes.c_void_p(dq.data_ptr()),
                                              ctypes.byref(rep), tr.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        n = rep.iterations
        return dq.cpu().numpy(), dict(status=st, newton_steps=n, wall_s=wall, s_per_step=wall / max(1, n),
                                      pcg_iterations=[int(x) for x in tr.reshape(-1, 4)[:n, 3]],
                                      pcg_failures=rep.pcg_failures, pcg_fallbacks=rep.pcg_fallbacks,
                                      objective=rep.energy, converged=bool(rep.converged),
                                      rel_gradient=rep.gradient_norm / rep.gradient_norm0 if rep.gradient_norm0 else 0.0)

    def variants(self):
        """(name, element precision, preconditioner, coarse correction, network handle or None)"""
        v = [("fp64_jacobi", 0, 0, 0, None), ("fp32_jacobi", 1, 0, 0, None), ("fp32_bic0", 1, 2, 0, None),
             ("fp32_sgs", 1, 5, 0, None), ("fp32full_sgs", 2, 5, 0, None), ("fp64_jacobi2l", 0, 7, 0, None),
             ("fp32_jacobi2l", 1, 7, 0, None), ("fp64_bic0_2l", 0, 2, 1, None), ("fp32_bic0_2l", 1, 2, 1, None),
             ("fp32_learned_untrained_2l", 1, self.net0.precond, 1, self.net0.h)]
        if self.net is not None:
            v.append(("fp32_learned_2l", 1, self.net.precond, 1, self.net.h))
        return v


# the FP64 judge of the Newton comparisons: FP64 elements, block IC(0) with the coarse correction, PCG to
# a 1e-12 relative residual -- within rounding of the reference's direct solve
JUDGE = ("fp64_judge", 0, 2, 1)


def newton_steps(p, mesh, steps, pcg_max, solve_tol=1e-6, solve_max_it=40, solve_variants=None):
    """(1) Newton-step wall time: `steps` Newton steps from the same start for every variant, PCG to a
    1e-6 relative residual, and every variant's displacement against the FP64 judge (the same steps with
    PCG to 1e-12).  (2) Time to solution: Newton to a 1e-6 relative gradient for the plain solver (FP64
    block-Jacobi PCG), the fastest per-step variants and the judge, with the converged displacements
    compared against the judge's."""
    nb = NewtonBench(p, mesh, pcg_max)
    res, qs = {}, {}
    try:
        qj, rj = nb.run(JUDGE[1], JUDGE[2], JUDGE[3], pcg_tol=1e-12, max_it=steps)
        res[JUDGE[0]] = rj
        for name, prec, precond, coarse, mgn in nb.variants():
            qs[name], res[name] = nb.run(prec, precond, coarse, pcg_tol=1e-6, max_it=steps, mgn=mgn)
        u = np.linalg.norm(qj - nb.q0)
        for name in qs:
            res[name]["rel_displacement_vs_fp64_judge"] = float(np.linalg.norm(qs[name] - qj) / u)
            res[name]["rel_displacement_vs_fp64_jacobi"] = float(np.linalg.norm(qs[name] - qs["fp64_jacobi"]) / u)
        base = res["fp64_jacobi"]["s_per_step"]
        for name in res:
            res[name]["speedup_vs_fp64_jacobi"] = base / res[name]["s_per_step"]
        # time to solution (relative gradient solve_tol)
        sol = {}
        if solve_variants is None:
            solve_variants = ["fp64_jacobi", "fp32_bic0_2l", "fp32_jacobi2l", "fp32_learned_untrained_2l"] + (
                ["fp32_learned_2l"] if nb.net is not None else [])
        qsj, sj = nb.run(JUDGE[1], JUDGE[2], JUDGE[3], pcg_tol=1e-12, max_it=solve_max_it, gtol=solve_tol)
        sol[JUDGE[0]] = sj
        uj = np.linalg.norm(qsj - nb.q0)
        table = {v[0]: v for v in nb.variants()}
        for name in solve_variants:
            _, prec, precond, coarse, mgn = table[name]
            qv, rv = nb.run(prec, precond, coarse, pcg_tol=1e-6, max_it=solve_max_it, gtol=solve_tol, mgn=mgn)
            rv["rel_displacement_vs_fp64_judge"] = float(np.linalg.norm(qv - qsj) / uj)
            sol[name] = rv
        for name in sol:
            sol[name]["speedup_vs_fp64_jacobi"] = sol["fp64_jacobi"]["wall_s"] / sol[name]["wall_s"]
    finally:
        nb.close()
    return dict(per_step=res, to_solution=dict(rel_gradient_tol=solve_tol, results=sol))


def config3_leg(stream):
    """Configuration 3 at full size: the jittered 48x48x192 beam (2.5M tets) with 70 / 20 / 10 %
    Neo-Hookean / StVK / corotational tets and per-tet materials, compressed to J in [0.2, 1.5].
    Times the FP64 and FP32 (precision form) assembly with per-tet models, with and without
    simple_psd_fix(1e-3), and reports the FP32 outputs' relative error against the FP64 ones."""
    import torch

    from bartels_amd import meshgen
    from bartels_amd import tetmesh as tm

    p = meshgen.config3()
    dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
    mesh = tm.TetMesh(p.V.shape[0], p.E, dX, vol, V=p.V)
    dev = torch.device("cuda:0")
    q = torch.tensor(p.q, device=dev)
    params = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device=dev)
    par32 = params.to(torch.float32)
    ids = torch.tensor(p.model_ids, device=dev)
    nnz, nV, nT = mesh.nnz, mesh.nV, mesh.nT
    psd = tm.simple_psd_fix(1e-3)
    M = tm.MODEL_PER_TET
    H = torch.empty(nnz, dtype=torch.float64, device=dev)
    H32 = torch.empty(nnz, dtype=torch.float32, device=dev)
    g, g32 = (torch.empty(3 * nV, dtype=torch.float64, device=dev) for _ in range(2))
    e, e32 = (torch.empty(1, dtype=torch.float64, device=dev) for _ in range(2))

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

    def rel(a, b):
        return float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))

    out = dict(workload=f"config 3: jittered 48x48x192 beam, mixed NH / StVK / corotational, J in [0.2, 1.5]",
               tets=nT, nnz=nnz)
    for tag, cb in (("psd", psd), ("nopsd", None)):
        out[f"dq2_{tag}_fp64_ms"] = timed(lambda: mesh.dq2_device(M, q, params, H, callback=cb, model_ids=ids,
                                                                  stream=stream))
        out[f"dq2_{tag}_fp32_ms"] = timed(lambda: mesh.dq2_f32_device(M, q, par32, H32, callback=cb, model_ids=ids,
                                                                      stream=stream))
        out[f"dq2_{tag}_fp32_rel_error"] = rel(H32, H)
    out["dq_fp64_ms"] = timed(lambda: mesh.dq_device(M, q, params, g, model_ids=ids, stream=stream))
    out["dq_fp32_ms"] = timed(lambda: mesh.dq_f32_device(M, q, par32, g32, model_ids=ids, stream=stream))
    out["dq_fp32_rel_error"] = rel(g32, g)
    out["q_fp64_ms"] = timed(lambda: mesh.q_device(M, q, params, e, model_ids=ids, stream=stream))
    out["q_fp32_ms"] = timed(lambda: mesh.q_f32_device(M, q, par32, e32, model_ids=ids, stream=stream))
    out["q_fp32_rel_error"] = rel(e32, e)
    out["dq2_psd_fp64_mtets_s"] = nT / (out["dq2_psd_fp64_ms"] * 1e-3) / 1e6
    out["dq2_psd_fp32_mtets_s"] = nT / (out["dq2_psd_fp32_ms"] * 1e-3) / 1e6
    mesh.close()
    return out


def _event_timer(stream):
    import torch

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

    return timed


def muscle_leg(p, mesh, stream):
    """The muscle-fibre model (Neo-Hookean plus a/2 |F d|^2 per tet) on the same 6.07M-tet beam and
    deformation: dq2 with and without simple_psd_fix(1e-3), FP64, fibres along the beam axis tilted per
    tet (a = 1e5 Pa)"""
    import torch

    from bartels_amd import tetmesh as tm

    nT = p.E.shape[0]
    rng = np.random.default_rng(5)
    d = np.c_[np.ones(nT), 0.2 * rng.standard_normal(nT), 0.2 * rng.standard_normal(nT)]
    d /= np.linalg.norm(d, axis=1)[:, None]
    par = tm.muscle_params(d, p.params, np.full(nT, 1e5))
    dev = torch.device("cuda:0")
    q = torch.tensor(p.q, device=dev)
    dpar = torch.tensor(np.asfortranarray(par).ravel(order="F"), device=dev)
    H = torch.empty(mesh.nnz, dtype=torch.float64, device=dev)
    timed = _event_timer(stream)
    M = tm.NEOHOOKEAN_MUSCLE
    ms_psd = timed(lambda: mesh.dq2_device(M, q, dpar, H, callback=tm.simple_psd_fix(1e-3), stream=stream))
    ms = timed(lambda: mesh.dq2_device(M, q, dpar, H, stream=stream))
    del H
    return dict(workload="NEOHOOKEAN_MUSCLE dq2 on the 64x64x256 bend, fibres along x tilted per tet, a = 1e5 Pa",
                dq2_psd_ms=ms_psd, dq2_psd_mtets_s=nT / (ms_psd * 1e-3) / 1e6, dq2_ms=ms,
                dq2_mtets_s=nT / (ms * 1e-3) / 1e6)


def triangle_leg(stream, cpu=True, n=1536):
    """The triangle siblings at scale, inputs resident on the device: the 2-D plane-strain Neo-Hookean
    strip of n x n vertices (2 (n - 1)^2 triangles) under a smooth bend + stretch -- q, dq, dq2 with and
    without simple_psd_fix(1e-3) -- and the StVK shell Hessian on the same strip embedded in 3-D (flat
    rest state, a sine bend out of the plane; per-element rest normal N = e_z, the current normal n and its
    exact derivative dndq formed on the host, as the reference's callers hand them over).  Algorithmic
    bytes per element: E, dphidX, area, params, the vertices' q once and every CSR value written once.
    The oracle's CPU time (the reference algorithm) on a 224 x 224 sample of the strip sits beside it."""
    import torch

    from bartels_amd import meshgen
    from bartels_amd import tetmesh as tm
    from bartels_amd import trimesh as trm

    h = 1e-3
    V, E = meshgen.strip2d(n, n, h)
    nV, nT = V.shape[0], E.shape[0]
    L = (n - 1) * h
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    dev = torch.device("cuda:0")
    timed = _event_timer(stream)

    def col(a):
        return torch.tensor(np.asfortranarray(a).ravel(order="F"), device=dev)

    out = dict(workload=f"{n} x {n} vertex strip, {nT} triangles", triangles=nT, vertices=nV)
    # tri2d: plane-strain Neo-Hookean
    x = V.copy()
    x[:, 0] *= 1.02
    x[:, 1] += 0.05 * L * np.sin(np.pi * V[:, 0] / L)
    par = meshgen.model_params(np.zeros(nT, np.uint8), lam, mu)
    dX = trm.linear_tri2dmesh_dphi_dX(V, E)
    area = meshgen.triangle_areas(V, E)
    m2 = trm.TriMeshDevice(2, nV, E)
    q, dXd, vol, pd = torch.tensor(x.ravel(), device=dev), col(dX), torch.tensor(area, device=dev), col(par)
    Hd = torch.empty(m2.nnz, dtype=torch.float64, device=dev)
    g = torch.empty(2 * nV, dtype=torch.float64, device=dev)
    e = torch.empty(1, dtype=torch.float64, device=dev)
    psd = tm.simple_psd_fix(1e-3)
    b_in = 12 + 48 + 8 + 16 + 16.0 * nV / nT
    r2 = dict(nnz=m2.nnz)
    r2["q_ms"] = timed(lambda: m2.q_device(0, q, dXd, vol, pd, e, stream=stream))
    r2["dq_ms"] = timed(lambda: m2.dq_device(0, q, dXd, vol, pd, g, stream=stream))
    r2["dq2_ms"] = timed(lambda: m2.dq2_device(0, q, dXd, vol, pd, Hd, stream=stream))
    r2["dq2_psd_ms"] = timed(lambda: m2.dq2_device(0, q, dXd, vol, pd, Hd, callback=psd, stream=stream))
    bytes_dq2 = (b_in + 8.0 * m2.nnz / nT) * nT
    r2["dq2_psd_mtris_s"] = nT / (r2["dq2_psd_ms"] * 1e-3) / 1e6
    r2["algorithmic_bytes"] = dict(q=b_in * nT, dq=(b_in + 16.0 * nV / nT) * nT, dq2=bytes_dq2)
    r2["hbm_frac"] = dict(q=b_in * nT / (r2["q_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          dq=(b_in + 16.0 * nV / nT) * nT / (r2["dq_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          dq2=bytes_dq2 / (r2["dq2_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          dq2_psd=bytes_dq2 / (r2["dq2_psd_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    out["tri2d_neohookean"] = r2
    del Hd, g
    m2.close()
    # tri3d: StVK shell
    V3 = np.c_[V, np.zeros(nV)]
    x3 = V3.copy()
    x3[:, 2] = 0.05 * L * np.sin(np.pi * V[:, 0] / L)
    X0, X1, X2 = (x3[E[:, k]] for k in range(3))
    e1, e2 = X1 - X0, X2 - X0
    c = np.cross(e1, e2)
    cn = np.linalg.norm(c, axis=1)
    nrm = c / cn[:, None]
    Pn = (np.eye(3)[None] - nrm[:, :, None] * nrm[:, None, :]) / cn[:, None, None]

    def skew(v):
        z = np.zeros(v.shape[0])
        return np.stack([np.stack([z, -v[:, 2], v[:, 1]], 1), np.stack([v[:, 2], z, -v[:, 0]], 1),
                         np.stack([-v[:, 1], v[:, 0], z], 1)], 1)

    dc1, dc2 = -skew(e2), skew(e1)  # d(e1 x e2)/d x1, / d x2
    dcs = [-(dc1 + dc2), dc1, dc2]
    dndq = np.concatenate([Pn @ d_ for d_ in dcs], axis=2)  # nT x 3 x 9, column 3a + c
    dX3 = trm.linear_tri3dmesh_dphi_dX(V3, E)
    par3 = meshgen.model_params(np.ones(nT, np.uint8), lam, mu)  # StVK [lambda, mu]
    m3 = trm.TriMeshDevice(3, nV, E)
    Nd = col(np.tile([0.0, 0.0, 1.0], (nT, 1)))
    nd = col(nrm)
    Dd = col(dndq.reshape(nT, 27, order="F"))
    q3, dX3d, p3 = torch.tensor(x3.ravel(), device=dev), col(dX3), col(par3)
    H3 = torch.empty(m3.nnz, dtype=torch.float64, device=dev)
    r3 = dict(nnz=m3.nnz)
    r3["dq2_ms"] = timed(lambda: m3.stvk_shell_dq2_device(q3, dX3d, Nd, nd, Dd, vol, p3, H3, stream=stream))
    r3["dq2_psd_ms"] = timed(lambda: m3.stvk_shell_dq2_device(q3, dX3d, Nd, nd, Dd, vol, p3, H3, callback=psd,
                                                             stream=stream))
    b3 = (12 + 72 + 24 + 24 + 216 + 8 + 16 + 24.0 * nV / nT + 8.0 * m3.nnz / nT) * nT
    r3["algorithmic_bytes"] = b3
    r3["hbm_frac"] = dict(dq2=b3 / (r3["dq2_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          dq2_psd=b3 / (r3["dq2_psd_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    r3["dq2_psd_mtris_s"] = nT / (r3["dq2_psd_ms"] * 1e-3) / 1e6
    out["tri3d_stvk_shell"] = r3
    del H3
    m3.close()
    if cpu:
        from oracle import oracle as O

        O.load(fast=True)
        ns = 224
        Vs, Es = meshgen.strip2d(ns, ns, h)
        xs = Vs.copy()
        xs[:, 0] *= 1.02
        xs[:, 1] += 0.05 * (ns - 1) * h * np.sin(np.pi * Vs[:, 0] / ((ns - 1) * h))
        dXs = trm.linear_tri2dmesh_dphi_dX(Vs, Es)
        ps = meshgen.model_params(np.zeros(Es.shape[0], np.uint8), lam, mu)
        t0 = time.perf_counter()
        O.tri2dmesh_dq2(0, Vs.shape[0], Es, xs.ravel(), dXs, meshgen.triangle_areas(Vs, Es), ps, psd=True, psd_tol=1e-3)
        dt = time.perf_counter() - t0
        out["cpu_oracle_tri2d_dq2_psd"] = dict(triangles=int(Es.shape[0]), seconds=dt,
                                                mtris_s=Es.shape[0] / dt / 1e6, threads=1,
                                                what="the oracle's restatement of the reference (dense B^T H B, "
                                                     "12x12-padded PSD, sort-and-fold), one host thread")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="256,64,64", help="vertices along x,y,z")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--gather-lanes", type=int, default=0, help="assembly kernel shape: 1, 2, 4 or 8 lanes per vertex "
                    "(0: the library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-config3", action="store_true", help="skip the 48x48x192 mixed-model assembly leg")
    ap.add_argument("--no-siblings", action="store_true",
                    help="skip the muscle-fibre model and the triangle siblings (tri2d, tri3d shell) legs")
    ap.add_argument("--newton-steps", type=int, default=2,
                    help="Newton steps per solver variant in the Newton-step comparison (0: skip)")
    ap.add_argument("--pcg-max", type=int, default=20000)
    ap.add_argument("--quick", action="store_true", help="timed steps only (for counter collection)")
    ap.add_argument("--newton-only", action="store_true",
                    help="only the Newton-step comparison (for the kernel profile of the solvers)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (reference point for the HBM byte counters)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="three calls each of dq2 + PSD, dq2, dq and q (FP64), then the FP32 dq2 + PSD and dq2, "
                         "and a 1 GiB device copy, for rocprofv3 (no timing output)")
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

    if args.profile_kernels:
        par32 = params.to(torch.float32)
        H32 = torch.empty(nnz, dtype=torch.float32, device=dev)
        calls = [step, lambda: mesh.dq2_device(tm.NEOHOOKEAN, q, params, H, stream=stream),
                 lambda: mesh.dq_device(tm.NEOHOOKEAN, q, params, g, stream=stream),
                 lambda: mesh.q_device(tm.NEOHOOKEAN, q, params, e, stream=stream),
                 lambda: mesh.dq2_f32_device(tm.NEOHOOKEAN, q, par32, H32, callback=psd, stream=stream),
                 lambda: mesh.dq2_f32_device(tm.NEOHOOKEAN, q, par32, H32, stream=stream)]
        for fn in calls:
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profiled": ["dq2_psd", "dq2", "dq", "q", "dq2_f32_psd", "dq2_f32", "copy_1GiB"],
                          "calls_each": 3, "tets": nT, "nnz": nnz}), flush=True)
        return

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

    def tune(lanes=0, xcd=-1, order=0):
        assert lib.bh_mesh_set_tuning(mesh.handle, ctypes.byref(_lib.BhTuning(lanes, xcd, 0, 0, order))) == 0

    ms_nopsd, k_nopsd = timed_kernels(lambda: mesh.dq2_device(tm.NEOHOOKEAN, q, params, H, stream=stream))
    gather_nopsd = k_nopsd["hessian_gather"]
    ms_dq = timed(lambda: mesh.dq_device(tm.NEOHOOKEAN, q, params, g, stream=stream))
    ms_q = timed(lambda: mesh.q_device(tm.NEOHOOKEAN, q, params, e, stream=stream))
    # the other kernel schedules (gather lanes x XCD grouping x visiting order) and the FP32 precision
    # form, same workload; none changes a result bit
    variants = {}
    par32 = params.to(torch.float32)
    H32 = torch.empty(nnz, dtype=torch.float32, device=dev)
    for lanes in (1, 2, 4, 8):
        for xcd, order in ((0, 2), (1, 2), (0, 1)):
            tune(lanes, xcd, order)
            key = f"lanes{lanes}_xcd{xcd}_{'morton' if order == 1 else 'vertex'}"
            variants[key + "_psd"] = timed_kernels(step)
            variants[key + "_nopsd"] = timed_kernels(lambda: mesh.dq2_device(tm.NEOHOOKEAN, q, params, H, stream=stream))
            variants[key + "_fp32_psd"] = timed_kernels(
                lambda: mesh.dq2_f32_device(tm.NEOHOOKEAN, q, par32, H32, callback=psd, stream=stream))
            variants[key + "_fp32_nopsd"] = timed_kernels(
                lambda: mesh.dq2_f32_device(tm.NEOHOOKEAN, q, par32, H32, stream=stream))
    tune(0, 0, 3)  # back to the library default (or the requested shape)
    assert lib.bh_mesh_set_gather_lanes(mesh.handle, args.gather_lanes) == 0
    del H32

    B = algorithmic_bytes(nV, nT, nnz)
    gather_ms = kt["hessian_gather"]
    achieved = B["dq2"] * nT / (gather_ms * 1e-3) / 1e9
    traffic = load_traffic("hessian_gather_psd")
    achieved_call = B["dq2"] * nT / (ms_step * 1e-3) / 1e9
    # the PSD pass is bound by FP64 VALU issue: its FP64 operations per tet from rocprofv3 counters
    # (64 lanes per issued wave instruction, FMA counted as 2), divided by its live duration
    psd_flops = load_traffic("psd_pass_fp64_flops_per_tet")
    psd_ms = kt.get("psd")
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
            "kernel": "k_hessian_recq (assembly gather from the PSD records)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_tet": B["dq2"],
            "call_achieved": achieved_call,
            "call_frac": achieved_call / HBM_PEAK_GBS,
            "call_note": "the same algorithmic bytes over the whole call (PSD pass + gather)",
            "note": "the HBM-side kernel of the call; the PSD pass (the longer of the two) is FP64-VALU "
                    "bound and is measured against the FP64 peak in roofline_psd_pass",
        },
        "roofline_psd_pass": None if not (psd_flops and psd_ms) else {
            "kernel": "k_psd_fused (reduced 9x9 M, eigenvalues, clamp, records)",
            "bound": "fp64 valu",
            "flops_per_tet": psd_flops,
            "achieved": psd_flops * nT / (psd_ms * 1e-3) / 1e12,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": psd_flops * nT / (psd_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
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
    if not args.no_config3:
        out["extra"]["config3"] = config3_leg(stream)
    if not args.no_siblings:
        out["extra"]["muscle"] = muscle_leg(p, mesh, stream)
        out["extra"]["triangles"] = triangle_leg(stream, cpu=not args.no_cpu_baseline)
    if args.newton_steps > 0:
        del H, g
        nres = newton_steps(p, mesh, args.newton_steps, args.pcg_max)
        out["extra"]["newton"] = nres
        # the fastest variant within 1e-5 relative displacement of the FP64 judge, per step and to solution
        for key, table, tcol in (("newton_best", nres["per_step"], "s_per_step"),
                                 ("newton_best_to_solution", nres["to_solution"]["results"], "wall_s")):
            ok = {k: v for k, v in table.items() if k != JUDGE[0] and v.get("rel_displacement_vs_fp64_judge", 1) < 1e-5
                  and v["status"] in (0, 4) and (key == "newton_best" or v["converged"])}
            if ok:
                best = max(ok, key=lambda k: ok[k]["speedup_vs_fp64_jacobi"])
                out["extra"][key] = dict(solver=best, seconds=ok[best][tcol],
                                         fp64_jacobi_seconds=table["fp64_jacobi"][tcol],
                                         speedup_vs_fp64_jacobi=ok[best]["speedup_vs_fp64_jacobi"],
                                         rel_displacement_vs_fp64_judge=ok[best]["rel_displacement_vs_fp64_judge"],
                                         pcg_tol=1e-6, what=tcol)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, dX, vol, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Backward-Euler implicit dynamics on the device (the reference's velocity-form objective,
matlab/examples/test_tet_materials.m:97-112, solved in position form): every step is one
bh_newton_static_device call with the inertia term 1/(2 dt^2) (q - q_pred)^T M (q - q_pred),
q_pred = q_t + dt v_t, then v_{t+1} = (q_{t+1} - q_t) / dt.

The falling clamped beam with per-tet Young's modulus log-uniform in [1e5, 1e7] Pa (meshgen.config5):
every variant solves each Newton direction by PCG to a 1e-6 relative residual with the Jacobi
fallback on; --judge steps the FP64 judge (FP64 elements, block IC(0) with the coarse correction,
PCG to 1e-12) in lockstep and reports each step's displacement difference against it.

    python -m bartels_amd.dynamics --steps 50 --variant fp32_learned_2l --judge
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib, meshgen
from . import tetmesh as tm


class BackwardEuler:
    def __init__(self, p):
        import torch

        self.lib = _lib.load()
        self.p = p
        nV = p.V.shape[0]
        dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
        self.mesh = tm.TetMesh(nV, p.E, dX, vol, V=p.V)
        dev = torch.device("cuda:0")
        fixed = np.zeros(3 * nV, np.uint8)
        fixed[(3 * p.fixed[:, None] + np.arange(3)).ravel()] = 1
        self.dfixed = torch.tensor(fixed, device=dev)
        self.dparams = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device=dev)
        dens = torch.full((p.E.shape[0],), p.density, dtype=torch.float64, device=dev)
        self.mass = torch.empty(self.mesh.nnz, dtype=torch.float64, device=dev)
        _lib.check(self.lib.bh_linear_tetmesh_mass_matrix_device(self.mesh.handle, ctypes.c_void_p(dens.data_ptr()),
                                                                 ctypes.c_void_p(self.mass.data_ptr()), None),
                   "mass_matrix")
        self.f = torch.zeros(3 * nV, dtype=torch.float64, device=dev)
        _lib.check(self.lib.bh_linear_tetmesh_const_ext_force_device(
            self.mesh.handle, (ctypes.c_double * 3)(*(p.density * p.gravity)), ctypes.c_void_p(self.f.data_ptr()),
            None), "const_ext_force")
        self.P = ctypes.c_void_p()
        _lib.check(self.lib.bh_precond_create(self.mesh.handle, ctypes.byref(self.P)), "bh_precond_create")
        self.x = torch.tensor(p.V.ravel(), device=dev)
        self.v = torch.zeros_like(self.x)

    def step(self, dt, precond=0, precision=0, mgn=None, gtol=1e-6, max_it=10, pcg_tol=1e-6, pcg_max=20000,
             psd=True, coarse=0, fallback=1):
        import torch

        lib = self.lib
        qp = self.x + dt * self.v
        o = _lib.BhNewtonOpts()
        lib.bh_newton_default_opts(ctypes.byref(o))
        o.model = 0
        o.psd = _lib.BhPsd(1 if psd else 0, 1e-3)
        o.gradient_tol = gtol
        o.relative_tol = 1
        o.max_iterations = max_it - 1
        o.pcg_tol = pcg_tol
        o.pcg_max_iterations = pcg_max
        o.precond = precond
        o.coarse_correction = coarse
        o.jacobi_fallback = fallback
        o.precond_handle = self.P
        o.element_precision = precision
        if mgn is not None:
            o.mgn = mgn
        o.d_mass_vals = ctypes.c_void_p(self.mass.data_ptr())
        o.d_q_pred = ctypes.c_void_p(qp.data_ptr())
        o.inertia = 1.0 / (dt * dt)
        q = qp.clone()
        rep = _lib.BhNewtonReport()
        st = lib.bh_newton_static_device(self.mesh.handle, ctypes.byref(o), ctypes.c_void_p(self.dparams.data_ptr()),
                                         ctypes.c_void_p(self.dfixed.data_ptr()), ctypes.c_void_p(self.f.data_ptr()),
                                         ctypes.c_void_p(q.data_ptr()), ctypes.byref(rep), None,
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if st not in (_lib.BH_OK, _lib.BH_MAXIT):
            raise _lib.BartelsError(st, "bh_newton_static_device")
        self.v = (q - self.x) / dt
        self.x = q
        return rep

    def close(self):
        if self.P:
            self.lib.bh_precond_destroy(self.P)
            self.P = ctypes.c_void_p()


# name: (preconditioner, element precision, coarse correction, network form or None, shipped weights)
VARIANTS = {"fp64_jacobi": (0, 0, 0, None, False), "fp32_jacobi": (0, 1, 0, None, False),
            "fp32_bic0": (2, 1, 0, None, False), "fp32_sgs": (5, 1, 0, None, False),
            "fp64_jacobi2l": (7, 0, 0, None, False), "fp32_jacobi2l": (7, 1, 0, None, False),
            "fp32_bic0_2l": (2, 1, 1, None, False), "fp32_learned_2l": (4, 1, 1, "tied", True),
            "fp32_learned_untrained_2l": (4, 1, 1, "tied", False), "fp32_learned_sgs": (6, 1, 0, "sgs", False)}
# the FP64 judge: FP64 elements, block IC(0) with the coarse correction, PCG to 1e-12
JUDGE = (2, 0, 1, 1e-12)


def run(steps, variant, grid=(256, 64, 64), weights=None, judge=False, log=print):
    """`steps` backward-Euler steps of the configuration-5 beam with one solver variant (and, with judge,
    the FP64 judge in lockstep); returns the final positions, per-step rows and the variant's wall time"""
    import torch

    from . import learned

    p = meshgen.config5(grid=grid)
    be = BackwardEuler(p)
    bj = BackwardEuler(p) if judge else None
    precond, precision, coarse, form, shipped = VARIANTS[variant]
    net = None
    if form is not None:
        w = weights or (learned.DEFAULT_WEIGHTS if shipped else None)
        net = learned.Network(weights=w, form=form)
    X = torch.tensor(p.V.ravel(), device="cuda")
    rows = []
    wall = 0.0
    for k in range(steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rep = be.step(p.dt, precond=precond, precision=precision, mgn=net.h if net else None, coarse=coarse)
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t
        wall += dt_s
        row = dict(step=k, newton=rep.iterations, pcg=rep.pcg_iterations_total, fallbacks=rep.pcg_fallbacks,
                   pcg_failures=rep.pcg_failures, converged=bool(rep.converged), s=dt_s,
                   tip_z=float(be.x[-1].item() - p.V.ravel()[-1]))
        if bj is not None:
            t = time.perf_counter()
            rj = bj.step(p.dt, precond=JUDGE[0], precision=JUDGE[1], coarse=JUDGE[2], pcg_tol=JUDGE[3])
            torch.cuda.synchronize()
            u = float(torch.linalg.norm(bj.x - X))
            row.update(judge_s=time.perf_counter() - t, judge_newton=rj.iterations, judge_pcg=rj.pcg_iterations_total,
                       rel_displacement_vs_fp64_judge=float(torch.linalg.norm(be.x - bj.x)) / u if u > 0 else 0.0)
        rows.append(row)
        log(json.dumps(rows[-1]))
    x = be.x.cpu().numpy()
    be.close()
    if bj is not None:
        bj.close()
    if net:
        net.close()
    return x, rows, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--variant", default="fp64_jacobi", choices=sorted(VARIANTS))
    ap.add_argument("--grid", default="256,64,64")
    ap.add_argument("--save", default="")
    ap.add_argument("--judge", action="store_true", help="step the FP64 judge in lockstep and compare")
    a = ap.parse_args()
    grid = tuple(int(v) for v in a.grid.split(","))
    x, rows, wall = run(a.steps, a.variant, grid=grid, judge=a.judge, log=lambda s: print(s, flush=True))
    summary = dict(variant=a.variant, grid=grid, steps=a.steps, wall_s=wall, s_per_step=wall / max(1, a.steps),
                   pcg_total=sum(r["pcg"] for r in rows), fallbacks=sum(r["fallbacks"] for r in rows),
                   pcg_failures=sum(r["pcg_failures"] for r in rows),
                   all_converged=all(r["converged"] for r in rows))
    if a.judge:
        summary.update(judge_s_per_step=sum(r["judge_s"] for r in rows) / max(1, a.steps),
                       max_rel_displacement_vs_fp64_judge=max(r["rel_displacement_vs_fp64_judge"] for r in rows),
                       final_rel_displacement_vs_fp64_judge=rows[-1]["rel_displacement_vs_fp64_judge"] if rows else 0)
    print(json.dumps(summary))
    if a.save:
        np.save(a.save, x)


if __name__ == "__main__":
    main()

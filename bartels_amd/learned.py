"""Training and evaluation of the learned preconditioner (the mesh-graph network of
csrc/mgn_impl.h) on the Newton systems that the plain solver generates.

Training problems: jittered beams of 24x24x96 ... 40x40x160 vertices, per-tet Young's modulus
log-uniform in [1e5, 1e7] Pa, Poisson ratio uniform in [0.30, 0.48], 70 / 20 / 10 % Neo-Hookean /
StVK / corotational tets, x = 0 clamped, gravity plus a random traction of magnitude up to 5e4 Pa in
a random direction on the free end.  Each problem is solved by the precision-form Newton with
block symmetric Gauss-Seidel PCG to 1e-6 relative residual; every Hessian it assembles is handed
(on the device, through bh_newton_opts.on_system) to a few Adam steps of the network on probe
systems: the Newton direction the solver found (a smooth vector: the modes CG converges on last)
and standard normal vectors (a Hutchinson estimate of the relative Frobenius error of M).

Two forms (bh_mgn_set_form): "sgs" (default), the Eisenstat form M = (Dt + L) Dt^-1 (Dt + L)^T with
the network writing the block diagonal Dt (Dt = D, block symmetric Gauss-Seidel, at zero decoders),
solved with BH_PRECOND_LEARNED_SGS; and "factor", a general block lower-triangular factor
(BH_PRECOND_LEARNED).

    python -m bartels_amd.learned train --solves 120 --out weights.f32
    python -m bartels_amd.learned eval --weights weights.f32 --problems 10
"""
import argparse
import ctypes
import json
import math
import os
import time

import numpy as np

from . import _lib, meshgen
from . import tetmesh as tm

H = meshgen.H_DEFAULT
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mgn_weights.f32")


def random_problem(grid, seed):
    """a training / evaluation problem on an (nx, ny, nz) vertex grid"""
    nx, ny, nz = grid
    V, E = meshgen.beam_grid(nx, ny, nz, H)
    V = meshgen.jitter(V, E, H, 0.2, seed)
    nT = E.shape[0]
    u = meshgen.uniform(seed + 1, 3 * nT).reshape(3, nT)
    Ey = np.exp(np.log(1e5) + (np.log(1e7) - np.log(1e5)) * u[0])
    nu = 0.30 + 0.18 * u[1]
    ids = np.where(u[2] < 0.7, 0, np.where(u[2] < 0.9, 1, 2)).astype(np.uint8)
    lam, mu = meshgen.emu_to_lame(Ey, nu)
    params = meshgen.model_params(ids, lam, mu)
    w = meshgen.uniform(seed + 2, 4)
    mag = 5e4 * w[0]
    th, ph = 2.0 * math.pi * w[1], math.acos(2.0 * w[2] - 1.0)
    t = mag * np.array([math.sin(ph) * math.cos(th), math.sin(ph) * math.sin(th), math.cos(ph)])
    nV = V.shape[0]
    end = np.nonzero(np.abs(V[:, 0] - V[:, 0].max()) <= 1e-3)[0]
    f_trac = np.zeros((nV, 3))
    f_trac[end] = t * ((ny - 1) * (nz - 1) * H * H / end.size)  # end-face area shared by its vertices
    return meshgen.Problem(V=V, E=E, params=params, model_ids=ids, fixed=meshgen.fixed_vertices(V),
                           f_trac=f_trac.ravel(), density=1000.0, gravity=np.array([0.0, 0.0, -9.8]))


class Solver:
    """device state of one problem: mesh, preconditioner lists, parameters, loads"""

    def __init__(self, p):
        import torch

        self.lib = _lib.load()
        self.p = p
        nV = p.V.shape[0]
        dX, vol = tm.linear_tetmesh_dphi_dX(p.V, p.E)
        self.mesh = tm.TetMesh(nV, p.E, dX, vol, V=p.V)
        self.P = ctypes.c_void_p()
        _lib.check(self.lib.bh_precond_create(self.mesh.handle, ctypes.byref(self.P)), "bh_precond_create")
        dev = torch.device("cuda:0")
        fixed = np.zeros(3 * nV, np.uint8)
        fixed[(3 * p.fixed[:, None] + np.arange(3)).ravel()] = 1
        self.dfixed = torch.tensor(fixed, device=dev)
        self.dparams = torch.tensor(np.asfortranarray(p.params).ravel(order="F"), device=dev)
        self.dids = torch.tensor(p.model_ids, device=dev) if p.model_ids is not None else None
        f = torch.zeros(3 * nV, dtype=torch.float64, device=dev)
        _lib.check(self.lib.bh_linear_tetmesh_const_ext_force_device(
            self.mesh.handle, (ctypes.c_double * 3)(*(p.density * p.gravity)), ctypes.c_void_p(f.data_ptr()), None),
            "const_ext_force")
        if "f_trac" in p:
            f += torch.tensor(p.f_trac, device=dev)
        self.df = f
        self.q0 = p.V.ravel().copy()

    def newton(self, precond, mgn=None, precision=1, max_it=4, pcg_tol=1e-6, pcg_max=5000, on_system=None,
               q0=None, psd=True, model=None, coarse=0):
        import torch

        lib = self.lib
        o = _lib.BhNewtonOpts()
        lib.bh_newton_default_opts(ctypes.byref(o))
        if self.dids is not None:
            o.model = 255
            o.d_model_ids = ctypes.c_void_p(self.dids.data_ptr())
        else:
            o.model = 0 if model is None else model
        o.psd = _lib.BhPsd(1 if psd else 0, 1e-3)
        o.gradient_tol = 1e-6
        o.relative_tol = 1
        o.max_iterations = max_it - 1
        o.pcg_tol = pcg_tol
        o.pcg_max_iterations = pcg_max
        o.precond = precond
        o.precond_handle = self.P
        o.element_precision = precision
        o.jacobi_fallback = 1
        o.coarse_correction = coarse
        if mgn is not None:
            o.mgn = mgn
        cb = _lib.ON_SYSTEM(on_system) if on_system else _lib.ON_SYSTEM()
        o.on_system = cb
        dq = torch.tensor(self.q0 if q0 is None else q0, device="cuda:0")
        rep = _lib.BhNewtonReport()
        trace = np.zeros(4 * (max_it + 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = lib.bh_newton_static_device(self.mesh.handle, ctypes.byref(o), ctypes.c_void_p(self.dparams.data_ptr()),
                                         ctypes.c_void_p(self.dfixed.data_ptr()), ctypes.c_void_p(self.df.data_ptr()),
                                         ctypes.c_void_p(dq.data_ptr()), ctypes.byref(rep),
                                         trace.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if st not in (_lib.BH_OK, _lib.BH_MAXIT):
            raise _lib.BartelsError(st, "bh_newton_static_device")
        return dq, rep, trace.reshape(-1, 4)[:rep.iterations], wall

    def close(self):
        if self.P:
            self.lib.bh_precond_destroy(self.P)
            self.P = ctypes.c_void_p()


FORMS = {"factor": 0, "sgs": 1, "tied": 2}
PRECOND = {"factor": 4, "sgs": 6, "tied": 4}  # BH_PRECOND_LEARNED, BH_PRECOND_LEARNED_SGS, BH_PRECOND_LEARNED


class Network:
    def __init__(self, seed=1, weights=None, form="sgs", probe_solution=True, energy_loss=True):
        self.lib = _lib.load()
        self.h = ctypes.c_void_p()
        _lib.check(self.lib.bh_mgn_create(seed, ctypes.byref(self.h)), "bh_mgn_create")
        self.form = form
        self.precond = PRECOND[form]
        _lib.check(self.lib.bh_mgn_set_form(self.h, FORMS[form], 1 if probe_solution else 0,
                                            1 if (energy_loss and form == "sgs") else 0), "bh_mgn_set_form")
        n = ctypes.c_int64()
        self.lib.bh_mgn_num_params(self.h, ctypes.byref(n))
        self.n = n.value
        if weights is not None:
            self.load(weights)

    def params(self):
        w = np.empty(self.n, np.float32)
        _lib.check(self.lib.bh_mgn_get_params(self.h, w.ctypes.data_as(ctypes.c_void_p), None), "get_params")
        return w

    def save(self, path):
        self.params().tofile(path)

    def load(self, path):
        w = np.fromfile(path, dtype=np.float32)
        if w.size != self.n:
            raise ValueError(f"{path}: {w.size} parameters, the network has {self.n}")
        _lib.check(self.lib.bh_mgn_set_params(self.h, w.ctypes.data_as(ctypes.c_void_p)), "set_params")

    def close(self):
        if self.h:
            self.lib.bh_mgn_destroy(self.h)
            self.h = ctypes.c_void_p()


def train(solves, out, seed=0, steps=4, lr=1e-3, newton_its=4, probes=4, form="sgs", probe_solution=True,
          log=print, ny_range=(24, 41)):
    net = Network(seed=seed + 12345, form=form, probe_solution=probe_solution)
    rng = np.random.default_rng(seed)
    history = []
    t0 = time.perf_counter()
    for i in range(solves):
        ny = int(rng.integers(*ny_range))
        grid = (4 * ny, ny, ny)
        p = random_problem(grid, seed=1000 + 7 * i + seed)
        S = Solver(p)
        losses = []

        def on_system(user, it, v32, v64, b, x):
            for k in range(steps):
                L = ctypes.c_double()
                st = net.lib.bh_mgn_train_step_device(net.h, S.P, ctypes.c_void_p(v32), S.dfixed.data_ptr(),
                                                      ctypes.c_void_p(x if probe_solution else b), probes,
                                                      (seed * 1000003 + 7919 * i + 31 * it + k) & (2**64 - 1),
                                                      lr, ctypes.byref(L), None)
                if st == _lib.BH_NONFINITE:  # the step was dropped (parameters unchanged)
                    skipped.append(1)
                    continue
                _lib.check(st, "mgn_train_step")
                losses.append(L.value)

        skipped = []
        _, rep, tr, wall = S.newton(5, precision=1, max_it=newton_its, on_system=on_system)
        S.close()
        history.append(dict(solve=i, grid=grid, newton=rep.iterations, pcg=rep.pcg_iterations_total,
                            loss_first=losses[0] if losses else None, loss_last=losses[-1] if losses else None,
                            wall=wall))
        log(f"solve {i} grid {grid}: newton {rep.iterations}, pcg {rep.pcg_iterations_total}, "
            f"loss {losses[0] if losses else float('nan'):.4f} -> {losses[-1] if losses else float('nan'):.4f}, "
            f"dropped {len(skipped)}, {wall:.1f} s (total {time.perf_counter() - t0:.0f} s)")
    net.save(out)
    net.close()
    return history


def evaluate(weights, problems=10, seed=777, newton_its=3, pcg_max=20000, form="sgs", coarse=0, log=print,
             grid=(128, 32, 32)):
    """held-out problems (32x32x128 by default): PCG iterations and time to 1e-6 over the first Newton steps,
    block-Jacobi PCG, block symmetric Gauss-Seidel and the learned preconditioner (all in the precision
    form); coarse = 1 adds the two-level coarse correction to the learned factor (and evaluates two-level
    block Jacobi beside it).  Every solve reports its inner PCG failures (after the Jacobi fallback)."""
    net = Network(weights=weights, form=form)
    rows = []
    for i in range(problems):
        p = random_problem(grid, seed=90000 + seed + 13 * i)
        S = Solver(p)
        _, rj, _, wj = S.newton(0, precision=1, max_it=newton_its, pcg_max=pcg_max)
        _, rl, _, wl = S.newton(net.precond, mgn=net.h, precision=1, max_it=newton_its, pcg_max=pcg_max,
                                coarse=coarse)
        _, rb, _, wb = S.newton(5, precision=1, max_it=newton_its, pcg_max=pcg_max)
        row = dict(problem=i, jacobi_pcg=rj.pcg_iterations_total, learned_pcg=rl.pcg_iterations_total,
                   sgs_pcg=rb.pcg_iterations_total, jacobi_s=wj, learned_s=wl, sgs_s=wb,
                   fallbacks=rl.pcg_fallbacks, learned_failures=rl.pcg_failures, newton=rl.iterations,
                   learned_rel_gradient=rl.gradient_norm / rl.gradient_norm0 if rl.gradient_norm0 else 0.0)
        if coarse:
            _, r2, _, w2 = S.newton(7, precision=1, max_it=newton_its, pcg_max=pcg_max)
            row.update(jacobi2l_pcg=r2.pcg_iterations_total, jacobi2l_s=w2)
        S.close()
        rows.append(row)
        log(json.dumps(rows[-1]))
    net.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("train")
    a.add_argument("--solves", type=int, default=120)
    a.add_argument("--steps", type=int, default=4)
    a.add_argument("--lr", type=float, default=1e-3)
    a.add_argument("--probes", type=int, default=4)
    a.add_argument("--seed", type=int, default=0)
    a.add_argument("--form", choices=sorted(FORMS), default="sgs")
    a.add_argument("--probe", choices=["solution", "rhs"], default="solution",
                   help="probe 0: the Newton direction (solution) or the scaled right-hand side")
    a.add_argument("--out", default=DEFAULT_WEIGHTS)
    b = sub.add_parser("eval")
    b.add_argument("--weights", default=DEFAULT_WEIGHTS)
    b.add_argument("--problems", type=int, default=10)
    b.add_argument("--form", choices=sorted(FORMS), default="sgs")
    args = ap.parse_args()
    log = lambda s: print(s, flush=True)  # noqa: E731
    if args.cmd == "train":
        train(args.solves, args.out, seed=args.seed, steps=args.steps, lr=args.lr, probes=args.probes,
              form=args.form, probe_solution=args.probe == "solution", log=log)
    else:
        rows = evaluate(args.weights, args.problems, form=args.form, log=log)
        j = np.array([r["jacobi_pcg"] for r in rows], float)
        l = np.array([r["learned_pcg"] for r in rows], float)
        b = np.array([r["sgs_pcg"] for r in rows], float)
        log(json.dumps(dict(mean_jacobi_pcg=j.mean(), mean_learned_pcg=l.mean(), mean_sgs_pcg=b.mean(),
                            sgs_over_learned=(b / l).mean(), jacobi_over_learned=(j / l).mean())))


if __name__ == "__main__":
    main()

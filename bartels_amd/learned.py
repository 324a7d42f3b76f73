"""Training and evaluation of the learned preconditioner (the mesh-graph network of
csrc/mgn_impl.h) on the Newton systems that the plain solver generates.

Problems: jittered beams of 96x24x24 ... 160x40x40 vertices, per-tet Young's modulus log-uniform in
[1e5, 1e7] Pa, Poisson ratio uniform in [0.30, 0.48], 70 / 20 / 10 % Neo-Hookean / StVK / corotational
tets, x = 0 clamped, gravity plus a random traction of magnitude up to 5e4 Pa in a random direction on
the free end.  Each problem is solved by the precision-form Newton with block symmetric Gauss-Seidel
PCG to 1e-6 relative residual, and the FP32 Hessians it assembles (handed over on the device through
bh_newton_opts.on_system) are the training systems.

The shipped weights (mgn_weights.f32) are the tied factor form (bh_mgn_set_form form 2: block
symmetric Gauss-Seidel as a factor at zero decoders, solved with BH_PRECOND_LEARNED and the two-level
coarse correction) trained on the Kaporin condition number (loss 3) of L L^T against each Hessian:
150 problems, their first three Newton systems each (450 systems), 30 epochs in minibatches of 8
systems, Adam at 3e-4, 8 Hutchinson probes per system:

    python -m bartels_amd.learned fit --problems 150 --epochs 30 --out weights.f32
    python -m bartels_amd.learned eval --weights weights.f32 --problems 6

The older online trainer (`train`: a few Adam steps per system as the solver meets it, residual /
energy / teacher losses) stays for the other forms: "sgs", the Eisenstat form M = (Dt + L) Dt^-1
(Dt + L)^T with the network writing the block diagonal Dt (BH_PRECOND_LEARNED_SGS), and "factor", a
general block lower-triangular factor (block Jacobi at zero decoders).
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
# the form the shipped weights were trained for (the tied factor: block SGS as a factor at zero decoders)
DEFAULT_FORM = "tied"


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


LOSSES = {"residual": 0, "energy": 1, "teacher": 2, "kaporin": 3}


class Network:
    def __init__(self, seed=1, weights=None, form="sgs", probe_solution=True, energy_loss=True, loss=None):
        """loss: "residual", "energy" (the Eisenstat form) or "teacher" (the factor forms: the IC(0) factor
        of each training Hessian as the target); default "energy" for the Eisenstat form with energy_loss,
        else "residual" """
        self.lib = _lib.load()
        self.h = ctypes.c_void_p()
        _lib.check(self.lib.bh_mgn_create(seed, ctypes.byref(self.h)), "bh_mgn_create")
        self.form = form
        self.precond = PRECOND[form]
        if loss is None:
            loss = "energy" if (energy_loss and form == "sgs") else "residual"
        _lib.check(self.lib.bh_mgn_set_form(self.h, FORMS[form], 1 if probe_solution else 0, LOSSES[loss]),
                   "bh_mgn_set_form")
        n = ctypes.c_int64()
        self.lib.bh_mgn_num_params(self.h, ctypes.byref(n))
        self.n = n.value
        if weights is not None:
            self.load(weights)

    def set_inference(self, bf16=True):
        """the MLPs of inference: BF16 on the matrix cores (the default) or the FP32 training forward"""
        _lib.check(self.lib.bh_mgn_set_inference(self.h, 1 if bf16 else 0), "bh_mgn_set_inference")

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
          log=print, ny_range=(24, 41), loss=None, init=None):
    net = Network(seed=seed + 12345, form=form, probe_solution=probe_solution, loss=loss)
    if init is not None:
        net.load(init)
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


def collect_systems(problems, seed=0, newton_its=3, ny_range=(24, 29), log=print):
    """the training set: the FP32 Newton systems (Hessian values on the device) that the plain solver (block
    symmetric Gauss-Seidel PCG, precision form) meets on `problems` random beams; [(Solver, values)]"""
    import torch

    rng = np.random.default_rng(seed)
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library is linked against
    systems = []
    for i in range(problems):
        ny = int(rng.integers(*ny_range))
        S = Solver(random_problem((4 * ny, ny, ny), seed=2000 + 7 * i + seed))
        caught = []

        def on_system(user, it, v32, v64, b, x):
            t = torch.empty(S.mesh.nnz, dtype=torch.float32, device="cuda")
            hip.hipMemcpy(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(v32), ctypes.c_size_t(4 * S.mesh.nnz), 3)
            caught.append(t)

        S.newton(5, precision=1, max_it=newton_its, on_system=on_system)
        torch.cuda.synchronize()
        S.mesh.trim()  # training needs the pattern and the preconditioner lists only
        systems += [(S, t) for t in caught]
        log(f"problem {i}: grid {(4 * ny, ny, ny)}, {len(caught)} systems")
    return systems


def train_batched(systems, out, epochs=30, batch=6, lr=3e-4, probes=8, form="tied", loss="kaporin", seed=0,
                  log=print, init=None):
    """minibatch training over a fixed set of Newton systems: per system the loss gradient is accumulated
    (bh_mgn_train_step_device with lr = 0), one Adam step per `batch` systems, `epochs` passes in a fresh
    random order each"""
    lib = _lib.load()
    net = Network(seed=seed + 12345, form=form, probe_solution=False, loss=loss)
    if init is not None:
        net.load(init)
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        order = rng.permutation(len(systems))
        losses, dropped = [], 0
        for bi, idx in enumerate(order):
            S, vals = systems[idx]
            L = ctypes.c_double()
            st = lib.bh_mgn_train_step_device(net.h, S.P, ctypes.c_void_p(vals.data_ptr()),
                                              ctypes.c_void_p(S.dfixed.data_ptr()), None, probes,
                                              (seed * 1000003 + 7919 * ep + 31 * bi) & (2**64 - 1), 0.0,
                                              ctypes.byref(L), None)
            if st == _lib.BH_NONFINITE:
                dropped += 1
                continue
            _lib.check(st, "mgn_train_step")
            losses.append(L.value)
            if (bi + 1) % batch == 0 or bi + 1 == len(order):
                _lib.check(lib.bh_mgn_adam_step(net.h, lr, 0.9, 0.999, 1e-8, None), "bh_mgn_adam_step")
        history.append(dict(epoch=ep, loss=float(np.mean(losses)) if losses else None, dropped=dropped))
        mean = f"{history[-1]['loss']:.5f}" if losses else "n/a"
        log(f"epoch {ep}: mean loss {mean}, dropped {dropped}")
    net.save(out)
    net.close()
    return history


def evaluate(weights, problems=10, seed=777, newton_its=3, pcg_max=20000, form="sgs", coarse=0, log=print,
             grid=(128, 32, 32), twin=False, bf16=True):
    """held-out problems (32x32x128 by default): PCG iterations and time to 1e-6 over the first Newton steps,
    block-Jacobi PCG, block symmetric Gauss-Seidel and the learned preconditioner (all in the precision
    form); coarse = 1 adds the two-level coarse correction to the learned factor (and evaluates two-level
    block Jacobi beside it).  Every solve reports its inner PCG failures (after the Jacobi fallback).
    twin: also the same form with zero decoders (the untrained network); bf16: the network's inference
    precision (BF16 on the matrix cores, else FP32)."""
    net = Network(weights=weights, form=form)
    net.set_inference(bf16)
    net0 = Network(form=form) if twin else None
    if net0 is not None:
        net0.set_inference(bf16)
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
        if net0 is not None:
            _, rt, _, wt = S.newton(net0.precond, mgn=net0.h, precision=1, max_it=newton_its, pcg_max=pcg_max,
                                    coarse=coarse)
            row.update(twin_pcg=rt.pcg_iterations_total, twin_s=wt, twin_fallbacks=rt.pcg_fallbacks)
        if coarse:
            _, r2, _, w2 = S.newton(7, precision=1, max_it=newton_its, pcg_max=pcg_max)
            _, r3, _, w3 = S.newton(2, precision=1, max_it=newton_its, pcg_max=pcg_max, coarse=1)
            row.update(jacobi2l_pcg=r2.pcg_iterations_total, jacobi2l_s=w2, bic0_2l_pcg=r3.pcg_iterations_total,
                       bic0_2l_s=w3)
        S.close()
        rows.append(row)
        log(json.dumps(rows[-1]))
    net.close()
    if net0 is not None:
        net0.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    f = sub.add_parser("fit", help="minibatch training over a fixed set of Newton systems (the shipped recipe)")
    f.add_argument("--problems", type=int, default=150)
    f.add_argument("--newton-its", type=int, default=3)
    f.add_argument("--epochs", type=int, default=30)
    f.add_argument("--batch", type=int, default=8)
    f.add_argument("--lr", type=float, default=3e-4)
    f.add_argument("--probes", type=int, default=8)
    f.add_argument("--form", choices=sorted(FORMS), default=DEFAULT_FORM)
    f.add_argument("--loss", choices=["kaporin", "teacher", "residual"], default="kaporin")
    f.add_argument("--data-seed", type=int, default=1)
    f.add_argument("--seed", type=int, default=0)
    f.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = sub.add_parser("train", help="online training, a few Adam steps per system as the solver meets it")
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
    b.add_argument("--problems", type=int, default=6)
    b.add_argument("--form", choices=sorted(FORMS), default=DEFAULT_FORM)
    b.add_argument("--coarse", type=int, default=1)
    b.add_argument("--seed", type=int, default=555)
    b.add_argument("--twin", action="store_true", help="also the untrained network (zero decoders)")
    b.add_argument("--fp32", action="store_true", help="FP32 inference instead of BF16 on the matrix cores")
    args = ap.parse_args()
    log = lambda s: print(s, flush=True)  # noqa: E731
    if args.cmd == "fit":
        systems = collect_systems(args.problems, seed=args.data_seed, newton_its=args.newton_its, ny_range=(24, 41),
                                  log=log)
        train_batched(systems, args.out, epochs=args.epochs, batch=args.batch, lr=args.lr, probes=args.probes,
                      form=args.form, loss=args.loss, seed=args.seed, log=log)
        for S, _ in systems:
            S.close()
    elif args.cmd == "train":
        train(args.solves, args.out, seed=args.seed, steps=args.steps, lr=args.lr, probes=args.probes,
              form=args.form, probe_solution=args.probe == "solution", log=log)
    else:
        rows = evaluate(args.weights, args.problems, form=args.form, coarse=args.coarse, seed=args.seed,
                        newton_its=2, log=log, twin=args.twin, bf16=not args.fp32)
        keys = ["jacobi_pcg", "sgs_pcg", "learned_pcg", "learned_s", "jacobi_s", "sgs_s"] + (
            ["jacobi2l_pcg", "bic0_2l_pcg", "jacobi2l_s", "bic0_2l_s"] if args.coarse else []) + (
            ["twin_pcg", "twin_s"] if args.twin else [])
        summary = {"problems": len(rows), "inference": "fp32" if args.fp32 else "bf16"}
        summary.update({"mean_" + k: float(np.mean([r[k] for r in rows])) for k in keys})
        summary["learned_fallbacks"] = int(sum(r["fallbacks"] for r in rows))
        summary["learned_failures"] = int(sum(r["learned_failures"] for r in rows))
        if args.twin:
            summary["learned_fewer_than_twin"] = int(sum(r["learned_pcg"] < r["twin_pcg"] for r in rows))
        if args.coarse:
            summary["learned_fewer_than_jacobi2l"] = int(sum(r["learned_pcg"] < r["jacobi2l_pcg"] for r in rows))
            summary["learned_faster_than_jacobi2l"] = int(sum(r["learned_s"] < r["jacobi2l_s"] for r in rows))
        log(json.dumps(summary))


if __name__ == "__main__":
    main()

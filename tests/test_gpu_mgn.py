"""The learned preconditioner's hand-written HIP kernels (bartels_amd/csrc/mgn_impl.h) against the
FP64 torch restatement in tests/mgn_reference.py: features, forward, loss, every parameter
gradient, the factor written into the preconditioner, and a short training run."""
import ctypes

import numpy as np
import pytest
import torch

import mgn_reference as R
from bartels_amd import meshgen

pytestmark = pytest.mark.gpu


def d2h(ptr, n, dtype):
    out = np.empty(n, dtype)
    if n:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipDeviceSynchronize()
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes),
                             2) == 0
    return out


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def system():
    from bartels_amd import _lib
    from bartels_amd import tetmesh as tm

    lib = _lib.load()
    V, E = meshgen.beam_grid(8, 5, 5, h=0.025)
    rng = np.random.default_rng(3)
    nV = V.shape[0]
    lam, mu = meshgen.emu_to_lame(2e6, 0.45)
    params = meshgen.model_params(np.zeros(E.shape[0], np.uint8), lam, mu)
    q = (V + 0.002 * rng.standard_normal(V.shape)).ravel()
    dX, vol = tm.linear_tetmesh_dphi_dX(V, E)
    mesh = tm.TetMesh(nV, E, dX, vol)
    dq = torch.tensor(q, device="cuda")
    dpar = torch.tensor(np.asfortranarray(params).ravel(order="F"), device="cuda", dtype=torch.float32)
    vals = torch.empty(mesh.nnz, dtype=torch.float32, device="cuda")
    mesh.dq2_f32_device(0, dq, dpar, vals, callback=tm.simple_psd_fix(1e-3))
    fixed = np.zeros(3 * nV, np.uint8)
    fixed[(3 * meshgen.fixed_vertices(V)[:, None] + np.arange(3)).ravel()] = 1
    P = ctypes.c_void_p()
    assert lib.bh_precond_create(mesh.handle, ctypes.byref(P)) == 0
    ptrs = [ctypes.c_void_p() for _ in range(5)]
    assert lib.bh_precond_lists_device(P, *[ctypes.byref(p) for p in ptrs]) == 0
    nc, nl = ctypes.c_int(), ctypes.c_int64()
    lib.bh_precond_info(P, ctypes.byref(nc), ctypes.byref(nl))
    lptr = d2h(ptrs[0].value, nV + 1, np.int32)
    lcol = d2h(ptrs[1].value, nl.value, np.int32)
    lslot = d2h(ptrs[2].value, nl.value, np.int32)
    dslot = d2h(ptrs[3].value, nV, np.int32)
    bp, bc, nb = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int64()
    lib.bh_mesh_block_pattern_device(mesh.handle, ctypes.byref(bp), ctypes.byref(bc), ctypes.byref(nb))
    bptr = d2h(bp.value, nV + 1, np.int32)
    bcol = d2h(bc.value, nb.value, np.int32)
    g = R.Graph(nV, bptr, bcol, lptr, lcol, lslot, dslot)
    hvals = vals.cpu().numpy().astype(np.float64)
    # a system H x = b of the masked Hessian: x random on the free DOFs, b = H x
    feats = R.features(g, hvals, fixed)
    H = np.zeros((3 * nV, 3 * nV))
    fx = fixed.astype(bool)
    for v in range(nV):
        for slot in range(bptr[v], bptr[v + 1]):
            n = bcol[slot]
            H[3 * v:3 * v + 3, 3 * n:3 * n + 3] = R.block(hvals, bptr, v, slot)
    H[fx, :] = 0.0
    H[:, fx] = 0.0
    H[fx, fx] = 1.0
    x = rng.standard_normal(3 * nV)
    x[fx] = 0.0
    b = H @ x
    yield dict(lib=lib, mesh=mesh, P=P, g=g, vals=vals, hvals=hvals, fixed=fixed,
               dfixed=torch.tensor(fixed, device="cuda"), feats=feats, H=H, x=x, b=b,
               db=torch.tensor(b, device="cuda"), dx=torch.tensor(x, device="cuda"))
    lib.bh_precond_destroy(P)


def make_mgn(lib, seed=7):
    M = ctypes.c_void_p()
    assert lib.bh_mgn_create(seed, ctypes.byref(M)) == 0
    n = ctypes.c_int64()
    lib.bh_mgn_num_params(M, ctypes.byref(n))
    return M, n.value


def _gradient_check(S, tied, tied_factor=False):
    lib = S["lib"]
    M, n = make_mgn(lib)
    try:
        assert lib.bh_mgn_set_form(M, 2 if tied_factor else (1 if tied else 0), 0, 0) == 0
        lay, ntot = R.layout()
        assert n == ntot
        rng = np.random.default_rng(1)
        w = np.empty(n, np.float32)
        assert lib.bh_mgn_get_params(M, w.ctypes.data_as(ctypes.c_void_p), None) == 0
        # the library's initialisation: zero decoders (block Jacobi at start)
        for k in ("Wdv", "bdv", "Wdo", "bdo"):
            o, sh = lay[k]
            assert not np.any(w[o:o + int(np.prod(sh))])
        # random decoders (and biases) so every gradient path carries signal
        for k, (o, sh) in lay.items():
            if k in ("Wdv", "Wdo"):
                w[o:o + int(np.prod(sh))] = 0.05 * rng.standard_normal(int(np.prod(sh)))
            elif k.endswith("b1") or k.endswith("b2") or k in ("bdv", "bdo"):
                w[o:o + int(np.prod(sh))] = 0.02 * rng.standard_normal(int(np.prod(sh)))
        assert lib.bh_mgn_set_params(M, w.ctypes.data_as(ctypes.c_void_p)) == 0
        loss = ctypes.c_double()
        st = lib.bh_mgn_loss_device(M, S["P"], vp(S["vals"]), vp(S["dfixed"]), vp(S["db"]), vp(S["dx"]), 1,
                                    ctypes.byref(loss), None)
        assert st == 0
        grad = np.empty(n, np.float32)
        lib.bh_mgn_get_params(M, None, grad.ctypes.data_as(ctypes.c_void_p))
        # features and decoder outputs of the same forward
        bufs = [ctypes.c_void_p() for _ in range(4)]
        lib.bh_mgn_debug_buffers(M, *[ctypes.byref(p) for p in bufs])
        g = S["g"]
        s_gpu = d2h(bufs[0].value, g.nV, np.float32)
        d_gpu = d2h(bufs[2].value, 6 * g.nV, np.float32).reshape(-1, 6)
        O_gpu = d2h(bufs[3].value, 9 * g.nP, np.float32).reshape(-1, 9)
        # reference
        Pt = {k: v.clone().requires_grad_(True) for k, v in R.unflatten(torch.tensor(w, dtype=torch.float64)).items()}
        L, d_ref, O_ref = R.loss(Pt, g, S["feats"], S["x"], S["b"], tied=tied, tied_factor=tied_factor)
        L.backward()
        g_ref = np.zeros(n)
        for k, (o, sh) in lay.items():
            if Pt[k].grad is not None:  # (the pair decoder is unused in the tied form)
                g_ref[o:o + int(np.prod(sh))] = Pt[k].grad.numpy().ravel()
        rs = np.abs(s_gpu - S["feats"][0]).max() / np.abs(S["feats"][0]).max()
        rd = np.abs(d_gpu - d_ref.detach().numpy()).max() / max(np.abs(d_ref.detach().numpy()).max(), 1e-30)
        rO = np.abs(O_gpu - O_ref.detach().numpy()).max() / max(np.abs(O_ref.detach().numpy()).max(), 1e-30)
        rl = abs(loss.value - L.item()) / L.item()
        print(f"scales {rs:.1e}, decoder node {rd:.1e}, pair {rO:.1e}, loss {loss.value:.6f} vs {L.item():.6f} ({rl:.1e})")
        assert rs < 1e-5 and rd < 1e-4 and rO < 1e-4 and rl < 1e-4
        worst = 0.0
        for k, (o, sh) in lay.items():
            a, r = grad[o:o + int(np.prod(sh))], g_ref[o:o + int(np.prod(sh))]
            if tied and k in ("Wdo", "bdo"):  # the pair decoder does not act in the tied form
                assert not np.any(a) and not np.any(r)
                continue
            e = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            worst = max(worst, e)
            assert e < 2e-3, (k, e, np.linalg.norm(r))
        print(f"worst per-tensor gradient error {worst:.1e} over {len(lay)} tensors")
    finally:
        lib.bh_mgn_destroy(M)


def test_mgn_forward_loss_and_gradient_match_reference(system):
    """general factor: features, decoder outputs, loss and every parameter gradient against FP64 torch"""
    _gradient_check(system, tied=False)


def test_mgn_tied_form_gradient_matches_reference(system):
    """Eisenstat (symmetric Gauss-Seidel) form: off-diagonal blocks Hn_vn Lhat_nn^{-T}, the gradient through
    them into the node decoder, against FP64 torch"""
    _gradient_check(system, tied=True)


def test_mgn_tied_factor_gradient_matches_reference(system):
    """tied factor form: off-diagonal blocks Hn_vn Lhat_nn^{-T} plus the pair decoder's output (block
    symmetric Gauss-Seidel at zero decoders), the gradient into both decoders, against FP64 torch"""
    _gradient_check(system, tied=False, tied_factor=True)


def test_mgn_energy_loss_gradient_matches_reference(system):
    """Eisenstat form with the energy-ratio loss (log x^T M x - log x^T H x)^2: loss and every parameter
    gradient against FP64 torch"""
    S = system
    lib = S["lib"]
    M, n = make_mgn(lib, seed=5)
    try:
        assert lib.bh_mgn_set_form(M, 1, 0, 1) == 0
        lay, _ = R.layout()
        rng = np.random.default_rng(4)
        w = np.empty(n, np.float32)
        lib.bh_mgn_get_params(M, w.ctypes.data_as(ctypes.c_void_p), None)
        for k, (o, sh) in lay.items():
            if k == "Wdv":
                w[o:o + int(np.prod(sh))] = 0.05 * rng.standard_normal(int(np.prod(sh)))
            elif k.endswith("b1") or k.endswith("b2") or k == "bdv":
                w[o:o + int(np.prod(sh))] = 0.02 * rng.standard_normal(int(np.prod(sh)))
        assert lib.bh_mgn_set_params(M, w.ctypes.data_as(ctypes.c_void_p)) == 0
        xs, bs = [S["x"]], [S["b"]]
        loss = ctypes.c_double()
        assert lib.bh_mgn_loss_device(M, S["P"], vp(S["vals"]), vp(S["dfixed"]), vp(S["db"]), vp(S["dx"]), 1,
                                      ctypes.byref(loss), None) == 0
        grad = np.empty(n, np.float32)
        lib.bh_mgn_get_params(M, None, grad.ctypes.data_as(ctypes.c_void_p))
        Pt = {k: v.clone().requires_grad_(True) for k, v in R.unflatten(torch.tensor(w, dtype=torch.float64)).items()}
        L, _ = R.energy_loss(Pt, S["g"], S["feats"], xs, bs)
        L.backward()
        rl = abs(loss.value - L.item()) / max(L.item(), 1e-30)
        print(f"energy loss {loss.value:.6e} vs {L.item():.6e} ({rl:.1e})")
        assert rl < 1e-4
        worst = 0.0
        for k, (o, sh) in lay.items():
            a = grad[o:o + int(np.prod(sh))]
            r = Pt[k].grad.numpy().ravel() if Pt[k].grad is not None else np.zeros_like(a, dtype=np.float64)
            if k in ("Wdo", "bdo"):
                assert not np.any(a) and not np.any(r)
                continue
            e = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            worst = max(worst, e)
            assert e < 2e-3, (k, e, np.linalg.norm(r))
        print(f"worst per-tensor gradient error {worst:.1e}")
    finally:
        lib.bh_mgn_destroy(M)


def test_mgn_factor_inference_matches_reference(system):
    """the factor the network writes into the preconditioner: (L L^T)^{-1} r by the colour sweeps equals
    a dense solve with the reference factor; at initialisation this is block Jacobi"""
    S = system
    lib = S["lib"]
    M, n = make_mgn(lib, seed=11)
    try:
        lay, _ = R.layout()
        rng = np.random.default_rng(2)
        w = np.empty(n, np.float32)
        lib.bh_mgn_get_params(M, w.ctypes.data_as(ctypes.c_void_p), None)
        for k in ("Wdv", "Wdo"):
            o, sh = lay[k]
            w[o:o + int(np.prod(sh))] = 0.02 * rng.standard_normal(int(np.prod(sh)))
        lib.bh_mgn_set_params(M, w.ctypes.data_as(ctypes.c_void_p))
        assert lib.bh_mgn_set_inference(M, 0) == 0  # the FP32 forward (BF16: the test below)
        assert lib.bh_mgn_factor_device(M, S["P"], vp(S["vals"]), vp(S["dfixed"]), None) == 0
        g = S["g"]
        s, C, xn, fp = (torch.as_tensor(a) for a in S["feats"])
        d, O = R.forward(R.unflatten(torch.tensor(w, dtype=torch.float64)), g, xn, fp)
        Lvv, Lk = R.factor(C, d, O)
        Ld = R.dense_factor(g, S["feats"][0], Lvv.detach().numpy(), Lk.detach().numpy(), S["fixed"])
        r = rng.standard_normal(3 * g.nV)
        z_ref = np.linalg.solve(Ld @ Ld.T, r)
        dr = torch.tensor(r, device="cuda")
        dz = torch.empty_like(dr)
        assert lib.bh_precond_apply_device(S["P"], vp(dr), vp(dz), None) == 0
        err = np.linalg.norm(dz.cpu().numpy() - z_ref) / np.linalg.norm(z_ref)
        print(f"learned factor application vs dense solve: {err:.1e}")
        assert err < 1e-3
    finally:
        lib.bh_mgn_destroy(M)


def _bf16_stage_check(S, M, n):
    """every MLP of the last forward against the FP64 torch restatement fed with the kernel's own FP32
    inputs of that stage, rounded to BF16 as the kernel rounds them: the remaining differences are
    the FP32 accumulation order and the rare hidden activation it moves across a BF16 rounding
    boundary"""
    lib, g = S["lib"], S["g"]
    w = np.empty(n, np.float32)
    lib.bh_mgn_get_params(M, w.ctypes.data_as(ctypes.c_void_p), None)
    Pw = R.unflatten(torch.tensor(w, dtype=torch.float64))
    nV, nP = g.nV, g.nP

    def lat(layer):
        p = [ctypes.c_void_p() for _ in range(3)]
        assert lib.bh_mgn_debug_latents(M, layer, *[ctypes.byref(x) for x in p]) == 0
        wh, we = (8, 10) if layer < 0 else (32, 32)
        h = torch.tensor(d2h(p[0].value, wh * nV, np.float32).reshape(nV, wh), dtype=torch.float64)
        e = torch.tensor(d2h(p[1].value, we * nP, np.float32).reshape(nP, we), dtype=torch.float64)
        m = (torch.tensor(d2h(p[2].value, 32 * nV, np.float32).reshape(nV, 32), dtype=torch.float64)
             if p[2].value else None)
        return h, e, m

    pv = torch.as_tensor(g.pv)
    pn = torch.as_tensor(g.lcol.astype(np.int64))
    xn, fp, _ = lat(-1)
    h, e, _ = lat(0)
    worst = {}

    def check(name, got, want):
        err = (torch.linalg.norm(got - want) / torch.linalg.norm(want)).item()
        worst[name] = err
        assert err < 1e-4, (name, err)

    check("node encoder", h, R.mlp(Pw, "nenc", xn, True))
    check("pair encoder", e, R.mlp(Pw, "penc", fp, True))
    for l in range(R.LAYERS):
        h1, e1, _ = lat(l + 1)
        _, _, m = lat(l)
        check(f"edge {l}", e1, e + R.mlp(Pw, f"edge{l}", torch.cat([e, h[pv], h[pn]], dim=1), True))
        check(f"messages {l}", m, torch.zeros_like(h).index_add(0, pv, e1).index_add(0, pn, e1))
        check(f"node {l}", h1, h + R.mlp(Pw, f"node{l}", torch.cat([h, m], dim=1), True))
        h, e = h1, e1
    return worst


def test_mgn_bf16_inference_matches_reference(system):
    """BF16 inference on the matrix cores (the default of bh_mgn_factor_device) with the shipped trained
    weights: each MLP against the BF16-rounding restatement on the kernel's own inputs; the decoder
    outputs of the whole network against the unrounded FP64 restatement (the size of the BF16 effect,
    reported) and the FP32 forward against the same (tight); the factor written from the BF16 outputs
    against a dense solve"""
    import os

    import bartels_amd

    S = system
    lib = S["lib"]
    M, n = make_mgn(lib, seed=5)
    try:
        w = np.fromfile(os.path.join(os.path.dirname(bartels_amd.__file__), "mgn_weights.f32"), np.float32)
        assert w.size == n
        lib.bh_mgn_set_params(M, w.ctypes.data_as(ctypes.c_void_p))
        assert lib.bh_mgn_set_form(M, 2, 0, 3) == 0  # the tied factor form the weights were trained in
        g = S["g"]
        outs = {}
        for prec in (0, 1):
            assert lib.bh_mgn_set_inference(M, prec) == 0
            assert lib.bh_mgn_factor_device(M, S["P"], vp(S["vals"]), vp(S["dfixed"]), None) == 0
            bufs = [ctypes.c_void_p() for _ in range(4)]
            lib.bh_mgn_debug_buffers(M, *[ctypes.byref(p) for p in bufs])
            outs[prec] = (d2h(bufs[2].value, 6 * g.nV, np.float32).reshape(-1, 6),
                          d2h(bufs[3].value, 9 * g.nP, np.float32).reshape(-1, 9))
        # the BF16 forward is the last one: its stages
        worst = _bf16_stage_check(S, M, n)
        print("BF16 MLP stages vs rounding restatement: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))
        s, C, xn, fp = (torch.as_tensor(a) for a in S["feats"])
        Pw = R.unflatten(torch.tensor(w, dtype=torch.float64))
        d, O = R.forward(Pw, g, xn, fp)
        _, Lk = R.factor(C, d, O, g, fp, add=True)  # the O buffer ends as Hn_vn Lhat_nn^-T + O_k
        d_f, O_f = d.detach().numpy(), Lk.detach().numpy().reshape(-1, 9)

        def rel(a, b):
            return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)

        (d_b, O_b), (d_32, O_32) = outs[1], outs[0]
        e_b = max(rel(d_b, d_f), rel(O_b, O_f))
        e_32 = max(rel(d_32, d_f), rel(O_32, O_f))
        print(f"decoder outputs vs unrounded FP64 restatement: BF16 inference {e_b:.1e}, FP32 inference {e_32:.1e}")
        assert e_32 < 1e-4 and e_b < 0.1
        # the factor written from the BF16 outputs (the device's own d and tied pair blocks)
        T = torch.zeros(g.nV, 3, 3, dtype=torch.float64)
        dd = torch.tensor(d_b, dtype=torch.float64)
        T[:, 0, 0], T[:, 1, 0], T[:, 1, 1] = torch.exp(dd[:, 0]), dd[:, 1], torch.exp(dd[:, 2])
        T[:, 2, 0], T[:, 2, 1], T[:, 2, 2] = dd[:, 3], dd[:, 4], torch.exp(dd[:, 5])
        Lvv = torch.bmm(C, T).numpy()
        Ld = R.dense_factor(g, S["feats"][0], Lvv, O_b.astype(np.float64).reshape(-1, 3, 3), S["fixed"])
        rng = np.random.default_rng(4)
        r = rng.standard_normal(3 * g.nV)
        z_ref = np.linalg.solve(Ld @ Ld.T, r)
        dr = torch.tensor(r, device="cuda")
        dz = torch.empty_like(dr)
        assert lib.bh_precond_apply_device(S["P"], vp(dr), vp(dz), None) == 0
        err = np.linalg.norm(dz.cpu().numpy() - z_ref) / np.linalg.norm(z_ref)
        print(f"BF16 learned factor application vs dense solve: {err:.1e}")
        assert err < 1e-3
    finally:
        lib.bh_mgn_destroy(M)


def test_mgn_training_reduces_loss(system):
    """Adam on one Newton system: the loss falls well below its block-Jacobi starting value"""
    S = system
    lib = S["lib"]
    M, n = make_mgn(lib, seed=3)
    try:
        losses = []
        for it in range(60):
            loss = ctypes.c_double()
            assert lib.bh_mgn_loss_device(M, S["P"], vp(S["vals"]), vp(S["dfixed"]), vp(S["db"]), vp(S["dx"]), 1,
                                          ctypes.byref(loss), None) == 0
            losses.append(loss.value)
            assert lib.bh_mgn_adam_step(M, 3e-3, 0.9, 0.999, 1e-8, None) == 0
        print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        assert np.isfinite(losses).all()
        assert losses[-1] < 0.7 * losses[0]
    finally:
        lib.bh_mgn_destroy(M)


def test_mgn_probe_training_step_reduces_frobenius_loss(system):
    """bh_mgn_train_step_device (probe systems of one Hessian, right-hand side first) lowers the
    relative Frobenius error of L L^T from its block-Jacobi value"""
    S = system
    lib = S["lib"]
    M, n = make_mgn(lib, seed=9)
    try:
        losses = []
        for it in range(80):
            loss = ctypes.c_double()
            assert lib.bh_mgn_train_step_device(M, S["P"], vp(S["vals"]), vp(S["dfixed"]), vp(S["db"]), 4, 100 + it,
                                                3e-3, ctypes.byref(loss), None) == 0
            losses.append(loss.value)
        print(f"probe loss {losses[0]:.4f} -> {losses[-1]:.4f} (min {min(losses):.4f})")
        assert np.isfinite(losses).all()
        assert 0.0 < losses[0] < 10.0
        assert np.mean(losses[-10:]) < 0.8 * losses[0]
    finally:
        lib.bh_mgn_destroy(M)

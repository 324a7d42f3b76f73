"""FP64 torch (CPU) restatement of the learned preconditioner's forward pass and training loss, used
by the GPU tests to check the hand-written HIP forward, backward and factor kernels
(bartels_amd/csrc/mgn_impl.h).  Test infrastructure only.

Conventions (same as the kernels): width 32, three message-passing layers, MLP = Linear -> ReLU ->
Linear with row-major [in][out] weights, a flat parameter vector in the order node encoder, pair
encoder, (edge MLP, node MLP) x 3, node decoder, pair decoder.
"""
import numpy as np
import torch

W = 32
LAYERS = 3


def layout():
    """name -> (offset, shape) of the flat parameter vector"""
    out, o = {}, 0

    def add(name, shape):
        nonlocal o
        out[name] = (o, shape)
        o += int(np.prod(shape))

    def mlp(prefix, fin):
        add(prefix + ".W1", (fin, W))
        add(prefix + ".b1", (W,))
        add(prefix + ".W2", (W, W))
        add(prefix + ".b2", (W,))

    mlp("nenc", 8)
    mlp("penc", 10)
    for l in range(LAYERS):
        mlp(f"edge{l}", 3 * W)
        mlp(f"node{l}", 2 * W)
    add("Wdv", (W, 6))
    add("bdv", (6,))
    add("Wdo", (W, 9))
    add("bdo", (9,))
    return out, o


def unflatten(flat):
    lay, n = layout()
    assert flat.numel() == n, (flat.numel(), n)
    return {k: flat[o:o + int(np.prod(sh))].view(*sh) for k, (o, sh) in lay.items()}


class Graph:
    """vertices and lower pairs of the colour order, from the preconditioner's lists"""

    def __init__(self, nV, bptr, bcol, lptr, lcol, lslot, dslot):
        self.nV = nV
        self.bptr, self.bcol = bptr, bcol
        self.lptr, self.lcol, self.lslot, self.dslot = lptr, lcol, lslot, dslot
        self.pv = np.repeat(np.arange(nV), np.diff(lptr))  # higher vertex of each pair
        self.nP = lcol.size


def block(vals, bptr, v, slot):
    """3x3 block at pattern index slot of block row v (row-major)"""
    b0, nb = bptr[v], bptr[v + 1] - bptr[v]
    r = slot - b0
    base = 9 * b0 + 3 * r
    return np.array([[vals[base + 3 * l * nb + k] for k in range(3)] for l in range(3)])


def features(g, vals, fixed):
    """s, C (nV x 3 x 3 lower), node features (nV x 8), pair features (nP x 10), all float64"""
    nV = g.nV
    fx = fixed.reshape(nV, 3).astype(bool) if fixed is not None else np.zeros((nV, 3), bool)

    def masked(v, slot, n):
        A = block(vals, g.bptr, v, slot).astype(np.float64)
        for l in range(3):
            for k in range(3):
                if fx[v, l] or fx[n, k]:
                    A[l, k] = 1.0 if (v == n and l == k) else 0.0
        return A

    s = np.zeros(nV)
    C = np.zeros((nV, 3, 3))
    xn = np.zeros((nV, 8))
    for v in range(nV):
        D = masked(v, g.dslot[v], v)
        s[v] = np.sqrt(max(np.trace(D) / 3.0, 1e-30))
        Dn = D / s[v] ** 2
        C[v] = np.linalg.cholesky(Dn)
        xn[v] = [Dn[0, 0], Dn[1, 1], Dn[2, 2], Dn[1, 0], Dn[2, 0], Dn[2, 1], fx[v].sum() / 3.0,
                 (g.bptr[v + 1] - g.bptr[v] - 1) / 16.0]
    fp = np.zeros((g.nP, 10))
    for k in range(g.nP):
        v, n = g.pv[k], g.lcol[k]
        A = masked(v, g.lslot[k], n) / (s[v] * s[n])
        fp[k, :9] = A.ravel()
        fp[k, 9] = np.log(s[v] / s[n])
    return s, C, xn, fp


def bf16(x):
    """x rounded to BF16 (nearest even), kept in x's dtype"""
    return x.to(torch.bfloat16).to(x.dtype)


def mlp(P, prefix, x, round_bf16=False):
    """Linear -> ReLU -> Linear; round_bf16: the MLP's input, weights and hidden activations rounded to
    BF16 (the biases and the sums stay wide), the arithmetic of the library's BF16 inference"""
    r = bf16 if round_bf16 else (lambda t: t)
    hid = torch.relu(r(x) @ r(P[prefix + ".W1"]) + P[prefix + ".b1"])
    return r(hid) @ r(P[prefix + ".W2"]) + P[prefix + ".b2"]


def forward(P, g, xn, fp, round_bf16=False):
    """decoder outputs d (nV x 6) and O (nP x 9)"""
    pv = torch.as_tensor(g.pv)
    pn = torch.as_tensor(g.lcol.astype(np.int64))
    h = mlp(P, "nenc", xn, round_bf16)
    e = mlp(P, "penc", fp, round_bf16)
    for l in range(LAYERS):
        e = e + mlp(P, f"edge{l}", torch.cat([e, h[pv], h[pn]], dim=1), round_bf16)
        m = torch.zeros_like(h).index_add(0, pv, e).index_add(0, pn, e)
        h = h + mlp(P, f"node{l}", torch.cat([h, m], dim=1), round_bf16)
    d = h @ P["Wdv"] + P["bdv"]
    O = e @ P["Wdo"] + P["bdo"]
    return d, O


def factor(C, d, O, g=None, fp=None, add=False):
    """normalised factor blocks: Lvv (nV x 3 x 3) = C T(d), Lk (nP x 3 x 3) = O; or, in the Eisenstat
    (symmetric Gauss-Seidel) form (g and the pair features fp given), Lk = Hn_vn Lvv_n^{-T} with Hn_vn
    the first 9 pair features, so that L L^T = (Dt + L_H) Dt^{-1} (Dt + L_H)^T, Dt = Lvv Lvv^T"""
    nV = d.shape[0]
    T = torch.zeros(nV, 3, 3, dtype=d.dtype)
    T[:, 0, 0] = torch.exp(d[:, 0])
    T[:, 1, 0] = d[:, 1]
    T[:, 1, 1] = torch.exp(d[:, 2])
    T[:, 2, 0] = d[:, 3]
    T[:, 2, 1] = d[:, 4]
    T[:, 2, 2] = torch.exp(d[:, 5])
    Lvv = torch.bmm(C, T)
    if fp is None:
        return Lvv, O.reshape(-1, 3, 3)
    pn = torch.as_tensor(g.lcol.astype(np.int64))
    Hn = fp[:, :9].reshape(-1, 3, 3)
    Xinv_t = torch.linalg.inv(Lvv[pn]).transpose(1, 2).contiguous()
    tied = torch.bmm(Hn, Xinv_t)
    # add: the tied factor form, the pair decoder's output added to the tied blocks
    return Lvv, (tied + O.reshape(-1, 3, 3)) if add else tied


def apply_LLt(g, Lvv, Lk, u):
    """L L^T u for per-vertex 3-vectors u (nV x 3)"""
    pv = torch.as_tensor(g.pv)
    pn = torch.as_tensor(g.lcol.astype(np.int64))
    y = torch.einsum("vji,vj->vi", Lvv, u)
    y = y.index_add(0, pn, torch.einsum("kji,kj->ki", Lk, u[pv]))
    z = torch.einsum("vij,vj->vi", Lvv, y)
    z = z.index_add(0, pv, torch.einsum("kij,kj->ki", Lk, y[pn]))
    return z


def loss(P, g, feats, x, b, tied=False, tied_factor=False):
    """|Lhat Lhat^T xh - bh|^2 / |bh|^2 with xh = s x, bh = b / s (tied: the Eisenstat form; tied_factor:
    the tied factor form, Lk = Hn_vn Lvv_n^{-T} + O)"""
    s, C, xn, fp = (torch.as_tensor(a) for a in feats)
    d, O = forward(P, g, xn, fp)
    if tied or tied_factor:
        Lvv, Lk = factor(C, d, O, g, fp, add=tied_factor)
        O = Lk.reshape(-1, 9)
    else:
        Lvv, Lk = factor(C, d, O)
    xh = torch.as_tensor(x.reshape(-1, 3)) * s[:, None]
    bh = torch.as_tensor(b.reshape(-1, 3)) / s[:, None]
    z = apply_LLt(g, Lvv, Lk, xh)
    return ((z - bh) ** 2).sum() / (bh ** 2).sum(), d, O


def dense_factor(g, s, Lvv, Lk, fixed=None):
    """the factor of H as a dense 3nV x 3nV lower-triangular-by-colour matrix: L = S Lhat, with
    identity rows and columns on the fixed DOFs (the masked system)"""
    nV = g.nV
    L = np.zeros((3 * nV, 3 * nV))
    for v in range(nV):
        L[3 * v:3 * v + 3, 3 * v:3 * v + 3] = s[v] * Lvv[v]
    for k in range(g.nP):
        v, n = g.pv[k], g.lcol[k]
        L[3 * v:3 * v + 3, 3 * n:3 * n + 3] = s[v] * Lk[k]
    if fixed is not None:
        fx = np.asarray(fixed).astype(bool)
        L[fx, :] = 0.0
        L[:, fx] = 0.0
        L[fx, fx] = 1.0
    return L


def energy_loss(P, g, feats, xs, bs):
    """Eisenstat form, energy-ratio loss: mean over probes of (log x^T M x - log x^T H x)^2 with
    M = (Dt + L) Dt^{-1} (Dt + L)^T in the normalised space, Dt = Lvv Lvv^T, L the lower pairs of Hn.
    x^T M x = sum_v [xh_v^T Dt_v xh_v + 2 xh_v^T z_v + z_v^T Dt_v^{-1} z_v], z_v = sum_{pairs (n, v)} Hn_nv^T xh_n"""
    s, C, xn, fp = (torch.as_tensor(a) for a in feats)
    d, O = forward(P, g, xn, fp)
    Lvv, _ = factor(C, d, O)
    D = torch.bmm(Lvv, Lvv.transpose(1, 2).contiguous())
    Dinv = torch.linalg.inv(D)
    pv = torch.as_tensor(g.pv)
    pn = torch.as_tensor(g.lcol.astype(np.int64))
    Hn = fp[:, :9].reshape(-1, 3, 3)
    total = 0.0
    for x, b in zip(xs, bs):
        xh = torch.as_tensor(x.reshape(-1, 3)) * s[:, None]
        z = torch.zeros_like(xh).index_add(0, pn, torch.einsum("kml,km->kl", Hn, xh[pv]))
        q = (torch.einsum("vi,vij,vj->v", xh, D, xh) + 2.0 * (xh * z).sum(1) +
             torch.einsum("vi,vij,vj->v", z, Dinv, z)).sum()
        a = float(np.dot(x, b))
        total = total + (torch.log(q) - np.log(a)) ** 2
    return total / len(xs), d

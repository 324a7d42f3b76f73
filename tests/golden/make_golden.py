"""Writes tests/golden/elements.npz: torch-FP64 autograd golden vectors of the constitutive models.

Energies are restated from the mathematics of the reference's energy functions (not from their
generated derivative code), then differentiated by autograd, so the golden gradient and Hessian
are independent of any hand-derived formula:
  NH (active)   psi = C (tr F^T F - 3 - 2 ln J) + D (J - 1)^2          src/psi_neohookean_F.cpp:31
  NH (legacy)   psi = C (J^{-2/3} tr F^T F - 3) + D (J - 1)^2          src/psi_neohookean_F.cpp:8-14
  StVK          psi = mu tr(E^2) + lambda/2 (tr E)^2, E = (F^T F - I)/2 src/psi_stvk_F.cpp:5-13
  corotational  psi = mu sum (s_i - 1)^2 + lambda/2 (sum s_i - 3)^2    src/psi_corotational_S.cpp:5-9
  ARAP          psi = k sum (s_i - 1)^2                                 src/psi_arap_S.cpp:5-11
For the stretch models the reference's 9x9 layout equals the Hessian evaluated at F^T
(src/d2psi_stretch_dF2.cpp:36-45); both are stored.  Samples have det F > 0, J in about
[0.2, 1.5] and singular-value gaps >= 0.03 (away from the reference's random perturbation).
vec(F) is column-major throughout.
"""
import os

import numpy as np
import torch

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elements.npz")
N = 48


def rot(rng):
    A = rng.standard_normal((3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def samples(rng):
    Fs = []
    while len(Fs) < N:
        s = np.sort(rng.uniform(0.55, 1.3, 3))[::-1]
        if np.min(np.diff(s[::-1])) < 0.03 or not (0.2 <= np.prod(s) <= 1.5):
            continue
        Fs.append(rot(rng) @ np.diag(s) @ rot(rng))
    return np.array(Fs)


def vecF(x):  # x: 9 column-major -> 3x3
    return x.reshape(3, 3).T


def psi_torch(model, f9, p):
    F = vecF(f9)
    if model in (0, 4):
        J = torch.det(F)
        I1 = (F * F).sum()
        if model == 0:
            return p[0] * (I1 - 3.0 - 2.0 * torch.log(J)) + p[1] * (J - 1.0) ** 2
        return p[0] * (J ** (-2.0 / 3.0) * I1 - 3.0) + p[1] * (J - 1.0) ** 2
    if model == 1:
        E = 0.5 * (F.T @ F - torch.eye(3, dtype=F.dtype))
        return p[1] * torch.trace(E @ E) + 0.5 * p[0] * torch.trace(E) ** 2
    s = torch.linalg.svdvals(F)
    if model == 2:
        return p[1] * ((s - 1.0) ** 2).sum() + 0.5 * p[0] * (s.sum() - 3.0) ** 2
    return p[0] * ((s - 1.0) ** 2).sum()


def main():
    rng = np.random.default_rng(20251203)
    out = {}
    for model in range(5):
        Fs = samples(rng)
        ps = np.stack([rng.uniform(1e4, 1e6, N), rng.uniform(1e4, 5e6, N)], axis=1)
        psi_v, P_v, H_v, Hr_v = [], [], [], []
        for F, p in zip(Fs, ps):
            f9 = torch.tensor(F.T.reshape(-1), dtype=torch.float64, requires_grad=True)
            pt = torch.tensor(p, dtype=torch.float64)
            fn = lambda x: psi_torch(model, x, pt)
            psi_v.append(fn(f9).item())
            P_v.append(torch.autograd.functional.jacobian(fn, f9).numpy())
            H_v.append(torch.autograd.functional.hessian(fn, f9).numpy().reshape(-1, order="F"))
            ft = torch.tensor(F.reshape(-1), dtype=torch.float64)  # vec(F^T) column-major = F row-major
            Hr_v.append(torch.autograd.functional.hessian(fn, ft).numpy().reshape(-1, order="F"))
        out[f"F{model}"] = np.array([F.T.reshape(-1) for F in Fs])
        out[f"p{model}"] = ps
        out[f"psi{model}"] = np.array(psi_v)
        out[f"P{model}"] = np.array(P_v)
        out[f"H{model}"] = np.array(H_v)
        out[f"Href{model}"] = np.array(Hr_v)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()

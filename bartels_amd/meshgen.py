"""Synthetic inputs of the benchmark configurations: beam grids, deformations and materials.

Everything is generated by a portable counter-based RNG (splitmix64 -> uniform in [0,1) ->
Box-Muller), so the same arrays can be regenerated bit-for-bit on any machine.

Beam convention: the long axis is x; vertex (ix, iy, iz) has index (ix*ny + iy)*nz + iz; vertices
with x == 0 are clamped; gravity is (0, 0, -9.8).  Each hex cell is split into 6 tets along its
(0,0,0)-(1,1,1) diagonal (Kuhn split, conforming), numbered type-major (all tets of the first
type, cell by cell, then the second type, ...), oriented so that the signed volume
vol = -(a-d).((b-d)x(c-d))/6 (igl::volume, as the reference uses in tests/test_assemble.cpp:27)
is positive.  Arrays follow the reference layouts: V nV x 3, E nT x 4 int32.
"""
import itertools
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
H_DEFAULT = 0.025


# ---- RNG ---------------------------------------------------------------------------------------
def splitmix64(seed, n, offset=0):
    """n outputs of the splitmix64 stream with the given seed (element i uses counter offset+i+1)."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, offset=0):
    """uniform doubles in [0, 1): top 53 bits of splitmix64."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n, offset=0):
    """standard normals by Box-Muller on consecutive uniform pairs."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, offset)
    u1, u2 = 1.0 - u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * math.pi * u2)
    out[1::2] = r * np.sin(2.0 * math.pi * u2)
    return out[:n]


# ---- materials ---------------------------------------------------------------------------------
def emu_to_lame(E, nu):
    """Young's modulus / Poisson ratio -> (lambda, mu).  Reference: matlab/emu_to_lame.m:17-18."""
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return lam, mu


def model_params(model_ids, lam, mu):
    """nT x 2 parameter matrix in each model's reference convention.
    Neo-Hookean [C, D] = [mu/2, lambda/2] (README.md:52-53); StVK / corotational [lambda, mu]
    (matlab/examples/test_tet_materials.m:69-76); ARAP [0.5 lambda, mu] as the MATLAB example
    passes it (test_tet_materials.m:80-82; only params(0) is read)."""
    ids = np.asarray(model_ids)
    lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), ids.shape)
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), ids.shape)
    p = np.empty((ids.shape[0], 2))
    nh = (ids == 0) | (ids == 4)
    p[nh, 0] = 0.5 * mu[nh]
    p[nh, 1] = 0.5 * lam[nh]
    sv = (ids == 1) | (ids == 2)
    p[sv, 0] = lam[sv]
    p[sv, 1] = mu[sv]
    ar = ids == 3
    p[ar, 0] = 0.5 * lam[ar]
    p[ar, 1] = mu[ar]
    return p


# ---- meshes ------------------------------------------------------------------------------------
def _kuhn_tets():
    """the 6 tets of the unit cube along the main diagonal, as corner bit patterns, positively oriented"""
    tets = []
    for perm in itertools.permutations(range(3)):
        path = [(0, 0, 0)]
        c = [0, 0, 0]
        for ax in perm:
            c = list(c)
            c[ax] = 1
            path.append(tuple(c))
        pts = np.array(path, dtype=np.float64)
        a, b, cc, d = pts
        vol = -np.dot(a - d, np.cross(b - d, cc - d)) / 6.0
        if vol < 0:
            path[1], path[2] = path[2], path[1]
        tets.append(path)
    return tets


def beam_grid(nx, ny, nz, h=H_DEFAULT):
    """Structured beam: nx x ny x nz vertices (x is the long axis), spacing h, 6 tets per cell."""
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    V = np.stack([ix.ravel() * h, iy.ravel() * h, iz.ravel() * h], axis=1).astype(np.float64)
    cx, cy, cz = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()

    def vid(dx, dy, dz):
        return ((cx + dx) * ny + (cy + dy)) * nz + (cz + dz)

    tets = _kuhn_tets()
    # type-major numbering: tet (type t, cell c) has index t * ncells + c, so the incident tets of
    # consecutive vertices (consecutive cells) are consecutive in every per-tet array
    E = np.empty((6, cx.size, 4), dtype=np.int32)
    for t, path in enumerate(tets):
        for k, (dx, dy, dz) in enumerate(path):
            E[t, :, k] = vid(dx, dy, dz)
    return V, E.reshape(-1, 4)


def strip2d(nx, ny, h=H_DEFAULT, seed=None, jitter_amount=0.0):
    """2-D triangle strip: nx x ny vertices (x the long axis), spacing h, each cell split along its
    rising diagonal into two counter-clockwise triangles; optional uniform jitter of the vertices by
    +-jitter_amount h (seeded), which keeps every triangle counter-clockwise for amounts below 0.25."""
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    V = np.stack([ix.ravel() * h, iy.ravel() * h], axis=1).astype(np.float64)
    cx, cy = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    cx, cy = cx.ravel(), cy.ravel()

    def vid(dx, dy):
        return (cx + dx) * ny + (cy + dy)

    E = np.concatenate([np.stack([vid(0, 0), vid(1, 0), vid(1, 1)], 1),
                        np.stack([vid(0, 0), vid(1, 1), vid(0, 1)], 1)]).astype(np.int32)
    if seed is not None and jitter_amount > 0:
        V = V + (2.0 * uniform(seed, V.size).reshape(V.shape) - 1.0) * jitter_amount * h
    return V, E


def triangle_areas(V, E):
    """areas of 2-D triangles (positive for counter-clockwise ones)"""
    a, b, c = (V[E[:, k]] for k in range(3))
    u, w = b - a, c - a
    return 0.5 * (u[:, 0] * w[:, 1] - u[:, 1] * w[:, 0])


def signed_volumes(V, E):
    a, b, c, d = (V[E[:, k]] for k in range(4))
    return -np.einsum("ij,ij->i", a - d, np.cross(b - d, c - d)) / 6.0


def jitter(V, E, h, amount, seed, max_rounds=50):
    """uniform jitter of +-amount*h per coordinate; vertices of non-positive tets are redrawn."""
    V = V.copy()
    V0 = V.copy()
    n = V.size
    V += (2.0 * uniform(seed, n).reshape(V.shape) - 1.0) * amount * h
    for r in range(1, max_rounds + 1):
        bad = signed_volumes(V, E) <= 0
        if not bad.any():
            return V
        vids = np.unique(E[bad].ravel())
        redraw = (2.0 * uniform(seed, 3 * vids.size, offset=n * (r + 1)).reshape(-1, 3) - 1.0) * amount * h
        V[vids] = V0[vids] + redraw
    raise RuntimeError("jitter: could not make every tet positive")


def read_mesh(path):
    """MEDIT .mesh reader (vertices and tetrahedra, returned 0-based): the library's C++ reader,
    bh_read_medit."""
    from .io import read_mesh as _read

    V, E = _read(path)
    return np.ascontiguousarray(V), np.ascontiguousarray(E)


# ---- the configurations --------------------------------------------------------------------------
class Problem(dict):
    """a config's arrays: V, E, q, params, model_ids (None = uniform), model, density, gravity, fixed."""

    __getattr__ = dict.__getitem__


def fixed_vertices(V, tol=1e-3):
    """vertices at the minimum x (reference find_min_vertices, src/find_min_vertices.cpp:5-21, tol 1e-3)."""
    return np.nonzero(np.abs(V[:, 0] - V[:, 0].min()) <= tol)[0]


def config1(grid=(40, 10, 10), h=H_DEFAULT, seed=0):
    """Plain-method CPU-runnable config: Neo-Hookean beam, 10x10x40 vertices, q = X + N(0, (1e-3 h)^2)."""
    nx, ny, nz = grid
    V, E = beam_grid(nx, ny, nz, h)
    lam, mu = emu_to_lame(2e6, 0.45)
    q = V.ravel() + normal(seed, V.size) * (1e-3 * h)
    ids = np.zeros(E.shape[0], dtype=np.uint8)
    return Problem(V=V, E=E, q=q, params=model_params(ids, lam, mu), model=0, model_ids=None,
                   density=1000.0, gravity=np.array([0.0, 0.0, -9.8]), fixed=fixed_vertices(V), h=h)


def bend_field(V, L, seed):
    """sum of 4 transverse sinusoids along x with amplitudes up to 0.05 L."""
    u = uniform(seed, 16)
    d = np.zeros_like(V)
    x = V[:, 0]
    for k in range(4):
        a = (2.0 * u[4 * k] - 1.0) * 0.05 * L
        f = 0.5 + 1.5 * u[4 * k + 1]
        ph = 2.0 * math.pi * u[4 * k + 2]
        ang = 2.0 * math.pi * u[4 * k + 3]
        s = a * np.sin(2.0 * math.pi * f * x / L + ph)
        d[:, 1] += s * math.cos(ang)
        d[:, 2] += s * math.sin(ang)
    return d


def config2(grid=(256, 64, 64), h=H_DEFAULT, seed=1):
    """Plain method at scale: 64x64x256 Neo-Hookean beam (6.07M tets) with a random smooth bend."""
    nx, ny, nz = grid
    V, E = beam_grid(nx, ny, nz, h)
    lam, mu = emu_to_lame(2e6, 0.45)
    L = (nx - 1) * h
    q = (V + bend_field(V, L, seed)).ravel()
    ids = np.zeros(E.shape[0], dtype=np.uint8)
    return Problem(V=V, E=E, q=q, params=model_params(ids, lam, mu), model=0, model_ids=None,
                   density=1000.0, gravity=np.array([0.0, 0.0, -9.8]), fixed=fixed_vertices(V), h=h)


def config3(grid=(192, 48, 48), h=H_DEFAULT, seed=3):
    """Precision form: jittered 48x48x192 beam, heterogeneous materials, mixed models
    (70% NH, 20% StVK, 10% corotational), compressed toward J in [0.2, 1.5]."""
    nx, ny, nz = grid
    V, E = beam_grid(nx, ny, nz, h)
    V = jitter(V, E, h, 0.2, seed)
    nT = E.shape[0]
    u = uniform(seed + 1, 3 * nT).reshape(3, nT)
    Ey = np.exp(np.log(1e5) + (np.log(1e7) - np.log(1e5)) * u[0])
    nu = 0.30 + 0.18 * u[1]
    ids = np.where(u[2] < 0.7, 0, np.where(u[2] < 0.9, 1, 2)).astype(np.uint8)
    lam, mu = emu_to_lame(Ey, nu)
    L = (nx - 1) * h
    Jx = 0.2 + 1.3 * (0.5 + 0.5 * np.sin(2.0 * math.pi * V[:, 0] / L))
    s = np.sqrt(Jx)
    c = V.mean(axis=0)
    q = V.copy()
    q[:, 1] = c[1] + s * (V[:, 1] - c[1])
    q[:, 2] = c[2] + s * (V[:, 2] - c[2])
    return Problem(V=V, E=E, q=q.ravel(), params=model_params(ids, lam, mu), model=255, model_ids=ids,
                   density=1000.0, gravity=np.array([0.0, 0.0, -9.8]), fixed=fixed_vertices(V), h=h)


def config5(grid=(256, 64, 64), h=H_DEFAULT, seed=5):
    """Full chosen form: the 64x64x256 Neo-Hookean beam at rest, x = 0 clamped, falling under gravity
    (backward Euler, dt = 0.01 s), with per-tet Young's modulus log-uniform in [1e5, 1e7] Pa
    (Poisson 0.45)."""
    nx, ny, nz = grid
    V, E = beam_grid(nx, ny, nz, h)
    nT = E.shape[0]
    u = uniform(seed, nT)
    Ey = np.exp(np.log(1e5) + (np.log(1e7) - np.log(1e5)) * u)
    lam, mu = emu_to_lame(Ey, 0.45)
    ids = np.zeros(nT, dtype=np.uint8)
    return Problem(V=V, E=E, q=V.ravel().copy(), params=model_params(ids, lam, mu), model=0, model_ids=None,
                   density=1000.0, gravity=np.array([0.0, 0.0, -9.8]), fixed=fixed_vertices(V), h=h, dt=0.01)

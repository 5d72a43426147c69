"""Mesh generators and seeded inputs for the FEM time step.

* ``bar_freudenthal`` -- regular cube grid split into 6 tets per cube (Freudenthal/Kuhn split), the mesh
  the benchmark configurations are quoted on.  Vertex ids are x-slowest, as in the reference generator
  ``unit_box`` (ssim/include/ssim/mesh/common_shapes.hpp:77-87).
* ``unit_box`` -- the reference's own 5-tet alternating split with *vertex* counts per axis
  (ssim/include/ssim/mesh/common_shapes.hpp:67-132), returned with the Python API's (3, N) / (4, M) layout
  (bindings/src/fem_host.cpp:216-219).
* ``tetra`` -- the single reference tetrahedron (ssim/include/ssim/mesh/common_shapes.hpp:134-145).

Random inputs come from a counter-based splitmix64 stream so that C, C++ and numpy produce identical bits:
``uniform(seed, stream, count)[i] = (splitmix64(seed * 2^32 + stream, i) >> 11) * 2^-53``.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform(seed: int, stream: int, count: int) -> np.ndarray:
    """``count`` doubles in [0, 1) from the splitmix64 stream (seed, stream)."""
    key = np.uint64((int(seed) * (1 << 32) + int(stream)) % (1 << 64))
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = key + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def bar_freudenthal(nx: int, ny: int, nz: int, h: float = 0.01):
    """Cube grid of nx×ny×nz cubes of edge h, 6 tets per cube, every tet with det Dm > 0.

    Returns ``(X [nv,3] float64, cells [nt,4] int32)``.  Vertex (i, j, k) has id ((i·(ny+1) + j)·(nz+1) + k).
    Cube (i, j, k) holds the 6 tets of the monotone paths 000 → e_p0 → e_p0+e_p1 → 111 for the permutations
    p of (x, y, z) in lexicographic order; odd permutations swap their 2nd and 3rd vertex.
    """
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([ii.ravel() * h, jj.ravel() * h, kk.ravel() * h], axis=1).astype(np.float64)

    def vid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    unit = np.eye(3, dtype=np.int64)
    tets = []
    for perm in itertools.permutations(range(3)):
        a = unit[perm[0]]
        b = unit[perm[0]] + unit[perm[1]]
        v0 = vid(ci, cj, ck)
        v1 = vid(ci + a[0], cj + a[1], ck + a[2])
        v2 = vid(ci + b[0], cj + b[1], ck + b[2])
        v3 = vid(ci + 1, cj + 1, ck + 1)
        inv = sum(1 for x in range(3) for y in range(x + 1, 3) if perm[x] > perm[y])
        if inv % 2 == 1:
            v1, v2 = v2, v1
        tets.append(np.stack([v0, v1, v2, v3], axis=1))
    cells = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return X, cells


def unit_box(nx: int = 2, ny: int = 2, nz: int = 2):
    """Reference ``unit_box`` (5 tets per cube, alternating), returned as ``(verts (3,N), cells (4,M))``."""
    X, cells = unit_box_rows(nx, ny, nz)
    return np.ascontiguousarray(X.T), np.ascontiguousarray(cells.T)


def unit_box_rows(nx: int = 2, ny: int = 2, nz: int = 2):
    """``unit_box`` with row layout: X [N,3] float64, cells [M,4] int32 (common_shapes.hpp:67-132)."""
    nv = nx * ny * nz
    X = np.zeros((nv, 3))
    idx = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                X[idx] = (i / (nx - 1), j / (ny - 1), k / (nz - 1))
                idx += 1
    cells = []
    for i in range(nx - 1):
        for j in range(ny - 1):
            for k in range(nz - 1):
                c = {
                    "000": i * ny * nz + j * nz + k,
                    "100": (i + 1) * ny * nz + j * nz + k,
                    "010": i * ny * nz + (j + 1) * nz + k,
                    "001": i * ny * nz + j * nz + k + 1,
                    "110": (i + 1) * ny * nz + (j + 1) * nz + k,
                    "101": (i + 1) * ny * nz + j * nz + k + 1,
                    "011": i * ny * nz + (j + 1) * nz + k + 1,
                    "111": (i + 1) * ny * nz + (j + 1) * nz + k + 1,
                }
                if i % 2 == 1:
                    for a, b in (("000", "100"), ("010", "110"), ("001", "101"), ("011", "111")):
                        c[a], c[b] = c[b], c[a]
                if j % 2 == 1:
                    for a, b in (("000", "010"), ("100", "110"), ("001", "011"), ("101", "111")):
                        c[a], c[b] = c[b], c[a]
                if k % 2 == 1:
                    for a, b in (("000", "001"), ("100", "101"), ("010", "011"), ("110", "111")):
                        c[a], c[b] = c[b], c[a]
                cells.append((c["000"], c["100"], c["010"], c["001"]))
                cells.append((c["010"], c["100"], c["110"], c["111"]))
                cells.append((c["100"], c["010"], c["001"], c["111"]))
                cells.append((c["100"], c["001"], c["101"], c["111"]))
                cells.append((c["011"], c["001"], c["010"], c["111"]))
    return X, np.asarray(cells, dtype=np.int32).reshape(-1, 4)


def tetra():
    """Reference single tet (common_shapes.hpp:134-145): X [4,3], cells [1,4]."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    return X, np.array([[0, 1, 2, 3]], dtype=np.int32)


def lame(E, nu):
    """(λ, μ) from Young's modulus and Poisson ratio (ssim/include/ssim/elast/basic_elast.hpp:11-20)."""
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam, mu


@dataclass
class BarCase:
    """One seeded bar problem: mesh, material, boundary, loading and initial state."""

    X: np.ndarray
    cells: np.ndarray
    lam: np.ndarray  # per tet
    mu: np.ndarray  # per tet
    rho: float
    dt: float
    dof_type: np.ndarray  # uint8 [nv*3]
    dbc: np.ndarray  # [nv*3]
    accel: np.ndarray  # [nv*3]
    u0: np.ndarray  # [nv*3]
    v0: np.ndarray  # [nv*3]
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nv(self):
        return self.X.shape[0]

    @property
    def nt(self):
        return self.cells.shape[0]


def cantilever_bar(nx=32, ny=8, nz=8, h=0.01, E=1e5, nu=0.45, rho=1000.0, dt=1.0 / 60, seed=0,
                   jitter=0.05, gravity=-9.8) -> BarCase:
    """Cantilever bar of nx×ny×nz Freudenthal cubes (long axis x), x = 0 face clamped, gravity along y.

    Initial deformation: every free vertex component jittered by U(−jitter·h, +jitter·h) from stream
    (seed, 1); clamped vertices start at rest (their Dirichlet value is 0).  Velocity 0.
    """
    X, cells = bar_freudenthal(nx, ny, nz, h)
    nv, nt = X.shape[0], cells.shape[0]
    lam, mu = lame(np.full(nt, E), np.full(nt, nu))
    clamped = X[:, 0] == 0.0
    dof = np.repeat(clamped.astype(np.uint8), 3)
    u0 = (uniform(seed, 1, 3 * nv) * 2 - 1) * jitter * h
    u0[dof == 1] = 0.0
    accel = np.zeros(3 * nv)
    accel[1::3] = gravity
    return BarCase(X, cells, lam, mu, rho, dt, dof, np.zeros(3 * nv), accel, u0, np.zeros(3 * nv),
                   name=f"bar{nx}x{ny}x{nz}", meta=dict(nx=nx, ny=ny, nz=nz, h=h, seed=seed))


def twisted_cube(n=64, h=0.01, rho=1000.0, dt=1.0 / 60, seed=0, max_twist_deg=90.0, jitter=0.1):
    """n³ Freudenthal cube with per-tet E log-uniform in [1e4, 1e6] Pa (stream 2), ν uniform in
    [0.3, 0.48] (stream 3), z = 0 face clamped, initial state a twist about the z axis through the cube centre
    by an angle growing linearly with z up to a seeded maximum ≤ max_twist_deg (stream 4), plus
    U(±jitter·h) per free component (stream 5)."""
    X, cells = bar_freudenthal(n, n, n, h)
    nv, nt = X.shape[0], cells.shape[0]
    E = np.exp(np.log(1e4) + uniform(seed, 2, nt) * (np.log(1e6) - np.log(1e4)))
    nu = 0.3 + uniform(seed, 3, nt) * 0.18
    lam, mu = lame(E, nu)
    clamped = X[:, 2] == 0.0
    dof = np.repeat(clamped.astype(np.uint8), 3)
    theta_max = np.deg2rad(max_twist_deg) * uniform(seed, 4, 1)[0]
    L = n * h
    c = 0.5 * L
    theta = theta_max * X[:, 2] / L
    dx, dy = X[:, 0] - c, X[:, 1] - c
    xr = c + np.cos(theta) * dx - np.sin(theta) * dy
    yr = c + np.sin(theta) * dx + np.cos(theta) * dy
    u = np.zeros((nv, 3))
    u[:, 0] = xr - X[:, 0]
    u[:, 1] = yr - X[:, 1]
    u = u.ravel() + (uniform(seed, 5, 3 * nv) * 2 - 1) * jitter * h
    u[dof == 1] = 0.0
    return BarCase(X, cells, lam, mu, rho, dt, dof, np.zeros(3 * nv), np.zeros(3 * nv), u, np.zeros(3 * nv),
                   name=f"twist{n}", meta=dict(n=n, h=h, seed=seed, theta_max=float(theta_max)))


def bar_batch(nmesh=256, seed=0, lo=4, hi=24, h=0.01, rho=1000.0, dt=1.0 / 60, jitter=0.05, gravity=-9.8):
    """Batch of independent cantilever bars of uneven size.

    Bar b: grid dims (nx, ny, nz) uniform integers in [lo, hi] (stream (seed, 100), three draws per bar), one
    E log-uniform in [1e4, 1e6] Pa (stream 101) and ν uniform in [0.3, 0.48] (stream 102) for all its tets,
    x = 0 end clamped, gravity along y, free components jittered by U(±jitter·h) (stream 1000 + b).
    """
    dims = lo + np.floor(uniform(seed, 100, 3 * nmesh) * (hi - lo + 1)).astype(np.int64)
    dims = np.minimum(dims, hi).reshape(nmesh, 3)
    E = np.exp(np.log(1e4) + uniform(seed, 101, nmesh) * (np.log(1e6) - np.log(1e4)))
    nu = 0.3 + uniform(seed, 102, nmesh) * 0.18
    cases = []
    for b in range(nmesh):
        nx, ny, nz = (int(d) for d in dims[b])
        X, cells = bar_freudenthal(nx, ny, nz, h)
        nv, nt = X.shape[0], cells.shape[0]
        lam, mu = lame(np.full(nt, E[b]), np.full(nt, nu[b]))
        clamped = X[:, 0] == 0.0
        dof = np.repeat(clamped.astype(np.uint8), 3)
        u0 = (uniform(seed, 1000 + b, 3 * nv) * 2 - 1) * jitter * h
        u0[dof == 1] = 0.0
        accel = np.zeros(3 * nv)
        accel[1::3] = gravity
        cases.append(BarCase(X, cells, lam, mu, rho, dt, dof, np.zeros(3 * nv), accel, u0, np.zeros(3 * nv),
                             name=f"batchbar{b}_{nx}x{ny}x{nz}",
                             meta=dict(nx=nx, ny=ny, nz=nz, h=h, seed=seed, E=float(E[b]), nu=float(nu[b]))))
    return cases


@dataclass
class RolloutCase:
    """A batch of bars for a differentiable rollout: per-tet log E, fixed ν, and a free-end target offset."""

    cases: list  # BarCase per bar (its lam / mu are the initial material)
    log_e: np.ndarray  # per tet, all bars concatenated
    nu: np.ndarray  # per tet
    targets: np.ndarray  # (nbar, 3) free-end offsets
    free_end: list  # free-end vertex indices per bar (x = max face)
    steps: int = 20


def rollout_batch(nbar=32, nx=64, ny=16, nz=16, h=0.01, nu=0.45, rho=1000.0, dt=1.0 / 60, seed=0, jitter=0.05,
                  gravity=-9.8, offset=0.02, steps=20):
    """Differentiable-rollout batch: nbar copies of the nx×ny×nz bar (x = 0 end clamped, gravity along y), each
    with per-tet log E uniform in [ln 1e4, ln 1e6] (stream (seed, 200 + b)), ν fixed, free components jittered by
    U(±jitter·h) (stream 300 + b), v₀ = 0, and a free-end target offset of norm `offset` in a seeded direction
    (stream 400, three normal draws per bar by Box-Muller from the uniform stream)."""
    X, cells = bar_freudenthal(nx, ny, nz, h)
    nv, nt = X.shape[0], cells.shape[0]
    clamped = X[:, 0] == 0.0
    dof = np.repeat(clamped.astype(np.uint8), 3)
    accel = np.zeros(3 * nv)
    accel[1::3] = gravity
    fe = np.nonzero(X[:, 0] == X[:, 0].max())[0]
    u1 = uniform(seed, 400, 6 * nbar).reshape(nbar, 6)
    z = np.sqrt(-2 * np.log(1 - u1[:, :3])) * np.cos(2 * np.pi * u1[:, 3:])  # Box-Muller normals
    targets = offset * z / np.linalg.norm(z, axis=1, keepdims=True)
    cases, log_e = [], []
    for b in range(nbar):
        le = np.log(1e4) + uniform(seed, 200 + b, nt) * (np.log(1e6) - np.log(1e4))
        lam, mu = lame(np.exp(le), np.full(nt, nu))
        u0 = (uniform(seed, 300 + b, 3 * nv) * 2 - 1) * jitter * h
        u0[dof == 1] = 0.0
        cases.append(BarCase(X, cells, lam, mu, rho, dt, dof, np.zeros(3 * nv), accel, u0, np.zeros(3 * nv),
                             name=f"rollout{b}_{nx}x{ny}x{nz}", meta=dict(nx=nx, ny=ny, nz=nz, h=h, seed=seed)))
        log_e.append(le)
    return RolloutCase(cases, np.concatenate(log_e), np.full(nbar * nt, nu), targets, [fe] * nbar, steps)

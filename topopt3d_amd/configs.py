"""Synthetic workloads: structured cantilever / column meshes with their boundary
conditions and seeded densities.

Randomness is SplitMix64 with seed = configuration index and u = (r >> 11) * 2^-53, so
every component of the build (tests, bench, oracle) draws the same numbers.

Mesh conventions follow the reference: nodes and elements x-fastest, then y, then z
(RecMesh.cpp:486-512, 565-597); DOF = 3*node + component (FEAnalysis.cpp:241).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """The first n outputs of SplitMix64 started from state `seed`."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = state
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """n i.i.d. uniform values in [lo, hi): u = (r >> 11) * 2^-53."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def node_index(nel, i, j, k):
    nx, ny, _ = nel
    return i + (nx + 1) * (j + (ny + 1) * k)


def face_fixed_dof(nel, axis: int, value: int) -> np.ndarray:
    """Per-DOF Dirichlet flags clamping every node with coordinate index `value` along
    `axis` in all three components."""
    nx, ny, nz = nel
    N = (nx + 1, ny + 1, nz + 1)
    idx = np.indices((N[2], N[1], N[0]))  # k, j, i
    coord = (idx[2], idx[1], idx[0])[axis]
    nodes = np.flatnonzero((coord == value).reshape(-1))
    fixed = np.zeros(3 * N[0] * N[1] * N[2], np.uint8)
    for c in range(3):
        fixed[3 * nodes + c] = 1
    return fixed


@dataclass
class Problem:
    name: str
    nel: tuple
    fixed_dof: np.ndarray
    F: np.ndarray
    x: np.ndarray
    rmin: float = 1.5
    penal: float = 3.0
    Emin: float = 1e-10
    mg_levels: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def ndof(self):
        nx, ny, nz = self.nel
        return 3 * (nx + 1) * (ny + 1) * (nz + 1)

    @property
    def nelem(self):
        nx, ny, nz = self.nel
        return nx * ny * nz


def cantilever(nel, x=None, rmin=1.5, penal=3.0, Emin=1e-10, mg_levels=0, seed=None,
               lo=0.01, hi=1.0, name="cantilever") -> Problem:
    """x = 0 face fully clamped; unit downward (-z) point load at the bottom-centre node
    (nx, ny/2, 0) of the x = max face."""
    nx, ny, nz = nel
    fixed = face_fixed_dof(nel, 0, 0)
    F = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    F[3 * node_index(nel, nx, ny // 2, 0) + 2] = -1.0
    ne = nx * ny * nz
    if x is None:
        x = uniform(seed, ne, lo, hi) if seed is not None else np.full(ne, 0.3)
    elif np.isscalar(x):
        x = np.full(ne, float(x))
    return Problem(name, tuple(nel), fixed, F, np.asarray(x, np.float64), rmin, penal, Emin,
                   mg_levels)


def cantilever_48x24x24(Emin=1e-10) -> Problem:
    """48x24x24 cantilever, uniform density 0.3, filter R = 1.5, 3-level GMG."""
    return cantilever((48, 24, 24), x=0.3, rmin=1.5, penal=3.0, Emin=Emin, mg_levels=3,
                      name="cantilever-48x24x24")


def cantilever_192x96x96(Emin=1e-9) -> Problem:
    """192x96x96 cantilever (5,447,811 DOF), densities U[0.01,1] seed 2, R = 1.5, p = 3,
    5-level V(2,2) Chebyshev GMG.  The densities are filtered inside tofe_set_design."""
    return cantilever((192, 96, 96), seed=2, lo=0.01, hi=1.0, rmin=1.5, penal=3.0, Emin=Emin,
                      mg_levels=5, name="cantilever-192x96x96")


def pressure_load(nel, axis: int, side: int, value) -> np.ndarray:
    """Load vector of a uniform pressure on the face coord[axis] = side * nel[axis] with the
    reference's per-face-node semantics (Inputs.cpp:126-147, 174-199): every node of every
    selected element face receives the full value vector, so a face node is loaded once per
    face it belongs to (4 inside the face, 2 on its edges, 1 at its corners)."""
    N = tuple(n + 1 for n in nel)
    value = np.asarray(value, np.float64)
    F = np.zeros((N[2], N[1], N[0], 3))
    others = [a for a in range(3) if a != axis]
    cnt = np.ones((N[2], N[1], N[0]))
    for a in others:
        m = np.full(N[a], 2.0)
        m[0] = m[-1] = 1.0
        shape = [1, 1, 1]
        shape[2 - a] = N[a]
        cnt = cnt * m.reshape(shape)
    sl = [slice(None)] * 3
    sl[2 - axis] = side * nel[axis]
    F[tuple(sl)] = cnt[tuple(sl)][..., None] * value
    return F.reshape(-1)


def column(nel, seed=3, lo=0.1, hi=1.0, rmin=2.0, penal=3.0, Emin=1e-9, mg_levels=0,
           name="column") -> Problem:
    """Linear-buckling column: the z = 0 face clamped, a uniform unit pressure (0, 0, -1)
    on the top face z = nz in the reference's per-face-node semantics, densities
    U[lo, hi) from SplitMix64(seed)."""
    nx, ny, nz = nel
    fixed = face_fixed_dof(nel, 2, 0)
    F = pressure_load(nel, 2, 1, (0.0, 0.0, -1.0))
    x = uniform(seed, nx * ny * nz, lo, hi)
    return Problem(name, tuple(nel), fixed, F, x, rmin, penal, Emin, mg_levels)


def column_128x64x64(Emin=1e-9) -> Problem:
    """128x64x64 column (1,635,075 DOF), bottom clamped, unit pressure on the top face,
    densities U[0.1,1] seed 3, filter R = 2.0, p = 3, 5-level GMG; 6 buckling modes."""
    return column((128, 64, 64), seed=3, lo=0.1, hi=1.0, rmin=2.0, penal=3.0, Emin=Emin,
                  mg_levels=5, name="column-128x64x64")


def vibration_160x80x80(Emin=1e-9) -> Problem:
    """Free vibration of the 160x80x80 cantilever (3,168,963 DOF): x = 0 face clamped,
    densities U[0.05,1] seed 5 through the R = 1.5 filter, SIMP p = 3 on the stiffness, the
    mass scaled by V = rho (p = 1) with density 1, 5-level GMG; 12 lowest modes."""
    pb = cantilever((160, 80, 80), seed=5, lo=0.05, hi=1.0, rmin=1.5, penal=3.0, Emin=Emin,
                    mg_levels=5, name="cantilever-160x80x80")
    pb.extra.update(density=1.0, nev=12)
    return pb


def batch_instances(n=32, sizes=tuple(range(16, 97, 2)), seed=4, nrhs=4, lo=0.05, hi=1.0,
                    Emin=1e-9):
    """Independent cantilevers of uneven size: instance i has nx_i drawn from `sizes`
    (SplitMix64(seed)), ny = nz = nx_i / 2, densities U[lo, hi) from SplitMix64(seed * 1000 + i),
    the x = 0 face clamped, and nrhs load cases: unit point loads at nodes of the x = max face
    drawn from SplitMix64(seed * 1000 + 500 + i), in directions (u0, u1, u2) / |u| with
    u_d uniform in [-1, 1) from the same stream."""
    pick = splitmix64(seed, n)
    out = []
    for i in range(n):
        nx = int(sizes[int(pick[i] % np.uint64(len(sizes)))])
        nel = (nx, nx // 2, nx // 2)
        pb = cantilever(nel, seed=seed * 1000 + i, lo=lo, hi=hi, Emin=Emin, mg_levels=0,
                        name=f"cantilever-{nel[0]}x{nel[1]}x{nel[2]}")
        r = uniform(seed * 1000 + 500 + i, 5 * nrhs)
        F = np.zeros((nrhs, pb.ndof))
        for k in range(nrhs):
            j = min(int(r[5 * k] * (nel[1] + 1)), nel[1])
            kz = min(int(r[5 * k + 1] * (nel[2] + 1)), nel[2])
            d = 2.0 * r[5 * k + 2:5 * k + 5] - 1.0
            d /= np.linalg.norm(d)
            F[k, 3 * node_index(nel, nx, j, kz):3 * node_index(nel, nx, j, kz) + 3] = d
        pb.F = F
        out.append(pb)
    return out


def cantilevers_batch32(nrhs=4) -> list:
    """32 cantilevers with nx in {16, 18, ..., 96} (seed 4), ny = nz = nx / 2, densities
    U[0.05,1], p = 3, 4 unit point-load cases each on the x = max face."""
    return batch_instances(32, tuple(range(16, 97, 2)), seed=4, nrhs=nrhs)

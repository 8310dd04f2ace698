"""Synthetic workloads: counter-based random streams and the Kuhn-split cube.

Morgana reads its meshes from gmsh 2.1 files (loadMesh.hpp:2065-2350): node
gid = 1-based file order, tet gid = 1-based order among type-4 elements,
geoId = second tag.  The generator below produces exactly such a mesh in
memory (and can write it as a gmsh 2.1 file), so that the numbering the
library and the oracle see is the one the reference loader would produce.

Random streams
--------------
Every random quantity has its own counter-based SplitMix64 stream, so any
entry can be regenerated independently of the others:

    key(seed, sid) = mix(seed * G + sid * H + 1)
    raw_i          = mix(key + (i + 1) * G)
    uniform_i      = (raw_i >> 11) * 2**-53                        in [0, 1)
    normal_i       = sqrt(-2 ln(1 - u_2i)) * cos(2 pi u_2i+1)      (Box-Muller)

with mix the SplitMix64 output function, G = 0x9E3779B97F4A7C15 and
H = 0xD1B54A32D192ED03.  Stream ids: 1 jitter, 2 k, 3 f, 4 rho, 5 mu,
6 density, 7 u_obs, 8 phi perturbation.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

G64 = np.uint64(0x9E3779B97F4A7C15)
H64 = np.uint64(0xD1B54A32D192ED03)

STREAM_JITTER = 1
STREAM_K = 2
STREAM_F = 3
STREAM_RHO = 4
STREAM_MU = 5
STREAM_DENSITY = 6
STREAM_UOBS = 7
STREAM_PHI = 8


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


def stream_raw(seed: int, sid: int, count: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        key = _mix(np.array([np.uint64(seed) * G64 + np.uint64(sid) * H64 + np.uint64(1)]))[0]
        ctr = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        return _mix(key + ctr * G64)


def stream_uniform(seed: int, sid: int, count: int, start: int = 0) -> np.ndarray:
    r = stream_raw(seed, sid, count, start)
    return (r >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def stream_normal(seed: int, sid: int, count: int) -> np.ndarray:
    u = stream_uniform(seed, sid, 2 * count)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


# Kuhn split: 6 tets per cube along the (0,0,0)->(1,1,1) diagonal; a tet is the
# monotone path p0 -> p0+e_a -> +e_b -> +e_c for a permutation (a,b,c).  Odd
# permutations give det < 0 and get their last two vertices swapped, because the
# reference integrates with the signed Jacobian and its loader asserts det > 0
# (meshDoctor3d.hpp:195-227).
KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_ODD = (False, True, True, False, False, True)


@dataclasses.dataclass
class TetMesh:
    n: int
    xyz: np.ndarray  # (V, 3) float64
    tets: np.ndarray  # (T, 4) int32, 1-based node gids
    btri: np.ndarray  # (B, 3) int32, 1-based boundary triangles
    bgeo: np.ndarray  # (B,) int32 geoIds: 1 x=0, 2 x=1, 3 y=0, 4 y=1, 5 z=0, 6 z=1

    @property
    def nV(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nT(self) -> int:
        return int(self.tets.shape[0])


def kuhn_cube(n: int, jitter: float = 0.0, seed: int = 0) -> TetMesh:
    """Unit cube, n^3 cells, 6 positively oriented Kuhn tets per cell.

    Interior nodes move by U(-jitter*h, +jitter*h) per coordinate (stream 1,
    counter 3*node + axis).  Node gid = 1 + i + (n+1)(j + (n+1)k).
    """
    h = 1.0 / n
    m = n + 1
    k3, j3, i3 = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    ijk = np.stack([i3.ravel(), j3.ravel(), k3.ravel()], axis=1)
    xyz = ijk.astype(np.float64) * h
    if jitter:
        interior = np.all((ijk > 0) & (ijk < n), axis=1)
        u = stream_uniform(seed, STREAM_JITTER, 3 * m ** 3).reshape(-1, 3)
        d = (2.0 * u - 1.0) * (jitter * h)
        xyz[interior] += d[interior]
    ck, cj, ci = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (ci + m * (cj + m * ck)).ravel().astype(np.int64)
    step = np.array([1, m, m * m], dtype=np.int64)
    tets = np.empty((base.size, 6, 4), dtype=np.int64)
    for t, (a, b, c) in enumerate(KUHN_PERMS):
        p0 = base
        p1 = p0 + step[a]
        p2 = p1 + step[b]
        p3 = p2 + step[c]
        if _ODD[t]:
            p2, p3 = p3, p2
        tets[:, t, 0], tets[:, t, 1], tets[:, t, 2], tets[:, t, 3] = p0, p1, p2, p3
    tets = (tets.reshape(-1, 4) + 1).astype(np.int32)
    btri, bgeo = _kuhn_boundary(n)
    return TetMesh(n=n, xyz=np.ascontiguousarray(xyz), tets=tets, btri=btri, bgeo=bgeo)


def _kuhn_boundary(n: int):
    """Boundary triangles of the Kuhn cube: each boundary square splits along
    the diagonal from its low corner to its high corner (the face of the tets
    touching it)."""
    m = n + 1

    def vid(i, j, k):
        return i + m * (j + m * k)

    tris, geos = [], []
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a, b = a.ravel(), b.ravel()
    for axis in range(3):
        for side in range(2):
            geo = 2 * axis + 1 + side
            c = 0 if side == 0 else n
            # the two in-face axes, in increasing order
            u_ax, v_ax = [ax for ax in range(3) if ax != axis]

            def P(du, dv):
                idx = [None, None, None]
                idx[axis] = np.full_like(a, c)
                idx[u_ax] = a + du
                idx[v_ax] = b + dv
                return vid(idx[0], idx[1], idx[2])

            p00, p10, p11, p01 = P(0, 0), P(1, 0), P(1, 1), P(0, 1)
            tris.append(np.stack([p00, p10, p11], axis=1))
            tris.append(np.stack([p00, p11, p01], axis=1))
            geos.append(np.full(2 * a.size, geo))
    btri = (np.concatenate(tris) + 1).astype(np.int32)
    bgeo = np.concatenate(geos).astype(np.int32)
    return btri, bgeo


def write_gmsh22(path: str, mesh: TetMesh) -> None:
    """gmsh 2.1/2.2 ASCII writer with the semantics loadMesh.hpp:2065-2350 reads:
    boundary triangles (type 2) then tets (type 4), tags (physical, geoId)."""
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.1 0 8\n$EndMeshFormat\n$Nodes\n")
        f.write(f"{mesh.nV}\n")
        for i, (x, y, z) in enumerate(mesh.xyz):
            f.write(f"{i + 1} {x:.17g} {y:.17g} {z:.17g}\n")
        f.write("$EndNodes\n$Elements\n")
        f.write(f"{mesh.btri.shape[0] + mesh.nT}\n")
        k = 1
        for t, g in zip(mesh.btri, mesh.bgeo):
            f.write(f"{k} 2 2 {g} {g} {t[0]} {t[1]} {t[2]}\n")
            k += 1
        for t in mesh.tets:
            f.write(f"{k} 4 2 1 1 {t[0]} {t[1]} {t[2]} {t[3]}\n")
            k += 1
        f.write("$EndElements\n")


def read_gmsh22(path: str) -> TetMesh:
    """Reads what write_gmsh22 writes (and gmsh 2.x ASCII tet meshes in general):
    node gid = file order, tets in file order, geoId = second tag."""
    with open(path) as f:
        lines = f.read().split("\n")
    i = lines.index("$Nodes")
    nv = int(lines[i + 1])
    xyz = np.array([[float(v) for v in lines[i + 2 + r].split()[1:4]] for r in range(nv)])
    i = lines.index("$Elements")
    ne = int(lines[i + 1])
    tets, btri, bgeo = [], [], []
    for r in range(ne):
        p = [int(v) for v in lines[i + 2 + r].split()]
        etype, ntag = p[1], p[2]
        tags, nodes = p[3:3 + ntag], p[3 + ntag:]
        if etype == 4:
            tets.append(nodes[:4])
        elif etype == 2:
            btri.append(nodes[:3])
            bgeo.append(tags[1] if ntag > 1 else tags[0])
    n = round(nv ** (1.0 / 3.0)) - 1
    return TetMesh(n=n, xyz=xyz, tets=np.array(tets, dtype=np.int32),
                   btri=np.array(btri, dtype=np.int32).reshape(-1, 3),
                   bgeo=np.array(bgeo, dtype=np.int32))


# Closed-form sizes of the Kuhn cube (exhaustive enumeration, n = 4..20).
def kuhn_counts(n: int) -> dict:
    return {
        "T": 6 * n ** 3,
        "V": (n + 1) ** 3,
        "E": 3 * n * (n + 1) ** 2 + 3 * n ** 2 * (n + 1) + n ** 3,
        "nnz_p1": 15 * n ** 3 + 21 * n ** 2 + 9 * n + 1,
        "nnz_p2": 230 * n ** 3 + 138 * n ** 2 + 24 * n + 1,
        "nnz_p1_int": 15 * n ** 3 - 69 * n ** 2 + 105 * n - 53,
        "nnz_p2_int": 230 * n ** 3 - 690 * n ** 2 + 732 * n - 273,
    }


@dataclasses.dataclass
class Problem:
    """One of the five benchmark workloads, fully materialised on the host."""
    name: str
    order: int
    mesh: TetMesh
    k: np.ndarray | None  # per-element stiffness coefficient (None = 1)
    mass: bool  # add the P1 mass matrix (K + M)
    f: np.ndarray  # nodal right-hand side on the full dof numbering
    dirichlet_geo: tuple


def config_problem(cfg: int, n: int | None = None, seed: int = 0) -> Problem:
    """Workload of configuration cfg (1, 2, 3, 5) at its size, or at size n."""
    if cfg == 1:
        n = n or 20
        mesh = kuhn_cube(n)
        nd = mesh.nV
        f = 2.0 * stream_uniform(seed, STREAM_F, nd) - 1.0
        return Problem("p1_stiff_mass_cube%d" % n, 1, mesh, None, True, f, (1, 2, 3, 4, 5, 6))
    if cfg in (2, 3):
        n = n or 64
        mesh = kuhn_cube(n, jitter=0.2, seed=seed)
        k = np.exp(stream_normal(seed, STREAM_K, mesh.nT))
        nd = mesh.nV + kuhn_counts(n)["E"]
        f = 2.0 * stream_uniform(seed, STREAM_F, nd) - 1.0
        return Problem("p2_diffusion_cube%d" % n, 2, mesh, k, False, f, (1, 2, 3, 4, 5, 6))
    if cfg == 5:
        n = n or 128
        mesh = kuhn_cube(n, jitter=0.2, seed=seed)
        k = np.exp(stream_normal(seed, STREAM_K, mesh.nT))
        f = 2.0 * stream_uniform(seed, STREAM_F, mesh.nV) - 1.0
        return Problem("p1_diffusion_cube%d" % n, 1, mesh, k, False, f, (1, 2, 3, 4, 5, 6))
    raise ValueError("no linear-solve workload for config %r" % cfg)


@dataclasses.dataclass
class DriftDiffusionProblem:
    """Configuration 4: P1 Poisson coefficients, Scharfetter-Gummel mobility and
    a state (phi, n) on a jittered Kuhn cube.  phi is prescribed (= x) on the
    faces x = 0 and x = 1 (geoIds 1, 2); the caller applies it with the
    library's Dirichlet mask (fix_state)."""
    name: str
    mesh: TetMesh
    eps: np.ndarray   # per element, U[1, 5]
    rho: np.ndarray   # per element, U[-1, 1]
    mu: np.ndarray    # per element, exp(N(0, 0.5^2))
    phi: np.ndarray   # per vertex, x + 0.1 U[-1, 1]
    dens: np.ndarray  # per vertex, exp(N(0, 1))
    dirichlet_geo: tuple = (1, 2)

    def fix_state(self, fixed: np.ndarray) -> None:
        """phi = x on the vertices the Dirichlet mask marks."""
        sel = np.asarray(fixed) != 0
        self.phi[sel] = self.mesh.xyz[sel, 0]


def config_drift_diffusion(n: int = 96, seed: int = 0) -> DriftDiffusionProblem:
    """Workload of configuration 4 at size n (96 by default)."""
    mesh = kuhn_cube(n, jitter=0.2, seed=seed)
    eps = 1.0 + 4.0 * stream_uniform(seed, STREAM_K, mesh.nT)
    rho = 2.0 * stream_uniform(seed, STREAM_RHO, mesh.nT) - 1.0
    mu = np.exp(0.5 * stream_normal(seed, STREAM_MU, mesh.nT))
    dens = np.exp(stream_normal(seed, STREAM_DENSITY, mesh.nV))
    phi = mesh.xyz[:, 0] + 0.1 * (2.0 * stream_uniform(seed, STREAM_PHI, mesh.nV) - 1.0)
    return DriftDiffusionProblem("p1_drift_diffusion_cube%d" % n, mesh, eps, rho, mu, phi, dens)

"""Synthetic workloads of the assembly path (meshes generated in code, fields from the counter RNG, seed 0).

  poisson_quad4(n=64)     -Lap u = 2 pi^2 sin(pi x) sin(pi y) on the unit square, n x n quad4, u = 0 on all edges
  poisson_hex8(n=128)     unit cube, n^3 hex8, f = 1, unit conductivity, u ~ U[-1, 1] per node, no Dirichlet DOFs
  neohookean_hex8(n=96)   unit cube, n^3 hex8, lambda = 1, mu = 0.5,
                          u = 0.05 sum_{k=1..4} a_k sin(k pi x) sin(k pi y) sin(k pi z), a_k ~ N(0, I3)
                          drawn with the first field seed whose continuous field has det(I + grad u) > 0.2
                          everywhere (seed 33; seed 0 inverts 7% of the elements, where ln J is undefined)
  j2_hex8(n=64)           unit cube, n^3 hex8, E = 200, nu = 0.3, sigma_y = 0.25, H = 2; 20 load steps of a
                          piecewise-linear random-direction macroscopic strain path (|eps| <= 0.01)
  neohookean_tet10(n=80)  unit cube, n^3 x 6 Kuhn tet10, interior vertices jittered +-0.2h (seed 0), NH as above
                          with amplitude 0.1 (field seed 1585 by the same rule; seed 0 inverts 28% of the elements)
"""
from dataclasses import dataclass, field

import numpy as np

from . import rng
from .mesh import StructuredMesh, kuhn_tet10_mesh, boundary_nodes
from .assembler import Poisson, NeoHookean, J2Plasticity


@dataclass
class Workload:
    name: str
    mesh: object
    physics: object
    nf: int
    U: np.ndarray                      # full field, NF x NN interleaved
    dirichlet_dofs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    extra: dict = field(default_factory=dict)

    @property
    def num_elements(self):
        return self.mesh.num_elements


FIELD_STREAM = 2
# first seeds (scanning 0, 1, 2, ...) whose sine field keeps det(I + grad u) > 0.2 on a 41^3 grid of the cube
FIELD_SEED = {0.05: 33, 0.1: 1585}


def sine_field_coefficients(seed, kmax=4):
    return rng.normal(seed, FIELD_STREAM, kmax * 3).reshape(kmax, 3)


def sine_field_min_det(amp, seed, pts=41):
    """min over a pts^3 grid of det(I + grad u) for the continuous sine field."""
    a = sine_field_coefficients(seed)
    g = np.linspace(0.0, 1.0, pts)
    X, Y, Z = (v.ravel() for v in np.meshgrid(g, g, g, indexing="ij"))
    H = np.zeros((X.size, 3, 3))
    for k in range(1, 5):
        kp = k * np.pi
        sx, sy, sz = np.sin(kp * X), np.sin(kp * Y), np.sin(kp * Z)
        cx, cy, cz = np.cos(kp * X), np.cos(kp * Y), np.cos(kp * Z)
        grad = np.stack([kp * cx * sy * sz, kp * sx * cy * sz, kp * sx * sy * cz], axis=1)
        H += amp * a[k - 1][None, :, None] * grad[:, None, :]
    return float(np.linalg.det(np.eye(3)[None] + H).min())


def _sine_field(X, amp, seed, stream=FIELD_STREAM, kmax=4):
    a = rng.normal(seed, stream, kmax * 3).reshape(kmax, 3)
    U = np.zeros_like(X)
    for k in range(1, kmax + 1):
        s = np.sin(k * np.pi * X[:, 0]) * np.sin(k * np.pi * X[:, 1]) * np.sin(k * np.pi * X[:, 2])
        U += amp * s[:, None] * a[k - 1][None, :]
    return U.ravel()


def poisson_quad4(n=64):
    mesh = StructuredMesh("quad", (0.0, 0.0), (1.0, 1.0), (n + 1, n + 1))
    nodes = np.unique(np.concatenate([mesh.nodeset_nodes[s] for s in ("bottom", "right", "top", "left")]))
    return Workload("poisson_quad4", mesh, Poisson("sines", 1.0), 1, np.zeros(mesh.num_nodes), nodes.astype(np.int64))


def poisson_hex8(n=128, seed=0):
    mesh = StructuredMesh("hex", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n + 1, n + 1, n + 1))
    U = 2.0 * rng.uniform(seed, 1, mesh.num_nodes) - 1.0
    return Workload("poisson_hex8", mesh, Poisson("constant", 1.0), 1, U)


def neohookean_hex8(n=96, seed=None, amp=0.05):
    seed = FIELD_SEED[amp] if seed is None else seed
    mesh = StructuredMesh("hex", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n + 1, n + 1, n + 1))
    return Workload("neohookean_hex8", mesh, NeoHookean(1.0, 0.5), 3, _sine_field(mesh.coords, amp, seed),
                    extra={"field_seed": seed})


def neohookean_tet10(n=80, seed=None, amp=0.1, mesh_seed=0):
    seed = FIELD_SEED[amp] if seed is None else seed
    mesh = kuhn_tet10_mesh(n, jitter=0.2, seed=mesh_seed)
    return Workload("neohookean_tet10", mesh, NeoHookean(1.0, 0.5), 3, _sine_field(mesh.coords, amp, seed),
                    extra={"field_seed": seed})


def j2_strain_path(seed=0, steps=20, max_norm=0.01, waypoints=4):
    """Macroscopic strain E(t_n), n = 1..steps: piecewise linear through 0 and `waypoints` random symmetric
    directions of norm max_norm (every E(t) has norm <= max_norm)."""
    g = rng.normal(seed, 4, waypoints * 6).reshape(waypoints, 6)
    W = [np.zeros((3, 3))]
    for w in g:
        S = np.array([[w[0], w[3], w[5]], [w[3], w[1], w[4]], [w[5], w[4], w[2]]])
        W.append(max_norm * S / np.linalg.norm(S))
    out = []
    per = steps // waypoints
    for n in range(1, steps + 1):
        seg = min((n - 1) // per, waypoints - 1)
        t = (n - seg * per) / per
        out.append((1 - t) * W[seg] + t * W[seg + 1])
    return out


def j2_perturbation(X, t, seed=0, amp=5e-4):
    """Smooth, boundary-nonvanishing perturbation amp * t * sum_k b_k cos(k pi x) cos(k pi y) cos(k pi z)."""
    b = rng.normal(seed, 5, 9).reshape(3, 3)
    U = np.zeros_like(X)
    for k in range(1, 4):
        c = np.cos(k * np.pi * X[:, 0]) * np.cos(k * np.pi * X[:, 1]) * np.cos(k * np.pi * X[:, 2])
        U += amp * t * c[:, None] * b[k - 1][None, :]
    return U


def j2_displacement(X, E, t, seed=0):
    return (X @ E.T + j2_perturbation(X, t, seed)).ravel()


def j2_hex8(n=64, seed=0, dirichlet=False):
    mesh = StructuredMesh("hex", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n + 1, n + 1, n + 1))
    path = j2_strain_path(seed)
    U = j2_displacement(mesh.coords, path[0], 1.0 / len(path), seed)
    wl = Workload("j2_hex8", mesh, J2Plasticity(), 3, U, extra={"path": path, "seed": seed})
    if dirichlet:
        nodes = boundary_nodes(mesh)
        wl.dirichlet_dofs = np.sort(np.concatenate([3 * (nodes - 1) + d + 1 for d in range(3)])).astype(np.int64)
    return wl


CONFIGS = {
    "poisson_quad4": poisson_quad4,
    "poisson_hex8": poisson_hex8,
    "neohookean_hex8": neohookean_hex8,
    "j2_hex8": j2_hex8,
    "neohookean_tet10": neohookean_tet10,
}

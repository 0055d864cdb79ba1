"""Meshes for the assembly path.

``StructuredMesh`` mirrors the reference's structured generator (src/meshes/StructuredMesh.jl:28-83):
``counts`` are NODE counts, node(i,j,k) = i + Nx(j-1) + NxNy(k-1), elements are ordered ex outer, ey, ez inner,
hex/quad local node order as at StructuredMesh.jl:116-123 / 246-249, coordinates via LinRange with the reference's
loop order (StructuredMesh.jl:433-484; exact for equal node counts per direction, like the reference).

``kuhn_tet10_mesh`` has no reference counterpart (the reference asserts on tets, StructuredMesh.jl:50-51): it splits
each cube of a structured grid into 6 Kuhn tetrahedra along the (0,0,0)-(1,1,1) diagonal, jitters interior vertices,
and places the 6 edge nodes at the edge midpoints after the jitter (straight-sided tet10, Exodus TETRA10 node order).
Every edge midpoint of this subdivision is a point of the 2x refined grid, so node ids are refined-grid ids.
"""
import numpy as np

from . import rng


def _linrange(a, b, n):
    # Base.LinRange getindex: lerpi(j, d, a, b) = (1 - j/d)*a + (j/d)*b
    t = np.arange(n, dtype=np.float64) / float(n - 1)
    return (1.0 - t) * a + t * b


class StructuredMesh:
    """StructuredMesh(element_type, mins, maxs, counts) with ``nodal_coords`` (nd x NN) and ``element_conns``."""

    def __init__(self, element_type, mins, maxs, counts):
        et = element_type.upper()
        if len(mins) != len(maxs) or len(maxs) != len(counts):
            raise ValueError("mins, maxs and counts must have equal length")
        for m1, m2 in zip(mins, maxs):
            if m1 >= m2:
                raise IndexError("dimension has negative or zero length")
        counts = tuple(int(c) for c in counts)
        if "HEX" in et:
            self.element_type = "HEX8"
            self.num_dims = 3
            self._hex8(mins, maxs, counts)
        elif "QUAD" in et:
            self.element_type = "QUAD4"
            self.num_dims = 2
            self._quad4(mins, maxs, counts)
        else:
            raise ValueError(f"Unsupported element type {element_type}")
        self.counts = counts

    def _quad4(self, mins, maxs, counts):
        Nx, Ny = counts
        xs = _linrange(mins[0], maxs[0], counts[0])
        ys = _linrange(mins[1], maxs[1], counts[1])
        # loops: ny over counts[1], nx over counts[2] (StructuredMesh.jl:476-477)
        ny, nx = np.meshgrid(np.arange(counts[0]), np.arange(counts[1]), indexing="ij")
        self.coords = np.stack([xs[nx.ravel()], ys[ny.ravel()]], axis=1)  # (NN, 2)
        ex, ey = np.meshgrid(np.arange(1, Nx), np.arange(1, Ny), indexing="ij")
        ex, ey = ex.ravel(), ey.ravel()
        node = lambda i, j: i + (j - 1) * Nx
        self.conn = np.stack([node(ex, ey), node(ex + 1, ey), node(ex + 1, ey + 1), node(ex, ey + 1)],
                             axis=1).astype(np.int64)
        self.nodeset_nodes = {
            "bottom": np.array([node(i, 1) for i in range(1, Nx + 1)], dtype=np.int64),
            "right": np.array([node(Nx, j) for j in range(1, Ny + 1)], dtype=np.int64),
            "top": np.array([node(i, Ny) for i in range(1, Nx + 1)], dtype=np.int64),
            "left": np.array([node(1, j) for j in range(1, Ny + 1)], dtype=np.int64),
        }

    def _hex8(self, mins, maxs, counts):
        Nx, Ny, Nz = counts
        xs = _linrange(mins[0], maxs[0], counts[0])
        ys = _linrange(mins[1], maxs[1], counts[1])
        zs = _linrange(mins[2], maxs[2], counts[2])
        # loops: nz over counts[1], ny over counts[2], nx over counts[3] (StructuredMesh.jl:438-440)
        nz, ny, nx = np.meshgrid(np.arange(counts[0]), np.arange(counts[1]), np.arange(counts[2]), indexing="ij")
        self.coords = np.stack([xs[nx.ravel()], ys[ny.ravel()], zs[nz.ravel()]], axis=1)
        ex, ey, ez = np.meshgrid(np.arange(1, Nx), np.arange(1, Ny), np.arange(1, Nz), indexing="ij")
        ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
        node = lambda i, j, k: i + Nx * (j - 1) + Nx * Ny * (k - 1)
        self.conn = np.stack([
            node(ex, ey, ez), node(ex + 1, ey, ez), node(ex + 1, ey + 1, ez), node(ex, ey + 1, ez),
            node(ex, ey, ez + 1), node(ex + 1, ey, ez + 1), node(ex + 1, ey + 1, ez + 1), node(ex, ey + 1, ez + 1)
        ], axis=1).astype(np.int64)
        I, K = np.meshgrid(np.arange(1, Nx + 1), np.arange(1, Nz + 1), indexing="ij")
        J2_, K2 = np.meshgrid(np.arange(1, Ny + 1), np.arange(1, Nz + 1), indexing="ij")
        I3, J3 = np.meshgrid(np.arange(1, Nx + 1), np.arange(1, Ny + 1), indexing="ij")
        # _three_dimensional_nsets (StructuredMesh.jl:452-470), vec of column-major comprehensions
        f = lambda a: a.ravel(order="F").astype(np.int64)
        self.nodeset_nodes = {
            "bottom": f(node(I, 1, K)), "right": f(node(Nx, J2_, K2)), "front": f(node(I3, J3, Nz)),
            "top": f(node(I, Ny, K)), "left": f(node(1, J2_, K2)), "back": f(node(I3, J3, 1)),
        }

    @property
    def nodal_coords(self):
        """nd x NN, the reference's H1Field layout."""
        return self.coords.T

    @property
    def element_conns(self):
        """nnpe x NE, 1-based, the reference's Connectivity layout."""
        return self.conn.T

    @property
    def num_nodes(self):
        return self.coords.shape[0]

    @property
    def num_elements(self):
        return self.conn.shape[0]


class TetMesh:
    def __init__(self, coords, conn, element_type="TET10"):
        self.coords = coords
        self.conn = conn
        self.element_type = element_type
        self.num_dims = 3
        self.nodeset_nodes = {}

    num_nodes = StructuredMesh.num_nodes
    num_elements = StructuredMesh.num_elements
    nodal_coords = StructuredMesh.nodal_coords
    element_conns = StructuredMesh.element_conns


class BlockMesh:
    """A mesh of several element blocks of one element type (the reference's FunctionSpace blocks, at most 16,
    src/FunctionSpaces.jl:63): ``blocks`` is an ordered list of (name, conn NE_b x NNPE, 1-based); elements are
    numbered block after block, so ``conn`` is the concatenation (Connectivity.data with per-block offsets,
    src/Fields.jl:129-188)."""

    def __init__(self, coords, blocks, element_type, nodeset_nodes=None, sideset_nodes=None):
        if not 1 <= len(blocks) <= 16:
            raise ValueError("a mesh holds 1..16 element blocks")
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.element_type = element_type
        self.num_dims = self.coords.shape[1]
        self.block_names = [b[0] for b in blocks]
        self.block_conns = [np.ascontiguousarray(b[1], dtype=np.int64) for b in blocks]
        self.conn = np.concatenate(self.block_conns, axis=0)
        self.nodeset_nodes = dict(nodeset_nodes or {})
        self.sideset_nodes = dict(sideset_nodes or {})

    @property
    def block_sizes(self):
        return [c.shape[0] for c in self.block_conns]

    num_nodes = StructuredMesh.num_nodes
    num_elements = StructuredMesh.num_elements


def split_into_blocks(mesh, labels, names=None):
    """The elements of `mesh` split into blocks by an integer label per element (blocks in ascending label order,
    element order kept within each block)."""
    labels = np.asarray(labels)
    keys = np.unique(labels)
    names = names or [f"block_{k}" for k in keys]
    blocks = [(nm, mesh.conn[labels == k]) for nm, k in zip(names, keys)]
    return BlockMesh(mesh.coords, blocks, mesh.element_type, getattr(mesh, "nodeset_nodes", {}),
                     getattr(mesh, "sideset_nodes", {}))


_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_PERM_SIGN = [1, -1, -1, 1, 1, -1]


def fan_tet10_mesh(m):
    """Tetrahedra around one vertex: a ring of m points in the z = 0 plane with an apex above and one below it, 2m TET10
    elements (Exodus TETRA10 node order). The centre vertex couples with 5m + 5 nodes: a row wider than the structured
    meshes' (165 column blocks at m = 32)."""
    V = [np.zeros(3), np.array([0.0, 0.0, 1.0]), np.array([0.0, 0.0, -1.0])]
    V += [np.array([np.cos(t), np.sin(t), 0.0]) for t in 2 * np.pi * np.arange(m) / m]
    tets = []
    for i in range(m):
        a, b = 3 + i, 3 + (i + 1) % m
        tets += [(0, a, b, 1), (0, b, a, 2)]  # both positively oriented
    mid = {}
    conn = []
    for t in tets:
        row = list(t)
        for a, b in [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]:  # TETRA10 edge midpoints
            key = (min(t[a], t[b]), max(t[a], t[b]))
            if key not in mid:
                mid[key] = len(V)
                V.append(0.5 * (V[key[0]] + V[key[1]]))
            row.append(mid[key])
        conn.append(row)
    return TetMesh(np.array(V), np.array(conn, dtype=np.int64) + 1)


def kuhn_tet10_mesh(n, jitter=0.2, seed=0, stream=7):
    """Unit cube, n^3 cubes x 6 Kuhn tets, interior vertices jittered U[-jitter*h, jitter*h] per coordinate."""
    h = 1.0 / n
    m = 2 * n + 1  # refined grid nodes per direction
    # vertex coordinates (base grid), jittered in the interior
    vi, vj, vk = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    vi, vj, vk = vi.ravel(order="F"), vj.ravel(order="F"), vk.ravel(order="F")  # i fastest
    vx = np.stack([vi * h, vj * h, vk * h], axis=1).astype(np.float64)
    interior = (vi > 0) & (vi < n) & (vj > 0) & (vj < n) & (vk > 0) & (vk < n)
    u = rng.uniform(seed, stream, 3 * vx.shape[0]).reshape(-1, 3)
    vx = vx + interior[:, None] * (jitter * h * (2.0 * u - 1.0))
    vid = lambda i, j, k: i + (n + 1) * (j + (n + 1) * k)
    rid = lambda I, J, K: I + m * (J + m * K)
    coords = np.zeros((m ** 3, 3))
    # every refined point (I,J,K): odd coordinates mark the edge direction; endpoints differ by -1/+1 there
    I, J, K = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    I, J, K = I.ravel(order="F"), J.ravel(order="F"), K.ravel(order="F")
    oi, oj, ok = I % 2, J % 2, K % 2
    a = vid((I - oi) // 2, (J - oj) // 2, (K - ok) // 2)
    b = vid((I + oi) // 2, (J + oj) // 2, (K + ok) // 2)
    coords[rid(I, J, K)] = 0.5 * (vx[a] + vx[b])
    # elements: cubes ex outer, ey, ez inner; 6 tets per cube
    ex, ey, ez = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    conn = np.zeros((n ** 3, 6, 10), dtype=np.int64)
    base = np.stack([2 * ex, 2 * ey, 2 * ez], axis=1)
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    for t, (p, sgn) in enumerate(zip(_PERMS, _PERM_SIGN)):
        e0 = np.zeros(3, dtype=np.int64); e0[p[0]] = 2
        e1 = e0.copy(); e1[p[1]] = 2
        verts = [np.zeros(3, dtype=np.int64), e0, e1, np.array([2, 2, 2])]
        if sgn < 0:
            verts[1], verts[2] = verts[2], verts[1]
        pts = [base + v for v in verts] + [base + (verts[i] + verts[j]) // 2 for i, j in edges]
        for c, pt in enumerate(pts):
            conn[:, t, c] = rid(pt[:, 0], pt[:, 1], pt[:, 2]) + 1
    return TetMesh(coords, conn.reshape(-1, 10))


def boundary_nodes(mesh, tol=1e-12):
    """1-based ids of nodes on the unit-cube (or unit-square) boundary."""
    X = mesh.coords
    on = np.zeros(X.shape[0], dtype=bool)
    for d in range(X.shape[1]):
        on |= (np.abs(X[:, d] - X[:, d].min()) < tol) | (np.abs(X[:, d] - X[:, d].max()) < tol)
    return np.nonzero(on)[0].astype(np.int64) + 1

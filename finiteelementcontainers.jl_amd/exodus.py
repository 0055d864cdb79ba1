"""Exodus II meshes and results (NetCDF classic), the FileMesh{ExodusDatabase} interface of the reference
(src/meshes/Exodus.jl:18-141) and its PostProcessor output (src/meshes/Exodus.jl:148-290), on the host.

Reading follows the reference's conventions:
- nodal coordinates in the file's own (positional) node numbering;
- connectivity 1-based in that numbering, one matrix per element block, keyed by block name;
- the node and element id maps kept aside;
- node sets sorted;
- side sets as (element, side) pairs plus the nodes of those sides. The side-node tables are the Exodus II
  conventions for QUAD4, HEX8 and TET10.

Writing produces a minimal Exodus II results file (coordinates, one element block, time steps, nodal variables) in
the layout exodiff and the Exodus readers expect. NetCDF I/O is scipy.io.netcdf_file (CDF-1/CDF-2).
"""
import numpy as np
from scipy.io import netcdf_file

# Exodus element type names -> element kind of this package
_ELEM_TYPES = {"QUAD": "QUAD4", "QUAD4": "QUAD4", "HEX": "HEX8", "HEX8": "HEX8", "TETRA10": "TET10", "TET10": "TET10"}

# Exodus II side -> local nodes (1-based), per element type
_SIDES = {
    "QUAD4": {1: (1, 2), 2: (2, 3), 3: (3, 4), 4: (4, 1)},
    "HEX8": {1: (1, 2, 6, 5), 2: (2, 3, 7, 6), 3: (3, 4, 8, 7), 4: (1, 5, 8, 4), 5: (1, 4, 3, 2), 6: (5, 6, 7, 8)},
    "TET10": {1: (1, 2, 4, 5, 9, 8), 2: (2, 3, 4, 6, 10, 9), 3: (1, 4, 3, 8, 10, 7), 4: (1, 3, 2, 7, 6, 5)},
}


def _names(f, var, n):
    if var not in f.variables:
        return [""] * n
    return [b"".join(row).strip(b"\x00").decode() for row in f.variables[var][:]]


class ExodusMesh:
    """One Exodus mesh file read into host arrays (FileMesh(ExodusMesh, file_name), src/meshes/Exodus.jl:18-21).

    Attributes: coords (NN x ND), node_id_map, elem_id_map, blocks {name: (element type, conn NE x NNPE 1-based,
    block element id map)}, nodeset_nodes {name: sorted nodes}, sideset_elems / sideset_sides {name: sorted by
    element}, sideset_nodes {name: unique sorted nodes}."""

    def __init__(self, file_name):
        self.file_name = file_name
        with netcdf_file(file_name, "r", mmap=False) as f:
            dims = f.dimensions
            nd = int(dims["num_dim"])
            self.num_dims = nd
            axes = ["coordx", "coordy", "coordz"][:nd]
            if all(a in f.variables for a in axes):
                self.coords = np.stack([np.asarray(f.variables[a][:], dtype=np.float64) for a in axes], axis=1)
            else:  # older files store one coord (num_dim, num_nodes) array
                self.coords = np.asarray(f.variables["coord"][:], dtype=np.float64).T.copy()
            nn = self.coords.shape[0]
            self.node_id_map = (np.asarray(f.variables["node_num_map"][:], dtype=np.int64)
                                if "node_num_map" in f.variables else np.arange(1, nn + 1))
            nblk = int(dims.get("num_el_blk", 0))
            bnames = _names(f, "eb_names", nblk)
            ids = np.asarray(f.variables["eb_prop1"][:], dtype=np.int64) if "eb_prop1" in f.variables else \
                np.arange(1, nblk + 1)
            self.blocks = {}
            start = 0
            for b in range(nblk):
                v = f.variables[f"connect{b + 1}"]
                etype = v._attributes.get("elem_type", b"").decode().upper()
                conn = np.asarray(v[:], dtype=np.int64)
                name = bnames[b] or f"block_{ids[b]}"
                self.blocks[name] = (_ELEM_TYPES.get(etype, etype), conn, np.arange(start + 1, start + len(conn) + 1))
                start += len(conn)
            ne = start
            self.elem_id_map = (np.asarray(f.variables["elem_num_map"][:], dtype=np.int64)
                                if "elem_num_map" in f.variables else np.arange(1, ne + 1))
            nns = int(dims.get("num_node_sets", 0))
            nsnames = _names(f, "ns_names", nns)
            nsids = np.asarray(f.variables["ns_prop1"][:], dtype=np.int64) if nns else []
            self.nodeset_nodes = {}
            for i in range(nns):
                name = nsnames[i] or f"nodeset_{nsids[i]}"
                self.nodeset_nodes[name] = np.sort(np.asarray(f.variables[f"node_ns{i + 1}"][:], dtype=np.int64))
            nss = int(dims.get("num_side_sets", 0))
            ssnames = _names(f, "ss_names", nss)
            ssids = np.asarray(f.variables["ss_prop1"][:], dtype=np.int64) if nss else []
            self.sideset_elems, self.sideset_sides, self.sideset_nodes = {}, {}, {}
            for i in range(nss):
                name = ssnames[i] or f"sideset_{ssids[i]}"
                el = np.asarray(f.variables[f"elem_ss{i + 1}"][:], dtype=np.int64)
                sd = np.asarray(f.variables[f"side_ss{i + 1}"][:], dtype=np.int64)
                perm = np.argsort(el, kind="stable")  # sortperm(sset.elements), src/meshes/Exodus.jl:87-90
                self.sideset_elems[name], self.sideset_sides[name] = el[perm], sd[perm]
                self.sideset_nodes[name] = self._side_nodes(el, sd)

    def _side_nodes(self, elems, sides):
        nodes = []
        for e, s in zip(elems, sides):
            etype, conn, local = self._element(int(e))
            nodes.extend(conn[local, np.asarray(_SIDES[etype][int(s)]) - 1])
        return np.unique(np.asarray(nodes, dtype=np.int64))

    def _element(self, e):
        """Block type, connectivity and row of global (1-based, file order) element e."""
        for etype, conn, ids in self.blocks.values():
            if ids[0] <= e <= ids[-1]:
                return etype, conn, e - ids[0]
        raise IndexError(e)

    @property
    def num_nodes(self):
        return self.coords.shape[0]


class UnstructuredMesh:
    """One element block of an Exodus mesh as a mesh for FunctionSpace / SparseMatrixAssembler (the reference's
    UnstructuredMesh over a FileMesh, src/meshes/UnstructuredMesh.jl). Node numbering is the file's."""

    def __init__(self, file_name, block=None):
        exo = ExodusMesh(file_name)
        if block is None:
            if len(exo.blocks) != 1:
                raise ValueError(f"{file_name} has blocks {list(exo.blocks)}; name one")
            block = next(iter(exo.blocks))
        etype, conn, _ = exo.blocks[block]
        if etype not in ("QUAD4", "HEX8", "TET10"):
            raise ValueError(f"element type {etype} is not assembled on the GPU")
        self.exodus = exo
        self.element_type = etype
        self.num_dims = exo.num_dims
        self.coords = exo.coords
        self.conn = conn
        self.nodeset_nodes = dict(exo.nodeset_nodes)
        self.sideset_nodes = dict(exo.sideset_nodes)

    @property
    def num_nodes(self):
        return self.coords.shape[0]

    @property
    def num_elements(self):
        return self.conn.shape[0]


def write_exodus(file_name, mesh, nodal_fields=None, times=None, title="fecx results"):
    """Exodus II results file (PostProcessor + write_times/write_field, src/meshes/Exodus.jl:148-290): coordinates,
    one element block, and nodal variables {name: array (ntimes, nn) or (nn,)} at `times`."""
    nodal_fields = nodal_fields or {}
    coords = np.asarray(mesh.coords, dtype=np.float64)
    conn = np.asarray(mesh.conn, dtype=np.int32)
    nn, nd = coords.shape
    ne, nnpe = conn.shape
    vals = {k: np.atleast_2d(np.asarray(v, dtype=np.float64)) for k, v in nodal_fields.items()}
    nt = max([v.shape[0] for v in vals.values()] + [len(times) if times is not None else 0])
    times = np.arange(nt, dtype=np.float64) if times is None else np.asarray(times, dtype=np.float64)
    etype = {"QUAD4": "QUAD4", "HEX8": "HEX8", "TET10": "TETRA10"}.get(mesh.element_type, mesh.element_type)

    def chars(strings, n):
        a = np.zeros((len(strings), n), dtype="S1")
        for i, s in enumerate(strings):
            b = s.encode()[:n - 1]
            a[i, :len(b)] = np.frombuffer(b, dtype="S1")
        return a

    with netcdf_file(file_name, "w", version=2) as f:
        f.api_version = np.float32(5.22)
        f.version = np.float32(5.22)
        f.floating_point_word_size = np.int32(8)
        f.file_size = np.int32(1)
        f.title = title
        f.createDimension("time_step", None)   # the record dimension comes first in this writer
        f.createDimension("len_string", 33)
        f.createDimension("len_line", 81)
        f.createDimension("four", 4)
        f.createDimension("len_name", 33)
        f.createDimension("num_dim", nd)
        f.createDimension("num_nodes", nn)
        f.createDimension("num_elem", ne)
        f.createDimension("num_el_blk", 1)
        f.createDimension("num_el_in_blk1", ne)
        f.createDimension("num_nod_per_el1", nnpe)
        tw = f.createVariable("time_whole", "d", ("time_step",))
        for i, ax in enumerate("xyz"[:nd]):
            v = f.createVariable(f"coord{ax}", "d", ("num_nodes",))
            v[:] = coords[:, i]
        cn = f.createVariable("coor_names", "c", ("num_dim", "len_name"))
        cn[:] = chars([c for c in "xyz"[:nd]], 33)
        st = f.createVariable("eb_status", "i", ("num_el_blk",))
        st[:] = [1]
        pr = f.createVariable("eb_prop1", "i", ("num_el_blk",))
        pr.name = "ID"
        pr[:] = [1]
        en = f.createVariable("eb_names", "c", ("num_el_blk", "len_name"))
        en[:] = chars(["block_1"], 33)
        c = f.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        c.elem_type = etype
        c[:] = conn
        nm = f.createVariable("node_num_map", "i", ("num_nodes",))
        nm[:] = np.arange(1, nn + 1, dtype=np.int32)
        em = f.createVariable("elem_num_map", "i", ("num_elem",))
        em[:] = np.arange(1, ne + 1, dtype=np.int32)
        if vals:
            f.createDimension("num_nod_var", len(vals))
            nv = f.createVariable("name_nod_var", "c", ("num_nod_var", "len_name"))
            nv[:] = chars(list(vals), 33)
            for i, (k, v) in enumerate(vals.items()):
                var = f.createVariable(f"vals_nod_var{i + 1}", "d", ("time_step", "num_nodes"))
                var[:v.shape[0], :] = v
        tw[:nt] = times[:nt]

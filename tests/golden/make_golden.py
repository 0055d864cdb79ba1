"""Writes tests/golden/poisson_gold_subset.npz from the reference's test/poisson/poisson.gold (data only).

poisson.gold is the Exodus result of the reference's Poisson regression test (test/poisson/TestPoisson.jl:1-22:
f = 2 pi^2 sin(pi x) sin(pi y), u = 0 on the four sidesets, 128 x 128 quad4 on the unit square). The fixture keeps
every 5th node: its coordinates and the gold nodal value u, read with scipy's NetCDF reader (no code from the
reference is executed or copied). Run once in a container that has the reference mounted:

    python tests/golden/make_golden.py /path/to/FiniteElementContainers.jl
"""
import os
import sys

import numpy as np
from scipy.io import netcdf_file


def main(ref_root):
    gold = os.path.join(ref_root, "test", "poisson", "poisson.gold")
    nc = netcdf_file(gold, "r", mmap=False)
    x = np.array(nc.variables["coordx"][:], dtype=np.float64)
    y = np.array(nc.variables["coordy"][:], dtype=np.float64)
    u = np.array(nc.variables["vals_nod_var1"][:], dtype=np.float64)
    u = u.reshape(-1, x.size)[-1]  # last time step
    nc.close()
    keep = np.arange(0, x.size, 5)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "poisson_gold_subset.npz")
    np.savez_compressed(out, x=x[keep], y=y[keep], u=u[keep], num_nodes=x.size, source="test/poisson/poisson.gold")
    print(f"wrote {out}: {keep.size} of {x.size} nodes, max u {u.max():.16g}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

"""Write wedge_example_mesh.npz: the mesh of the reference's shipped example
examples/wedge (wedge.cgns, Mach 2.5 over a 5 degree wedge), as this
package's CGNS reader (freecfd_amd/cgns.py, Grid::readCGNS's rules) returns
it: node coordinates, cells, and BC node sets in the order BC_1..BC_4.
Data only, so that the example's GPU test runs where the reference tree is
absent.  Usage: python tests/golden/make_wedge_example_mesh.py [REFERENCE_ROOT]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    from freecfd_amd import cgns
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "examples", "wedge", "wedge.cgns")
    raw = cgns.read_cgns(src, dimension=2)
    out = os.path.join(HERE, "wedge_example_mesh.npz")
    np.savez_compressed(out, xyz=np.asarray(raw["xyz"], np.float64), cell_ptr=np.asarray(raw["cell_ptr"], np.int32),
                        cell_nodes=np.asarray(raw["cell_nodes"], np.int32), bc_ptr=np.asarray(raw["bc_ptr"], np.int32),
                        bc_nodes=np.asarray(raw["bc_nodes"], np.int32), bc_names=np.array(raw["bc_names"]),
                        dimension=np.int32(raw["dimension"]))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()

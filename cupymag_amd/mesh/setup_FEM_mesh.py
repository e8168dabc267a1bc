"""FEM mesh handler (cupymag/mesh/setup_FEM_mesh.py:33-116).

Builds the grid named by the parameters (Hex grid or .nas tet mesh), the
periodic node -> DOF map and the defect DOFs, and keeps host (numpy) copies;
``*_cp`` device copies are torch tensors made on first use.  Unlike the
reference (class-level DefDOF / _mesh_computed flags), every instance owns
its mesh, so several simulations can live in one process.
"""
import numpy as np

from ..utils.defect_shapes import get_defect_shape_function
from ..utils.final_assembly import extract_defect_dofs
from .gridHex import HexGrid
from .gridTet import TetraGrid
from .periodic import structured_dims


class FEMMesh:
    def __init__(self, P, mesh_arrays=None):
        """``mesh_arrays = (coords, tets)`` feeds a tet mesh held in memory
        (the arrays read_nas would return) instead of ``P.mesh_file``."""
        self.P = P
        self.grid_type = P.grid_type
        if self.grid_type == "Hex":
            self.fem_grid = HexGrid(P.Nx, P.Ny, P.Nz, P.Ndx, P.Ndy, P.Ndz, P.defect_center)
        elif self.grid_type == "Tet":
            shape_fn = get_defect_shape_function(P.defect_shape, P.shape_params)
            self.fem_grid = TetraGrid(None if mesh_arrays is not None else P.mesh_file, shape_fn,
                                      P.defect_center, arrays=mesh_arrays)
        else:
            raise NotImplementedError(f"Grid type '{self.grid_type}' is not supported.")
        self.node_coords_np, self.elements_np = self.fem_grid.std_fem_mesh()
        self.global_id_np = self.fem_grid.build_periodic_node_map(self.node_coords_np)
        self.n_dof = int(self.global_id_np.max()) + 1
        self.n_nodes = self.node_coords_np.shape[0]
        self.n_elements = self.elements_np.shape[0]
        self.DefDOF = extract_defect_dofs(self.elements_np, self.global_id_np)
        self.grid_dims = structured_dims(self.node_coords_np, self.global_id_np)

    def get_mesh_data(self):
        return {"node_coords_np": self.node_coords_np, "elements_np": self.elements_np,
                "global_id_np": self.global_id_np, "n_dof": self.n_dof, "n_nodes": self.n_nodes,
                "n_elements": self.n_elements}

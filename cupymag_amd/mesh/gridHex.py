"""Structured hexahedral grid (cupymag/mesh/gridHex.py:68-213).

Unit cubes on [0,Nx]x[0,Ny]x[0,Nz]; node id i*(Ny+1)*(Nz+1) + j*(Nz+1) + k;
elements in i-major order (k fastest) with corner order
(i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k) then the same at k+1; defect flag =
cell centre inside the half-open box [c - Nd/2, c + Nd/2) (c defaults to
the domain centre).  Built with array operations instead of the
reference's per-element Python loop.
"""
import numpy as np

from .periodic import periodic_node_map


class HexGrid:
    def __init__(self, Nx, Ny, Nz, Ndx, Ndy, Ndz, defect_center=None):
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.Ndx, self.Ndy, self.Ndz = Ndx, Ndy, Ndz
        self.defect_center = defect_center
        self.x = np.linspace(0, Nx, Nx + 1)
        self.y = np.linspace(0, Ny, Ny + 1)
        self.z = np.linspace(0, Nz, Nz + 1)

    def std_fem_mesh(self):
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        i3, j3, k3 = np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), np.arange(Nz + 1), indexing="ij")
        coords = np.stack((self.x[i3.ravel()], self.y[j3.ravel()], self.z[k3.ravel()]), axis=1)
        if self.defect_center is None:
            cx, cy, cz = Nx / 2.0, Ny / 2.0, Nz / 2.0
        else:
            cx, cy, cz = (float(v) for v in self.defect_center)
        i, j, k = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()

        def nid(a, b, c):
            return a * (Ny + 1) * (Nz + 1) + b * (Nz + 1) + c

        el = np.zeros((Nx * Ny * Nz, 9), dtype=np.int32)
        el[:, 0] = nid(i, j, k)
        el[:, 1] = nid(i + 1, j, k)
        el[:, 2] = nid(i + 1, j + 1, k)
        el[:, 3] = nid(i, j + 1, k)
        el[:, 4] = nid(i, j, k + 1)
        el[:, 5] = nid(i + 1, j, k + 1)
        el[:, 6] = nid(i + 1, j + 1, k + 1)
        el[:, 7] = nid(i, j + 1, k + 1)
        xc, yc, zc = i + 0.5, j + 0.5, k + 0.5
        el[:, 8] = ((xc >= cx - self.Ndx / 2.0) & (xc < cx + self.Ndx / 2.0) &
                    (yc >= cy - self.Ndy / 2.0) & (yc < cy + self.Ndy / 2.0) &
                    (zc >= cz - self.Ndz / 2.0) & (zc < cz + self.Ndz / 2.0)).astype(np.int32)
        return coords, el

    def build_periodic_node_map(self, node_coords, tol=1e-12):
        lo = (self.x[0], self.y[0], self.z[0])
        hi = (self.x[-1], self.y[-1], self.z[-1])
        return periodic_node_map(node_coords, lo, hi, tol)

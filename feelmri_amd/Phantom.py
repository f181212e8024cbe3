"""Drop-in for the reference's ``feelmri.Phantom.FEMPhantom``
(python/feelmri/Phantom.py:49-422) on one GPU, without meshio, pymetis or MPI.

* reading: ``mesh_io.load_phantom`` -- the time-series reader when the file
  holds one, the single mesh otherwise (Phantom.py:76-139); nodes scaled by
  ``scale_factor`` and cast to ``dtype``;
* distribution: the reference partitions the element dual graph with METIS
  over its MPI ranks (Phantom.py:194-239).  Here ``n_parts`` / ``part``
  select one part of ``feelmri_amd.partition``'s partition (1 / 0 by
  default: the whole mesh), with the same fields -- ``local_elements``,
  ``local_nodes``, ``local_to_global_nodes`` / ``_elems``, ``partitioning``;
* ``read_data(fr)``: a frame's node and cell data, cast to ``dtype``, the
  displacement and velocity scaled by ``scale_factor`` and the acceleration
  by its square (Phantom.py:250-272);
* ``create_submesh(markers)``, ``to_submesh``, ``orient`` / ``reorient``,
  ``bounding_box`` as in Phantom.py:141-192, :274-377;
* ``mass_matrix(lumped)`` / ``moving_mass_matrix``: the consistent P1 mass
  of the local part on the GPU (``Assemble.MassAssemble``), or its row sums
  as a diagonal (``Assemble.lumped_mass``, the GPU gather) -- Phantom.py:379-422.
  P1 tetrahedra only, like ``Assemble``.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import mesh_io, partition


class FEMPhantom:
    def __init__(self, path: str = "", scale_factor: float = 1.0, displacement_label: str = "displacement",
                 velocity_label: str = "velocity", acceleration_label: str = "acceleration",
                 pressure_label: str = "pressure", dtype=np.float32, n_parts: int = 1, part: int = 0,
                 partition_method: str = "rcb"):
        if not 0 <= part < n_parts:
            raise ValueError(f"part {part} outside [0, {n_parts})")
        self.path = path
        self.scale_factor = scale_factor
        self.displacement_label = displacement_label
        self.velocity_label = velocity_label
        self.acceleration_label = acceleration_label
        self.pressure_label = pressure_label
        self.dtype = dtype
        self.n_parts, self.part, self.partition_method = int(n_parts), int(part), partition_method
        ph = mesh_io.load_phantom(path, scale_factor=scale_factor, dtype=dtype)
        self.reader, self.Nfr = ph.reader, ph.n_frames
        self.point_data = ph.point_data if ph.reader is None else None
        self.cell_data = ph.cell_data if ph.reader is None else None
        self.cell_type = ph.cell_type
        self.global_elements = ph.elements
        self.global_nodes = ph.nodes
        self.global_shape = self.global_nodes.shape
        self.local_elements = ph.elements
        self.local_nodes = ph.nodes
        self.local_shape = self.global_nodes.shape
        self.bbox = self.bounding_box()
        self.distribute_mesh()

    def bounding_box(self):
        return np.min(self.global_nodes, axis=0), np.max(self.global_nodes, axis=0)

    def distribute_mesh(self):
        """This process's part of the element partition (Phantom.py:194-239)."""
        elems = np.asarray(self.global_elements)
        member = partition.partition_elements(elems, np.asarray(self.global_nodes, dtype=np.float64),
                                              self.n_parts, self.partition_method)
        lm = partition.local_mesh(elems, member, self.part)
        self.local_elements = lm.elems
        self.local_nodes = self.global_nodes[lm.nodes_l2g]
        self.local_shape = self.local_nodes.shape
        self.local_to_global_nodes = lm.nodes_l2g
        self.local_to_global_elems = lm.elems_l2g
        self.partitioning = np.asarray(member).reshape(-1, 1)

    def read_data(self, fr: int):
        """Node and cell data of frame ``fr`` (Phantom.py:250-272)."""
        if self.reader is None:
            raise ValueError(f"{self.path} holds no time series")
        _, self.point_data, self.cell_data = self.reader.read_data(fr)
        for key in self.point_data:
            self.point_data[key] = self.point_data[key].astype(self.dtype)
        for key, s in ((self.displacement_label, self.scale_factor), (self.velocity_label, self.scale_factor),
                       (self.acceleration_label, self.scale_factor ** 2), (self.pressure_label, 1.0)):
            if key in self.point_data:
                self.point_data[key] *= s

    def create_submesh(self, markers, refine: bool = False):
        """Keep the elements where ``markers`` is true and the nodes they use,
        renumbered; the original mesh is kept (Phantom.py:141-192)."""
        if refine:
            raise NotImplementedError("submesh refinement")
        sub = self.global_elements[np.asarray(markers, dtype=bool)]
        keep = np.unique(sub)
        mapped = -np.ones(self.global_nodes.shape[0], dtype=np.int64)
        mapped[keep] = np.arange(keep.size)
        self._global_nodes, self._global_elements, self._global_shape = (self.global_nodes, self.global_elements,
                                                                          self.global_shape)
        self.global_nodes = self.global_nodes[keep]
        self.global_elements = mapped[sub].astype(self.global_elements.dtype)
        self.mesh_to_submesh_nodes = keep
        self.global_shape = self.global_nodes.shape
        self.distribute_mesh()

    def to_submesh(self, data, global_mesh: bool = False):
        """Node data of the original mesh on the submesh nodes
        (Phantom.py:274-309)."""
        if not hasattr(self, "mesh_to_submesh_nodes"):
            raise ValueError("Submesh not created. Please create a submesh first using `create_submesh`.")
        if data.shape[0] != self._global_nodes.shape[0]:
            raise NotImplementedError("cell data to submesh conversion")
        idx = self.mesh_to_submesh_nodes if global_mesh else self.mesh_to_submesh_nodes[self.local_to_global_nodes]
        return data[idx, ...]

    def orient(self, MPS_ori, LOC):
        """Move into the slice frame: (x - LOC) @ MPS_ori (Phantom.py:358-366)."""
        R = np.asarray(MPS_ori).astype(self.dtype)
        loc = np.asarray(getattr(LOC, "m", LOC)).astype(self.dtype)
        self.global_nodes = (self.global_nodes - loc) @ R
        self.local_nodes = (self.local_nodes - loc) @ R

    def reorient(self, MPS_ori, LOC):
        """Undo ``orient``: x @ MPS_ori.T + LOC (Phantom.py:368-376)."""
        R = np.asarray(MPS_ori).astype(self.dtype)
        loc = np.asarray(getattr(LOC, "m", LOC)).astype(self.dtype)
        self.global_nodes = self.global_nodes @ R.T + loc
        self.local_nodes = self.local_nodes @ R.T + loc

    def _mass(self, nodes, lumped: bool):
        from . import Assemble

        if self.cell_type != "tetra":
            raise NotImplementedError(f"mass matrices of {self.cell_type} cells (P1 tetrahedra only)")
        if lumped:
            return Assemble.lumped_mass(self.local_elements, nodes)
        return Assemble.MassAssemble(self.local_elements, nodes)

    def mass_matrix(self, lumped: bool = False, use_submesh: bool = False, quadrature_order: int = 2):
        """The local part's P1 mass matrix (Phantom.py:379-399); lumped: its
        row sums on the diagonal."""
        return self._mass(self.local_nodes, lumped)

    def moving_mass_matrix(self, local_nodes, lumped: bool = False, use_submesh: bool = False,
                           quadrature_order: int = 2):
        """As ``mass_matrix`` at displaced local node positions (Phantom.py:401-422)."""
        return self._mass(np.asarray(local_nodes), lumped)


def lumped_weights(M) -> np.ndarray:
    """The diagonal of a lumped mass matrix as a vector."""
    return np.asarray(sp.csr_matrix(M).diagonal()).ravel()

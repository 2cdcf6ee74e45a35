"""Mesh set-up: the native quadtree-forest builder of liboceano_dg.so, seen from Python.

Replaces GridIn::read_msh + refine_global + MatrixFree's face lists (main.cc:617-663,
OceanDG.h:438-470) for the Cartesian meshes of the test cases.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib


class Mesh:
    """Axis-aligned quadrilateral mesh with conforming, hanging and boundary faces.

    Arrays (numpy, C-contiguous):
      geom      [n_cells, 4]  x0, y0, hx, hy
      level     [n_cells]     refinement level
      inner     [n_inner, 6]  cell_m, face_m, subface_m, cell_p, face_p, subface_p
      boundary  [n_bdry, 3]   cell, face, boundary_id
    """

    def __init__(self, handle):
        self._h = handle
        self._export()

    @classmethod
    def rectangle(cls, x0, x1, y0, y1, nx, ny, periodic=(False, False),
                  boundary_ids=(0, 0, 0, 0)):
        h = C.c_void_p()
        ids = (C.c_int32 * 4)(*boundary_ids)
        check(lib().oc_mesh_create_rectangle(x0, x1, y0, y1, nx, ny, int(periodic[0]),
                                             int(periodic[1]), ids, C.byref(h)),
              "oc_mesh_create_rectangle")
        return cls(h)

    @classmethod
    def read_msh(cls, path, periodic=(False, False)):
        """Gmsh 4.1 deck of a uniform Cartesian grid (GridIn::read_msh, main.cc:621-637)."""
        h = C.c_void_p()
        check(lib().oc_mesh_read_msh(str(path).encode(), int(periodic[0]), int(periodic[1]),
                                     C.byref(h)), "oc_mesh_read_msh")
        return cls(h)

    def refine_global(self, times=1):
        check(lib().oc_mesh_refine_global(self._h, times), "oc_mesh_refine_global")
        self._export()
        return self

    def refine_fraction(self, fraction, seed):
        check(lib().oc_mesh_refine_fraction(self._h, fraction, seed), "oc_mesh_refine_fraction")
        self._export()
        return self

    def _export(self):
        info = np.zeros(4, np.int64)
        check(lib().oc_mesh_info(self._h, info.ctypes.data_as(_capi.PI64)), "oc_mesh_info")
        n, ni, nb, nh = (int(v) for v in info)
        self.n_hanging = nh
        self.geom = np.zeros((n, 4))
        self.level = np.zeros(n, np.int32)
        self.inner = np.zeros((ni, 6), np.int32)
        self.boundary = np.zeros((nb, 3), np.int32)
        check(lib().oc_mesh_export(self._h, _capi.dptr(self.geom), _capi.iptr(self.level),
                                   _capi.iptr(self.inner), _capi.iptr(self.boundary)),
              "oc_mesh_export")

    @property
    def n_cells(self):
        return self.geom.shape[0]

    def __del__(self):
        try:
            if self._h:
                lib().oc_mesh_destroy(self._h)
                self._h = None
        except Exception:
            pass

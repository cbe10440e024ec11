"""MI355X ghost-cell AB2/CN cylinder step: host tables + HIP (gfx950) library via a C ABI.

Importing the package does not load the GPU library; the first handle does, and
raises if libghostflow.so is missing.
"""
from .tables import (CircleSDF, GhostCellGeometry, GhostOperators, RectangleFactor, SchurFactor, extension,
                     rectangle_matrix)
from .flow import MAC_CylinderFlow, GhostCellIntegrator, gaussian_v_perturbation

__all__ = ['CircleSDF', 'GhostCellGeometry', 'GhostOperators', 'RectangleFactor', 'SchurFactor', 'extension',
           'rectangle_matrix', 'MAC_CylinderFlow', 'GhostCellIntegrator', 'gaussian_v_perturbation']

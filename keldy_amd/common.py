"""keldy.common counterpart (python/keldy/common_desc.py): contour points,
contour ordering and the Sobol generator."""
from ._host import ext as _ext

ContourPtT = _ext.common.ContourPtT
Sobol = _ext.common.Sobol
compare_3way = _ext.common.compare_3way
jk2008_matrices = _ext.common.jk2008_matrices

__all__ = ["ContourPtT", "Sobol", "compare_3way", "jk2008_matrices"]

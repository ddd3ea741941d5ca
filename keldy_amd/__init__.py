"""keldy_amd -- MI355X-native Keldysh-sum integrand and quasi-Monte-Carlo driver.

Module layout mirrors the reference's Python package (keldy.common,
keldy.warpers, keldy.impurity_oneband); the compute path is the HIP library
behind the C-ABI in include/keldy_amd.h.
"""
__version__ = "0.1.0"

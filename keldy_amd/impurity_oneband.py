"""keldy.impurity_oneband counterpart (python/keldy/impurity_oneband_desc.py).

Model parameters are passed as a dict with the reference's keys (beta,
bias_V, eps_d, Gamma, alpha, half_bandwidth, bath_type, time_max,
nr_time_points_gf, ft_method) and defaults (model.hpp:37-50).
"""
from ._host import ext as _ext

_m = _ext.impurity_oneband
GfIndexT = _m.GfIndexT
G0ModelOmega = _m.G0ModelOmega
G0Model = _m.G0Model
G0KeldyshContourT = _m.G0KeldyshContourT
IntegrandGDirect = _m.IntegrandGDirect
ComputeChargeQDirect = _m.ComputeChargeQDirect
ComputeCurrentJDirect = _m.ComputeCurrentJDirect
IntegratorDirect = _m.IntegratorDirect
ComputeChargeQDirectTime = _m.ComputeChargeQDirectTime
IntegrandGDirectTime = _m.IntegrandGDirectTime
Binner1 = _m.Binner1
SparseBinner1 = _m.SparseBinner1
IntegrandGKernel = _m.IntegrandGKernel
IntegrandGKernelSingleOmega = _m.IntegrandGKernelSingleOmega
ComputeGfKernel = _m.ComputeGfKernel
ComputeGfKernelSingleOmega = _m.ComputeGfKernelSingleOmega
Binner_1_1 = _m.Binner_1_1
SparseBinner_1_1 = _m.SparseBinner_1_1
run_batch = _m.run_batch
train_flow_warper_conditioned = _m.train_flow_warper_conditioned

__all__ = ["GfIndexT", "G0ModelOmega", "G0Model", "G0KeldyshContourT", "IntegrandGDirect",
           "ComputeChargeQDirect", "ComputeCurrentJDirect", "IntegratorDirect", "ComputeChargeQDirectTime",
           "IntegrandGDirectTime", "Binner1", "SparseBinner1", "IntegrandGKernel", "IntegrandGKernelSingleOmega",
           "ComputeGfKernel", "ComputeGfKernelSingleOmega", "Binner_1_1", "SparseBinner_1_1", "run_batch",
           "train_flow_warper_conditioned"]

"""ctypes binding of the C-ABI in include/keldy_amd.h (libkeldy_amd.so).

This is the same boundary a ctypes/cgo/JNI caller of the reference path would
bind (see the header for the reference interface each entry point replaces).
The library is loaded from keldy_amd/lib/; importing this module on a machine
without the built library raises at load time -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _build

KD_OK = 0
KD_ERR_ARG, KD_ERR_HIP, KD_ERR_STATE, KD_ERR_RANGE, KD_ERR_NODEVICE = -1, -2, -3, -4, -5
KD_MAX_ORDER = 12
KD_PREC_FP64, KD_PREC_FP32_LU, KD_PREC_FP32_LU_ALL = 0, 1, 2

KD_MAP_UNIFORM, KD_MAP_EXPONENTIAL, KD_MAP_INVERSE, KD_MAP_INVERSE_SQ = 0, 1, 2, 3
KD_MAP_INVERSE_POW, KD_MAP_NEAREST, KD_MAP_HYBRID = 4, 5, 6
KD_MAP_PROJECTION = 7
KD_WARP_IDENTITY, KD_WARP_PLASMA, KD_WARP_PRODUCT, KD_WARP_FLOW = 0, 1, 2, 3
KD_ALGO_DEFAULT, KD_ALGO_CONFIG, KD_ALGO_TREE, KD_ALGO_PAIR, KD_ALGO_SPLIT, KD_ALGO_WSPLIT = -1, 0, 1, 2, 3, 4
KD_ALGO_CTREE = 5
KD_TIMING_NONE, KD_TIMING_TOTAL, KD_TIMING_KERNELS = 0, 1, 2
KD_OBS_DIRECT, KD_OBS_KERNEL, KD_OBS_KERNEL_OMEGA = 0, 1, 2


class GfIndex(C.Structure):
    _fields_ = [("time", C.c_double), ("k_idx", C.c_int32), ("spin", C.c_int32),
                ("orbital", C.c_int32), ("pad", C.c_int32)]


class Map1d(C.Structure):
    _fields_ = [("kind", C.c_int32), ("power", C.c_int32), ("u_max", C.c_double),
                ("w", C.c_double), ("norm", C.c_double), ("n_pts", C.c_int32),
                ("n_inv", C.c_int32), ("t_off", C.c_int64), ("f_off", C.c_int64),
                ("F_off", C.c_int64), ("c_off", C.c_int64)]


class WarperDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_maps", C.c_int32), ("t_max", C.c_double),
                ("map_offset", C.c_int64)]


class Accum(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double), ("n_points", C.c_uint64),
                ("n_in_domain", C.c_uint64), ("n_fallback", C.c_uint64)]


# every symbol the header declares
SYMBOLS = ["kd_create", "kd_destroy", "kd_last_error", "kd_device_count", "kd_set_g0_table",
           "kd_set_integrand", "kd_set_precision", "kd_eval_batch", "kd_eval_batch_device",
           "kd_set_warper_chain", "kd_eval_warped_batch", "kd_set_sobol", "kd_get_sobol_state",
           "kd_run", "kd_get_accum", "kd_reset_accum", "kd_sample_device", "kd_set_timing", "kd_last_timing",
           "kd_flops_per_point", "kd_flow_n_params", "kd_flow_init", "kd_flow_set_params",
           "kd_flow_get_params", "kd_flow_map", "kd_flow_loss_grad", "kd_flow_collect", "kd_flow_train",
           "kd_autotune", "kd_set_launch_shape", "kd_get_launch_shape", "kd_run_batch",
           "kd_set_time_bins", "kd_get_time_bins", "kd_set_algorithm", "kd_get_algorithm",
           "kd_pilot_histograms", "kd_set_kernel_integrand", "kd_set_kernel_omega_integrand",
           "kd_set_kernel_bins", "kd_get_kernel_bins", "kd_eval_kernel_batch",
           "kd_eval_batch_device_noalloc", "kd_flow_set_conditioning", "kd_flow_n_params_cond",
           "kd_flow_loss_grad_cond", "kd_flow_get_samples", "kd_flow_set_samples"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _build.lib_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built: run __graft_entry__.build() (no CPU fallback)")
        L = C.CDLL(path)
        dp, ip, up = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)
        vp = C.c_void_p
        L.kd_create.argtypes = [C.POINTER(vp), C.c_int]
        L.kd_destroy.argtypes = [vp]
        L.kd_last_error.argtypes = [vp]
        L.kd_last_error.restype = C.c_char_p
        L.kd_device_count.argtypes = [C.POINTER(C.c_int)]
        L.kd_set_g0_table.argtypes = [vp, dp, dp, C.c_int32, C.c_int32, C.c_double, C.c_double]
        L.kd_set_integrand.argtypes = [vp, C.c_int32, C.POINTER(GfIndex), C.POINTER(GfIndex), C.c_double]
        L.kd_set_precision.argtypes = [vp, C.c_int32]
        L.kd_eval_batch.argtypes = [vp, dp, C.c_int64, C.c_int32, dp, ip]
        L.kd_eval_batch_device.argtypes = [vp, vp, C.c_int64, C.c_int32, vp, vp]
        L.kd_eval_batch_device_noalloc.argtypes = [vp, vp, C.c_int64, C.c_int32, vp, vp]
        L.kd_set_warper_chain.argtypes = [vp, C.POINTER(WarperDesc), C.c_int32, C.POINTER(Map1d),
                                          C.c_int32, dp, C.c_int64]
        L.kd_eval_warped_batch.argtypes = [vp, dp, C.c_int64, C.c_int32, dp, ip, dp]
        L.kd_set_sobol.argtypes = [vp, C.c_int32, up, up, C.c_uint32]
        L.kd_get_sobol_state.argtypes = [vp, up]
        L.kd_run.argtypes = [vp, C.c_uint64, C.POINTER(Accum)]
        L.kd_get_accum.argtypes = [vp, C.POINTER(Accum)]
        L.kd_reset_accum.argtypes = [vp]
        L.kd_sample_device.argtypes = [vp, C.c_uint64, vp, vp]
        L.kd_set_timing.argtypes = [vp, C.c_int32]
        L.kd_last_timing.argtypes = [vp, dp, dp, dp, ip]
        L.kd_flops_per_point.argtypes = [vp, dp]
        fp = C.POINTER(C.c_float)
        L.kd_flow_n_params.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int64)]
        L.kd_flow_init.argtypes = [vp, C.c_int32, C.c_int32, C.c_uint64]
        L.kd_flow_set_params.argtypes = [vp, C.c_int32, C.c_int32, fp, C.c_int64]
        L.kd_flow_get_params.argtypes = [vp, fp, C.c_int64]
        L.kd_flow_map.argtypes = [vp, dp, C.c_int64, dp, dp]
        L.kd_flow_loss_grad.argtypes = [vp, dp, dp, C.c_int64, C.c_double, fp, dp]
        L.kd_flow_collect.argtypes = [vp, C.c_int64, C.c_double, dp]
        L.kd_flow_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_float, dp]
        L.kd_flow_set_conditioning.argtypes = [vp, C.c_int32, C.c_double]
        L.kd_flow_n_params_cond.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64)]
        L.kd_flow_loss_grad_cond.argtypes = [vp, dp, dp, dp, C.c_int64, C.c_double, fp, dp]
        L.kd_flow_get_samples.argtypes = [vp, dp, dp, C.c_int64]
        L.kd_flow_set_samples.argtypes = [vp, dp, dp, dp, C.c_int64]
        L.kd_autotune.argtypes = [vp, C.c_int64, ip, ip, dp]
        L.kd_set_launch_shape.argtypes = [vp, C.c_int32, C.c_int32]
        L.kd_get_launch_shape.argtypes = [vp, ip, ip]
        L.kd_run_batch.argtypes = [C.POINTER(vp), C.POINTER(C.c_uint64), C.c_int32, C.POINTER(Accum), dp]
        L.kd_set_time_bins.argtypes = [vp, C.c_double, C.c_double, C.c_int32]
        L.kd_set_algorithm.argtypes = [vp, C.c_int32]
        L.kd_pilot_histograms.argtypes = [vp, up, up, C.c_int64, C.c_int32, C.c_int32, dp, C.POINTER(C.c_uint64)]
        L.kd_get_algorithm.argtypes = [vp, ip]
        L.kd_get_time_bins.argtypes = [vp, dp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        u64p = C.POINTER(C.c_uint64)
        L.kd_set_kernel_integrand.argtypes = [vp, C.c_int32, C.POINTER(GfIndex)]
        L.kd_set_kernel_omega_integrand.argtypes = [vp, C.c_int32, C.POINTER(GfIndex), C.c_double]
        L.kd_set_kernel_bins.argtypes = [vp, C.c_double, C.c_double, C.c_int32]
        L.kd_get_kernel_bins.argtypes = [vp, dp, u64p, u64p]
        L.kd_eval_kernel_batch.argtypes = [vp, dp, C.c_int64, C.c_int32, C.c_int32, dp, ip, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class KdError(RuntimeError):
    pass


class Context:
    """Owner of one kd_ctx (one device, one stream)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().kd_create(C.byref(self.h), device)
        if rc != KD_OK:
            raise KdError(f"kd_create failed ({rc}): no usable HIP device")

    def _check(self, rc):
        if rc != KD_OK:
            msg = lib().kd_last_error(self.h).decode()
            raise KdError(msg or f"keldy_amd error {rc}")

    def close(self):
        if self.h:
            lib().kd_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up ----------------------------------------------------------
    def set_g0_table(self, lesser, greater, time_max, alpha=0.0):
        les = np.ascontiguousarray(lesser, dtype=np.complex128)
        gre = np.ascontiguousarray(greater, dtype=np.complex128)
        if les.ndim == 1:
            les = les.reshape(1, 1, -1)
            gre = gre.reshape(1, 1, -1)
        self._g0 = (les, gre)
        self._check(lib().kd_set_g0_table(self.h, _dp(les.view(np.float64)), _dp(gre.view(np.float64)),
                                          les.shape[0], les.shape[-1], time_max, alpha))

    def set_integrand(self, order, A, B, cutoff=0.0):
        def gi(x):
            return x if isinstance(x, GfIndex) else GfIndex(float(x[0]), int(x[2]), int(x[1]),
                                                             int(x[3]) if len(x) > 3 else 0, 0)
        self._check(lib().kd_set_integrand(self.h, order, C.byref(gi(A)), C.byref(gi(B)), cutoff))
        self.order = order

    def set_precision(self, prec):
        self._check(lib().kd_set_precision(self.h, prec))

    def set_warper_chain(self, descs, maps, tables):
        nd, nm = len(descs), len(maps)
        darr = (WarperDesc * max(nd, 1))(*descs)
        marr = (Map1d * max(nm, 1))(*maps)
        tab = np.ascontiguousarray(tables if len(tables) else np.zeros(1), dtype=np.float64)
        self._tab = tab
        self._check(lib().kd_set_warper_chain(self.h, darr, nd, marr, nm, _dp(tab), len(tables)))

    def set_sobol(self, Csr, shfr, k=0):
        Csr = np.ascontiguousarray(Csr, dtype=np.uint32)
        shfr = np.ascontiguousarray(shfr, dtype=np.uint32)
        self._check(lib().kd_set_sobol(self.h, shfr.shape[0], Csr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       shfr.ctypes.data_as(C.POINTER(C.c_uint32)), k))

    # -- evaluation ------------------------------------------------------
    def eval_batch(self, u, keep_u_hypercube=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        n = u.shape[0]
        out = np.empty((n, 2))
        dom = np.empty(n, dtype=np.int32)
        self._check(lib().kd_eval_batch(self.h, _dp(u), n, int(keep_u_hypercube), _dp(out),
                                        dom.ctypes.data_as(C.POINTER(C.c_int32))))
        return out[:, 0] + 1j * out[:, 1], dom

    def eval_warped_batch(self, l, keep_u_hypercube=True):
        l = np.ascontiguousarray(l, dtype=np.float64)
        n = l.shape[0]
        out = np.empty((n, 2))
        dom = np.empty(n, dtype=np.int32)
        jac = np.empty(n)
        self._check(lib().kd_eval_warped_batch(self.h, _dp(l), n, int(keep_u_hypercube), _dp(out),
                                               dom.ctypes.data_as(C.POINTER(C.c_int32)), _dp(jac)))
        return out[:, 0] + 1j * out[:, 1], dom, jac

    def run(self, n_steps):
        acc = Accum()
        self._check(lib().kd_run(self.h, n_steps, C.byref(acc)))
        return acc

    def accum(self):
        acc = Accum()
        self._check(lib().kd_get_accum(self.h, C.byref(acc)))
        return acc

    def reset_accum(self):
        self._check(lib().kd_reset_accum(self.h))

    def sobol_state(self):
        k = C.c_uint32()
        self._check(lib().kd_get_sobol_state(self.h, C.byref(k)))
        return k.value

    def set_timing(self, level):
        """KD_TIMING_NONE, KD_TIMING_TOTAL or KD_TIMING_KERNELS (the default)."""
        self._check(lib().kd_set_timing(self.h, int(level)))

    def last_timing(self):
        """(total_ms, lu_ms, prep_ms, launches) of the last device call."""
        t, lu, pr = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int32()
        self._check(lib().kd_last_timing(self.h, C.byref(t), C.byref(lu), C.byref(pr), C.byref(n)))
        return t.value, lu.value, pr.value, n.value

    def flops_per_point(self):
        f = C.c_double()
        self._check(lib().kd_flops_per_point(self.h, C.byref(f)))
        return f.value

    # -- time binning ----------------------------------------------------
    def set_time_bins(self, xmin, xmax, nbins):
        self._check(lib().kd_set_time_bins(self.h, xmin, xmax, nbins))
        self.nbins = nbins

    def time_bins(self):
        """-> (complex sums per bin, values added per bin, dropped count)"""
        v = np.empty((self.nbins, 2))
        cnt = np.empty(self.nbins, dtype=np.uint64)
        d = C.c_uint64()
        self._check(lib().kd_get_time_bins(self.h, _dp(v), cnt.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(d)))
        return v[:, 0] + 1j * v[:, 1], cnt, d.value

    # -- kernel method ---------------------------------------------------
    def set_kernel_integrand(self, order, X, omega=None):
        """integrand_g_kernel (omega None) or integrand_g_kernel_single_omega at X"""
        if omega is None:
            self._check(lib().kd_set_kernel_integrand(self.h, order, C.byref(X)))
        else:
            self._check(lib().kd_set_kernel_omega_integrand(self.h, order, C.byref(X), float(omega)))
        self.order = order

    def set_kernel_bins(self, xmin, xmax, nbins):
        self._check(lib().kd_set_kernel_bins(self.h, xmin, xmax, nbins))
        self.nbins = nbins

    def kernel_bins(self):
        """-> (complex sums [nbins][2], entries added [nbins][2], dropped)"""
        v = np.empty((self.nbins, 2, 2))
        cnt = np.empty((self.nbins, 2), dtype=np.uint64)
        d = C.c_uint64()
        self._check(lib().kd_get_kernel_bins(self.h, _dp(v), cnt.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(d)))
        return v[..., 0] + 1j * v[..., 1], cnt, d.value

    def eval_kernel_batch(self, pts, warped=False, keep_u_hypercube=True):
        """-> (entries [n][2^order][order] complex, in_domain [n], jacobian [n])"""
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        n, order = pts.shape
        out = np.empty((n, 1 << order, order, 2))
        dom = np.empty(n, dtype=np.int32)
        jac = np.empty(n)
        self._check(lib().kd_eval_kernel_batch(self.h, _dp(pts), n, int(warped), int(keep_u_hypercube), _dp(out),
                                               dom.ctypes.data_as(C.POINTER(C.c_int32)), _dp(jac)))
        return out[..., 0] + 1j * out[..., 1], dom, jac

    # -- launch shape ----------------------------------------------------
    def autotune(self, sample_points):
        """-> (lanes per point, waves per CU, LU ms per 2^20 points)"""
        l, w, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._check(lib().kd_autotune(self.h, sample_points, C.byref(l), C.byref(w), C.byref(ms)))
        return l.value, w.value, ms.value

    def set_launch_shape(self, lanes, waves_per_cu):
        self._check(lib().kd_set_launch_shape(self.h, lanes, waves_per_cu))

    def set_algorithm(self, algo):
        self._check(lib().kd_set_algorithm(self.h, algo))

    def algorithm(self):
        a = C.c_int32()
        self._check(lib().kd_get_algorithm(self.h, C.byref(a)))
        return a.value

    def launch_shape(self):
        l, w = C.c_int32(), C.c_int32()
        self._check(lib().kd_get_launch_shape(self.h, C.byref(l), C.byref(w)))
        return l.value, w.value

    # -- learned flow warper ---------------------------------------------
    flow_nc = 0

    def flow_set_conditioning(self, n_cond, value=0.0):
        self._check(lib().kd_flow_set_conditioning(self.h, n_cond, float(value)))
        self.flow_nc = n_cond

    def flow_init(self, order, n_layers, seed=0):
        self._check(lib().kd_flow_init(self.h, order, n_layers, seed))
        self.flow_shape = (order, n_layers, self.flow_nc)

    def flow_set_params(self, order, n_layers, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(lib().kd_flow_set_params(self.h, order, n_layers, p.ctypes.data_as(C.POINTER(C.c_float)),
                                             p.size))
        self.flow_shape = (order, n_layers, self.flow_nc)

    def flow_get_params(self):
        p = np.empty(flow_n_params(*self.flow_shape), dtype=np.float32)
        self._check(lib().kd_flow_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))
        return p

    def flow_map(self, l):
        l = np.ascontiguousarray(l, dtype=np.float64)
        x = np.empty_like(l)
        jac = np.empty(l.shape[0])
        self._check(lib().kd_flow_map(self.h, _dp(l), l.shape[0], _dp(x), _dp(jac)))
        return x, jac

    def flow_loss_grad(self, x, w, w_mean=None, cond=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        g = np.empty(flow_n_params(*self.flow_shape), dtype=np.float32)
        loss = C.c_double()
        cp = None
        if cond is not None:
            cond = np.ascontiguousarray(np.broadcast_to(cond, (x.shape[0],)), dtype=np.float64)
            cp = _dp(cond)
        self._check(lib().kd_flow_loss_grad_cond(self.h, _dp(x), _dp(w), cp, x.shape[0],
                                                 float(w.mean() if w_mean is None else w_mean),
                                                 g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return loss.value, g

    def flow_get_samples(self, n, order):
        x = np.empty((n, order))
        w = np.empty(n)
        self._check(lib().kd_flow_get_samples(self.h, _dp(x), _dp(w), n))
        return x, w

    def flow_set_samples(self, x, w, cond=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        cp = None
        if cond is not None:
            cond = np.ascontiguousarray(cond, dtype=np.float64)
            cp = _dp(cond)
        self._check(lib().kd_flow_set_samples(self.h, _dp(x), _dp(w), cp, x.shape[0]))

    def flow_collect(self, n_samples, weight_power=1.0):
        wm = C.c_double()
        self._check(lib().kd_flow_collect(self.h, n_samples, weight_power, C.byref(wm)))
        return wm.value

    def flow_train(self, steps, batch, lr, lr_final):
        losses = np.empty(max(steps, 1))
        self._check(lib().kd_flow_train(self.h, steps, batch, lr, lr_final, _dp(losses)))
        return losses[:steps]


def run_batch(ctxs, n_steps):
    """One hipGraph launch over several contexts -> (list of Accum, ms)."""
    n = len(ctxs)
    hs = (C.c_void_p * n)(*[c.h for c in ctxs])
    ns = (C.c_uint64 * n)(*[int(x) for x in n_steps])
    acc = (Accum * n)()
    ms = C.c_double()
    rc = lib().kd_run_batch(hs, ns, n, acc, C.byref(ms))
    if rc != KD_OK:
        raise KdError(lib().kd_last_error(ctxs[0].h).decode() or f"keldy_amd error {rc}")
    return list(acc), ms.value


def flow_n_params(order, n_layers, n_cond=0):
    n = C.c_int64()
    if lib().kd_flow_n_params_cond(order, n_layers, n_cond, C.byref(n)) != KD_OK:
        raise KdError("bad flow shape")
    return n.value


def device_count():
    n = C.c_int()
    lib().kd_device_count(C.byref(n))
    return n.value

"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  It wraps oracle/_build/libkeldy_oracle.so
(the plain-C restatement in keldy_oracle.c) and oracle/g0_oracle.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libkeldy_oracle.so")
REF_SOBOL = os.path.join(HERE, "_ref", "sobol_ref")

KO_MAX_DIM = 32
MAP_KINDS = {"uniform": 0, "exponential": 1, "inverse": 2, "inverse_square": 3,
             "inverse_power": 4, "nearest": 5, "hybrid": 6}
WARP_IDENTITY, WARP_PLASMA, WARP_PRODUCT = 0, 1, 2


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


class _Sobol(C.Structure):
    _fields_ = [("dim", C.c_int), ("k", C.c_uint32),
                ("Csr", C.c_uint32 * (KO_MAX_DIM * 31)), ("shfr", C.c_uint32 * KO_MAX_DIM),
                ("cur", C.c_uint32 * KO_MAX_DIM)]


class _GfIndex(C.Structure):
    _fields_ = [("time", C.c_double), ("k_idx", C.c_int), ("timesplit", C.c_int),
                ("spin", C.c_int), ("orbital", C.c_int)]


class _G0(C.Structure):
    _fields_ = [("n_pts", C.c_int), ("n_orb", C.c_int), ("t_min", C.c_double),
                ("dt", C.c_double), ("alpha", C.c_double),
                ("lesser", C.POINTER(C.c_double)), ("greater", C.POINTER(C.c_double))]


class _Map1d(C.Structure):
    _fields_ = [("kind", C.c_int), ("power", C.c_int), ("u_max", C.c_double), ("w", C.c_double),
                ("n_pts", C.c_int), ("n_inv", C.c_int), ("t", C.POINTER(C.c_double)),
                ("f", C.POINTER(C.c_double)), ("F", C.POINTER(C.c_double)),
                ("inv_c", C.POINTER(C.c_double)), ("norm", C.c_double)]


class _Warper(C.Structure):
    _fields_ = [("kind", C.c_int), ("t_max", C.c_double), ("maps", C.POINTER(_Map1d)),
                ("n_maps", C.c_int)]


class _Integrand(C.Structure):
    _fields_ = [("g0", _G0), ("A", _GfIndex), ("B", _GfIndex), ("cutoff", C.c_double)]


class _Accum(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double), ("n_points", C.c_uint64),
                ("n_in_domain", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def load(path):
    """The oracle library at `path` with its ctypes signatures (lib() loads the
    restatement; tests load variants built with a planted mistake)."""
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    L.ko_mt64_seed.argtypes = [C.c_void_p, C.c_uint64]
    L.ko_mt64_next.argtypes = [C.c_void_p]
    L.ko_mt64_next.restype = C.c_uint64
    L.ko_jk2008_matrices.argtypes = [C.POINTER(C.c_uint32), C.c_int, C.c_int]
    L.ko_sobol_init.argtypes = [C.POINTER(_Sobol), C.c_int, C.c_uint32, C.c_int, C.c_int,
                                C.c_int64]
    L.ko_sobol_next.argtypes = [C.POINTER(_Sobol), dp]
    L.ko_sobol_set_state.argtypes = [C.POINTER(_Sobol), C.c_uint32]
    L.ko_compare_3way.argtypes = [C.POINTER(_GfIndex), C.POINTER(_GfIndex)]
    L.ko_g0_eval.argtypes = [C.POINTER(_G0), C.POINTER(_GfIndex), C.POINTER(_GfIndex), C.c_int, dp]
    L.ko_map1d_reverse.argtypes = [C.POINTER(_Map1d), C.c_double, dp]
    L.ko_map1d_reverse.restype = C.c_double
    L.ko_warp_reverse.argtypes = [C.POINTER(_Warper), C.c_int, C.c_int, dp, dp]
    L.ko_warp_reverse.restype = C.c_double
    L.ko_steffen_coeffs.argtypes = [dp, dp, C.c_int, dp]
    L.ko_steffen_eval.argtypes = [dp, dp, C.c_int, C.c_double]
    L.ko_steffen_eval.restype = C.c_double
    L.ko_integrand_eval.argtypes = [C.POINTER(_Integrand), C.c_int, dp, C.c_int, dp]
    L.ko_integrand_eval_no_det.argtypes = [C.POINTER(_Integrand), C.c_int, dp, C.c_int, dp]
    L.ko_det.argtypes = [dp, C.c_int, dp]
    L.ko_eval_warped.argtypes = [C.POINTER(_Integrand), C.POINTER(_Warper), C.c_int, C.c_int,
                                 dp, C.c_int, dp]
    L.ko_run.argtypes = [C.POINTER(_Integrand), C.POINTER(_Warper), C.c_int, C.POINTER(_Sobol),
                         C.c_uint64, C.POINTER(_Accum)]
    L.ko_run_parallel.argtypes = [C.POINTER(_Integrand), C.POINTER(_Warper), C.c_int,
                                  C.POINTER(_Sobol), C.c_uint64, C.c_int, C.POINTER(_Accum)]
    L.ko_run_intrapoint.argtypes = [C.POINTER(_Integrand), C.POINTER(_Warper), C.c_int,
                                    C.POINTER(_Sobol), C.c_uint64, C.c_int, C.POINTER(_Accum)]
    L.ko_first_row_expansion.argtypes = [dp, C.c_int, dp, dp]
    L.ko_kernel_eval.argtypes = [C.POINTER(_G0), C.POINTER(_GfIndex), C.c_int, dp, C.c_int, dp]
    L.ko_kernel_omega_eval.argtypes = [C.POINTER(_G0), C.POINTER(_GfIndex), C.c_int, dp, C.c_int,
                                       C.c_double, dp]
    u64p = C.POINTER(C.c_uint64)
    L.ko_kernel_run.argtypes = [C.POINTER(_G0), C.POINTER(_GfIndex), C.POINTER(_Warper), C.c_int,
                                C.POINTER(_Sobol), C.c_uint64, C.c_double, C.c_int, dp, u64p, u64p, u64p]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# --------------------------------------------------------------------------
def mt64(seed, count):
    g = C.create_string_buffer(312 * 8 + 16)
    lib().ko_mt64_seed(g, seed)
    return [lib().ko_mt64_next(g) for _ in range(count)]


def jk2008(dim, m=31):
    out = (C.c_uint32 * (dim * m))()
    if lib().ko_jk2008_matrices(out, dim, m) != 0:
        raise ValueError("dimension out of range")
    return np.array(out, dtype=np.uint32).reshape(dim, m)


class Sobol:
    """keldy::sobol restated (qrng.hpp:49-114)."""

    def __init__(self, dim, state_seed=0, do_shift=False, do_scramble=False, rng_seed_shift=0):
        self.s = _Sobol()
        if lib().ko_sobol_init(C.byref(self.s), dim, state_seed, int(do_shift), int(do_scramble),
                               rng_seed_shift) != 0:
            raise ValueError("bad Sobol dimension")
        self.dim = dim

    def next(self):
        x = np.empty(self.dim)
        if lib().ko_sobol_next(C.byref(self.s), _dp(x)) != 0:
            raise RuntimeError("More sample points requested then generator can hold.")
        return x

    def points(self, n):
        return np.array([self.next() for _ in range(n)])

    def tables(self):
        Csr = np.array(self.s.Csr, dtype=np.uint32)[: self.dim * 31].reshape(self.dim, 31)
        shfr = np.array(self.s.shfr, dtype=np.uint32)[: self.dim]
        return Csr, shfr


def ref_sobol_points(dim, state_seed, do_shift, do_scramble, rng_seed_shift, n):
    """Points from the reference's own Sobol header (oracle/_ref/sobol_ref)."""
    out = subprocess.run([REF_SOBOL, str(dim), str(state_seed), str(int(do_shift)),
                          str(int(do_scramble)), str(rng_seed_shift), str(n)],
                         check=True, capture_output=True, text=True).stdout
    return np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])


# --------------------------------------------------------------------------
class G0Table:
    """A g0 time table as consumed by the oracle (model.hpp:238-268)."""

    def __init__(self, lesser, greater, time_max, alpha=0.0):
        lesser = np.ascontiguousarray(lesser, dtype=np.complex128)
        greater = np.ascontiguousarray(greater, dtype=np.complex128)
        if lesser.ndim == 1:
            lesser = lesser.reshape(1, 1, -1)
            greater = greater.reshape(1, 1, -1)
        self.n_orb = lesser.shape[0]
        self.n_pts = lesser.shape[-1]
        self.lesser = lesser.view(np.float64).copy()
        self.greater = greater.view(np.float64).copy()
        self.time_max = time_max
        self.alpha = alpha
        self.c = _G0(self.n_pts, self.n_orb, -time_max, 2 * time_max / (self.n_pts - 1), alpha,
                     _dp(self.lesser), _dp(self.greater))


def gf_index(time, spin=0, k_idx=0, timesplit=0, orbital=0):
    return _GfIndex(time, k_idx, timesplit, spin, orbital)


def g0_eval(g0, a, b, internal=True):
    """g0_keldysh_contour_t::operator() (model.hpp:244-265) on a G0Table."""
    o = np.zeros(2)
    lib().ko_g0_eval(C.byref(g0.c), C.byref(a), C.byref(b), int(internal), _dp(o))
    return complex(o[0], o[1])


class Map1d:
    def __init__(self, kind, u_max, w=1.0, power=1, t=None, f=None, F=None, n_inv=None,
                 inv_c=None, norm=1.0):
        self.keep = []
        self.c = _Map1d()
        self.c.kind = MAP_KINDS[kind]
        self.c.u_max = u_max
        self.c.w = w
        self.c.power = power
        self.c.norm = norm
        for name, arr in (("t", t), ("f", f), ("F", F), ("inv_c", inv_c)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                self.keep.append(a)
                setattr(self.c, name, _dp(a))
        if t is not None:
            self.c.n_pts = len(t)
            self.c.n_inv = n_inv if n_inv is not None else len(t)

    def reverse(self, l):
        x = np.zeros(1)
        j = lib().ko_map1d_reverse(C.byref(self.c), l, _dp(x))
        return x[0], j


def nearest_map(times, f1):
    """warper_product_1d_simple_interp_nearest_t set-up (product_1d_simple.hpp:219-287)."""
    times = np.asarray(times, dtype=np.float64)
    f1 = np.array(f1, dtype=np.float64)
    n = len(times)
    F = np.zeros(n)
    cut = n - 1
    found = False
    for i in range(n - 2, -1, -1):
        F[i] = F[i + 1] - 0.5 * (f1[i + 1] + f1[i]) * (times[i + 1] - times[i])
        found = found or (F[i] != 0)
        if not found:
            cut = i
    if not found:
        raise RuntimeError("Model function is flat zero.")
    norm = -F[0]
    F /= norm
    f1 /= norm
    F[0] = -1
    return Map1d("nearest", times[-1], t=times, f=f1, F=F, n_inv=cut + 1), norm


def hybrid_map(power, w, u_max, n):
    """warper_product_1d_simple_interp_hybrid_t set-up (product_1d_simple.hpp:424-491)
    for the f1(t) = w^p/(w+t)^p kernels used by
    make_product_1d_inverse_cube_alternate_interp_hybrid (product_1d.hpp:291-310)."""

    def fu(t):
        num = w
        for _ in range(1, power):
            num *= w
        d = w + t
        den = d
        for _ in range(1, power):
            den *= d
        return num / den

    delta = u_max / (n - 1)
    times = np.array([delta * i for i in range(n)])
    times[-1] = u_max
    F = np.zeros(n)
    cut = n - 1
    found = False
    f2d = fu(times[n - 1])
    for i in range(n - 2, -1, -1):
        f0d = f2d
        f1d = fu(0.5 * (times[i + 1] + times[i]))
        f2d = fu(times[i])
        dl = delta / 6. * (f0d + 4 * f1d + f2d)
        F[i] = F[i + 1] - dl
        found = found or (F[i] != 0)
        if not found:
            cut = i
    norm = -F[0]
    F /= norm
    F[0] = -1
    n_inv = cut + 1
    coeffs = np.zeros((n_inv - 1) * 4)
    Fi = np.ascontiguousarray(F[:n_inv])
    ti = np.ascontiguousarray(times[:n_inv])
    lib().ko_steffen_coeffs(_dp(Fi), _dp(ti), n_inv, _dp(coeffs))
    return Map1d("hybrid", u_max, w=w, power=power, t=times, F=F, n_inv=n_inv, inv_c=coeffs,
                 norm=norm)


class Warper:
    def __init__(self, kind, t_max=0.0, maps=None):
        self.c = _Warper()
        self.c.kind = kind
        self.c.t_max = t_max
        self.maps = maps or []
        if maps:
            self.arr = (_Map1d * len(maps))(*[m.c for m in maps])
            self.c.maps = C.cast(self.arr, C.POINTER(_Map1d))
            self.c.n_maps = len(maps)


def plasma(t_max):
    return Warper(WARP_PLASMA, t_max)


def product(maps):
    return Warper(WARP_PRODUCT, 0.0, maps if isinstance(maps, list) else [maps])


class Train:
    def __init__(self, warpers):
        self.warpers = list(warpers)
        self.arr = (_Warper * max(1, len(self.warpers)))(*[w.c for w in self.warpers])

    def reverse(self, l):
        l = np.ascontiguousarray(l, dtype=np.float64)
        u = np.empty_like(l)
        j = lib().ko_warp_reverse(self.arr, len(self.warpers), len(l), _dp(l), _dp(u))
        return u, j


class Integrand:
    """integrand_g_direct (wick_direct.hpp:34-65)."""

    def __init__(self, g0, A, B, cutoff=0.0):
        self.g0 = g0
        self.c = _Integrand(g0.c, A, B, cutoff)

    def __call__(self, times, keep_u_hypercube=True):
        t = np.ascontiguousarray(times, dtype=np.float64)
        out = np.zeros(2)
        dom = lib().ko_integrand_eval(C.byref(self.c), len(t), _dp(t), int(keep_u_hypercube), _dp(out))
        return complex(out[0], out[1]), dom

    def eval_no_det(self, times, keep_u_hypercube=True):
        t = np.ascontiguousarray(times, dtype=np.float64)
        out = np.zeros(2)
        dom = lib().ko_integrand_eval_no_det(C.byref(self.c), len(t), _dp(t), int(keep_u_hypercube),
                                             _dp(out))
        if dom < 0:
            raise RuntimeError("Not supported for order > 2")
        return complex(out[0], out[1]), dom

    def batch(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        vals = np.empty(len(u), dtype=np.complex128)
        dom = np.empty(len(u), dtype=np.int32)
        for i in range(len(u)):
            vals[i], dom[i] = self(u[i])
        return vals, dom


def charge_integrand(g0, t, cutoff=0.0):
    """compute_charge_Q_direct's integrand (compute_obs.hpp:53-56)."""
    return Integrand(g0, gf_index(t, 0, 0), gf_index(t, 0, 1), cutoff)


def det(m):
    a = np.ascontiguousarray(m, dtype=np.complex128).copy()
    out = np.zeros(2)
    lib().ko_det(_dp(a.view(np.float64)), a.shape[0], _dp(out))
    return complex(out[0], out[1])


def eval_warped(integrand, train, l, keep_u_hypercube=True):
    l = np.ascontiguousarray(l, dtype=np.float64)
    out = np.zeros(2)
    dom = lib().ko_eval_warped(C.byref(integrand.c), train.arr, len(train.warpers), len(l), _dp(l),
                               int(keep_u_hypercube), _dp(out))
    return complex(out[0], out[1]), dom


def run(integrand, train, sobol, n_steps, threads=1, intrapoint=False):
    """integrator::run restated; returns (sum, n_points, n_in_domain).
    threads > 1 splits the points over the threads (contiguous Sobol ranges),
    or, with intrapoint, parallelises inside each point over its Wick rows and
    configurations as the reference does (wick_direct.cpp:81,93,110)."""
    acc = _Accum()
    if intrapoint:
        rc = lib().ko_run_intrapoint(C.byref(integrand.c), train.arr, len(train.warpers), C.byref(sobol.s),
                                     n_steps, threads, C.byref(acc))
    elif threads == 1:
        rc = lib().ko_run(C.byref(integrand.c), train.arr, len(train.warpers), C.byref(sobol.s),
                          n_steps, C.byref(acc))
    else:
        rc = lib().ko_run_parallel(C.byref(integrand.c), train.arr, len(train.warpers),
                                   C.byref(sobol.s), n_steps, threads, C.byref(acc))
    if rc != 0:
        raise RuntimeError("More sample points requested then generator can hold.")
    return complex(acc.re, acc.im), acc.n_points, acc.n_in_domain


# --------------------------------------------------------------------------
# kernel method (wick_kernel.cpp:42-138, wick_kernel_single_omega.cpp:41-140)
def first_row_expansion(m):
    """det_expansion.hpp:44-74: (first row of the comatrix, determinant)."""
    a = np.ascontiguousarray(m, dtype=np.complex128).copy()
    x = np.zeros(a.shape[0], dtype=np.complex128)
    d = np.zeros(2)
    if lib().ko_first_row_expansion(_dp(a.view(np.float64)), a.shape[0], _dp(x.view(np.float64)), _dp(d)) != 0:
        raise RuntimeError("lapack::getrf failed")
    return x, complex(d[0], d[1])


class KernelIntegrand:
    """integrand_g_kernel (wick_kernel.hpp:43-60) with fixed external point X."""

    def __init__(self, g0, X):
        self.g0 = g0
        self.X = X

    def entries(self, times, keep_u_hypercube=True):
        """(values [2^n][n] complex, in_domain): entry (c, i) belongs to vertex
        time times[i] and Keldysh index bit i of c."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        n = len(t)
        out = np.zeros((1 << n) * n * 2)
        dom = lib().ko_kernel_eval(C.byref(self.g0.c), C.byref(self.X), n, _dp(t), int(keep_u_hypercube), _dp(out))
        if dom < 0:
            raise RuntimeError("lapack::getrf failed")
        return out.view(np.complex128).reshape(1 << n, n), dom

    def __call__(self, times, keep_u_hypercube=True):
        """The reference's sparse binner: [((time, k_idx), value)] in accumulate order."""
        vals, dom = self.entries(times, keep_u_hypercube)
        if not dom:
            return [], 0
        n = len(times)
        return [((float(times[i]), (c >> i) & 1), complex(vals[c, i])) for c in range(1 << n) for i in range(n)], dom

    def omega(self, times, omega, keep_u_hypercube=True):
        t = np.ascontiguousarray(times, dtype=np.float64)
        out = np.zeros(2)
        dom = lib().ko_kernel_omega_eval(C.byref(self.g0.c), C.byref(self.X), len(t), _dp(t), int(keep_u_hypercube),
                                         float(omega), _dp(out))
        return complex(out[0], out[1]), dom


def kernel_run(kintegrand, train, sobol, n_steps, xmax, nbins):
    """compute_gf_kernel's run (compute_obs.hpp:105-117) into binner_t<1,1>:
    (bin sums [nbins][2] complex, counts [nbins][2], dropped, n_in_domain)."""
    bins = np.zeros(nbins * 4)
    counts = np.zeros(nbins * 2, dtype=np.uint64)
    dropped, ndom = C.c_uint64(0), C.c_uint64(0)
    u64p = C.POINTER(C.c_uint64)
    rc = lib().ko_kernel_run(C.byref(kintegrand.g0.c), C.byref(kintegrand.X), train.arr, len(train.warpers),
                             C.byref(sobol.s), n_steps, float(xmax), int(nbins), _dp(bins),
                             counts.ctypes.data_as(u64p), C.byref(dropped), C.byref(ndom))
    if rc != 0:
        raise RuntimeError("kernel run failed (generator exhausted or singular matrix)")
    return bins.view(np.complex128).reshape(nbins, 2), counts.reshape(nbins, 2), dropped.value, ndom.value

"""warper_projection_t restated in numpy (TEST INFRASTRUCTURE ONLY).

Only tests/ import this module, as the checker of the host construction and
the device application of the projection warper.  It follows
c++/keldy/warpers/projection.hpp:176-382 (gather_data, populate_sigmas with a
fixed sigma, populate_functions, the map and Jacobian evaluations), with the
pieces it calls restated from their published definitions:
- binner_t<1, 0, double> (binner.hpp:147-352): bins of [xmin, xmax], a value
  at xmax in the last bin, values outside dropped; the per-bin mean;
- kernel_smoothing / local_linear_reg (projection.hpp:53-78,
  gsl_fit_linear_wrap.hpp:66-113): at every bin a Gaussian-weighted linear
  least-squares fit of log(data) over the kernel window, bins with no data
  weighted zero;
- gsl_interp_linear and gsl_interp_akima (GSL interpolation; Akima 1970 with
  GSL's end extension of the slopes, m_-1 = 2 m_0 - m_1, m_-2 = 3 m_0 - 2 m_1
  and the mirror at the far end);
- TRIQS gf<retime> evaluation between mesh points: linear interpolation on
  the uniform mesh.
"""
import numpy as np

from oracle import oracle as O


def bin_means(values_axis, nbins, xmin, xmax):
    """binner_t<1, 0, double>: per-bin mean of the values (0 where a bin is empty)."""
    size = (xmax - xmin) / nbins
    data = np.zeros(nbins)
    count = np.zeros(nbins, dtype=np.int64)
    for x, v in values_axis:
        if not (xmin <= x <= xmax):
            continue
        i = nbins - 1 if x == xmax else int((x - xmin) / size)
        data[i] += v
        count[i] += 1
    out = np.where(count > 0, data / np.maximum(count, 1), data)
    return out, size


def weighted_line(x, y, w):
    """Weighted least-squares line y ~ c0 + c1 x over the points of nonzero weight."""
    m = w > 0
    x, y, w = x[m], y[m], w[m]
    W = w.sum()
    xm, ym = (w * x).sum() / W, (w * y).sum() / W
    c1 = (w * (x - xm) * (y - ym)).sum() / (w * (x - xm) ** 2).sum()
    return ym - c1 * xm, c1


def kernel_smoothing(y_in, sigma, log=True):
    """Local linear regression of log(y_in) with a Gaussian kernel of width
    sigma (in bins), exponentiated (log=False, a planted mistake: of y_in)."""
    y_in = np.asarray(y_in, dtype=np.float64)
    N = len(y_in)
    x = np.arange(N, dtype=np.float64)
    with np.errstate(divide="ignore"):
        ly = np.log(y_in) if log else y_in
    H = max(min(int(3 * sigma), N), 1)
    K = 2 * H + 1
    kernel = np.exp(-(((np.arange(K) - H) / sigma) ** 2) / 2)
    dead = y_in <= 0
    out = np.empty(N)
    for i in range(N):
        j = np.arange(max(0, i - H), min(N, i + H + 1))
        w = kernel[j - i + H] * (~dead[j])
        c0, c1 = weighted_line(x[j], ly[j], w)
        out[i] = c0 + x[i] * c1
    return np.exp(out) if log else out


def loo_cost(y, sigma, upper):
    """The cost that optimize_sigma minimises (projection.hpp:82-123): the
    squared error of the smoothing of y with the kernel's centre weight set to
    zero (each bin predicted from its neighbours), the kernel's half width
    fixed by `upper`, |sigma| floored at 0.3 bins, weights scaled so that the
    first neighbour weighs 1."""
    y = np.asarray(y, dtype=np.float64)
    N = len(y)
    x = np.arange(N, dtype=np.float64)
    s = max(abs(sigma), 0.3)
    dead = y <= 0
    with np.errstate(divide="ignore"):
        ly = np.where(dead, 0.0, np.log(np.where(dead, 1.0, y)))
    H = max(min(int(3 * upper), N), 1)
    kernel = np.exp(-(((np.arange(2 * H + 1) - H) / s) ** 2) / 2)
    kernel[H] = 0.0
    kernel = kernel / kernel[H + 1]
    out = np.empty(N)
    for i in range(N):
        j = np.arange(max(0, i - H), min(N, i + H + 1))
        w = kernel[j - i + H] * (~dead[j])
        c0, c1 = weighted_line(x[j], ly[j], w)
        out[i] = c0 + x[i] * c1
    return float(np.sum((np.exp(out) - y) ** 2))


def linear_interp(xa, ya, x):
    """gsl_interp_linear with the binary-search index (x inside [xa[0], xa[-1]])."""
    i = int(np.searchsorted(xa, x, side="right")) - 1
    i = min(max(i, 0), len(xa) - 2)
    return ya[i] + (x - xa[i]) / (xa[i + 1] - xa[i]) * (ya[i + 1] - ya[i])


def akima_coeffs(x, y, flat_ends=False):
    """Per-interval (b, c, d) of Akima's spline through (x, y), non-periodic:
    slopes m_i of the intervals, extended by two on each side; the knot slope
    weights the neighbouring slopes by the jumps of the slopes beyond them."""
    n = len(x)
    m = np.empty(n + 3)  # m[k + 2] = slope of interval k, k = -2 .. n
    m[2:n + 1] = np.diff(y) / np.diff(x)
    m[0] = 3 * m[2] - 2 * m[3]
    m[1] = 2 * m[2] - m[3]
    m[n + 1] = 2 * m[n] - m[n - 1]
    m[n + 2] = 3 * m[n] - 2 * m[n - 1]
    if flat_ends:  # a planted mistake, for the tests
        m[0] = m[1] = m[2]
        m[n + 1] = m[n + 2] = m[n]
    b, c, d = np.zeros(n - 1), np.zeros(n - 1), np.zeros(n - 1)
    for i in range(n - 1):
        mi = lambda k: m[i + k + 2]  # slope of interval i + k
        ne = abs(mi(1) - mi(0)) + abs(mi(-1) - mi(-2))
        if ne == 0.0:
            b[i], c[i], d[i] = mi(0), 0.0, 0.0
            continue
        h = x[i + 1] - x[i]
        ne_next = abs(mi(2) - mi(1)) + abs(mi(0) - mi(-1))
        a_i = abs(mi(-1) - mi(-2)) / ne
        if ne_next == 0.0:
            t_next = mi(0)
        else:
            a_n = abs(mi(0) - mi(-1)) / ne_next
            t_next = (1 - a_n) * mi(0) + a_n * mi(1)
        b[i] = (1 - a_i) * mi(-1) + a_i * mi(0)
        c[i] = (3 * mi(0) - 2 * b[i] - t_next) / h
        d[i] = (b[i] + t_next - 2 * mi(0)) / (h * h)
    return b, c, d


def akima_eval(x, y, coeffs, xv):
    b, c, d = coeffs
    i = int(np.searchsorted(x, xv, side="right")) - 1
    i = min(max(i, 0), len(x) - 2)
    dx = xv - x[i]
    return y[i] + dx * (b[i] + dx * (c[i] + dx * d[i]))


def mesh_eval(vals, t):
    """A function on the uniform mesh of [0, 1] with len(vals) points, evaluated
    at t by linear interpolation between the neighbouring mesh points."""
    n = len(vals)
    a = t * (n - 1)
    i = min(max(int(np.floor(a)), 0), n - 2)
    w = a - i
    return (1 - w) * vals[i] + w * vals[i + 1]


class Projection:
    """warper_projection_t(warped_integrand, order, num_bins, nr_samples, sigma,
    optimize_sigma=False) or, with `samples` = (l [n][order], |values| [n]),
    the same construction from given pilot samples.  `plant` (tests only)
    names a deliberate mistake: "unit_bins" (pilot binned on [0, 1], not on
    [-d/2, 1 + d/2]), "flat_ends" (Akima slopes extended flat at the ends),
    "linear_smooth" (smoothing of the values, not their logarithm), "left_rect"
    (left-rectangle CDF), "half_sigma" (sigma taken in units of half a bin)."""

    def __init__(self, warped_integrand, order, num_bins, nr_samples, sigma, samples=None, plant=None):
        if samples is None:
            l = O.Sobol(order, 0).points(nr_samples)
            vals = np.array([abs(warped_integrand(li)) for li in l])
        else:
            l, vals = samples
        self.order, self.num_bins = order, num_bins
        delta = 1.0 / (num_bins - 1.0)
        mesh = np.array([k * delta for k in range(num_bins)])
        self.fn, self.F, self.Finv, self.norm, self.xi, self.sigmas = [], [], [], [], [], []
        n_inv = 1 + 5 * (num_bins - 1)
        for a in range(order):
            lo, hi = (0.0, 1.0) if plant == "unit_bins" else (-delta / 2, 1 + delta / 2)
            data, size = bin_means(zip(l[:, a], vals), num_bins, lo, hi)
            self.xi.append(data)
            self.sigmas.append(sigma)
            s_bins = sigma / size * (2.0 if plant == "half_sigma" else 1.0)
            if plant == "linear_smooth":
                sm = kernel_smoothing(data, s_bins, log=False)
            else:
                sm = kernel_smoothing(data, s_bins)
            times = lo + (np.arange(num_bins) + 0.5) * size
            times[0], times[-1] = 0.0, 1.0
            fn = np.array([linear_interp(times, sm, t) for t in mesh])
            F = np.zeros(num_bins)
            for k in range(1, num_bins):
                F[k] = mesh_eval(fn, mesh[k] - (delta if plant == "left_rect" else delta / 2)) * delta
            F = np.cumsum(F)
            norm = F[-1]
            F = F / norm
            co = akima_coeffs(F, mesh, flat_ends=plant == "flat_ends")
            dl = 1.0 / (n_inv - 1)
            Finv = np.array([akima_eval(F, mesh, co, j * dl) for j in range(n_inv)])
            self.fn.append(fn)
            self.F.append(F)
            self.Finv.append(Finv)
            self.norm.append(norm)

    def ui_from_li(self, l):
        return np.array([mesh_eval(self.Finv[i], li) for i, li in enumerate(l)])

    def li_from_ui(self, u):
        return np.array([mesh_eval(self.F[i], ui) for i, ui in enumerate(u)])

    def jacobian_reverse(self, l):
        u = self.ui_from_li(l)
        return float(np.prod([self.norm[i] / mesh_eval(self.fn[i], u[i]) for i in range(len(l))]))

    def jacobian_forward(self, u):
        return float(np.prod([mesh_eval(self.fn[i], ui) for i, ui in enumerate(u)]))

    def map_reverse(self, l):
        return self.ui_from_li(l), self.jacobian_reverse(l)

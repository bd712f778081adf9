"""B-spline radial basis (host-side setup, numpy).

Follows pyprop/bspline/BSpline.py:5-235 (breakpoints, continuity, knots,
per-interval Gauss-Legendre rule, global quadrature grid) and
core/transform/bspline/bspline.cpp:26-161,313-466 (B-spline values and
derivatives, basis-pair integrals <B_i| f d^d |B_j> over the common support,
with the (-1)^d factor when the pair is transposed).  The values here are
computed with a vectorised de Boor recursion over all quadrature points at
once instead of the reference's per-point recursive calls; the integrals are
the same sums.
"""
import numpy as np


# ---------------------------------------------------------------- breakpoints
# pyprop/bspline/breakpointsequences.py.  The formulas define the grid (a
# config names a sequence and its parameters), so they are the reference's;
# the code evaluates them as numpy arrays and returns lists, as the reference.
def _exp_points(r0, r1, n, gamma):
    """n points from r0 to r1 with spacings growing by exp(gamma/(n-1))
    (breakpointsequences.py:6-28)."""
    s = np.arange(n, dtype=np.float64) / float(n - 1)
    return r0 + (r1 - r0) * np.expm1(gamma * s) / np.expm1(gamma)


def _closed_frange(start, stop, step):
    """matplotlib.mlab.frange(start, stop, step) as the reference imports it
    through pylab: the closed interval, round((stop - start)/step) + 1 points
    start + i*step (Python 2 round, half away from zero)."""
    q = (stop - start) / step
    npts = int(np.sign(q) * np.floor(abs(q) + 0.5)) + 1
    return start + step * np.arange(max(npts, 0), dtype=np.float64)


def LinearBreakpointSequence(rmin, rmax, n):
    """n equidistant points on [rmin, rmax] (breakpointsequences.py:31-38)."""
    return list(rmin + (rmax - rmin) * np.arange(n) / float(n - 1))


def ExponentialBreakpointSequence(rmin, rmax, n, gamma):
    """breakpointsequences.py:6-28."""
    return list(_exp_points(rmin, rmax, n, gamma))


def ExponentialLinearBreakpointSequence(rmin, rpartition, rmax, n, gamma):
    """n exponential points on [rmin, rpartition], then points spaced like
    the last two of them from rpartition + h up to rmax, both ends included
    (breakpointsequences.py:72-90)."""
    inner = _exp_points(rmin, rpartition, n, gamma)
    h = inner[-1] - inner[-2]
    return list(inner) + list(_closed_frange(rpartition + h, rmax, h))


def CenterExponentialLinearBreakpointSequence(rmin, rpartition, rmax, n, gamma):
    """The exponential-linear sequence from the centre (rmin + rmax)/2
    outwards, mirrored through x -> -x (breakpointsequences.py:92-121; the
    mirror is about zero, so the grid is centred on (rmin + rmax)/2 only when
    that is zero)."""
    c = (rmax + rmin) / 2.0
    right = ExponentialLinearBreakpointSequence(c, rpartition, rmax, n, gamma)
    return [-x for x in reversed(right[1:])] + right


def QuadraticLinearBreakpointSequence(rmin, rmax, n, joinPoint):
    """joinPoint - 1 points on a parabola from rmin, then n - joinPoint + 1
    equidistant points to rmax; r0 is chosen so the slope is continuous at
    the join, 2 alpha (joinPoint - 1) = beta (breakpointsequences.py:41-70)."""
    rmin, rmax, j = float(rmin), float(rmax), int(joinPoint)
    r0 = (rmax * (j - 1) + rmin * (n - j)) / (2.0 * n - j - 1.0)
    alpha = (r0 - rmin) / float(j - 1) ** 2
    beta = (rmax - r0) / float(n - j)
    quad = rmin + alpha * np.arange(j - 1, dtype=np.float64) ** 2
    lin = r0 + beta * np.arange(n - j + 1, dtype=np.float64)
    return list(quad) + list(lin)


def gauss_legendre_rule(n):
    """GaussQuadratureRule (gausslegendrequadrature.py:3-42): Golub-Welsch on the
    Legendre Jacobi matrix, nodes ascending, weights 2 v0^2."""
    beta = np.array([1.0 / (2.0 * np.sqrt(1.0 - (2.0 * k) ** -2)) for k in range(1, n)])
    J = np.diag(beta, -1) + np.diag(beta, 1)
    E, V = np.linalg.eigh(J)
    I = np.argsort(E)
    return E[I], 2.0 * V[0, I] ** 2


class BSplineBasis:
    """Radial B-spline basis; attributes mirror core.BSpline's accessors."""

    def __init__(self, breakpoints, order, continuity, quad_order_additional=0,
                 left_point_continuity=None, right_point_continuity=None):
        self.MaxSplineOrder = k = int(order)
        bp = np.asarray(breakpoints, dtype=np.float64)
        self.Breakpoints = bp
        nbp = len(bp)
        cont = np.full(nbp, k - 1, dtype=int)
        if continuity == "vanilla":
            cont[0] = cont[-1] = 0
        elif continuity == "zero":
            cont[0] = cont[-1] = 1
        elif continuity == "custom":
            cont[0] = left_point_continuity
            cont[-1] = right_point_continuity
        else:
            raise Exception("Unknown continuity type!")
        self.Continuity = cont
        knots = []
        for i in range(nbp):
            knots += [bp[i]] * (k - cont[i])
        self.Knots = np.array(knots)
        self.NumberOfBSplines = len(self.Knots) - k
        nq = k + int(quad_order_additional)
        self.Nodes, self.Weights = gauss_legendre_rule(nq)
        self.nq = nq
        # global grid (BSpline.py:183-235): x = (xi-1)(b-a)/2 + b
        a = bp[:-1][:, None]
        b = bp[1:][:, None]
        self.QuadratureGridGlobal = ((self.Nodes[None, :] - 1.0) * (b - a) / 2.0 + b).reshape(-1)
        self.QuadratureWeightsGlobal = ((b - a) / 2.0 * self.Weights[None, :]).reshape(-1)
        self._tables()
        self._overlap = None

    # ------------------------------------------------------------ tables
    def _tables(self):
        """Nonzero B_i, B_i', B_i'' at every global quadrature point.

        span[q]: first B-spline index whose support contains point q; the k
        nonzero functions are span..span+k-1.  Cox-de Boor recursion with the
        reference's derivative formulas (bspline.cpp:82-161)."""
        k = self.MaxSplineOrder
        t = self.Knots
        x = self.QuadratureGridGlobal
        G = len(x)
        mu = np.searchsorted(t, x, side="right") - 1          # t[mu] <= x < t[mu+1]
        self.Span = mu - (k - 1)
        # values of order p: B_{mu-p+1..mu, p}
        vals = {1: np.ones((G, 1))}
        for p in range(2, k + 1):
            prev = vals[p - 1]
            cur = np.zeros((G, p))
            for r in range(p):
                i = mu - p + 1 + r
                if r >= 1:   # term from B_{i,p-1} (index r-1 in prev)
                    den = t[i + p - 1] - t[i]
                    cur[:, r] += np.where(den > 0, (x - t[i]) / np.where(den > 0, den, 1.0), 0.0) * prev[:, r - 1]
                if r <= p - 2:  # term from B_{i+1,p-1} (index r in prev)
                    den = t[i + p] - t[i + 1]
                    cur[:, r] += np.where(den > 0, (t[i + p] - x) / np.where(den > 0, den, 1.0), 0.0) * prev[:, r]
            vals[p] = cur
        self.B = vals[k]
        # first derivative: (k-1)[B_{i,k-1}/(t_{i+k-1}-t_i) - B_{i+1,k-1}/(t_{i+k}-t_{i+1})]
        prev = vals[k - 1]
        d1 = np.zeros((G, k))
        for r in range(k):
            i = mu - k + 1 + r
            if r >= 1:
                den = t[i + k - 1] - t[i]
                d1[:, r] += np.where(den > 0, (k - 1.0) / np.where(den > 0, den, 1.0), 0.0) * prev[:, r - 1]
            if r <= k - 2:
                den = t[i + k] - t[i + 1]
                d1[:, r] -= np.where(den > 0, (k - 1.0) / np.where(den > 0, den, 1.0), 0.0) * prev[:, r]
        self.D1 = d1
        # second derivative (bspline.cpp:115-161)
        d2 = np.zeros((G, k))
        if k >= 3:
            pp = vals[k - 2]   # B_{mu-k+3..mu, k-2}

            def Bkm2(i):  # value of B_{i,k-2} for vector of indices i
                r = i - (mu - k + 3)
                ok = (r >= 0) & (r <= k - 3)
                return np.where(ok, pp[np.arange(G), np.clip(r, 0, k - 3)], 0.0)

            nk = len(t)

            def T(idx):
                return t[np.clip(idx, 0, nk - 1)]

            def safe(num, den):
                return np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)

            for r in range(k):
                i = mu - k + 1 + r
                B1 = Bkm2(i) * safe(k - 2.0, T(i + k - 2) - T(i))
                B2 = Bkm2(i + 1) * safe(k - 2.0, T(i + k - 1) - T(i + 1))
                B3 = Bkm2(i + 2) * safe(k - 2.0, T(i + k) - T(i + 2))
                d2[:, r] = safe(k - 1.0, T(i + k - 1) - T(i)) * (B1 - B2) - safe(k - 1.0, T(i + k) - T(i + 1)) * (B2 - B3)
        self.D2 = d2

    def table(self, d):
        return {0: self.B, 1: self.D1, 2: self.D2}[d]

    # ------------------------------------------------------------ integrals
    def _projector(self, d):
        """Sparse (G, N*(2k-1)) map from weighted samples w*f on the global
        quadrature grid to BandNH storage of <B_i| f d^d |B_j>."""
        cache = self.__dict__.setdefault("_proj_cache", {})
        if d in cache:
            return cache[d]
        import scipy.sparse
        k = self.MaxSplineOrder
        N = self.NumberOfBSplines
        s = k - 1
        bw = 2 * s + 1
        Bt = self.B
        Dt = self.table(d)
        span = self.Span
        G = Bt.shape[0]
        sign = (-1.0) ** d
        rows, cols, vals = [], [], []
        q = np.arange(G)
        for a_ in range(k):
            for b_ in range(k):
                i = span + a_
                j = span + b_
                valid = (i >= 0) & (i < N) & (j >= 0) & (j < N)
                if a_ <= b_:
                    c = Bt[:, a_] * Dt[:, b_]          # element (i, j), i <= j
                else:
                    c = sign * (Bt[:, b_] * Dt[:, a_])  # i > j: (-1)^d * B_j D_i
                idx = j * bw + s + i - j
                rows.append(q[valid])
                cols.append(idx[valid])
                vals.append(c[valid])
        M = scipy.sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                    shape=(G, N * bw))
        cache[d] = M
        return M

    def project_bandnh(self, f, d=0, device=False):
        """<B_i| f d^d |B_j> in BandNH storage, shape (..., N*(2k-1)).

        f: samples on the global quadrature grid, shape (..., G); element
        (i,j) is stored at j*(2k-1) + (k-1)+i-j (GeometryInfo.py:433-454).
        For i > j the reference integrates the transposed pair and multiplies
        by (-1)^d (bspline.cpp:380-392).  device: the sparse projection runs
        on the GPU (pp_projector_apply, the same sums) and the result comes
        back as a numpy array."""
        f = np.asarray(f)
        lead = f.shape[:-1]
        fw = (f * self.QuadratureWeightsGlobal).reshape(-1, f.shape[-1])
        M = self._projector(d)
        if device:
            from . import device as dev
            cache = self.__dict__.setdefault("_dev_proj", {})
            if d not in cache:
                cache[d] = dev.Projector(M.T.tocsr())
            out = cache[d].apply(np.ascontiguousarray(fw, dtype=np.complex128))
        else:
            out = np.asarray((M.T @ fw.T).T)
        return out.reshape(lead + (M.shape[1],))

    def overlap_upper(self):
        """S in upper hermitian band storage (N, k): [j*k + (k-1)+i-j] = S(i,j), i<=j
        (overlapmatrix.cpp:22-32)."""
        if self._overlap is None:
            k = self.MaxSplineOrder
            N = self.NumberOfBSplines
            s = k - 1
            band = self.project_bandnh(np.ones_like(self.QuadratureGridGlobal), 0)
            up = np.zeros(N * k, dtype=np.complex128)
            for j in range(N):
                for i in range(max(0, j - s), j + 1):
                    up[j * k + s + i - j] = band[j * (2 * s + 1) + s + i - j]
            self._overlap = up
        return self._overlap

    def overlap_dense(self):
        k = self.MaxSplineOrder
        N = self.NumberOfBSplines
        s = k - 1
        up = self.overlap_upper()
        S = np.zeros((N, N))
        for j in range(N):
            for i in range(max(0, j - s), j + 1):
                S[i, j] = S[j, i] = up[j * k + s + i - j].real
        return S


def InitBSpline(conf):
    """InitBSpline (pyprop/bspline/BSpline.py:5-11) from a representation section."""
    t = conf.bpstype
    if t == "linear":
        bps = LinearBreakpointSequence(conf.xmin, conf.xmax, conf.xsize)
    elif t == "exponential":
        bps = ExponentialBreakpointSequence(conf.xmin, conf.xmax, conf.xsize, conf.gamma)
    elif t == "quadraticlinear":
        bps = QuadraticLinearBreakpointSequence(conf.xmin, conf.xmax, conf.xsize, conf.joinpoint)
    elif t == "exponentiallinear":
        bps = ExponentialLinearBreakpointSequence(conf.xmin, conf.xpartition, conf.xmax, conf.xsize, conf.gamma)
    elif t == "centerexponentiallinear":
        bps = CenterExponentialLinearBreakpointSequence(conf.xmin, conf.xpartition, conf.xmax, conf.xsize,
                                                        conf.gamma)
    else:
        raise NameError("Sequence not recognized!")
    return BSplineBasis(bps, conf.order, conf.continuity,
                        getattr(conf, "quad_order_additional", 0),
                        getattr(conf, "left_point_continuity", None),
                        getattr(conf, "right_point_continuity", None))

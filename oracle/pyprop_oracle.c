/*
 * pyprop_oracle.c -- CPU restatement of pyprop's tensor-potential Krylov
 * propagation path, in complex128, no third-party BLAS.  OpenMP runs only
 * over independent outputs (angular rows, fibers, whole dot products,
 * elements), so every thread count gives bitwise the result of one thread.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing in the product.  Every function cites the reference file:line it
 * follows (paths relative to the pyprop source tree).
 *
 * Contents
 *   Gauss-Legendre rule ........ pyprop/bspline/gausslegendrequadrature.py:3-42
 *   B-spline basis ............. core/transform/bspline/bspline.cpp:26-161,313-466
 *   angular (reduced sph.) ..... core/transform/reducedspherical/reducedsphericaltools.cpp:284-365
 *   band Cholesky / solves ..... core/representation/overlapmatrix.cpp:9-37,137-149 (zpbtrf, ztbsv, ztbmv)
 *   tensor multiply ............ core/tensorpotential/tensorpotentialmultiply_generator.py:302-763
 *   balanced matvec ............ pyprop/tensorpotential/BasisPropagator.py:131-159
 *   pAMP Krylov step ........... core/krylov/pamp/pamp/pamp.h:271-669
 *   Expokit zgexpv / zgpadm .... core/krylov/expokit/f2c/zgexpv.cpp:32-586, zgpadm.cpp:28-285
 */
#include <complex.h>
#include <omp.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

typedef double complex zc;

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Gauss-Legendre on [-1,1], nodes ascending.                          */
/* gausslegendrequadrature.py:25-42 solves the Jacobi eigenproblem;    */
/* Newton on P_n gives the same rule to rounding.                      */
/* ------------------------------------------------------------------ */
ORC_API int orc_gauss_legendre(int n, double *x, double *w)
{
    if (n < 1) return -1;
    for (int i = 0; i < n; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        /* recompute derivative at converged node */
        double p1 = 1.0, p2 = 0.0;
        for (int j = 1; j <= n; j++) {
            double p3 = p2; p2 = p1;
            p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        x[n - 1 - i] = z;               /* cos decreasing in i -> ascending */
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* B-spline basis: knots from breakpoints + continuity                 */
/* (pyprop/bspline/BSpline.py:133-161), Cox-de Boor with the eps guard  */
/* (bspline.cpp:26-76), derivatives (bspline.cpp:82-161).               */
/* ------------------------------------------------------------------ */
typedef struct {
    int order;        /* k */
    int nbp;          /* breakpoints */
    int nknots;
    int N;            /* number of B-splines = nknots - k (BSpline.py:33) */
    int nq;           /* Gauss points per interval */
    double *bp, *knots, *gx, *gw;
    double eps;
} orc_bspline;

static double bs_eval(const orc_bspline *b, double x, int k, int i)
{
    const double *t = b->knots;
    if (k == 1) return (t[i] <= x && x < t[i + 1]) ? 1.0 : 0.0;
    double r = 0.0;
    double B00 = bs_eval(b, x, k - 1, i);
    double B01 = bs_eval(b, x, k - 1, i + 1);
    if (B00 > b->eps) r += (x - t[i]) / (t[i + k - 1] - t[i]) * B00;
    if (B01 > b->eps) r += (t[i + k] - x) / (t[i + k] - t[i + 1]) * B01;
    return r;
}

static double bs_d1(const orc_bspline *b, double x, int k, int i)
{
    const double *t = b->knots;
    double r = 0.0;
    double B1 = bs_eval(b, x, k - 1, i), B2 = bs_eval(b, x, k - 1, i + 1);
    if (B1 > b->eps) r += (k - 1.0) / (t[i + k - 1] - t[i]) * B1;
    if (B2 > b->eps) r -= (k - 1.0) / (t[i + k] - t[i + 1]) * B2;
    return r;
}

static double bs_d2(const orc_bspline *b, double x, int k, int i)
{
    const double *t = b->knots;
    double r = 0.0;
    double B1 = bs_eval(b, x, k - 2, i), B2 = bs_eval(b, x, k - 2, i + 1), B3 = bs_eval(b, x, k - 2, i + 2);
    if (B1 > b->eps) B1 *= (k - 2.0) / (t[i + k - 2] - t[i]);
    if (B2 > b->eps) B2 *= (k - 2.0) / (t[i + k - 1] - t[i + 1]);
    if (B3 > b->eps) B3 *= (k - 2.0) / (t[i + k] - t[i + 2]);
    if (B1 > b->eps || B2 > b->eps) r += (k - 1.0) / (t[i + k - 1] - t[i]) * (B1 - B2);
    if (B2 > b->eps || B3 > b->eps) r -= (k - 1.0) / (t[i + k] - t[i + 1]) * (B2 - B3);
    return r;
}

ORC_API double orc_bspline_value(const orc_bspline *b, double x, int i, int deriv)
{
    if (deriv == 0) return bs_eval(b, x, b->order, i);
    if (deriv == 1) return bs_d1(b, x, b->order, i);
    return bs_d2(b, x, b->order, i);
}

/* continuity: value per breakpoint, endpoints given, interior = k-1      */
/* ('zero' -> ends 1, 'vanilla' -> ends 0; BSpline.py:102-122)            */
ORC_API orc_bspline *orc_bspline_create(const double *bp, int nbp, int order,
                                        int cont_left, int cont_right, int quad_add)
{
    orc_bspline *b = calloc(1, sizeof(*b));
    b->order = order; b->nbp = nbp; b->eps = 1e-15;
    b->bp = malloc(sizeof(double) * nbp);
    memcpy(b->bp, bp, sizeof(double) * nbp);
    int cap = nbp * order;
    b->knots = malloc(sizeof(double) * cap);
    int nk = 0;
    for (int i = 0; i < nbp; i++) {
        int cont = (i == 0) ? cont_left : (i == nbp - 1) ? cont_right : order - 1;
        int mult = order - cont;
        for (int j = 0; j < mult; j++) b->knots[nk++] = bp[i];
    }
    b->nknots = nk;
    b->N = nk - order;
    b->nq = order + quad_add;
    b->gx = malloc(sizeof(double) * b->nq);
    b->gw = malloc(sizeof(double) * b->nq);
    orc_gauss_legendre(b->nq, b->gx, b->gw);
    return b;
}

ORC_API void orc_bspline_destroy(orc_bspline *b)
{
    if (!b) return;
    free(b->bp); free(b->knots); free(b->gx); free(b->gw); free(b);
}

ORC_API int orc_bspline_count(const orc_bspline *b) { return b->N; }
ORC_API int orc_bspline_grid_size(const orc_bspline *b) { return b->nq * (b->nbp - 1); }

/* global quadrature grid and weights (BSpline.py:183-235):             */
/* x = (xi-1)(b-a)/2 + b, w = (b-a)/2 * wi (bspline.h ScaleAndTranslate) */
ORC_API void orc_bspline_grid(const orc_bspline *b, double *x, double *w)
{
    int idx = 0;
    for (int j = 0; j < b->nbp - 1; j++) {
        double a = b->bp[j], c = b->bp[j + 1];
        for (int q = 0; q < b->nq; q++) {
            x[idx] = (b->gx[q] - 1.0) * (c - a) / 2.0 + c;
            w[idx] = (c - a) / 2.0 * b->gw[q];
            idx++;
        }
    }
}

/* index of the breakpoint equal to a knot value */
static int bp_index(const orc_bspline *b, double v)
{
    for (int i = 0; i < b->nbp; i++) if (fabs(b->bp[i] - v) < b->eps) return i;
    return -1;
}

/* BSplineGlobalOverlapIntegral (bspline.cpp:373-466): <B_i| f d^d |B_j>, */
/* pair reordered to i<j with factor (-1)^d, summed over the common       */
/* support with the per-interval Gauss rule; f sampled on the global grid */
static zc bs_global_integral(const orc_bspline *b, const zc *fgrid, int d, int i, int j)
{
    double factor = 1.0;
    if (i > j) { int t = i; i = j; j = t; factor = pow(-1.0, d); }
    if (j - i >= b->order) return 0.0;
    double lo = b->knots[j], hi = b->knots[i + b->order];
    int blo = bp_index(b, lo), bhi = bp_index(b, hi);
    zc acc = 0.0;
    for (int iv = blo; iv < bhi; iv++) {
        double a = b->bp[iv], c = b->bp[iv + 1];
        for (int q = 0; q < b->nq; q++) {
            double x = (b->gx[q] - 1.0) * (c - a) / 2.0 + c;
            double w = (c - a) / 2.0 * b->gw[q];
            double Bl = bs_eval(b, x, b->order, i);
            double Br = (d == 0) ? bs_eval(b, x, b->order, j)
                       : (d == 1) ? bs_d1(b, x, b->order, j) : bs_d2(b, x, b->order, j);
            acc += w * fgrid[iv * b->nq + q] * Bl * Br;
        }
    }
    return factor * acc;
}

ORC_API void orc_bspline_integral(const orc_bspline *b, const double *fre, const double *fim,
                                  int d, int i, int j, double *out)
{
    int G = orc_bspline_grid_size(b);
    zc *f = malloc(sizeof(zc) * G);
    for (int g = 0; g < G; g++) f[g] = fre[g] + I * (fim ? fim[g] : 0.0);
    zc v = bs_global_integral(b, f, d, i, j);
    out[0] = creal(v); out[1] = cimag(v);
    free(f);
}

/* Project f into BandNH storage: (N, 2s+1), element (i,j) at           */
/* j*(2s+1) + s+i-j, out-of-range slots get the pair (0,N-1) value 0     */
/* (GeometryInfo.py:433-454, basis_bspline.h:11-47).                     */
ORC_API void orc_bspline_project_bandnh(const orc_bspline *b, const double *fre, const double *fim,
                                        int d, double *out /* complex interleaved N*(2s+1) */)
{
    int N = b->N, s = b->order - 1, bw = 2 * s + 1;
    int G = orc_bspline_grid_size(b);
    zc *f = malloc(sizeof(zc) * G);
    for (int g = 0; g < G; g++) f[g] = fre[g] + I * (fim ? fim[g] : 0.0);
    zc *o = (zc *)out;
    for (int p = 0; p < N * bw; p++) o[p] = bs_global_integral(b, f, d, 0, N - 1);
    for (int i = 0; i < N; i++)
        for (int j = i - s; j <= i + s; j++) {
            if (j < 0 || j >= N) continue;
            o[j * bw + s + i - j] = bs_global_integral(b, f, d, i, j);
        }
    free(f);
}

/* Overlap in upper hermitian band storage (overlapmatrix.cpp:22-32):  */
/* ab[j*(kd+1) + kd+i-j] = S(i,j), i<=j; evaluator = f==1, d=0.        */
ORC_API void orc_bspline_overlap_upper(const orc_bspline *b, double *out /* complex N*(kd+1) */)
{
    int N = b->N, kd = b->order - 1;
    int G = orc_bspline_grid_size(b);
    zc *f = malloc(sizeof(zc) * G);
    for (int g = 0; g < G; g++) f[g] = 1.0;
    zc *o = (zc *)out;
    memset(o, 0, sizeof(zc) * N * (kd + 1));
    for (int i = 0; i < N; i++)
        for (int j = i; j < N && j <= i + kd; j++)
            o[j * (kd + 1) + kd + i - j] = bs_global_integral(b, f, 0, i, j);
    free(f);
}

/* ------------------------------------------------------------------ */
/* LAPACK/BLAS band routines, upper storage, column j at ab[j*(kd+1)]  */
/* element (i,j) at ab[j*(kd+1) + kd + i - j] for max(0,j-kd)<=i<=j.   */
/* ------------------------------------------------------------------ */
#define UB(ab, kd, i, j) ((ab)[(size_t)(j) * ((kd) + 1) + (kd) + (i) - (j)])

/* zpbtrf 'U': S = U^H U (blitzlapack.h:843-856 -> LAPACK zpbtrf).     */
/* Returns LAPACK info (0 ok, >0 leading minor not positive definite). */
ORC_API int orc_zpbtrf_upper(int N, int kd, double *abd)
{
    zc *ab = (zc *)abd;
    for (int j = 0; j < N; j++) {
        /* U(i,j) for i<j already final (computed row-wise below) */
        double d = creal(UB(ab, kd, j, j));
        int i0 = j - kd < 0 ? 0 : j - kd;
        for (int i = i0; i < j; i++) d -= creal(conj(UB(ab, kd, i, j)) * UB(ab, kd, i, j));
        if (d <= 0.0) return j + 1;
        d = sqrt(d);
        UB(ab, kd, j, j) = d;
        /* row j of U: U(j,c) = (S(j,c) - sum_i conj(U(i,j)) U(i,c)) / U(j,j) */
        for (int c = j + 1; c < N && c <= j + kd; c++) {
            zc s = UB(ab, kd, j, c);
            int ii0 = c - kd < 0 ? 0 : c - kd;
            for (int i = ii0; i < j; i++) s -= conj(UB(ab, kd, i, j)) * UB(ab, kd, i, c);
            UB(ab, kd, j, c) = s / d;
        }
    }
    return 0;
}

/* ztbsv 'U', trans 'N' (x := U^-1 x) or 'C' (x := U^-H x), non-unit,  */
/* strided fiber (piram/blitzblas.h:454-486).                          */
static void ztbsv_u(int conjt, int N, int kd, const zc *ab, zc *x, long inc)
{
    if (!conjt) {
        for (int j = N - 1; j >= 0; j--) {
            zc t = x[j * inc] / UB(ab, kd, j, j);
            x[j * inc] = t;
            int i0 = j - kd < 0 ? 0 : j - kd;
            for (int i = i0; i < j; i++) x[i * inc] -= t * UB(ab, kd, i, j);
        }
    } else {
        for (int j = 0; j < N; j++) {
            zc t = x[j * inc];
            int i0 = j - kd < 0 ? 0 : j - kd;
            for (int i = i0; i < j; i++) t -= conj(UB(ab, kd, i, j)) * x[i * inc];
            x[j * inc] = t / conj(UB(ab, kd, j, j));
        }
    }
}

/* ztbmv 'U', trans 'N': x := U x */
static void ztbmv_un(int N, int kd, const zc *ab, zc *x, long inc)
{
    for (int i = 0; i < N; i++) {
        zc t = 0.0;
        int j1 = i + kd < N - 1 ? i + kd : N - 1;
        for (int j = i; j <= j1; j++) t += UB(ab, kd, i, j) * x[j * inc];
        x[i * inc] = t;
    }
}

/* zhbmv 'U': y := S x with S in upper hermitian band storage */
static void zhbmv_u(int N, int kd, const zc *ab, const zc *x, zc *y, long inc)
{
    for (int i = 0; i < N; i++) {
        zc t = 0.0;
        int j0 = i - kd < 0 ? 0 : i - kd, j1 = i + kd < N - 1 ? i + kd : N - 1;
        for (int j = j0; j <= j1; j++) {
            zc a = (j >= i) ? UB(ab, kd, i, j) : conj(UB(ab, kd, j, i));
            t += a * x[j * inc];
        }
        y[i * inc] = t;
    }
}

/* apply along rank r of a C-ordered tensor (MapToRank3, blitztricks.h:63-107) */
typedef enum { OP_SOLVE_N = 0, OP_SOLVE_C = 1, OP_MUL_N = 2 } orc_band_op;

ORC_API void orc_band_apply(int op, int rank, int R, const int *shape, int kd,
                            const double *abd, double *xd)
{
    const zc *ab = (const zc *)abd;
    zc *x = (zc *)xd;
    long pre = 1, post = 1;
    for (int i = 0; i < rank; i++) pre *= shape[i];
    for (int i = rank + 1; i < R; i++) post *= shape[i];
    int N = shape[rank];
    /* fibers are independent: any thread count gives the same result */
    #pragma omp parallel for collapse(2) schedule(static)
    for (long a = 0; a < pre; a++)
        for (long c = 0; c < post; c++) {
            zc *f = x + a * N * post + c;
            if (op == OP_SOLVE_N) ztbsv_u(0, N, kd, ab, f, post);
            else if (op == OP_SOLVE_C) ztbsv_u(1, N, kd, ab, f, post);
            else ztbmv_un(N, kd, ab, f, post);
        }
}

/* ------------------------------------------------------------------ */
/* Angular basis (reducedsphericaltools.cpp:284-365), m = 0 or general */
/* ------------------------------------------------------------------ */
/* P~[q*L + l], D~[q*L + l] for l = |m|..lmax, q = 0..lmax (x ascending). */
ORC_API int orc_angular_tables(int lmax, int m, double *xq, double *wq, double *P, double *D)
{
    int nth = lmax + 1;
    int am = abs(m);
    int L = lmax + 1 - am;
    orc_gauss_legendre(nth, xq, wq);
    for (int q = 0; q < nth; q++) {
        double x = cos(acos(xq[q]));     /* theta grid round trip, shtools.cpp:394 */
        double somx2 = sqrt((1.0 - x) * (1.0 + x));
        /* P_m^m */
        double pmm = 1.0, fact = 1.0;
        for (int k = 0; k < am; k++) { pmm *= -fact * somx2; fact += 2.0; }
        double plm_prev = 0.0, plm = pmm;
        for (int l = am; l <= lmax; l++) {
            if (l == am) { plm = pmm; plm_prev = 0.0; }
            else if (l == am + 1) { plm_prev = pmm; plm = x * (2.0 * am + 1.0) * pmm; }
            else {
                double pll = (x * (2.0 * l - 1.0) * plm - (l + am - 1.0) * plm_prev) / (l - am);
                plm_prev = plm; plm = pll;
            }
            double lf = 1.0;
            for (int t = l - am + 1; t <= l + am; t++) lf *= t;  /* (l+|m|)!/(l-|m|)! */
            double sign = (m > 0) ? pow(-1.0, m) : 1.0;
            double norm = sign * sqrt((2.0 * l + 1.0) / (4.0 * M_PI) / lf) * sqrt(2.0 * M_PI);
            P[q * L + (l - am)] = plm * norm;
        }
        for (int l = am; l <= lmax; l++) {
            int li = l - am;
            if (li > 0) {
                double nrm = sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) * (double)(l - am) / (double)(l + am));
                D[q * L + li] = l * x * P[q * L + li] - (l + m) * nrm * P[q * L + li - 1];
            } else D[q * L + li] = 0.0;
        }
    }
    return L;
}

/* ------------------------------------------------------------------ */
/* Tensor multiply: dest += scaling * (x_r V_r) src                     */
/* Storage ids per rank follow the generator's snippets:               */
/*  0 Ident (:436-462), 1 Diag (:403-433), 2 Simp (:302-363),           */
/*  3 Herm (:370-398), 4 BandNH (:622-763), 5 Band (:467-618).          */
/* ------------------------------------------------------------------ */
enum { ST_IDENT = 0, ST_DIAG = 1, ST_SIMP = 2, ST_HERM = 3, ST_BANDNH = 4, ST_BAND = 5 };

/* flags of an entry: conjugate V; the innermost Band rank's zhbmv call     */
/* (:562-575) replaces the conjugation the outer ranks asked for by its own */
/* (ENT_OWN); zhbmv reads only the real part of the diagonal (ENT_REAL).    */
enum { ENT_CONJ = 1, ENT_OWN = 2, ENT_REAL = 4 };
typedef struct { int row, col; long pidx; int flags; } orc_entry;

static int build_entries(int storage, int potExt, int srcExt, const int *pairs, int npairs,
                         int innermost, orc_entry **out)
{
    int cap = 16 + 2 * potExt + 2 * npairs + srcExt * 4;
    if (storage == ST_BANDNH || storage == ST_BAND) cap = potExt * 2 + 16;
    orc_entry *e = malloc(sizeof(orc_entry) * cap);
    int n = 0;
    switch (storage) {
    case ST_IDENT:
        for (int r = 0; r < srcExt; r++) e[n++] = (orc_entry){ r, r, 0, 0 };
        break;
    case ST_DIAG:
        for (int i = 0; i < potExt; i++) e[n++] = (orc_entry){ i, i, i, 0 };
        break;
    case ST_SIMP:
        for (int i = 0; i < potExt; i++) e[n++] = (orc_entry){ pairs[2 * i], pairs[2 * i + 1], i, 0 };
        break;
    case ST_HERM:
        for (int i = 0; i < potExt; i++) {
            e[n++] = (orc_entry){ pairs[2 * i], pairs[2 * i + 1], i, 0 };
            e[n++] = (orc_entry){ pairs[2 * i + 1], pairs[2 * i], i, ENT_CONJ };
        }
        break;
    case ST_BANDNH: {
        int N = srcExt, bandCount = potExt / N, s = (bandCount - 1) / 2;
        for (int col = 0; col < N; col++)
            for (int row = col - s; row <= col + s; row++) {
                if (row < 0 || row >= N) continue;
                e[n++] = (orc_entry){ row, col, (long)col * bandCount + s + row - col, 0 };
            }
        break;
    }
    case ST_BAND: {
        /* band storage (N, k): slot j*k + d, d = 0..k-1, belongs to the pair */
        /* (j, j+d).  Innermost rank, zhbmv 'L' (:562-575): A[j+d][j] = V,     */
        /* A[j][j+d] = conj V, diagonal Re V, the outer conjugation dropped.   */
        /* Outer rank, the explicit loops (:513-551): A[j][j+d] = V,           */
        /* A[j+d][j] = conj V, complex diagonal.                               */
        int N = srcExt, k = potExt / N;
        for (int j = 0; j < N; j++)
            for (int d = 0; d < k && j + d < N; d++) {
                long slot = (long)j * k + d;
                if (innermost) {
                    if (d == 0) { e[n++] = (orc_entry){ j, j, slot, ENT_OWN | ENT_REAL }; continue; }
                    e[n++] = (orc_entry){ j + d, j, slot, ENT_OWN };
                    e[n++] = (orc_entry){ j, j + d, slot, ENT_OWN | ENT_CONJ };
                } else {
                    e[n++] = (orc_entry){ j, j + d, slot, 0 };
                    if (d > 0) e[n++] = (orc_entry){ j + d, j, slot, ENT_CONJ };
                }
            }
        break;
    }
    default: free(e); return -1;
    }
    *out = e;
    return n;
}

/* pairs[r] may be NULL for storages without pairs; potShape = extents of V */
ORC_API int orc_tp_multiply(int R, const int *storage, const int *potShape, const double *potd,
                            double scaling, const int *shape, const double *srcd, double *dstd,
                            const int *const *pairs)
{
    const zc *V = (const zc *)potd;
    const zc *src = (const zc *)srcd;
    zc *dst = (zc *)dstd;
    if (R < 1 || R > 3) return -1;
    orc_entry *E[3] = { 0, 0, 0 };
    int ne[3] = { 1, 1, 1 };
    for (int r = 0; r < R; r++) {
        ne[r] = build_entries(storage[r], potShape[r], shape[r], pairs ? pairs[r] : NULL,
                              potShape[r], r == R - 1, &E[r]);
        if (ne[r] < 0) { for (int q = 0; q < r; q++) free(E[q]); return -2; }
    }
    long dstStride[3], potStride[3];
    dstStride[R - 1] = 1; potStride[R - 1] = 1;
    for (int r = R - 2; r >= 0; r--) {
        dstStride[r] = dstStride[r + 1] * shape[r + 1];
        potStride[r] = potStride[r + 1] * potShape[r + 1];
    }
    orc_entry dummy = { 0, 0, 0, 0 };
    /* rank-0 entries grouped by destination row in their original order:  */
    /* each thread owns whole rows, so every destination element sums its   */
    /* contributions in the serial order (result independent of threads)   */
    int nrow0 = shape[0];
    int *rp = calloc((size_t)nrow0 + 1, sizeof(int)), *idx = malloc(sizeof(int) * (ne[0] > 0 ? ne[0] : 1));
    for (int a = 0; a < ne[0]; a++) rp[E[0][a].row + 1]++;
    for (int r0 = 0; r0 < nrow0; r0++) rp[r0 + 1] += rp[r0];
    {
        int *fill = malloc(sizeof(int) * (nrow0 > 0 ? nrow0 : 1));
        for (int r0 = 0; r0 < nrow0; r0++) fill[r0] = rp[r0];
        for (int a = 0; a < ne[0]; a++) idx[fill[E[0][a].row]++] = a;
        free(fill);
    }
    #pragma omp parallel for schedule(dynamic)
    for (int r0 = 0; r0 < nrow0; r0++)
    for (int q = rp[r0]; q < rp[r0 + 1]; q++) {
        const int a = idx[q];
        orc_entry e0 = E[0][a];
        for (int b = 0; b < (R > 1 ? ne[1] : 1); b++) {
            orc_entry e1 = R > 1 ? E[1][b] : dummy;
            for (int c = 0; c < (R > 2 ? ne[2] : 1); c++) {
                orc_entry e2 = R > 2 ? E[2][c] : dummy;
                long di = e0.row * dstStride[0] + (R > 1 ? e1.row * dstStride[1] : 0) + (R > 2 ? e2.row * dstStride[2] : 0);
                long si = e0.col * dstStride[0] + (R > 1 ? e1.col * dstStride[1] : 0) + (R > 2 ? e2.col * dstStride[2] : 0);
                long pi = e0.pidx * potStride[0] + (R > 1 ? e1.pidx * potStride[1] : 0) + (R > 2 ? e2.pidx * potStride[2] : 0);
                const orc_entry *last = R == 1 ? &e0 : (R == 2 ? &e1 : &e2);
                int cj = (e0.flags ^ e1.flags ^ e2.flags) & ENT_CONJ;
                if (last->flags & ENT_OWN) cj = last->flags & ENT_CONJ;
                zc v = cj ? conj(V[pi]) : V[pi];
                if (last->flags & ENT_REAL) v = creal(v);
                dst[di] += v * scaling * src[si];
            }
        }
    }
    free(rp);
    free(idx);
    for (int r = 0; r < R; r++) free(E[r]);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Hamiltonian: list of tensor potentials + overlap factors per rank.  */
/* ------------------------------------------------------------------ */
#define ORC_MAXTERMS 32
#define ORC_FACTORED (-1)        /* storage id marking a factored two-electron term */

/* Factored two-electron term on a (C, N, N) state (no reference layout):  */
/* sum_p coef_p |a_p><b_p| (x) A (x) B (kind 0, A, B real row-major bands,  */
/* element (i, j) at i*bw + s + j - i) or (x) R (kind 1, R[i1][d1][i2][d2]). */
/* Rows are kept in CSR order of a_p for the channel-parallel apply.        */
typedef struct {
    int kind, P, N, s;
    int *rowptr, *col;           /* C+1, P */
    zc *coef;                    /* P */
    double *A, *B, *R;
} orc_fterm;

typedef struct {
    int R;
    int shape[3];
    long n;
    int kd[3];
    zc *U[3];                    /* Cholesky factor per rank, NULL if orthogonal */
    int nterms;
    int storage[ORC_MAXTERMS][3];
    int potShape[ORC_MAXTERMS][3];
    zc *pot[ORC_MAXTERMS];
    int *pairs[ORC_MAXTERMS][3];
    double scaling[ORC_MAXTERMS];
    orc_fterm *ft[ORC_MAXTERMS];
    zc *tmp;
    long matvecs;
} orc_ham;

ORC_API orc_ham *orc_ham_create(int R, const int *shape)
{
    orc_ham *h = calloc(1, sizeof(*h));
    h->R = R; h->n = 1;
    for (int r = 0; r < R; r++) { h->shape[r] = shape[r]; h->n *= shape[r]; }
    h->tmp = malloc(sizeof(zc) * h->n);
    return h;
}

ORC_API void orc_ham_set_overlap(orc_ham *h, int rank, int kd, const double *Uband)
{
    int N = h->shape[rank];
    free(h->U[rank]);
    h->kd[rank] = kd;
    h->U[rank] = malloc(sizeof(zc) * N * (kd + 1));
    memcpy(h->U[rank], Uband, sizeof(zc) * N * (kd + 1));
}

ORC_API int orc_ham_add_term(orc_ham *h, const int *storage, const int *potShape, const double *pot,
                             const int *const *pairs)
{
    if (h->nterms >= ORC_MAXTERMS) return -1;
    int t = h->nterms++;
    long sz = 1;
    for (int r = 0; r < h->R; r++) {
        h->storage[t][r] = storage[r];
        h->potShape[t][r] = potShape[r];
        sz *= potShape[r];
        h->pairs[t][r] = NULL;
        if (pairs && pairs[r]) {
            h->pairs[t][r] = malloc(sizeof(int) * 2 * potShape[r]);
            memcpy(h->pairs[t][r], pairs[r], sizeof(int) * 2 * potShape[r]);
        }
    }
    h->pot[t] = malloc(sizeof(zc) * sz);
    memcpy(h->pot[t], pot, sizeof(zc) * sz);
    h->scaling[t] = 1.0;
    return t;
}

/* Add a factored term (rank-3 states only); returns its term index, whose */
/* scaling is set by orc_ham_set_scaling like any other term.              */
ORC_API int orc_ham_add_factored(orc_ham *h, int kind, int P, const int *pairs, const double *coef,
                                 const double *A, const double *B, const double *R, int s)
{
    if (h->nterms >= ORC_MAXTERMS || h->R != 3 || h->shape[1] != h->shape[2] || h->shape[0] <= 0) return -1;
    const int C = h->shape[0], N = h->shape[1], bw = 2 * s + 1;
    for (int p = 0; p < 2 * P; p++)
        if (pairs[p] < 0 || pairs[p] >= C) return -1;
    int t = h->nterms++;
    orc_fterm *f = calloc(1, sizeof(*f));
    f->kind = kind; f->P = P; f->N = N; f->s = s;
    f->rowptr = calloc(C + 1, sizeof(int));
    f->col = malloc(sizeof(int) * (P > 0 ? P : 1));
    f->coef = malloc(sizeof(zc) * (P > 0 ? P : 1));
    for (int p = 0; p < P; p++) f->rowptr[pairs[2 * p] + 1]++;
    for (int c = 0; c < C; c++) f->rowptr[c + 1] += f->rowptr[c];
    int *fill = calloc(C, sizeof(int));
    for (int p = 0; p < P; p++) {
        const int a = pairs[2 * p], q = f->rowptr[a] + fill[a]++;
        f->col[q] = pairs[2 * p + 1];
        f->coef[q] = coef[2 * p] + I * coef[2 * p + 1];
    }
    free(fill);
    if (kind == 0) {
        f->A = malloc(sizeof(double) * N * bw); memcpy(f->A, A, sizeof(double) * N * bw);
        f->B = malloc(sizeof(double) * N * bw); memcpy(f->B, B, sizeof(double) * N * bw);
    } else {
        const size_t cnt = (size_t)N * bw * N * bw;
        f->R = malloc(sizeof(double) * cnt); memcpy(f->R, R, sizeof(double) * cnt);
    }
    h->storage[t][0] = h->storage[t][1] = h->storage[t][2] = ORC_FACTORED;
    h->pot[t] = NULL;
    for (int r = 0; r < 3; r++) h->pairs[t][r] = NULL;
    h->ft[t] = f;
    h->scaling[t] = 1.0;
    return t;
}

/* dst += sc * term(src) for a factored term, parallel over output channels */
/* (each thread owns whole output channels, so the result does not depend   */
/* on the thread count).                                                    */
static void fterm_apply(const orc_fterm *f, int C, double sc, const zc *src, zc *dst)
{
    const int N = f->N, s = f->s, bw = 2 * s + 1;
    const long NN = (long)N * N;
    if (f->kind == 0) {
        #pragma omp parallel
        {
            zc *W = malloc(sizeof(zc) * NN);
            #pragma omp for schedule(dynamic)
            for (int a = 0; a < C; a++) {
                for (int q = f->rowptr[a]; q < f->rowptr[a + 1]; q++) {
                    const zc *y = src + (long)f->col[q] * NN;
                    for (int j1 = 0; j1 < N; j1++)          /* W = y B^T */
                        for (int i2 = 0; i2 < N; i2++) {
                            zc acc = 0.0;
                            for (int d = 0; d < bw; d++) {
                                const int j2 = i2 + d - s;
                                if (j2 >= 0 && j2 < N) acc += f->B[(long)i2 * bw + d] * y[(long)j1 * N + j2];
                            }
                            W[(long)j1 * N + i2] = acc;
                        }
                    const zc cf = sc * f->coef[q];
                    for (int i1 = 0; i1 < N; i1++)          /* dst += cf A W */
                        for (int i2 = 0; i2 < N; i2++) {
                            zc acc = 0.0;
                            for (int d = 0; d < bw; d++) {
                                const int j1 = i1 + d - s;
                                if (j1 >= 0 && j1 < N) acc += f->A[(long)i1 * bw + d] * W[(long)j1 * N + i2];
                            }
                            dst[(long)a * NN + (long)i1 * N + i2] += cf * acc;
                        }
                }
            }
            free(W);
        }
        return;
    }
    /* multipole: U_b = R (.) y_b once per source channel, then the mixing */
    char *need = calloc(C, 1);
    for (int q = 0; q < f->rowptr[C]; q++) need[f->col[q]] = 1;
    zc *Ub = malloc(sizeof(zc) * NN * C);
    #pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < C; b++) {
        if (!need[b]) continue;
        const zc *y = src + (long)b * NN;
        zc *u = Ub + (long)b * NN;
        for (int i1 = 0; i1 < N; i1++)
            for (int i2 = 0; i2 < N; i2++) {
                zc acc = 0.0;
                for (int d1 = 0; d1 < bw; d1++) {
                    const int j1 = i1 + d1 - s;
                    if (j1 < 0 || j1 >= N) continue;
                    const double *Rr = f->R + (((long)i1 * bw + d1) * N + i2) * bw;
                    for (int d2 = 0; d2 < bw; d2++) {
                        const int j2 = i2 + d2 - s;
                        if (j2 >= 0 && j2 < N) acc += Rr[d2] * y[(long)j1 * N + j2];
                    }
                }
                u[(long)i1 * N + i2] = acc;
            }
    }
    #pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < C; a++)
        for (int q = f->rowptr[a]; q < f->rowptr[a + 1]; q++) {
            const zc cf = sc * f->coef[q];
            const zc *u = Ub + (long)f->col[q] * NN;
            zc *d = dst + (long)a * NN;
            for (long e = 0; e < NN; e++) d[e] += cf * u[e];
        }
    free(Ub);
    free(need);
}

ORC_API void orc_ham_set_scaling(orc_ham *h, const double *s)
{
    for (int t = 0; t < h->nterms; t++) h->scaling[t] = s[t];
}

ORC_API void orc_ham_destroy(orc_ham *h)
{
    if (!h) return;
    for (int t = 0; t < h->nterms; t++) {
        free(h->pot[t]);
        for (int r = 0; r < h->R; r++) free(h->pairs[t][r]);
        if (h->ft[t]) {
            orc_fterm *f = h->ft[t];
            free(f->rowptr); free(f->col); free(f->coef); free(f->A); free(f->B); free(f->R); free(f);
        }
    }
    for (int r = 0; r < h->R; r++) free(h->U[r]);
    free(h->tmp);
    free(h);
}

/* threads of the OpenMP regions (timing of the all-core CPU baseline) */
ORC_API void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
ORC_API int orc_get_threads(void) { return omp_get_max_threads(); }

/* sum of tensor potentials (Propagator.py:61-66 after dest = 0) */
static void ham_potentials(orc_ham *h, const zc *src, zc *dst)
{
    memset(dst, 0, sizeof(zc) * h->n);
    for (int t = 0; t < h->nterms; t++) {
        if (h->ft[t]) { fterm_apply(h->ft[t], h->shape[0], h->scaling[t], src, dst); continue; }
        orc_tp_multiply(h->R, h->storage[t], h->potShape[t], (const double *)h->pot[t],
                        h->scaling[t], h->shape, (const double *)src, (double *)dst,
                        (const int *const *)h->pairs[t]);
    }
}

/* MultiplyHamiltonianBalancedOverlap (BasisPropagator.py:131-159):    */
/* dst = U^-H (sum_t s_t V_t) U^-1 src; src left unchanged.            */
ORC_API void orc_ham_balanced(orc_ham *h, const double *srcd, double *dstd)
{
    zc *x = h->tmp;
    memcpy(x, srcd, sizeof(zc) * h->n);
    for (int r = 0; r < h->R; r++)
        if (h->U[r]) orc_band_apply(OP_SOLVE_N, r, h->R, h->shape, h->kd[r], (const double *)h->U[r], (double *)x);
    ham_potentials(h, x, (zc *)dstd);
    for (int r = 0; r < h->R; r++)
        if (h->U[r]) orc_band_apply(OP_SOLVE_C, r, h->R, h->shape, h->kd[r], (const double *)h->U[r], dstd);
    h->matvecs++;
}

/* plain H src (no overlap) */
ORC_API void orc_ham_potentials(orc_ham *h, const double *srcd, double *dstd)
{
    ham_potentials(h, (const zc *)srcd, (zc *)dstd);
}

/* S-weighted inner product <a|S|b> (combinedrepresentation.cpp:423-474) */
ORC_API void orc_ham_inner(orc_ham *h, const double *ad, const double *bd, double *out)
{
    zc *sb = malloc(sizeof(zc) * h->n);
    memcpy(sb, bd, sizeof(zc) * h->n);
    zc *t = malloc(sizeof(zc) * h->n);
    for (int r = 0; r < h->R; r++) {
        if (!h->U[r]) continue;
        /* S = U^H U: multiply by U then by U^H via band storage */
        long pre = 1, post = 1;
        for (int i = 0; i < r; i++) pre *= h->shape[i];
        for (int i = r + 1; i < h->R; i++) post *= h->shape[i];
        int N = h->shape[r], kd = h->kd[r];
        /* rebuild S band from U: S(i,j) = sum_k conj(U(k,i)) U(k,j) */
        zc *S = calloc((size_t)N * (kd + 1), sizeof(zc));
        for (int j = 0; j < N; j++)
            for (int i = (j - kd < 0 ? 0 : j - kd); i <= j; i++) {
                zc s = 0.0;
                for (int k = (j - kd < 0 ? 0 : j - kd); k <= i; k++)
                    s += conj(UB(h->U[r], kd, k, i)) * UB(h->U[r], kd, k, j);
                UB(S, kd, i, j) = s;
            }
        memcpy(t, sb, sizeof(zc) * h->n);
        for (long a = 0; a < pre; a++)
            for (long c = 0; c < post; c++)
                zhbmv_u(N, kd, S, t + a * N * post + c, sb + a * N * post + c, post);
        free(S);
    }
    const zc *a = (const zc *)ad;
    zc s = 0.0;
    for (long i = 0; i < h->n; i++) s += conj(a[i]) * sb[i];
    out[0] = creal(s); out[1] = cimag(s);
    free(sb); free(t);
}

/* ------------------------------------------------------------------ */
/* small dense complex helpers (column-major m x m)                    */
/* ------------------------------------------------------------------ */
static void zmm(int m, const zc *A, const zc *B, zc *C)
{
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) {
            zc s = 0.0;
            for (int k = 0; k < m; k++) s += A[i + k * m] * B[k + j * m];
            C[i + j * m] = s;
        }
}

/* zgetrf (partial pivoting) + zgetrs on nrhs columns; returns info */
static int zlu_solve(int m, zc *A, zc *B, int nrhs, int *piv)
{
    for (int k = 0; k < m; k++) {
        int p = k;
        double best = cabs(A[k + k * m]);
        /* izamax uses |re|+|im| */
        best = fabs(creal(A[k + k * m])) + fabs(cimag(A[k + k * m]));
        for (int i = k + 1; i < m; i++) {
            double v = fabs(creal(A[i + k * m])) + fabs(cimag(A[i + k * m]));
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (A[p + k * m] == 0.0) return k + 1;
        if (p != k)
            for (int j = 0; j < m; j++) { zc t = A[k + j * m]; A[k + j * m] = A[p + j * m]; A[p + j * m] = t; }
        zc inv = 1.0 / A[k + k * m];
        for (int i = k + 1; i < m; i++) A[i + k * m] *= inv;
        for (int j = k + 1; j < m; j++) {
            zc a = A[k + j * m];
            for (int i = k + 1; i < m; i++) A[i + j * m] -= A[i + k * m] * a;
        }
    }
    for (int c = 0; c < nrhs; c++) {
        zc *b = B + c * m;
        for (int k = 0; k < m; k++) if (piv[k] != k) { zc t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
        for (int k = 0; k < m; k++) for (int i = k + 1; i < m; i++) b[i] -= A[i + k * m] * b[k];
        for (int k = m - 1; k >= 0; k--) {
            b[k] /= A[k + k * m];
            for (int i = 0; i < k; i++) b[i] -= A[i + k * m] * b[k];
        }
    }
    return 0;
}

/* pAMP::PadeExponential (pamp.h:470-564): irreducible (p,p) Pade,     */
/* out = +-(I + 2 (Q-P)^-1 P) with P odd / Q even Horner polynomials.  */
ORC_API int orc_pade_exp(int m, const double *Ad, double *outd, int order)
{
    const zc *A = (const zc *)Ad;
    zc *out = (zc *)outd;
    size_t sz = sizeof(zc) * m * m;
    zc *A2 = malloc(sz), *P = calloc(m * m, sizeof(zc)), *Q = calloc(m * m, sizeof(zc)), *T = malloc(sz);
    int *piv = malloc(sizeof(int) * m);
    double c[64];
    zmm(m, A, A, A2);
    c[0] = 1.0;
    for (int i = 0; i < order; i++) c[i + 1] = c[i] * (double)(order - i) / (double)((i + 1) * (2 * order - i));
    for (int i = 0; i < m; i++) { Q[i + i * m] = c[order]; P[i + i * m] = c[order - 1]; }
    int odd = 1;
    for (int i = order - 2; i >= 0; i--) {
        zc *X = odd ? Q : P;
        zmm(m, X, A2, T);
        memcpy(X, T, sz);
        for (int r = 0; r < m; r++) X[r + r * m] += c[i];
        odd = !odd;
    }
    if (odd) { zmm(m, Q, A, T); memcpy(Q, T, sz); }
    else { zmm(m, P, A, T); memcpy(P, T, sz); }
    for (int i = 0; i < m * m; i++) Q[i] -= P[i];
    int info = zlu_solve(m, Q, P, m, piv);
    if (info == 0) {
        for (int i = 0; i < m * m; i++) out[i] = 2.0 * P[i];
        for (int i = 0; i < m; i++) out[i + i * m] += 1.0;
        if (odd) for (int i = 0; i < m * m; i++) out[i] = -out[i];
    }
    free(A2); free(P); free(Q); free(T); free(piv);
    return info;
}

/* pAMP::ScalingPadeExponential (pamp.h:430-467) with scalingOrder = -1:  */
/* s = max(0, ceil(log2(max|A_ij|))), A/2^s, Pade(6), square s times.     */
ORC_API int orc_scaling_pade_exp(int m, const double *Ad, double *outd)
{
    zc *A = malloc(sizeof(zc) * m * m), *T = malloc(sizeof(zc) * m * m);
    memcpy(A, Ad, sizeof(zc) * m * m);
    double mx = 0.0;
    for (int i = 0; i < m * m; i++) if (cabs(A[i]) > mx) mx = cabs(A[i]);
    int s = mx > 0.0 ? (int)ceil(log2(mx)) : 0;
    if (s < 0) s = 0;
    double sc = pow(2.0, s);
    for (int i = 0; i < m * m; i++) A[i] /= sc;
    zc *E = (zc *)outd;
    int info = orc_pade_exp(m, (const double *)A, outd, 6);
    for (int i = 0; i < s && info == 0; i++) { zmm(m, E, E, T); memcpy(E, T, sizeof(zc) * m * m); }
    free(A); free(T);
    return info;
}

/* ------------------------------------------------------------------ */
/* pAMP::PropagateVector (pamp.h:613-669), Euclidean inner products,  */
/* CGS (pamp.h:293-350), optional re-orthogonalisation (:353-400).    */
/* ------------------------------------------------------------------ */
/* exponential of the m x m Krylov matrix A = -i dt H (column-major): out = */
/* exp(A) by pAMP's scaling + Pade, or by a caller-supplied routine (the    */
/* eigenvector exponential, pamp.h:567-607, needs LAPACK's zgeev, which     */
/* this file does not restate: oracle.py passes numpy's).                   */
typedef int (*orc_expfn)(int m, const double *A, double *out);

typedef struct {
    int m;
    int reorth;
    long n;
    zc *V;     /* m x n */
    zc *r, *tmp;
    zc *H, *Hexp;
    double err_estimate;
    int happy;
    int basis_used;
    orc_expfn expfn;   /* NULL: orc_scaling_pade_exp */
} orc_pamp;

static zc zdotc(long n, const zc *a, const zc *b)
{
    zc s = 0.0;
    for (long i = 0; i < n; i++) s += conj(a[i]) * b[i];
    return s;
}

/* |a| as 64 fixed blocks summed in order: the same result on any thread count */
static double vnorm(long n, const zc *a)
{
    enum { NB = 64 };
    double part[NB];
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < NB; b++) {
        const long i0 = n * b / NB, i1 = n * (b + 1) / NB;
        double s = 0.0;
        for (long i = i0; i < i1; i++) s += creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
        part[b] = s;
    }
    double s = 0.0;
    for (int b = 0; b < NB; b++) s += part[b];
    return sqrt(s);
}

ORC_API orc_pamp *orc_pamp_create(long n, int m, int reorth)
{
    orc_pamp *p = calloc(1, sizeof(*p));
    p->m = m; p->n = n; p->reorth = reorth;
    p->V = malloc(sizeof(zc) * n * m);
    p->r = malloc(sizeof(zc) * n);
    p->tmp = malloc(sizeof(zc) * n);
    p->H = malloc(sizeof(zc) * m * m);
    p->Hexp = malloc(sizeof(zc) * m * m);
    return p;
}

ORC_API void orc_pamp_destroy(orc_pamp *p)
{
    if (!p) return;
    free(p->V); free(p->r); free(p->tmp); free(p->H); free(p->Hexp); free(p);
}

ORC_API double orc_pamp_error_estimate(const orc_pamp *p) { return p->err_estimate; }
ORC_API int orc_pamp_happy(const orc_pamp *p) { return p->happy; }

/* H is kept column-major: H[row + col*m] */
ORC_API int orc_pamp_step(orc_pamp *p, orc_ham *h, double *psid, double dt_re, double dt_im)
{
    long n = p->n;
    int m = p->m;
    zc *psi = (zc *)psid, *V = p->V, *r = p->r, *H = p->H;
    zc dt = dt_re + I * dt_im;
    double tol = sqrt(DBL_EPSILON);
    memcpy(r, psi, sizeof(zc) * n);
    double norm = vnorm(n, r);
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) r[i] *= 1.0 / norm;
    memset(V, 0, sizeof(zc) * n * m);
    memset(H, 0, sizeof(zc) * m * m);
    p->happy = 0;
    /* PerformInitialArnoldiStep (:271-290) */
    memcpy(V, r, sizeof(zc) * n);
    orc_ham_balanced(h, (double *)V, (double *)r);
    zc alpha = zdotc(n, V, r);
    for (long i = 0; i < n; i++) r[i] -= alpha * V[i];
    H[0] = alpha;
    int j = 0;
    zc ov[256], ov3[256];
    while (j < m - 1 && !p->happy) {
        /* PerformArnoldiStep (:293-350) */
        double beta = vnorm(n, r);
        if (fabs(beta) < tol) { p->happy = 1; break; }
        zc *vj1 = V + (size_t)(j + 1) * n;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) { r[i] *= 1.0 / beta; vj1[i] = r[i]; }
        H[(j + 1) + j * m] = beta;
        orc_ham_balanced(h, (double *)vj1, (double *)r);
        double origNorm = vnorm(n, r);
        /* one whole dot per thread and independent elements: the result */
        /* does not depend on the thread count                           */
        #pragma omp parallel for schedule(dynamic)
        for (int k = 0; k <= j + 1; k++) ov[k] = zdotc(n, V + (size_t)k * n, r);
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            zc t = 0.0;
            for (int k = 0; k <= j + 1; k++) t += V[(size_t)k * n + i] * ov[k];
            r[i] -= t;
        }
        double resNorm = vnorm(n, r);
        if (p->reorth) {
            int cnt = 0;
            while (fabs(resNorm) < 0.717 * fabs(origNorm)) {
                cnt++;
                if (cnt > 3) break;
                for (int k = 0; k <= j + 1; k++) ov3[k] = zdotc(n, V + (size_t)k * n, r);
                for (long i = 0; i < n; i++) {
                    zc t = 0.0;
                    for (int k = 0; k <= j + 1; k++) t += V[(size_t)k * n + i] * ov3[k];
                    r[i] -= t;
                }
                for (int k = 0; k <= j + 1; k++) ov[k] += ov3[k];
                origNorm = resNorm;
                resNorm = vnorm(n, r);
            }
        }
        for (int k = 0; k <= j + 1; k++) H[k + (j + 1) * m] = ov[k];
        j++;
    }
    p->basis_used = j + 1;
    int mm = m;   /* pAMP exponentiates the full (zero-padded) m x m matrix */
    for (int i = 0; i < mm * mm; i++) H[i] *= -dt * I;
    int info = p->expfn ? p->expfn(mm, (const double *)H, (double *)p->Hexp)
                        : orc_scaling_pade_exp(mm, (const double *)H, (double *)p->Hexp);
    if (info) return info;
    /* psi = norm * V * exp(H) e1 (:657-662) */
    zc *v = p->Hexp;   /* first column */
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        zc t = 0.0;
        for (int k = 0; k < m; k++) t += V[(size_t)k * n + i] * v[k];
        psi[i] = t * norm;
    }
    p->err_estimate = cabs(v[m - 1]);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Problem loop with PamPropagator, balanced overlap                  */
/* (Problem.py:158-173, PamPropagator.py:50-59):                       */
/* psi~ = U psi ; pAMP step on U^-H H(t_n) U^-1 ; psi = U^-1 psi~.     */
/* scalings[step*nterms + t] = time factor of term t at t_n.           */
/* ------------------------------------------------------------------ */
ORC_API int orc_propagate_pamp_ex(orc_ham *h, double *psid, int nsteps, double dt,
                                  const double *scalings, int m, int reorth, orc_expfn expfn)
{
    orc_pamp *p = orc_pamp_create(h->n, m, reorth);
    p->expfn = expfn;
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; s++) {
        orc_ham_set_scaling(h, scalings + (size_t)s * h->nterms);
        for (int r = 0; r < h->R; r++)
            if (h->U[r]) orc_band_apply(OP_MUL_N, r, h->R, h->shape, h->kd[r], (const double *)h->U[r], psid);
        rc = orc_pamp_step(p, h, psid, dt, 0.0);
        for (int r = 0; r < h->R; r++)
            if (h->U[r]) orc_band_apply(OP_SOLVE_N, r, h->R, h->shape, h->kd[r], (const double *)h->U[r], psid);
    }
    orc_pamp_destroy(p);
    return rc;
}

ORC_API int orc_propagate_pamp(orc_ham *h, double *psid, int nsteps, double dt,
                               const double *scalings, int m, int reorth)
{
    return orc_propagate_pamp_ex(h, psid, nsteps, dt, scalings, m, reorth, NULL);
}

/* ------------------------------------------------------------------ */
/* Expokit (Sidje, ACM TOMS 24(1):130-156, 1998) as translated by f2c  */
/* in core/krylov/expokit/f2c/zgpadm.cpp:28-285 and zgexpv.cpp:32-586. */
/* ------------------------------------------------------------------ */

/* zgpadm: exp(t H) for an m x m complex matrix (column-major, ldh = m), */
/* irreducible Pade(ideg) with ns = max(0, int(log2 ||tH||_inf) + 2).    */
ORC_API int orc_zgpadm(int ideg, int m, double t, const double *Hd, double *outd, int *ns_out)
{
    const zc *H = (const zc *)Hd;
    zc *out = (zc *)outd;
    size_t sz = sizeof(zc) * m * m;
    double hnorm = 0.0;
    for (int i = 0; i < m; i++) {
        double rs = 0.0;
        for (int j = 0; j < m; j++) rs += cabs(H[i + j * m]);
        if (rs > hnorm) hnorm = rs;
    }
    hnorm = fabs(t * hnorm);
    if (hnorm == 0.0) return -1;                     /* "null H" stop */
    int ns = (int)(log(hnorm) / log(2.0)) + 2;
    if (ns < 0) ns = 0;
    double scale = t / pow(2.0, ns);
    double scale2 = scale * scale;
    double coef[32];
    coef[0] = 1.0;
    for (int k = 1; k <= ideg; k++)
        coef[k] = coef[k - 1] * (double)(ideg + 1 - k) / (double)(k * (2 * ideg + 1 - k));
    zc *H2 = malloc(sz), *P = calloc(m * m, sizeof(zc)), *Q = calloc(m * m, sizeof(zc)), *F = malloc(sz);
    int *piv = malloc(sizeof(int) * m);
    zmm(m, H, H, H2);
    for (int e = 0; e < m * m; e++) H2[e] *= scale2;
    for (int j = 0; j < m; j++) { P[j + j * m] = coef[ideg - 1]; Q[j + j * m] = coef[ideg]; }
    int iodd = 1;
    for (int k = ideg - 1; k > 0; k--) {
        zc *U = iodd ? Q : P;
        zmm(m, U, H2, F);
        for (int j = 0; j < m; j++) F[j + j * m] += coef[k - 1];
        memcpy(U, F, sz);
        iodd = 1 - iodd;
    }
    if (iodd) { zmm(m, Q, H, F); for (int e = 0; e < m * m; e++) Q[e] = scale * F[e]; }
    else { zmm(m, P, H, F); for (int e = 0; e < m * m; e++) P[e] = scale * F[e]; }
    for (int e = 0; e < m * m; e++) Q[e] -= P[e];
    int info = zlu_solve(m, Q, P, m, piv);
    if (info) { free(H2); free(P); free(Q); free(F); free(piv); return info; }
    for (int e = 0; e < m * m; e++) P[e] *= 2.0;
    for (int j = 0; j < m; j++) P[j + j * m] += 1.0;
    if (ns == 0 && iodd) for (int e = 0; e < m * m; e++) P[e] = -P[e];
    else for (int k = 0; k < ns; k++) { zmm(m, P, P, F); memcpy(P, F, sz); }
    memcpy(out, P, sz);
    if (ns_out) *ns_out = ns;
    free(H2); free(P); free(Q); free(F); free(piv);
    return 0;
}

/* f2c helpers: i_dnnt rounds half away from zero, d_int truncates. */
static double x_roundstep(double t)
{
    double sqr1 = sqrt(0.1);
    double e = log10(t) - sqr1;
    int ie = (int)(e >= 0 ? floor(e + 0.5) : -floor(0.5 - e)) - 1;
    double p1 = pow(10.0, ie);
    return trunc(t / p1 + 0.55) * p1;
}

typedef struct { orc_ham *h; int imag_time; zc *tmp; } orc_exp_ctx;

/* MultiplyHamiltonian (krylovcommon.h:15-78): out = -i H^ in (real time), -H^ in (imaginary) */
static void exp_matvec(orc_exp_ctx *c, const zc *in, zc *out)
{
    orc_ham_balanced(c->h, (const double *)in, (double *)out);
    long n = c->h->n;
    if (!c->imag_time) for (long i = 0; i < n; i++) out[i] *= -I;
    else for (long i = 0; i < n; i++) out[i] = -out[i];
}

static double dznrm2(long n, const zc *x) { return sqrt(creal(zdotc(n, x, x))); }

/* zgexpv: w = exp(t A) v with A the balanced operator times -i (or -1).   */
/* Returns iflag (0 ok, 1 max 500 steps); stats[0..4] = nmult, nexph,     */
/* nstep, nreject, mbrkdwn.                                                */
ORC_API int orc_zgexpv(orc_ham *h, int m, double t, const double *vd, double *wd, double tol, double anorm,
                       int imag_time, int *stats)
{
    long n = h->n;
    const zc *v = (const zc *)vd;
    zc *w = (zc *)wd;
    int mh = m + 2;
    zc *V = calloc((size_t)n * (m + 1) + n, sizeof(zc));
    zc *Hm = calloc((size_t)mh * mh, sizeof(zc));
    zc *E = malloc(sizeof(zc) * mh * mh);
    orc_exp_ctx ctx = { h, imag_time, NULL };
    int k1 = 2, mbrkdwn = m, nmult = 0, nreject = 0, nexph = 0, nstep = 0, iflag = 0;
    double t_out = fabs(t), t_now = 0.0, t_new = 0.0, avnorm = 0.0;
    double eps = DBL_EPSILON;   /* the 4/3 loop of zgexpv.cpp:277-283 gives 2^-52 */
    if (tol <= eps) tol = sqrt(eps);
    double rndoff = eps * anorm, break_tol = 1e-7;
    double sgn = t >= 0 ? 1.0 : -1.0;
    memcpy(w, v, sizeof(zc) * n);
    double beta = dznrm2(n, w);
    double sqr1 = sqrt(0.1);
    double xm = 1.0 / m;
    double p2 = tol * pow((m + 1) / 2.72, m + 1) * sqrt((m + 1) * 6.28);
    t_new = (1.0 / anorm) * pow(p2 / (4.0 * beta * anorm), xm);
    (void)sqr1;
    t_new = x_roundstep(t_new);
    double err_loc = 0.0, t_step = 0.0;
    while (t_now < t_out) {
        nstep++;
        t_step = fmin(t_out - t_now, t_new);
        double p1 = 1.0 / beta;
        for (long i = 0; i < n; i++) V[i] = p1 * w[i];
        memset(Hm, 0, sizeof(zc) * mh * mh);
        k1 = 2; mbrkdwn = m;
        int j;
        for (j = 1; j <= m; j++) {
            nmult++;
            zc *vj = V + (size_t)(j - 1) * n, *vj1 = V + (size_t)j * n;
            exp_matvec(&ctx, vj, vj1);
            for (int i = 1; i <= j; i++) {
                zc hij = zdotc(n, V + (size_t)(i - 1) * n, vj1);
                for (long q = 0; q < n; q++) vj1[q] -= hij * V[(size_t)(i - 1) * n + q];
                Hm[(i - 1) + (j - 1) * mh] = hij;
            }
            double hj1j = dznrm2(n, vj1);
            if (hj1j <= break_tol) {
                k1 = 0; mbrkdwn = j; t_step = t_out - t_now;
                break;
            }
            Hm[j + (j - 1) * mh] = hj1j;
            for (long q = 0; q < n; q++) vj1[q] *= 1.0 / hj1j;
        }
        if (k1 != 0) {
            nmult++;
            exp_matvec(&ctx, V + (size_t)m * n, V + (size_t)(m + 1) * n);
            avnorm = dznrm2(n, V + (size_t)(m + 1) * n);
        }
        Hm[(m + 1) + m * mh] = 1.0;
        int ireject = 0;
        double xmloc = 1.0 / m;
        for (;;) {
            nexph++;
            int mx = mbrkdwn + k1;
            /* zgpadm on the leading mx x mx block of the (mh x mh) H */
            zc *Hs = malloc(sizeof(zc) * mx * mx);
            for (int c = 0; c < mx; c++) for (int r = 0; r < mx; r++) Hs[r + c * mx] = Hm[r + c * mh];
            int ns;
            int info = orc_zgpadm(6, mx, sgn * t_step, (const double *)Hs, (double *)Hs, &ns);
            for (int c = 0; c < mx; c++) for (int r = 0; r < mx; r++) E[r + c * mh] = Hs[r + c * mx];
            free(Hs);
            if (info) { iflag = -9; goto done; }
            if (k1 == 0) err_loc = tol;
            else {
                double q1 = cabs(E[m]) * beta;
                double q2 = cabs(E[m + 1]) * beta * avnorm;
                if (q1 > 10.0 * q2) { err_loc = q2; xmloc = 1.0 / m; }
                else if (q1 > q2) { err_loc = q1 * q2 / (q1 - q2); xmloc = 1.0 / m; }
                else { err_loc = q1; xmloc = 1.0 / (m - 1); }
            }
            if (k1 != 0 && err_loc > 1.2 * t_step * tol) {
                t_step = 0.9 * t_step * pow(t_step * tol / err_loc, xmloc);
                t_step = x_roundstep(t_step);
                ireject++; nreject++;
                continue;
            }
            break;
        }
        int mx = mbrkdwn + (k1 - 1 > 0 ? k1 - 1 : 0);
        for (long i = 0; i < n; i++) {
            zc s = 0.0;
            for (int k = 0; k < mx; k++) s += V[(size_t)k * n + i] * E[k];
            w[i] = beta * s;
        }
        beta = dznrm2(n, w);
        t_new = 0.9 * t_step * pow(t_step * tol / err_loc, xmloc);
        t_new = x_roundstep(t_new);
        err_loc = fmax(err_loc, rndoff);
        t_now += t_step;
        if (nstep >= 500 && t_now < t_out) { iflag = 1; break; }
    }
done:
    if (stats) { stats[0] = nmult; stats[1] = nexph; stats[2] = nstep; stats[3] = nreject; stats[4] = mbrkdwn; }
    free(V); free(Hm); free(E);
    return iflag;
}

/* ExpokitPropagator::AdvanceStep (expokitpropagator.cpp:36-103) inside the */
/* balanced-overlap wrapping of KrylovPropagator.py:41-45.  The time passed  */
/* to zgexpv is |dt| (the reference's line 61 computes dt^2 for real dt).    */
ORC_API int orc_propagate_expokit(orc_ham *h, double *psid, int nsteps, double dt, const double *scalings,
                                  int m, double tol, double anorm, int *stats)
{
    zc *w = malloc(sizeof(zc) * h->n);
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; s++) {
        orc_ham_set_scaling(h, scalings + (size_t)s * h->nterms);
        for (int r = 0; r < h->R; r++)
            if (h->U[r]) orc_band_apply(OP_MUL_N, r, h->R, h->shape, h->kd[r], (const double *)h->U[r], psid);
        rc = orc_zgexpv(h, m, fabs(dt), psid, (double *)w, tol, anorm, 0, stats);
        memcpy(psid, w, sizeof(zc) * h->n);
        for (int r = 0; r < h->R; r++)
            if (h->U[r]) orc_band_apply(OP_SOLVE_N, r, h->R, h->shape, h->kd[r], (const double *)h->U[r], psid);
    }
    free(w);
    return rc;
}

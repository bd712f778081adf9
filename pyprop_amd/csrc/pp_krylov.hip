// pAMP Krylov propagation on MI355X: one persistent cooperative launch runs
// whole time steps of
//
//   psi~ = U psi;  fixed-m Arnoldi on  H^ = U^-H (sum_t s_t(t_n) V_t) U^-1
//   with classical Gram-Schmidt;  exp(-i dt H_m) by scaling + Pade(6);
//   psi~ <- |psi~| V exp(-i dt H_m) e1;  psi = U^-1 psi~
//
// Reference: pyprop/propagator/PamPropagator.py:50-75 (balanced overlap
// wrapping), pyprop/tensorpotential/BasisPropagator.py:131-159 (the matvec),
// core/krylov/pamp/pamp/pamp.h:271-350 (Arnoldi/CGS), :353-400 (optional
// re-orthogonalisation, 0.717 criterion, <= 3 passes), :430-564 (scaling and
// Pade), :613-669 (PropagateVector); Problem.py:158-173 (H held at t_n).
//
// Work split: the wavefunction is (L, N) with the angular rank outer and the
// B-spline rank contiguous; workgroup a owns fiber a for the whole run.  The
// Arnoldi vectors, the residual and the Gram-Schmidt projections of fiber a
// never leave workgroup a (residual in LDS, V fiber-major in HBM/L2).  Only
// U^-1 w (needed by the neighbours through the angular couplings) and the
// per-fiber partial inner products cross workgroups, so each Arnoldi
// iteration costs two grid barriers.  Every workgroup reduces the partial
// sums itself in one fixed order, so branch decisions (breakdown,
// re-orthogonalisation) are bitwise identical across the grid.
#include "pp_common.h"
#include <hip/hip_cooperative_groups.h>
#include <cstring>
#include <cmath>

namespace cg = cooperative_groups;

#define PP_MAXTERMS 8
#define PP_MAXM 64
#define PP_THREADS 256

struct pp_kterm {
  const int *rowptr;  // L+1
  const int *ent;     // (col, pidx, conj) per entry
  const zd *pot;      // (P0, N*bw) reference BandNH layout
};

struct pp_pamp_args {
  int L, N, m, s, bw, nterms, nsteps, reorth, padN;
  double dt_re, dt_im;
  pp_kterm term[PP_MAXTERMS];
  const double *scal;  // nsteps x nterms
  pp_overlap ov;
  zd *psi;             // L x N, in/out
  zd *V;               // L x m x N (fiber-major)
  zd *y;               // L x N, U^-1 w exchanged between workgroups
  double *part;        // L x pstr partial sums
  int pstr;
  zd *Hm;              // m x m column-major, owned by workgroup 0
  zd *cvec;            // m, exp(-i dt H) e1 broadcast
  double *diag;        // nsteps x 4: error estimate, basis used, |psi~|, re-orth passes
  long long *tlog;     // optional: wall-clock stamps of workgroup 0 during step 0
};

struct pp_pamp {
  int L, N, m, reorth, s, bw;
  const pp_overlap *ov;
  int nterms;
  pp_kterm term[PP_MAXTERMS];
  int *csr_dev[PP_MAXTERMS];
  zd *V, *y, *Hm, *cvec;
  double *part, *diag, *scal;
  int diag_cap, scal_cap, pstr;
  long long *tlog;
  int timing;
};

#define PP_TLOG_MAX 512

// ------------------------------------------------------------------ helpers
__device__ inline double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic reduction over workgroups: out[c] = sum_{a=0}^{L-1} part[a*pstr + off + c]
// for c < ncomp, summed in the same fixed order by every workgroup.
__device__ inline void grid_partial_sum(const double *part, int pstr, int off, int ncomp, int L,
                                        double *out_lds, double *scratch_lds)
{
  const int tid = threadIdx.x;
  int G = PP_THREADS / ncomp;
  if (G > L) G = L;
  if (G < 1) G = 1;
  if (tid < G * ncomp) {
    int c = tid % ncomp, g = tid / ncomp;
    double acc = 0.0;
    for (int a = g; a < L; a += G) acc += part[(size_t)a * pstr + off + c];
    scratch_lds[g * ncomp + c] = acc;
  }
  __syncthreads();
  if (tid < ncomp) {
    double acc = 0.0;
    for (int g = 0; g < G; g++) acc += scratch_lds[g * ncomp + tid];
    out_lds[tid] = acc;
  }
  __syncthreads();
}

// per-workgroup sum of cnt values, each thread contributing val(k) for k<cnt
// (wave shuffles, then LDS across the 4 waves); result in red_out[k].
template <typename F>
__device__ inline void block_sums(int cnt, F val, double *wave_lds, double *red_out)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int k = 0; k < cnt; k++) {
    double v = wave_sum(val(k));
    if (lane == 0) wave_lds[wave * cnt + k] = v;
  }
  __syncthreads();
  if (tid < cnt) {
    double acc = 0.0;
    for (int w = 0; w < PP_THREADS / 64; w++) acc += wave_lds[w * cnt + tid];
    red_out[tid] = acc;
  }
  __syncthreads();
}

// out[2k], out[2k+1] = <V_k, W> over one fiber for k < nk (V_k in global
// memory, W in LDS); with extra_norm also out[2nk] = |W|^2.  Wave w owns
// k = w, w+4, ...; each lane streams a strided slice of the fiber so the
// loads of one k are independent and pipeline.
__device__ inline void fiber_dots(const zd *V, const zd *W, int nk, int N, bool extra_norm, double *out)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ntot = nk + (extra_norm ? 1 : 0);
  for (int k = wave; k < ntot; k += PP_THREADS / 64) {
    double re = 0.0, im = 0.0;
    if (k < nk) {
      const zd *vk = V + (size_t)k * N;
#pragma unroll 4
      for (int i = lane; i < N; i += 64) {
        zd v = vk[i], w = W[i];
        re = fma(v.x, w.x, re); re = fma(v.y, w.y, re);
        im = fma(v.x, w.y, im); im = fma(-v.y, w.x, im);
      }
    } else {
#pragma unroll 4
      for (int i = lane; i < N; i += 64) { zd w = W[i]; re = fma(w.x, w.x, re); re = fma(w.y, w.y, re); }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) { out[2 * k] = re; if (k < nk) out[2 * k + 1] = im; }
  }
  __syncthreads();
}

// ------------------------------------------------------------------ Pade on one workgroup
// pAMP::ScalingPadeExponential + PadeExponential (pamp.h:430-564), column-major
// m x m complex matrices in LDS; result column 0 -> cvec.
__device__ void pade_block(int m, const zd *Hg, double dt_re, double dt_im, zd *lds, zd *cvec, double *diag)
{
  const int tid = threadIdx.x;
  const int mm = m * m;
  zd *A = lds, *A2 = lds + mm, *P = lds + 2 * mm, *Q = lds + 3 * mm, *T = lds + 4 * mm;
  double *red = (double *)(lds + 5 * mm);
  int *ipiv = (int *)(red + 8);
  const zd fac = zmk(dt_im, -dt_re);  // -i * dt
  for (int e = tid; e < mm; e += PP_THREADS) A[e] = zmul(fac, Hg[e]);
  __syncthreads();
  // max |A_ij|
  double mx = 0.0;
  for (int e = tid; e < mm; e += PP_THREADS) mx = fmax(mx, sqrt(zabs2(A[e])));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  int sc = 0;
  if (mx > 0.0) sc = (int)ceil(log2(mx));
  if (sc < 0) sc = 0;
  double inv = 1.0 / pow(2.0, (double)sc);
  __syncthreads();
  for (int e = tid; e < mm; e += PP_THREADS) A[e] = zscale(A[e], inv);
  __syncthreads();
  auto matmul = [&](const zd *X, const zd *Y, zd *Z) {
    for (int e = tid; e < mm; e += PP_THREADS) {
      int i = e % m, j = e / m;
      zd acc = zmk(0.0, 0.0);
      for (int k = 0; k < m; k++) zfma(acc, X[i + k * m], Y[k + j * m]);
      Z[e] = acc;
    }
    __syncthreads();
  };
  auto copy = [&](const zd *X, zd *Z) {
    for (int e = tid; e < mm; e += PP_THREADS) Z[e] = X[e];
    __syncthreads();
  };
  matmul(A, A, A2);
  const int order = 6;
  double c[order + 1];
  c[0] = 1.0;
  for (int i = 0; i < order; i++) c[i + 1] = c[i] * (double)(order - i) / (double)((i + 1) * (2 * order - i));
  for (int e = tid; e < mm; e += PP_THREADS) {
    int i = e % m, j = e / m;
    Q[e] = zmk(i == j ? c[order] : 0.0, 0.0);
    P[e] = zmk(i == j ? c[order - 1] : 0.0, 0.0);
  }
  __syncthreads();
  bool odd = true;
  for (int i = order - 2; i >= 0; i--) {
    zd *X = odd ? Q : P;
    matmul(X, A2, T);
    for (int e = tid; e < mm; e += PP_THREADS) {
      zd v = T[e];
      if (e % m == e / m) v.x += c[i];
      X[e] = v;
    }
    __syncthreads();
    odd = !odd;
  }
  if (odd) { matmul(Q, A, T); copy(T, Q); }
  else { matmul(P, A, T); copy(T, P); }
  for (int e = tid; e < mm; e += PP_THREADS) Q[e] = zsub(Q[e], P[e]);
  __syncthreads();
  // LU with partial pivoting (izamax: |re|+|im|) applied to Q and the RHS P
  for (int k = 0; k < m; k++) {
    if (tid == 0) {
      int p = k;
      double best = fabs(Q[k + k * m].x) + fabs(Q[k + k * m].y);
      for (int i = k + 1; i < m; i++) {
        double v = fabs(Q[i + k * m].x) + fabs(Q[i + k * m].y);
        if (v > best) { best = v; p = i; }
      }
      ipiv[0] = p;
      ipiv[1] = (best == 0.0);
    }
    __syncthreads();
    int p = ipiv[0];
    if (ipiv[1]) { if (tid == 0) diag[0] = -1.0; return; }
    if (p != k)
      for (int j = tid; j < 2 * m; j += PP_THREADS) {
        zd *M = j < m ? Q : P;
        int jj = j < m ? j : j - m;
        zd t = M[k + jj * m]; M[k + jj * m] = M[p + jj * m]; M[p + jj * m] = t;
      }
    __syncthreads();
    zd piv = Q[k + k * m];
    double dn = zabs2(piv);
    zd pinv = zmk(piv.x / dn, -piv.y / dn);
    for (int i = k + 1 + tid; i < m; i += PP_THREADS) Q[i + k * m] = zmul(Q[i + k * m], pinv);
    __syncthreads();
    int rem = m - k - 1;
    for (int e = tid; e < rem * rem; e += PP_THREADS) {
      int i = k + 1 + e % rem, j = k + 1 + e / rem;
      zd v = Q[i + j * m];
      zd l = Q[i + k * m], u = Q[k + j * m];
      v = zsub(v, zmul(l, u));
      Q[i + j * m] = v;
    }
    // forward elimination of the RHS column block for step k
    for (int e = tid; e < rem * m; e += PP_THREADS) {
      int i = k + 1 + e % rem, j = e / rem;
      P[i + j * m] = zsub(P[i + j * m], zmul(Q[i + k * m], P[k + j * m]));
    }
    __syncthreads();
  }
  // back substitution U X = P
  for (int k = m - 1; k >= 0; k--) {
    zd piv = Q[k + k * m];
    double dn = zabs2(piv);
    zd pinv = zmk(piv.x / dn, -piv.y / dn);
    for (int j = tid; j < m; j += PP_THREADS) P[k + j * m] = zmul(P[k + j * m], pinv);
    __syncthreads();
    for (int e = tid; e < k * m; e += PP_THREADS) {
      int i = e % k, j = e / k;
      P[i + j * m] = zsub(P[i + j * m], zmul(Q[i + k * m], P[k + j * m]));
    }
    __syncthreads();
  }
  // E = +-(I + 2X)
  for (int e = tid; e < mm; e += PP_THREADS) {
    zd v = zscale(P[e], 2.0);
    if (e % m == e / m) v.x += 1.0;
    if (odd) v = zmk(-v.x, -v.y);
    A[e] = v;
  }
  __syncthreads();
  for (int s = 0; s < sc; s++) { matmul(A, A, T); copy(T, A); }
  for (int i = tid; i < m; i += PP_THREADS) cvec[i] = A[i];
  if (tid == 0) diag[0] = sqrt(zabs2(A[m - 1]));
  __syncthreads();
}

// ------------------------------------------------------------------ the step kernel
template <int KD, int MAXB>
__global__ void __launch_bounds__(PP_THREADS) pp_pamp_kernel(pp_pamp_args A)
{
  cg::grid_group grid = cg::this_grid();
  extern __shared__ zd lds[];
  const int N = A.N, L = A.L, m = A.m, a = blockIdx.x, tid = threadIdx.x;
  const int padN = A.padN;
  zd *W = lds;                 // residual w of fiber a (stays in LDS across barriers)
  zd *X = lds + padN;          // solve / scratch buffer
  zd *Yn = lds + 2 * padN;     // staged neighbour fiber
  double *wsum = (double *)(Yn + N);          // 4 waves x (2m+4)
  double *hred = wsum + 4 * (2 * m + 4);      // reduced values (2m+4)
  double *gscr = hred + (2 * m + 4);          // grid-sum scratch (PP_THREADS)
  double *hacc = gscr + PP_THREADS;           // accumulated projections (2m) when re-orthogonalising
  const double tol = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON), pamp.h:75
  const pp_overlap &ov = A.ov;
  zd *Va = A.V + (size_t)a * m * N;
  double *mypart = A.part + (size_t)a * A.pstr;
  int tcount = 0;
  auto stamp = [&](int step) {
    if (A.tlog && step == 0 && a == 0 && tid == 0 && tcount < PP_TLOG_MAX) A.tlog[tcount++] = (long long)wall_clock64();
  };
  const int NRM = 2 * m;        // slot of |w|^2 in the partial sums
  const int ONRM = 2 * m + 1;   // slot of the pre-projection norm (re-orthogonalisation)

  for (int step = 0; step < A.nsteps; step++) {
    // ---- psi~ = U psi (MultiplySqrtOverlap, PamPropagator.py:53) ; y = U^-1 psi~
    for (int i = tid; i < N; i += PP_THREADS) X[i] = A.psi[(size_t)a * N + i];
    __syncthreads();
    for (int i = tid; i < N; i += PP_THREADS) {
      zd acc = zmk(0.0, 0.0);
      int j1 = i + KD < N - 1 ? i + KD : N - 1;
      for (int j = i; j <= j1; j++) zfmar(acc, ov.Uband[(size_t)j * (KD + 1) + KD + i - j], X[j]);
      W[i] = acc;
    }
    __syncthreads();
    block_sums(1, [&](int) {
      double s = 0.0;
      for (int i = tid; i < N; i += PP_THREADS) s += zabs2(W[i]);
      return s; }, wsum, hred);
    if (tid == 0) mypart[NRM] = hred[0];
    for (int i = tid; i < N; i += PP_THREADS) X[i] = W[i];
    __syncthreads();
    if (tid < 64) pp_band_solve_wave<KD, MAXB>(ov.solveN, X, false, N);
    __syncthreads();
    for (int i = tid; i < N; i += PP_THREADS) A.y[(size_t)a * N + i] = X[i];
    stamp(step);
    grid.sync();
    stamp(step);

    // Every workgroup executes the same number of grid barriers whatever the
    // data: breakdown and re-orthogonalisation only predicate work, so a
    // disagreement between workgroups could never leave one waiting.
    double beta0 = 0.0, beta = 0.0;
    int mEff = m;
    bool broken = false;
    int reorth_passes = 0;
    for (int j = 0; j < m; j++) {
      // ---------------- phase A: beta_j, v_j = w/beta, w' = U^-H H U^-1 v_j, <v_k, w'>
      const int ncomp = 2 * (j + 1);
      grid_partial_sum(A.part, A.pstr, NRM, 1, L, hred, gscr);
      beta = sqrt(hred[0]);
      if (j == 0) beta0 = beta;
      else if (!broken) {
        if (fabs(beta) < tol) { broken = true; mEff = j; }   // happy breakdown (pamp.h:305-312)
        else if (a == 0 && tid == 0) A.Hm[j + (j - 1) * m] = zmk(beta, 0.0);
      }
      if (!broken) {
        const double inv = 1.0 / beta;
        for (int i = tid; i < N; i += PP_THREADS) {
          zd v = zscale(W[i], inv);
          W[i] = v;
          Va[(size_t)j * N + i] = v;
        }
        for (int i = tid; i < padN; i += PP_THREADS) X[i] = zmk(0.0, 0.0);
        __syncthreads();
        for (int t = 0; t < A.nterms; t++) {
          const double st = A.scal[(size_t)step * A.nterms + t] * inv;
          const pp_kterm T = A.term[t];
          const int e0 = T.rowptr[a], e1 = T.rowptr[a + 1];
          for (int e = e0; e < e1; e++) {
            const int col = T.ent[3 * e], pidx = T.ent[3 * e + 1], cj = T.ent[3 * e + 2];
            for (int i = tid; i < N; i += PP_THREADS) Yn[i] = A.y[(size_t)col * N + i];
            __syncthreads();
            const zd *Vp = T.pot + (size_t)pidx * N * A.bw;
            for (int i = tid; i < N; i += PP_THREADS) {
              zd acc = zmk(0.0, 0.0);
              int d0 = -A.s, d1 = A.s;
              if (i + d0 < 0) d0 = -i;
              if (i + d1 > N - 1) d1 = N - 1 - i;
              for (int d = d0; d <= d1; d++) {
                zd v = Vp[(size_t)(i + d) * A.bw + A.s - d];
                if (cj) zfmac(acc, v, Yn[i + d]); else zfma(acc, v, Yn[i + d]);
              }
              zd x = X[i];
              x.x = fma(st, acc.x, x.x);
              x.y = fma(st, acc.y, x.y);
              X[i] = x;
            }
            __syncthreads();
          }
        }
        stamp(step);
        if (tid < 64) pp_band_solve_wave<KD, MAXB>(ov.solveC, X, true, N);
        __syncthreads();
        stamp(step);
        // w' -> W ; partial <v_k, w'> for k <= j ; |w'|^2 for re-orthogonalisation
        for (int i = tid; i < N; i += PP_THREADS) W[i] = X[i];
        __syncthreads();
        fiber_dots(Va, W, j + 1, N, A.reorth != 0, hred);
        for (int c = tid; c < ncomp; c += PP_THREADS) mypart[c] = hred[c];
        if (A.reorth && tid == 0) mypart[ONRM] = hred[ncomp];
      }
      stamp(step);
      grid.sync();
      stamp(step);

      // ---------------- phase B: h = sum_a partials ; w' -= V h ; |w'|^2 ; y = U^-1 w'
      if (!broken) {
        grid_partial_sum(A.part, A.pstr, 0, ncomp, L, hred, gscr);
        if (a == 0 && tid <= j) A.Hm[tid + j * m] = zmk(hred[2 * tid], hred[2 * tid + 1]);
      }
      if (j == m - 1) break;   // the last residual is not used (pamp.h:419-426)
      if (!broken) {
        for (int i = tid; i < N; i += PP_THREADS) {
          zd w = W[i];
#pragma unroll 4
          for (int k = 0; k <= j; k++) {
            zd h = zmk(hred[2 * k], hred[2 * k + 1]);
            zd v = Va[(size_t)k * N + i];
            w = zsub(w, zmul(v, h));
          }
          W[i] = w;
        }
        __syncthreads();
      }
      if (A.reorth) {
        // PerformOrthogonalization (pamp.h:353-400): repeat CGS while the
        // residual shrank below 0.717 of the previous norm, at most 3 times.
        double orig = 0.0;
        bool stop = broken;
        if (!broken) {
          for (int c = tid; c < ncomp; c += PP_THREADS) hacc[c] = hred[c];
          grid_partial_sum(A.part, A.pstr, ONRM, 1, L, hred, gscr);
          orig = sqrt(hred[0]);
          __syncthreads();
        }
        for (int pass = 0; pass < 4; pass++) {
          if (!stop) {
            block_sums(1, [&](int) { double s = 0.0; for (int i = tid; i < N; i += PP_THREADS) s += zabs2(W[i]); return s; }, wsum, hred);
            if (tid == 0) mypart[NRM] = hred[0];
          }
          grid.sync();
          bool doit = false;
          double res = 0.0;
          if (!stop) {
            grid_partial_sum(A.part, A.pstr, NRM, 1, L, hred, gscr);
            res = sqrt(hred[0]);
            __syncthreads();
            if (fabs(res) < 0.717 * fabs(orig) && pass < 3) doit = true;
            else stop = true;
          }
          if (doit) {
            fiber_dots(Va, W, j + 1, N, false, hred);
            for (int c = tid; c < ncomp; c += PP_THREADS) mypart[c] = hred[c];
          }
          grid.sync();
          if (doit) {
            grid_partial_sum(A.part, A.pstr, 0, ncomp, L, hred, gscr);
            for (int i = tid; i < N; i += PP_THREADS) {
              zd w = W[i];
              for (int k = 0; k <= j; k++) w = zsub(w, zmul(Va[(size_t)k * N + i], zmk(hred[2 * k], hred[2 * k + 1])));
              W[i] = w;
            }
            __syncthreads();
            for (int c = tid; c < ncomp; c += PP_THREADS) hacc[c] += hred[c];
            __syncthreads();
            if (a == 0 && tid <= j) A.Hm[tid + j * m] = zmk(hacc[2 * tid], hacc[2 * tid + 1]);
            orig = res;
            reorth_passes++;
            __syncthreads();
          }
        }
        grid.sync();   // everyone has read NRM before it is rewritten below
      }
      if (!broken) {
        block_sums(1, [&](int) {
          double s = 0.0;
          for (int i = tid; i < N; i += PP_THREADS) s += zabs2(W[i]);
          return s; }, wsum, hred);
        if (tid == 0) mypart[NRM] = hred[0];
        for (int i = tid; i < padN; i += PP_THREADS) X[i] = (i < N) ? W[i] : zmk(0.0, 0.0);
        __syncthreads();
        stamp(step);
        if (tid < 64) pp_band_solve_wave<KD, MAXB>(ov.solveN, X, false, N);
        __syncthreads();
        for (int i = tid; i < N; i += PP_THREADS) A.y[(size_t)a * N + i] = X[i];
      }
      stamp(step);
      grid.sync();
      stamp(step);
    }

    // ---------------- exponentiate on workgroup 0, broadcast exp(-i dt H) e1
    if (a == 0) {
      pade_block(m, A.Hm, A.dt_re, A.dt_im, lds, A.cvec, A.diag + (size_t)step * 4);
      if (tid == 0) {
        A.diag[(size_t)step * 4 + 1] = mEff;
        A.diag[(size_t)step * 4 + 2] = beta0;
        A.diag[(size_t)step * 4 + 3] = reorth_passes;
      }
      // reset the Hessenberg matrix for the next step
      __syncthreads();
      for (int e = tid; e < m * m; e += PP_THREADS) A.Hm[e] = zmk(0.0, 0.0);
    }
    stamp(step);
    grid.sync();
    stamp(step);
    // ---------------- psi~ = |psi~| V c ; psi = U^-1 psi~ (SolveSqrtOverlap)
    for (int i = tid; i < padN; i += PP_THREADS) {
      zd acc = zmk(0.0, 0.0);
      if (i < N)
        for (int k = 0; k < mEff; k++) zfma(acc, Va[(size_t)k * N + i], A.cvec[k]);
      X[i] = zscale(acc, beta0);
    }
    __syncthreads();
    if (tid < 64) pp_band_solve_wave<KD, MAXB>(ov.solveN, X, false, N);
    __syncthreads();
    for (int i = tid; i < N; i += PP_THREADS) A.psi[(size_t)a * N + i] = X[i];
    __syncthreads();
    stamp(step);
  }
  if (A.tlog && a == 0 && tid == 0) A.tlog[PP_TLOG_MAX] = tcount;
}

// ------------------------------------------------------------------ host API
extern "C" int pp_pamp_create(int L, int N, int m, int reorth, const pp_overlap *ov, pp_pamp **out)
{
  if (!out || !ov || L < 1 || N < 1 || m < 1) return PP_EARG;
  if (ov->N != N) return PP_EARG;
  if (m > PP_MAXM) return PP_EUNSUP;
  pp_pamp *p = new pp_pamp();
  memset(p, 0, sizeof(*p));
  p->L = L; p->N = N; p->m = m; p->reorth = reorth ? 1 : 0; p->ov = ov;
  p->pstr = 2 * m + 4;
  p->s = -1;
  PP_HIP(hipMalloc(&p->V, sizeof(zd) * (size_t)L * m * N));
  PP_HIP(hipMalloc(&p->y, sizeof(zd) * (size_t)L * N));
  PP_HIP(hipMalloc(&p->Hm, sizeof(zd) * (size_t)m * m));
  PP_HIP(hipMalloc(&p->cvec, sizeof(zd) * (size_t)m));
  PP_HIP(hipMalloc(&p->part, sizeof(double) * (size_t)L * p->pstr));
  PP_HIP(hipMemset(p->V, 0, sizeof(zd) * (size_t)L * m * N));
  PP_HIP(hipMemset(p->Hm, 0, sizeof(zd) * (size_t)m * m));
  *out = p;
  return PP_OK;
}

// One tensor potential: angular storage (Ident/Diag/Simp/Herm) with P0
// compressed angular indices (pairs host int32 P0x2 for Simp/Herm), radial
// BandNH band of total width bw (=2s+1) per B-spline, device data (P0, N*bw).
extern "C" int pp_pamp_add_term(pp_pamp *p, int angular_storage, int P0, const int *pairs, int radial_storage,
                                int bw, const void *pot)
{
  if (!p || !pot || P0 < 1) return PP_EARG;
  if (p->nterms >= PP_MAXTERMS) return PP_EUNSUP;
  if (radial_storage != PP_BANDNH) return PP_EUNSUP;
  if (bw < 1 || (bw % 2) == 0) return PP_EARG;
  if (p->s >= 0 && bw != p->bw) return PP_EUNSUP;
  pp_rank_csr csr;
  int rc = pp_build_csr(angular_storage, P0, p->L, pairs, csr);
  if (rc) return rc;
  std::vector<int> host(csr.rowptr);
  host.insert(host.end(), csr.ent.begin(), csr.ent.end());
  int *dev = nullptr;
  PP_HIP(hipMalloc(&dev, host.size() * sizeof(int)));
  PP_HIP(hipMemcpy(dev, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice));
  int t = p->nterms++;
  p->csr_dev[t] = dev;
  p->term[t].rowptr = dev;
  p->term[t].ent = dev + csr.rowptr.size();
  p->term[t].pot = (const zd *)pot;
  p->bw = bw;
  p->s = (bw - 1) / 2;
  return PP_OK;
}

static size_t pamp_lds_bytes(const pp_pamp *p, int padN)
{
  size_t fib = sizeof(zd) * (size_t)(2 * padN + p->N) +
               sizeof(double) * (size_t)(4 * (2 * p->m + 4) + (2 * p->m + 4) + PP_THREADS + 2 * p->m);
  size_t pade = sizeof(zd) * (size_t)(5 * p->m * p->m) + sizeof(double) * 8 + 2 * sizeof(int) + 16;
  return fib > pade ? fib : pade;
}

// Propagate psi (device, L x N complex128) through nsteps steps of length dt.
// scal: host array nsteps x nterms with each term's time factor at t_n.
extern "C" int pp_pamp_propagate(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im,
                                 const double *scal, hipStream_t stream)
{
  if (!p || !psi || nsteps < 0 || (!scal && p->nterms > 0)) return PP_EARG;
  if (nsteps == 0) return PP_OK;
  if (p->nterms == 0) return PP_EARG;
  const pp_overlap *ov = p->ov;
  if (nsteps * p->nterms > p->scal_cap) {
    if (p->scal) (void)hipFree(p->scal);
    p->scal_cap = nsteps * p->nterms;
    PP_HIP(hipMalloc(&p->scal, sizeof(double) * p->scal_cap));
  }
  if (nsteps > p->diag_cap) {
    if (p->diag) (void)hipFree(p->diag);
    p->diag_cap = nsteps;
    PP_HIP(hipMalloc(&p->diag, sizeof(double) * 4 * (size_t)nsteps));
  }
  PP_HIP(hipMemcpyAsync(p->scal, scal, sizeof(double) * nsteps * p->nterms, hipMemcpyHostToDevice, stream));
  pp_pamp_args A;
  memset(&A, 0, sizeof(A));
  A.L = p->L; A.N = p->N; A.m = p->m; A.s = p->s; A.bw = p->bw; A.nterms = p->nterms;
  A.nsteps = nsteps; A.reorth = p->reorth;
  A.padN = ov->solveN.P * ov->solveN.B;
  if (A.padN < p->N) A.padN = p->N;
  A.dt_re = dt_re; A.dt_im = dt_im;
  for (int t = 0; t < p->nterms; t++) A.term[t] = p->term[t];
  A.scal = p->scal;
  A.ov = *ov;
  A.psi = (zd *)psi; A.V = p->V; A.y = p->y; A.part = p->part; A.pstr = p->pstr;
  A.Hm = p->Hm; A.cvec = p->cvec; A.diag = p->diag;
  A.tlog = p->timing ? p->tlog : nullptr;
  size_t lds = pamp_lds_bytes(p, A.padN);
  void *args[] = {&A};
  int maxb = ov->solveN.B;
  bool launched = false;
  hipError_t err = hipSuccess;
  auto go = [&](const void *fn) -> hipError_t {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipLaunchCooperativeKernel(fn, dim3(p->L), dim3(PP_THREADS), args, (unsigned)lds, stream);
  };
  if (maxb <= 16) {
    PP_KD_SWITCH(ov->kd, err = go((const void *)pp_pamp_kernel<KD_, 16>); launched = true)
  } else {
    PP_KD_SWITCH(ov->kd, err = go((const void *)pp_pamp_kernel<KD_, 32>); launched = true)
  }
  if (!launched) return PP_EUNSUP;
  if (err != hipSuccess) return PP_EHIP;
  return PP_OK;
}

// diag: host nsteps x 4 (error estimate |c_{m-1}|, basis used, |psi~|, re-orth passes)
// of the last pp_pamp_propagate call; synchronises the stream.
extern "C" int pp_pamp_diagnostics(pp_pamp *p, int nsteps, double *diag, hipStream_t stream)
{
  if (!p || !diag || nsteps > p->diag_cap) return PP_EARG;
  PP_HIP(hipMemcpyAsync(diag, p->diag, sizeof(double) * 4 * (size_t)nsteps, hipMemcpyDeviceToHost, stream));
  PP_HIP(hipStreamSynchronize(stream));
  return PP_OK;
}

// Profiling hook: when enabled, the next launches record wall-clock stamps
// (100 MHz counter) of workgroup 0 around every grid barrier of step 0.
extern "C" int pp_pamp_set_timing(pp_pamp *p, int enable)
{
  if (!p) return PP_EARG;
  if (enable && !p->tlog) {
    PP_HIP(hipMalloc(&p->tlog, sizeof(long long) * (PP_TLOG_MAX + 1)));
    PP_HIP(hipMemset(p->tlog, 0, sizeof(long long) * (PP_TLOG_MAX + 1)));
  }
  p->timing = enable ? 1 : 0;
  return PP_OK;
}

// Copy the stamps of the last timed launch (host, max entries); *count = stamps written.
extern "C" int pp_pamp_timings(pp_pamp *p, long long *out, int max, int *count)
{
  if (!p || !out || !count || !p->tlog) return PP_EARG;
  long long buf[PP_TLOG_MAX + 1];
  PP_HIP(hipDeviceSynchronize());
  PP_HIP(hipMemcpy(buf, p->tlog, sizeof(buf), hipMemcpyDeviceToHost));
  int n = (int)buf[PP_TLOG_MAX];
  if (n > max) n = max;
  for (int i = 0; i < n; i++) out[i] = buf[i];
  *count = n;
  return PP_OK;
}

extern "C" int pp_pamp_destroy(pp_pamp *p)
{
  if (!p) return PP_OK;
  if (p->tlog) (void)hipFree(p->tlog);
  for (int t = 0; t < p->nterms; t++) if (p->csr_dev[t]) (void)hipFree(p->csr_dev[t]);
  if (p->V) (void)hipFree(p->V);
  if (p->y) (void)hipFree(p->y);
  if (p->Hm) (void)hipFree(p->Hm);
  if (p->cvec) (void)hipFree(p->cvec);
  if (p->part) (void)hipFree(p->part);
  if (p->diag) (void)hipFree(p->diag);
  if (p->scal) (void)hipFree(p->scal);
  delete p;
  return PP_OK;
}

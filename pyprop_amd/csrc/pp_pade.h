// Shared device helpers of the Krylov engines: agent-scope exchange loads and
// stores, and the scaling + Pade(6) exponential of the small Hessenberg
// matrix on one workgroup of PP_THREADS threads.
#pragma once
#include "pp_common.h"

#ifndef PP_THREADS
#define PP_THREADS 256
#endif

// ------------------------------------------------------------------ exchanged data
__device__ inline void xst(double *p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline double xld(const double *p) { return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void zst(zd *p, zd v) { xst(&p->x, v.x); xst(&p->y, v.y); }
__device__ inline zd zld(const zd *p) { return zmk(xld(&p->x), xld(&p->y)); }

// ------------------------------------------------------------------ Pade on one workgroup
// pAMP::ScalingPadeExponential + PadeExponential (pamp.h:430-564): A = -i dt H
// scaled by 2^-s with s = max(0, ceil(log2 max|A_ij|)), Pade(6) numerator
// and denominator by Horner in A^2, X = (Q - P)^-1 P, E = +-(I + 2X), then s
// squarings; the first column of E goes to cvec_out (exchanged).  Column-major
// m x m complex matrices in the dynamic LDS at double offset `off`.  The
// linear solve is Gauss-Jordan with the same partial pivoting as zgesv
// (largest |re|+|im|, first index on ties) on ping-pong buffers, one barrier
// per column.
// normscale: 0 scales by the largest |A_ij| as pAMP does; 1 scales until the
// 1-norm of A / 2^s is at most 1/2, which keeps the Pade(6) error near
// rounding and makes the step a smooth function of its inputs (the largest
// element can be several times smaller than the norm, and a change of s
// then moves the result by the Pade error).
static __device__ __attribute__((noinline)) void pade_block(int m, const zd *Hg, double dt_re, double dt_im, int off,
                                                     zd *cvec_out, double *diag, int normscale)
{
  extern __shared__ double pp_lds[];
  zd *lds = (zd *)(pp_lds + off);
  const int tid = threadIdx.x;
  const int mm = m * m;
  zd *A = lds, *A2 = lds + mm, *P = lds + 2 * mm, *Q = lds + 3 * mm, *T = lds + 4 * mm;
  zd *G0 = lds + 5 * mm, *G1 = G0 + 2 * mm;          // Gauss-Jordan buffers (m x 2m)
  double *red = (double *)(G1 + 2 * mm);
  // element e = tid + r*PP_THREADS of an m x m matrix: row i = e % m, column j = e / m
  const int i0 = tid % m, j0 = tid / m, di = PP_THREADS % m, dj = PP_THREADS / m;
  const zd fac = zmk(dt_im, -dt_re);  // -i * dt
  for (int e = tid; e < mm; e += PP_THREADS) A[e] = zmul(fac, Hg[e]);
  __syncthreads();
  double mx = 0.0;
  if (normscale) {   // largest column sum of |A_ij|
    if (tid < m)
      for (int i = 0; i < m; i++) mx += sqrt(zabs2(A[i + tid * m]));
  } else {
    for (int e = tid; e < mm; e += PP_THREADS) mx = fmax(mx, sqrt(zabs2(A[e])));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = red[0];
  for (int w = 1; w < PP_THREADS / 64; w++) mx = fmax(mx, red[w]);
  int sc = 0;
  if (mx > 0.0) sc = (int)ceil(log2(mx)) + (normscale ? 1 : 0);
  if (sc < 0) sc = 0;
  const double inv = 1.0 / pow(2.0, (double)sc);
  for (int e = tid; e < mm; e += PP_THREADS) A[e] = zscale(A[e], inv);
  __syncthreads();
  auto matmul = [&](const zd *X, const zd *Y, zd *Z) {   // Z = X Y
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      const zd *xr = X + i, *yc = Y + j * m;
      zd a0 = zmk(0.0, 0.0), a1 = zmk(0.0, 0.0), a2 = zmk(0.0, 0.0), a3 = zmk(0.0, 0.0);
      int k = 0;
      for (; k + 3 < m; k += 4) {   // four independent chains keep 8 LDS reads in flight
        const zd x0 = xr[k * m], x1 = xr[(k + 1) * m], x2 = xr[(k + 2) * m], x3 = xr[(k + 3) * m];
        const zd y0 = yc[k], y1 = yc[k + 1], y2 = yc[k + 2], y3 = yc[k + 3];
        zfma(a0, x0, y0);
        zfma(a1, x1, y1);
        zfma(a2, x2, y2);
        zfma(a3, x3, y3);
      }
      for (; k < m; k++) zfma(a0, xr[k * m], yc[k]);
      Z[e] = zadd(zadd(a0, a1), zadd(a2, a3));
      i += di; j += dj;
      if (i >= m) { i -= m; j++; }
    }
    __syncthreads();
  };
  matmul(A, A, A2);
  const int order = 6;
  double c[order + 1];
  c[0] = 1.0;
  for (int i = 0; i < order; i++) c[i + 1] = c[i] * (double)(order - i) / (double)((i + 1) * (2 * order - i));
  {
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      Q[e] = zmk(i == j ? c[order] : 0.0, 0.0);
      P[e] = zmk(i == j ? c[order - 1] : 0.0, 0.0);
      i += di; j += dj;
      if (i >= m) { i -= m; j++; }
    }
  }
  __syncthreads();
  bool odd = true;
  for (int q = order - 2; q >= 0; q--) {
    zd *X = odd ? Q : P;
    matmul(X, A2, T);
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      zd v = T[e];
      if (i == j) v.x += c[q];
      X[e] = v;
      i += di; j += dj;
      if (i >= m) { i -= m; j++; }
    }
    __syncthreads();
    odd = !odd;
  }
  zd *Pm = odd ? Q : P;   // the odd polynomial gets the extra factor A
  matmul(Pm, A, T);
  // Gauss-Jordan on [Q - P | P] (m x 2m, column-major)
  for (int e = tid; e < mm; e += PP_THREADS) {
    const zd qv = odd ? T[e] : Q[e];       // Q after its final update
    const zd pv = odd ? P[e] : T[e];       // P after its final update
    G0[e] = zsub(qv, pv);
    G0[mm + e] = pv;
  }
  __syncthreads();
  zd *gc = G0, *gn = G1;
  const int w2 = 2 * m * m;
  for (int k = 0; k < m; k++) {
    // every wavefront finds the pivot itself: rows k..m-1 of column k, one
    // per lane, reduced by shuffles (first index on ties, like izamax)
    int p = m;
    double best = -1.0;
    const int lane = tid & 63;
    for (int i = k + lane; i < m; i += 64) {
      const double v = fabs(gc[i + k * m].x) + fabs(gc[i + k * m].y);
      if (v > best) { best = v; p = i; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int op = __shfl_xor(p, o, 64);
      if (ob > best || (ob == best && op < p)) { best = ob; p = op; }
    }
    if (best == 0.0) {   // singular denominator (pamp.h:544-550)
      if (tid == 0) diag[0] = -1.0;
      return;
    }
    const zd piv = gc[p + k * m];
    const double dn = zabs2(piv);
    const zd pinv = zmk(piv.x / dn, -piv.y / dn);
    int i = i0, j = j0;
    for (int e = tid; e < w2; e += PP_THREADS) {
      if (j >= k) {   // eliminated columns are not read again
        const zd rk = zmul(gc[p + j * m], pinv);   // new pivot row
        zd v;
        if (i == k) v = rk;
        else {
          const int si = (i == p) ? k : i;       // row swap k <-> p
          v = zsub(gc[si + j * m], zmul(gc[si + k * m], rk));
        }
        gn[e] = v;
      }
      i += di; j += dj;
      if (i >= m) { i -= m; j++; }
    }
    __syncthreads();
    zd *t = gc; gc = gn; gn = t;
  }
  // X = right half of gc;  E = +-(I + 2X)
  const bool neg = odd;
  if (sc == 0) {   // only the first column of E is needed
    for (int i = tid; i < m; i += PP_THREADS) {
      zd v = zscale(gc[mm + i], 2.0);
      if (i == 0) v.x += 1.0;
      if (neg) v = zmk(-v.x, -v.y);
      zst(cvec_out + i, v);
      if (i == m - 1) diag[0] = sqrt(zabs2(v));
    }
    __syncthreads();
    return;
  }
  {
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      zd v = zscale(gc[mm + e], 2.0);
      if (i == j) v.x += 1.0;
      if (neg) v = zmk(-v.x, -v.y);
      A[e] = v;
      i += di; j += dj;
      if (i >= m) { i -= m; j++; }
    }
  }
  __syncthreads();
  for (int s = 0; s < sc; s++) {
    matmul(A, A, T);
    for (int e = tid; e < mm; e += PP_THREADS) A[e] = T[e];
    __syncthreads();
  }
  for (int i = tid; i < m; i += PP_THREADS) zst(cvec_out + i, A[i]);
  if (tid == 0) diag[0] = sqrt(zabs2(A[m - 1]));
  __syncthreads();
}


// Shared device helpers of the Krylov engines: agent-scope exchange loads and
// stores, and the scaling + Pade(6) exponential of the small Hessenberg
// matrix on one workgroup of PP_THREADS threads.
#pragma once
#include "pp_common.h"

#ifndef PP_THREADS
#define PP_THREADS 256
#endif
// Gauss-Jordan elements a thread of wavefronts 1.. holds in registers per
// elimination step (the rest, when a step has more, four at a time)
#ifndef PP_PADE_NU
#define PP_PADE_NU 6
#endif

// ------------------------------------------------------------------ exchanged data
__device__ inline void xst(double *p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline double xld(const double *p) { return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void zst(zd *p, zd v) { xst(&p->x, v.x); xst(&p->y, v.y); }
__device__ inline zd zld(const zd *p) { return zmk(xld(&p->x), xld(&p->y)); }

// ------------------------------------------------------------------ Pade on one workgroup
// pAMP::ScalingPadeExponential + PadeExponential (pamp.h:430-564): A = -i dt H
// scaled by 2^-s with s = max(0, ceil(log2 max|A_ij|)), the Pade(6)
// numerator and denominator, X = (Q - P)^-1 P, E = I + 2X, then s
// squarings; the first column of E goes to cvec_out (exchanged).  Column-
// major m x m complex matrices in the dynamic LDS at double offset `off`.
// The even and odd parts are formed from the powers A^2, A^4, A^6 (Q = c0 +
// c2 A^2 + c4 A^4 + c6 A^6, P = A (c1 + c3 A^2 + c5 A^4): four matrix
// products instead of Horner's seven, the same polynomials).  The linear
// solve is Gauss-Jordan with zgesv's partial pivoting (largest |re|+|im|,
// first index on ties) on ping-pong buffers, one barrier per column; the
// wavefront that forms the next column also picks its pivot, so a step
// starts with the pivot's value and row already in LDS (m <= 64).
// The last squaring forms only the first column.  Matrix products run on
// v_mfma_f64_16x16x4f64.
// normscale: 0 scales by the largest |A_ij| as pAMP does; 1 scales until the
// 1-norm of A / 2^s is at most 1/2, which keeps the Pade(6) error near
// rounding and makes the step a smooth function of its inputs (the largest
// element can be several times smaller than the norm, and a change of s
// then moves the result by the Pade error); 2 selects the eigenvector
// exponential (eig_block) instead of Pade.
static __device__ __attribute__((noinline)) void eig_block(int m, const zd *Hg, double dt_re, double dt_im, int off,
                                                    zd *cvec_out, double *diag);

typedef double pp_d4 __attribute__((ext_vector_type(4)));

// Doubles of dynamic LDS the exponential of an m x m Krylov matrix needs at
// its offset: five m x m complex buffers (Pade) or the eigenvector form's two
// padded matrices and rotation table, then the wavefront maxima, two
// Gauss-Jordan pivot slots and its row permutation.
__host__ __device__ inline int pp_pade_lds_doubles(int m)
{
  const int n2 = m + (m & 1), np = n2 / 2;
  const int pade = 2 * 5 * m * m, eig = 2 * 2 * n2 * n2 + 2 * np + 2 * np + np + 2;
  return (pade > eig ? pade : eig) + PP_THREADS / 64 + 8 + 32 + 2;
}

// Z = X Y (+ c on the diagonal when diag), m x m complex, column-major, LDS:
// 16 x 16 output tiles, one per wavefront at a time, with the real and
// imaginary parts accumulated by v_mfma_f64_16x16x4f64 (FP64 products and
// sums).  Operand lanes: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k;
// D[4r + l/16][l%16] comes back in lane l, register r.  Ends with a barrier.
__device__ __attribute__((always_inline)) inline void pade_mm(int m, const zd *X, const zd *Y, zd *Z, double c, bool diag)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int nt = (m + 15) >> 4;
  for (int t = wave; t < nt * nt; t += PP_THREADS / 64) {
    const int ti = t % nt, tj = t / nt;
    const int i = ti * 16 + li, j = tj * 16 + li;
    pp_d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < m; k0 += 4) {
      const int k = k0 + lk;
      const zd x = (i < m && k < m) ? X[i + k * m] : zmk(0.0, 0.0);
      const zd y = (k < m && j < m) ? Y[k + j * m] : zmk(0.0, 0.0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, ci, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, y.y, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.x, ci, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = ti * 16 + 4 * r + lk, col = tj * 16 + li;
      if (row < m && col < m) {
        zd v = zmk(cr[r], ci[r]);
        if (diag && row == col) v.x += c;
        Z[row + col * m] = v;
      }
    }
  }
  __syncthreads();
}

// Lane l's value of d (l uniform).
__device__ __attribute__((always_inline)) inline double pade_readlane(double d, int l)
{
  const long long b = __double_as_longlong(d);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Wavefront maximum (MAX) or minimum of x, reduced into lane 63 by DPP moves
// (quad swaps, half-row and row mirrors, then the row broadcasts) and read
// from there.  Lanes a move leaves out (row mask) combine with themselves.
template <bool MAX>
__device__ __attribute__((always_inline)) inline unsigned int pade_wave_ext(unsigned int x)
{
#define PADE_EXT_LEVEL(CTRL, RMASK)                                                          \
  {                                                                                        \
    const unsigned int o = (unsigned int)__builtin_amdgcn_update_dpp((int)x, (int)x, CTRL, RMASK, 0xF, false); \
    x = MAX ? (o > x ? o : x) : (o < x ? o : x);                                           \
  }
  PADE_EXT_LEVEL(0xB1, 0xF)    // quad_perm [1,0,3,2]
  PADE_EXT_LEVEL(0x4E, 0xF)    // quad_perm [2,3,0,1]
  PADE_EXT_LEVEL(0x141, 0xF)   // row_half_mirror
  PADE_EXT_LEVEL(0x140, 0xF)   // row_mirror
  PADE_EXT_LEVEL(0x142, 0xA)   // row_bcast:15 into rows 1, 3
  PADE_EXT_LEVEL(0x143, 0xC)   // row_bcast:31 into rows 2, 3
#undef PADE_EXT_LEVEL
  return (unsigned int)__builtin_amdgcn_readlane((int)x, 63);
}

// Partial pivoting of one Gauss-Jordan column, one wavefront, lane r holding
// row r of m.  Rows stay where they are; pos is the row's place in zgesv's
// row-swapped order.  The pivot of step k is zgesv's izamax over places
// >= k: the largest |re| + |im|, the first place on ties.  |re| + |im| >= 0,
// so its bits order like its value: the largest high word, then the largest
// low word among those, then the smallest place among exact ties.  Lane 0
// stores the pivot's value and row in slot[0..2]; every lane then applies
// the swap of places k and the pivot's, and at the last step stores
// row_of[pos] = r.
__device__ __attribute__((always_inline)) inline void pade_pivot(zd v, bool valid, int k, int m, int &pos, double *slot,
                                                                 int *row_of)
{
  const int lane = threadIdx.x & 63;
  const bool cand = valid && pos >= k;
  const unsigned long long b = (unsigned long long)__double_as_longlong(cand ? fabs(v.x) + fabs(v.y) : 0.0);
  const unsigned int hi = (unsigned int)(b >> 32), lo = (unsigned int)b;
  const unsigned int mh = pade_wave_ext<true>(hi);
  const unsigned int ml = pade_wave_ext<true>(hi == mh ? lo : 0u);
  const int pk = (int)pade_wave_ext<false>(cand && hi == mh && lo == ml ? (unsigned int)pos : 64u);
  const unsigned long long at = __ballot(cand && pos == pk);   // bits of valid lanes only: r < m
  const int r = at ? (int)__builtin_ctzll(at) : 0;
  const double vx = pade_readlane(v.x, r), vy = pade_readlane(v.y, r);
  if (lane == 0) { slot[0] = vx; slot[1] = vy; slot[2] = (double)r; }
  if (at) {
    if (pos == k) pos = pk;
    if (lane == r) pos = k;
  }
  if (k == m - 1 && valid) row_of[pos] = lane;
}

static __device__ __attribute__((noinline)) void pade_block(int m, const zd *Hg, double dt_re, double dt_im, int off,
                                                     zd *cvec_out, double *diag, int normscale)
{
  if (normscale == 2) {
    eig_block(m, Hg, dt_re, dt_im, off, cvec_out, diag);
    return;
  }
  extern __shared__ double pp_lds[];
  zd *lds = (zd *)(pp_lds + off);
  const int tid = threadIdx.x;
  const int mm = m * m;
  // five m x m buffers; Gauss-Jordan ping-pongs between [A | A2] and
  // [A4 | A6] (m x 2m each) once Q and P are formed, and the squarings use Wt
  // and A, so the exponential needs 5 m^2 complex values in all
  zd *A = lds, *A2 = lds + mm, *A4 = lds + 2 * mm, *A6 = lds + 3 * mm, *Wt = lds + 4 * mm;
  zd *G0 = A, *G1 = A4;                               // Gauss-Jordan buffers (m x 2m)
  double *red = (double *)(lds + 5 * mm);              // wavefront maxima
  double *pvs = red + PP_THREADS / 64;                 // pivot slots: value, row (x2)
  int *row_of = (int *)(pvs + 8);                      // solution row -> Gauss-Jordan row (64)
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
  const int order = 6;
  double c[order + 1];
  c[0] = 1.0;
  for (int i = 0; i < order; i++) c[i + 1] = c[i] * (double)(order - i) / (double)((i + 1) * (2 * order - i));
  pade_mm(m, A, A, A2, 0.0, false);
  pade_mm(m, A2, A2, A4, 0.0, false);
  pade_mm(m, A4, A2, A6, 0.0, false);
  // even part Q (over A4), the odd part's cofactor c1 + c3 A^2 + c5 A^4 (Wt)
  for (int e = tid; e < mm; e += PP_THREADS) {
    const int i = e % m, j = e / m;
    const zd a2 = A2[e], a4 = A4[e], a6 = A6[e];
    zd w = zadd(zscale(a4, c[5]), zscale(a2, c[3]));
    zd q = zadd(zadd(zscale(a6, c[6]), zscale(a4, c[4])), zscale(a2, c[2]));
    if (i == j) { w.x += c[1]; q.x += c[0]; }
    Wt[e] = w;
    A4[e] = q;
  }
  __syncthreads();
  pade_mm(m, A, Wt, A6, 0.0, false);   // P = A (c1 + c3 A^2 + c5 A^4)
  // Gauss-Jordan on [Q - P | P] (m x 2m, column-major) in [A | A2] (A and A2
  // are no longer needed) with zgesv's partial pivoting, rows left in place:
  // row r ends as row pos(r) of the solution, row_of inverts that.  Step k
  // updates columns k+1 .. 2m-1 only (column k is not read again).
  // Wavefront 0 forms the new column k+1 (lane r: row r, m <= 64) and picks
  // the next pivot from its registers, leaving (value, row) in one of two LDS
  // slots before the step's barrier; the other wavefronts update columns
  // k+2 .. 2m-1.
  for (int e = tid; e < mm; e += PP_THREADS) {
    const zd pv = A6[e];
    G0[e] = zsub(A4[e], pv);
    G0[mm + e] = pv;
  }
  int pos = tid & 63;   // wavefront 0: the place of this lane's row
  if (tid < 64) pade_pivot(tid < m ? zsub(A4[tid], A6[tid]) : zmk(0.0, 0.0), tid < m, 0, m, pos, pvs, row_of);
  __syncthreads();
  zd *gc = G0, *gn = G1;
  constexpr int T1 = PP_THREADS - 64;
  const int t1 = tid - 64, ti = (t1 + m) % m, tj = (t1 + m) / m, di = T1 % m, dj = T1 / m, jl = 2 * m - 1;
  constexpr int NU = PP_PADE_NU;
  for (int k = 0; k < m; k++) {
    const int nel = (2 * m - k - 2) * m;   // elements of columns k+2 .. 2m-1
    const double *sl = pvs + (k & 1) * 4;
    const zd piv = zmk(sl[0], sl[1]);
    const int p = __builtin_amdgcn_readfirstlane((int)sl[2]);
    if (tid < 64) {
      const int i = min(tid, m - 1);
      const zd gb0 = gc[i + (k + 1) * m], gk0 = gc[i + k * m];
      if (piv.x == 0.0 && piv.y == 0.0) {   // singular denominator (pamp.h:544-550)
        if (tid == 0) diag[0] = -1.0;
        return;
      }
      const double dn = zabs2(piv);
      const zd pinv = zmk(piv.x / dn, -piv.y / dn);
      const zd rk = zmul(gc[p + (k + 1) * m], pinv);   // pivot row / pivot
      const zd v = i == p ? rk : zsub(gb0, zmul(gk0, rk));
      if (tid < m) gn[i + (k + 1) * m] = v;
      if (k + 1 < m) pade_pivot(v, tid < m, k + 1, m, pos, pvs + ((k + 1) & 1) * 4, row_of);
    } else {
      // this thread's first NU elements: the operands that do not depend on
      // the pivot, loaded with it (columns past the block's end clamped to
      // its last)
      int iu[NU], ju[NU];
      zd ga[NU], gb[NU], gk[NU];
      int i = ti, j = k + 1 + tj;
#pragma unroll
      for (int u = 0; u < NU; u++) {
        iu[u] = i; ju[u] = min(j, jl);
        gb[u] = gc[i + ju[u] * m];
        gk[u] = gc[i + k * m];
        i += di; j += dj;
        if (i >= m) { i -= m; j++; }
      }
      if (piv.x == 0.0 && piv.y == 0.0) return;   // singular: wavefront 0 reports it
      const double dn = zabs2(piv);
      const zd pinv = zmk(piv.x / dn, -piv.y / dn);
#pragma unroll
      for (int u = 0; u < NU; u++) ga[u] = gc[p + ju[u] * m];
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const zd rk = zmul(ga[u], pinv);
        const zd v = iu[u] == p ? rk : zsub(gb[u], zmul(gk[u], rk));
        if (t1 + u * T1 < nel) gn[iu[u] + ju[u] * m] = v;
      }
      // the rest (more than NU * T1 elements): four elements at a time
      for (int r0 = t1 + NU * T1; r0 < nel; r0 += 4 * T1) {
        zd ga2[4], gb2[4], gk2[4];
        int iv[4], jv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          iv[u] = i; jv[u] = min(j, jl);
          ga2[u] = gc[p + jv[u] * m];
          gb2[u] = gc[i + jv[u] * m];
          gk2[u] = gc[i + k * m];
          i += di; j += dj;
          if (i >= m) { i -= m; j++; }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (r0 + u * T1 < nel) {
            const zd rk = zmul(ga2[u], pinv);
            gn[iv[u] + jv[u] * m] = iv[u] == p ? rk : zsub(gb2[u], zmul(gk2[u], rk));
          }
      }
    }
    __syncthreads();
    zd *t = gc; gc = gn; gn = t;
  }
  // X = right half of gc;  E = I + 2X
  if (sc == 0) {   // only the first column of E is needed
    for (int i = tid; i < m; i += PP_THREADS) {
      zd v = zscale(gc[mm + row_of[i]], 2.0);
      if (i == 0) v.x += 1.0;
      zst(cvec_out + i, v);
      if (i == m - 1) diag[0] = sqrt(zabs2(v));
    }
    __syncthreads();
    return;
  }
  zd *E = Wt, *F = gc == G0 ? G1 : G0;   // free: Wt and the left half of the other Gauss-Jordan buffer
  for (int e = tid; e < mm; e += PP_THREADS) {
    const int i = e % m, j = e / m;
    zd v = zscale(gc[mm + row_of[i] + j * m], 2.0);
    if (i == j) v.x += 1.0;
    E[e] = v;
  }
  __syncthreads();
  for (int s = 0; s + 1 < sc; s++) {   // all squarings but the last, ping-pong
    pade_mm(m, E, E, F, 0.0, false);
    zd *t = E; E = F; F = t;
  }
  // the last squaring: only its first column, E (E e1)
  for (int i = tid; i < m; i += PP_THREADS) {
    zd acc = zmk(0.0, 0.0);
    for (int k = 0; k < m; k++) zfma(acc, E[i + k * m], E[k]);
    zst(cvec_out + i, acc);
    if (i == m - 1) diag[0] = sqrt(zabs2(acc));
  }
  __syncthreads();
}


// ------------------------------------------------------------------ eigenvector exponential
// pAMP::EigenvectorExponential (pamp.h:567-607, krylov_exponentiation_method
// = 1): exp(A) e1 = R exp(Lambda) R^H e1 with A = -i dt H and R its unit
// eigenvectors.  pAMP's H is the Arnoldi matrix of a Hermitian operator,
// Hermitian up to the rounding of the Arnoldi inner products, so A = -i dt H
// is normal and R is unitary: the eigenvectors and eigenvalues come from the
// Hermitian part (H + H^H)/2 by the cyclic Jacobi method (the order-m matrix
// padded to even order n2; per sweep n2-1 rounds of n2/2 disjoint rotations
// in round-robin pairing, applied as one column pass over A and R and one row
// pass over A; sweeps until the off-diagonal norm is below n2 eps of the
// Frobenius norm -- 6 to 8 sweeps at m = 12..40 -- at most 20), then c_k = sum_n R_kn exp(-i dt lambda_n)
// conj(R_0n).  Column-major n2 x n2 matrices in the dynamic LDS at double
// offset `off` (2 n2^2 complex entries and the rotation table: less than
// pade_block's region).
static __device__ double eig_block_sum(double v, double *red)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < PP_THREADS / 64; w++) t += red[w];
  return t;
}

static __device__ __attribute__((noinline)) void eig_block(int m, const zd *Hg, double dt_re, double dt_im, int off,
                                                    zd *cvec_out, double *diag)
{
  extern __shared__ double pp_lds[];
  const int tid = threadIdx.x;
  const int n2 = m + (m & 1);
  const int np = n2 / 2;
  zd *A = (zd *)(pp_lds + off);
  zd *R = A + n2 * n2;
  zd *ph = R + n2 * n2;                       // e^{i phi} of each rotation
  double *rc = (double *)(ph + np), *rs = rc + np;
  int *rp = (int *)(rs + np), *rq = rp + np;
  double *red = (double *)(rq + np);          // rp, rq: 2 np ints = np doubles, so still 8-byte aligned
  for (int e = tid; e < n2 * n2; e += PP_THREADS) {
    const int i = e % n2, j = e / n2;
    zd v = zmk(0.0, 0.0);
    if (i < m && j < m) {   // (H + H^H)/2
      const zd a = Hg[i + j * m], b = Hg[j + i * m];
      v = zmk(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    }
    A[e] = v;
    R[e] = zmk(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  double fro = 0.0;
  for (int e = tid; e < n2 * n2; e += PP_THREADS) fro += zabs2(A[e]);
  fro = eig_block_sum(fro, red);
  const double tol = n2 * 2.220446049250313e-16;
  const double stop = fro * tol * tol;         // off-diagonal norm <= n2 eps of the Frobenius norm
  for (int sweep = 0; sweep < 20; sweep++) {
    double offd = 0.0;
    for (int e = tid; e < n2 * n2; e += PP_THREADS)
      if (e % n2 != e / n2) offd += zabs2(A[e]);
    offd = eig_block_sum(offd, red);
    if (!(offd > stop)) break;
    for (int r = 0; r < n2 - 1; r++) {
      if (tid < np) {   // pair tid of round r: the circle method with index 0 fixed
        const int a = tid == 0 ? 0 : 1 + (tid - 1 + r) % (n2 - 1);
        const int b = 1 + (n2 - 2 - tid + r) % (n2 - 1);
        const int p = a < b ? a : b, q = a < b ? b : a;
        const zd apq = A[p + q * n2];
        const double g = sqrt(zabs2(apq));
        double c = 1.0, sn = 0.0;
        zd e = zmk(1.0, 0.0);
        if (q < m && g > 1e-300) {
          // J = diag(1, e^{-i phi}) [[c, s], [-s, c]] zeroes A_pq of the
          // Hermitian 2 x 2 block [[a_pp, |a_pq| e^{i phi}], [c.c., a_qq]]
          e = zmk(apq.x / g, apq.y / g);
          const double tau = (A[q + q * n2].x - A[p + p * n2].x) / (2.0 * g);
          const double t = fabs(tau) > 1e150 ? 0.5 / tau
                                             : (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          sn = t * c;
        }
        rp[tid] = p; rq[tid] = q; rc[tid] = c; rs[tid] = sn; ph[tid] = e;
      }
      __syncthreads();
      for (int e = tid; e < np * n2; e += PP_THREADS) {   // columns: A J, R J
        const int k = e / n2, i = e % n2;
        const int p = rp[k], q = rq[k];
        const double c = rc[k], sn = rs[k];
        const zd em = zconj(ph[k]);
        zd x = A[i + p * n2], y = zmul(A[i + q * n2], em);
        A[i + p * n2] = zsub(zscale(x, c), zscale(y, sn));
        A[i + q * n2] = zadd(zscale(x, sn), zscale(y, c));
        x = R[i + p * n2]; y = zmul(R[i + q * n2], em);
        R[i + p * n2] = zsub(zscale(x, c), zscale(y, sn));
        R[i + q * n2] = zadd(zscale(x, sn), zscale(y, c));
      }
      __syncthreads();
      for (int e = tid; e < np * n2; e += PP_THREADS) {   // rows: J^H (A J)
        const int k = e / n2, j = e % n2;
        const int p = rp[k], q = rq[k];
        const double c = rc[k], sn = rs[k];
        const zd x = A[p + j * n2], y = zmul(A[q + j * n2], ph[k]);
        A[p + j * n2] = zsub(zscale(x, c), zscale(y, sn));
        A[q + j * n2] = zadd(zscale(x, sn), zscale(y, c));
      }
      __syncthreads();
    }
  }
  // c_k = sum_n R_kn exp(-i dt lambda_n) conj(R_0n)
  for (int k = tid; k < m; k += PP_THREADS) {
    zd acc = zmk(0.0, 0.0);
    for (int n = 0; n < m; n++) {
      const double lam = A[n + n * n2].x;
      const double mag = exp(dt_im * lam), ang = -dt_re * lam;
      const zd ex = zmk(mag * cos(ang), mag * sin(ang));
      zfma(acc, R[k + n * n2], zmul(ex, zconj(R[n * n2])));
    }
    zst(cvec_out + k, acc);
    if (k == m - 1) diag[0] = sqrt(zabs2(acc));
  }
  __syncthreads();
}

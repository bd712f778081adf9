// Shared device/host helpers for the pyprop_amd HIP library (gfx950).
// Complex amplitudes are stored interleaved (re, im) exactly like
// std::complex<double> / numpy complex128, so device buffers are
// layout-compatible with the reference's blitz::Array<cplx, R> data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <vector>

#define PP_OK 0
#define PP_EARG (-1)       // bad argument (shape, storage id, null pointer)
#define PP_EUNSUP (-2)     // combination not implemented on this path
#define PP_EHIP (-3)       // HIP runtime error
#define PP_ENUM 1          // numerical failure (Cholesky info, singular Pade denominator)

#define PP_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return PP_EHIP;              \
  } while (0)

typedef double2 zd;

__host__ __device__ inline zd zmk(double a, double b) { return make_double2(a, b); }
__host__ __device__ inline zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd zmul(zd a, zd b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline zd zconj(zd a) { return zmk(a.x, -a.y); }
__host__ __device__ inline zd zscale(zd a, double s) { return zmk(a.x * s, a.y * s); }
// a += b * c
__host__ __device__ inline void zfma(zd &a, zd b, zd c)
{
  a.x = fma(b.x, c.x, a.x);
  a.x = fma(-b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y);
  a.y = fma(b.y, c.x, a.y);
}
// a += conj(b) * c
__host__ __device__ inline void zfmac(zd &a, zd b, zd c)
{
  a.x = fma(b.x, c.x, a.x);
  a.x = fma(b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y);
  a.y = fma(-b.y, c.x, a.y);
}
// a += s * c (s real)
__host__ __device__ inline void zfmar(zd &a, double s, zd c)
{
  a.x = fma(s, c.x, a.x);
  a.y = fma(s, c.y, a.y);
}
__host__ __device__ inline double zabs2(zd a) { return a.x * a.x + a.y * a.y; }

// Storage ids of the reference generator (tensorpotentialmultiply_generator.py:1336).
enum pp_storage { PP_IDENT = 0, PP_DIAG = 1, PP_SIMP = 2, PP_HERM = 3, PP_BANDNH = 4, PP_BAND = 5 };

// Chunked parallel solver for one real upper-banded triangular factor U
// (kd super-diagonals), both for U x = b (backward) and U^T x = b (forward,
// = U^H x = b since U is real).  Rows are padded to P*B, B a compile-time
// chunk length, one wavefront lane per chunk (P <= 32).  Each lane solves
// its chunk locally; the chunk couplings form a kd-dimensional affine
// recurrence (suffix for U, prefix for U^T) solved by a log-depth scan whose
// data-independent matrix products are precomputed on the host; a second
// local pass with the coupling folded into the right-hand side gives x.
// Padded rows are identity rows with zero right-hand side.
//
// Packed plan (doubles), read by lane c at [..]*P + c (conflict-free LDS):
//   Ui    [(r*(kd+1)+d)*P + c] : d=0: 1/U(i,i); d>0: U(i, i+d) (also across chunks)
//   NlevN [((l*kd+i)*kd+k)*P + c] suffix-scan matrices of U x = b
//   NlevC [((l*kd+i)*kd+k)*P + c] prefix-scan matrices of U^T x = b
struct pp_band_plan {
  int N, kd, B, P, levels;
  int ndoubles;      // size of the packed plan
  double *packed;    // device copy
};

// Overlap factor on the device: S = U^H U for one B-spline rank.
struct pp_overlap {
  int N, kd;
  double *Uband;       // device, [j*(kd+1) + kd+i-j] = U(i,j), real
  pp_band_plan plan;
};

// U x = b for one fiber; x (LDS, >= P*B complex, zero beyond N) holds b on
// entry and x on exit.  Executed by the 64 lanes of the calling wavefront
// (lanes >= P idle).  Each lane keeps a sliding window of kd unknowns in
// registers; the rows touching a chunk boundary are peeled off so the
// steady-state loop is branch-free.
struct pp_win_ops {
  template <int KD>
  __device__ static inline void shift_in(zd (&win)[KD], zd v)
  {
#pragma unroll
    for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
    win[0] = v;
  }
};

// one backward row: acc = b - sum_{d valid} U(r, r+d) win[d-1] - coupling; returns acc / U(r,r)
template <int KD, int B, bool EDGE>
__device__ inline zd pp_back_row(const double *Ui, int P, int c, int r, zd b, const zd (&win)[KD], const zd *t)
{
  zd acc = b;
  if (EDGE && t) {
#pragma unroll
    for (int k = 0; k < KD; k++) {
      const int d = B + k - r;
      if (d >= 1 && d <= KD) { const double u = Ui[(r * (KD + 1) + d) * P + c]; acc.x = fma(-u, t[k].x, acc.x); acc.y = fma(-u, t[k].y, acc.y); }
    }
  }
#pragma unroll
  for (int d = KD; d >= 1; d--)
    if (!EDGE || r + d < B) { const double u = Ui[(r * (KD + 1) + d) * P + c]; acc.x = fma(-u, win[d - 1].x, acc.x); acc.y = fma(-u, win[d - 1].y, acc.y); }
  return zscale(acc, Ui[(r * (KD + 1)) * P + c]);
}

template <int KD, int B>
__device__ inline void pp_back_pass(const double *Ui, int P, int c, zd *x, zd (&win)[KD], const zd *t, bool write)
{
#pragma unroll
  for (int k = 0; k < KD; k++) win[k] = zmk(0.0, 0.0);
  // peeled rows B-1 .. B-KD (window partly empty, coupling to the next chunk)
#pragma unroll
  for (int r = B - 1; r >= B - KD; r--) {
    zd v = pp_back_row<KD, B, true>(Ui, P, c, r, x[c * B + r], win, t);
    pp_win_ops::shift_in<KD>(win, v);
    if (write) x[c * B + r] = v;
  }
#pragma unroll 4
  for (int r = B - KD - 1; r >= 0; r--) {
    zd v = pp_back_row<KD, B, false>(Ui, P, c, r, x[c * B + r], win, t);
    pp_win_ops::shift_in<KD>(win, v);
    if (write) x[c * B + r] = v;
  }
}

template <int KD, int B>
__device__ inline void pp_solve_upper(const double *pk, int P, int levels, zd *x, int N)
{
  const int c = threadIdx.x & 63;
  if (c >= P) return;
  const double *Ui = pk;
  const double *Nl = pk + B * (KD + 1) * P;
  zd win[KD];
  pp_back_pass<KD, B>(Ui, P, c, x, win, nullptr, false);   // window ends on rows 0..kd-1
  zd a[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) a[k] = win[k];                // row 0 went in last: win[k] = y of row k
  for (int l = 0; l < levels; l++) {
    const int s = 1 << l;
    zd an[KD];
#pragma unroll
    for (int k = 0; k < KD; k++) { an[k].x = __shfl_down(a[k].x, s, 64); an[k].y = __shfl_down(a[k].y, s, 64); }
    if (c + s < P) {
#pragma unroll
      for (int i = 0; i < KD; i++) {
        zd acc = a[i];
#pragma unroll
        for (int k = 0; k < KD; k++) { const double nv = Nl[((l * KD + i) * KD + k) * P + c]; acc.x = fma(nv, an[k].x, acc.x); acc.y = fma(nv, an[k].y, acc.y); }
        a[i] = acc;
      }
    }
  }
  zd t[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) {
    t[k].x = __shfl_down(a[k].x, 1, 64); t[k].y = __shfl_down(a[k].y, 1, 64);
    if (c + 1 >= P) t[k] = zmk(0.0, 0.0);
  }
  pp_back_pass<KD, B>(Ui, P, c, x, win, t, true);
  (void)N;
}

// one forward row of U^T x = b: acc = b - sum_{d valid} U(r-d, r) win[d-1] - coupling
template <int KD, int B, bool EDGE>
__device__ inline zd pp_fwd_row(const double *Ui, int P, int c, int r, zd b, const zd (&win)[KD], const zd *t)
{
  zd acc = b;
  if (EDGE && t && c > 0) {
#pragma unroll
    for (int k = 0; k < KD; k++) {
      if (k >= r) {
        const int rr = B - KD + k, d = r + KD - k;
        const double u = Ui[(rr * (KD + 1) + d) * P + c - 1];
        acc.x = fma(-u, t[k].x, acc.x); acc.y = fma(-u, t[k].y, acc.y);
      }
    }
  }
#pragma unroll
  for (int d = KD; d >= 1; d--)
    if (!EDGE || r - d >= 0) { const double u = Ui[((r - d) * (KD + 1) + d) * P + c]; acc.x = fma(-u, win[d - 1].x, acc.x); acc.y = fma(-u, win[d - 1].y, acc.y); }
  return zscale(acc, Ui[(r * (KD + 1)) * P + c]);
}

template <int KD, int B>
__device__ inline void pp_fwd_pass(const double *Ui, int P, int c, zd *x, zd (&win)[KD], const zd *t, bool write)
{
#pragma unroll
  for (int k = 0; k < KD; k++) win[k] = zmk(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < KD; r++) {
    zd v = pp_fwd_row<KD, B, true>(Ui, P, c, r, x[c * B + r], win, t);
    pp_win_ops::shift_in<KD>(win, v);
    if (write) x[c * B + r] = v;
  }
#pragma unroll 4
  for (int r = KD; r < B; r++) {
    zd v = pp_fwd_row<KD, B, false>(Ui, P, c, r, x[c * B + r], win, t);
    pp_win_ops::shift_in<KD>(win, v);
    if (write) x[c * B + r] = v;
  }
}

// U^T x = b (forward substitution) with the same packed plan.
template <int KD, int B>
__device__ inline void pp_solve_lowerT(const double *pk, int P, int levels, zd *x, int N)
{
  const int c = threadIdx.x & 63;
  if (c >= P) return;
  const double *Ui = pk;
  const double *Nl = pk + B * (KD + 1) * P + levels * KD * KD * P;
  zd win[KD];
  pp_fwd_pass<KD, B>(Ui, P, c, x, win, nullptr, false);    // window ends on rows B-1 .. B-kd
  zd a[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) a[k] = win[KD - 1 - k];      // z of row B-kd+k
  for (int l = 0; l < levels; l++) {
    const int s = 1 << l;
    zd an[KD];
#pragma unroll
    for (int k = 0; k < KD; k++) { an[k].x = __shfl_up(a[k].x, s, 64); an[k].y = __shfl_up(a[k].y, s, 64); }
    if (c - s >= 0) {
#pragma unroll
      for (int i = 0; i < KD; i++) {
        zd acc = a[i];
#pragma unroll
        for (int k = 0; k < KD; k++) { const double nv = Nl[((l * KD + i) * KD + k) * P + c]; acc.x = fma(nv, an[k].x, acc.x); acc.y = fma(nv, an[k].y, acc.y); }
        a[i] = acc;
      }
    }
  }
  zd t[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) {
    t[k].x = __shfl_up(a[k].x, 1, 64); t[k].y = __shfl_up(a[k].y, 1, 64);
    if (c == 0) t[k] = zmk(0.0, 0.0);
  }
  pp_fwd_pass<KD, B>(Ui, P, c, x, win, t, true);
  (void)N;
}

// Chunk length for N rows: the smallest of 8/16/32/64 giving P <= 16 chunks
// (B >= kd); 16 chunks keep the packed plan small enough to live in LDS next
// to the Krylov buffers of a fiber of up to 1024 B-splines.
inline int pp_plan_chunk(int N, int kd)
{
  for (int B : {8, 16, 32, 64})
    if (B >= kd && (N + B - 1) / B <= 16) return B;
  return 0;
}

#define PP_B_SWITCH(Bv, ...) \
  switch (Bv) {              \
    case 8: { constexpr int B_ = 8; __VA_ARGS__; } break; \
    case 16: { constexpr int B_ = 16; __VA_ARGS__; } break; \
    case 32: { constexpr int B_ = 32; __VA_ARGS__; } break; \
    case 64: { constexpr int B_ = 64; __VA_ARGS__; } break; \
    default: break; }

// Dispatch helper over the supported super-diagonal counts (B-spline order-1).
#define PP_KD_SWITCH(kd, ...) \
  switch (kd) {                      \
    case 4: { constexpr int KD_ = 4; __VA_ARGS__; } break; \
    case 5: { constexpr int KD_ = 5; __VA_ARGS__; } break; \
    case 6: { constexpr int KD_ = 6; __VA_ARGS__; } break; \
    case 8: { constexpr int KD_ = 8; __VA_ARGS__; } break; \
    default: break; }

int pp_plan_build(pp_band_plan *pl, int N, int kd, const double *Uhost, int B);
void pp_plan_free(pp_band_plan *pl);

// Per-destination-row entry lists for one rank's storage id (host side).
struct pp_rank_csr {
  std::vector<int> rowptr;   // extent+1
  std::vector<int> ent;      // 3 ints per entry: source col, potential index, conjugate flag
};
int pp_build_csr(int storage, int potExt, int srcExt, const int *pairs, pp_rank_csr &csr);

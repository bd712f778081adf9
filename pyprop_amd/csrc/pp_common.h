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

// Chunked parallel solver for one real upper-banded triangular factor
// (kd super-diagonals).  Rows are split into P <= 64 chunks of B rows,
// one wavefront lane per chunk; the chunk couplings form a kd-dimensional
// affine recurrence solved by a log-depth suffix scan whose (data
// independent) matrices are precomputed here.  Padded rows (N..P*B-1) are
// identity rows with zero right-hand side.
struct pp_band_plan {
  int N, kd, B, P, levels;
  double *ucoef;  // [(r*(kd+1)+d)*P + c]   d=0: 1/U(i,i); d>0: U(i,i+d) if r+d<B
  double *G;      // [(r*kd+k)*P + c]       G_c = U_cc^-1 U_{c,c+1}[:, 0:kd]
  double *Nlev;   // [((l*kd+i)*kd+k)*P + c] scan matrices per level
};

// Overlap factor on the device: S = U^H U for one B-spline rank.
struct pp_overlap {
  int N, kd;
  double *Uband;       // device, [j*(kd+1) + kd+i-j] = U(i,j), real
  pp_band_plan solveN; // U x = b
  pp_band_plan solveC; // U^H x = b (solved as a reversed upper system)
};

// Device-side scan/solve for one fiber held in LDS, executed by wavefront 0.
// x_lds (length >= P*B, complex) holds b on entry and x on exit.
template <int KD, int MAXB>
__device__ inline void pp_band_solve_wave(const pp_band_plan &pl, zd *x_lds, bool reversed, int N)
{
  const int lane = threadIdx.x;  // caller guarantees threadIdx.x < 64
  const int P = pl.P, B = pl.B;
  const int c = lane;
  const bool active = c < P;
  zd y[MAXB];
  // local backward solve of chunk c
#pragma unroll
  for (int r = MAXB - 1; r >= 0; r--) {
    if (r < B) {
      int row = c * B + r;
      zd b = zmk(0.0, 0.0);
      if (active && row < N) b = x_lds[reversed ? (N - 1 - row) : row];
      zd acc = b;
#pragma unroll
      for (int d = 1; d <= KD; d++) {
        if (r + d < MAXB && r + d < B) {
          double u = active ? pl.ucoef[(r * (KD + 1) + d) * P + c] : 0.0;
          acc.x = fma(-u, y[r + d].x, acc.x);
          acc.y = fma(-u, y[r + d].y, acc.y);
        }
      }
      double inv = active ? pl.ucoef[(r * (KD + 1)) * P + c] : 0.0;
      y[r] = zscale(acc, inv);
    } else {
      y[r] = zmk(0.0, 0.0);
    }
  }
  // suffix scan of t_c = a_c - M_c t_{c+1}
  zd a[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) a[k] = y[k];
  for (int l = 0; l < pl.levels; l++) {
    int s = 1 << l;
    zd an[KD];
#pragma unroll
    for (int k = 0; k < KD; k++) {
      an[k].x = __shfl_down(a[k].x, s, 64);
      an[k].y = __shfl_down(a[k].y, s, 64);
    }
    if (active && c + s < P) {
#pragma unroll
      for (int i = 0; i < KD; i++) {
        zd acc = a[i];
#pragma unroll
        for (int k = 0; k < KD; k++) {
          double nv = pl.Nlev[((l * KD + i) * KD + k) * P + c];
          acc.x = fma(nv, an[k].x, acc.x);
          acc.y = fma(nv, an[k].y, acc.y);
        }
        a[i] = acc;
      }
    }
  }
  // t_{c+1} = a_{c+1}
  zd t[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) {
    t[k].x = __shfl_down(a[k].x, 1, 64);
    t[k].y = __shfl_down(a[k].y, 1, 64);
    if (c + 1 >= P) t[k] = zmk(0.0, 0.0);
  }
  // fix-up x_c = y_c - G_c t_{c+1}
#pragma unroll
  for (int r = 0; r < MAXB; r++) {
    if (r < B) {
      int row = c * B + r;
      zd acc = y[r];
#pragma unroll
      for (int k = 0; k < KD; k++) {
        double g = active ? pl.G[(r * KD + k) * P + c] : 0.0;
        acc.x = fma(-g, t[k].x, acc.x);
        acc.y = fma(-g, t[k].y, acc.y);
      }
      if (active && row < N) x_lds[reversed ? (N - 1 - row) : row] = acc;
    }
  }
}

// Dispatch helper over the supported super-diagonal counts (B-spline order-1).
#define PP_KD_SWITCH(kd, ...) \
  switch (kd) {                      \
    case 4: { constexpr int KD_ = 4; __VA_ARGS__; } break; \
    case 5: { constexpr int KD_ = 5; __VA_ARGS__; } break; \
    case 6: { constexpr int KD_ = 6; __VA_ARGS__; } break; \
    case 8: { constexpr int KD_ = 8; __VA_ARGS__; } break; \
    default: break; }

int pp_plan_build(pp_band_plan *pl, int N, int kd, const double *Uhost, bool reversed);
void pp_plan_free(pp_band_plan *pl);

// Per-destination-row entry lists for one rank's storage id (host side).
struct pp_rank_csr {
  std::vector<int> rowptr;   // extent+1
  std::vector<int> ent;      // 3 ints per entry: source col, potential index, conjugate flag
};
int pp_build_csr(int storage, int potExt, int srcExt, const int *pairs, pp_rank_csr &csr);

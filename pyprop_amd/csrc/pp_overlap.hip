// Overlap-matrix operations for one B-spline rank on MI355X.
//
// Reference: core/representation/overlapmatrix.cpp:9-37 (upper band storage,
// zpbtrf), :137-149 (ztbmv / ztbsv per fiber), :207-229 (tensor fibers via
// MapToRank3).  The Cholesky factor of a B-spline overlap matrix is real, so
// the device keeps it as doubles; solves run one wavefront per fiber with the
// chunked scan of pp_band_solve_wave (see pp_common.h).
#include "pp_common.h"
#include <vector>
#include <cmath>
#include <cstring>
#include <cstdlib>

// ---------------------------------------------------------------- host: zpbtrf 'U'
// S = U^H U in place, upper band storage ab[j*(kd+1) + kd+i-j]; complex
// interleaved.  Returns LAPACK-style info.
extern "C" int pp_band_cholesky_upper(int N, int kd, double *ab)
{
  if (N < 1 || kd < 0 || !ab) return PP_EARG;
  auto re = [&](int i, int j) -> double & { return ab[2 * ((size_t)j * (kd + 1) + kd + i - j)]; };
  auto im = [&](int i, int j) -> double & { return ab[2 * ((size_t)j * (kd + 1) + kd + i - j) + 1]; };
  for (int j = 0; j < N; j++) {
    double d = re(j, j);
    for (int i = (j - kd < 0 ? 0 : j - kd); i < j; i++) d -= re(i, j) * re(i, j) + im(i, j) * im(i, j);
    if (d <= 0.0) return j + 1;
    d = std::sqrt(d);
    re(j, j) = d;
    im(j, j) = 0.0;
    for (int c = j + 1; c < N && c <= j + kd; c++) {
      double sr = re(j, c), si = im(j, c);
      for (int i = (c - kd < 0 ? 0 : c - kd); i < j; i++) {
        // s -= conj(U(i,j)) * U(i,c)
        double ar = re(i, j), ai = -im(i, j), br = re(i, c), bi = im(i, c);
        sr -= ar * br - ai * bi;
        si -= ar * bi + ai * br;
      }
      re(j, c) = sr / d;
      im(j, c) = si / d;
    }
  }
  return 0;
}

// ---------------------------------------------------------------- host: plan
int pp_plan_build(pp_band_plan *pl, int N, int kd, const double *U, bool reversed)
{
  // U: real band, [j*(kd+1) + kd+i-j]
  auto getU = [&](int i, int j) -> double {
    if (i < 0 || j < i || j - i > kd) return 0.0;
    if (i >= N) return (i == j) ? 1.0 : 0.0;  // padded identity rows
    if (j >= N) return 0.0;
    int ii = i, jj = j;
    if (reversed) { ii = N - 1 - j; jj = N - 1 - i; }
    return U[(size_t)jj * (kd + 1) + kd + ii - jj];
  };
  int B = (N + 63) / 64;
  if (B < kd) B = kd;
  if (B < 1) B = 1;
  int P = (N + B - 1) / B;
  int levels = 0;
  while ((1 << levels) < P) levels++;
  pl->N = N; pl->kd = kd; pl->B = B; pl->P = P; pl->levels = levels;
  std::vector<double> uc((size_t)B * (kd + 1) * P, 0.0), G((size_t)B * kd * P, 0.0),
      Nl((size_t)(levels > 0 ? levels : 1) * kd * kd * P, 0.0);
  for (int c = 0; c < P; c++)
    for (int r = 0; r < B; r++) {
      int row = c * B + r;
      uc[(size_t)(r * (kd + 1)) * P + c] = 1.0 / getU(row, row);
      for (int d = 1; d <= kd; d++)
        if (r + d < B) uc[(size_t)(r * (kd + 1) + d) * P + c] = getU(row, row + d);
    }
  // G_c = U_cc^-1 C_c, C_c[r][k] = U(cB+r, (c+1)B+k)
  std::vector<double> col(B);
  std::vector<std::vector<double>> M((size_t)P, std::vector<double>((size_t)kd * kd, 0.0));
  for (int c = 0; c + 1 < P; c++) {
    for (int k = 0; k < kd; k++) {
      for (int r = 0; r < B; r++) col[r] = getU(c * B + r, (c + 1) * B + k);
      for (int r = B - 1; r >= 0; r--) {
        double s = col[r];
        for (int d = 1; d <= kd && r + d < B; d++) s -= getU(c * B + r, c * B + r + d) * col[r + d];
        col[r] = s / getU(c * B + r, c * B + r);
      }
      for (int r = 0; r < B; r++) G[(size_t)(r * kd + k) * P + c] = col[r];
      for (int i = 0; i < kd; i++) M[c][(size_t)i * kd + k] = col[i];
    }
  }
  // level matrices: N0_c = -M_c ; N_{l+1,c} = N_{l,c} N_{l,c+2^l}
  std::vector<std::vector<double>> cur((size_t)P, std::vector<double>((size_t)kd * kd));
  for (int c = 0; c < P; c++)
    for (int e = 0; e < kd * kd; e++) cur[c][e] = -M[c][e];
  for (int l = 0; l < levels; l++) {
    for (int c = 0; c < P; c++)
      for (int i = 0; i < kd; i++)
        for (int k = 0; k < kd; k++) Nl[(size_t)((l * kd + i) * kd + k) * P + c] = cur[c][(size_t)i * kd + k];
    int s = 1 << l;
    std::vector<std::vector<double>> nxt = cur;
    for (int c = 0; c + s < P; c++)
      for (int i = 0; i < kd; i++)
        for (int k = 0; k < kd; k++) {
          double a = 0.0;
          for (int q = 0; q < kd; q++) a += cur[c][(size_t)i * kd + q] * cur[c + s][(size_t)q * kd + k];
          nxt[c][(size_t)i * kd + k] = a;
        }
    cur.swap(nxt);
  }
  PP_HIP(hipMalloc(&pl->ucoef, uc.size() * sizeof(double)));
  PP_HIP(hipMalloc(&pl->G, G.size() * sizeof(double)));
  PP_HIP(hipMalloc(&pl->Nlev, Nl.size() * sizeof(double)));
  PP_HIP(hipMemcpy(pl->ucoef, uc.data(), uc.size() * sizeof(double), hipMemcpyHostToDevice));
  PP_HIP(hipMemcpy(pl->G, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
  PP_HIP(hipMemcpy(pl->Nlev, Nl.data(), Nl.size() * sizeof(double), hipMemcpyHostToDevice));
  return PP_OK;
}

void pp_plan_free(pp_band_plan *pl)
{
  if (pl->ucoef) (void)hipFree(pl->ucoef);
  if (pl->G) (void)hipFree(pl->G);
  if (pl->Nlev) (void)hipFree(pl->Nlev);
  pl->ucoef = pl->G = pl->Nlev = nullptr;
}

// ---------------------------------------------------------------- overlap object
// Takes S in upper hermitian band storage (host, complex interleaved, N*(kd+1)),
// factorises it (overlapmatrix.cpp:35-37) and uploads U plus both solve plans.
extern "C" int pp_overlap_create(int N, int kd, const double *S_upper, pp_overlap **out, int *info)
{
  if (!out || !S_upper || N < 1 || kd < 1) return PP_EARG;
  if (!(kd == 4 || kd == 5 || kd == 6 || kd == 8) || N > 64 * 32) return PP_EUNSUP;
  std::vector<double> ab(S_upper, S_upper + (size_t)2 * N * (kd + 1));
  int inf = pp_band_cholesky_upper(N, kd, ab.data());
  if (info) *info = inf;
  if (inf) return PP_ENUM;
  std::vector<double> U((size_t)N * (kd + 1));
  for (size_t e = 0; e < U.size(); e++) {
    if (ab[2 * e + 1] != 0.0) return PP_EUNSUP;  // complex overlap factors are not on this path
    U[e] = ab[2 * e];
  }
  pp_overlap *ov = new pp_overlap();
  memset(ov, 0, sizeof(*ov));
  ov->N = N; ov->kd = kd;
  int rc;
  if (hipMalloc(&ov->Uband, U.size() * sizeof(double)) != hipSuccess) { delete ov; return PP_EHIP; }
  if (hipMemcpy(ov->Uband, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { delete ov; return PP_EHIP; }
  if ((rc = pp_plan_build(&ov->solveN, N, kd, U.data(), false)) != PP_OK) return rc;
  if ((rc = pp_plan_build(&ov->solveC, N, kd, U.data(), true)) != PP_OK) return rc;
  *out = ov;
  return PP_OK;
}

extern "C" int pp_overlap_factor(const pp_overlap *ov, double *U_upper_host)
{
  if (!ov || !U_upper_host) return PP_EARG;
  PP_HIP(hipMemcpy(U_upper_host, ov->Uband, sizeof(double) * ov->N * (ov->kd + 1), hipMemcpyDeviceToHost));
  return PP_OK;
}

extern "C" int pp_overlap_destroy(pp_overlap *ov)
{
  if (!ov) return PP_OK;
  if (ov->Uband) (void)hipFree(ov->Uband);
  pp_plan_free(&ov->solveN);
  pp_plan_free(&ov->solveC);
  delete ov;
  return PP_OK;
}

// ---------------------------------------------------------------- device ops
enum { PP_OV_SOLVE_SQRT = 0, PP_OV_SOLVE_SQRT_CONJ = 1, PP_OV_MUL_SQRT = 2, PP_OV_MUL_SQRT_CONJ = 3,
       PP_OV_SOLVE = 4, PP_OV_MUL = 5 };

// x := U x (conj=0) or U^H x (conj=1) for a fiber in LDS (in place, block-wide).
__device__ inline void pp_band_mul_block(const double *Ub, int N, int kd, zd *x, zd *tmp, bool conj)
{
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    zd acc = zmk(0.0, 0.0);
    if (!conj) {
      int j1 = i + kd < N - 1 ? i + kd : N - 1;
      for (int j = i; j <= j1; j++) zfmar(acc, Ub[(size_t)j * (kd + 1) + kd + i - j], x[j]);
    } else {
      int j0 = i - kd < 0 ? 0 : i - kd;
      for (int j = j0; j <= i; j++) zfmar(acc, Ub[(size_t)i * (kd + 1) + kd + j - i], x[j]);
    }
    tmp[i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] = tmp[i];
  __syncthreads();
}

template <int KD, int MAXB>
__global__ void __launch_bounds__(256) pp_overlap_kernel(pp_overlap ov, int op, zd *x, long post, long nfib)
{
  extern __shared__ zd lds[];
  const int N = ov.N;
  zd *xs = lds;
  zd *tmp = lds + ov.solveN.P * ov.solveN.B;
  long f = blockIdx.x;
  if (f >= nfib) return;
  long pre = f / post, pst = f % post;
  zd *base = x + pre * (long)N * post + pst;
  for (int i = threadIdx.x; i < N; i += blockDim.x) xs[i] = base[(long)i * post];
  __syncthreads();
  auto solve = [&](bool conj) {
    if (threadIdx.x < 64) pp_band_solve_wave<KD, MAXB>(conj ? ov.solveC : ov.solveN, xs, conj, N);
    __syncthreads();
  };
  switch (op) {
    case PP_OV_SOLVE_SQRT: solve(false); break;
    case PP_OV_SOLVE_SQRT_CONJ: solve(true); break;
    case PP_OV_MUL_SQRT: pp_band_mul_block(ov.Uband, N, ov.kd, xs, tmp, false); break;
    case PP_OV_MUL_SQRT_CONJ: pp_band_mul_block(ov.Uband, N, ov.kd, xs, tmp, true); break;
    case PP_OV_SOLVE: solve(true); solve(false); break;     // S^-1 = U^-1 U^-H (overlapmatrix.cpp:131-135)
    case PP_OV_MUL:                                         // S = U^H U
      pp_band_mul_block(ov.Uband, N, ov.kd, xs, tmp, false);
      pp_band_mul_block(ov.Uband, N, ov.kd, xs, tmp, true);
      break;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) base[(long)i * post] = xs[i];
}

// Apply `op` along rank `rank` of a C-ordered complex tensor x (device).
extern "C" int pp_overlap_apply(const pp_overlap *ov, int op, int rank, int R, const int *shape, void *x,
                                hipStream_t stream)
{
  if (!ov || !shape || !x || rank < 0 || rank >= R || R < 1 || R > 8) return PP_EARG;
  if (shape[rank] != ov->N) return PP_EARG;
  if (op < 0 || op > 5) return PP_EARG;
  long pre = 1, post = 1;
  for (int i = 0; i < rank; i++) pre *= shape[i];
  for (int i = rank + 1; i < R; i++) post *= shape[i];
  long nfib = pre * post;
  if (nfib == 0) return PP_OK;
  size_t lds = sizeof(zd) * (size_t)(ov->solveN.P * ov->solveN.B + ov->N);
  int maxb = ov->solveN.B > ov->solveC.B ? ov->solveN.B : ov->solveC.B;
  bool launched = false;
  if (maxb <= 16) {
    PP_KD_SWITCH(ov->kd, {
      hipLaunchKernelGGL((pp_overlap_kernel<KD_, 16>), dim3(nfib), dim3(256), lds, stream, *ov, op, (zd *)x, post, nfib);
      launched = true; })
  } else {
    PP_KD_SWITCH(ov->kd, {
      hipLaunchKernelGGL((pp_overlap_kernel<KD_, 32>), dim3(nfib), dim3(256), lds, stream, *ov, op, (zd *)x, post, nfib);
      launched = true; })
  }
  if (!launched) return PP_EUNSUP;
  PP_HIP(hipGetLastError());
  return PP_OK;
}

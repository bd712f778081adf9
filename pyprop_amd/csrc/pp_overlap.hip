// Overlap-matrix operations for one B-spline rank on MI355X.
//
// Reference: core/representation/overlapmatrix.cpp:9-37 (upper band storage,
// zpbtrf), :137-149 (ztbmv / ztbsv per fiber), :207-229 (tensor fibers via
// MapToRank3).  The Cholesky factor of a B-spline overlap matrix is real, so
// the device keeps it as doubles; solves run one workgroup per fiber with the
// chunked block-inverse solver pp_band_solve (pp_common.h).
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
// Chunk plan of U for pp_band_solve (pp_common.h): per chunk the dense
// triangular inverse U_cc^-1, the coupling columns G_c = U_cc^-1 C_c and
// G'_c = U_cc^-T C'_c, and the level matrices of both boundary scans.
int pp_plan_build(pp_band_plan *pl, int N, int kd, const double *U, int B)
{
  // U: real band, [j*(kd+1) + kd+i-j]; rows/cols >= N are identity padding
  auto getU = [&](int i, int j) -> double {
    if (i < 0 || j < i || j - i > kd) return 0.0;
    if (i >= N) return (i == j) ? 1.0 : 0.0;
    if (j >= N) return 0.0;
    return U[(size_t)j * (kd + 1) + kd + i - j];
  };
  int lb = 0;
  while ((1 << lb) < B) lb++;
  if (B < kd || B < 8 || (1 << lb) != B) return PP_EUNSUP;
  const int P = (N + B - 1) / B;
  if (P > 64) return PP_EUNSUP;
  int levels = 0;
  while ((1 << levels) < P) levels++;
  const int npad = P * B;
  pl->N = N; pl->kd = kd; pl->B = B; pl->lb = lb; pl->P = P; pl->levels = levels;
  const size_t nL = (size_t)levels * kd * kd * P, nB = (size_t)P * B * B, nG = (size_t)kd * npad;
  std::vector<double> pk(2 * nL + 2 * nB + 2 * nG, 0.0);
  double *NlN = pk.data(), *NlC = NlN + nL, *Bi = NlC + nL, *BiT = Bi + nB, *G = BiT + nB, *GT = G + nG;
  std::vector<double> col(B);
  auto back = [&](int c) {   // col := U_cc^-1 col
    for (int r = B - 1; r >= 0; r--) {
      double s = col[r];
      for (int d = 1; d <= kd && r + d < B; d++) s -= getU(c * B + r, c * B + r + d) * col[r + d];
      col[r] = s / getU(c * B + r, c * B + r);
    }
  };
  auto fwd = [&](int c) {    // col := U_cc^-T col
    for (int r = 0; r < B; r++) {
      double s = col[r];
      for (int d = 1; d <= kd && r - d >= 0; d++) s -= getU(c * B + r - d, c * B + r) * col[r - d];
      col[r] = s / getU(c * B + r, c * B + r);
    }
  };
  std::vector<std::vector<double>> MN((size_t)P, std::vector<double>((size_t)kd * kd, 0.0)), MC = MN;
  for (int c = 0; c < P; c++) {
    for (int j = 0; j < B; j++) {   // column j of U_cc^-1
      std::fill(col.begin(), col.end(), 0.0);
      col[j] = 1.0;
      back(c);
      for (int i = 0; i < B; i++) {
        Bi[((size_t)c * B + j) * B + i] = col[i];    // (i, j)
        BiT[((size_t)c * B + i) * B + j] = col[i];   // (j, i) of the transpose
      }
    }
    for (int k = 0; k < kd; k++) {
      if (c + 1 < P) {   // U x = b: C_c[r][k] = U(cB+r, (c+1)B+k)
        for (int r = 0; r < B; r++) col[r] = getU(c * B + r, (c + 1) * B + k);
        back(c);
        for (int i = 0; i < B; i++) G[(size_t)k * npad + c * B + i] = col[i];
        for (int i = 0; i < kd; i++) MN[c][(size_t)i * kd + k] = col[i];
      }
      if (c > 0) {       // U^T x = b: C'_c[r][k] = U((c-1)B+B-kd+k, cB+r)
        for (int r = 0; r < B; r++) col[r] = getU((c - 1) * B + B - kd + k, c * B + r);
        fwd(c);
        for (int i = 0; i < B; i++) GT[(size_t)k * npad + c * B + i] = col[i];
        for (int i = 0; i < kd; i++) MC[c][(size_t)i * kd + k] = col[B - kd + i];
      }
    }
  }
  // scan matrices: N0_c = -M_c; suffix N_{l+1,c} = N_{l,c} N_{l,c+2^l}; prefix uses c-2^l
  auto levelize = [&](std::vector<std::vector<double>> &M, double *out, int dir) {
    std::vector<std::vector<double>> cur = M;
    for (auto &m : cur) for (auto &v : m) v = -v;
    for (int l = 0; l < levels; l++) {
      for (int c = 0; c < P; c++)
        for (int e = 0; e < kd * kd; e++) out[(size_t)(l * kd * kd + e) * P + c] = cur[c][e];
      int s = 1 << l;
      std::vector<std::vector<double>> nxt = cur;
      for (int c = 0; c < P; c++) {
        int o = c + dir * s;
        if (o < 0 || o >= P) continue;
        for (int i = 0; i < kd; i++)
          for (int k = 0; k < kd; k++) {
            double a = 0.0;
            for (int q = 0; q < kd; q++) a += cur[c][(size_t)i * kd + q] * cur[o][(size_t)q * kd + k];
            nxt[c][(size_t)i * kd + k] = a;
          }
      }
      cur.swap(nxt);
    }
  };
  levelize(MN, NlN, +1);
  levelize(MC, NlC, -1);
  pl->nlev = (int)(2 * nL);
  PP_HIP(hipMalloc(&pl->packed, pk.size() * sizeof(double)));
  PP_HIP(hipMemcpy(pl->packed, pk.data(), pk.size() * sizeof(double), hipMemcpyHostToDevice));
  pl->nlevs = pl->packed;
  pl->binv = pl->packed + 2 * nL;
  pl->binvT = pl->binv + nB;
  pl->G = pl->binvT + nB;
  pl->GT = pl->G + nG;
  return PP_OK;
}

void pp_plan_free(pp_band_plan *pl)
{
  if (pl->packed) (void)hipFree(pl->packed);
  pl->packed = nullptr;
}

// ---------------------------------------------------------------- overlap object
// Takes S in upper hermitian band storage (host, complex interleaved, N*(kd+1)),
// factorises it (overlapmatrix.cpp:35-37) and uploads U plus both solve plans.
extern "C" int pp_overlap_create_chunked(int N, int kd, const double *S_upper, int B, pp_overlap **out, int *info);
extern "C" int pp_overlap_destroy(pp_overlap *ov);

extern "C" int pp_overlap_create(int N, int kd, const double *S_upper, pp_overlap **out, int *info)
{
  return pp_overlap_create_chunked(N, kd, S_upper, 0, out, info);
}

// Same with an explicit solver chunk length B (a power of two >= max(kd, 8),
// at most 64 chunks); B = 0 picks the default of pp_plan_chunk.  The
// autotuner times the candidates.
extern "C" int pp_overlap_create_chunked(int N, int kd, const double *S_upper, int B, pp_overlap **out, int *info)
{
  if (!out || !S_upper || N < 1 || kd < 1 || B < 0) return PP_EARG;
  if (kd < PP_KD_MIN || kd > PP_KD_MAX) return PP_EUNSUP;
  if (B == 0) B = pp_plan_chunk(N, kd);
  if (B == 0 || B < kd || B < 8 || (B & (B - 1)) || (N + B - 1) / B > 64) return PP_EUNSUP;
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
  if (hipMemcpy(ov->Uband, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { pp_overlap_destroy(ov); return PP_EHIP; }
  if ((rc = pp_plan_build(&ov->plan, N, kd, U.data(), B)) != PP_OK) { pp_overlap_destroy(ov); return rc; }
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
  pp_plan_free(&ov->plan);
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

template <int KD>
__global__ void __launch_bounds__(256) pp_overlap_kernel(pp_overlap ov, int op, zd *x, long post, long nfib)
{
  extern __shared__ double lds_d[];
  const pp_band_plan &pl = ov.plan;
  const int N = ov.N, npad = pl.P * pl.B;
  double *nl = lds_d;
  zd *xs = (zd *)(lds_d + ((pl.nlev + 1) & ~1));   // npad, zero beyond N
  zd *xp = xs + npad;                               // npad
  zd *tb = xp + npad;                               // P*KD
  long f = blockIdx.x;
  if (f >= nfib) return;
  for (int i = threadIdx.x; i < pl.nlev; i += blockDim.x) nl[i] = pl.nlevs[i];
  long pre = f / post, pst = f % post;
  zd *base = x + pre * (long)N * post + pst;
  for (int i = threadIdx.x; i < npad; i += blockDim.x) xs[i] = i < N ? base[(long)i * post] : zmk(0.0, 0.0);
  __syncthreads();
  auto solve = [&](bool conj) {
    if (conj) pp_band_solve<KD, false, 256>(pl, nl, xs, xp, tb);
    else pp_band_solve<KD, true, 256>(pl, nl, xs, xp, tb);
    for (int i = threadIdx.x; i < npad; i += blockDim.x) xs[i] = xp[i];
    __syncthreads();
  };
  switch (op) {
    case PP_OV_SOLVE_SQRT: solve(false); break;
    case PP_OV_SOLVE_SQRT_CONJ: solve(true); break;
    case PP_OV_MUL_SQRT: pp_band_mul_block(ov.Uband, N, ov.kd, xs, xp, false); break;
    case PP_OV_MUL_SQRT_CONJ: pp_band_mul_block(ov.Uband, N, ov.kd, xs, xp, true); break;
    case PP_OV_SOLVE: solve(true); solve(false); break;     // S^-1 = U^-1 U^-H (overlapmatrix.cpp:131-135)
    case PP_OV_MUL:                                         // S = U^H U
      pp_band_mul_block(ov.Uband, N, ov.kd, xs, xp, false);
      pp_band_mul_block(ov.Uband, N, ov.kd, xs, xp, true);
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
  const pp_band_plan &pl = ov->plan;
  size_t lds = sizeof(double) * (size_t)((pl.nlev + 1) & ~1) + sizeof(zd) * (size_t)(2 * pl.P * pl.B + pl.P * ov->kd);
  bool launched = false;
  hipError_t err = hipSuccess;
  PP_KD_SWITCH(ov->kd, {
    const void *fn = (const void *)pp_overlap_kernel<KD_>;
    err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err == hipSuccess) {
      hipLaunchKernelGGL((pp_overlap_kernel<KD_>), dim3(nfib), dim3(256), lds, stream, *ov, op, (zd *)x, post, nfib);
      err = hipGetLastError();
    }
    launched = true; })
  if (!launched) return PP_EUNSUP;
  if (err != hipSuccess) return PP_EHIP;
  return PP_OK;
}

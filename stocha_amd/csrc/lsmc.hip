// lsmc.hip -- Longstaff-Schwartz American option (lsmc.rs:25-123) on the
// device.
//
// Layout: the GBM paths (mu = r, q = 0, block-split streams as gbm.rs) are
// written date-major, P[j][i] = S_{j+1} of path i, so every per-date pass
// reads one contiguous column with coalesced loads.  Cash flows cf[i] stay
// FP64.  Per exercise date d = steps-1 .. 1 (lsmc.rs:50-99):
//   pass A   exercise at d+1 (with the coefficients solved there), discount,
//            ITM count and shifted sums of (S-K), (S-K)^2   -> block partials
//   reduce A one block: ITM count, mean, population std (floor 1e-12), skip
//            flag when count < ncols+1 (lsmc.rs:62-64,67-71)
//   pass B   FP64 moments of x = (S-mean)/std over ITM paths: sum x^k for
//            k = 0..2*deg and sum cf*x^k for k = 0..deg      -> block partials
//   solve    one block: reduce, form the (deg+1)^2 normal equations, solve in
//            FP64 by Cholesky; this replaces faer's QR least squares
//            (lsmc.rs:83-84) -- same minimiser, different rounding.
// After date 1: exercise at 1, final discount, mean and two-pass variance
// (lsmc.rs:102-112).  All reductions are fixed-order.
#include "host.hpp"

namespace stocha {

constexpr int kLsBlock = 256;
constexpr int kMaxCols = 5;                      // poly_degree <= 4
constexpr int kMaxMom = 2 * (kMaxCols - 1) + 1;  // sum x^0..x^8
constexpr int kMaxAccB = kMaxMom + kMaxCols;

struct DateState {
  double mean, std;
  double coeff[kMaxCols];
  uint64_t itm;
  int skip;  // 1: no regression / exercise at this date
};

__device__ __forceinline__ double intrinsic(double s, double k, int is_put) {
  double v = is_put ? k - s : s - k;
  return v > 0.0 ? v : 0.0;
}

// compiler-rt __powidf2, the lowering of f64::powi (lsmc.rs:78,94)
__device__ __forceinline__ double powi_rt(double a, int b) {
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// ------------------------------------------------------------ generation --
template <bool kFp32>
__global__ __launch_bounds__(kLsBlock) void lsmc_gen_kernel(const GbmConst mc, const StreamDesc sd,
                                                            const uint64_t n, const uint64_t steps,
                                                            void* __restrict__ paths,
                                                            double* __restrict__ cf, const double k,
                                                            const int is_put) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    Pcg g{stream_at(&sd, i)};
    const U128 inc = sd.inc;
    double last;
    if constexpr (kFp32) {
      Gbm32 m(mc, false);
      float* P = (float*)paths;
      float s = 0.f;
      for (uint64_t j = 0; j < steps; ++j) {
        const float ls = m.step(g, inc, &zig);
        s = (float)mc.s0 * __builtin_amdgcn_exp2f(ls * 1.44269504088896341f);
        P[j * n + i] = s;
      }
      last = s;
    } else {
      Gbm64 m(mc, false);
      double* P = (double*)paths;
      for (uint64_t j = 0; j < steps; ++j) P[j * n + i] = m.step(g, inc, &zig);
      last = m.s;
    }
    cf[i] = intrinsic(last, k, is_put);  // lsmc.rs:45-47
  }
}

template <bool kFp32>
__device__ __forceinline__ double spot_at(const void* paths, uint64_t n, uint64_t date, uint64_t i) {
  // date in 1..steps -> column date-1
  if constexpr (kFp32)
    return (double)((const float*)paths)[(date - 1) * n + i];
  else
    return ((const double*)paths)[(date - 1) * n + i];
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NA>
__device__ __forceinline__ void block_store(double (&acc)[NA], int na, double* __restrict__ part) {
  __shared__ double red[kLsBlock / 64][kMaxAccB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    if (a < na) {
      double w = wave_sum_d(acc[a]);
      if (lane == 0) red[wave][a] = w;
    }
  }
  __syncthreads();
  if (threadIdx.x < na) {
    double s = 0.0;
    for (int w = 0; w < kLsBlock / 64; ++w) s += red[w][threadIdx.x];
    part[(uint64_t)blockIdx.x * na + threadIdx.x] = s;
  }
}

// Exercise at date `ex_date` (if it had a regression), then discount and
// gather ITM statistics at date `date` (0 = final pass: mean of cf).
template <bool kFp32>
__global__ __launch_bounds__(kLsBlock) void lsmc_pass_a(const void* __restrict__ paths,
                                                        double* __restrict__ cf, const uint64_t n,
                                                        const uint64_t ex_date,
                                                        const DateState* __restrict__ ds,
                                                        const uint64_t date, const double k,
                                                        const int is_put, const int ncols,
                                                        const double discount,
                                                        double* __restrict__ part) {
  double acc[3] = {0.0, 0.0, 0.0};
  const DateState* ex = ex_date ? &ds[ex_date] : nullptr;
  const bool do_ex = ex && !ex->skip;
  double mean = 0.0, sd = 1.0, c[kMaxCols] = {0, 0, 0, 0, 0};
  if (do_ex) {
    mean = ex->mean;
    sd = ex->std;
    for (int j = 0; j < ncols; ++j) c[j] = ex->coeff[j];
  }
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    double v = cf[i];
    if (do_ex) {  // lsmc.rs:91-98
      const double sp = spot_at<kFp32>(paths, n, ex_date, i);
      const double exv = intrinsic(sp, k, is_put);
      if (exv > 0.0) {
        const double x = (sp - mean) / sd;
        double cont = 0.0;
        for (int j = 0; j < ncols; ++j) cont += c[j] * powi_rt(x, j);
        if (exv > (cont > 0.0 ? cont : 0.0)) v = exv;
      }
    }
    v *= discount;  // lsmc.rs:52-54 (or :102-104 on the final pass)
    cf[i] = v;
    if (date) {
      const double sp = spot_at<kFp32>(paths, n, date, i);
      if (intrinsic(sp, k, is_put) > 0.0) {
        const double u = sp - k;
        acc[0] += 1.0;
        acc[1] += u;
        acc[2] += u * u;
      }
    } else {
      acc[0] += v;
    }
  }
  block_store<3>(acc, date ? 3 : 1, part);
}

// Fixed-order reduction of n_blocks x na partials into out[0..na) (one block).
__device__ void reduce_block_partials(const double* __restrict__ part, int n_blocks, int na,
                                      double* out) {
  __shared__ double buf[kLsBlock];
  for (int a = 0; a < na; ++a) {
    double s = 0.0;
    for (int b = threadIdx.x; b < n_blocks; b += kLsBlock) s += part[(uint64_t)b * na + a];
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int o = kLsBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) buf[threadIdx.x] += buf[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[a] = buf[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kLsBlock) void lsmc_reduce_a(const double* __restrict__ part,
                                                          const int n_blocks, const uint64_t date,
                                                          const double k, const int ncols,
                                                          DateState* __restrict__ ds) {
  __shared__ double tot[3];
  reduce_block_partials(part, n_blocks, 3, tot);
  if (threadIdx.x == 0) {
    DateState& d = ds[date];
    const double m = tot[0];
    d.itm = (uint64_t)m;
    if (d.itm < (uint64_t)ncols + 1) {
      d.skip = 1;
      d.mean = 0.0;
      d.std = 1.0;
    } else {
      const double mu = tot[1] / m;  // mean of (S - K)
      double var = tot[2] / m - mu * mu;
      var = var > 0.0 ? var : 0.0;
      double sdv = sqrt(var);
      d.mean = mu + k;
      d.std = sdv > 1e-12 ? sdv : 1e-12;
      d.skip = 0;
    }
  }
}

template <bool kFp32>
__global__ __launch_bounds__(kLsBlock) void lsmc_pass_b(const void* __restrict__ paths,
                                                        const double* __restrict__ cf,
                                                        const uint64_t n, const uint64_t date,
                                                        const DateState* __restrict__ ds,
                                                        const double k, const int is_put,
                                                        const int ncols,
                                                        double* __restrict__ part) {
  const DateState& d = ds[date];
  if (d.skip) return;
  const int nm = 2 * (ncols - 1) + 1;
  double acc[kMaxAccB];
#pragma unroll
  for (int a = 0; a < kMaxAccB; ++a) acc[a] = 0.0;
  const double mean = d.mean, sd = d.std;
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    const double sp = spot_at<kFp32>(paths, n, date, i);
    if (intrinsic(sp, k, is_put) > 0.0) {
      const double x = (sp - mean) / sd;
      const double y = cf[i];
      double xp = 1.0;
#pragma unroll
      for (int a = 0; a < kMaxMom; ++a) {
        if (a < nm) {
          acc[a] += xp;
          if (a < ncols) acc[kMaxMom + a] += y * xp;
          xp *= x;
        }
      }
    }
  }
  // pack: [0, nm) moments, [nm, nm+ncols) cross moments
  double packed[kMaxAccB];
#pragma unroll
  for (int a = 0; a < kMaxAccB; ++a) packed[a] = 0.0;
#pragma unroll
  for (int a = 0; a < kMaxMom; ++a)
    if (a < nm) packed[a] = acc[a];
#pragma unroll
  for (int a = 0; a < kMaxCols; ++a)
    if (a < ncols) packed[nm + a] = acc[kMaxMom + a];
  block_store<kMaxAccB>(packed, nm + ncols, part);
}

__global__ __launch_bounds__(kLsBlock) void lsmc_solve(const double* __restrict__ part,
                                                       const int n_blocks, const uint64_t date,
                                                       const int ncols,
                                                       DateState* __restrict__ ds) {
  DateState& d = ds[date];
  if (d.skip) return;
  const int nm = 2 * (ncols - 1) + 1;
  __shared__ double tot[kMaxAccB];
  reduce_block_partials(part, n_blocks, nm + ncols, tot);
  if (threadIdx.x == 0) {
    // Normal equations G c = b, G[i][j] = sum x^(i+j), b[i] = sum y x^i.
    double G[kMaxCols][kMaxCols], L[kMaxCols][kMaxCols], b[kMaxCols], w[kMaxCols];
    for (int i = 0; i < ncols; ++i) {
      for (int j = 0; j < ncols; ++j) G[i][j] = tot[i + j];
      b[i] = tot[nm + i];
    }
    // Cholesky G = L L^T (G is SPD when the ITM spots take >= ncols values)
    bool ok = true;
    for (int i = 0; i < ncols && ok; ++i) {
      for (int j = 0; j <= i; ++j) {
        double s = G[i][j];
        for (int q = 0; q < j; ++q) s -= L[i][q] * L[j][q];
        if (i == j) {
          if (!(s > 0.0)) {
            ok = false;
            break;
          }
          L[i][i] = sqrt(s);
        } else {
          L[i][j] = s / L[j][j];
        }
      }
    }
    if (!ok) {
      // Rank-deficient basis (all ITM spots equal): regress on the constant.
      for (int i = 0; i < ncols; ++i) d.coeff[i] = 0.0;
      d.coeff[0] = b[0] / G[0][0];
      return;
    }
    for (int i = 0; i < ncols; ++i) {
      double s = b[i];
      for (int q = 0; q < i; ++q) s -= L[i][q] * w[q];
      w[i] = s / L[i][i];
    }
    for (int i = ncols - 1; i >= 0; --i) {
      double s = w[i];
      for (int q = i + 1; q < ncols; ++q) s -= L[q][i] * d.coeff[q];
      d.coeff[i] = s / L[i][i];
    }
  }
}

// Final variance pass: sum (cf - mean)^2 (lsmc.rs:109).
__global__ __launch_bounds__(kLsBlock) void lsmc_var_pass(const double* __restrict__ cf,
                                                          const uint64_t n,
                                                          const double* __restrict__ mean_sum,
                                                          double* __restrict__ part) {
  const double mean = mean_sum[0] / (double)n;
  double acc[1] = {0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    const double dlt = cf[i] - mean;
    acc[0] += dlt * dlt;
  }
  block_store<1>(acc, 1, part);
}

__global__ void lsmc_reduce_scalar(const double* __restrict__ part, int n_blocks,
                                   double* __restrict__ out) {
  __shared__ double tot[1];
  reduce_block_partials(part, n_blocks, 1, tot);
  if (threadIdx.x == 0) out[0] = tot[0];
}

template <bool kFp32>
static int lsmc_impl(const GbmConst& gc, const StreamDesc& sd, uint64_t n, uint64_t steps,
                     double k, double discount, int is_put, int ncols, double* price,
                     double* std_err, hipStream_t st, const Err& err) {
  const size_t esz = kFp32 ? sizeof(float) : sizeof(double);
  uint64_t nb64 = (n + kLsBlock - 1) / kLsBlock;
  const int nb = (int)(nb64 > 2048 ? 2048 : nb64);
  DevBuf paths, cf, part, ds, scal;
  if (int rc = paths.alloc(esz * n * steps, err)) return rc;
  if (int rc = cf.alloc(sizeof(double) * n, err)) return rc;
  if (int rc = part.alloc(sizeof(double) * (size_t)nb * kMaxAccB, err)) return rc;
  if (int rc = ds.alloc(sizeof(DateState) * (steps + 1), err)) return rc;
  if (int rc = scal.alloc(sizeof(double) * 2, err)) return rc;
  STOCHA_HIP(hipMemsetAsync(ds.p, 0, sizeof(DateState) * (steps + 1), st));

  KernelTimer timer(st);
  timer.start();
  lsmc_gen_kernel<kFp32><<<nb, kLsBlock, 0, st>>>(gc, sd, n, steps, paths.p, cf.as<double>(), k,
                                                  is_put);
  STOCHA_HIP(hipGetLastError());
  DateState* dsp = ds.as<DateState>();
  uint64_t ex_date = 0;
  for (uint64_t date = steps - 1; date >= 1; --date) {
    lsmc_pass_a<kFp32><<<nb, kLsBlock, 0, st>>>(paths.p, cf.as<double>(), n, ex_date, dsp, date,
                                                k, is_put, ncols, discount, part.as<double>());
    lsmc_reduce_a<<<1, kLsBlock, 0, st>>>(part.as<double>(), nb, date, k, ncols, dsp);
    lsmc_pass_b<kFp32><<<nb, kLsBlock, 0, st>>>(paths.p, cf.as<double>(), n, date, dsp, k,
                                                is_put, ncols, part.as<double>());
    lsmc_solve<<<1, kLsBlock, 0, st>>>(part.as<double>(), nb, date, ncols, dsp);
    ex_date = date;
    if (date == 1) break;
  }
  // exercise at the last regression date, final discount, sum of cf
  lsmc_pass_a<kFp32><<<nb, kLsBlock, 0, st>>>(paths.p, cf.as<double>(), n, ex_date, dsp, 0, k,
                                              is_put, ncols, discount, part.as<double>());
  lsmc_reduce_scalar<<<1, kLsBlock, 0, st>>>(part.as<double>(), nb, scal.as<double>());
  lsmc_var_pass<<<nb, kLsBlock, 0, st>>>(cf.as<double>(), n, scal.as<double>(), part.as<double>());
  lsmc_reduce_scalar<<<1, kLsBlock, 0, st>>>(part.as<double>(), nb, scal.as<double>() + 1);
  STOCHA_HIP(hipGetLastError());
  timer.stop(0);
  double h[2];
  STOCHA_HIP(hipMemcpyAsync(h, scal.p, sizeof(h), hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  const double mean = h[0] / (double)n;
  *price = mean;
  *std_err = n <= 1 ? 0.0 : sqrt((h[1] / (double)(n - 1)) / (double)n);
  return STOCHA_OK;
}

int run_lsmc(const GbmConst& gc, const StreamDesc& sd, uint64_t n_paths, uint64_t steps,
             double k, double discount, int is_put, int ncols, int precision, double* price,
             double* std_err, hipStream_t st, const Err& err) {
  if (precision == STOCHA_FP32)
    return lsmc_impl<true>(gc, sd, n_paths, steps, k, discount, is_put, ncols, price, std_err,
                           st, err);
  return lsmc_impl<false>(gc, sd, n_paths, steps, k, discount, is_put, ncols, price, std_err, st,
                          err);
}

}  // namespace stocha

// multi.hip -- correlated multi-asset GBM (multi_gbm.rs:22-113).
//
// FP64 path-storing form (the drop-in multi_gbm): one thread per output row,
// the reference's arithmetic -- d normals per step in asset order (negated
// for antithetic rows), then for each asset i the sequential dot
// eps = sum_{k<=i} L[i][k] z[k] starting from 0.0 and s_i *= exp(drift_i +
// diffusion_i * eps).  The per-thread z and s vectors live in a scratch
// buffer laid out [asset][row] so the per-asset accesses of a wavefront are
// coalesced; L, drift and diffusion are uniform and read with scalar loads.
// With corr = [[1]] this is bit-for-bit the GBM stepper (same block size, L=1).
#include <math.h>

#include <string>
#include <vector>

#include "multi.hpp"

namespace stocha {

int cholesky_host(const double* a, uint64_t n, double* l, const Err& err) {
  for (uint64_t i = 0; i < n * n; ++i) l[i] = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    for (uint64_t j = 0; j <= i; ++j) {
      double sum = a[i * n + j];
      for (uint64_t k = 0; k < j; ++k) sum -= l[i * n + k] * l[j * n + k];
      if (i == j) {
        if (sum <= 0.0) {
          std::string v = rust_f64(sum);
          return err.set(STOCHA_EINVAL,
                         "matrix is not positive-definite (diagonal element %llu = %s)",
                         (unsigned long long)i, v.c_str());
        }
        l[i * n + j] = sqrt(sum);
      } else {
        l[i * n + j] = sum / l[j * n + j];
      }
    }
  }
  return STOCHA_OK;
}

constexpr int kMultiBlock = 128;

__global__ __launch_bounds__(kMultiBlock) void multi_paths_kernel(
    const StreamDesc sd, const uint64_t d, const uint64_t steps, const uint64_t n_rows,
    const uint64_t base_rows, const double* __restrict__ L, const double* __restrict__ s0,
    const double* __restrict__ drift, const double* __restrict__ diff,
    double* __restrict__ zs, double* __restrict__ ss, double* __restrict__ out) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t row = (uint64_t)blockIdx.x * kMultiBlock + threadIdx.x;
  if (row >= n_rows) return;
  const bool mirror = row >= base_rows;
  const uint64_t path = mirror ? row - base_rows : row;
  Pcg g{stream_at(&sd, path)};
  const U128 inc = sd.inc;
  double* o = out + row * (steps + 1) * d;
  for (uint64_t i = 0; i < d; ++i) {
    ss[i * n_rows + row] = s0[i];
    o[i] = s0[i];
  }
  for (uint64_t j = 0; j < steps; ++j) {
    for (uint64_t i = 0; i < d; ++i) {  // sample_into (normal.rs:70-74)
      const double z = zig_normal64(g, inc, &zig);
      zs[i * n_rows + row] = mirror ? -z : z;
    }
    for (uint64_t i = 0; i < d; ++i) {  // multi_gbm.rs:67-74
      double eps = 0.0;
      for (uint64_t k = 0; k <= i; ++k) eps += L[i * d + k] * zs[k * n_rows + row];
      const double s = ss[i * n_rows + row] * exp(drift[i] + diff[i] * eps);
      ss[i * n_rows + row] = s;
      o[(j + 1) * d + i] = s;
    }
  }
}

int run_multi_gbm_paths(const double* s0, const double* mu, const double* sigma,
                        const double* corr, uint64_t d, double t, uint64_t steps,
                        uint64_t n_paths, uint64_t seed, int antithetic, double* out, int where,
                        hipStream_t st, const Err& err) {
  std::vector<double> L(d * d), consts(3 * d);
  if (int rc = cholesky_host(corr, d, L.data(), err)) return rc;
  const double dt = t / (double)steps;
  const double sqrt_dt = sqrt(dt);
  for (uint64_t i = 0; i < d; ++i) {
    consts[i] = s0[i];
    consts[d + i] = (mu[i] - 0.5 * sigma[i] * sigma[i]) * dt;  // multi_gbm.rs:29-31
    consts[2 * d + i] = sigma[i] * sqrt_dt;                     // :32
  }
  const uint64_t base = antithetic ? (n_paths + 1) / 2 : n_paths;
  const uint64_t block = (steps * d + 1024) * 2;  // multi_gbm.rs:40
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);

  DevBuf dL, dc, dz, ds, dout;
  if (int rc = dL.alloc(sizeof(double) * d * d, err)) return rc;
  if (int rc = dc.alloc(sizeof(double) * 3 * d, err)) return rc;
  if (int rc = dz.alloc(sizeof(double) * d * n_paths, err)) return rc;
  if (int rc = ds.alloc(sizeof(double) * d * n_paths, err)) return rc;
  const size_t bytes = sizeof(double) * n_paths * (steps + 1) * d;
  double* target = out;
  if (where != STOCHA_DEVICE) {
    if (int rc = dout.alloc(bytes, err)) return rc;
    target = dout.as<double>();
  }
  STOCHA_HIP(hipMemcpyAsync(dL.p, L.data(), sizeof(double) * d * d, hipMemcpyHostToDevice, st));
  STOCHA_HIP(hipMemcpyAsync(dc.p, consts.data(), sizeof(double) * 3 * d, hipMemcpyHostToDevice, st));
  const uint64_t blocks = (n_paths + kMultiBlock - 1) / kMultiBlock;
  KernelTimer timer(st);
  timer.start();
  multi_paths_kernel<<<(unsigned)blocks, kMultiBlock, 0, st>>>(
      sd, d, steps, n_paths, base, dL.as<double>(), dc.as<double>(), dc.as<double>() + d,
      dc.as<double>() + 2 * d, dz.as<double>(), ds.as<double>(), target);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  if (where != STOCHA_DEVICE)
    STOCHA_HIP(hipMemcpyAsync(out, target, bytes, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

}  // namespace stocha

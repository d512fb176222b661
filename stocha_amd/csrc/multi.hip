// multi.hip -- correlated multi-asset GBM (multi_gbm.rs:22-113).
//
// FP64 path-storing form (the drop-in multi_gbm): one thread per output row,
// the reference's arithmetic -- d normals per step in asset order (negated
// for antithetic rows), then for each asset i the sequential dot
// eps = sum_{k<=i} L[i][k] z[k] starting from 0.0 and s_i *= exp(drift_i +
// diffusion_i * eps).  The normals come from the queue schedule
// (drivers.hpp): each queue step hands every lane one normal, and the model
// collects d of them before it applies the step, so a wave never waits on a
// lane's slow normal.  The per-row z and s vectors live in workspace buffers
// laid out [asset][row] (coalesced per asset across a wavefront); L, drift
// and diffusion are uniform and read with scalar loads.  Up to kMultiTile
// values per row are staged in an LDS tile and written row segment by row
// segment (d <= kMultiTile; larger d writes each row's d values directly).
// With corr = [[1]] this is bit-for-bit the GBM stepper (same block size, L=1).
#include <math.h>

#include <string>
#include <vector>

#include "multi.hpp"
#include "drivers.hpp"

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

constexpr int kMultiBlock = 256;
constexpr int kMultiTile = 16;  // values per row staged before a flush

struct MultiArgs {
  const double* L;      // d x d row-major (lower triangle used)
  const double* s0;     // d
  const double* drift;  // d: (mu_i - 0.5 sigma_i^2) dt
  const double* diff;   // d: sigma_i sqrt(dt)
  double* zs;           // [d][stride] this step's normals of each row
  double* ss;           // [d][stride] each row's prices
  double* out;          // (n_rows, steps+1, d)
  uint64_t n_rows, base_rows, stride;
  uint32_t d, steps;
};

// The queue model: one normal per apply; the d-th completes a step
// (multi_gbm.rs:60-76).  Every lane of the block completes its steps on the
// same queue steps, so the tile flushes' barriers line up.
template <bool kTiled>
struct MultiModel {
  const MultiArgs a;
  const uint64_t row;  // < a.stride (rows past n_rows use padding and write nothing)
  const bool mirror, valid;
  double (*tile)[kMultiBlock + 1];
  uint32_t k = 0, step = 0, tpos = 0, tcap = 0;
  uint64_t col0 = 0;  // output column (flattened step x asset) of the tile's first value
  __device__ __forceinline__ MultiModel(const MultiArgs& args, uint64_t r, bool m, bool v,
                                        double (*t)[kMultiBlock + 1])
      : a(args), row(r), mirror(m), valid(v), tile(t) {
    tcap = kTiled ? (kMultiTile / a.d) * a.d : 0;  // whole steps per tile
    for (uint32_t i = 0; i < a.d; ++i) {
      a.ss[(uint64_t)i * a.stride + row] = a.s0[i];
      if (kTiled) tile[i][threadIdx.x] = a.s0[i];
    }
    tpos = kTiled ? a.d : 0;
    if (!kTiled && valid)
      for (uint32_t i = 0; i < a.d; ++i) a.out[row * (a.steps + 1ull) * a.d + i] = a.s0[i];
    // invariant before every step: the step's d values fit (tpos + d <= tcap)
    if (kTiled && tpos + a.d > tcap) flush();
  }
  // (LDS-only barriers: the row stores stay in flight across them)
  __device__ __forceinline__ void flush() {
    lds_barrier();
    const uint64_t row0 = (uint64_t)blockIdx.x * kMultiBlock;
    const uint64_t left = a.n_rows - row0;
    const int rows_here = left < (uint64_t)kMultiBlock ? (int)left : kMultiBlock;
    const int n_el = rows_here * (int)tpos;
    const uint64_t width = (a.steps + 1ull) * a.d;
    for (int e = threadIdx.x; e < n_el; e += kMultiBlock) {
      const int r = e / (int)tpos;
      const int c = e - r * (int)tpos;
      a.out[(row0 + r) * width + col0 + c] = tile[c][r];
    }
    lds_barrier();
    col0 += tpos;
    tpos = 0;
  }
  __device__ __forceinline__ void apply(double z) {
    a.zs[(uint64_t)k * a.stride + row] = mirror ? -z : z;  // sample_into (normal.rs:70-74)
    if (++k < a.d) return;
    k = 0;
    ++step;
    for (uint32_t i = 0; i < a.d; ++i) {  // multi_gbm.rs:67-74
      double eps = 0.0;
      for (uint32_t kk = 0; kk <= i; ++kk) eps += a.L[i * a.d + kk] * a.zs[(uint64_t)kk * a.stride + row];
      double* sp = &a.ss[(uint64_t)i * a.stride + row];
      const double s = *sp * exp(a.drift[i] + a.diff[i] * eps);
      *sp = s;
      if (kTiled) {
        tile[tpos + i][threadIdx.x] = s;
      } else if (valid) {
        a.out[(row * (a.steps + 1ull) + step) * a.d + i] = s;
      }
    }
    if (kTiled) {
      tpos += a.d;
      if (tpos + a.d > tcap || step == a.steps) flush();
    }
  }
};

template <bool kTiled>
__global__ __launch_bounds__(kMultiBlock) void multi_feed_kernel(const MultiArgs a,
                                                                 const StreamDesc sd) {
  using Src = GbmSrc<false>;
  __shared__ ZigLds64 zig;
  __shared__ double q[8][kMultiBlock];
  __shared__ double tile[kTiled ? kMultiTile : 1][kMultiBlock + 1];
  zig_load(&zig);
  const uint64_t row = (uint64_t)blockIdx.x * kMultiBlock + threadIdx.x;
  // rows past n_rows step a valid stream (path 0) in padding slots and write nothing
  const bool valid = row < a.n_rows;
  const bool mirror = valid && row >= a.base_rows;
  const uint64_t path = !valid ? 0 : (mirror ? row - a.base_rows : row);
  Pcg g{stream_at(&sd, path)};
  MultiModel<kTiled> m(a, row, mirror, valid, tile);
  __syncthreads();
  Src src(a);
  queue_run<Src, 8, kMultiBlock>(src, m, g, sd.inc, a.steps * a.d, &zig, q);
}

int run_multi_gbm_paths(const double* s0, const double* mu, const double* sigma,
                        const double* corr, uint64_t d, double t, uint64_t steps,
                        uint64_t n_paths, uint64_t seed, int antithetic, double* out, int where,
                        hipStream_t st, const Err& err) {
  std::vector<double> L(d * d), consts(3 * d);
  if (int rc = cholesky_host(corr, d, L.data(), err)) return rc;
  if (steps * d > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps x d must be below 2^32");
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

  // workspace: [host-bound output] | L | consts | z | s  (256-byte aligned)
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const uint64_t blocks = (n_paths + kMultiBlock - 1) / kMultiBlock;
  const uint64_t stride = blocks * kMultiBlock;
  const size_t out_bytes = sizeof(double) * n_paths * (steps + 1) * d;
  const size_t o_l = where == STOCHA_DEVICE ? 0 : up(out_bytes);
  const size_t o_c = o_l + up(sizeof(double) * d * d);
  const size_t o_z = o_c + up(sizeof(double) * 3 * d);
  const size_t o_s = o_z + up(sizeof(double) * d * stride);
  Workspace* ws = nullptr;
  if (int rc = workspace(o_s + up(sizeof(double) * d * stride), 0, &ws, err)) return rc;
  char* wb = static_cast<char*>(ws->dev);
  double* target = where == STOCHA_DEVICE ? out : reinterpret_cast<double*>(wb);
  MultiArgs a;
  a.L = reinterpret_cast<double*>(wb + o_l);
  a.s0 = reinterpret_cast<double*>(wb + o_c);
  a.drift = a.s0 + d;
  a.diff = a.s0 + 2 * d;
  a.zs = reinterpret_cast<double*>(wb + o_z);
  a.ss = reinterpret_cast<double*>(wb + o_s);
  a.out = target;
  a.n_rows = n_paths;
  a.base_rows = base;
  a.stride = stride;
  a.d = (uint32_t)d;
  a.steps = (uint32_t)steps;
  STOCHA_HIP(hipMemcpyAsync(wb + o_l, L.data(), sizeof(double) * d * d, hipMemcpyHostToDevice, st));
  STOCHA_HIP(hipMemcpyAsync(wb + o_c, consts.data(), sizeof(double) * 3 * d,
                            hipMemcpyHostToDevice, st));
  if (blocks > 0x7fffffffull) return err.set(STOCHA_EINVAL, "n_paths too large for one launch");
  KernelTimer timer(st);
  timer.start();
  if (d <= (uint64_t)kMultiTile)
    multi_feed_kernel<true><<<(unsigned)blocks, kMultiBlock, 0, st>>>(a, sd);
  else
    multi_feed_kernel<false><<<(unsigned)blocks, kMultiBlock, 0, st>>>(a, sd);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  if (where != STOCHA_DEVICE)
    STOCHA_HIP(hipMemcpyAsync(out, target, out_bytes, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

}  // namespace stocha

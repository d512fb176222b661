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
// For d <= kRegD (multi_reg_kernel) a step's normals and the row's prices sit
// in registers instead: the queue fill asks each lane for d normals at once,
// and the step reads them from the ring in asset order.
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

// Writes the tile's tpos values of each of the wavefront's rows (output
// columns col0 .. col0+tpos-1) and empties the tile; every lane of the
// wavefront calls it together.  Each wavefront owns its 64 rows of the tile
// (no block barrier).  Lane l owns tile column l % kMultiTile and rows
// l / kMultiTile + kFlushRows * j: each pass stores kFlushRows row segments
// of up to kMultiTile consecutive values (16 lanes per 128-byte segment), one
// pointer add per store.
constexpr uint32_t kFlushRows = 64 / kMultiTile;
static_assert(64 % kMultiTile == 0, "flush layout");
__device__ __forceinline__ void flush_tile(const MultiArgs& a, double (*tile)[kMultiBlock + 1],
                                           uint32_t& tpos, uint64_t& col0) {
  wave_lds_sync();
  const uint32_t wbase = threadIdx.x & ~63u, lane = threadIdx.x & 63u;
  const uint64_t row0 = (uint64_t)blockIdx.x * kMultiBlock + wbase;
  const uint32_t rows_here =
      a.n_rows <= row0 ? 0u : a.n_rows - row0 < 64u ? (uint32_t)(a.n_rows - row0) : 64u;
  const uint64_t width = (a.steps + 1ull) * a.d;
  const uint32_t c = lane % kMultiTile;
  const uint32_t r0 = lane / kMultiTile;
  if (c < tpos) {
    double* p = a.out + (row0 + r0) * width + col0 + c;
    const uint64_t pass = (uint64_t)kFlushRows * width;
    for (uint32_t r = r0; r < rows_here; r += kFlushRows, p += pass) *p = tile[c][wbase + r];
  }
  wave_lds_sync();
  col0 += tpos;
  tpos = 0;
}

// The queue model: one normal per apply; the d-th completes a step
// (multi_gbm.rs:60-76).  Every lane of a wavefront completes its steps on the
// same queue steps, so the wavefront flushes its tile rows together.
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
  __device__ __forceinline__ void flush() { flush_tile(a, tile, tpos, col0); }
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
  if ((uint64_t)blockIdx.x * kMultiBlock + (threadIdx.x & ~63u) >= a.n_rows) return;  // no row
  Src src(a);
  queue_run<Src, 8, kMultiBlock>(src, m, g, sd.inc, a.steps * a.d, &zig, q);
}

// d <= kRegD: one kernel per d (template D), with the step's normals z and
// the row's prices s in registers, and L's lower triangle, drift and
// diffusion too: held in VGPRs for the whole loop (as scalars they would not
// fit the SGPR file at d = 8 and spill).  With d fixed at compile time the d
// exp chains of a step are independent straight-line code the compiler
// interleaves.  A queue fill leaves every lane at least D normals in its ring
// (kRegSlots - 1 entries at most); the step takes them in draw order, so each
// row consumes its stream exactly as the one-normal-per-step form does.  The
// arithmetic is multi_gbm.rs:67-74's, term for term: eps from 0.0 by products
// and sums in k order, then s * exp(drift + diffusion * eps).
constexpr uint32_t kRegD = 8;
constexpr uint32_t kRegSlots = 16;

template <uint32_t D>
__global__ __launch_bounds__(kMultiBlock) void multi_reg_kernel(const MultiArgs a,
                                                                const StreamDesc sd) {
  using Src = GbmSrc<false>;
  static_assert(D >= 1 && D <= kRegD && kRegSlots > kRegD && kMultiTile >= kRegD,
                "ring and tile sizes");
  constexpr uint32_t kCap = (kMultiTile / D) * D;  // whole steps per tile (>= 2 steps)
  __shared__ ZigLds64 zig;
  __shared__ double q[kRegSlots][kMultiBlock];
  __shared__ double tile[kMultiTile][kMultiBlock + 1];
  zig_load(&zig);
  const uint64_t row = (uint64_t)blockIdx.x * kMultiBlock + threadIdx.x;
  const bool valid = row < a.n_rows;
  const bool mirror = valid && row >= a.base_rows;
  const uint64_t path = !valid ? 0 : (mirror ? row - a.base_rows : row);
  const U128 st0 = stream_at(&sd, path);
  double s[D], lt[D * (D + 1) / 2], dr[D], df[D];
#pragma unroll
  for (uint32_t i = 0; i < D; ++i) {
#pragma unroll
    for (uint32_t k = 0; k <= i; ++k) {
      lt[i * (i + 1) / 2 + k] = a.L[i * D + k];
      asm("" : "+v"(lt[i * (i + 1) / 2 + k]));
    }
    dr[i] = a.drift[i];
    df[i] = a.diff[i];
    asm("" : "+v"(dr[i]), "+v"(df[i]));
    s[i] = a.s0[i];
    tile[i][threadIdx.x] = s[i];
  }
  uint32_t tpos = D;
  uint64_t col0 = 0;
  __syncthreads();
  if ((uint64_t)blockIdx.x * kMultiBlock + (threadIdx.x & ~63u) >= a.n_rows) return;  // no row
  Src src(a);
  const LcgInc incv = vgpr_increment(sd.inc);
  DrawAhead<Src> da(st0, &zig);
  uint32_t head = 0, cnt = 0;
  for (uint32_t step = 1; step <= a.steps; ++step) {
    queue_fill_ahead<Src, kRegSlots, kMultiBlock>(src, da, incv, &zig, q, head, cnt, D);
    double z[D];
#pragma unroll
    for (uint32_t k = 0; k < D; ++k) {
      const double v = q[(head + k) & (kRegSlots - 1)][threadIdx.x];
      z[k] = mirror ? -v : v;  // sample_into (normal.rs:70-74)
    }
    head = (head + D) & (kRegSlots - 1);
    cnt -= D;
#pragma unroll
    for (uint32_t i = 0; i < D; ++i) {
      double eps = 0.0;
#pragma unroll
      for (uint32_t k = 0; k <= i; ++k) eps += lt[i * (i + 1) / 2 + k] * z[k];
      s[i] = s[i] * exp(dr[i] + df[i] * eps);
      tile[tpos + i][threadIdx.x] = s[i];
    }
    tpos += D;
    if (tpos + D > kCap || step == a.steps) flush_tile(a, tile, tpos, col0);
  }
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
  const uint64_t zs_len = d <= kRegD ? 0 : d * stride;  // per-row z and s in registers
  const size_t o_z = o_c + up(sizeof(double) * 3 * d);
  const size_t o_s = o_z + up(sizeof(double) * zs_len);
  Workspace* ws = nullptr;
  if (int rc = workspace(o_s + up(sizeof(double) * zs_len), 0, &ws, err)) return rc;
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
  const dim3 grid((unsigned)blocks), blk(kMultiBlock);
  if (d == 1) multi_reg_kernel<1><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 2) multi_reg_kernel<2><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 3) multi_reg_kernel<3><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 4) multi_reg_kernel<4><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 5) multi_reg_kernel<5><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 6) multi_reg_kernel<6><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 7) multi_reg_kernel<7><<<grid, blk, 0, st>>>(a, sd);
  else if (d == 8) multi_reg_kernel<8><<<grid, blk, 0, st>>>(a, sd);
  else if (d <= (uint64_t)kMultiTile)
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

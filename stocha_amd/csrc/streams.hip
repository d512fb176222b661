// streams.hip -- the random-number primitives on their own: raw PCG64DXSM
// draws and ziggurat normals of the block-split streams (pcg64.rs:23-43,
// normal.rs:18-74), one thread per path, output [path][draw] row-major.
// The steppers inline the same device functions (rng.hpp).
#include "host.hpp"

namespace stocha {

constexpr int kStreamBlock = 128;

__global__ __launch_bounds__(kStreamBlock) void stream_u64_kernel(const StreamDesc sd,
                                                                  const uint64_t path0,
                                                                  const uint64_t n_paths,
                                                                  const uint64_t n_draws,
                                                                  uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * kStreamBlock + threadIdx.x;
  if (i >= n_paths) return;
  Pcg g{stream_at(&sd, path0 + i)};
  for (uint64_t j = 0; j < n_draws; ++j) out[i * n_draws + j] = g.next(sd.inc);
}

// precision 64: FP64 normals; 32: FP32 normals (stored as double).  Also
// writes the number of u64 draws each path consumed.
template <bool kFp32>
__global__ __launch_bounds__(kStreamBlock) void stream_normal_kernel(
    const StreamDesc sd, const uint64_t path0, const uint64_t n_paths, const uint64_t n_draws,
    double* __restrict__ out, uint64_t* __restrict__ consumed) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kStreamBlock + threadIdx.x;
  if (i >= n_paths) return;
  Pcg g{stream_at(&sd, path0 + i)};
  const U128 start = g.s;
  for (uint64_t j = 0; j < n_draws; ++j)
    out[i * n_draws + j] = kFp32 ? (double)zig_normal32(g, sd.inc, &zig)
                                 : zig_normal64(g, sd.inc, &zig);
  if (consumed) {
    // distance from start: step a copy until it meets the final state
    Pcg h{start};
    uint64_t c = 0;
    while ((h.s.lo != g.s.lo || h.s.hi != g.s.hi) && c < 64 * n_draws + 64) {
      (void)h.next(sd.inc);
      ++c;
    }
    consumed[i] = c;
  }
}

}  // namespace stocha

using namespace stocha;

extern "C" int stocha_stream_u64(uint64_t seed, uint64_t block, uint64_t path0, uint64_t n_paths,
                                 uint64_t n_draws, uint64_t* out, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (n_paths == 0 || n_draws == 0) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);
  DevBuf d;
  const size_t bytes = sizeof(uint64_t) * n_paths * n_draws;
  if (int rc = d.alloc(bytes, err)) return rc;
  stream_u64_kernel<<<(unsigned)((n_paths + kStreamBlock - 1) / kStreamBlock), kStreamBlock, 0,
                      st>>>(sd, path0, n_paths, n_draws, d.as<uint64_t>());
  STOCHA_HIP(hipGetLastError());
  STOCHA_HIP(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

extern "C" int stocha_stream_normals(uint64_t seed, uint64_t block, uint64_t path0,
                                     uint64_t n_paths, uint64_t n_draws, int precision,
                                     double* out, uint64_t* consumed, char* errbuf,
                                     size_t errlen) {
  Err err{errbuf, errlen};
  if (precision != STOCHA_FP64 && precision != STOCHA_FP32)
    return err.set(STOCHA_EINVAL, "precision must be 32 or 64");
  if (n_paths == 0 || n_draws == 0) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);
  DevBuf d, c;
  const size_t bytes = sizeof(double) * n_paths * n_draws;
  if (int rc = d.alloc(bytes, err)) return rc;
  if (int rc = c.alloc(sizeof(uint64_t) * n_paths, err)) return rc;
  const unsigned grid = (unsigned)((n_paths + kStreamBlock - 1) / kStreamBlock);
  if (precision == STOCHA_FP32)
    stream_normal_kernel<true><<<grid, kStreamBlock, 0, st>>>(sd, path0, n_paths, n_draws,
                                                              d.as<double>(), c.as<uint64_t>());
  else
    stream_normal_kernel<false><<<grid, kStreamBlock, 0, st>>>(sd, path0, n_paths, n_draws,
                                                               d.as<double>(), c.as<uint64_t>());
  STOCHA_HIP(hipGetLastError());
  STOCHA_HIP(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, st));
  if (consumed)
    STOCHA_HIP(hipMemcpyAsync(consumed, c.p, sizeof(uint64_t) * n_paths, hipMemcpyDeviceToHost,
                              st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

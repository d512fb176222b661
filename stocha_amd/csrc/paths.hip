// paths.hip -- path-storing FP64 steppers (the drop-in gbm / heston /
// merton_jump_diffusion outputs).
//
// One thread owns one output row (one path, or its antithetic mirror) and
// steps it in registers with the reference's arithmetic (models.hpp).  The
// output is the reference's row-major (n_rows, steps+1) matrix, so a thread's
// consecutive values are 8 bytes apart and neighbouring threads are a whole
// row apart.  To make the HBM writes coalesced the block stages kTile columns
// of its 256 rows in LDS and then writes them row by row: 16 consecutive
// lanes cover one 128-byte row segment.
#include "host.hpp"

namespace stocha {

constexpr int kPathBlock = 256;
constexpr int kTile = 16;

template <class Model, class Const>
__global__ __launch_bounds__(kPathBlock) void paths_kernel(const Const mc, const StreamDesc sd,
                                                           const uint64_t n_rows,
                                                           const uint64_t base_rows,
                                                           const uint64_t steps,
                                                           double* __restrict__ out) {
  __shared__ ZigLds zig;
  __shared__ double tile[kTile][kPathBlock + 1];
  zig_load(&zig);

  const uint64_t row0 = (uint64_t)blockIdx.x * kPathBlock;
  const uint64_t row = row0 + threadIdx.x;
  // Threads past the last row still take part in the barriers; they step a
  // valid stream (path 0) and their values are never written.
  const bool valid = row < n_rows;
  const bool mirror = valid && row >= base_rows;
  const uint64_t path = !valid ? 0 : (mirror ? row - base_rows : row);
  Pcg g{stream_at(&sd, path)};
  const U128 inc = sd.inc;
  Model m(mc, mirror);
  __syncthreads();

  const uint64_t width = steps + 1;
  const uint64_t rows_here = n_rows - row0 < (uint64_t)kPathBlock ? n_rows - row0 : kPathBlock;
  for (uint64_t c0 = 0; c0 < width; c0 += kTile) {
    const int cols = (int)((width - c0) < (uint64_t)kTile ? (width - c0) : kTile);
    for (int j = 0; j < cols; ++j) {
      const uint64_t col = c0 + j;
      tile[j][threadIdx.x] = col == 0 ? mc.s0 : m.step(g, inc, &zig);
    }
    __syncthreads();
    // Row-major write-out: element e of the tile is (row e / cols, col e % cols).
    const int n_el = (int)rows_here * cols;
    for (int e = threadIdx.x; e < n_el; e += kPathBlock) {
      const int r = e / cols;
      const int c = e - r * cols;
      out[(row0 + r) * width + c0 + c] = tile[c][r];
    }
    __syncthreads();
  }
}

int launch_paths(int model, const void* mconst, const StreamDesc& sd, uint64_t n_rows,
                 uint64_t base_rows, uint64_t steps, double* d_out, hipStream_t st,
                 const Err& err) {
  const uint64_t blocks = (n_rows + kPathBlock - 1) / kPathBlock;
  if (blocks > 0x7fffffffull) return err.set(STOCHA_EINVAL, "n_paths too large for one launch");
  dim3 grid((unsigned)blocks), block(kPathBlock);
  KernelTimer timer(st);
  timer.start();
  switch (model) {
    case kGbm:
      paths_kernel<Gbm64, GbmConst><<<grid, block, 0, st>>>(*(const GbmConst*)mconst, sd, n_rows,
                                                            base_rows, steps, d_out);
      break;
    case kHeston:
      paths_kernel<Heston64, HestonConst><<<grid, block, 0, st>>>(
          *(const HestonConst*)mconst, sd, n_rows, base_rows, steps, d_out);
      break;
    case kMerton:
      paths_kernel<Merton64, MertonConst><<<grid, block, 0, st>>>(
          *(const MertonConst*)mconst, sd, n_rows, base_rows, steps, d_out);
      break;
    default:
      return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  }
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  return STOCHA_OK;
}

}  // namespace stocha

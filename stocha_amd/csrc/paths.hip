// paths.hip -- path-storing FP64 steppers (the drop-in gbm / heston /
// merton_jump_diffusion / hull_white outputs).
//
// One thread owns one output row (one path, or its antithetic mirror) and
// steps it with the reference's arithmetic (models.hpp) on the queue schedule
// (drivers.hpp: per-lane ziggurat state machine, steps applied by the whole
// wave).  The output is the reference's row-major (n_rows, steps+1) matrix,
// so a thread's consecutive values are 8 bytes apart and neighbouring threads
// are a whole row apart.  To make the HBM writes coalesced each wavefront
// stages kTile columns of its 64 rows in LDS and then writes them row by row:
// 8 consecutive lanes cover one 64-byte row segment.  Lane l owns tile column
// l % kTile and rows l / kTile + kWaveRows * j, so a flush is kTile passes of
// one LDS read, one store and one pointer add.  The tile of a wavefront is
// its own (no block barrier: the wavefronts of a block run their steps
// independently).
#include "host.hpp"
#include "drivers.hpp"

namespace stocha {

constexpr int kPathBlock = 256;
constexpr int kTile = 8;
constexpr int kWaveRows = 64 / kTile;  // rows written per pass of a flush

// Step hook: column step+1 of the wavefront's rows goes to the LDS tile; a
// full tile (or the last column) is written out row-major.  Runs on every
// lane of the wavefront after each step.
struct TileHook {
  double (*tile)[kPathBlock + 1];
  double* out;
  uint64_t row0, width;  // row0: the wavefront's first row
  uint32_t rows_here;    // the wavefront's rows below n_rows (0..64)
  __device__ __forceinline__ void flush(uint64_t c0, int cols) const {
    wave_lds_sync();
    const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
    const int c = (int)(lane % kTile);
    const uint32_t r0 = lane / kTile;
    if (c < cols) {
      double* p = out + (row0 + r0) * width + c0 + c;
      const uint64_t pass = (uint64_t)kWaveRows * width;
      for (uint32_t r = r0; r < rows_here; r += kWaveRows, p += pass) *p = tile[c][wbase + r];
    }
    wave_lds_sync();
  }
  template <class M>
  __device__ __forceinline__ void operator()(uint32_t step, const M& m) const {
    const uint64_t col = (uint64_t)step + 1;
    const int j = (int)(col % kTile);
    tile[j][threadIdx.x] = m.terminal();
    if (j == kTile - 1 || col == width - 1) flush(col - j, j + 1);
  }
};

// kAhead: the fill runs the generator one draw ahead (queue_run_ahead); for
// the normal-only sources, whose table read per draw is the fill loop's
// longest wait at these few waves per SIMD.
template <class Src, class Model, class Const, uint32_t kSlots, bool kAhead = false>
__global__ __launch_bounds__(kPathBlock) void paths_kernel(const Const mc, const StreamDesc sd,
                                                           const uint64_t n_rows,
                                                           const uint64_t base_rows,
                                                           const uint32_t steps,
                                                           double* __restrict__ out) {
  using Rec = typename Src::Rec;
  __shared__ typename Src::Tables zig;
  __shared__ Rec q[kSlots][kPathBlock];
  __shared__ double tile[kTile][kPathBlock + 1];
  zig_load(&zig);

  const uint64_t row0 = (uint64_t)blockIdx.x * kPathBlock;
  const uint64_t row = row0 + threadIdx.x;
  // Lanes past the last row in a wavefront with rows still take part in the
  // steps; they step a valid stream (path 0) and their values are never
  // written.  Wavefronts with no row leave after the table load.
  const bool valid = row < n_rows;
  const bool mirror = valid && row >= base_rows;
  const uint64_t path = !valid ? 0 : (mirror ? row - base_rows : row);
  const U128 s0 = stream_at(&sd, path);
  tile[0][threadIdx.x] = mc.s0;
  __syncthreads();

  TileHook hook;
  hook.tile = tile;
  hook.out = out;
  hook.row0 = row0 + (threadIdx.x & ~63u);
  hook.width = (uint64_t)steps + 1;
  hook.rows_here = n_rows <= hook.row0 ? 0u
                   : n_rows - hook.row0 < 64u ? (uint32_t)(n_rows - hook.row0) : 64u;
  if (hook.rows_here == 0) return;  // wave-uniform
  if constexpr (kAhead) {
    Src src(mc);
    Model m(mc, mirror);
    queue_run_ahead<Src, kSlots, kPathBlock>(src, m, s0, sd.inc, steps, &zig, q, hook);
  } else {
    Pcg g{s0};
    (void)queue_path<Src, Model, kSlots, kPathBlock>(mc, g, sd.inc, steps, &zig, q, mirror, hook);
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
  if (steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps must be below 2^32");
  const uint32_t st32 = (uint32_t)steps;
  switch (model) {
    case kGbm:
      paths_kernel<GbmSrc<false>, Gbm64, GbmConst, 8, true><<<grid, block, 0, st>>>(
          *(const GbmConst*)mconst, sd, n_rows, base_rows, st32, d_out);
      break;
    case kHeston:
      paths_kernel<HestonSrc<false>, Heston64, HestonConst, 8, true><<<grid, block, 0, st>>>(
          *(const HestonConst*)mconst, sd, n_rows, base_rows, st32, d_out);
      break;
    case kMerton:
      paths_kernel<MertonSrc<false, Merton64>, Merton64, MertonConst, 4><<<grid, block, 0, st>>>(
          *(const MertonConst*)mconst, sd, n_rows, base_rows, st32, d_out);
      break;
    case kHullWhite:
      paths_kernel<GbmSrc<false>, HullWhite64, HullWhiteConst, 8, true><<<grid, block, 0, st>>>(
          *(const HullWhiteConst*)mconst, sd, n_rows, base_rows, st32, d_out);
      break;
    case kHestonQe:
      paths_kernel<HestonQeSrc<HestonQe64Path>, HestonQe64Path, HestonQeConst, 4>
          <<<grid, block, 0, st>>>(*(const HestonQeConst*)mconst, sd, n_rows, base_rows, st32,
                                   d_out);
      break;
    default:
      return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  }
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  return STOCHA_OK;
}

}  // namespace stocha

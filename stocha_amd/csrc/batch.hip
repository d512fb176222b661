// batch.hip -- thousands of uneven pricing instances in one persistent
// launch.
//
// An instance is one European call on GBM, Heston (full-truncation Euler) or
// Merton paths with its own parameters, n_paths and steps, priced on the
// reference's block-split streams of `seed` (each instance sees exactly the
// paths the single-instance call would), FP32 log-space stepping, FP64 sums.
//
// Work is cut into tiles of at most `tile` consecutive paths of one instance.
// The host orders tiles by cost (steps x model weight, longest first) and a
// persistent grid -- sized to what is resident on the device -- pulls tile
// indices from a device counter.  Each tile's payoff sum and sum of squares
// go to its own slot, and a second kernel adds every instance's tile slots in
// tile order, so results do not depend on which block ran which tile.
// Within a tile every lane runs its own path on the queue schedule
// (drivers.hpp: per-lane ziggurat state machine, steps applied by the whole
// wave); the model switch is uniform over the block.
//
// The three launches (counter reset, persistent kernel, per-instance
// reduction) are captured once in a hipGraph and replayed; the counter is
// cleared by a kernel, not a memset node.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "host.hpp"
#include "drivers.hpp"

namespace stocha {

struct BatchInstDev {
  int model;
  uint32_t steps;
  uint32_t desc;  // index into the stream-descriptor table
  uint32_t pad;
  double strike;
  GbmConst g;
  HestonConst h;
  MertonConst m;
};

struct BatchTile {
  uint32_t inst;
  uint32_t count;
  uint64_t path0;
};

template <class T>
__device__ __forceinline__ T bwave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}


template <int kBlock>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchInstDev* __restrict__ inst,
                                                       const BatchTile* __restrict__ tiles,
                                                       const uint32_t n_tiles,
                                                       const StreamDesc* __restrict__ descs,
                                                       uint32_t* __restrict__ counter,
                                                       double* __restrict__ tile_sums) {
  __shared__ ZigLds32 zig;
  __shared__ uint32_t cur;
  __shared__ double red[kBlock / 64][2];
  // step-input rings: 64 bytes per lane (16 GBM/Heston normals, 8 Merton entries)
  __shared__ float2 qbuf[8][kBlock];
  zig_load(&zig);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) cur = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = cur;
    if (t >= n_tiles) break;  // every block exits once the queue is drained
    const BatchTile tl = tiles[t];
    const BatchInstDev& in = inst[tl.inst];
    const StreamDesc& sd = descs[in.desc];
    double s1 = 0.0, s2 = 0.0;
    const U128 inc = sd.inc;
    // every lane runs a path so that the queue schedule sees whole
    // wavefronts; only valid lanes count
    // path k + kBlock's first state is k's moved by one affine map (lanes
    // past the tile draw from their own unused streams)
    const Jump next = stream_stride(&sd, kBlock);
    U128 first = stream_at(&sd, tl.path0 + threadIdx.x);
    for (uint32_t j = 0; j * kBlock < tl.count; ++j) {
      const uint32_t k = j * kBlock + threadIdx.x;
      const bool valid = k < tl.count;
      Pcg g{first};
      first = add128(mul128(next.a, first), next.c);
      double st;
      if (in.model == kGbm)
        st = queue_path<GbmSrc<true>, Gbm32, 16, kBlock>(in.g, g, inc, in.steps, &zig,
                                                        (GbmSrc<true>::Rec(*)[kBlock])qbuf);
      else if (in.model == kHeston)
        st = queue_path<HestonSrc<true>, Heston32, 16, kBlock>(
            in.h, g, inc, in.steps, &zig, (HestonSrc<true>::Rec(*)[kBlock])qbuf);
      else
        st = queue_path<MertonSrc<true, Merton32>, Merton32, 8, kBlock>(
            in.m, g, inc, in.steps, &zig, (MertonSrc<true, Merton32>::Rec(*)[kBlock])qbuf);
      if (valid) {
        double pay = st - in.strike;
        pay = pay > 0.0 ? pay : 0.0;
        s1 += pay;
        s2 += pay * pay;
      }
    }
    s1 = bwave_sum(s1);
    s2 = bwave_sum(s2);
    if (lane == 0) {
      red[wave][0] = s1;
      red[wave][1] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
      tile_sums[2 * (uint64_t)t + threadIdx.x] = s;
    }
  }
}

__global__ void batch_reset_kernel(uint32_t* counter) { *counter = 0u; }

// out[2*i], out[2*i+1] = sum over instance i's tiles (in tile-id order)
__global__ __launch_bounds__(256) void batch_reduce_kernel(const uint32_t* __restrict__ first,
                                                           const uint32_t* __restrict__ tile_ids,
                                                           const uint32_t n_inst,
                                                           const double* __restrict__ tile_sums,
                                                           double* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_inst) return;
  double s1 = 0.0, s2 = 0.0;
  for (uint32_t j = first[i]; j < first[i + 1]; ++j) {
    const uint32_t t = tile_ids[j];
    s1 += tile_sums[2 * (uint64_t)t];
    s2 += tile_sums[2 * (uint64_t)t + 1];
  }
  out[2 * i] = s1;
  out[2 * i + 1] = s2;
}

// ------------------------------------------------------------------ host --
struct BatchPlan {
  uint32_t n_inst = 0, n_tiles = 0;
  int block = 256;
  uint32_t paths_per_lane = 1;
  unsigned grid = 0;
  DevBuf inst, tiles, descs, counter, tile_sums, first, tile_ids, out;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<double> host_out;
  std::vector<uint64_t> n_paths;
  std::vector<double> disc;
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static double model_weight(int model) {
  // relative cost per step: u64 draws per step of the model (1.03 / 2.06 / ~2.03)
  return model == kGbm ? 1.0 : model == kHeston ? 2.0 : 1.97;
}

template <int kBlock>
static void batch_enqueue(BatchPlan& P) {
  batch_reset_kernel<<<1, 1, 0, P.stream>>>(P.counter.as<uint32_t>());
  batch_kernel<kBlock><<<P.grid, kBlock, 0, P.stream>>>(
      P.inst.as<BatchInstDev>(), P.tiles.as<BatchTile>(), P.n_tiles, P.descs.as<StreamDesc>(),
      P.counter.as<uint32_t>(), P.tile_sums.as<double>());
  batch_reduce_kernel<<<(P.n_inst + 255) / 256, 256, 0, P.stream>>>(
      P.first.as<uint32_t>(), P.tile_ids.as<uint32_t>(), P.n_inst, P.tile_sums.as<double>(),
      P.out.as<double>());
}

template <int kBlock>
static int batch_occupancy(int* blocks_per_cu) {
  return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_kernel<kBlock>,
                                                           kBlock, 0);
}

int batch_plan_create(const stocha_batch_instance* in, uint64_t n, uint64_t seed, int block,
                      int paths_per_lane, int waves_per_simd, BatchPlan** out, const Err& err) {
  if (block != 64 && block != 128 && block != 256)
    return err.set(STOCHA_EINVAL, "block must be 64, 128 or 256");
  if (paths_per_lane < 1 || paths_per_lane > 16)
    return err.set(STOCHA_EINVAL, "paths_per_lane must be in [1, 16]");
  auto P = std::make_unique<BatchPlan>();
  P->n_inst = (uint32_t)n;
  P->block = block;
  P->paths_per_lane = (uint32_t)paths_per_lane;
  const uint64_t tile = (uint64_t)block * (uint64_t)paths_per_lane;

  // instances -> device constants; one stream descriptor per distinct block size
  std::vector<BatchInstDev> hi(n);
  std::vector<StreamDesc> hd;
  std::vector<uint64_t> blocks;
  P->n_paths.resize(n);
  P->disc.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    const stocha_batch_instance& b = in[i];
    BatchInstDev& d = hi[i];
    memset(&d, 0, sizeof d);
    d.model = b.model;
    d.steps = b.steps;
    d.strike = b.strike;
    if (b.model == kGbm) {
      d.g = gbm_const(b.s0, b.mu, b.sigma, b.t, b.steps, 0.0);
    } else if (b.model == kHeston) {
      const double p[7] = {b.s0, b.v0, b.mu, b.kappa, b.theta, b.xi, b.rho};
      d.h = heston_const(p, b.t, b.steps);
    } else {
      const double p[6] = {b.s0, b.mu, b.sigma, b.lambda_, b.mu_j, b.sigma_j};
      d.m = merton_const(p, b.t, b.steps);
    }
    const uint64_t blk = block_size(b.model, b.steps);
    auto it = std::find(blocks.begin(), blocks.end(), blk);
    if (it == blocks.end()) {
      blocks.push_back(blk);
      hd.emplace_back();
      make_stream_desc(seed, blk, &hd.back());
      d.desc = (uint32_t)(hd.size() - 1);
    } else {
      d.desc = (uint32_t)(it - blocks.begin());
    }
    P->n_paths[i] = b.n_paths;
    P->disc[i] = exp(-b.r * b.t);
  }

  // tiles, longest first
  std::vector<BatchTile> ht;
  std::vector<double> cost;
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t p0 = 0; p0 < in[i].n_paths; p0 += tile) {
      BatchTile t;
      t.inst = (uint32_t)i;
      t.path0 = p0;
      t.count = (uint32_t)std::min<uint64_t>(tile, in[i].n_paths - p0);
      ht.push_back(t);
      // a tile costs one path per lane per round: steps x draws/step x rounds
      cost.push_back((double)in[i].steps * model_weight(in[i].model) *
                     std::ceil((double)t.count / (double)block));
    }
  std::vector<uint32_t> order(ht.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<BatchTile> sorted(ht.size());
  for (size_t k = 0; k < order.size(); ++k) sorted[k] = ht[order[k]];
  P->n_tiles = (uint32_t)sorted.size();
  // CSR: each instance's tiles in ascending tile id
  std::vector<uint32_t> first(n + 1, 0), ids(sorted.size());
  for (const BatchTile& t : sorted) first[t.inst + 1]++;
  for (uint64_t i = 0; i < n; ++i) first[i + 1] += first[i];
  {
    std::vector<uint32_t> fill(first.begin(), first.end() - 1);
    for (uint32_t k = 0; k < sorted.size(); ++k) ids[fill[sorted[k].inst]++] = k;
  }

  if (int rc = P->inst.alloc(sizeof(BatchInstDev) * n, err)) return rc;
  if (int rc = P->tiles.alloc(sizeof(BatchTile) * sorted.size(), err)) return rc;
  if (int rc = P->descs.alloc(sizeof(StreamDesc) * hd.size(), err)) return rc;
  if (int rc = P->counter.alloc(sizeof(uint32_t), err)) return rc;
  if (int rc = P->tile_sums.alloc(sizeof(double) * 2 * sorted.size(), err)) return rc;
  if (int rc = P->first.alloc(sizeof(uint32_t) * (n + 1), err)) return rc;
  if (int rc = P->tile_ids.alloc(sizeof(uint32_t) * ids.size(), err)) return rc;
  if (int rc = P->out.alloc(sizeof(double) * 2 * n, err)) return rc;
  STOCHA_HIP(hipMemcpy(P->inst.p, hi.data(), sizeof(BatchInstDev) * n, hipMemcpyHostToDevice));
  STOCHA_HIP(hipMemcpy(P->tiles.p, sorted.data(), sizeof(BatchTile) * sorted.size(),
                       hipMemcpyHostToDevice));
  STOCHA_HIP(hipMemcpy(P->descs.p, hd.data(), sizeof(StreamDesc) * hd.size(),
                       hipMemcpyHostToDevice));
  STOCHA_HIP(hipMemcpy(P->first.p, first.data(), sizeof(uint32_t) * (n + 1),
                       hipMemcpyHostToDevice));
  STOCHA_HIP(hipMemcpy(P->tile_ids.p, ids.data(), sizeof(uint32_t) * ids.size(),
                       hipMemcpyHostToDevice));

  // persistent grid: resident blocks per CU (capped by the requested waves/SIMD)
  int dev = 0, cus = 0, bpc = 0;
  STOCHA_HIP(hipGetDevice(&dev));
  STOCHA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  hipError_t oe = block == 64 ? (hipError_t)batch_occupancy<64>(&bpc)
                  : block == 128 ? (hipError_t)batch_occupancy<128>(&bpc)
                                 : (hipError_t)batch_occupancy<256>(&bpc);
  STOCHA_HIP(oe);
  const int waves_per_block = block / 64;
  int want = waves_per_simd > 0 ? (4 * waves_per_simd) / waves_per_block : bpc;
  if (want < 1) want = 1;
  if (want > bpc) want = bpc;
  P->grid = (unsigned)std::max(1, cus * want);
  if (P->grid > P->n_tiles) P->grid = P->n_tiles;

  // capture the three launches into a graph on a private stream
  STOCHA_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  STOCHA_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
  if (block == 64)
    batch_enqueue<64>(*P);
  else if (block == 128)
    batch_enqueue<128>(*P);
  else
    batch_enqueue<256>(*P);
  STOCHA_HIP(hipStreamEndCapture(P->stream, &P->graph));
  STOCHA_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
  P->host_out.resize(2 * n);
  *out = P.release();
  return STOCHA_OK;
}

int batch_plan_run(BatchPlan* P, double* price, double* std_err, const Err& err) {
  KernelTimer timer(P->stream);
  timer.start();
  STOCHA_HIP(hipGraphLaunch(P->exec, P->stream));
  timer.stop(1);
  STOCHA_HIP(hipMemcpyAsync(P->host_out.data(), P->out.p, sizeof(double) * 2 * P->n_inst,
                            hipMemcpyDeviceToHost, P->stream));
  STOCHA_HIP(hipStreamSynchronize(P->stream));
  for (uint32_t i = 0; i < P->n_inst; ++i) {
    const double n = (double)P->n_paths[i];
    const double s1 = P->host_out[2 * i], s2 = P->host_out[2 * i + 1];
    price[i] = P->disc[i] * s1 / n;
    if (std_err) {
      double var = n > 1 ? (s2 - s1 * s1 / n) / (n - 1.0) : 0.0;
      var = var > 0.0 ? var : 0.0;
      std_err[i] = P->disc[i] * sqrt(var / n);
    }
  }
  return STOCHA_OK;
}

void batch_plan_destroy(BatchPlan* P) { delete P; }

uint32_t batch_plan_tiles(const BatchPlan* P) { return P->n_tiles; }
unsigned batch_plan_grid(const BatchPlan* P) { return P->grid; }

}  // namespace stocha

// basket.hip -- correlated multi-asset GBM priced in-register: a weighted
// basket call max(sum_a w_a S_T,a - K, 0) on the terminal prices, with the
// pathwise adjoint of the price with respect to every asset's s0, mu, sigma.
//
// Streams and draws are multi_gbm's (multi_gbm.rs:40-64): path i uses the
// seed's stream advanced by i*(steps*d + 1024)*2 and draws d normals per step,
// asset-minor.  The payoff only needs terminal values, and with a constant
// Cholesky factor L
//     log S_T,a = log s0_a + steps*drift_a + diffusion_a * (L W)_a,
//     W = sum over steps of z_t,
// so a path keeps the running sum W (d values in registers) and forms E = L W
// once at the end (VALU dot products with L read through the scalar cache;
// or, in the FP32 form, v_mfma_f32_16x16x4f32 tiles over a wave's 64
// paths).  This equals the reference's per-step L z_t products in exact
// arithmetic; rounding differs by the reassociation only.  Both precisions
// run in two kernels: the draws, then payoff and adjoint from W kept
// asset-major in the workspace.
//
// Adjoint of one path (reverse mode through the payoff):
//   B = sum_a w_a S_a,  Bbar = 1{B > K},  Sbar_a = w_a Bbar
//   dP/ds0_a    = Sbar_a S_a / s0_a
//   dP/dmu_a    = Sbar_a S_a T
//   dP/dsigma_a = Sbar_a S_a (-sigma_a T + sqrt(dt) E_a)
// (for d = 1 these are greeks.rs:388-418's delta and vega).  The backward
// pass of the autograd wrapper replays the streams with this kernel instead
// of storing paths.
//
// The draws run on the queue schedule (drivers.hpp): each lane's steps*d
// normals pass through a per-lane ring, and the wavefront takes them out in
// chunks of the same assets on every lane.
#include "host.hpp"
#include "drivers.hpp"

namespace stocha {

constexpr int kMaxAssets = 64;

struct BasketDev {
  uint32_t d, steps;
  uint64_t n_paths;
  double strike, t, sqrt_dt;
  const double* L;      // d x d row-major (FP64)
  const float* Lf;      // 64 x 64 row-major FP32, zero outside the d x d factor
  const double* s0;     // d
  const double* lmean;  // d: steps * drift_a (log-drift over the horizon)
  const double* diff;   // d: sigma_a * sqrt(dt)
  const double* sigma;  // d
  const double* w;      // d
};

template <class T>
__device__ __forceinline__ T wave_sum_t(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Payoff and adjoint from the MFMA accumulators acc[mt][nt] (lane l: E of
// path 16*mt + 4*(l/16) + r, asset 16*nt + l%16).
typedef float f4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void basket_mfma_payoff(const BasketDev& bd, const f4_t (&acc)[4][4],
                                                   uint64_t wave_path0, double& psum, double& psq,
                                                   double (&gd)[4], double (&gm)[4],
                                                   double (&gv)[4], bool greeks) {
  const uint32_t d = bd.d;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int nnt = (int)((d + 15) / 16);
  // S_T per (path, asset) held by this lane, basket partial per path
  float bp[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) bp[mt][r] = 0.f;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const uint32_t a = 16 * nt + li;
    if (nt < nnt && a < d) {
      const float s0 = (float)bd.s0[a], lm = (float)bd.lmean[a], df = (float)bd.diff[a];
      const float w = (float)bd.w[a];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float st = s0 * __expf(fmaf(df, acc[mt][nt][r], lm));
          bp[mt][r] = fmaf(w, st, bp[mt][r]);
        }
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = bp[mt][r];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      bp[mt][r] = v;  // basket value of path 16*mt + 4*lg + r (all 16 lanes)
    }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint64_t path = wave_path0 + 16 * mt + 4 * lg + r;
      const bool valid = path < bd.n_paths;
      const double B = (double)bp[mt][r];
      const bool itm = valid && B > bd.strike;
      if (li == 0 && itm) {
        const double pay = B - bd.strike;
        psum += pay;
        psq += pay * pay;
      }
      if (greeks && itm) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
          const uint32_t a = 16 * nt + li;
          if (nt < nnt && a < d) {
            const float e = acc[mt][nt][r];
            const double st = (double)((float)bd.s0[a] *
                                       __expf(fmaf((float)bd.diff[a], e, (float)bd.lmean[a])));
            const double sb = bd.w[a] * st;
            gd[nt] += sb / bd.s0[a];
            gm[nt] += sb * bd.t;
            gv[nt] += sb * (-bd.sigma[a] * bd.t + bd.sqrt_dt * e);
          }
        }
      }
    }
}

// FP32 epilogue on the matrix cores, for one wavefront's 64 paths at once.
// E (paths x assets) = W^T L^T with v_mfma_f32_16x16x4f32: A = W^T tile
// (lane l: path 16*mt + l%16, asset 4*kc + l/16), B = L^T tile (lane l:
// L[16*nt + l%16][4*kc + l/16]), C lane l: path 16*mt + 4*(l/16) + r,
// asset 16*nt + l%16.  k-chunks above the diagonal block are skipped (L is
// lower triangular).  The basket sum of a path is then a reduction over the
// 16 lanes that share l/16, and each lane keeps the Greek sums of its four
// assets (16*nt + l%16) across all paths.
// W(k, path) gives asset k's sum of the path (an LDS column or the
// asset-major workspace).
template <class WAcc>
__device__ __forceinline__ void basket_mfma_epilogue_w(const BasketDev& bd, const WAcc& W,
                                                       uint64_t wave_path0, double& psum,
                                                       double& psq, double (&gd)[4],
                                                       double (&gm)[4], double (&gv)[4],
                                                       bool greeks) {
  const uint32_t d = bd.d;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int nk = (int)((d + 3) / 4), nnt = (int)((d + 15) / 16);
  f4_t acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    if (nt < nnt) {
      const int kend = nk < 4 * nt + 4 ? nk : 4 * nt + 4;
      for (int kc = 0; kc < kend; ++kc) {
        const int k = 4 * kc + lg;
        const float b = bd.Lf[(16 * nt + li) * 64 + k];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          const float a = (uint32_t)k < d ? W(k, wave_path0 + 16 * mt + li) : 0.f;
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[mt][nt], 0, 0, 0);
        }
      }
    }
  }
  basket_mfma_payoff(bd, acc, wave_path0, psum, psq, gd, gm, gv, greeks);
}

// ------------------------------------------------------------- two phases --
// Phase 1: the draws, on the queue schedule (drivers.hpp): one path per lane,
// the lanes' normals pass through per-lane rings in LDS and the wave takes
// them out in chunks -- chunk c of a step is assets kChunk*c .. kChunk*c+n-1,
// n = min(kChunk, d - kChunk*c) -- so that the asset of every normal taken is
// the same on all lanes of the wave.  W (the running sum of each asset's
// normals) then lives in registers, indexed by constants under a
// wave-uniform branch on the chunk; LDS holds only the tables and the rings
// (kSumsSlots entries per path), so occupancy is set by the registers (W is
// kChunk*NC values).  W goes to ws[a * np + p] (asset-major, so both these
// stores and the loads of phase 2 are coalesced).  Phase 2, the payoff and
// adjoint, is a separate kernel so that none of its LDS or registers count
// against the draws.  Both run over slices of at most kSlicePaths paths, which
// bounds the workspace at kSlicePaths*d values.
constexpr int kSumsBlock = 256;
constexpr int kChunk = 8;
// Ring entries per lane.  FP32: 32 lets lanes that draw ahead bank normals
// across chunks (16 costs 6% more draw-loop passes and 45% more slow-block
// passes); FP64: 16, which keeps a 256-thread block's LDS at 51 KB.
template <bool kFp32, int NC>
constexpr uint32_t kSumsSlots = kFp32 ? 32 : 16;
constexpr int kRegBlock = 256;
constexpr uint64_t kSlicePaths = 1ull << 22;
constexpr uint64_t kPayoffBlocks = 4096;  // phase-2 grid cap (grid-stride beyond)

// waves per SIMD the phase-1 register budget allows (W is kChunk*NC values,
// two registers each in FP64, beside the draw loop's own)
template <bool kFp32, int NC>
constexpr int sums_waves() {
  return kFp32 ? (NC > 4 ? 4 : (NC > 2 ? 5 : 6)) : (NC > 4 ? 2 : 3);
}

// Paths p0 .. p0+np-1 of the launch (np = bd.n_paths of the slice); lane
// p - p0 of the grid runs path p.
template <bool kFp32, int NC>
__global__ __launch_bounds__(kSumsBlock) __attribute__((amdgpu_waves_per_eu(
    sums_waves<kFp32, NC>()))) void basket_sums_kernel(const BasketDev bd, const StreamDesc sd,
                                                       const uint64_t p0,
                                                       typename ZigWave<kFp32>::Real* __restrict__
                                                           ws) {
  using Src = GbmSrc<kFp32>;
  using Real = typename Src::Real;
  __shared__ typename Src::Tables zig;
  constexpr uint32_t kSlots = kSumsSlots<kFp32, NC>;
  __shared__ Real q[kSlots][kSumsBlock];
  zig_load(&zig);
  __syncthreads();
  const uint64_t np = bd.n_paths;
  // whole wavefronts past the last path leave (wave-uniform; no barrier below)
  if ((uint64_t)blockIdx.x * kSumsBlock + (threadIdx.x & ~63u) >= np) return;
  const uint64_t p = (uint64_t)blockIdx.x * kSumsBlock + threadIdx.x;
  const bool valid = p < np;
  Src src(bd);
  Pcg g{stream_at(&sd, p0 + (valid ? p : 0))};  // lanes past the end replay a path, unstored
  const LcgInc incv = vgpr_increment(sd.inc);
  const uint32_t d = bd.d;
  const uint32_t nc = (d + kChunk - 1) / kChunk;  // <= NC
  Real W[kChunk * NC];
#pragma unroll
  for (int i = 0; i < kChunk * NC; ++i) W[i] = (Real)0;
  uint32_t head = 0, cnt = 0;
  for (uint32_t step = 0; step < bd.steps; ++step) {
    for (uint32_t c = 0; c < nc; ++c) {
      const uint32_t n = min((uint32_t)kChunk, d - kChunk * c);
      queue_fill<Src, kSlots, kSumsBlock>(src, g, incv, &zig, q, head, cnt, n);
      // all kChunk ring reads are issued before the first add (entries past
      // n are read and dropped); a whole chunk from a chunk-aligned head
      // reads kChunk consecutive rows at constant offsets
      Real z[kChunk];
      const bool whole = n == (uint32_t)kChunk && (head & (kChunk - 1)) == 0;
      if (whole) {
        const Real* row = &q[head][threadIdx.x];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) z[j] = row[j * kSumsBlock];
      } else {
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const Real v = q[(head + j) & (kSlots - 1)][threadIdx.x];
          z[j] = (uint32_t)j < n ? v : (Real)0;
        }
      }
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) {
        if (cc == (int)c) {
#pragma unroll
          for (int j = 0; j < kChunk; ++j) W[kChunk * cc + j] += z[j];
        }
      }
      head = (head + n) & (kSlots - 1);
      cnt -= n;
    }
  }
  if (!valid) return;
#pragma unroll
  for (int a = 0; a < kChunk * NC; ++a)
    if ((uint32_t)a < d) ws[(uint64_t)a * np + p] = W[a];
}

// Phase 2: payoff and adjoint from W (ws, asset-major over the slice's np =
// bd.n_paths paths).  Each wave takes 64 consecutive paths, grid-stride.
// kMfma (FP32): E = W^T L^T on v_mfma_f32_16x16x4f32 (basket_mfma_epilogue_w,
// A operands loaded straight from ws); otherwise per lane
// E_i = sum_{k<=i} L[i][k] W[k] on the VALU with L through the scalar cache
// (FP32: fmaf over the FP32 factor; FP64: products then sums, in k order).
// accumulator layout per block: [0] sum payoff, [1] sum payoff^2,
// then d x {dP/ds0, dP/dmu, dP/dsigma}
template <class Real>
struct GlobalW {
  const Real* ws;
  uint64_t n;
  __device__ __forceinline__ Real operator()(int k, uint64_t path) const {
    return path < n ? ws[(uint64_t)k * n + path] : (Real)0;
  }
};

template <class Real>
__device__ __forceinline__ Real basket_e(const BasketDev& bd, const GlobalW<Real>& W, uint32_t i,
                                         uint64_t p) {
  Real e = (Real)0;
  for (uint32_t k = 0; k <= i; ++k) {
    if constexpr (std::is_same<Real, float>::value)
      e = fmaf(bd.Lf[i * 64 + k], W((int)k, p), e);
    else
      e += bd.L[i * bd.d + k] * W((int)k, p);
  }
  return e;
}

template <class Real, bool kGreeks, bool kMfma>
__global__ __launch_bounds__(kRegBlock) void basket_payoff_kernel(const BasketDev bd,
                                                                  const Real* __restrict__ ws,
                                                                  double* __restrict__ partials) {
  static_assert(!kMfma || std::is_same<Real, float>::value, "the matrix-core epilogue is FP32");
  __shared__ double red[kRegBlock / 64][2 + 3 * kMaxAssets];
  const uint32_t d = bd.d;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const GlobalW<Real> W{ws, bd.n_paths};
  double psum = 0.0, psq = 0.0;
  double gd = 0.0, gm = 0.0, gv = 0.0;  // VALU: lane a owns asset a's sums
  double md[4] = {0, 0, 0, 0}, mm[4] = {0, 0, 0, 0}, mv[4] = {0, 0, 0, 0};  // MFMA
  for (uint64_t wave_path0 = (uint64_t)blockIdx.x * kRegBlock + (threadIdx.x & ~63u);
       wave_path0 < bd.n_paths; wave_path0 += (uint64_t)gridDim.x * kRegBlock) {
    if constexpr (kMfma) {
      basket_mfma_epilogue_w(bd, W, wave_path0, psum, psq, md, mm, mv, kGreeks);
    } else {
      const uint64_t p = wave_path0 + lane;
      const bool valid = p < bd.n_paths;
      double B = 0.0;
#pragma unroll 1
      for (uint32_t i = 0; i < d; ++i) {
        const Real e = basket_e(bd, W, i, p);
        B += bd.w[i] * (bd.s0[i] * exp(bd.lmean[i] + bd.diff[i] * (double)e));
      }
      const bool itm = valid && B > bd.strike;
      const double pay = itm ? B - bd.strike : 0.0;
      psum += pay;
      psq += pay * pay;
      if constexpr (kGreeks) {
#pragma unroll 1
        for (uint32_t i = 0; i < d; ++i) {
          double cd = 0.0, cm = 0.0, cv = 0.0;
          if (itm) {
            const Real e = basket_e(bd, W, i, p);
            const double sb = bd.w[i] * (bd.s0[i] * exp(bd.lmean[i] + bd.diff[i] * (double)e));
            cd = sb / bd.s0[i];
            cm = sb * bd.t;
            cv = sb * (-bd.sigma[i] * bd.t + bd.sqrt_dt * (double)e);
          }
          cd = wave_sum_t(cd);
          cm = wave_sum_t(cm);
          cv = wave_sum_t(cv);
          if ((uint32_t)lane == i) {
            gd += cd;
            gm += cm;
            gv += cv;
          }
        }
      }
    }
  }
  if constexpr (kMfma && kGreeks) {
    // lanes l, l^16, l^32, l^48 hold the same four assets: fold them, then
    // lane li (< 16) owns assets li, 16+li, 32+li, 48+li
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      md[nt] += __shfl_xor(md[nt], 16, 64);
      md[nt] += __shfl_xor(md[nt], 32, 64);
      mm[nt] += __shfl_xor(mm[nt], 16, 64);
      mm[nt] += __shfl_xor(mm[nt], 32, 64);
      mv[nt] += __shfl_xor(mv[nt], 16, 64);
      mv[nt] += __shfl_xor(mv[nt], 32, 64);
    }
  }
  psum = wave_sum_t(psum);
  psq = wave_sum_t(psq);
  if (lane == 0) {
    red[wave][0] = psum;
    red[wave][1] = psq;
  }
  if constexpr (kGreeks && !kMfma) {
    if ((uint32_t)lane < d) {
      red[wave][2 + 3 * lane] = gd;
      red[wave][3 + 3 * lane] = gm;
      red[wave][4 + 3 * lane] = gv;
    }
  }
  if constexpr (kGreeks && kMfma) {
    if (lane < 16) {
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const uint32_t a = 16 * nt + lane;
        if (a < d) {
          red[wave][2 + 3 * a] = md[nt];
          red[wave][3 + 3 * a] = mm[nt];
          red[wave][4 + 3 * a] = mv[nt];
        }
      }
    }
  }
  __syncthreads();
  const int na = kGreeks ? 2 + 3 * (int)d : 2;
  for (int i = threadIdx.x; i < na; i += kRegBlock) {
    double sum = 0.0;
    for (int wv = 0; wv < kRegBlock / 64; ++wv) sum += red[wv][i];
    partials[(uint64_t)blockIdx.x * na + i] = sum;
  }
}

template <bool kFp32>
static void launch_sums(const BasketDev& bd, const StreamDesc& sd, uint64_t p0,
                        typename ZigWave<kFp32>::Real* ws, hipStream_t st) {
  const unsigned nb = (unsigned)((bd.n_paths + kSumsBlock - 1) / kSumsBlock);
  const uint32_t nc = (bd.d + kChunk - 1) / kChunk;
  if (nc <= 1)
    basket_sums_kernel<kFp32, 1><<<nb, kSumsBlock, 0, st>>>(bd, sd, p0, ws);
  else if (nc <= 2)
    basket_sums_kernel<kFp32, 2><<<nb, kSumsBlock, 0, st>>>(bd, sd, p0, ws);
  else if (nc <= 4)
    basket_sums_kernel<kFp32, 4><<<nb, kSumsBlock, 0, st>>>(bd, sd, p0, ws);
  else
    basket_sums_kernel<kFp32, 8><<<nb, kSumsBlock, 0, st>>>(bd, sd, p0, ws);
}

// Both phases over slices of the paths; per-block partial sums of all
// slices are reduced together at the end.
template <bool kFp32, bool kMfma>
static int basket_two_phase(const BasketDev& bd, const StreamDesc& sd, bool greeks,
                            double* h_out, int na, hipStream_t st, const Err& err) {
  using Real = typename ZigWave<kFp32>::Real;
  const uint64_t slice = bd.n_paths < kSlicePaths ? bd.n_paths : kSlicePaths;
  const uint64_t n_slices = (bd.n_paths + slice - 1) / slice;
  uint64_t nbp = (slice + kRegBlock - 1) / kRegBlock;  // phase-2 blocks per slice
  if (nbp > kPayoffBlocks) nbp = kPayoffBlocks;
  const uint64_t n_part = n_slices * nbp;
  if (n_part > 0x7fffffffull) return err.set(STOCHA_EINVAL, "n_paths too large for one call");
  Workspace* wsp = nullptr;
  const size_t w_bytes = ((sizeof(Real) * bd.d * slice + 255) / 256) * 256;
  const size_t part_bytes = sizeof(double) * n_part * na;
  if (int rc = workspace(w_bytes + part_bytes + sizeof(double) * na, sizeof(double) * na, &wsp,
                         err))
    return rc;
  Real* ws = static_cast<Real*>(wsp->dev);
  double* part = reinterpret_cast<double*>(static_cast<char*>(wsp->dev) + w_bytes);
  double* out = part + n_part * na;
  KernelTimer timer(st);
  timer.start();
  for (uint64_t s = 0; s < n_slices; ++s) {
    BasketDev bs = bd;
    const uint64_t p0 = s * slice;
    bs.n_paths = bd.n_paths - p0 < slice ? bd.n_paths - p0 : slice;
    launch_sums<kFp32>(bs, sd, p0, ws, st);
    STOCHA_HIP(hipGetLastError());
    double* ps = part + s * nbp * na;
    if (greeks)
      basket_payoff_kernel<Real, true, kMfma><<<(unsigned)nbp, kRegBlock, 0, st>>>(bs, ws, ps);
    else
      basket_payoff_kernel<Real, false, kMfma><<<(unsigned)nbp, kRegBlock, 0, st>>>(bs, ws, ps);
    STOCHA_HIP(hipGetLastError());
  }
  timer.stop(2 * n_slices);
  if (int rc = launch_reduce_partials(part, (int)n_part, na, out, st, err)) return rc;
  STOCHA_HIP(hipMemcpyAsync(wsp->host, out, sizeof(double) * na, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < na; ++i) h_out[i] = wsp->host[i];
  return STOCHA_OK;
}

int run_basket(const double* s0, const double* mu, const double* sigma, const double* L,
               const double* w, uint64_t d, double strike, double t, uint64_t steps,
               uint64_t n_paths, uint64_t seed, int precision, bool greeks, bool mfma,
               double* h_sums, hipStream_t st, const Err& err) {
  const double dt = t / (double)steps;
  const double sqrt_dt = sqrt(dt);
  // device copy of the per-asset constants: L (f64), L (f32), s0, lmean, diff, sigma, w
  const size_t nd = (size_t)d;
  std::vector<double> hd(nd * nd + 5 * nd);
  std::vector<float> hf(64 * 64, 0.f);
  for (size_t i = 0; i < nd; ++i)
    for (size_t k = 0; k < nd; ++k) {
      hd[i * nd + k] = L[i * nd + k];
      hf[i * 64 + k] = (float)L[i * nd + k];
    }
  double* c = hd.data() + nd * nd;
  for (size_t a = 0; a < nd; ++a) {
    const double drift = (mu[a] - 0.5 * sigma[a] * sigma[a]) * dt;  // multi_gbm.rs:29-31
    c[a] = s0[a];
    c[nd + a] = (double)steps * drift;
    c[2 * nd + a] = sigma[a] * sqrt_dt;  // multi_gbm.rs:32
    c[3 * nd + a] = sigma[a];
    c[4 * nd + a] = w[a];
  }
  DevBuf dd, df;
  if (int rc = dd.alloc(sizeof(double) * hd.size(), err)) return rc;
  if (int rc = df.alloc(sizeof(float) * hf.size(), err)) return rc;
  STOCHA_HIP(hipMemcpyAsync(dd.p, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, st));
  STOCHA_HIP(hipMemcpyAsync(df.p, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice, st));
  BasketDev bd;
  bd.d = (uint32_t)d;
  bd.steps = (uint32_t)steps;
  bd.n_paths = n_paths;
  bd.strike = strike;
  bd.t = t;
  bd.sqrt_dt = sqrt_dt;
  bd.L = dd.as<double>();
  bd.Lf = df.as<float>();
  const double* dc = dd.as<double>() + nd * nd;
  bd.s0 = dc;
  bd.lmean = dc + nd;
  bd.diff = dc + 2 * nd;
  bd.sigma = dc + 3 * nd;
  bd.w = dc + 4 * nd;
  StreamDesc sd;
  make_stream_desc(seed, (steps * d + 1024) * 2, &sd);  // multi_gbm.rs:40
  const int na = greeks ? 2 + 3 * (int)d : 2;
  if (precision == STOCHA_FP32) {
    if (mfma) return basket_two_phase<true, true>(bd, sd, greeks, h_sums, na, st, err);
    return basket_two_phase<true, false>(bd, sd, greeks, h_sums, na, st, err);
  }
  return basket_two_phase<false, false>(bd, sd, greeks, h_sums, na, st, err);
}

}  // namespace stocha

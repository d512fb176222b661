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
// so a path keeps the running sum W (one LDS column per lane, d entries) and
// forms E = L W once at the end (VALU dot products with L read through the
// scalar cache; or, in the FP32 form, v_mfma_f32_16x16x4f32 tiles over a
// wave's 64 paths).  This equals the reference's per-step L z_t products in
// exact arithmetic; rounding differs by the reassociation only.  The FP32
// form runs in two kernels (the draws, then payoff and adjoint from W kept
// asset-major in the workspace); the FP64 form in one.
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
// Each lane runs the flat draw-feed state machine (zigsm.hpp) over its
// path's steps*d normals; lanes re-synchronise only at path boundaries.
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

// accumulator layout per block: [0] sum payoff, [1] sum payoff^2,
// then d x {dP/ds0, dP/dmu, dP/dsigma}
template <bool kFp32, int kBlock, bool kGreeks, bool kMfma>
__global__ __launch_bounds__(kBlock) void basket_kernel(const BasketDev bd, const StreamDesc sd,
                                                        double* __restrict__ partials) {
  static_assert(!kMfma || kFp32, "the matrix-core epilogue is FP32");
  using Real = typename std::conditional<kFp32, float, double>::type;
  __shared__ ZigLds zig;
  __shared__ Real W[kMaxAssets][kBlock];
  __shared__ double red[kBlock / 64][2 + 3 * kMaxAssets];
  zig_load(&zig);
  __syncthreads();

  const uint32_t d = bd.d;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double psum = 0.0, psq = 0.0;
  double gd = 0.0, gm = 0.0, gv = 0.0;  // VALU epilogue: lane a owns asset a's sums
  double md[4] = {0, 0, 0, 0}, mm[4] = {0, 0, 0, 0}, mv[4] = {0, 0, 0, 0};  // MFMA epilogue
  const U128 inc = sd.inc;
  const uint64_t total = (uint64_t)bd.steps * d;

  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < bd.n_paths;
       base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < bd.n_paths;
    for (uint32_t a = 0; a < d; ++a) W[a][threadIdx.x] = (Real)0;
    Pcg g{stream_at(&sd, valid ? p : 0)};
    ZigFeed<kFp32> zf;
    uint32_t a = 0;
    for (uint64_t k = 0; k < total;) {
      const uint64_t u = g.next(inc);
      Real x;
      if (zf.feed(u, &zig, x)) {
        W[a][threadIdx.x] += x;
        a = (a + 1 == d) ? 0 : a + 1;
        ++k;
      }
    }
    if constexpr (kMfma) {
      const uint64_t wp0 = base + (threadIdx.x & ~63);
      const auto lds_w = [&](int k, uint64_t path) { return W[k][(int)(path - base)]; };
      basket_mfma_epilogue_w(bd, lds_w, wp0, psum, psq, md, mm, mv,
                                   kGreeks);
      continue;
    }
    // epilogue: E = L W, S_T, basket payoff (per lane)
    double B = 0.0;
    for (uint32_t i = 0; i < d; ++i) {
      Real e = (Real)0;
      for (uint32_t k = 0; k <= i; ++k) {
        if constexpr (kFp32)
          e = fmaf(bd.Lf[i * 64 + k], W[k][threadIdx.x], e);
        else
          e += bd.L[i * d + k] * W[k][threadIdx.x];
      }
      const double s = bd.s0[i] * exp(bd.lmean[i] + bd.diff[i] * (double)e);
      B += bd.w[i] * s;
    }
    const double pay = valid && B > bd.strike ? B - bd.strike : 0.0;
    psum += pay;
    psq += pay * pay;
    if constexpr (kGreeks) {
      const bool itm = valid && B > bd.strike;
      for (uint32_t i = 0; i < d; ++i) {
        double cd = 0.0, cm = 0.0, cv = 0.0;
        if (itm) {
          Real e = (Real)0;
          for (uint32_t k = 0; k <= i; ++k) {
            if constexpr (kFp32)
              e = fmaf(bd.Lf[i * 64 + k], W[k][threadIdx.x], e);
            else
              e += bd.L[i * d + k] * W[k][threadIdx.x];
          }
          const double s = bd.s0[i] * exp(bd.lmean[i] + bd.diff[i] * (double)e);
          const double sb = bd.w[i] * s;  // Sbar_a * S_a
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
  // block reduction: price sums over all lanes, greek sums over waves
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
  for (int i = threadIdx.x; i < na; i += kBlock) {
    double s = 0.0;
    for (int wv = 0; wv < kBlock / 64; ++wv) s += red[wv][i];
    partials[(uint64_t)blockIdx.x * na + i] = s;
  }
}

// ------------------------------------------------------- FP32, two phases --
// Phase 1 (FP32): the draws.  One path per thread, the per-lane draw-feed
// loop over its steps*d normals with W (the running sum of the normals of
// each asset) in an LDS column; W then goes to ws[a * n + p] (asset-major,
// so both these stores and the loads of phase 2 are coalesced).  W takes
// 256 bytes of LDS per path, so a 512-thread block (146 KB with the tables)
// is the most a CU holds: the payoff and adjoint run as a separate kernel
// so that none of their LDS or registers count against this one.
constexpr int kSumsBlock = 512;
constexpr int kRegBlock = 256;

__global__ __launch_bounds__(kSumsBlock) void basket_sums_kernel(const BasketDev bd,
                                                                 const StreamDesc sd,
                                                                 float* __restrict__ ws) {
  __shared__ ZigLds zig;
  __shared__ float W[kMaxAssets][kSumsBlock];
  zig_load(&zig);
  const uint32_t d = bd.d;
  for (uint32_t a = 0; a < d; ++a) W[a][threadIdx.x] = 0.f;
  __syncthreads();
  const uint64_t p = (uint64_t)blockIdx.x * kSumsBlock + threadIdx.x;
  if (p >= bd.n_paths) return;  // no block barrier below
  Pcg g{stream_at(&sd, p)};
  const U128 inc = sd.inc;
  ZigFeed<true> zf;
  const uint64_t total = (uint64_t)bd.steps * d;
  uint32_t a = 0;
  for (uint64_t k = 0; k < total;) {
    const uint64_t u = g.next(inc);
    float x;
    if (zf.feed(u, &zig, x)) {
      W[a][threadIdx.x] += x;
      a = (a + 1 == d) ? 0 : a + 1;
      ++k;
    }
  }
  for (uint32_t k2 = 0; k2 < d; ++k2) ws[(uint64_t)k2 * bd.n_paths + p] = W[k2][threadIdx.x];
}

// Phase 2 (FP32): payoff and adjoint from W (ws, asset-major).  Each wave
// takes 64 consecutive paths.  kMfma: E = W^T L^T on v_mfma_f32_16x16x4f32
// (basket_mfma_epilogue, A operands loaded straight from ws); otherwise per
// lane E_i = sum_k L[i][k] W[k] on the VALU with L through the scalar cache.
struct GlobalW {
  const float* ws;
  uint64_t n;
  __device__ __forceinline__ float operator()(int k, uint64_t path) const {
    return path < n ? ws[(uint64_t)k * n + path] : 0.f;
  }
};

template <bool kGreeks, bool kMfma>
__global__ __launch_bounds__(kRegBlock) void basket_payoff_kernel(const BasketDev bd,
                                                                  const float* __restrict__ ws,
                                                                  double* __restrict__ partials) {
  __shared__ double red[kRegBlock / 64][2 + 3 * kMaxAssets];
  const uint32_t d = bd.d;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t wave_path0 = (uint64_t)blockIdx.x * kRegBlock + (threadIdx.x & ~63);
  const GlobalW W{ws, bd.n_paths};
  double psum = 0.0, psq = 0.0;
  double gd = 0.0, gm = 0.0, gv = 0.0;  // VALU: lane a owns asset a's sums
  double md[4] = {0, 0, 0, 0}, mm[4] = {0, 0, 0, 0}, mv[4] = {0, 0, 0, 0};  // MFMA
  if constexpr (kMfma) {
    basket_mfma_epilogue_w(bd, W, wave_path0, psum, psq, md, mm, mv, kGreeks);
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
  } else {
    const uint64_t p = wave_path0 + lane;
    const bool valid = p < bd.n_paths;
    double B = 0.0;
#pragma unroll 1
    for (uint32_t i = 0; i < d; ++i) {
      float e = 0.f;
      for (uint32_t k = 0; k <= i; ++k) e = fmaf(bd.Lf[i * 64 + k], W(k, p), e);
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
          float e = 0.f;
          for (uint32_t k = 0; k <= i; ++k) e = fmaf(bd.Lf[i * 64 + k], W(k, p), e);
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

template <bool kMfma>
static int basket_reg_launch(const BasketDev& bd, const StreamDesc& sd, bool greeks,
                             double* h_out, int na, hipStream_t st, const Err& err) {
  const uint64_t nb = (bd.n_paths + kRegBlock - 1) / kRegBlock;  // phase 2: 1 path / thread
  const uint64_t nb1 = (bd.n_paths + kSumsBlock - 1) / kSumsBlock;
  if (nb > 0x7fffffffull) return err.set(STOCHA_EINVAL, "n_paths too large for one launch");
  Workspace* wsp = nullptr;
  const size_t w_bytes = ((sizeof(float) * bd.d * bd.n_paths + 255) / 256) * 256;
  const size_t part_bytes = sizeof(double) * nb * na;
  if (int rc = workspace(w_bytes + part_bytes + sizeof(double) * na, sizeof(double) * na, &wsp,
                         err))
    return rc;
  float* ws = static_cast<float*>(wsp->dev);
  double* part = reinterpret_cast<double*>(static_cast<char*>(wsp->dev) + w_bytes);
  double* out = part + nb * na;
  KernelTimer timer(st);
  timer.start();
  basket_sums_kernel<<<(unsigned)nb1, kSumsBlock, 0, st>>>(bd, sd, ws);
  STOCHA_HIP(hipGetLastError());
  if (greeks)
    basket_payoff_kernel<true, kMfma><<<(unsigned)nb, kRegBlock, 0, st>>>(bd, ws, part);
  else
    basket_payoff_kernel<false, kMfma><<<(unsigned)nb, kRegBlock, 0, st>>>(bd, ws, part);
  STOCHA_HIP(hipGetLastError());
  timer.stop(2);
  if (int rc = launch_reduce_partials(part, (int)nb, na, out, st, err)) return rc;
  STOCHA_HIP(hipMemcpyAsync(wsp->host, out, sizeof(double) * na, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < na; ++i) h_out[i] = wsp->host[i];
  return STOCHA_OK;
}

template <bool kFp32, int kBlock, bool kMfma>
static int basket_launch(const BasketDev& bd, const StreamDesc& sd, bool greeks, double* h_out,
                         int na, hipStream_t st, const Err& err) {
  uint64_t nb = (bd.n_paths + kBlock - 1) / kBlock;
  if (nb > 4096) nb = 4096;
  DevBuf part, out;
  if (int rc = part.alloc(sizeof(double) * nb * na, err)) return rc;
  if (int rc = out.alloc(sizeof(double) * na, err)) return rc;
  KernelTimer timer(st);
  timer.start();
  if (greeks)
    basket_kernel<kFp32, kBlock, true, kMfma>
        <<<(unsigned)nb, kBlock, 0, st>>>(bd, sd, part.as<double>());
  else
    basket_kernel<kFp32, kBlock, false, kMfma>
        <<<(unsigned)nb, kBlock, 0, st>>>(bd, sd, part.as<double>());
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  if (int rc = launch_reduce_partials(part.as<double>(), (int)nb, na, out.as<double>(), st, err))
    return rc;
  STOCHA_HIP(hipMemcpyAsync(h_out, out.p, sizeof(double) * na, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
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
    if (mfma) return basket_reg_launch<true>(bd, sd, greeks, h_sums, na, st, err);
    return basket_reg_launch<false>(bd, sd, greeks, h_sums, na, st, err);
  }
  return basket_launch<false, 64, false>(bd, sd, greeks, h_sums, na, st, err);
}

}  // namespace stocha

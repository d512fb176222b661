// bulk.hip -- bulk draws of ONE sequential PCG64DXSM stream: the RNG class's
// uniform / standard_normal / normal (src/lib.rs:75-184, pcg64.rs:33-39,
// normal.rs:18-79).
//
// The reference fills its output from a single generator, one value after the
// other.  Uniforms take exactly one draw each, so value i is the draw at
// stream position i: each thread jumps to its chunk (the affine LCG map,
// host tables) and writes a tile through LDS so that stores coalesce.
//
// Normals take a variable number of draws (a ziggurat attempt, a wedge
// uniform, tail pairs), so where normal k starts depends on every draw before
// it.  The positions are found in parallel by speculative chunk decoding:
//   1. spec:    the stream is cut into chunks of kChunk positions; a thread
//               decodes normals from its chunk's first position as if a normal
//               started there, recording the starts it meets inside the chunk
//               (count, and a 64-bit mask of which positions are starts) and
//               the first start at or after the chunk's end (its exit);
//   2. resolve: the true first start in chunk c is the exit of chunk c-1.  If
//               that position is on c's speculative chain (the mask), the two
//               chains coincide from there; otherwise the thread walks the true
//               chain and the speculative one together until they meet (a few
//               draws: most starts are one draw apart).  A chunk whose chains do
//               not meet before its end changes its exit, and the pass repeats
//               with the corrected entries (a cascade that has not occurred in
//               the tests; it is bounded by the number of chunks);
//   3. scan:    exclusive sum of the per-chunk counts = each chunk's first
//               output index;
//   4. emit:    each chunk decodes its normals again from its true entry and
//               writes them; the thread holding the last requested normal
//               records where it ends (the draws consumed).
// Values are the reference's (ZigFeed consumes exactly the reference's draws);
// the host advances the generator by the draws consumed.
#include <hipcub/hipcub.hpp>

#include <math.h>

#include <vector>

#include "host.hpp"
#include "multi.hpp"
#include "zigsm.hpp"

namespace stocha {
namespace {

constexpr int kBulkBlock = 256;
constexpr int kUniPer = 16;  // uniforms per thread (one LDS tile row)
constexpr int kChunk = 64;   // stream positions per normal-decoding chunk

__global__ __launch_bounds__(kBulkBlock) void seq_uniform_kernel(const StreamDesc sd,
                                                                 const uint64_t n,
                                                                 double* __restrict__ out) {
  __shared__ double tile[kBulkBlock][kUniPer + 1];
  const uint64_t chunk = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kBulkBlock * kUniPer;
  Pcg g{stream_at(&sd, chunk)};
#pragma unroll
  for (int j = 0; j < kUniPer; ++j) tile[threadIdx.x][j] = (double)(g.next(sd.inc) >> 11) * 0x1p-53;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kUniPer; ++i) {
    const int e = i * kBulkBlock + threadIdx.x;
    const uint64_t idx = base + (uint64_t)e;
    if (idx < n) out[idx] = tile[e / kUniPer][e % kUniPer];
  }
}

// One normal from stream position p (g at p): the value; p and g move past it.
__device__ __forceinline__ double decode_one(Pcg& g, const U128 inc, const ZigLds* z,
                                             uint64_t& p) {
  ZigFeed<false> zf;
  double x;
  for (;;) {
    const uint64_t u = g.next(inc);
    ++p;
    if (zf.feed(u, z, x, true)) return x;
  }
}

// The generator at position p of chunk c (p - c * kChunk single steps on from
// the chunk's first position).
__device__ __forceinline__ Pcg state_at(const StreamDesc& sd, uint64_t c, uint64_t p) {
  Pcg g{stream_at(&sd, c)};
  for (uint64_t q = c * kChunk; q < p; ++q) g.s = g.stepped(sd.inc);
  return g;
}

__global__ __launch_bounds__(kBulkBlock) void norm_spec_kernel(const StreamDesc sd,
                                                               const uint64_t n_chunks,
                                                               uint32_t* __restrict__ cnt,
                                                               uint64_t* __restrict__ mask,
                                                               uint64_t* __restrict__ exit) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c >= n_chunks) return;
  const uint64_t s = c * kChunk, end = s + kChunk;
  Pcg g{stream_at(&sd, c)};
  uint64_t p = s, m = 0;
  uint32_t k = 0;
  while (p < end) {
    ++k;
    m |= 1ull << (p - s);
    (void)decode_one(g, sd.inc, &zig, p);
  }
  cnt[c] = k;
  mask[c] = m;
  exit[c] = p;
}

// prev_exit[c-1] is chunk c's entry (0 for chunk 0); prev_exit[c] the exit
// assumed for c, which the pass checks.
__global__ __launch_bounds__(kBulkBlock) void norm_resolve_kernel(
    const StreamDesc sd, const uint64_t n_chunks, const uint32_t* __restrict__ spec_cnt,
    const uint64_t* __restrict__ spec_mask, const uint64_t* __restrict__ spec_exit,
    const uint64_t* __restrict__ prev_exit, uint64_t* __restrict__ cnt,
    uint64_t* __restrict__ texit, unsigned* __restrict__ changed) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c >= n_chunks) return;
  const uint64_t s = c * kChunk, end = s + kChunk;
  const uint64_t e = c == 0 ? 0 : prev_exit[c - 1];
  const uint64_t sx = spec_exit[c];
  uint64_t k, x;
  if (e >= end) {  // a normal spans the whole chunk
    k = 0;
    x = e;
  } else if ((spec_mask[c] >> (e - s)) & 1ull) {
    k = spec_cnt[c] - (uint32_t)__popcll(spec_mask[c] & ((1ull << (e - s)) - 1ull));
    x = sx;
  } else {
    // walk the true chain (a) and the speculative one (b) until they meet
    Pcg ga = state_at(sd, c, e), gb{stream_at(&sd, c)};
    uint64_t a = e, b = s, ka = 0, kb = 0;
    for (;;) {
      if (a == b) {
        k = ka + (spec_cnt[c] - kb);
        x = sx;
        break;
      }
      if (a >= end && b >= end) {
        k = ka;
        x = a;
        break;
      }
      if (a < b) {
        ++ka;
        (void)decode_one(ga, sd.inc, &zig, a);
      } else {
        ++kb;
        (void)decode_one(gb, sd.inc, &zig, b);
      }
    }
  }
  cnt[c] = k;
  texit[c] = x;
  if (x != prev_exit[c]) atomicOr(changed, 1u);  // not yet a fixed point
}

__global__ __launch_bounds__(kBulkBlock) void norm_emit_kernel(
    const StreamDesc sd, const uint64_t n_chunks, const uint64_t* __restrict__ prev_exit,
    const uint64_t* __restrict__ cnt, const uint64_t* __restrict__ off, const uint64_t n,
    const double loc, const double scale, const int scaled, double* __restrict__ out,
    uint64_t* __restrict__ consumed) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c >= n_chunks) return;
  const uint64_t o = off[c], k = cnt[c];
  if (k == 0 || o >= n) return;
  const uint64_t e = c == 0 ? 0 : prev_exit[c - 1];
  Pcg g = state_at(sd, c, e);
  uint64_t p = e;
  for (uint64_t j = 0; j < k && o + j < n; ++j) {
    const double z = decode_one(g, sd.inc, &zig, p);
    out[o + j] = scaled ? loc + scale * z : z;  // sample_scaled (normal.rs:77-79)
    if (o + j == n - 1) *consumed = p;
  }
}

typedef unsigned __int128 u128h_t;

// StreamDesc of one stream from (state, inc) with jumps of block * 2^k.
void seq_desc(const uint64_t st[2], const uint64_t inc[2], uint64_t block, StreamDesc* d) {
  d->base = U128{st[0], st[1]};
  d->inc = U128{inc[0], inc[1]};
  const u128h_t i = ((u128h_t)inc[1] << 64) | inc[0];
  u128h a, cc;
  lcg_power(i, (u128h)block, &a, &cc);
  for (int k = 0; k < kJumpBits; ++k) {
    d->jump[k].a = to_u128(a);
    d->jump[k].c = to_u128(cc);
    cc = a * cc + cc;
    a = a * a;
  }
}

void advance_state(const uint64_t st[2], const uint64_t inc[2], uint64_t delta, uint64_t out[2]) {
  const u128h_t i = ((u128h_t)inc[1] << 64) | inc[0];
  u128h a, cc;
  lcg_power(i, (u128h)delta, &a, &cc);
  const u128h_t s = ((u128h_t)st[1] << 64) | st[0];
  const u128h_t r = a * s + cc;
  out[0] = (uint64_t)r;
  out[1] = (uint64_t)(r >> 64);
}

int place(double* out, int where, size_t bytes, double** dev, Workspace** ws, size_t extra,
          const Err& err) {
  if (where == STOCHA_DEVICE) {
    *dev = out;
    return workspace(extra, 0, ws, err);
  }
  const size_t off = (bytes + 255) & ~(size_t)255;
  if (int rc = workspace(off + extra, 0, ws, err)) return rc;
  *dev = static_cast<double*>((*ws)->dev);
  return STOCHA_OK;
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// n normals of the stream (state, inc) on the device: into ext (device) when
// given, else at the start of the workspace; `front` bytes at the start of the
// workspace are left to the caller (they hold the normals when ext is null).
// *ws_out is the workspace the call used; *used the draws consumed.
int decode_normals(const uint64_t state[2], const uint64_t inc[2], uint64_t n, double loc,
                   double scale, int scaled, size_t front, double* ext, Workspace** ws_out,
                   uint64_t* used, const Err& err) {
  hipStream_t st = globals().stream;
  StreamDesc sd;
  seq_desc(state, inc, kChunk, &sd);
  // positions to decode: ~1.0296 draws per normal, with margin; grown if short
  uint64_t n_chunks = (n + n / 25 + 2 * kChunk) / kChunk + 1;
  for (int attempt = 0; attempt < 8; ++attempt) {
    if (n_chunks > 0x7fffffffull) return err.set(STOCHA_EINVAL, "size too large");
    const unsigned grid = (unsigned)((n_chunks + kBulkBlock - 1) / kBulkBlock);
    const size_t nc = n_chunks;
    size_t tmp = 0;
    STOCHA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const uint64_t*)nullptr,
                                                (uint64_t*)nullptr, (int)nc, st));
    const size_t extra = al256(4 * nc) + 6 * al256(8 * nc) + 256 + al256(tmp);
    Workspace* ws = nullptr;
    if (int rc = workspace(al256(front) + extra, 0, &ws, err)) return rc;
    double* dev = ext ? ext : static_cast<double*>(ws->dev);
    char* w = static_cast<char*>(ws->dev) + al256(front);
    uint32_t* spec_cnt = reinterpret_cast<uint32_t*>(w);
    w += al256(4 * nc);
    uint64_t* spec_mask = reinterpret_cast<uint64_t*>(w);
    w += al256(8 * nc);
    uint64_t* spec_exit = reinterpret_cast<uint64_t*>(w);
    w += al256(8 * nc);
    uint64_t* ex[2] = {reinterpret_cast<uint64_t*>(w),
                       reinterpret_cast<uint64_t*>(w + al256(8 * nc))};
    w += 2 * al256(8 * nc);
    uint64_t* cnt = reinterpret_cast<uint64_t*>(w);
    w += al256(8 * nc);
    uint64_t* off = reinterpret_cast<uint64_t*>(w);
    w += al256(8 * nc);
    unsigned* flag = reinterpret_cast<unsigned*>(w);
    uint64_t* cons = reinterpret_cast<uint64_t*>(w + 64);
    w += 256;
    void* scan_tmp = w;

    KernelTimer timer(st);
    timer.start();
    norm_spec_kernel<<<grid, kBulkBlock, 0, st>>>(sd, n_chunks, spec_cnt, spec_mask, spec_exit);
    STOCHA_HIP(hipGetLastError());
    // resolve until no chunk's exit moves (the speculative exits first)
    const uint64_t* prev = spec_exit;
    int cur = 0;
    uint64_t rounds = 0;
    for (;;) {
      STOCHA_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned), st));
      norm_resolve_kernel<<<grid, kBulkBlock, 0, st>>>(sd, n_chunks, spec_cnt, spec_mask,
                                                       spec_exit, prev, cnt, ex[cur], flag);
      STOCHA_HIP(hipGetLastError());
      unsigned h = 0;
      STOCHA_HIP(hipMemcpyAsync(&h, flag, sizeof(unsigned), hipMemcpyDeviceToHost, st));
      STOCHA_HIP(hipStreamSynchronize(st));
      if (!h) break;
      if (++rounds > n_chunks)
        return err.set(STOCHA_EDEVICE, "normal chunk resolution did not settle");
      prev = ex[cur];
      cur ^= 1;
    }
    // prev: the exits whose entries the final resolve used
    STOCHA_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tmp, cnt, off, (int)nc, st));
    uint64_t last_off = 0, last_cnt = 0;
    STOCHA_HIP(hipMemcpyAsync(&last_off, off + nc - 1, 8, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipMemcpyAsync(&last_cnt, cnt + nc - 1, 8, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipStreamSynchronize(st));
    if (last_off + last_cnt < n) {  // too few positions: decode more
      n_chunks += n_chunks / 4 + 4;
      continue;
    }
    norm_emit_kernel<<<grid, kBulkBlock, 0, st>>>(sd, n_chunks, prev, cnt, off, n, loc, scale,
                                                  scaled, dev, cons);
    STOCHA_HIP(hipGetLastError());
    timer.stop(3);
    STOCHA_HIP(hipMemcpyAsync(used, cons, 8, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipStreamSynchronize(st));
    *ws_out = ws;
    return STOCHA_OK;
  }
  return err.set(STOCHA_EDEVICE, "could not decode enough normals");
}

// normal_cdf.rs:3-21 (Abramowitz & Stegun 26.2.17)
__device__ __forceinline__ double norm_cdf_2617(double x) {
  const double p = 0.2316419;
  const double b0 = 0.319381530, b1 = -0.356563782, b2 = 1.781477937, b3 = -1.821255978,
               b4 = 1.330274429;
  const double t = 1.0 / (1.0 + p * fabs(x));
  const double poly = t * (b0 + t * (b1 + t * (b2 + t * (b3 + t * b4))));
  const double pdf = exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
  const double cdf_pos = 1.0 - pdf * poly;
  return x >= 0.0 ? cdf_pos : 1.0 - cdf_pos;
}

// gaussian.rs:58-70: x_j = sum_{k<=j} L[j][k] z_k from 0.0 in k order
__global__ __launch_bounds__(kBulkBlock) void gauss_copula_kernel(const double* __restrict__ z,
                                                                  const double* __restrict__ L,
                                                                  const uint64_t n,
                                                                  const uint32_t d,
                                                                  double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (i >= n) return;
  const double* zi = z + i * d;
  for (uint32_t j = 0; j < d; ++j) {
    double xj = 0.0;
    for (uint32_t k = 0; k <= j; ++k) xj += L[(uint64_t)j * d + k] * zi[k];
    out[i * d + j] = norm_cdf_2617(xj);
  }
}

}  // namespace
}  // namespace stocha

using namespace stocha;

extern "C" {

// pcg64.rs:23-30 for a 128-bit seed: (state, increment) of Pcg64Dxsm::new.
int stocha_rng_new(uint64_t seed_lo, uint64_t seed_hi, uint64_t* state, uint64_t* inc) {
  const u128h_t s = ((u128h_t)seed_hi << 64) | seed_lo;
  const u128h_t stream = s + (((u128h_t)0xda3e39cb94b95bdbull << 64) | 0x7ef55da5c0374f5aull);
  const u128h_t i = (stream << 1) | 1;
  u128h_t x = s + i;
  x = x * (u128h_t)kPcgMul + i;
  state[0] = (uint64_t)x;
  state[1] = (uint64_t)(x >> 64);
  inc[0] = (uint64_t)i;
  inc[1] = (uint64_t)(i >> 64);
  return STOCHA_OK;
}

// RNG.uniform (lib.rs:108-112): n values of next_f64 from (state, inc);
// new_state = the state after n draws.
int stocha_rng_uniform(const uint64_t* state, const uint64_t* inc, uint64_t n, double* out,
                       int where, uint64_t* new_state, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  advance_state(state, inc, n, new_state);
  if (n == 0) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  hipStream_t st = globals().stream;
  StreamDesc sd;
  seq_desc(state, inc, kUniPer, &sd);
  const uint64_t blocks = (n + (uint64_t)kBulkBlock * kUniPer - 1) / ((uint64_t)kBulkBlock * kUniPer);
  if (blocks > 0x7fffffffull) return err.set(STOCHA_EINVAL, "size too large for one launch");
  Workspace* ws = nullptr;
  double* dev = nullptr;
  if (int rc = place(out, where, sizeof(double) * n, &dev, &ws, 0, err)) return rc;
  KernelTimer timer(st);
  timer.start();
  seq_uniform_kernel<<<(unsigned)blocks, kBulkBlock, 0, st>>>(sd, n, dev);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  if (where != STOCHA_DEVICE)
    STOCHA_HIP(hipMemcpyAsync(out, dev, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

// RNG.standard_normal / RNG.normal (lib.rs:114-145): n ziggurat normals of
// the stream from (state, inc), scaled to loc + scale z when scaled != 0;
// new_state = the state after the draws they consumed (*consumed, optional).
int stocha_rng_normal(const uint64_t* state, const uint64_t* inc, uint64_t n, double loc,
                      double scale, int scaled, double* out, int where, uint64_t* new_state,
                      uint64_t* consumed, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (scaled && scale <= 0.0) return err.set(STOCHA_EINVAL, "scale must be positive");
  if (n == 0) {
    new_state[0] = state[0];
    new_state[1] = state[1];
    if (consumed) *consumed = 0;
    return STOCHA_OK;
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  hipStream_t st = globals().stream;
  const bool host = where != STOCHA_DEVICE;
  Workspace* ws = nullptr;
  uint64_t used = 0;
  if (int rc = decode_normals(state, inc, n, loc, scale, scaled, host ? al256(8 * n) : 0,
                              host ? nullptr : out, &ws, &used, err))
    return rc;
  if (host)
    STOCHA_HIP(hipMemcpyAsync(out, ws->dev, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  advance_state(state, inc, used, new_state);
  if (consumed) *consumed = used;
  return STOCHA_OK;
}

// copula/gaussian.rs:15-73 (gaussian_copula_samples, binding lib.rs:538-558):
// row i = Phi(L z_i) with z_i the normals i*d .. i*d+d-1 of the stream
// Pcg64Dxsm::new(seed); corr is (rows, cols) row-major on the host.
int stocha_gaussian_copula(const double* corr, uint64_t rows, uint64_t cols, uint64_t n_samples,
                           uint64_t seed, double* out, int where, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  const uint64_t d = rows;
  if (cols != d) return err.set(STOCHA_EINVAL, "correlation matrix must be square");
  for (uint64_t i = 0; i < d; ++i) {
    const double v = corr[i * d + i];
    if (fabs(v - 1.0) > 1e-10)
      return err.set(STOCHA_EINVAL,
                     "correlation matrix diagonal element [%llu][%llu] must be 1.0, got %s",
                     (unsigned long long)i, (unsigned long long)i, rust_f64(v).c_str());
    for (uint64_t j = i + 1; j < d; ++j) {
      const double a = corr[i * d + j], b = corr[j * d + i];
      if (fabs(a - b) > 1e-10)
        return err.set(STOCHA_EINVAL,
                       "correlation matrix is not symmetric: [%llu,%llu]=%s vs [%llu,%llu]=%s",
                       (unsigned long long)i, (unsigned long long)j, rust_f64(a).c_str(),
                       (unsigned long long)j, (unsigned long long)i, rust_f64(b).c_str());
    }
  }
  std::vector<double> L(d * d, 0.0);
  if (d > 0)
    if (int rc = cholesky_host(corr, d, L.data(), err)) return rc;
  if (!out) return STOCHA_OK;  // validate only
  const uint64_t n = n_samples * d;
  if (n == 0) return STOCHA_OK;
  if (d > 0xffffffffull) return err.set(STOCHA_EINVAL, "dimension too large");
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  hipStream_t st = globals().stream;
  uint64_t state[2], inc[2], used = 0;
  stocha_rng_new(seed, 0, state, inc);
  // front of the workspace: normals, L, and the output when it goes to the host
  const bool host = where != STOCHA_DEVICE;
  const size_t z_bytes = al256(8 * n), l_bytes = al256(8 * d * d);
  Workspace* ws = nullptr;
  if (int rc = decode_normals(state, inc, n, 0.0, 1.0, 0, z_bytes + l_bytes + (host ? al256(8 * n) : 0),
                              nullptr, &ws, &used, err))
    return rc;
  char* base = static_cast<char*>(ws->dev);
  const double* z = reinterpret_cast<const double*>(base);
  double* dl = reinterpret_cast<double*>(base + z_bytes);
  double* dout = host ? reinterpret_cast<double*>(base + z_bytes + l_bytes) : out;
  STOCHA_HIP(hipMemcpyAsync(dl, L.data(), sizeof(double) * d * d, hipMemcpyHostToDevice, st));
  const uint64_t blocks = (n_samples + kBulkBlock - 1) / kBulkBlock;
  if (blocks > 0x7fffffffull) return err.set(STOCHA_EINVAL, "n_samples too large for one launch");
  gauss_copula_kernel<<<(unsigned)blocks, kBulkBlock, 0, st>>>(z, dl, n_samples, (uint32_t)d,
                                                               dout);
  STOCHA_HIP(hipGetLastError());
  if (host) STOCHA_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

}  // extern "C"

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

#include <math.h>

#include <vector>

#include "host.hpp"
#include "scan.hpp"
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
__device__ __forceinline__ double decode_one(Pcg& g, const U128 inc, const ZigLds64* z,
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
  __shared__ ZigLds64 zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c >= n_chunks) return;
  const uint64_t s = c * kChunk, end = s + kChunk;
  Pcg g{stream_at(&sd, c)};
  uint64_t p = s, m = 0;
  uint32_t k = 0;
  // one draw per iteration (the lanes of a wave stay in step; a per-normal
  // loop would make every lane wait for the slowest lane's normal)
  ZigFeed<false> zf;
  bool start = true;  // a normal starts at p
  while (!(start && p >= end)) {
    if (start) {
      ++k;
      m |= 1ull << (p - s);
    }
    double x;
    start = zf.feed(g.next(sd.inc), &zig, x, true);
    ++p;
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
  __shared__ ZigLds64 zig;
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
  __shared__ ZigLds64 zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c >= n_chunks) return;
  const uint64_t o = off[c], k = cnt[c];
  if (k == 0 || o >= n) return;
  const uint64_t e = c == 0 ? 0 : prev_exit[c - 1];
  Pcg g = state_at(sd, c, e);
  uint64_t p = e, j = 0;
  ZigFeed<false> zf;  // one draw per iteration, as in the speculative pass
  while (j < k && o + j < n) {
    double z;
    const bool done = zf.feed(g.next(sd.inc), &zig, z, true);
    ++p;
    if (done) {
      out[o + j] = scaled ? loc + scale * z : z;  // sample_scaled (normal.rs:77-79)
      if (o + j == n - 1) *consumed = p;
      ++j;
    }
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
    const size_t tmp = 8 * scan_tmp_words(nc);
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
    if (int rc = exclusive_sum_u64(cnt, off, nc, static_cast<uint64_t*>(scan_tmp), st, err))
      return rc;
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

// ---------------------------------------------------------------------------
// Student-t copula (copula/student_t.rs:135-169) on one stream new(seed).
// Sample i draws a chi-square(nu) = 2 gamma(nu/2) (Marsaglia-Tsang,
// student_t.rs:94-119: a normal, a uniform unless 1 + c x <= 0, repeated
// until accepted), then d normals.  Every sample starts in the same state,
// so the draws a sample takes, len(p), depend only on its start position p.
// The starts are found in parallel:
//   1. draws:  the raw stream, positions [0, n_tot), into memory;
//   2. len:    len(p) for every position p < n_pos (one sample decoded per
//              position; kLenBad if it runs past n_tot);
//   3. maps:   positions in chunks of kTChunk; per chunk and entry offset e
//              (the first start in the chunk is at chunk + e) the exit
//              offset into the next chunk and the number of starts;
//   4. entries: each block of 256 chunks composes their maps (a tree); one
//              thread applies the blocks' compositions in order to entry 0
//              (chunk 0 enters at 0), giving each block's entry; each block
//              then walks its own chunks' maps from it (maps in LDS).  The
//              counts' exclusive sum gives each chunk's first sample index;
//   5. emit:   each chunk lists its starts; one thread per sample decodes it
//              (the normals and sqrt(nu / chi2));
//   6. cdf:    u_ij = t_cdf((L z_i)_j sqrt(nu / chi2_i), nu).
// A sample longer than kTChunk draws (d above ~50) overflows the maps; then
// the starts come from pointer doubling over len(p) instead (same values).
constexpr int kTChunk = 64;
constexpr uint32_t kLenBad = 0xffffu;
// largest dimension: a sample takes ~2.1 + 1.03 d draws, far below kLenBad
constexpr uint64_t kTMaxDim = 32768;
constexpr uint8_t kMapOver = 0xff;  // exit offset >= kTChunk
constexpr uint8_t kMapBad = 0xfe;   // ran past the decoded positions

struct __align__(16) TMap {
  uint8_t e[kTChunk];
};
// r = first a, then b (codes >= kTChunk stay), on maps in LDS (16-byte
// aligned rows): a and r move as four 16-byte words, b is read per entry
__device__ __forceinline__ void t_compose(const uint8_t* __restrict__ a,
                                          const uint8_t* __restrict__ b, uint8_t* __restrict__ r) {
#pragma unroll
  for (int w = 0; w < kTChunk / 16; ++w) {
    const uint4 av = reinterpret_cast<const uint4*>(a)[w];
    const uint32_t in[4] = {av.x, av.y, av.z, av.w};
    uint32_t out[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t o = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t x = (in[q] >> (8 * k)) & 0xffu;
        const uint32_t y = x >= (uint32_t)kTChunk ? x : (uint32_t)b[x];
        o |= y << (8 * k);
      }
      out[q] = o;
    }
    reinterpret_cast<uint4*>(r)[w] = make_uint4(out[0], out[1], out[2], out[3]);
  }
}

__device__ __forceinline__ void t_copy_row(const uint8_t* __restrict__ a, uint8_t* __restrict__ r) {
#pragma unroll
  for (int w = 0; w < kTChunk / 16; ++w)
    reinterpret_cast<uint4*>(r)[w] = reinterpret_cast<const uint4*>(a)[w];
}

// row i = the identity map (entry e -> e)
__device__ __forceinline__ void t_ident_row(uint8_t* __restrict__ r) {
#pragma unroll
  for (int w = 0; w < kTChunk / 16; ++w) {
    const uint32_t b0 = 16u * w;
    const uint32_t x = (b0) | (b0 + 1) << 8 | (b0 + 2) << 16 | (b0 + 3) << 24;
    reinterpret_cast<uint4*>(r)[w] = make_uint4(x, x + 0x04040404u, x + 0x08080808u,
                                                x + 0x0c0c0c0cu);
  }
}

struct TParams {
  double nu, d, c;  // d = nu/2 - 1/3, c = 1 / sqrt(9 d) (student_t.rs:104-105)
  double lb;        // log B(nu/2, 1/2) by the reference's Lanczos log-gamma
  uint32_t dim;
};

__global__ __launch_bounds__(kBulkBlock) void seq_u64_kernel(const StreamDesc sd,
                                                             const uint64_t n,
                                                             uint64_t* __restrict__ out) {
  __shared__ uint64_t tile[kBulkBlock][kUniPer + 1];
  const uint64_t chunk = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kBulkBlock * kUniPer;
  Pcg g{stream_at(&sd, chunk)};
#pragma unroll
  for (int j = 0; j < kUniPer; ++j) tile[threadIdx.x][j] = g.next(sd.inc);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kUniPer; ++i) {
    const int e = i * kBulkBlock + threadIdx.x;
    const uint64_t idx = base + (uint64_t)e;
    if (idx < n) out[idx] = tile[e / kUniPer][e % kUniPer];
  }
}

// One normal from the draws at q (q moves past it); false if it needs draws
// at or past n.
__device__ __forceinline__ bool normal_from(const uint64_t* __restrict__ U, uint64_t n,
                                            uint64_t& q, const ZigLds64* z, double& x) {
  ZigFeed<false> zf;
  for (;;) {
    if (q >= n) return false;
    if (zf.feed(U[q++], z, x, true)) return true;
  }
}

// The sample starting at p: its draw count (kLenBad when it runs past n);
// with kKeep its normals into zo and sqrt(nu / chi2) into *scale.
template <bool kKeep>
__device__ uint32_t t_sample(const uint64_t* __restrict__ U, uint64_t n, uint64_t p,
                             const TParams& tp, const ZigLds64* z, double* __restrict__ zo,
                             double* __restrict__ scale) {
  uint64_t q = p;
  double x, v;
  for (;;) {  // student_t.rs:106-118
    if (!normal_from(U, n, q, z, x)) return kLenBad;
    const double vi = 1.0 + tp.c * x;
    if (vi <= 0.0) continue;
    v = vi * vi * vi;
    if (q >= n) return kLenBad;
    const double u = (double)(U[q++] >> 11) * 0x1p-53;  // pcg64.rs:37-39
    if (u < 1.0 - 0.0331 * (x * x) * (x * x)) break;
    if (log(u) < 0.5 * x * x + tp.d * (1.0 - v + log(v))) break;
  }
  if constexpr (kKeep) *scale = sqrt(tp.nu / (2.0 * (tp.d * v)));  // :156-157
  for (uint32_t j = 0; j < tp.dim; ++j) {
    double zz;
    if (!normal_from(U, n, q, z, zz)) return kLenBad;
    if constexpr (kKeep) zo[j] = zz;
  }
  const uint64_t len = q - p;
  return len >= kLenBad ? kLenBad : (uint32_t)len;
}

// t_sample's draw count as one flat loop, one draw per iteration, so the
// lanes of a wave stay in step (nested data-dependent loops had them in
// different places of their samples).  phase 0: a gamma normal, 1: the gamma
// uniform, 2 + j: the sample's normal j.
__device__ uint32_t t_sample_len(const uint64_t* __restrict__ U, uint64_t n, uint64_t p,
                                 const TParams& tp, const ZigLds64* z) {
  ZigFeed<false> zf;
  const uint32_t end = 2u + tp.dim;
  uint32_t phase = 0;
  double x = 0.0;  // the gamma normal under test
  uint64_t q = p;
  while (phase < end) {
    if (q >= n) return kLenBad;
    const uint64_t u = U[q++];
    if (phase == 1u) {  // student_t.rs:110-117
      const double uf = (double)(u >> 11) * 0x1p-53;
      const double vi = 1.0 + tp.c * x;
      const double v = vi * vi * vi;
      bool ok = uf < 1.0 - 0.0331 * (x * x) * (x * x);
      if (!ok) ok = log(uf) < 0.5 * x * x + tp.d * (1.0 - v + log(v));
      phase = ok ? 2u : 0u;
    } else {
      double y;
      if (zf.feed(u, z, y, true)) {
        if (phase == 0u) {
          x = y;
          phase = (1.0 + tp.c * y <= 0.0) ? 0u : 1u;  // :108-109
        } else {
          ++phase;
        }
      }
    }
  }
  const uint64_t l = q - p;
  return l >= kLenBad ? kLenBad : (uint32_t)l;
}

__global__ __launch_bounds__(kBulkBlock) void t_len_kernel(const uint64_t* __restrict__ U,
                                                           const uint64_t n_tot,
                                                           const uint64_t n_pos,
                                                           const TParams tp,
                                                           uint16_t* __restrict__ len) {
  __shared__ ZigLds64 zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t p = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (p >= n_pos) return;
  len[p] = (uint16_t)t_sample_len(U, n_tot, p, tp, &zig);
}

// thread (chunk, entry offset): walk the starts to the chunk's end
__global__ __launch_bounds__(kBulkBlock) void t_map_kernel(const uint16_t* __restrict__ len,
                                                           const uint64_t n_chunks,
                                                           TMap* __restrict__ maps,
                                                           uint8_t* __restrict__ cnt) {
  const uint64_t t = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  const uint64_t c = t / kTChunk;
  const uint32_t e = (uint32_t)(t % kTChunk);
  if (c >= n_chunks) return;
  const uint64_t end = (c + 1) * kTChunk;
  uint64_t p = c * kTChunk + e;
  uint32_t k = 0;
  uint8_t code = 0;
  while (p < end) {
    const uint32_t l = len[p];
    if (l == kLenBad) {
      code = kMapBad;
      break;
    }
    p += l;
    ++k;
  }
  if (code == 0) code = p - end < (uint64_t)kTChunk ? (uint8_t)(p - end) : kMapOver;
  maps[c].e[e] = code;
  cnt[c * kTChunk + e] = (uint8_t)k;
}

// Level 1: each block composes its kBulkBlock chunks' maps in order (a tree
// of pairwise compositions in LDS) into agg[block].
__global__ __launch_bounds__(kBulkBlock) void t_agg_kernel(const TMap* __restrict__ maps,
                                                           const uint64_t n_chunks,
                                                           TMap* __restrict__ agg) {
  __shared__ __align__(16) uint8_t m[2][kBulkBlock][kTChunk];
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c < n_chunks)
    t_copy_row(maps[c].e, m[0][threadIdx.x]);
  else
    t_ident_row(m[0][threadIdx.x]);  // identity past the last chunk
  int cur = 0;
  for (int w = kBulkBlock / 2; w >= 1; w >>= 1) {  // row t = rows 2t, 2t+1 composed
    __syncthreads();
    if ((int)threadIdx.x < w)
      t_compose(m[cur][2 * threadIdx.x], m[cur][2 * threadIdx.x + 1], m[cur ^ 1][threadIdx.x]);
    cur ^= 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) t_copy_row(m[cur][0], agg[blockIdx.x].e);
}

// Level 2 (one block): xin[b] = the entry of block b's first chunk, the
// blocks' totals applied in order to entry 0 (chunk 0 enters at 0); tiles of
// kBulkBlock totals staged in LDS, walked by one thread.
__global__ __launch_bounds__(kBulkBlock) void t_xin_kernel(const TMap* __restrict__ agg,
                                                           const uint64_t n_blocks,
                                                           uint8_t* __restrict__ xin) {
  __shared__ __align__(16) uint8_t m[kBulkBlock][kTChunk];
  uint32_t x = 0;
  for (uint64_t base = 0; base < n_blocks; base += kBulkBlock) {
    const uint64_t b = base + threadIdx.x;
    __syncthreads();
    if (b < n_blocks) t_copy_row(agg[b].e, m[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t k = n_blocks - base < (uint64_t)kBulkBlock ? n_blocks - base : kBulkBlock;
      for (uint64_t j = 0; j < k; ++j) {
        xin[base + j] = (uint8_t)x;
        x = x < (uint32_t)kTChunk ? m[j][x] : x;
      }
    }
  }
}

// Level 3: each block walks its chunks from its entry: the true entry of
// every chunk (thread 0, maps in LDS), then its count of starts; flag bit 0:
// a chunk's entry overflowed, bit 1: ran past the decoded positions
__global__ __launch_bounds__(kBulkBlock) void t_entry_kernel(const TMap* __restrict__ maps,
                                                             const uint8_t* __restrict__ xin,
                                                             const uint8_t* __restrict__ cnt,
                                                             const uint64_t n_chunks,
                                                             uint8_t* __restrict__ entry,
                                                             uint64_t* __restrict__ cnt_true,
                                                             unsigned* __restrict__ flag) {
  __shared__ __align__(16) uint8_t m[kBulkBlock][kTChunk];
  __shared__ uint8_t ent[kBulkBlock];
  const uint64_t c0 = (uint64_t)blockIdx.x * kBulkBlock, c = c0 + threadIdx.x;
  if (c < n_chunks) t_copy_row(maps[c].e, m[threadIdx.x]);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t k = n_chunks - c0 < (uint64_t)kBulkBlock ? n_chunks - c0 : kBulkBlock;
    uint32_t x = xin[blockIdx.x];
    for (uint64_t j = 0; j < k; ++j) {
      ent[j] = (uint8_t)x;
      x = x < (uint32_t)kTChunk ? m[j][x] : x;
    }
  }
  __syncthreads();
  if (c >= n_chunks) return;
  const uint8_t e = ent[threadIdx.x];
  entry[c] = e;
  if (e < kTChunk) {
    cnt_true[c] = cnt[c * kTChunk + e];
  } else {
    cnt_true[c] = 0;
    atomicOr(flag, e == kMapOver ? 1u : 2u);
  }
}

// each chunk lists its starts (sample index off[c] + k)
__global__ __launch_bounds__(kBulkBlock) void t_starts_kernel(const uint16_t* __restrict__ len,
                                                              const uint64_t n_chunks,
                                                              const uint8_t* __restrict__ entry,
                                                              const uint64_t* __restrict__ cnt_true,
                                                              const uint64_t* __restrict__ off,
                                                              const uint64_t n_samples,
                                                              uint64_t* __restrict__ starts) {
  const uint64_t c = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (c >= n_chunks || entry[c] >= kTChunk) return;
  uint64_t p = c * kTChunk + entry[c];
  const uint64_t o = off[c], k = cnt_true[c];
  for (uint64_t j = 0; j < k && o + j < n_samples; ++j) {
    starts[o + j] = p;
    p += len[p];
  }
}

// Samples above kTChunk draws (the maps overflow): the starts by pointer
// doubling.  J0(p) = p + len(p) is the next start after a start at p (kNoPos
// where len is bad or the next start lies at or past n_pos); doubling a table
// (J(J(p))) six times gives J6 = J0^64, six more J12 = J0^4096.  One thread
// walks J12 from 0 (every 4096th start), one thread per such start walks J6
// (every 64th), one per those walks J0 and writes the starts; a walk that
// meets kNoPos records where the chain ends (*found, atomicMin).
constexpr uint32_t kNoPos = 0xffffffffu;
constexpr int kJumpLog = 6;  // J6 = J0^(2^kJumpLog)
constexpr uint64_t kJumpSpan = 1ull << kJumpLog;

__global__ __launch_bounds__(kBulkBlock) void t_next_kernel(const uint16_t* __restrict__ len,
                                                            const uint64_t n_pos,
                                                            uint32_t* __restrict__ J) {
  const uint64_t p = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (p >= n_pos) return;
  const uint32_t l = len[p];
  J[p] = (l == kLenBad || p + l >= n_pos) ? kNoPos : (uint32_t)(p + l);
}

__global__ __launch_bounds__(kBulkBlock) void t_double_kernel(const uint32_t* __restrict__ Jin,
                                                              const uint64_t n_pos,
                                                              uint32_t* __restrict__ Jout) {
  const uint64_t p = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (p >= n_pos) return;
  const uint32_t q = Jin[p];
  Jout[p] = q == kNoPos ? kNoPos : Jin[q];
}

// level 2: one thread, every (kJumpSpan^2)-th start into top[] (kNoPos past the end)
__global__ void t_walk_top_kernel(const uint32_t* __restrict__ J12, const uint64_t n_top,
                                  uint32_t* __restrict__ top) {
  uint32_t p = 0;
  for (uint64_t i = 0; i < n_top; ++i) {
    top[i] = p;
    if (p != kNoPos) p = J12[p];
  }
}

// level 1: thread i walks J6 from top[i]: mid[i * kJumpSpan + k]
__global__ __launch_bounds__(kBulkBlock) void t_walk_mid_kernel(const uint32_t* __restrict__ top,
                                                                const uint64_t n_top,
                                                                const uint32_t* __restrict__ J6,
                                                                const uint64_t n_mid,
                                                                uint32_t* __restrict__ mid) {
  const uint64_t i = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (i >= n_top) return;
  uint32_t p = top[i];
  for (uint64_t k = 0; k < kJumpSpan; ++k) {
    const uint64_t m = i * kJumpSpan + k;
    if (m >= n_mid) return;
    mid[m] = p;
    if (p != kNoPos) p = J6[p];
  }
}

// level 0: thread m walks J0 from mid[m] and writes starts m*kJumpSpan + k;
// the first missing start (past n_pos, or a sample that runs past the decoded
// draws) bounds *found
__global__ __launch_bounds__(kBulkBlock) void t_walk_low_kernel(
    const uint32_t* __restrict__ mid, const uint64_t n_mid, const uint32_t* __restrict__ J0,
    const uint16_t* __restrict__ len, const uint64_t n_samples, uint64_t* __restrict__ starts,
    unsigned long long* __restrict__ found) {
  const uint64_t m = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (m >= n_mid) return;
  uint32_t p = mid[m];
  for (uint64_t k = 0; k < kJumpSpan; ++k) {
    const uint64_t s = m * kJumpSpan + k;
    if (s >= n_samples) return;
    if (p == kNoPos || len[p] == kLenBad) {
      atomicMin(found, (unsigned long long)s);
      return;
    }
    starts[s] = p;
    p = J0[p];
  }
}

__global__ __launch_bounds__(kBulkBlock) void t_emit_kernel(const uint64_t* __restrict__ U,
                                                            const uint64_t n_tot,
                                                            const uint64_t* __restrict__ starts,
                                                            const uint64_t n_samples,
                                                            const TParams tp,
                                                            double* __restrict__ Z,
                                                            double* __restrict__ S) {
  __shared__ ZigLds64 zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (i >= n_samples) return;
  (void)t_sample<true>(U, n_tot, starts[i], tp, &zig, Z + i * tp.dim, S + i);
}

// student_t.rs:19-63: I_x(a, b) with log B(a, b) = lb (the same for (b, a)):
// 1 - I_{1-x}(b, a) past the mean, else the continued fraction (modified
// Lentz, EPS 3e-7, at most 200 steps)
__device__ double t_ibeta(double x, double a, double b, double lb) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  bool flip = false;
  if (x > (a + 1.0) / (a + b + 2.0)) {
    flip = true;
    x = 1.0 - x;
    const double t = a;
    a = b;
    b = t;
  }
  const double front = exp(a * log(x) + b * log(1.0 - x) - lb) / a;
  const double eps = 3e-7, fpmin = 1e-30;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0;
  double d = fmax(1.0 - qab * x / qap, fpmin);
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 200; ++m) {
    const double mf = (double)m, m2 = 2.0 * mf;
    double aa = mf * (b - mf) * x / ((qam + m2) * (a + m2));
    d = fmax(fabs(1.0 + aa * d), fpmin);
    c = fmax(fabs(1.0 + aa / c), fpmin);
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + mf) * (qab + mf) * x / ((a + m2) * (qap + m2));
    d = fmax(fabs(1.0 + aa * d), fpmin);
    c = fmax(fabs(1.0 + aa / c), fpmin);
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < eps) break;
  }
  const double r = front * h;
  return flip ? 1.0 - r : r;
}

// student_t.rs:8-16 and :160-165: x_j = sum_{k<=j} L[j][k] z_k from 0.0 in
// k order, u = t_cdf(x_j scale, nu)
__global__ __launch_bounds__(kBulkBlock) void t_cdf_kernel(const double* __restrict__ Z,
                                                           const double* __restrict__ S,
                                                           const double* __restrict__ L,
                                                           const uint64_t n, const TParams tp,
                                                           double* __restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * kBulkBlock + threadIdx.x;
  if (t >= n) return;
  const uint64_t i = t / tp.dim;
  const uint32_t j = (uint32_t)(t % tp.dim);
  const double* zi = Z + i * tp.dim;
  double xj = 0.0;
  for (uint32_t k = 0; k <= j; ++k) xj += L[(uint64_t)j * tp.dim + k] * zi[k];
  const double x = xj * S[i];
  const double zz = tp.nu / (tp.nu + x * x);
  const double ib = t_ibeta(zz, tp.nu / 2.0, 0.5, tp.lb);
  out[t] = x >= 0.0 ? 1.0 - 0.5 * ib : 0.5 * ib;
}

// student_t.rs:67-91 on the host (the platform libm, as the reference's)
double lgamma_lanczos_host(double x) {
  static const double C[9] = {0.99999999999980993,  676.5203681218851,     -1259.1392167224028,
                              771.32342877765313,   -176.61502916214059,   12.507343278686905,
                              -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
  if (x < 0.5) return log(M_PI) - log(sin(M_PI * x)) - lgamma_lanczos_host(1.0 - x);
  x -= 1.0;
  double acc = C[0];
  for (int i = 1; i < 9; ++i) acc += C[i] / (x + (double)i);
  const double t = x + 7.0 + 0.5;
  return 0.5 * log(2.0 * M_PI) + (x + 0.5) * log(t) - t + log(acc);
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
  size_t bytes = 0;
  if (int rc = check_output(out, where, n, 1, sizeof(double), "out", &bytes, err)) return rc;
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
  size_t bytes = 0;
  if (int rc = check_output(out, where, n, 1, sizeof(double), "out", &bytes, err)) return rc;
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
  size_t bytes = 0;
  if (int rc = check_output(out, where, n_samples, d, sizeof(double), "out", &bytes, err))
    return rc;
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

extern "C" {

// copula/student_t.rs:135-169 (student_t_copula_samples, binding
// lib.rs:563-578): row i = t_cdf((L z_i) sqrt(nu / chi2_i), nu) on the stream
// Pcg64Dxsm::new(seed); corr is (rows, cols) row-major on the host.
int stocha_student_t_copula(const double* corr, uint64_t rows, uint64_t cols, double nu,
                            uint64_t n_samples, uint64_t seed, double* out, int where,
                            char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  // lib.rs:570-572 (a NaN nu is refused too: the reference's gamma loop
  // would never accept)
  if (!(nu > 2.0)) return err.set(STOCHA_EINVAL, "nu must be > 2");
  const uint64_t d = rows;
  if (cols != d) return err.set(STOCHA_EINVAL, "correlation matrix must be square");
  std::vector<double> L(d * d, 0.0);
  if (d > 0)
    if (int rc = cholesky_host(corr, d, L.data(), err)) return rc;  // gaussian.rs:75-95
  if (!out) return STOCHA_OK;  // validate only
  size_t bytes = 0;
  if (int rc = check_output(out, where, n_samples, d, sizeof(double), "out", &bytes, err))
    return rc;
  const uint64_t n = n_samples * d;
  if (n == 0) return STOCHA_OK;
  // per-position sample lengths are 16-bit (0xffff marks a position whose
  // sample runs past the decoded draws): ~1.03 d + 2 draws per sample fit
  // with a wide margin up to this dimension
  if (d > kTMaxDim)
    return err.set(STOCHA_EINVAL, "the Student-t copula supports dimensions up to %llu, got %llu",
                   (unsigned long long)kTMaxDim, (unsigned long long)d);
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  hipStream_t st = globals().stream;
  uint64_t state[2], inc[2];
  stocha_rng_new(seed, 0, state, inc);
  StreamDesc sd;
  seq_desc(state, inc, kUniPer, &sd);
  TParams tp;
  tp.nu = nu;
  tp.d = nu / 2.0 - 1.0 / 3.0;
  tp.c = 1.0 / sqrt(9.0 * tp.d);
  tp.lb = lgamma_lanczos_host(nu / 2.0) + lgamma_lanczos_host(0.5) -
          lgamma_lanczos_host(nu / 2.0 + 0.5);  // student_t.rs:28
  tp.dim = (uint32_t)d;
  const bool host = where != STOCHA_DEVICE;
  // draws per sample ~2.1 + 1.03 d; positions with margin, grown if short
  uint64_t n_chunks = (uint64_t)((double)n_samples * (2.4 + 1.08 * (double)d)) / kTChunk + 4;
  // the doubling tables of the long-sample walk: a sample takes at least d + 2
  // draws, so below d = 16 it overflows a 64-draw chunk only after ~25
  // rejected gamma attempts; the tables are then added on the retry
  bool walk_tables = d >= 16;
  for (int attempt = 0; attempt < 8; ++attempt) {
    const uint64_t n_pos = n_chunks * kTChunk;
    const uint64_t n_tot = n_pos + 16 * (d + 8) + 1024;  // the last samples' draws
    if (n_tot > (1ull << 40)) return err.set(STOCHA_EINVAL, "size too large");
    if (walk_tables && n_pos >= (uint64_t)kNoPos)
      return err.set(STOCHA_EINVAL, "size too large");
    const size_t nc = n_chunks;
    const uint64_t n_blk = (nc + kBulkBlock - 1) / kBulkBlock;
    const size_t tmp = 8 * scan_tmp_words(nc);
    const uint64_t n_mid = (n_samples + kJumpSpan - 1) / kJumpSpan;
    const uint64_t n_top = (n_mid + kJumpSpan - 1) / kJumpSpan;
    const size_t jt = walk_tables ? 4 * n_pos : 0;  // each of J0, J6, and two doubling buffers
    const size_t sz[] = {8 * n_tot,     2 * n_pos,     sizeof(TMap) * nc, 0,
                         kTChunk * nc,  nc,            8 * nc,            8 * nc,
                         256,           tmp,           8 * n_samples,     8 * n,
                         8 * n_samples, 8 * d * d,     host ? 8 * n : 0,
                         sizeof(TMap) * n_blk, n_blk,
                         jt,            jt,            jt,                jt,
                         walk_tables ? 4 * n_mid : 0, walk_tables ? 4 * n_top : 0};
    constexpr int kRegions = 23;
    size_t o[kRegions], total = 0;
    for (int i = 0; i < kRegions; ++i) {
      o[i] = total;
      total += al256(sz[i]);
    }
    Workspace* ws = nullptr;
    if (int rc = workspace(total, 0, &ws, err)) return rc;
    char* b = static_cast<char*>(ws->dev);
    uint64_t* U = reinterpret_cast<uint64_t*>(b + o[0]);
    uint16_t* len = reinterpret_cast<uint16_t*>(b + o[1]);
    TMap* maps = reinterpret_cast<TMap*>(b + o[2]);
    uint8_t* cnt = reinterpret_cast<uint8_t*>(b + o[4]);
    uint8_t* entry = reinterpret_cast<uint8_t*>(b + o[5]);
    uint64_t* cnt_true = reinterpret_cast<uint64_t*>(b + o[6]);
    uint64_t* off = reinterpret_cast<uint64_t*>(b + o[7]);
    unsigned* flag = reinterpret_cast<unsigned*>(b + o[8]);
    unsigned long long* found = reinterpret_cast<unsigned long long*>(b + o[8] + 64);
    void* scan_tmp = b + o[9];
    uint64_t* starts = reinterpret_cast<uint64_t*>(b + o[10]);
    double* Z = reinterpret_cast<double*>(b + o[11]);
    double* S = reinterpret_cast<double*>(b + o[12]);
    double* dl = reinterpret_cast<double*>(b + o[13]);
    double* dout = host ? reinterpret_cast<double*>(b + o[14]) : out;
    TMap* agg = reinterpret_cast<TMap*>(b + o[15]);
    uint8_t* xin = reinterpret_cast<uint8_t*>(b + o[16]);
    uint32_t* J0 = reinterpret_cast<uint32_t*>(b + o[17]);
    uint32_t* J6 = reinterpret_cast<uint32_t*>(b + o[18]);
    uint32_t* T1 = reinterpret_cast<uint32_t*>(b + o[19]);
    uint32_t* T2 = reinterpret_cast<uint32_t*>(b + o[20]);
    uint32_t* mid = reinterpret_cast<uint32_t*>(b + o[21]);
    uint32_t* top = reinterpret_cast<uint32_t*>(b + o[22]);

    const auto blocks = [](uint64_t items, uint64_t per) { return (unsigned)((items + per - 1) / per); };
    KernelTimer timer(st);
    timer.start();
    seq_u64_kernel<<<blocks(n_tot, (uint64_t)kBulkBlock * kUniPer), kBulkBlock, 0, st>>>(sd, n_tot,
                                                                                        U);
    t_len_kernel<<<blocks(n_pos, kBulkBlock), kBulkBlock, 0, st>>>(U, n_tot, n_pos, tp, len);
    t_map_kernel<<<blocks(n_pos, kBulkBlock), kBulkBlock, 0, st>>>(len, n_chunks, maps, cnt);
    STOCHA_HIP(hipGetLastError());
    t_agg_kernel<<<(unsigned)n_blk, kBulkBlock, 0, st>>>(maps, n_chunks, agg);
    t_xin_kernel<<<1, kBulkBlock, 0, st>>>(agg, n_blk, xin);
    STOCHA_HIP(hipMemsetAsync(flag, 0, 256, st));
    t_entry_kernel<<<(unsigned)n_blk, kBulkBlock, 0, st>>>(maps, xin, cnt, n_chunks, entry,
                                                            cnt_true, flag);
    STOCHA_HIP(hipGetLastError());
    if (int rc = exclusive_sum_u64(cnt_true, off, nc, static_cast<uint64_t*>(scan_tmp), st, err))
      return rc;
    unsigned h_flag = 0;
    uint64_t last_off = 0, last_cnt = 0;
    STOCHA_HIP(hipMemcpyAsync(&h_flag, flag, 4, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipMemcpyAsync(&last_off, off + nc - 1, 8, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipMemcpyAsync(&last_cnt, cnt_true + nc - 1, 8, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipStreamSynchronize(st));
    if (h_flag & 1u) {  // a sample above kTChunk draws: the starts by pointer doubling
      if (!walk_tables) {
        walk_tables = true;
        continue;
      }
      const unsigned pb = blocks(n_pos, kBulkBlock);
      t_next_kernel<<<pb, kBulkBlock, 0, st>>>(len, n_pos, J0);
      // J0 -> T1 -> T2 -> T1 -> T2 -> T1 -> J6 (x64), then J6 -> T1 -> T2 -> T1
      // -> T2 -> T1 -> T2 (x4096): J12 ends in T2
      const uint32_t* src = J0;
      for (int r = 0; r < 2 * kJumpLog; ++r) {
        uint32_t* dst = r == kJumpLog - 1 ? J6 : (r % 2 == 0 ? T1 : T2);
        t_double_kernel<<<pb, kBulkBlock, 0, st>>>(src, n_pos, dst);
        src = dst;
      }
      const unsigned long long n_all = n_samples;
      STOCHA_HIP(hipMemcpyAsync(found, &n_all, 8, hipMemcpyHostToDevice, st));
      t_walk_top_kernel<<<1, 1, 0, st>>>(T2, n_top, top);
      t_walk_mid_kernel<<<blocks(n_top, kBulkBlock), kBulkBlock, 0, st>>>(top, n_top, J6, n_mid,
                                                                          mid);
      t_walk_low_kernel<<<blocks(n_mid, kBulkBlock), kBulkBlock, 0, st>>>(mid, n_mid, J0, len,
                                                                          n_samples, starts,
                                                                          found);
      STOCHA_HIP(hipGetLastError());
      unsigned long long h_found = 0;
      STOCHA_HIP(hipMemcpyAsync(&h_found, found, 8, hipMemcpyDeviceToHost, st));
      STOCHA_HIP(hipStreamSynchronize(st));
      if (h_found < n_samples) {
        n_chunks += n_chunks / 4 + 4;
        continue;
      }
    } else {
      if (last_off + last_cnt < n_samples) {  // too few positions: decode more
        n_chunks += n_chunks / 4 + 4;
        continue;
      }
      t_starts_kernel<<<blocks(nc, kBulkBlock), kBulkBlock, 0, st>>>(len, n_chunks, entry,
                                                                     cnt_true, off, n_samples,
                                                                     starts);
    }
    t_emit_kernel<<<blocks(n_samples, kBulkBlock), kBulkBlock, 0, st>>>(U, n_tot, starts,
                                                                        n_samples, tp, Z, S);
    STOCHA_HIP(hipMemcpyAsync(dl, L.data(), 8 * d * d, hipMemcpyHostToDevice, st));
    t_cdf_kernel<<<blocks(n, kBulkBlock), kBulkBlock, 0, st>>>(Z, S, dl, n, tp, dout);
    STOCHA_HIP(hipGetLastError());
    timer.stop(4);
    if (host) STOCHA_HIP(hipMemcpyAsync(out, dout, 8 * n, hipMemcpyDeviceToHost, st));
    STOCHA_HIP(hipStreamSynchronize(st));
    return STOCHA_OK;
  }
  return err.set(STOCHA_EDEVICE, "could not decode enough samples");
}

}  // extern "C"

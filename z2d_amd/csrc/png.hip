// png.hip - device PNG export for gfx950 (the reference's writeToPNGFile,
// src/export_png.zig:56-360).
//
// The scanline stream (filter byte 0 + encoded row) is exactly the
// reference's: rgb/xrgb as r,g,b; argb/rgba demultiplied in integers
// (pixel_vector.RGBA16.demultiply, src/internal/pixel_vector.zig:27-49);
// alpha8 as is; alpha4/2/1 packed MSB-first, rows padded to a byte
// (src/export_png.zig:265-295); the sRGB profile re-encodes the color
// channels with pow(c/255, 1/2.2) (:330-350).  The zlib stream around it is
// this library's own: the reference uses Zig's std.compress.flate at its
// default level, so the compressed bytes differ while every decoder yields
// the same stream.  Chunk framing follows :100-145, :393-446.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "png.h"
#include "render.h"

namespace z2d {
namespace {

constexpr uint32_t kAdlerMod = 65521;
constexpr uint32_t kCrcPoly = 0xEDB88320u;  // CRC-32 (PNG, zlib), reflected
constexpr uint32_t kSegWords = PNG_SEG / 32;
constexpr uint32_t kMatchWords = kSegWords + 2;  // + the word holding the segment end
constexpr uint32_t kOutWords = ((PNG_SEG + 5) * 8 + 1024) / 32 + 2;

// ---------------------------------------------------------------------------
// k_png_rows: one thread per pixel (4-byte formats, alpha8) or per output
// byte (alpha4/2/1), grid-stride.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_png_rows(PngJob J) {
  __shared__ uint8_t lut[256];
  const bool srgb = J.profile == Z2D_PNG_PROFILE_SRGB;
  if (srgb) {  // decodeRGBAVecRaw -> applyGammaVec -> encodeRGBAVecRaw per 8-bit value
    const float v = (float)threadIdx.x / 255.0f;
    lut[threadIdx.x] = (uint8_t)f2u8(255.0f * zpowf(v, 1.0f / 2.2f));
    __syncthreads();
  }
  const uint32_t fmt = J.format;
  const uint64_t W = (uint64_t)J.width, L = J.row_bytes + 1;
  const uint64_t per_row = fmt <= Z2D_FORMAT_ALPHA8 ? W : J.row_bytes;
  const uint64_t total = per_row * (uint64_t)J.height;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    const uint64_t y = i / per_row, x = i - y * per_row;
    uint8_t *row = J.raw + y * L;
    if (x == 0) row[0] = 0;  // filter type 0
    if (fmt <= Z2D_FORMAT_RGBA) {
      const uint32_t v = reinterpret_cast<const uint32_t *>(J.surface)[y * W + x];
      const bool rgba_order = fmt == Z2D_FORMAT_RGBA || fmt == Z2D_FORMAT_RGB;
      uint32_t r = rgba_order ? (v & 0xFF) : ((v >> 16) & 0xFF);
      const uint32_t g = (v >> 8) & 0xFF;
      uint32_t b = rgba_order ? ((v >> 16) & 0xFF) : (v & 0xFF);
      uint32_t gg = g;
      uint32_t a = v >> 24;
      const bool alpha = fmt == Z2D_FORMAT_ARGB || fmt == Z2D_FORMAT_RGBA;
      if (alpha) {
        if (a == 0) {
          r = gg = b = 0;
        } else {
          r = r * 255u / a;
          gg = gg * 255u / a;
          b = b * 255u / a;
        }
      }
      if (srgb) {
        r = lut[r];
        gg = lut[gg];
        b = lut[b];
        if (alpha) a = f2u8(255.0f * ((float)a / 255.0f));
      }
      uint8_t *o = row + 1 + x * (alpha ? 4 : 3);
      o[0] = (uint8_t)r;
      o[1] = (uint8_t)gg;
      o[2] = (uint8_t)b;
      if (alpha) o[3] = (uint8_t)a;
    } else if (fmt == Z2D_FORMAT_ALPHA8) {
      row[1 + x] = J.surface[y * W + x];
    } else {
      const uint32_t bits = fmt == Z2D_FORMAT_ALPHA4 ? 4u : fmt == Z2D_FORMAT_ALPHA2 ? 2u : 1u;
      const uint32_t per = 8u / bits, mask = (1u << bits) - 1u;
      uint32_t byte = 0;
      for (uint32_t j = 0; j < per; j++) {
        const uint64_t px = x * per + j;
        if (px >= W) break;
        const uint64_t bit = (y * W + px) * bits;
        const uint32_t v = (J.surface[bit >> 3] >> (bit & 7)) & mask;
        byte |= v << (8u - bits - j * bits);
      }
      row[1 + x] = (uint8_t)byte;
    }
  }
}

// ---------------------------------------------------------------------------
// k_png_deflate: one wave per segment.
//
// Matching: for each candidate distance d a bitmap marks the positions whose
// byte differs from the byte d earlier (positions past the segment end, and
// those with no byte d earlier, count as differing).  A position where some
// candidate repeats three bytes is "good".  The wave parses greedily: literal
// runs up to the next good position, and at a good position each candidate
// lane measures its run from its bitmap; the longest (earliest candidate on
// ties) becomes a match of up to 258 bytes.  The parse is kept as match
// tokens and symbol counts; the block is then coded with Huffman codes built
// for the segment (RFC 1951 3.2.7: a bitonic sort of the counts, one lane's
// two-queue merge, lengths up to 15 / 7 bits, else the fixed codes), with
// the fixed codes, or stored, whichever is smallest.  Codes go into an LDS
// bit buffer with atomic ORs; literal runs are coded 64 bytes at a time (one
// lane each, offsets by a wave prefix sum).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put_bits(uint32_t *w, uint32_t pos, uint32_t v, uint32_t nb) {
  const uint32_t k = pos >> 5, sh = pos & 31u;
  atomicOr(&w[k], v << sh);
  if (sh + nb > 32u) atomicOr(&w[k + 1], v >> (32u - sh));
}

// Length symbol (257..285) of a match length 3..258 and its extra bits.
__device__ __forceinline__ uint32_t len_sym(uint32_t len, uint32_t &ext, uint32_t &val) {
  ext = 0;
  val = 0;
  if (len == 258u) return 285u;
  const uint32_t v = len - 3u;
  if (v < 8u) return 257u + v;
  const uint32_t e = 31u - __builtin_clz(v) - 2u;
  ext = e;
  val = v & ((1u << e) - 1u);
  return 257u + 4u * (e + 1u) + ((v >> e) & 3u);
}
// Distance symbol (0..29) of a distance 1..32768 and its extra bits.
__device__ __forceinline__ uint32_t dist_sym(uint32_t dist, uint32_t &ext, uint32_t &val) {
  ext = 0;
  val = 0;
  const uint32_t dv = dist - 1u;
  if (dv < 4u) return dv;
  const uint32_t e = 31u - __builtin_clz(dv) - 1u;
  ext = e;
  val = dv & ((1u << e) - 1u);
  return 2u * (e + 1u) + ((dv >> e) & 1u);
}

constexpr uint32_t kMaxTok = PNG_SEG / 3 + 2;  // matches in a segment
static_assert(4 * kMaxTok <= PNG_SEG_CAP, "token words must fit the segment's staging bytes");
constexpr uint32_t kNoKey = 0xFFFFFFFFu;

struct DeflateShared {
  union {
    struct {  // matching, until the parse ends
      uint32_t mm[PNG_MAX_CAND][kMatchWords];
      uint32_t good[kMatchWords];
    };
    struct {  // code construction, after it
      uint32_t key[512];  // sort keys: count << 9 | symbol
      union {
        uint16_t hw[576];  // two-queue weights (a segment's counts sum to <= PNG_SEG + 1)
        uint8_t hdep[576];  // node depths, once the merge is done with the weights
      };
      uint16_t hpar[576];
    };
    uint8_t cl[320];           // literal/length then distance code lengths, for the run coding
    uint32_t outw[kOutWords];  // the coded block, last
  };
  uint32_t lfreq[288], dfreq[32], cfreq[20];
  uint32_t lcode[288], dcode[32], ccode[20];  // bit-reversed code | length << 16
  uint8_t llen[288], dlen[32], clen[20];
  uint32_t rle[320];  // code-length symbols: symbol | extra << 8 | extra bits << 16
  uint32_t litw[kSegWords], mstw[kSegWords];  // literal positions, match starts
  uint32_t n_rle;
};

// Ascending bitonic sort of 64 * R keys held in registers, element
// lane * R + r in v[r]; strides below R swap within a lane, larger ones
// exchange with lane ^ (stride / R).
template <int R>
__device__ __forceinline__ void wave_sort(uint32_t (&v)[R], uint32_t lane) {
#pragma unroll
  for (uint32_t size = 2; size <= 64u * R; size <<= 1) {
#pragma unroll
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride < (uint32_t)R) {
#pragma unroll
        for (uint32_t r = 0; r < (uint32_t)R; r++) {
          if (r & stride) continue;
          const uint32_t e = lane * R + r;
          const bool up = (e & size) == 0;
          const uint32_t a = v[r], b = v[r ^ stride];
          const bool sw = (a > b) == up;
          v[r] = sw ? b : a;
          v[r ^ stride] = sw ? a : b;
        }
      } else {
#pragma unroll
        for (uint32_t r = 0; r < (uint32_t)R; r++) {
          const uint32_t e = lane * R + r;
          const bool up = (e & size) == 0, low = (e & stride) == 0;
          const uint32_t a = v[r], b = (uint32_t)__shfl_xor((int)a, (int)(stride / R), 64);
          v[r] = (low == up) ? min(a, b) : max(a, b);
        }
      }
    }
  }
}

// Code lengths for `nsym` counts (64 * R >= nsym), limited to `maxlen`;
// returns false (lengths unusable) when a code would be longer.  Keys
// count << 9 | symbol are sorted on the wave, one lane merges the two queues
// (leaves, then internal nodes in order of creation), and depths follow by
// pointer jumping up the parent links.
template <int R>
__device__ bool huff_lengths(DeflateShared &S, const uint32_t *freq, uint32_t nsym,
                             uint32_t maxlen, uint8_t *lens, uint32_t lane) {
  uint32_t v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t i = lane * R + r;
    v[r] = (i < nsym && freq[i]) ? (freq[i] << 9 | i) : kNoKey;
  }
  wave_sort<R>(v, lane);
  uint32_t cnt = 0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    S.key[lane * R + r] = v[r];
    cnt += v[r] != kNoKey;
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  const uint32_t m = cnt;
  for (uint32_t i = lane; i < nsym; i += 64) lens[i] = 0;
  __syncthreads();
  if (m == 0) return true;
  if (m == 1) {
    if (lane == 0) lens[S.key[0] & 511u] = 1;
    __syncthreads();
    return true;
  }
  const uint32_t root = 2 * m - 2;
  for (uint32_t k = lane; k < m; k += 64) S.hw[k] = (uint16_t)(S.key[k] >> 9);
  __syncthreads();
  if (lane == 0) {
    uint32_t i1 = 0, i2 = m, nxt = m;
    uint32_t w1 = S.hw[0], w2 = 0;
    while (nxt <= root) {
      uint32_t sum = 0;
      for (int t = 0; t < 2; t++) {
        uint32_t pick;
        if (i1 < m && (i2 >= nxt || w1 <= w2)) {
          pick = i1++;
          sum += w1;
          w1 = i1 < m ? S.hw[i1] : 0u;
        } else {
          pick = i2++;
          sum += w2;
          w2 = i2 < nxt ? S.hw[i2] : 0u;
        }
        S.hpar[pick] = (uint16_t)nxt;
      }
      S.hw[nxt] = (uint16_t)sum;
      if (i2 == nxt) w2 = sum;  // the internal queue was empty
      nxt++;
    }
    S.hpar[root] = (uint16_t)root;
  }
  __syncthreads();
  // depth by pointer jumping: after j rounds hpar[k] is the 2^j-th ancestor
  // (or the root) and hdep[k] the distance to it; five rounds settle every
  // depth up to 32, and a node not yet at the root is deeper than that
  constexpr int NR = 9;  // nodes per lane (2 * 286 - 1 <= 64 * 9)
#pragma unroll
  for (int j = 0; j < NR; j++) {
    const uint32_t k = lane + 64u * j;
    if (k <= root) S.hdep[k] = k == root ? 0 : 1;
  }
  __syncthreads();
  for (int round = 0; round < 5; round++) {
    uint32_t na[NR], nd[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) {
      const uint32_t k = lane + 64u * j;
      if (k <= root) {
        const uint32_t a = S.hpar[k];
        na[j] = S.hpar[a];
        nd[j] = (uint32_t)S.hdep[k] + S.hdep[a];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NR; j++) {
      const uint32_t k = lane + 64u * j;
      if (k <= root) {
        S.hpar[k] = (uint16_t)na[j];
        S.hdep[k] = (uint8_t)min(nd[j], 255u);
      }
    }
    __syncthreads();
  }
  bool ok = true;
  for (uint32_t k = lane; k < m; k += 64) {
    const uint32_t d = S.hdep[k];
    if (S.hpar[k] != root || d > maxlen) ok = false;
    lens[S.key[k] & 511u] = (uint8_t)d;
  }
  __syncthreads();
  return __ballot(!ok) == 0;
}

// Canonical codes from lengths (RFC 1951 3.2.2), bit-reversed for LSB-first
// output, packed as code | length << 16: a symbol's code is the first code of
// its length plus the number of symbols of that length before it.
__device__ void huff_codes(const uint8_t *lens, uint32_t nsym, uint32_t *codes, uint32_t lane) {
  uint32_t bl[16];
#pragma unroll
  for (int L = 0; L < 16; L++) bl[L] = 0;
  for (uint32_t c0 = 0; c0 < nsym; c0 += 64) {
    const uint32_t i = c0 + lane;
    const uint32_t l = i < nsym ? lens[i] : 0u;
#pragma unroll
    for (int L = 1; L < 16; L++) bl[L] += (uint32_t)__popcll(__ballot(l == (uint32_t)L));
  }
  uint32_t next[16];
  uint32_t code = 0;
  next[0] = 0;
#pragma unroll
  for (int L = 1; L < 16; L++) {
    code = (code + (L > 1 ? bl[L - 1] : 0u)) << 1;
    next[L] = code;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t c0 = 0; c0 < nsym; c0 += 64) {
    const uint32_t i = c0 + lane;
    const uint32_t l = i < nsym ? lens[i] : 0u;
    uint32_t out = 0;
#pragma unroll
    for (int L = 1; L < 16; L++) {
      const uint64_t msk = __ballot(l == (uint32_t)L);
      if (l == (uint32_t)L) {
        const uint32_t c = next[L] + (uint32_t)__popcll(msk & below);
        out = (__builtin_bitreverse32(c) >> (32 - L)) | ((uint32_t)L << 16);
      }
      next[L] += (uint32_t)__popcll(msk);
    }
    if (i < nsym) codes[i] = out;
  }
  __syncthreads();
}

// Code-length symbols for a run of L lengths v (RFC 1951 3.2.7: 16 repeats
// the previous length 3-6 times, 17 / 18 give 3-10 / 11-138 zeros), as
// symbol | extra << 8 | extra bits << 16 from out[k]; returns their count.
// With out null it only counts.
__device__ uint32_t rle_run(uint32_t v, uint32_t L, uint32_t *out, uint32_t k) {
  uint32_t c = 0;
  auto put = [&](uint32_t x) {
    if (out) out[k + c] = x;
    c++;
  };
  if (v == 0) {
    while (L >= 11) {
      const uint32_t r = min(L, 138u);
      put(18u | ((r - 11u) << 8) | (7u << 16));
      L -= r;
    }
    if (L >= 3) {
      put(17u | ((L - 3u) << 8) | (3u << 16));
      L = 0;
    }
    for (; L > 0; L--) put(0u);
  } else {
    put(v);
    L--;
    while (L >= 3) {
      const uint32_t r = min(L, 6u);
      put(16u | ((r - 3u) << 8) | (2u << 16));
      L -= r;
    }
    for (; L > 0; L--) put(v);
  }
  return c;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_png_deflate(PngJob J) {
  __shared__ DeflateShared S;
  const uint32_t lane = threadIdx.x;
  const uint32_t s = blockIdx.x;
  const uint64_t base = (uint64_t)s * PNG_SEG;
  const uint32_t n = (uint32_t)min((uint64_t)PNG_SEG, J.raw_len - base);
  const bool final_seg = s + 1 == J.n_seg;
  const uint32_t nw = (n >> 5) + 2;  // bitmap words written (<= kMatchWords)
  const uint32_t nc = J.n_cand;
  uint8_t *dst = J.seg + (uint64_t)s * PNG_SEG_CAP;
  // match tokens (pos << 11 | (len - 3) << 3 | candidate) wait in the
  // segment's staging bytes until the block is coded over them
  uint32_t *tok = reinterpret_cast<uint32_t *>(dst);
  for (uint32_t i = lane; i < 288; i += 64) S.lfreq[i] = i == 256 ? 1u : 0u;  // one end of block
  for (uint32_t w = lane; w < kSegWords; w += 64) {
    S.litw[w] = 0;
    S.mstw[w] = 0;
  }
  if (lane < 32) S.dfreq[lane] = 0;
  if (lane < 20) S.cfreq[lane] = 0;

  // bitmaps and the Adler-32 partial sums of the segment
  uint64_t s1 = 0, s2 = 0;
  for (uint32_t w = lane; w < nw; w += 64) {
    const uint32_t p0 = w * 32u;
    const uint64_t g0 = base + p0;
    const uint4 *cp = reinterpret_cast<const uint4 *>(J.raw + g0);
    const uint4 c0 = cp[0], c1 = cp[1];
    const uint32_t cur[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    const uint32_t valid = p0 >= n ? 0u : (n - p0 >= 32u ? ~0u : ((1u << (n - p0)) - 1u));
    for (uint32_t k = 0; k < 32; k++) {
      if (valid >> k & 1u) {
        const uint32_t b = (cur[k >> 2] >> ((k & 3u) * 8u)) & 0xFFu;
        s1 += b;
        s2 += (uint64_t)(n - (p0 + k)) * b;
      }
    }
    for (uint32_t c = 0; c < nc; c++) {
      const uint64_t d = J.cand[c];
      uint32_t m = 0;
      if (g0 + 32u <= d) {
        m = ~0u;
      } else if (g0 >= d) {
        const uint64_t src = g0 - d;
        const uint32_t *sp = reinterpret_cast<const uint32_t *>(J.raw + (src & ~3ull));
        const uint32_t sh = (uint32_t)(src & 3u) * 8u;
        uint32_t prev = sp[0];
        for (uint32_t k = 0; k < 8; k++) {
          const uint32_t nxt = sp[k + 1];
          const uint32_t v = sh ? ((prev >> sh) | (nxt << (32u - sh))) : prev;
          prev = nxt;
          const uint32_t x = v ^ cur[k];
          for (uint32_t j = 0; j < 4; j++)
            if ((x >> (8u * j)) & 0xFFu) m |= 1u << (4u * k + j);
        }
      } else {
        for (uint32_t k = 0; k < 32; k++) {
          const uint64_t gp = g0 + k;
          const uint32_t b = (cur[k >> 2] >> ((k & 3u) * 8u)) & 0xFFu;
          if (gp < d || J.raw[gp - d] != b) m |= 1u << k;
        }
      }
      S.mm[c][w] = m | ~valid;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (lane == 0) {
    J.seg_sum[2 * s] = (uint32_t)(s1 % kAdlerMod);
    J.seg_sum[2 * s + 1] = (uint32_t)(s2 % kAdlerMod);
  }
  __syncthreads();
  for (uint32_t w = lane; w + 1 < nw; w += 64) {
    uint32_t gm = 0;
    for (uint32_t c = 0; c < nc; c++) {
      const uint32_t m = S.mm[c][w], m1 = S.mm[c][w + 1];
      gm |= ~(m | ((m >> 1) | (m1 << 31)) | ((m >> 2) | (m1 << 30)));
    }
    S.good[w] = gm;
  }
  __syncthreads();

  // greedy parse into match tokens and the literal / match-start bitmaps
  const uint32_t ngw = (n + 31u) >> 5;
  uint32_t p = 0, nt = 0;
  while (p < n) {
    uint32_t q = n;
    for (uint32_t wb = p >> 5; wb < ngw; wb += 64) {
      const uint32_t wi = wb + lane;
      uint32_t g = wi < ngw ? S.good[wi] : 0u;
      if (wi == (p >> 5)) g &= ~0u << (p & 31u);
      const uint64_t bal = __ballot(g != 0u);
      if (bal) {
        const uint32_t first = (uint32_t)__builtin_ctzll(bal);
        const uint32_t gw = __shfl(g, (int)first, 64);
        q = min(n, (wb + first) * 32u + (uint32_t)__builtin_ctz(gw));
        break;
      }
    }
    if (q > p) {  // literal run [p, q)
      for (uint32_t wi = (p >> 5) + lane; wi <= ((q - 1u) >> 5); wi += 64) {
        const uint32_t lo = max(p, wi * 32u) - wi * 32u, hi = min(q, wi * 32u + 32u) - wi * 32u;
        atomicOr(&S.litw[wi], (hi >= 32u ? ~0u : ((1u << hi) - 1u)) & (~0u << lo));
      }
      p = q;
      continue;
    }
    uint32_t len = 0;
    if (lane < nc) {
      uint32_t w = p >> 5;
      uint32_t m = S.mm[lane][w] & (~0u << (p & 31u));
      while (m == 0u && w * 32u + 32u - p < 258u) m = S.mm[lane][++w];
      const uint32_t e = m ? w * 32u + (uint32_t)__builtin_ctz(m) : w * 32u + 32u;
      len = min(e - p, 258u);
    }
    uint32_t key = lane < nc ? (len << 4) | (15u - lane) : 0u;
    for (int off = 32; off > 0; off >>= 1) key = max(key, (uint32_t)__shfl_xor(key, off, 64));
    const uint32_t blen = key >> 4, c = 15u - (key & 15u);
    if (blen < 3u) {  // not reached for a good position; kept as a literal
      if (lane == 0) atomicOr(&S.litw[p >> 5], 1u << (p & 31u));
      p += 1;
      continue;
    }
    if (lane == 0) {
      tok[nt] = (p << 11) | ((blen - 3u) << 3) | c;
      atomicOr(&S.mstw[p >> 5], 1u << (p & 31u));
    }
    nt++;
    p += blen;
  }
  __syncthreads();

  // From here each lane owns the 64 positions [64 lane, 64 lane + 64): its
  // literals (bytes as 16 words), its match starts (tokens tb.. in order).
  const uint64_t LB = S.litw[2 * lane] | (uint64_t)S.litw[2 * lane + 1] << 32;
  const uint64_t MB = S.mstw[2 * lane] | (uint64_t)S.mstw[2 * lane + 1] << 32;
  const uint32_t *r32 = reinterpret_cast<const uint32_t *>(J.raw + base) + lane * 16u;
  uint32_t tb = (uint32_t)__popcll(MB);
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(tb, off, 64);
    if (lane >= (uint32_t)off) tb += y;
  }
  tb -= (uint32_t)__popcll(MB);
  // symbol counts and extra bits
  uint32_t extra = 0;
  {
    uint32_t rw[16];
#pragma unroll
    for (int j = 0; j < 16; j++) rw[j] = r32[j];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const uint32_t lb = (uint32_t)(LB >> (4 * j)) & 15u;
      for (uint32_t b = 0; b < 4; b++)
        if (lb >> b & 1u) atomicAdd(&S.lfreq[(rw[j] >> (8u * b)) & 0xFFu], 1u);
    }
    uint64_t m = MB;
    for (uint32_t t = tb; m; m &= m - 1ull, t++) {
      const uint32_t tk = tok[t];
      uint32_t le, lv, de, dv;
      const uint32_t ls = len_sym(((tk >> 3) & 0xFFu) + 3u, le, lv);
      const uint32_t ds = dist_sym(J.cand[tk & 7u], de, dv);
      atomicAdd(&S.lfreq[ls], 1u);
      atomicAdd(&S.dfreq[ds], 1u);
      extra += le + de;
    }
  }
  for (int off = 32; off > 0; off >>= 1) extra += __shfl_xor(extra, off, 64);
  __syncthreads();

  // block sizes: fixed codes, the segment's own codes, stored
  uint64_t fixed_bits = 0;
  for (uint32_t i = lane; i < 286; i += 64) {
    const uint32_t fl = i < 144 ? 8u : i < 256 ? 9u : i < 280 ? 7u : 8u;
    fixed_bits += (uint64_t)S.lfreq[i] * fl;
  }
  if (lane < 30) fixed_bits += (uint64_t)S.dfreq[lane] * 5u;
  for (int off = 32; off > 0; off >>= 1) fixed_bits += __shfl_xor(fixed_bits, off, 64);
  fixed_bits += 3 + extra;
  uint64_t dyn_bits = ~0ull;
  uint32_t hlit = 257, hdist = 1, hclen = 4;
  const bool ok_l = huff_lengths<8>(S, S.lfreq, 286, 15, S.llen, lane);
  const bool ok_d = huff_lengths<1>(S, S.dfreq, 30, 15, S.dlen, lane);
  if (ok_l && ok_d) {
    for (uint32_t i = lane; i < 286; i += 64)
      if (S.llen[i]) hlit = max(hlit, i + 1);
    if (lane < 30 && S.dlen[lane]) hdist = lane + 1;
    for (int off = 32; off > 0; off >>= 1) {
      hlit = max(hlit, (uint32_t)__shfl_xor((int)hlit, off, 64));
      hdist = max(hdist, (uint32_t)__shfl_xor((int)hdist, off, 64));
    }
    // the code lengths, run-length coded: runs of equal lengths over the
    // literal/length then distance lengths, each run coded by the lane that
    // holds its first position, in position order (a prefix sum per chunk)
    const uint32_t total = hlit + hdist;
    for (uint32_t i = lane; i < 320; i += 64)
      S.cl[i] = (uint8_t)(i < hlit ? S.llen[i] : i < total ? S.dlen[i - hlit] : 0xFFu);
    __syncthreads();
    uint64_t st[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const uint32_t i = 64u * c + lane;
      st[c] = __ballot(i < total && (i == 0 || S.cl[i] != S.cl[i - 1]));
    }
#pragma unroll
    for (int c = 0; c < 5; c++)
      if (total >> 6 == (uint32_t)c) st[c] |= 1ull << (total & 63u);  // end of the last run
    uint32_t nr = 0;
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const uint32_t i = 64u * c + lane;
      uint32_t L = 0;
      if (st[c] >> lane & 1ull && i < total) {
        uint32_t nxt = 0;
        bool found = false;
#pragma unroll
        for (int c2 = c; c2 < 5; c2++) {
          const uint64_t mk = c2 == c ? (lane == 63 ? 0ull : st[c2] & (~0ull << (lane + 1))) : st[c2];
          if (!found && mk) {
            nxt = 64u * c2 + (uint32_t)__builtin_ctzll(mk);
            found = true;
          }
        }
        L = nxt - i;
      }
      const uint32_t v = i < 320 ? S.cl[i] : 0u;
      const uint32_t cnt = L ? rle_run(v, L, nullptr, 0) : 0u;
      uint32_t x = cnt;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
      }
      if (L) rle_run(v, L, S.rle, nr + x - cnt);
      nr += (uint32_t)__shfl((int)x, 63, 64);
    }
    if (lane == 0) S.n_rle = nr;
    __syncthreads();
    for (uint32_t k = lane; k < S.n_rle; k += 64) atomicAdd(&S.cfreq[S.rle[k] & 0xFFu], 1u);
    __syncthreads();
    if (huff_lengths<1>(S, S.cfreq, 19, 7, S.clen, lane)) {
      constexpr uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
      const uint32_t used = lane < 19 && S.clen[order[lane]] ? lane + 1 : 0u;
      hclen = max(4u, used);
      uint64_t bits = 0;
      for (uint32_t k = lane; k < S.n_rle; k += 64)
        bits += S.clen[S.rle[k] & 0xFFu] + ((S.rle[k] >> 16) & 0xFFu);
      for (uint32_t i = lane; i < 286; i += 64) bits += (uint64_t)S.lfreq[i] * S.llen[i];
      if (lane < 30) bits += (uint64_t)S.dfreq[lane] * S.dlen[lane];
      for (int off = 32; off > 0; off >>= 1) {
        bits += __shfl_xor(bits, off, 64);
        hclen = max(hclen, (uint32_t)__shfl_xor((int)hclen, off, 64));
      }
      dyn_bits = bits + 3 + 14 + 3ull * hclen + extra;
    }
  }
  // 0 stored, 1 fixed, 2 own codes; the sync flush or final padding is the
  // same for the coded forms
  uint32_t mode = fixed_bits <= dyn_bits ? 1u : 2u;
  const uint64_t coded = mode == 1u ? fixed_bits : dyn_bits;
  const uint64_t coded_bytes = (coded + 7 + (final_seg ? 0 : 3 + 7 + 32)) / 8;
  if (coded_bytes >= (uint64_t)n + 5u) mode = 0;

  if (mode == 0) {
    if (lane == 0) {
      dst[0] = final_seg ? 1 : 0;  // BFINAL, BTYPE=00
      dst[1] = (uint8_t)(n & 0xFFu);
      dst[2] = (uint8_t)(n >> 8);
      dst[3] = (uint8_t)(~n & 0xFFu);
      dst[4] = (uint8_t)((~n >> 8) & 0xFFu);
      J.seg_len[s] = n + 5u;
    }
    for (uint32_t i = lane; i < n; i += 64) dst[5 + i] = J.raw[base + i];
    return;
  }
  for (uint32_t w = lane; w < kOutWords; w += 64) S.outw[w] = 0;
  __syncthreads();
  // code tables
  if (mode == 1u) {
    for (uint32_t i = lane; i < 288; i += 64)
      S.llen[i] = (uint8_t)(i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8);
    if (lane < 30) S.dlen[lane] = 5;
    __syncthreads();
    huff_codes(S.llen, 288, S.lcode, lane);
    huff_codes(S.dlen, 30, S.dcode, lane);
  } else {
    huff_codes(S.llen, 286, S.lcode, lane);
    huff_codes(S.dlen, 30, S.dcode, lane);
    huff_codes(S.clen, 19, S.ccode, lane);
  }
  uint32_t bp = 3;
  if (lane == 0) put_bits(S.outw, 0, (final_seg ? 1u : 0u) | (mode << 1), 3);
  if (mode == 2u) {
    constexpr uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
    if (lane == 0) put_bits(S.outw, bp, (hlit - 257u) | (hdist - 1u) << 5 | (hclen - 4u) << 10, 14);
    if (lane < hclen) put_bits(S.outw, bp + 14 + 3 * lane, S.clen[order[lane]], 3);
    bp += 14 + 3 * hclen;
    for (uint32_t k0 = 0; k0 < S.n_rle; k0 += 64) {
      const uint32_t k = k0 + lane;
      uint32_t r = 0, cc = 0, nb = 0;
      if (k < S.n_rle) {
        r = S.rle[k];
        cc = S.ccode[r & 0xFFu];
        nb = (cc >> 16) + ((r >> 16) & 0xFFu);
      }
      uint32_t x = nb;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
      }
      if (nb) put_bits(S.outw, bp + x - nb, (cc & 0xFFFFu) | (((r >> 8) & 0xFFu) << (cc >> 16)), nb);
      bp += (uint32_t)__shfl((int)x, 63, 64);
    }
  }
  // each lane's bits, their offsets (a wave prefix sum), then each lane
  // codes its positions in order
  {
    uint32_t rw[16];
#pragma unroll
    for (int j = 0; j < 16; j++) rw[j] = r32[j];
    uint32_t nb = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const uint32_t lb = (uint32_t)(LB >> (4 * j)) & 15u;
      for (uint32_t b = 0; b < 4; b++)
        if (lb >> b & 1u) nb += S.lcode[(rw[j] >> (8u * b)) & 0xFFu] >> 16;
    }
    uint64_t m = MB;
    for (uint32_t t = tb; m; m &= m - 1ull, t++) {
      const uint32_t tk = tok[t];
      uint32_t le, lv, de, dv;
      const uint32_t ls = len_sym(((tk >> 3) & 0xFFu) + 3u, le, lv);
      const uint32_t ds = dist_sym(J.cand[tk & 7u], de, dv);
      nb += (S.lcode[ls] >> 16) + le + (S.dcode[ds] >> 16) + de;
    }
    uint32_t x = nb;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, 64);
      if (lane >= (uint32_t)off) x += y;
    }
    uint32_t b = bp + x - nb;
    bp += (uint32_t)__shfl((int)x, 63, 64);
    uint32_t t = tb;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const uint32_t lb = (uint32_t)(LB >> (4 * j)) & 15u, mb = (uint32_t)(MB >> (4 * j)) & 15u;
      for (uint32_t k = 0; k < 4; k++) {
        if (lb >> k & 1u) {
          const uint32_t lc = S.lcode[(rw[j] >> (8u * k)) & 0xFFu];
          put_bits(S.outw, b, lc & 0xFFFFu, lc >> 16);
          b += lc >> 16;
        } else if (mb >> k & 1u) {
          const uint32_t tk = tok[t++];
          uint32_t le, lv, de, dv;
          const uint32_t ls = len_sym(((tk >> 3) & 0xFFu) + 3u, le, lv);
          const uint32_t ds = dist_sym(J.cand[tk & 7u], de, dv);
          const uint32_t lc = S.lcode[ls], dc = S.dcode[ds];
          const uint32_t l1 = lc >> 16, d1 = dc >> 16;
          put_bits(S.outw, b, (lc & 0xFFFFu) | (lv << l1), l1 + le);
          put_bits(S.outw, b + l1 + le, (dc & 0xFFFFu) | (dv << d1), d1 + de);
          b += l1 + le + d1 + de;
        }
      }
    }
  }
  __syncthreads();
  {
    const uint32_t ec = S.lcode[256];
    if (lane == 0) put_bits(S.outw, bp, ec & 0xFFFFu, ec >> 16);
    bp += ec >> 16;
  }
  if (!final_seg) {
    bp += 3;  // empty stored block: BFINAL=0, BTYPE=00, then LEN=0, NLEN=0xFFFF on a byte
    bp = (bp + 7u) & ~7u;
    if (lane == 0) put_bits(S.outw, bp, 0xFFFF0000u, 32);
    bp += 32;
  } else {
    bp = (bp + 7u) & ~7u;
  }
  __syncthreads();
  const uint32_t nbytes = bp / 8u;
  uint32_t *d32 = reinterpret_cast<uint32_t *>(dst);
  for (uint32_t w = lane; w < (nbytes + 3u) / 4u; w += 64) d32[w] = S.outw[w];
  if (lane == 0) J.seg_len[s] = nbytes;
}

// ---------------------------------------------------------------------------
// k_png_layout: one block.  Segment offsets, Adler-32 (each segment's sums
// S1 = sum b_i and S2 = sum (n - i) b_i fold into the running (a, b) as
// a += S1, b += n*a + S2), lengths, and whether the output fits.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_png_layout(PngJob J) {
  __shared__ uint64_t part[1024];
  __shared__ uint32_t pa[1024], pb[1024], pn[1024];
  const uint32_t t = threadIdx.x, ns = J.n_seg;
  const uint32_t per = (ns + 1023u) / 1024u;
  const uint32_t i0 = min(t * per, ns), i1 = min(i0 + per, ns);
  uint64_t sum = 0;
  uint32_t A = 0, B = 0, N = 0;
  for (uint32_t i = i0; i < i1; i++) {
    sum += J.seg_len[i];
    const uint32_t n =
        (uint32_t)(min((uint64_t)PNG_SEG, J.raw_len - (uint64_t)i * PNG_SEG) % kAdlerMod);
    B = (uint32_t)((B + (uint64_t)n * A + J.seg_sum[2 * i + 1]) % kAdlerMod);
    A = (A + J.seg_sum[2 * i]) % kAdlerMod;
    N = (N + n) % kAdlerMod;
  }
  part[t] = sum;
  pa[t] = A;
  pb[t] = B;
  pn[t] = N;
  __syncthreads();
  // inclusive scan of the byte counts, and the ordered combination of the
  // Adler sums (X then Y: A = Ax + Ay, B = Bx + Ny * Ax + By, N = Nx + Ny),
  // both by doubling steps over the 1024 partials
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    uint64_t v = 0;
    uint32_t xa = 0, xb = 0, xn = 0;
    const bool has = t >= d;
    if (has) {
      v = part[t - d];
      xa = pa[t - d];
      xb = pb[t - d];
      xn = pn[t - d];
    }
    __syncthreads();
    if (has) {
      part[t] += v;
      pb[t] = (uint32_t)((xb + (uint64_t)pn[t] * xa + pb[t]) % kAdlerMod);
      pa[t] = (xa + pa[t]) % kAdlerMod;
      pn[t] = (xn + pn[t]) % kAdlerMod;
    }
    __syncthreads();
  }
  if (t == 1023) {
    const uint64_t run = part[1023];
    // the Adler state starts at (a, b) = (1, 0): the stream's N bytes add N
    // times the initial a to b
    const uint32_t a = (1u + pa[1023]) % kAdlerMod;
    const uint32_t b = (uint32_t)(((uint64_t)pb[1023] + (J.raw_len % kAdlerMod)) % kAdlerMod);
    const uint64_t zlib = 2 + run + 4;
    const uint64_t nidat = (zlib + PNG_IDAT - 1) / PNG_IDAT;
    const uint64_t file = J.head_len + zlib + 12 * nidat + 12;
    J.seg_off[ns] = run;
    J.meta[0] = file;
    J.meta[1] = zlib;
    J.meta[2] = ((uint64_t)b << 16) | a;
    J.meta[3] = file > J.out_cap ? 1 : 0;
  }
  __syncthreads();
  uint64_t run = part[t] - sum;  // exclusive
  for (uint32_t i = i0; i < i1; i++) {
    J.seg_off[i] = run;
    run += J.seg_len[i];
  }
}

__device__ __forceinline__ uint64_t file_pos(const PngJob &J, uint64_t z) {
  return J.head_len + (z / PNG_IDAT) * (PNG_IDAT + 12) + 8 + z % PNG_IDAT;
}

// k_png_copy: one block per segment; block 0 adds the zlib header (CMF 0x78:
// deflate, 32 KiB window; FLG 0x01: fastest, check bits) and the Adler-32.
__global__ void __launch_bounds__(256) k_png_copy(PngJob J) {
  if (J.meta[3]) return;
  const uint32_t s = blockIdx.x, t = threadIdx.x;
  const uint8_t *src = J.seg + (uint64_t)s * PNG_SEG_CAP;
  const uint64_t z0 = 2 + J.seg_off[s];
  const uint32_t len = J.seg_len[s];
  for (uint32_t i = t; i < len; i += 256) J.out[file_pos(J, z0 + i)] = src[i];
  if (s == 0 && t < 6) {
    const uint64_t zlib = J.meta[1];
    const uint32_t adler = (uint32_t)J.meta[2];
    uint64_t z;
    uint8_t v;
    if (t < 2) {
      z = t;
      v = t == 0 ? 0x78 : 0x01;
    } else {
      z = zlib - 4 + (t - 2);
      v = (uint8_t)(adler >> (24u - 8u * (t - 2)));
    }
    J.out[file_pos(J, z)] = v;
  }
}

// CRC-32 polynomial arithmetic for combining the CRCs of pieces (the GF(2)
// method published with zlib's crc32_combine): multiply modulo the reflected
// polynomial, and x^(8n) from a table of x^(2^k).  Both tables are built at
// compile time.
__host__ __device__ constexpr uint32_t multmodp(uint32_t a, uint32_t b) {
  uint32_t m = 1u << 31, p = 0;
  for (;;) {
    if (a & m) {
      p ^= b;
      if ((a & (m - 1u)) == 0) break;
    }
    m >>= 1;
    b = (b & 1u) ? (b >> 1) ^ kCrcPoly : b >> 1;
  }
  return p;
}
struct CrcTables {
  uint32_t tab[256];   // byte-at-a-time table
  uint32_t x2n[32];    // x^(2^k) mod P
  uint32_t piece[256];  // x^(8n) mod P for the bytes after thread i's piece of a full chunk
  uint32_t full;       // x^(8n) mod P, n = PNG_IDAT
  constexpr uint32_t pow8n(uint64_t n) const {  // x^(8n) mod P
    uint32_t p = 1u << 31;                      // x^0
    for (uint32_t k = 3; n; n >>= 1, k++)
      if (n & 1u) p = multmodp(x2n[k & 31u], p);
    return p;
  }
  constexpr CrcTables() : tab(), x2n(), piece(), full(0) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1u) ? (c >> 1) ^ kCrcPoly : c >> 1;
      tab[i] = c;
    }
    uint32_t p = 1u << 30;  // x^1
    x2n[0] = p;
    for (int k = 1; k < 32; k++) {
      p = multmodp(p, p);
      x2n[k] = p;
    }
    for (uint32_t i = 0; i < 256; i++) piece[i] = pow8n((uint64_t)(PNG_IDAT / 256) * (255 - i));
    full = pow8n(PNG_IDAT);
  }
};
static_assert(PNG_IDAT % 256 == 0, "a full IDAT chunk splits into 256 equal pieces");
__constant__ CrcTables kCrc = CrcTables();

// x^(8n) mod P: appending n zero bytes to a CRC register multiplies it by this.
__device__ uint32_t x8nmodp(uint64_t n) {
  uint32_t p = 1u << 31;  // x^0
  uint32_t k = 3;
  while (n) {
    if (n & 1u) p = multmodp(kCrc.x2n[k & 31u], p);
    n >>= 1;
    k++;
  }
  return p;
}

__device__ uint32_t crc_update(const uint32_t *tab, uint32_t c, const uint8_t *p, uint32_t n) {
  uint32_t i = 0;
  for (; i + 16 <= n; i += 16) {  // 16 independent loads in flight, then the table steps
    uint8_t b[16];
#pragma unroll
    for (int k = 0; k < 16; k++) b[k] = p[i + k];
#pragma unroll
    for (int k = 0; k < 16; k++) c = tab[(c ^ b[k]) & 0xFFu] ^ (c >> 8);
  }
  for (; i < n; i++) c = tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
  return c;
}

__device__ void put_be32(uint8_t *o, uint32_t v) {
  o[0] = (uint8_t)(v >> 24);
  o[1] = (uint8_t)(v >> 16);
  o[2] = (uint8_t)(v >> 8);
  o[3] = (uint8_t)v;
}

// A whole small chunk (length, type, data, CRC) at o; returns its size.
__device__ uint32_t put_chunk(const uint32_t *tab, uint8_t *o, const char *type,
                              const uint8_t *data, uint32_t n) {
  put_be32(o, n);
  for (int i = 0; i < 4; i++) o[4 + i] = (uint8_t)type[i];
  for (uint32_t i = 0; i < n; i++) o[8 + i] = data[i];
  const uint32_t crc = crc_update(tab, 0xFFFFFFFFu, o + 4, 4 + n) ^ 0xFFFFFFFFu;
  put_be32(o + 8 + n, crc);
  return 12 + n;
}

// k_png_crc: block b < IDAT count frames chunk b (256 threads: each thread's
// piece CRC register, started at zero, is shifted to the chunk's end and the
// 256 terms XOR together, the type bytes' register shifted likewise; the
// shifts of a full chunk come from a compile-time table); the last
// block writes the signature, IHDR, gAMA and IEND (src/export_png.zig:86-145,
// :422-446).
__global__ void __launch_bounds__(256) k_png_crc(PngJob J) {
  __shared__ uint32_t tab[256];
  __shared__ uint32_t part[4];
  if (J.meta[3]) return;
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  tab[tid] = kCrc.tab[tid];
  __syncthreads();
  const uint64_t file = J.meta[0], zlib = J.meta[1];
  const uint64_t nidat = (zlib + PNG_IDAT - 1) / PNG_IDAT;
  const uint32_t b = blockIdx.x;
  if (b == gridDim.x - 1) {
    if (tid != 0) return;
    uint8_t *o = J.out;
    const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    for (int i = 0; i < 8; i++) o[i] = sig[i];
    uint8_t ihdr[13];
    put_be32(ihdr, (uint32_t)J.width);
    put_be32(ihdr + 4, (uint32_t)J.height);
    const uint32_t f = J.format;
    ihdr[8] = f == Z2D_FORMAT_ALPHA4 ? 4 : f == Z2D_FORMAT_ALPHA2 ? 2 : f == Z2D_FORMAT_ALPHA1 ? 1 : 8;
    ihdr[9] = (f == Z2D_FORMAT_ARGB || f == Z2D_FORMAT_RGBA) ? 6
              : (f == Z2D_FORMAT_XRGB || f == Z2D_FORMAT_RGB) ? 2
                                                              : 0;
    ihdr[10] = ihdr[11] = ihdr[12] = 0;  // deflate, filter method 0, no interlace
    uint32_t at = 8 + put_chunk(tab, o + 8, "IHDR", ihdr, 13);
    if (J.profile != Z2D_PNG_PROFILE_NONE) {
      const float gamma = J.profile == Z2D_PNG_PROFILE_SRGB ? 2.2f : 1.0f;
      uint8_t g[4];
      put_be32(g, (uint32_t)((1.0f / gamma) * 100000.0f));
      at += put_chunk(tab, o + at, "gAMA", g, 4);
    }
    put_chunk(tab, o + file - 12, "IEND", nullptr, 0);
    return;
  }
  if (b >= nidat) return;
  const uint64_t co = J.head_len + (uint64_t)b * (PNG_IDAT + 12);
  const uint32_t len = (uint32_t)min((uint64_t)PNG_IDAT, zlib - (uint64_t)b * PNG_IDAT);
  uint8_t *o = J.out + co;
  const uint32_t P = (len + 255u) / 256u;
  const uint32_t lo = min(tid * P, len), hi = min(lo + P, len);
  uint32_t term = 0;
  if (hi > lo) {
    const uint32_t after = len == PNG_IDAT ? kCrc.piece[tid] : x8nmodp(len - hi);
    term = multmodp(after, crc_update(tab, 0u, o + 8 + lo, hi - lo));
  }
  for (int off = 32; off > 0; off >>= 1) term ^= __shfl_xor(term, off, 64);
  if (lane == 0) part[tid >> 6] = term;
  __syncthreads();
  if (tid == 0) {
    term = part[0] ^ part[1] ^ part[2] ^ part[3];
    const uint8_t type[4] = {'I', 'D', 'A', 'T'};
    const uint32_t head = crc_update(tab, 0xFFFFFFFFu, type, 4);
    const uint32_t crc = ~(multmodp(len == PNG_IDAT ? kCrc.full : x8nmodp(len), head) ^ term);
    put_be32(o, len);
    for (int i = 0; i < 4; i++) o[4 + i] = type[i];
    put_be32(o + 8 + len, crc);
  }
}

}  // namespace
}  // namespace z2d

using namespace z2d;

hipError_t z2d_launch_png(const PngJob &J, hipStream_t st) {
  const uint64_t per_row = J.format <= Z2D_FORMAT_ALPHA8 ? (uint64_t)J.width : J.row_bytes;
  const uint64_t units = per_row * (uint64_t)J.height;
  const uint32_t rows_grid = (uint32_t)std::min<uint64_t>((units + 255) / 256, 16384);
  hipLaunchKernelGGL(k_png_rows, dim3(rows_grid), dim3(256), 0, st, J);
  hipLaunchKernelGGL(k_png_deflate, dim3(J.n_seg), dim3(64), 0, st, J);
  hipLaunchKernelGGL(k_png_layout, dim3(1), dim3(1024), 0, st, J);
  hipLaunchKernelGGL(k_png_copy, dim3(J.n_seg), dim3(256), 0, st, J);
  hipLaunchKernelGGL(k_png_crc, dim3(J.max_idat + 1), dim3(256), 0, st, J);
  return hipGetLastError();
}

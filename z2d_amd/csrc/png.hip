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
// candidate repeats three bytes is "good".  The wave then parses greedily:
// literal runs up to the next good position are coded 64 bytes at a time
// (one lane each, offsets by a wave prefix sum); at a good position each
// candidate lane measures its run from its bitmap and the longest (earliest
// candidate on ties) becomes a match of up to 258 bytes.  Codes go into an
// LDS bit buffer with atomic ORs.  The fixed-Huffman form is dropped for a
// stored block as soon as it grows past the stored size.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put_bits(uint32_t *w, uint32_t pos, uint32_t v, uint32_t nb) {
  const uint32_t k = pos >> 5, sh = pos & 31u;
  atomicOr(&w[k], v << sh);
  if (sh + nb > 32u) atomicOr(&w[k + 1], v >> (32u - sh));
}

// Fixed literal/length code (RFC 1951 3.2.6), bit-reversed for LSB-first
// packing; returns the bit count.
__device__ __forceinline__ uint32_t lit_code(uint32_t b, uint32_t &code) {
  if (b < 144u) {
    code = __builtin_bitreverse32(0x30u + b) >> 24;
    return 8u;
  }
  code = __builtin_bitreverse32(0x190u + b - 144u) >> 23;
  return 9u;
}

// A match of `len` (3..258) at distance `dist` (1..32768) as one bit string.
__device__ __forceinline__ uint32_t match_code(uint32_t len, uint32_t dist, uint32_t &bits) {
  uint32_t sym, lext = 0, lval = 0;
  if (len == 258u) {
    sym = 285u;
  } else {
    const uint32_t v = len - 3u;
    if (v < 8u) {
      sym = 257u + v;
    } else {
      const uint32_t e = 31u - __builtin_clz(v) - 2u;
      sym = 257u + 4u * (e + 1u) + ((v >> e) & 3u);
      lext = e;
      lval = v & ((1u << e) - 1u);
    }
  }
  uint32_t lcode, lnb;
  if (sym <= 279u) {
    lcode = __builtin_bitreverse32(sym - 256u) >> 25;
    lnb = 7u;
  } else {
    lcode = __builtin_bitreverse32(0xC0u + sym - 280u) >> 24;
    lnb = 8u;
  }
  uint32_t dc, dext = 0, dval = 0;
  const uint32_t dv = dist - 1u;
  if (dv < 4u) {
    dc = dv;
  } else {
    const uint32_t e = 31u - __builtin_clz(dv) - 1u;
    dc = 2u * (e + 1u) + ((dv >> e) & 1u);
    dext = e;
    dval = dv & ((1u << e) - 1u);
  }
  const uint32_t dcode = __builtin_bitreverse32(dc) >> 27;
  uint32_t v = lcode;
  uint32_t n = lnb;
  v |= lval << n;
  n += lext;
  v |= dcode << n;
  n += 5u;
  v |= dval << n;
  n += dext;
  bits = n;  // <= 8 + 5 + 5 + 13 = 31
  return v;
}

__global__ void __launch_bounds__(64) k_png_deflate(PngJob J) {
  __shared__ uint32_t mm[PNG_MAX_CAND][kMatchWords];
  __shared__ uint32_t good[kMatchWords];
  __shared__ uint32_t outw[kOutWords];
  const uint32_t lane = threadIdx.x;
  const uint32_t s = blockIdx.x;
  const uint64_t base = (uint64_t)s * PNG_SEG;
  const uint32_t n = (uint32_t)min((uint64_t)PNG_SEG, J.raw_len - base);
  const bool final_seg = s + 1 == J.n_seg;
  const uint32_t nw = (n >> 5) + 2;  // bitmap words written (<= kMatchWords)
  const uint32_t nc = J.n_cand;
  for (uint32_t w = lane; w < kOutWords; w += 64) outw[w] = 0;

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
      mm[c][w] = m | ~valid;
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
      const uint32_t m = mm[c][w], m1 = mm[c][w + 1];
      gm |= ~(m | ((m >> 1) | (m1 << 31)) | ((m >> 2) | (m1 << 30)));
    }
    good[w] = gm;
  }
  __syncthreads();

  // greedy parse and fixed-Huffman coding
  const uint32_t limit = (n + 5u) * 8u;  // bits of the stored form
  const uint32_t ngw = (n + 31u) >> 5;
  if (lane == 0) outw[0] = (final_seg ? 1u : 0u) | (1u << 1);  // BFINAL, BTYPE=01
  uint32_t bp = 3;
  uint32_t p = 0;
  bool over = false;
  while (p < n) {
    uint32_t q = n;
    for (uint32_t wb = p >> 5; wb < ngw; wb += 64) {
      const uint32_t wi = wb + lane;
      uint32_t g = wi < ngw ? good[wi] : 0u;
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
      for (uint32_t i0 = p; i0 < q; i0 += 64) {
        const uint32_t i = i0 + lane;
        uint32_t code = 0, nb = 0;
        if (i < q) nb = lit_code(J.raw[base + i], code);
        uint32_t x = nb;
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t y = __shfl_up(x, off, 64);
          if (lane >= (uint32_t)off) x += y;
        }
        const uint32_t tot = __shfl(x, 63, 64);
        if (nb) put_bits(outw, bp + x - nb, code, nb);
        bp += tot;
        if (bp > limit) {
          over = true;
          break;
        }
      }
      if (over) break;
      p = q;
      continue;
    }
    uint32_t len = 0;
    if (lane < nc) {
      uint32_t w = p >> 5;
      uint32_t m = mm[lane][w] & (~0u << (p & 31u));
      while (m == 0u && w * 32u + 32u - p < 258u) m = mm[lane][++w];
      const uint32_t e = m ? w * 32u + (uint32_t)__builtin_ctz(m) : w * 32u + 32u;
      len = min(e - p, 258u);
    }
    uint32_t key = lane < nc ? (len << 4) | (15u - lane) : 0u;
    for (int off = 32; off > 0; off >>= 1) key = max(key, (uint32_t)__shfl_xor(key, off, 64));
    const uint32_t blen = key >> 4, c = 15u - (key & 15u);
    if (blen < 3u) {  // not reached for a good position; kept as a literal
      uint32_t code;
      const uint32_t nb = lit_code(J.raw[base + p], code);
      if (lane == 0) put_bits(outw, bp, code, nb);
      bp += nb;
      p += 1;
    } else {
      uint32_t nb;
      const uint32_t v = match_code(blen, J.cand[c], nb);
      if (lane == 0) put_bits(outw, bp, v, nb);
      bp += nb;
      p += blen;
    }
    if (bp > limit) {
      over = true;
      break;
    }
  }
  if (!over) {
    bp += 7;  // end of block (code 0000000)
    if (!final_seg) {
      bp += 3;  // empty stored block: BFINAL=0, BTYPE=00, then LEN=0, NLEN=0xFFFF on a byte
      bp = (bp + 7u) & ~7u;
      if (lane == 0) put_bits(outw, bp, 0xFFFF0000u, 32);
      bp += 32;
    } else {
      bp = (bp + 7u) & ~7u;
    }
    if (bp / 8u >= n + 5u) over = true;
  }
  __syncthreads();
  uint8_t *dst = J.seg + (uint64_t)s * PNG_SEG_CAP;
  if (!over) {
    const uint32_t nbytes = bp / 8u;
    uint32_t *d32 = reinterpret_cast<uint32_t *>(dst);
    for (uint32_t w = lane; w < (nbytes + 3u) / 4u; w += 64) d32[w] = outw[w];
    if (lane == 0) J.seg_len[s] = nbytes;
  } else {
    if (lane == 0) {
      dst[0] = final_seg ? 1 : 0;  // BFINAL, BTYPE=00
      dst[1] = (uint8_t)(n & 0xFFu);
      dst[2] = (uint8_t)(n >> 8);
      dst[3] = (uint8_t)(~n & 0xFFu);
      dst[4] = (uint8_t)((~n >> 8) & 0xFFu);
      J.seg_len[s] = n + 5u;
    }
    for (uint32_t i = lane; i < n; i += 64) dst[5 + i] = J.raw[base + i];
  }
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
  uint32_t tab[256];  // byte-at-a-time table
  uint32_t x2n[32];   // x^(2^k) mod P
  constexpr CrcTables() : tab(), x2n() {
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
  }
};
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

// k_png_crc: block b < IDAT count frames chunk b (one wave: each lane's
// piece CRC register, started at zero, is shifted to the chunk's end and the
// 64 terms XOR together, the type bytes' register shifted likewise); the last
// block writes the signature, IHDR, gAMA and IEND (src/export_png.zig:86-145,
// :422-446).
__global__ void __launch_bounds__(64) k_png_crc(PngJob J) {
  __shared__ uint32_t tab[256];
  if (J.meta[3]) return;
  const uint32_t lane = threadIdx.x;
  for (uint32_t i = lane; i < 256; i += 64) tab[i] = kCrc.tab[i];
  __syncthreads();
  const uint64_t file = J.meta[0], zlib = J.meta[1];
  const uint64_t nidat = (zlib + PNG_IDAT - 1) / PNG_IDAT;
  const uint32_t b = blockIdx.x;
  if (b == gridDim.x - 1) {
    if (lane != 0) return;
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
  const uint32_t P = (len + 63u) / 64u;
  const uint32_t lo = min(lane * P, len), hi = min(lo + P, len);
  uint32_t term = 0;
  if (hi > lo) term = multmodp(x8nmodp(len - hi), crc_update(tab, 0u, o + 8 + lo, hi - lo));
  for (int off = 32; off > 0; off >>= 1) term ^= __shfl_xor(term, off, 64);
  if (lane == 0) {
    const uint8_t type[4] = {'I', 'D', 'A', 'T'};
    const uint32_t head = crc_update(tab, 0xFFFFFFFFu, type, 4);
    const uint32_t crc = ~(multmodp(x8nmodp(len), head) ^ term);
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
  hipLaunchKernelGGL(k_png_crc, dim3(J.max_idat + 1), dim3(64), 0, st, J);
  return hipGetLastError();
}

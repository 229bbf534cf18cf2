// png.h - device PNG export (src/export_png.zig:56-360): the job record shared
// by the host entry point (host.cpp) and the kernels (png.hip).
//
// Pipeline, all on the device:
//   k_png_rows     surface -> scanline stream (filter byte 0 + encoded row),
//                  the bytes the reference feeds its zlib compressor
//   k_png_deflate  one wave per PNG_SEG-byte segment of the stream: greedy
//                  LZ77 over a few candidate distances (pixel and row
//                  neighbours) coded with Huffman codes built for the
//                  segment, the fixed codes, or a stored block, whichever is
//                  smallest; each segment ends on a byte (empty stored block)
//                  so segments concatenate
//   k_png_layout   segment offsets (exclusive scan), Adler-32 of the stream
//                  from per-segment sums, zlib and file lengths
//   k_png_copy     segments into the file, framed as IDAT chunks of
//                  PNG_IDAT zlib bytes, plus the zlib header and trailer
//   k_png_crc      one block per IDAT chunk: CRC-32 by pieces combined in
//                  GF(2); one more wave writes signature, IHDR, gAMA, IEND
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace z2d {

constexpr uint32_t PNG_SEG = 4096;               // stream bytes per deflate segment
constexpr uint32_t PNG_SEG_CAP = 5504;  // staging bytes per segment: >= 4 * (PNG_SEG / 3 + 2) token words, >= stored size
constexpr uint32_t PNG_IDAT = 16384;             // zlib bytes per IDAT chunk
constexpr uint32_t PNG_MAX_CAND = 8;             // LZ77 candidate distances
constexpr uint32_t PNG_RAW_PAD = PNG_SEG + 256;  // readable bytes past the stream's end

struct PngJob {
  const uint8_t *surface;
  uint32_t format;
  uint32_t profile;  // z2d_png_profile
  int32_t width, height;
  uint64_t row_bytes;  // encoded bytes per row, without the filter byte
  uint64_t raw_len;    // height * (1 + row_bytes)
  uint8_t *raw;        // raw_len + PNG_RAW_PAD bytes
  uint32_t n_seg;
  uint32_t n_cand;
  uint32_t cand[PNG_MAX_CAND];  // match distances, most preferred first
  uint8_t *seg;                 // n_seg * PNG_SEG_CAP staging bytes
  uint32_t *seg_len;            // n_seg compressed lengths
  uint32_t *seg_sum;            // 2 * n_seg Adler-32 partial sums
  uint64_t *seg_off;            // n_seg + 1 offsets
  uint64_t *meta;  // [0] file bytes, [1] zlib bytes, [2] Adler-32, [3] 1 if out_cap is short
  uint8_t *out;
  uint64_t out_cap;
  uint32_t head_len;   // signature + IHDR (+ gAMA)
  uint32_t max_idat;   // IDAT chunk count bound (grid of k_png_crc)
};

}  // namespace z2d

// Dword spans of a wave's contiguous particle rows.
//
// zpc stores each particle attribute as an array of rows (x: 3 floats, F and
// C: 9 floats per particle). A wave that owns particles q0 .. q0+cnt-1 moves
// their rows as one dword range [a, b) = [W q0, W (q0 + cnt)). The range is
// split into at most 3 leading dwords, float4 vectors on 16-byte boundaries,
// and at most 3 trailing dwords, so every global access of the wave is a
// full 16-byte vector except a few edge lanes. The rows pass through a small
// LDS stage; `shift` places the vector part of the stage on 16-byte
// boundaries too. This header has no HIP dependency so the span arithmetic
// can be checked on the host.
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#define ZS_HD __host__ __device__
#else
#define ZS_HD
#endif

namespace zsmpm {

struct RowSpan {
  int64_t a;     // first dword (relative to the attribute's base pointer)
  int64_t b;     // one past the last dword
  int64_t a4;    // first dword on a 16-byte boundary, clamped to b
  int64_t nvec;  // float4 vectors in [a4, a4 + 4 nvec)
  int32_t shift; // stage index of dword d is (d - a) + shift
};

// base_addr: the attribute's byte address (4-byte aligned)
ZS_HD inline RowSpan row_span(uint64_t base_addr, int64_t q0, int32_t cnt, int32_t W) {
  RowSpan r;
  r.a = (int64_t)W * q0;
  r.b = r.a + (int64_t)W * cnt;
  const int64_t mis = (int64_t)((base_addr >> 2) & 3u);  // dword offset of base within 16 bytes
  int64_t a4 = r.a + ((4 - ((mis + r.a) & 3)) & 3);
  if (a4 > r.b) a4 = r.b;
  r.a4 = a4;
  r.nvec = (r.b - a4) / 4;
  r.shift = (int32_t)((4 - ((a4 - r.a) & 3)) & 3);
  return r;
}

// floats a per-wave stage must hold for rows of W floats (64 rows + shift)
constexpr int32_t stage_floats(int32_t W) { return 64 * W + 4; }

}  // namespace zsmpm

// Host check of zsmpm::row_span (zpc_amd/csrc/rowspan.hpp), the dword split
// the wave-cooperative row loads and stores use: emulates the 64 lanes of
// wave_load_rows / wave_store_rows and checks that every dword of the rows is
// moved exactly once, inside the rows, that vector accesses sit on 16-byte
// boundaries in memory and in the stage, and that the stage stays in bounds.
#include <cstdio>
#include <vector>

#include "../zpc_amd/csrc/rowspan.hpp"

using namespace zsmpm;

static int check(uint64_t base, int64_t q0, int cnt, int W) {
  const RowSpan r = row_span(base, q0, cnt, W);
  const int64_t a = (int64_t)W * q0, b = a + (int64_t)W * cnt;
  if (r.a != a || r.b != b) return 1;
  std::vector<int> seen(W * 64 + 8, 0);
  const int stage_n = stage_floats(W);
  auto touch = [&](int64_t d, int64_t st) -> int {
    if (d < a || d >= b) return 2;
    if (st < 0 || st >= stage_n) return 3;
    if (st != (d - a) + r.shift) return 4;
    seen[d - a] += 1;
    return 0;
  };
  const int64_t head = r.a4 - r.a;
  if (head < 0 || head > 3) return 5;
  // RowLoad keeps (16 W + 63) / 64 float4 per lane: every vector index fits
  const int64_t kvec = (16 * W + 63) / 64;
  if (r.nvec > kvec * 64) return 11;
  for (int lane = 0; lane < 64; ++lane) {
    if (lane < head)
      if (int e = touch(r.a + lane, r.shift + lane)) return e;
    for (int64_t j = lane; j < r.nvec; j += 64) {
      const int64_t d = r.a4 + 4 * j;
      if (((base + 4 * (uint64_t)d) & 15u) != 0) return 6;  // 16-byte aligned in memory
      const int64_t st = r.shift + head + 4 * j;
      if (st % 4 != 0) return 7;                            // and in the stage
      for (int c = 0; c < 4; ++c)
        if (int e = touch(d + c, st + c)) return e;
    }
    const int64_t t0 = r.a4 + 4 * r.nvec;
    if (r.b - t0 < 0 || r.b - t0 > 3) return 8;
    if (lane < r.b - t0)
      if (int e = touch(t0 + lane, r.shift + (t0 - r.a) + lane)) return e;
  }
  for (int64_t i = 0; i < b - a; ++i)
    if (seen[i] != 1) return 9;
  // each lane's row lies inside the stage
  for (int lane = 0; lane < cnt; ++lane)
    if (r.shift + lane * W + W > stage_n) return 10;
  return 0;
}

int main() {
  long cases = 0;
  const int widths[] = {1, 3, 9};
  for (int W : widths)
    for (uint64_t mis = 0; mis < 4; ++mis)
      for (int64_t q0 = 0; q0 < 40; ++q0)
        for (int cnt = 1; cnt <= 64; ++cnt) {
          const uint64_t base = 0x7f0000000000ull + 4 * mis;
          if (int e = check(base, q0, cnt, W)) {
            std::printf("FAIL W=%d mis=%llu q0=%lld cnt=%d code=%d\n", W, (unsigned long long)mis,
                        (long long)q0, cnt, e);
            return 1;
          }
          ++cases;
        }
  // large row numbers (offsets beyond 2^31 dwords)
  for (int W : widths)
    for (int cnt = 1; cnt <= 64; cnt += 7)
      if (int e = check(0x7f0000000004ull, (int64_t)3 << 30, cnt, W)) {
        std::printf("FAIL large W=%d cnt=%d code=%d\n", W, cnt, e);
        return 1;
      }
  std::printf("ok %ld\n", cases);
  return 0;
}

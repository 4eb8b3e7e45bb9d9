// Lists the floats above 2^18 pi for which fdlibm's float reduction
// (__kernel_rem_pio2f, modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:234-420)
// takes more than 11 columns (J = 12 or 13): arguments within about 2^-24 of a
// multiple of pi/2, where the kernels' column masks for the last frame words
// are used.  Written as a byte-by-byte restatement of the column loop
// (carry the column sums q[i] into bytes from column J up, quadrant,
// complement, "more columns" while the leading bytes vanish), independent of
// include/nt2hip/math.hpp; the 2/pi bytes come from the oracle's table.
// Output: one line per float, "J bits" (bits as 0x%08x), every float with
// J = 13 and every 8th with J = 12.  Used by make_fdlibm_wide_columns.py.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../oracle/nt2ref.hpp"

static uint8_t P[200];

static int columns(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  const int e0 = (int)(b >> 23) - 134;
  const uint32_t m = (b & 0x7fffff) | 0x800000;
  const uint32_t t0 = m >> 16, t1 = (m >> 8) & 255, t2 = m & 255;
  const int jv = (e0 - 3) / 8;
  const int q0 = e0 - 8 * (jv + 1);
  uint32_t q[25];
  for (int i = 0; i <= 24; ++i) {
    const int p = jv + i;
    q[i] = t0 * P[p] + t1 * (p >= 1 ? P[p - 1] : 0) + t2 * (p >= 2 ? P[p - 2] : 0);
  }
  int J = 9;
  for (;;) {
    uint32_t B[25] = {0}, c = 0;
    for (int i = J; i >= 1; --i) {
      const uint32_t v = q[i] + c;
      B[i] = v & 255;
      c = v >> 8;
    }
    const uint32_t top = q[0] + c;
    uint32_t zf = q0 < 0 ? top & ((1u << -q0) - 1) : 0;
    int ih = 0;
    if (q0 > 0) {
      B[1] &= (1u << (8 - q0)) - 1;
      ih = (int)(B[1] >> (7 - q0));
    } else if (q0 == 0) {
      ih = (int)(B[1] >> 7);
    } else if (zf >= (1u << (-q0 - 1))) {
      ih = 2;
    }
    if (ih) {
      uint32_t borrow = 0;
      for (int i = J; i >= 1; --i) {
        const uint32_t j = B[i];
        B[i] = borrow ? 255 - j : (j ? 256 - j : 0);
        borrow |= j != 0;
      }
      if (q0 > 0) B[1] &= 0xff >> q0;
      if (ih == 2) zf = (1u << -q0) - zf - borrow;
    }
    if (zf != 0) return J;
    uint32_t lead = 0;
    for (int i = 1; i <= J - 9; ++i) lead |= B[i];
    if (lead != 0) return J;
    int s = 0;
    for (int i = J; i >= 1; --i)
      if (B[i]) s = i;
    if (s == 0 || 9 + s > 24) return -1;
    J = 9 + s;
  }
}

int main() {
  for (int i = 0; i < 66; ++i) {
    const int32_t w = nt2ref::trig::two_over_pi_24[i];
    P[3 * i] = (w >> 16) & 255;
    P[3 * i + 1] = (w >> 8) & 255;
    P[3 * i + 2] = w & 255;
  }
  float lo = 0x1.921fb6p+19f;
  uint32_t b0;
  std::memcpy(&b0, &lo, 4);
  long n12 = 0;
  for (uint64_t b = b0 + 1; b <= 0x7f7fffffu; ++b) {
    float x;
    const uint32_t u = (uint32_t)b;
    std::memcpy(&x, &u, 4);
    const int J = columns(x);
    if (J == 13 || J < 0 || (J == 12 && (n12++ % 8) == 0)) std::printf("%d 0x%08x\n", J, u);
  }
  return 0;
}

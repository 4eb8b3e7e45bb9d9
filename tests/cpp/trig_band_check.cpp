// Exhaustive check of the float sin/cos band (2^6 pi, 2^28] of the EXACT
// policy.  The GPU reduces every float of the band with two double FMAs:
//   k = rint(x * 2/pi),  r = fma(-k, P1, x),  r = fma(-k, P2, r),  xr = (float)r
// with pi/2 = P1 + P2 + O(2^-107).  This program compares, on every float
// of the band:
//  * up to 2^18 pi (float), with NT2's own reduction as the oracle restates
//    it (trig_reduction.hpp:94-247: the double medium reduction of
//    generic/rem_pio2_medium.hpp:68-83, fdlibm's double medium for the last
//    float) — xr's bits and the quadrant must agree;
//  * above, up to 2^28, with the correctly rounded x mod pi/2 and the
//    nearest quadrant, computed in binary128 (error below 2^-84; a float
//    whose reduced argument lies within 2^-80 of a rounding midpoint or of
//    pi/4 would be reported as undecided).
// It prints the counts; the test requires no difference and nothing
// undecided.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <quadmath.h>

#include "../../oracle/nt2ref.hpp"

static float fbits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

static int two_fma(float x, float* xr) {
  const double xd = (double)x;
  const double k = std::rint(xd * 0x1.45f306dc9c883p-1);
  double r = std::fma(-k, 0x1.921fb54442d18p+0, xd);
  r = std::fma(-k, 0x1.1a62633145c07p-54, r);
  *xr = (float)r;
  return ((int32_t)k) & 3;
}

// correctly rounded reduction in binary128; returns -1 if undecided
static int quad_reduce(float x, float* xr) {
  const __float128 pio2 = M_PI_2q;
  const __float128 xq = x;
  const __float128 kq = rintq(xq / pio2);
  const __float128 rq = xq - kq * pio2;
  if (fabsq(fabsq(rq) - pio2 / 2) < 0x1p-80Q) return -1;
  const float c = (float)rq;
  const float c2 = std::nextafter(c, rq > (__float128)c ? INFINITY : -INFINITY);
  if (fabsq(rq - ((__float128)c + (__float128)c2) / 2) < 0x1p-80Q) return -1;
  *xr = c;
  return (int)(((int64_t)kq) & 3);
}

int main() {
  const uint32_t lo = ubits(0x1.921fb6p+7f) + 1, mid = ubits(0x1.921fb6p+19f), hi = ubits(0x1.0p+28f);
  long n1 = 0, bad1 = 0;
  for (uint32_t b = lo; b <= mid; ++b) {
    const float x = fbits(b);
    float ref, a;
    const int qr = nt2ref::trig::reduce(x, ref);
    const int qa = two_fma(x, &a);
    ++n1;
    if (qr != qa || ubits(ref) != ubits(a)) {
      if (bad1 < 5) std::printf("differ from NT2 at %a: %a q%d vs %a q%d\n", x, ref, qr, a, qa);
      ++bad1;
    }
  }
  long n2 = 0, bad2 = 0, und = 0;
#pragma omp parallel for schedule(dynamic, 65536) reduction(+ : n2, bad2, und)
  for (int64_t b = (int64_t)mid + 1; b <= (int64_t)hi; ++b) {
    const float x = fbits((uint32_t)b);
    float ref, a;
    const int qr = quad_reduce(x, &ref);
    const int qa = two_fma(x, &a);
    ++n2;
    if (qr < 0) { ++und; continue; }
    if (qr != qa || ubits(ref) != ubits(a)) ++bad2;
  }
  std::printf("NT2 part: %ld floats in (%a, %a]: %ld differ\n", n1, 0x1.921fb6p+7f, fbits(mid), bad1);
  std::printf("correctly rounded part: %ld floats in (%a, %a]: %ld differ, %ld undecided\n", n2, fbits(mid),
              fbits(hi), bad2, und);
  return (bad1 == 0 && bad2 == 0 && und == 0) ? 0 : 1;
}

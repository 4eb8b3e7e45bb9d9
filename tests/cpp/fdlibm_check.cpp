// Exhaustive check of the EXACT float sin/cos reduction above 2^18 pi.
// include/nt2hip/math.hpp's reduce_fdlibm_f (the code the GPU kernels run,
// compiled here for the host from the same header) is compared with the
// oracle's statement-by-statement restatement of fdlibm's
// __ieee754_rem_pio2f (oracle/rem_pio2.cpp, after NT2's
// modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:83-233)
// on every finite float above 2^18 pi: the quadrant (n & 3) and the bits of
// the head y[0] must agree.  An optional argument walks every k-th float
// instead.  The double reduction above 2^20 pi/2 (reduce_fdlibm_d against
// the oracle's __ieee754_rem_pio2, rem_pio2.hpp:484-913) is compared on
// 2^16 random doubles of every binade from 2^20 to 2^1023, on the doubles
// nearest to k pi/2 (k pi/2 rounded from binary128, and two neighbours each side)
// for k up to 2^60, and around the double closest to a multiple of pi/2
// (6381956970095103 2^797, |x mod pi/2| ~ 4.7e-19).  Prints the counts; exit
// status 0 when nothing differs.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <random>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/nt2hip/math.hpp"
#include "../../oracle/nt2ref.hpp"

static float fbits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static uint64_t dbits(double d) { uint64_t u; std::memcpy(&u, &d, 8); return u; }
static double dfrom(uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }

// one double against the oracle; returns 1 if they differ
static int check_d(double x, std::atomic<int>& shown) {
  double y[2];
  const int nq = nt2ref::trig::ieee754_rem_pio2d(x, y) & 3;
  double a;
  const int qa = nt2h::reduce_fdlibm_d(x, &a);
  if (qa == nq && dbits(a) == dbits(y[0])) return 0;
  if (shown++ < 3) std::printf("double differs at %a: oracle %a q%d, kernel %a q%d\n", x, y[0], nq, a, qa);
  return 1;
}

static int double_checks(unsigned nt) {
  std::atomic<uint64_t> n{0}, bad{0};
  std::atomic<int> shown{0};
  std::vector<std::thread> th;
  const double lo = 0x1.921fb54442d18p+20;  // 2^20 pi/2
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      uint64_t ln = 0, lbad = 0;
      std::mt19937_64 rng(12345 + t);
      // random mantissas in every binade
      for (int e = 20 + (int)t; e <= 1023; e += (int)nt)
        for (int i = 0; i < (1 << 16); ++i) {
          const double x = dfrom(((uint64_t)(e + 1023) << 52) | (rng() & 0xfffffffffffffull));
          if (!(x > lo)) continue;
          ++ln;
          lbad += check_d(x, shown);
        }
      // the doubles nearest to k pi/2 and their neighbours
      // pi/2 to about 107 bits as a binary128 sum of two doubles
      const __float128 pio2 = (__float128)0x1.921fb54442d18p+0 + (__float128)0x1.1a62633145c07p-54;
      for (int i = 0; i < (1 << 18); ++i) {
        const int bitsk = 21 + (int)(rng() % 40);
        const uint64_t k = (rng() >> (64 - bitsk)) | (1ull << (bitsk - 1));
        const double c = (double)((__float128)k * pio2);
        for (int d = -2; d <= 2; ++d) {
          const double x = dfrom(dbits(c) + (int64_t)d);
          if (!(x > lo)) continue;
          ++ln;
          lbad += check_d(x, shown);
        }
      }
      if (t == 0) {  // the double closest to a multiple of pi/2
        const double w = std::ldexp(6381956970095103.0, 797);
        for (int d = -64; d <= 64; ++d) {
          ++ln;
          lbad += check_d(dfrom(dbits(w) + (int64_t)d), shown);
        }
        ++ln;
        lbad += check_d(0x1.fffffffffffffp+1023, shown);
      }
      n += ln;
      bad += lbad;
    });
  for (auto& x : th) x.join();
  std::printf("fdlibm double reduction: %llu doubles above 2^20 pi/2 (sampled): %llu differ\n",
              (unsigned long long)n.load(), (unsigned long long)bad.load());
  return bad.load() == 0 ? 0 : 1;
}

int main(int argc, char** argv) {
  const uint32_t stride = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 0) : 1u;
  const uint32_t lo = ubits(nt2h::kFdlibmLoF) + 1, hi = 0x7f7fffffu;
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  std::atomic<uint64_t> n{0}, bad{0}, nband{0}, fast{0}, fastbad{0};
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      uint64_t ln = 0, lbad = 0, lband = 0, lfast = 0, lfastbad = 0;
      for (uint64_t b = lo + (uint64_t)t * stride; b <= hi; b += (uint64_t)nt * stride) {
        const float x = fbits((uint32_t)b);
        float y[2];
        const int nq = nt2ref::trig::ieee754_rem_pio2f(x, y) & 3;
        float a;
        const int qa = nt2h::reduce_fdlibm_f(x, &a);
        ++ln;
        if (x <= nt2h::kFdlibmFastHiF) {  // the fast path, where it decides, must agree too
          float af;
          int qf;
          ++lband;
          if (nt2h::reduce_fdlibm_f_fast(x, &af, &qf)) {
            ++lfast;
            if (qf != nq || ubits(af) != ubits(y[0])) {
              if (lfastbad < 3) std::printf("fast path differs at %a: oracle %a q%d, fast %a q%d\n", x, y[0], nq, af, qf);
              ++lfastbad;
            }
          }
        }
        if (qa != nq || ubits(a) != ubits(y[0])) {
          if (lbad < 3) std::printf("differ at %a: oracle %a q%d, kernel %a q%d\n", x, y[0], nq, a, qa);
          ++lbad;
        }
      }
      n += ln;
      bad += lbad;
      nband += lband;
      fast += lfast;
      fastbad += lfastbad;
    });
  for (auto& t : th) t.join();
  std::printf("fdlibm float reduction: %llu floats in (%a, %a] (stride %u): %llu differ\n",
              (unsigned long long)n.load(), nt2h::kFdlibmLoF, fbits(hi), stride, (unsigned long long)bad.load());
  std::printf("fast path on (%a, %a]: %llu floats, %llu decided (%.3f%%), %llu differ\n", nt2h::kFdlibmLoF,
              nt2h::kFdlibmFastHiF, (unsigned long long)nband.load(), (unsigned long long)fast.load(),
              100.0 * (double)fast.load() / (double)(nband.load() ? nband.load() : 1), (unsigned long long)fastbad.load());
  const int drc = double_checks(nt);
  return (bad.load() == 0 && fastbad.load() == 0 && drc == 0) ? 0 : 1;
}

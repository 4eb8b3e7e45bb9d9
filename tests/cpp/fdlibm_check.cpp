// Exhaustive check of the EXACT float sin/cos reduction above 2^18 pi.
// include/nt2hip/math.hpp's reduce_fdlibm_f (the code the GPU kernels run,
// compiled here for the host from the same header) is compared with the
// oracle's statement-by-statement restatement of fdlibm's
// __ieee754_rem_pio2f (oracle/rem_pio2.cpp, after NT2's
// modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:83-233)
// on every finite float above 2^18 pi: the quadrant (n & 3) and the bits of
// the head y[0] must agree.  An optional argument walks every k-th float
// instead.  Prints the counts; exit status 0 when nothing differs.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/nt2hip/math.hpp"
#include "../../oracle/nt2ref.hpp"

static float fbits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

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
  return (bad.load() == 0 && fastbad.load() == 0) ? 0 : 1;
}

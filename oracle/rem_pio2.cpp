// Large-argument pi/2 reduction for the oracle — TEST INFRASTRUCTURE ONLY.
//
// Restates the fdlibm reduction that NT2 vendors in
// modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:83-913
// (__ieee754_rem_pio2f / __kernel_rem_pio2f for float, __ieee754_rem_pio2 /
// __kernel_rem_pio2 for double; Sun Microsystems 1993, freely distributable).
// The published method (Payne & Hanek 1983, as fdlibm organises it): split x
// into K-bit integer chunks, multiply by the K-bit chunks of 2/pi that matter
// at x's exponent, read the quadrant from the integer part, and multiply the
// fraction by pi/2 held in K-bit pieces.  The float variant uses K=8 and float
// arithmetic, the double variant K=24 and double arithmetic; one template
// below expresses both so that each step runs in the reference's type and
// order.  The medium-size paths (|x| <= 2^7 pi/2 float, 2^19 pi/2 double) are
// the 3-stage Cody-Waite with cancellation checks of the same files.
#include "nt2ref.hpp"

#include <cmath>

namespace nt2ref {
namespace trig {

// 2/pi in 24-bit words (the binary expansion of 2/pi; 66 words).
const int32_t two_over_pi_24[66] = {
    0xA2F983, 0x6E4E44, 0x1529FC, 0x2757D1, 0xF534DD, 0xC0DB62, 0x95993C, 0x439041, 0xFE5163, 0xABDEBB,
    0xC561B7, 0x246E3A, 0x424DD2, 0xE00649, 0x2EEA09, 0xD1921C, 0xFE1DEB, 0x1CB129, 0xA73EE8, 0x8235F5,
    0x2EBB44, 0x84E99C, 0x7026B4, 0x5F7E41, 0x3991D6, 0x398353, 0x39F49C, 0x845F8B, 0xBDF928, 0x3B1FF8,
    0x97FFDE, 0x05980F, 0xEF2F11, 0x8B5A0A, 0x6D1F6D, 0x367ECF, 0x27CB09, 0xB74F46, 0x3F669E, 0x5FEA2D,
    0x7527BA, 0xC7EBE5, 0xF17B3D, 0x0739F7, 0x8A5292, 0xEA6BFB, 0x5FB11F, 0x8D5D08, 0x560330, 0x46FC7B,
    0x6BABF0, 0xCFBC20, 0x9AF436, 0x1DA9E3, 0x91615E, 0xE61B08, 0x659985, 0x5F14A0, 0x68408D, 0xFFD880,
    0x4D7327, 0x310606, 0x1556CA, 0x73A8C9, 0x60E27B, 0xC08C6B,
};

namespace {

// The same bit string in 8-bit chunks (198 bytes), built once.
struct Bytes2oPi {
  int32_t v[198];
  Bytes2oPi() {
    for (int i = 0; i < 66; ++i) {
      v[3 * i + 0] = (two_over_pi_24[i] >> 16) & 0xff;
      v[3 * i + 1] = (two_over_pi_24[i] >> 8) & 0xff;
      v[3 * i + 2] = two_over_pi_24[i] & 0xff;
    }
  }
};
const Bytes2oPi& bytes_2opi() { static const Bytes2oPi b; return b; }

// pi/2 in K-bit pieces, each exactly representable in T.
const uint32_t kPio2F8[11] = {0x3fc90000, 0x39f00000, 0x37da0000, 0x33a20000, 0x2e840000, 0x2b500000,
                              0x27c20000, 0x22d00000, 0x1fc40000, 0x1bc60000, 0x17440000};
const uint64_t kPio2D24[8] = {0x3FF921FB40000000ull, 0x3E74442D00000000ull, 0x3CF8469880000000ull,
                              0x3B78CC5160000000ull, 0x39F01B8380000000ull, 0x387A252040000000ull,
                              0x36E3822280000000ull, 0x3569F31D00000000ull};

template <class T> struct PH;
template <> struct PH<float> {
  static const int K = 8;
  static const int jk = 9;  // init_jk[prec=2] of the float kernel
  static float pio2(int k) { return f_bits(kPio2F8[k]); }
  static const int32_t* ipio2() { return bytes_2opi().v; }
};
template <> struct PH<double> {
  static const int K = 24;
  static const int jk = 4;  // init_jk[prec=2] of the double kernel
  static double pio2(int k) { return d_bits(kPio2D24[k]); }
  static const int32_t* ipio2() { return two_over_pi_24; }
};

// Payne-Hanek core with prec = 2 (two output words), as both callers use it.
template <class T> int32_t kernel_rem_pio2(const T* x, T* y, int e0, int nx) {
  const int K = PH<T>::K;
  const T twoK = (T)(1 << K);
  const T twonK = (T)1 / twoK;
  const int32_t* ipio2 = PH<T>::ipio2();
  const int jk = PH<T>::jk, jp = jk;
  int32_t iq[20];
  T f[20], fq[20], q[20];

  const int jx = nx - 1;
  int jv = (e0 - 3) / K;
  if (jv < 0) jv = 0;
  int q0 = e0 - K * (jv + 1);

  {  // f[0..jx+jk] = ipio2[jv-jx .. jv+jk] (zero before the table)
    int j = jv - jx, m = jx + jk;
    for (int i = 0; i <= m; ++i, ++j) f[i] = (j < 0) ? T(0) : (T)ipio2[j];
  }
  for (int i = 0; i <= jk; ++i) {
    T fw = 0;
    for (int j = 0; j <= jx; ++j) fw += x[j] * f[jx + i - j];
    q[i] = fw;
  }

  int jz = jk;
  int32_t n, ih;
  T z;
  for (;;) {
    // distill q[] into K-bit integers, most significant last
    int i = 0;
    z = q[jz];
    for (int j = jz; j > 0; ++i, --j) {
      T fw = (T)((int32_t)(twonK * z));
      iq[i] = (int32_t)(z - twoK * fw);
      z = q[j - 1] + fw;
    }
    z = std::ldexp(z, q0);
    z -= (T)8 * std::floor(z * (T)0.125);
    n = (int32_t)z;
    z -= (T)n;
    ih = 0;
    if (q0 > 0) {
      int32_t t = iq[jz - 1] >> (K - q0);
      n += t;
      iq[jz - 1] -= t << (K - q0);
      ih = iq[jz - 1] >> (K - 1 - q0);
    } else if (q0 == 0) {
      ih = iq[jz - 1] >> (K - 1);
    } else if (z >= (T)0.5) {
      ih = 2;
    }
    if (ih > 0) {  // fraction > 1/2: take 1 - fraction
      n += 1;
      int32_t carry = 0;
      for (int k = 0; k < jz; ++k) {
        int32_t j = iq[k];
        if (carry == 0) {
          if (j != 0) { carry = 1; iq[k] = (1 << K) - j; }
        } else {
          iq[k] = ((1 << K) - 1) - j;
        }
      }
      if (q0 > 0) {
        if (q0 == 1) iq[jz - 1] &= (1 << (K - 1)) - 1;
        else if (q0 == 2) iq[jz - 1] &= (1 << (K - 2)) - 1;
      }
      if (ih == 2) {
        z = (T)1 - z;
        if (carry != 0) z -= std::ldexp((T)1, q0);
      }
    }
    if (z == (T)0) {  // possible total cancellation: add terms and redo
      int32_t j = 0;
      for (int k = jz - 1; k >= jk; --k) j |= iq[k];
      if (j == 0) {
        int k = 1;
        while (iq[jk - k] == 0) ++k;
        for (int ii = jz + 1; ii <= jz + k; ++ii) {
          f[jx + ii] = (T)ipio2[jv + ii];
          T fw = 0;
          for (int jj = 0; jj <= jx; ++jj) fw += x[jj] * f[jx + ii - jj];
          q[ii] = fw;
        }
        jz += k;
        continue;
      }
    }
    break;
  }

  if (z == (T)0) {  // chop trailing zero chunks
    jz -= 1; q0 -= K;
    while (iq[jz] == 0) { jz--; q0 -= K; }
  } else {
    z = std::ldexp(z, -q0);
    if (z >= twoK) {
      T fw = (T)((int32_t)(twonK * z));
      iq[jz] = (int32_t)(z - twoK * fw);
      jz += 1; q0 += K;
      iq[jz] = (int32_t)fw;
    } else {
      iq[jz] = (int32_t)z;
    }
  }
  {
    T fw = std::ldexp((T)1, q0);
    for (int i = jz; i >= 0; --i) { q[i] = fw * (T)iq[i]; fw *= twonK; }
  }
  for (int i = jz; i >= 0; --i) {
    T fw = 0;
    for (int k = 0; k <= jp && k <= jz - i; ++k) fw += PH<T>::pio2(k) * q[i + k];
    fq[jz - i] = fw;
  }
  {  // prec 2: head and tail
    T fw = 0;
    for (int i = jz; i >= 0; --i) fw += fq[i];
    y[0] = (ih == 0) ? fw : -fw;
    fw = fq[0] - fw;
    for (int i = 1; i <= jz; ++i) fw += fq[i];
    y[1] = (ih == 0) ? fw : -fw;
  }
  return n & 7;
}

}  // namespace

int32_t kernel_rem_pio2f(const float* x, float* y, int e0, int nx) { return kernel_rem_pio2<float>(x, y, e0, nx); }
int32_t kernel_rem_pio2d(const double* x, double* y, int e0, int nx) { return kernel_rem_pio2<double>(x, y, e0, nx); }

// __ieee754_rem_pio2f (rem_pio2.hpp:83-231)
int32_t ieee754_rem_pio2f(float x, float* y) {
  static const uint32_t npio2_hw[32] = {
      0x3fc90f00, 0x40490f00, 0x4096cb00, 0x40c90f00, 0x40fb5300, 0x4116cb00, 0x412fed00, 0x41490f00,
      0x41623100, 0x417b5300, 0x418a3a00, 0x4196cb00, 0x41a35c00, 0x41afed00, 0x41bc7e00, 0x41c90f00,
      0x41d5a000, 0x41e23100, 0x41eec200, 0x41fb5300, 0x4203f200, 0x420a3a00, 0x42108300, 0x4216cb00,
      0x421d1400, 0x42235c00, 0x4229a500, 0x422fed00, 0x42363600, 0x423c7e00, 0x4242c700, 0x42490f00};
  const float invpio2 = f_bits(0x3f22f984), pio2_1 = f_bits(0x3fc90f80), pio2_1t = f_bits(0x37354443),
              pio2_2 = f_bits(0x37354400), pio2_2t = f_bits(0x2e85a308), pio2_3 = f_bits(0x2e85a300),
              pio2_3t = f_bits(0x248d3132);
  int32_t hx = (int32_t)bits_of(x);
  int32_t ix = hx & 0x7fffffff;
  if (ix <= 0x3f490fd8) { y[0] = x; y[1] = 0; return 0; }
  if (ix < 0x4016cbe4) {
    float z;
    if (hx > 0) {
      z = x - pio2_1;
      if ((ix & 0xfffffff0) != 0x3fc90fd0) { y[0] = z - pio2_1t; y[1] = (z - y[0]) - pio2_1t; }
      else { z -= pio2_2; y[0] = z - pio2_2t; y[1] = (z - y[0]) - pio2_2t; }
      return 1;
    }
    z = x + pio2_1;
    if ((ix & 0xfffffff0) != 0x3fc90fd0) { y[0] = z + pio2_1t; y[1] = (z - y[0]) + pio2_1t; }
    else { z += pio2_2; y[0] = z + pio2_2t; y[1] = (z - y[0]) + pio2_2t; }
    return -1;
  }
  if (ix <= 0x43490f80) {
    float t = std::fabs(x);
    int32_t n = (int32_t)(t * invpio2 + 0.5f);
    float fn = (float)n;
    float r = t - fn * pio2_1;
    float w = fn * pio2_1t;
    if (n < 32 && (uint32_t)(ix & 0xffffff00) != npio2_hw[n - 1]) {
      y[0] = r - w;
    } else {
      int32_t j = ix >> 23;
      y[0] = r - w;
      uint32_t high = bits_of(y[0]);
      int32_t i = j - ((high >> 23) & 0xff);
      if (i > 8) {
        t = r; w = fn * pio2_2; r = t - w; w = fn * pio2_2t - ((t - r) - w); y[0] = r - w;
        high = bits_of(y[0]);
        i = j - ((high >> 23) & 0xff);
        if (i > 25) { t = r; w = fn * pio2_3; r = t - w; w = fn * pio2_3t - ((t - r) - w); y[0] = r - w; }
      }
    }
    y[1] = (r - y[0]) - w;
    if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
    return n;
  }
  if (ix >= 0x7f800000) { y[0] = y[1] = x - x; return 0; }
  int32_t e0 = (ix >> 23) - 134;
  float z = f_bits((uint32_t)(ix - (e0 << 23)));
  float tx[3];
  for (int i = 0; i < 2; ++i) { tx[i] = (float)((int32_t)z); z = (z - tx[i]) * 256.0f; }
  tx[2] = z;
  int nx = 3;
  while (tx[nx - 1] == 0.0f) nx--;
  int32_t n = kernel_rem_pio2<float>(tx, y, e0, nx);
  if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
  return n;
}

// __ieee754_rem_pio2 (rem_pio2.hpp:484-622)
int32_t ieee754_rem_pio2d(double x, double* y) {
  static const int32_t npio2_hw[32] = {
      0x3FF921FB, 0x400921FB, 0x4012D97C, 0x401921FB, 0x401F6A7A, 0x4022D97C, 0x4025FDBB, 0x402921FB,
      0x402C463A, 0x402F6A7A, 0x4031475C, 0x4032D97C, 0x40346B9C, 0x4035FDBB, 0x40378FDB, 0x403921FB,
      0x403AB41B, 0x403C463A, 0x403DD85A, 0x403F6A7A, 0x40407E4C, 0x4041475C, 0x4042106C, 0x4042D97C,
      0x4043A28C, 0x40446B9C, 0x404534AC, 0x4045FDBB, 0x4046C6CB, 0x40478FDB, 0x404858EB, 0x404921FB};
  const double invpio2 = d_bits(0x3FE45F306DC9C883ull), pio2_1 = d_bits(0x3FF921FB54400000ull),
               pio2_1t = d_bits(0x3DD0B4611A626331ull), pio2_2 = d_bits(0x3DD0B4611A600000ull),
               pio2_2t = d_bits(0x3BA3198A2E037073ull), pio2_3 = d_bits(0x3BA3198A2E000000ull),
               pio2_3t = d_bits(0x397B839A252049C1ull);
  uint64_t bx = bits_of(x);
  int32_t hx = (int32_t)(bx >> 32);
  int32_t ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) { y[0] = x; y[1] = 0; return 0; }
  if (ix < 0x4002d97c) {
    double z;
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) { y[0] = z - pio2_1t; y[1] = (z - y[0]) - pio2_1t; }
      else { z -= pio2_2; y[0] = z - pio2_2t; y[1] = (z - y[0]) - pio2_2t; }
      return 1;
    }
    z = x + pio2_1;
    if (ix != 0x3ff921fb) { y[0] = z + pio2_1t; y[1] = (z - y[0]) + pio2_1t; }
    else { z += pio2_2; y[0] = z + pio2_2t; y[1] = (z - y[0]) + pio2_2t; }
    return -1;
  }
  if (ix <= 0x413921fb) {
    double t = std::fabs(x);
    int32_t n = (int32_t)(t * invpio2 + 0.5);
    double fn = (double)n;
    double r = t - fn * pio2_1;
    double w = fn * pio2_1t;
    if ((n < 32) && (n > 0) && (ix != npio2_hw[n - 1])) {
      y[0] = r - w;
    } else {
      int32_t j = ix >> 20;
      y[0] = r - w;
      uint32_t high = (uint32_t)(bits_of(y[0]) >> 32);
      int32_t i = j - ((high >> 20) & 0x7ff);
      if (i > 16) {
        t = r; w = fn * pio2_2; r = t - w; w = fn * pio2_2t - ((t - r) - w); y[0] = r - w;
        high = (uint32_t)(bits_of(y[0]) >> 32);
        i = j - ((high >> 20) & 0x7ff);
        if (i > 49) { t = r; w = fn * pio2_3; r = t - w; w = fn * pio2_3t - ((t - r) - w); y[0] = r - w; }
      }
    }
    y[1] = (r - y[0]) - w;
    if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
    return n;
  }
  if (ix >= 0x7ff00000) { y[0] = y[1] = x - x; return 0; }
  int32_t e0 = (ix >> 20) - 1046;
  uint64_t zb = ((uint64_t)(uint32_t)(ix - (e0 << 20)) << 32) | (bx & 0xffffffffull);
  double z = d_bits(zb);
  double tx[3];
  for (int i = 0; i < 2; ++i) { tx[i] = (double)((int32_t)z); z = (z - tx[i]) * 16777216.0; }
  tx[2] = z;
  int nx = 3;
  while (tx[nx - 1] == 0.0) nx--;
  int32_t n = kernel_rem_pio2<double>(tx, y, e0, nx);
  if (hx < 0) { y[0] = -y[0]; y[1] = -y[1]; return -n; }
  return n;
}

}  // namespace trig
}  // namespace nt2ref

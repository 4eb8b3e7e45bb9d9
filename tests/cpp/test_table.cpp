// C++ front-end check: NT2-style table expressions on the GPU through
// include/nt2hip/table.hpp, compared with the oracle (libnt2ref.so) on the
// CPU.  Prints one "PASS name" / "FAIL name ..." line per check and exits
// non-zero on any failure.  Built by tests/cpp/Makefile (with the nd_*.cpp
// units), run by tests/test_gpu_frontend.py.
#include "common.hpp"

namespace nt2 = nt2hip;  // NT2's demo bodies below compile as written
static int failures = 0;
void report(const char* name, bool ok, const char* detail) {
  std::printf("%s %s %s\n", ok ? "PASS" : "FAIL", name, detail);
  if (!ok) ++failures;
}
// the N-D checks (nd_*.cpp)
void nd_checks_3d();
void nd_checks_4d();
void nd_checks_paths();
void nd_checks_edges();
void nd_checks_f64();
void nd_checks_into_forms();
void fused_fold_speed_check();

// ---------------------------------------------------------------------------
// NT2's own demo expressions, bodies copied unchanged (nt2:: is nt2hip::).
// demo/blackscholes/nt2/blackandscholes_nt2.cpp:53-60
template <typename T> struct blackandscholes_demo {
  nt2::table<T> Sa, Xa, Ta, ra, va, R;
  void operator()() {
    nt2::table<T> da = nt2::sqrt(Ta);
    nt2::table<T> d1 = nt2::log(Sa/Xa) + (nt2::fma(nt2::sqr(va),T(0.5),ra)*Ta)/(va*da);
    nt2::table<T> d2 = nt2::fnms(va,da,d1);

    R = nt2::fnms(Xa*nt2::exp(-ra*Ta),nt2::fastnormcdf(d2),Sa*nt2::fastnormcdf(d1));
  }
};
// demo/blackscholes/nt2_merged/blackandscholes_nt2_merged.cpp:57-64
template <typename T> struct blackandscholes_merged_demo {
  nt2::table<T> Sa, Xa, Ta, ra, va, R, da, d1, d2;
  void operator()() {
    nt2::tie(da,d1,d2,R) = nt2::tie ( nt2::sqrt(Ta)
                                    , nt2::log(Sa/Xa) + (nt2::fma(nt2::sqr(va),T(0.5),ra)*Ta)/(va*da)
                                    , nt2::fnms(va,da,d1)
                                    , nt2::fnms(Xa*nt2::exp(-ra*Ta),nt2::fastnormcdf(d2),Sa*nt2::fastnormcdf(d1))
                                    );
  }
};
// demo/saxpy/nt2/saxpy_nt2.cpp:39-42
template <typename T> struct axpy_demo {
  T alpha;
  nt2::table<T> X, Y;
  void operator()() { Y = alpha * X + Y; }
};
// demo/sdot/nt2/sdot_nt2.cpp:35-38
template <typename T> struct dot_demo {
  nt2::table<T> X, Y;
  T res_;
  void operator()() { res_ = nt2::globalsum(X*Y); }
};

// The oracle's composition of the same functors (NT2's scalar path: every
// operation rounded, fma/fnms as a product then a sum, each functor the
// oracle's restatement): the demo's four loop nests per element, with the
// merged demo's tie reading da, d1 and d2 as they were before the call.
struct bs_terms {
  std::vector<float> da, d1, d2, R;
};
static std::vector<float> map_ref(int op, const std::vector<float>& x) {
  std::vector<float> y(x.size());
  nt2ref_map_f32(op, x.data(), y.data(), (int64_t)x.size());
  return y;
}
static bs_terms bs_demo_ref(const std::vector<float>& S, const std::vector<float>& X, const std::vector<float>& T,
                            const std::vector<float>& r, const std::vector<float>& v, const std::vector<float>* da0,
                            const std::vector<float>* d10, const std::vector<float>* d20) {
  const size_t n = S.size();
  bs_terms o;
  o.da = map_ref(2, T);  // sqrt
  const std::vector<float>& da = da0 ? *da0 : o.da;
  std::vector<float> q(n), mrt(n);
  for (size_t i = 0; i < n; ++i) { q[i] = S[i] / X[i]; mrt[i] = -r[i] * T[i]; }
  std::vector<float> lq = map_ref(1, q), e = map_ref(0, mrt);
  o.d1.resize(n);
  o.d2.resize(n);
  for (size_t i = 0; i < n; ++i) {
    float sq = v[i] * v[i];
    float f = sq * 0.5f;
    f = f + r[i];  // fma(sqr(v), 0.5, r): two roundings
    float num = f * T[i];
    float den = v[i] * da[i];
    float t = num / den;
    o.d1[i] = lq[i] + t;
  }
  const std::vector<float>& d1 = d10 ? *d10 : o.d1;
  for (size_t i = 0; i < n; ++i) {
    float p = v[i] * da[i];
    o.d2[i] = d1[i] - p;  // fnms(v, da, d1) = d1 - v*da
  }
  const std::vector<float>& d2 = d20 ? *d20 : o.d2;
  std::vector<float> n1 = map_ref(10, d1), n2 = map_ref(10, d2);  // fastnormcdf
  o.R.resize(n);
  for (size_t i = 0; i < n; ++i) {
    float a = X[i] * e[i];
    float b = S[i] * n1[i];
    float p = a * n2[i];
    o.R[i] = b - p;  // fnms(a, n2, b)
  }
  return o;
}

static void demo_checks() {
  {  // the demo's own inputs (Sa(i) = ... = T(i), i = 1..4096) and C2-like ones
    for (int which = 0; which < 2; ++which) {
      const size_t n = which == 0 ? 4096 : 100003;
      std::vector<float> S(n), X(n), T(n), r(n), v(n);
      if (which == 0) {
        for (size_t i = 1; i <= n; ++i) S[i - 1] = X[i - 1] = T[i - 1] = r[i - 1] = v[i - 1] = float(i);
      } else {
        S = uniform(n, 5.f, 30.f, 21); X = uniform(n, 1.f, 100.f, 22); T = uniform(n, 0.25f, 10.f, 23);
        r.assign(n, 0.02f); v.assign(n, 0.30f);
      }
      blackandscholes_demo<float> d{table<float>(S), table<float>(X), table<float>(T), table<float>(r),
                                    table<float>(v), table<float>()};
      d();
      bs_terms want = bs_demo_ref(S, X, T, r, v, nullptr, nullptr, nullptr);
      report(which == 0 ? "demo_blackscholes_nt2_demo_inputs_bit_exact" : "demo_blackscholes_nt2_bit_exact",
             same_bits(d.R.to_host(), want.R));
    }
  }
  {  // merged demo: one loop nest, four outputs; the tie reads da, d1, d2 as
     // they were before the call (NT2's semantics), so after four calls from
     // arbitrary values R equals the four-loop demo's R bit for bit
    const size_t n = 100003;
    std::vector<float> S = uniform(n, 5.f, 30.f, 31), X = uniform(n, 1.f, 100.f, 32), T = uniform(n, 0.25f, 10.f, 33);
    std::vector<float> r(n, 0.02f), v(n, 0.30f);
    std::vector<float> da0 = uniform(n, 0.5f, 3.f, 34), d10 = uniform(n, -2.f, 2.f, 35), d20 = uniform(n, -2.f, 2.f, 36);
    blackandscholes_merged_demo<float> m;
    m.Sa = table<float>(S); m.Xa = table<float>(X); m.Ta = table<float>(T); m.ra = table<float>(r);
    m.va = table<float>(v); m.R = table<float>(n); m.da = table<float>(da0); m.d1 = table<float>(d10);
    m.d2 = table<float>(d20);
    bool ok = true;
    std::vector<float> a = da0, b = d10, c = d20;
    for (int call = 0; call < 4; ++call) {
      m();
      bs_terms want = bs_demo_ref(S, X, T, r, v, &a, &b, &c);
      ok = ok && same_bits(m.da.to_host(), want.da) && same_bits(m.d1.to_host(), want.d1) &&
           same_bits(m.d2.to_host(), want.d2) && same_bits(m.R.to_host(), want.R);
      a = want.da; b = want.d1; c = want.d2;
    }
    report("demo_blackscholes_merged_tie4_stale_reads_bit_exact", ok);
    bs_terms plain = bs_demo_ref(S, X, T, r, v, nullptr, nullptr, nullptr);
    report("demo_blackscholes_merged_converges_to_demo", same_bits(m.R.to_host(), plain.R));
  }
  {  // saxpy demo: X(i) = Y(i) = T(i-1); known answer (alpha+1)(i-1), bit-exact vs the oracle
    const size_t n = size_t(1) << 20;
    std::vector<float> x(n);
    for (size_t i = 1; i <= n; ++i) x[i - 1] = float(i - 1);
    axpy_demo<float> d{2.5f, table<float>(x), table<float>(x)};
    d();
    std::vector<float> want = x;
    nt2ref_saxpy_f32(2.5f, x.data(), want.data(), (int64_t)n);
    auto got = d.Y.to_host();
    bool known = true;
    for (size_t i = 0; i < n && known; ++i) known = got[i] == 3.5f * float(i);
    report("demo_saxpy_bit_exact_and_known_answer", same_bits(got, want) && known);
  }
  {  // sdot demo: X(i) = Y(i) = T(i-1); below 4096 the products and the closed
     // form sum (n-1) n (2n-1) / 6 are exact, so the result is its rounding
    for (size_t n : {size_t(4096), size_t(1) << 20}) {
      std::vector<float> x(n);
      for (size_t i = 1; i <= n; ++i) x[i - 1] = float(i - 1);
      dot_demo<float> d{table<float>(x), table<float>(x), 0.f};
      d();
      double ex[2];
      nt2ref_exact_dot_f32(x.data(), x.data(), (int64_t)n, ex);
      float want = (float)(ex[0] + ex[1]);
      bool ok = std::fabs(d.res_ - want) <= std::nextafter(want, INFINITY) - want;
      if (n == 4096) {
        double m = (double)n;
        ok = ok && d.res_ == (float)((m - 1) * m * (2 * m - 1) / 6);
      }
      report(n == 4096 ? "demo_sdot_closed_form" : "demo_sdot_within_1ulp", ok);
    }
  }
  {  // fma / fms / fnms / fnma nodes: two roundings under exact, FMA under fused
    const int64_t n = 65536 + 3;
    auto ha = uniform(n, -3.f, 3.f, 41), hb = uniform(n, -3.f, 3.f, 42), hc = uniform(n, -3.f, 3.f, 43);
    table<float> A(ha), B(hb), C(hc), Z;
    Z = nt2::fma(A, B, C) - nt2::fms(A, 2.0f, C) + nt2::fnms(A, B, C) * nt2::fnma(1.5f, B, C);
    std::vector<float> want(n), fz(n);
    for (int64_t i = 0; i < n; ++i) {
      float p = ha[i] * hb[i], f1 = p + hc[i];
      float q = ha[i] * 2.0f, f2 = q - hc[i];
      float f3 = hc[i] - p;
      float u = 1.5f * hb[i], f4 = -hc[i] - u;
      float s1 = f1 - f2, s2 = f3 * f4;
      want[i] = s1 + s2;
    }
    report("fma_fms_fnms_fnma_exact_bit_exact", same_bits(Z.to_host(), want));
    nt2::assign<nt2hip::fused>(Z, nt2::fma(A, B, C));
    auto g = Z.to_host();
    bool ok = true;
    for (int64_t i = 0; i < n && ok; ++i) ok = g[i] == std::fmaf(ha[i], hb[i], hc[i]);
    report("fma_fused_single_rounding", ok);
  }
  {  // tie of tables (a swap), a tie whose output is read through sx(), and the fused tie
    const int64_t rows = 33, cols = 9;
    auto h1 = uniform(rows * cols, -1.f, 1.f, 51), h2 = uniform(rows * cols, -1.f, 1.f, 52), h3 = uniform(cols, 1.f, 2.f, 53);
    table<float> P(h1, rows, cols), Q(h2, rows, cols), W(h3, 1, cols), M(h1, rows, cols);
    nt2::tie(P, Q) = nt2::tie(Q, P);
    bool ok = same_bits(P.to_host(), h2) && same_bits(Q.to_host(), h1);
    // W is both an output and the sx() row of a child: staged, so every
    // element reads the row as it was before the call
    nt2::tie(M, W) = nt2::tie(M * nt2::sx(W), M + 1.0f);
    auto hm = M.to_host(), hw = W.to_host();
    ok = ok && W.rows() == rows && W.cols() == cols;
    for (int64_t i = 0; i < rows * cols && ok; ++i) ok = hm[i] == h1[i] * h3[i / rows] && hw[i] == h1[i] + 1.0f;
    report("tie_tables_swap_and_staged_sx_output", ok);
    table<float> A(h1, rows, cols), B(h2, rows, cols), C, D, E;
    nt2::assign<nt2hip::fused>(nt2::tie(C, D, E), nt2::tie(nt2::fma(A, B, 1.0f), A - B, nt2::sqrt(nt2::abs(A))));
    auto hc = C.to_host(), hd = D.to_host(), he = E.to_host();
    bool ok2 = true;
    for (int64_t i = 0; i < rows * cols && ok2; ++i)
      ok2 = hc[i] == std::fmaf(h1[i], h2[i], 1.0f) && hd[i] == h1[i] - h2[i] && he[i] == std::sqrt(std::fabs(h1[i]));
    report("tie3_fused", ok2);
  }
}

int main() {
  demo_checks();
  nd_checks_3d();
  nd_checks_4d();
  nd_checks_paths();
  nd_checks_edges();
  nd_checks_f64();
  nd_checks_into_forms();
  fused_fold_speed_check();
  const int64_t n = (1 << 20) + 7;
  auto hx = uniform(n, -1.f, 1.f, 1), hy = uniform(n, -1.f, 1.f, 2);
  table<float> X(hx), Y(hy);

  {  // Y = a*X + Y lowers to nt2h_saxpy_f32; bit-exact vs NT2
    Y = 2.5f * X + Y;
    std::vector<float> want = hy;
    nt2ref_saxpy_f32(2.5f, hx.data(), want.data(), n);
    report("saxpy_lowered_bit_exact", same_bits(Y.to_host(), want));
    Y = table<float>(hy);
  }
  {  // a single functor node lowers to nt2h_map_f32 (EXACT): bit-exact
    table<float> Z;
    Z = exp(X);
    std::vector<float> want(n);
    nt2ref_map_f32(0, hx.data(), want.data(), n);
    report("exp_lowered_bit_exact", same_bits(Z.to_host(), want));
  }
  {  // an arbitrary tree: one generated kernel, NT2 two-rounding semantics
    table<float> Z;
    Z = exp(-sqr(X) / 2.0f) * cos(3.0f * X + Y) + tanh(Y);
    std::vector<float> e(n), c(n), t(n), arg(n), q(n), want(n);
    for (int64_t i = 0; i < n; ++i) { q[i] = -(hx[i] * hx[i]) / 2.0f; float a = 3.0f * hx[i]; arg[i] = a + hy[i]; }
    nt2ref_map_f32(0, q.data(), e.data(), n);
    nt2ref_map_f32(8, arg.data(), c.data(), n);
    nt2ref_map_f32(6, hy.data(), t.data(), n);
    for (int64_t i = 0; i < n; ++i) { float p = e[i] * c[i]; want[i] = p + t[i]; }
    report("fused_tree_bit_exact", same_bits(Z.to_host(), want));
  }
  {  // globalsum / dot lower to the compensated folds
    double ex[2];
    nt2ref_exact_dot_f32(hx.data(), hy.data(), n, ex);
    float d = nt2hip::dot(X, Y);
    float want = (float)(ex[0] + ex[1]);
    float sp = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
    char buf[128];
    std::snprintf(buf, sizeof buf, "got %.9g want %.9g", d, want);
    report("dot_within_1ulp", std::fabs(d - want) <= sp, buf);
    nt2ref_exact_sum_f32(hx.data(), n, ex);
    float s = nt2hip::globalsum(X);
    want = (float)(ex[0] + ex[1]);
    sp = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
    report("globalsum_within_1ulp", std::fabs(s - want) <= sp);
    // a generic fold of a tree
    float s2 = nt2hip::globalsum(X * X + 1.0f);
    double ref = 0;
    for (int64_t i = 0; i < n; ++i) { float p = hx[i] * hx[i]; ref += (double)(p + 1.0f); }
    std::snprintf(buf, sizeof buf, "got %.9g want %.12g", s2, ref);
    report("globalsum_tree", std::fabs(s2 - ref) <= 1e-6 * std::fabs(ref), buf);
    double s3 = nt2hip::globalsum(nt2hip::table<double>(std::vector<double>(1000, 0.5)) * 2.0);
    std::snprintf(buf, sizeof buf, "got %.17g want 1000", s3);
    report("globalsum_tree_f64", s3 == 1000.0, buf);
  }
  {  // tie: both outputs from one loop nest, inputs read before outputs written
    table<float> A(hx), B(hy);
    nt2hip::tie(A, B) = nt2hip::tie(A + B, A - B);
    auto ha = A.to_host(), hb = B.to_host();
    bool ok = true;
    for (int64_t i = 0; i < n && ok; ++i) ok = (ha[i] == hx[i] + hy[i]) && (hb[i] == hx[i] - hy[i]);
    report("tie_reads_before_writes", ok);
  }
  {  // 2-D: sum along dimension 1 (inner_fold) and 2 (outer_fold), sx broadcast
    const int64_t rows = 1000, cols = 37;
    auto hm = uniform(rows * cols, -2.f, 2.f, 3), hs = uniform(cols, 0.5f, 2.f, 4), hk = uniform(cols, 0.1f, 10.f, 5);
    table<float> M(hm, rows, cols), Sx(hs, 1, cols), Kx(hk, 1, cols);
    table<float> z = nt2hip::sum(exp(-sqr(M) / (2.0f * sqr(nt2hip::sx(Sx)))) * cos(nt2hip::sx(Kx) * M), 1);
    table<float> zr = nt2hip::sum(M * M, 2);
    auto hz = z.to_host(), hzr = zr.to_host();
    bool ok1 = z.rows() == 1 && z.cols() == cols, ok2 = zr.rows() == rows && zr.cols() == 1;
    for (int64_t j = 0; j < cols && ok1; ++j) {
      double ref = 0, mag = 0;
      for (int64_t i = 0; i < rows; ++i) {
        double x = hm[j * rows + i], e = std::exp(-x * x / (2.0 * hs[j] * hs[j]));
        ref += e * std::cos((double)hk[j] * x);
        mag += e;
      }
      ok1 = std::fabs(hz[j] - ref) <= 1e-5 * mag;
    }
    for (int64_t i = 0; i < rows && ok2; ++i) {
      double ref = 0;
      for (int64_t j = 0; j < cols; ++j) ref += (double)hm[j * rows + i] * hm[j * rows + i];
      ok2 = std::fabs(hzr[i] - ref) <= 1e-6 * ref;
    }
    report("colfold_sum_dim1_sx", ok1);
    report("rowfold_sum_dim2", ok2);
  }
  {  // elementwise trees through the vector body: sx rows that straddle a quad
     // (37 rows), a scalar, a quad-free tail, and a double table — bit-exact vs
     // the same two-rounding composition on the CPU
    const int64_t rows = 37, cols = 7;
    auto hm = uniform(rows * cols, -3.f, 3.f, 13), hs = uniform(cols, 0.5f, 2.f, 14);
    table<float> M(hm, rows, cols), Sx(hs, 1, cols);
    table<float> Z;
    Z = M * nt2hip::sx(Sx) + exp(M) - 0.5f;
    std::vector<float> e(rows * cols), want(rows * cols);
    nt2ref_map_f32(0, hm.data(), e.data(), rows * cols);
    for (int64_t i = 0; i < rows * cols; ++i) {
      float a = hm[i] * hs[i / rows];
      float b = a + e[i];
      want[i] = b - 0.5f;
    }
    report("vector_tree_sx_straddling_quads", same_bits(Z.to_host(), want));
    const int64_t nd = 4099;
    std::vector<double> hd(nd);
    for (int64_t i = 0; i < nd; ++i) hd[i] = -4.0 + 8.0 * (double)i / (double)nd;
    table<double> D(hd, nd, 1);
    table<double> E;
    E = exp(D) * D + 1.0;
    std::vector<double> ed(nd), wd(nd);
    nt2ref_map_f64(0, hd.data(), ed.data(), nd);
    for (int64_t i = 0; i < nd; ++i) {
      double a = ed[i] * hd[i];
      wd[i] = a + 1.0;
    }
    auto gd = E.to_host();
    report("vector_tree_f64", std::memcmp(gd.data(), wd.data(), sizeof(double) * nd) == 0);
  }
  {  // Black-Scholes through tie (5 -> 2 streams): bit-exact vs NT2's scalar composition
    const int64_t m = 100003;
    auto hS = uniform(m, 5.f, 30.f, 6), hK = uniform(m, 1.f, 100.f, 7), hT = uniform(m, 0.25f, 10.f, 8);
    std::vector<float> hr(m, 0.02f), hv(m, 0.30f), wc(m), wp(m);
    table<float> S(hS), K(hK), T(hT), r(hr), v(hv), call, put;
    nt2hip::tie(call, put) = nt2hip::blackscholes(S, K, T, r, v);
    nt2ref_bs_fwd_f32(hS.data(), hK.data(), hT.data(), hr.data(), hv.data(), wc.data(), wp.data(), m);
    report("blackscholes_tie_bit_exact", same_bits(call.to_host(), wc) && same_bits(put.to_host(), wp));
    auto hw = uniform(m, -1.f, 1.f, 9);
    table<float> w(hw), price, dS, dK, dT, dr, dsig;
    auto tot = nt2hip::blackscholes_adjoint(S, K, T, r, v, w, price, dS, dK, dT, dr, dsig);
    auto hdr = dr.to_host();
    double sdr = 0, adr = 0;
    for (float x : hdr) { sdr += x; adr += std::fabs(x); }
    report("blackscholes_adjoint_totals", std::fabs(tot.dP_dr - sdr) <= 1e-9 * adr);
  }
  {  // errors surface as exceptions
    bool threw = false;
    try {
      table<float> A(10), B(11), C;
      C = A + B;
    } catch (const nt2hip::error& e) {
      threw = e.code == NT2H_ERR_INVALID_ARG;
    }
    report("size_mismatch_throws", threw);
  }
  {  // maximum / minimum along each dimension: bit-exact vs NT2's fold, NaNs included
    const int64_t R = 300, C = 7;
    auto h = uniform(R * C, -4.f, 4.f, 9);
    h[5] = NAN; h[R + 17] = NAN; h[2 * R + R - 1] = NAN; h[3 * R + 1] = -0.0f;
    table<float> A(h, R, C);
    bool ok = true;
    for (int dim = 1; dim <= 3; ++dim) {
      int64_t lo = dim == 1 ? 1 : (dim == 2 ? R : R * C), mid = dim == 1 ? R : (dim == 2 ? C : 1),
              hi = dim == 1 ? C : 1;
      for (int op : {NT2H_FOLD_MAXIMUM, NT2H_FOLD_MINIMUM}) {
        std::vector<float> want(lo * hi);
        nt2ref_fold_lmh_f32(op, h.data(), lo, mid, hi, want.data());
        table<float> got = op == NT2H_FOLD_MAXIMUM ? nt2hip::maximum(A, dim) : nt2hip::minimum(A, dim);
        ok = ok && same_bits(got.to_host(), want);
      }
    }
    std::vector<float> want(1);
    nt2ref_fold_lmh_f32(NT2H_FOLD_MAXIMUM, h.data(), 1, R * C, 1, want.data());
    ok = ok && bits(nt2hip::globalmax(A)) == bits(want[0]);
    report("maximum_minimum_dims_bit_exact", ok);
  }
  {  // globalmax/globalprod/maximum tables of NT2's unit tests
    std::vector<float> a{1, 2, 3, 4, 5, 6, 7, 8, 9};
    table<float> A(a, 3, 3);
    bool ok = nt2hip::globalmax(A) == 9.f;
    std::vector<float> b{1, 2, 3, 4};
    ok = ok && nt2hip::globalprod(table<float>(b, 2, 2)) == 24.f;
    std::vector<float> y(15);
    for (int j = 1; j <= 3; ++j)
      for (int i = 1; i <= 5; ++i) y[(i - 1) + 5 * (j - 1)] = (float)(i + 10 * j);
    table<float> Y(y, 5, 3);
    auto m1 = nt2hip::maximum(Y).to_host();        // first non-singleton: dim 1
    auto m2 = nt2hip::maximum(Y, 2).to_host();
    ok = ok && m1 == std::vector<float>{15, 25, 35} && m2 == std::vector<float>{31, 32, 33, 34, 35};
    auto mn = nt2hip::mean(Y, 1).to_host();
    ok = ok && std::fabs(mn[0] - 13.f) < 1e-5f && std::fabs(mn[2] - 33.f) < 1e-5f;
    report("fold_unit_tables", ok);
  }
  {  // cumsum along 1 and 2 of reshape(1:30, 3, 10): NT2's cumsum test values; cummax bit-exact
    std::vector<float> a(30);
    for (int i = 0; i < 30; ++i) a[i] = (float)(i + 1);
    table<float> A(a, 3, 10);
    std::vector<float> w1(30), w2(30);
    nt2ref_scan_lmh_f32(NT2H_SCAN_CUMSUM, a.data(), 1, 3, 10, w1.data());
    nt2ref_scan_lmh_f32(NT2H_SCAN_CUMSUM, a.data(), 3, 10, 1, w2.data());
    bool ok = nt2hip::cumsum(A).to_host() == w1 && nt2hip::cumsum(A, 2).to_host() == w2 && w1[8] == 24.f;
    auto h = uniform(5000, -1.f, 1.f, 11);
    h[100] = NAN;
    table<float> H(h, 5000, 1);
    std::vector<float> wm(5000);
    nt2ref_scan_lmh_f32(NT2H_SCAN_CUMMAX, h.data(), 1, 5000, 1, wm.data());
    ok = ok && same_bits(nt2hip::cummax(H).to_host(), wm);
    // an expression operand is evaluated first
    table<float> E = nt2hip::cummin(-H);
    std::vector<float> nh(5000), wn(5000);
    for (int i = 0; i < 5000; ++i) nh[i] = -h[i];
    nt2ref_scan_lmh_f32(NT2H_SCAN_CUMMIN, nh.data(), 1, 5000, 1, wn.data());
    ok = ok && same_bits(E.to_host(), wn);
    report("scans_unit_and_bit_exact", ok);
  }
  {  // double tables: maximum along 2 and cumsum along 1 of i + 10 j
    std::vector<double> y(15);
    for (int j = 1; j <= 3; ++j)
      for (int i = 1; i <= 5; ++i) y[(i - 1) + 5 * (j - 1)] = (double)(i + 10 * j);
    table<double> Y(y, 5, 3);
    bool ok = nt2hip::maximum(Y, 2).to_host() == std::vector<double>{31, 32, 33, 34, 35};
    auto c = nt2hip::cumsum(Y).to_host();
    ok = ok && c[4] == 11 + 12 + 13 + 14 + 15 && nt2hip::globalmin(Y) == 11.0;
    report("double_folds_scans", ok);
  }
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}

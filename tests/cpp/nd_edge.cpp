// Front-end edge cases: empty tables and empty dimensions, one-element
// tables, and non-finite data in the fused reductions and scans of trees
// (sums of data holding +-inf give +-inf, as NT2's recurrence does).
#include <cfloat>

#include "common.hpp"

void nd_checks_edges() {
  using nt2hip::extent;
  {  // empty tables: a tree, its reductions and scans
    bool ok = true;
    try {
      table<float> A(0), B(0), C;
      C = nt2hip::exp(A) * 2.0f + B;
      ok = ok && C.numel() == 0;
      ok = ok && nt2hip::globalsum(A) == 0.0f && nt2hip::globalsum(A * 2.0f) == 0.0f;
      ok = ok && nt2hip::globalmax(A) == -FLT_MAX && nt2hip::globalprod(A * 2.0f) == 1.0f;
      ok = ok && nt2hip::cumsum(A).numel() == 0 && nt2hip::cumsum(A + 1.0f).numel() == 0;
      ok = ok && std::isnan(nt2hip::globalmean(A));
    } catch (const nt2hip::error& e) {
      std::printf("  %s\n", e.what());
      ok = false;
    }
    report("edge_empty_tables", ok);
  }
  {  // an empty dimension: folds along it give the neutral element, folds
     // along the others an empty result
    bool ok = true;
    try {
      table<float> E(0, 5);
      auto s = nt2hip::sum(E, 1), m = nt2hip::maximum(E + 1.0f, 1), a = nt2hip::mean(E, 1), t = nt2hip::sum(E, 2);
      ok = s.numel() == 5 && m.numel() == 5 && a.numel() == 5 && t.numel() == 0;
      for (float v : s.to_host()) ok = ok && v == 0.0f;
      for (float v : m.to_host()) ok = ok && v == -FLT_MAX;
      for (float v : a.to_host()) ok = ok && std::isnan(v);
      ok = ok && nt2hip::cumsum(E * 3.0f, 1).numel() == 0;
    } catch (const nt2hip::error& e) {
      std::printf("  %s\n", e.what());
      ok = false;
    }
    report("edge_empty_dimension", ok);
  }
  {  // one element: a tree (NT2's exp), its sum and its scan
    std::vector<float> h{0.75f}, e(1);
    table<float> X(h), Y;
    Y = nt2hip::exp(X) * 2.0f;
    nt2ref_map_f32(0, h.data(), e.data(), 1);
    const float want = e[0] * 2.0f;
    bool ok = bits(Y.to_host()[0]) == bits(want);
    ok = ok && nt2hip::globalsum(nt2hip::exp(X) * 2.0f) == want && nt2hip::cumsum(nt2hip::exp(X) * 2.0f).to_host()[0] == want;
    report("edge_one_element", ok);
  }
  {  // +-inf and NaN in the folds and scans of trees (the fold kernels'
     // compensation must not turn an infinite sum into NaN)
    const int64_t R = 1000, C = 6;
    auto h = uniform(R * C, 0.5f, 1.5f, 61);
    h[0 * R + 17] = INFINITY;                              // column 0: +inf
    h[1 * R + 999] = -INFINITY;                            // column 1: -inf in the last row
    h[2 * R + 5] = NAN;                                    // column 2: NaN
    h[3 * R + 1] = INFINITY, h[3 * R + 900] = -INFINITY;   // column 3: both
    for (int64_t i = 0; i < R; ++i) h[4 * R + i] = 1e37f;  // column 4: the sum overflows
    table<float> A(h, R, C);
    auto sums = nt2hip::sum(A * 1.0f, 1).to_host();
    auto norms = nt2hip::norm2(A + 0.0f, 1).to_host();
    auto cs = nt2hip::cumsum(A * 1.0f, 1).to_host();
    bool ok = sums[0] == INFINITY && sums[1] == -INFINITY && std::isnan(sums[2]) && std::isnan(sums[3]) &&
              sums[4] == INFINITY && std::isfinite(sums[5]);
    ok = ok && norms[0] == INFINITY && norms[1] == INFINITY && std::isnan(norms[2]) && norms[3] == INFINITY &&
         norms[4] == INFINITY && std::isfinite(norms[5]);
    ok = ok && std::isfinite(cs[16]) && cs[17] == INFINITY && cs[R - 1] == INFINITY;  // column 0
    ok = ok && std::isfinite(cs[R + 998]) && cs[R + 999] == -INFINITY;                // column 1
    ok = ok && std::isfinite(cs[3 * R + 0]) && cs[3 * R + 1] == INFINITY && std::isnan(cs[3 * R + 900]);
    ok = ok && nt2hip::globalsum(A * 1.0f) != nt2hip::globalsum(A * 1.0f);  // NaN (column 2)
    table<float> P(std::vector<float>(h.begin(), h.begin() + R), R, 1);      // column 0 alone
    ok = ok && nt2hip::globalsum(P) == INFINITY && nt2hip::globalsum(P * 1.0f) == INFINITY &&
         nt2hip::dot(P, P) == INFINITY;
    report("edge_nonfinite_folds_and_scans_of_trees", ok);
  }
}

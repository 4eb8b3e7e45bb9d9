// N-D front-end checks: a double tree, the into-forms fold<OP>/scan<OP>,
// and the speed of a fused fold against the fold of the materialised table.
#include "common.hpp"

// double: (5, 4, 3) with a plane expanded along the first dimension
void nd_checks_f64() {
  const Ext4 sh{{5, 4, 3, 1}}, op{{1, 4, 3, 1}};
  std::vector<double> hd(sh.n()), hp(op.n());
  for (int64_t i = 0; i < sh.n(); ++i) hd[i] = std::sin(0.37 * (double)i) * 3.0;
  for (int64_t i = 0; i < op.n(); ++i) hp[i] = 1.0 + 0.125 * (double)i;
  table<double> D(hd, nt2hip::extent(5, 4, 3)), Pd(hp, nt2hip::extent(1, 4, 3));
  auto e = D * 2.0 + nt2hip::sx(Pd);
  std::vector<double> want(sh.n());
  for (int64_t i = 0; i < sh.n(); ++i) {
    double a = hd[i] * 2.0;
    want[i] = a + hp[expanded_index(i, sh, op)];
  }
  reduce_checks<double>("nd_folds_scans_of_tree_f64", e, want, sh);
}

// into-forms reuse the output; an output that is an operand is staged
void nd_checks_into_forms() {
  table<float> X(uniform(4096, -1.f, 1.f, 40), 64, 64), z(64, 1);
  float* before = z.data();
  nt2hip::fold<NT2H_FOLD_MAXIMUM>(z, nt2hip::abs(X), 2);
  bool ok = z.data() == before && z.rows() == 64 && z.cols() == 1;
  table<float> m = nt2hip::maximum(nt2hip::abs(X), 2);
  ok = ok && same_bits(z.to_host(), m.to_host());
  table<float> c = X;
  nt2hip::scan<NT2H_SCAN_CUMSUM>(c, c * 2.0f, 1);
  table<float> c2 = nt2hip::cumsum(X * 2.0f, 1);
  ok = ok && same_bits(c.to_host(), c2.to_host());
  report("nd_fold_scan_into_forms", ok);
}

// sum(f(X), 2) of a 16384 x 8192 table fused (the tree evaluated where the
// fold loads) against nt2h_fold_f32 over the materialised f(X): kernel time
// only (HIP events, outputs preallocated), best of 15 interleaved runs.
void fused_fold_speed_check() {
  const int64_t R = 16384, C = 8192;
  table<float> X(uniform(R * C, -1.f, 1.f, 41), R, C);
  auto f = nt2hip::abs(X) * 2.0f - 1.0f;
  table<float> F;
  F = f;
  table<float> zf(R, 1), zm(R, 1);
  hipEvent_t a, b;
  hip_ok(hipEventCreate(&a));
  hip_ok(hipEventCreate(&b));
  auto time = [&](bool fused) {
    hip_ok(hipEventRecord(a, nullptr));
    if (fused) nt2hip::fold<NT2H_FOLD_SUM>(zf, f, 2);
    else check(nt2h_fold_f32(NT2H_FOLD_SUM, F.data(), R, C, 1, zm.data(), nullptr), "nt2h_fold_f32");
    hip_ok(hipEventRecord(b, nullptr));
    hip_ok(hipEventSynchronize(b));
    float ms = 0;
    hip_ok(hipEventElapsedTime(&ms, a, b));
    return ms;
  };
  for (int w = 0; w < 3; ++w) { time(true); time(false); }
  float bf = 1e30f, bm = 1e30f;
  for (int r = 0; r < 15; ++r) {
    bf = std::min(bf, time(true));
    bm = std::min(bm, time(false));
  }
  const bool same = same_bits(zf.to_host(), zm.to_host());
  char buf[160];
  std::snprintf(buf, sizeof buf, "fused %.1f us, materialised-table fold %.1f us, ratio %.3f", 1e3 * bf, 1e3 * bm,
                bf / bm);
  report("fused_sum_dim2_16384x8192_within_10pct", same && bf <= 1.10f * bm, buf);
  hip_ok(hipEventDestroy(a));
  hip_ok(hipEventDestroy(b));
}

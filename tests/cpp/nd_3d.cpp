// N-D front-end checks: a 3-D tree with sx() along every dimension, and
// sx(tag, a, b).
#include "common.hpp"

// 3-D: sx along every dimension, (37, 9, 5) against planes, a column slab and a pillar
void nd_checks_3d() {
  const Ext4 sh{{37, 9, 5, 1}}, oP{{1, 9, 5, 1}}, oQ{{37, 1, 5, 1}}, oR{{1, 1, 5, 1}}, oC{{37, 9, 1, 1}};
  auto ha = uniform(sh.n(), -2.f, 2.f, 31), hp = uniform(oP.n(), 0.5f, 1.5f, 32), hq = uniform(oQ.n(), -1.f, 1.f, 33),
       hr = uniform(oR.n(), -1.f, 1.f, 34), hc = uniform(oC.n(), -1.f, 1.f, 35);
  ha[7] = NAN;  // a NaN through every fold and scan
  table<float> A(ha, nt2hip::extent(37, 9, 5)), P(hp, nt2hip::extent(1, 9, 5)), Q(hq, nt2hip::extent(37, 1, 5)),
      R(hr, nt2hip::extent(1, 1, 5)), C(hc, nt2hip::extent(37, 9));
  auto e = A * nt2hip::sx(P) + nt2hip::sx(Q) - nt2hip::sx(R) * nt2hip::sx(C);
  std::vector<float> want(sh.n());
  for (int64_t i = 0; i < sh.n(); ++i) {
    float a = ha[i] * hp[expanded_index(i, sh, oP)];
    float b = a + hq[expanded_index(i, sh, oQ)];
    float c = hr[expanded_index(i, sh, oR)] * hc[expanded_index(i, sh, oC)];
    want[i] = b - c;
  }
  table<float> Z;
  Z = e;
  bool ok = Z.dim(1) == 37 && Z.dim(2) == 9 && Z.dim(3) == 5 && Z.dim(4) == 1 && same_bits(Z.to_host(), want);
  report("nd_sx_every_dimension_bit_exact", ok);
  reduce_checks<float>("nd_folds_scans_of_tree_3d", e, want, sh);
  // sx(tag, a, b): NT2's bsxfun, a column against a row
  table<float> col(std::vector<float>(hq.begin(), hq.begin() + 37), 37, 1),
      row(std::vector<float>(hp.begin(), hp.begin() + 9), 1, 9);
  table<float> O;
  O = nt2hip::sx(nt2hip::tag::plus_(), col, row);
  std::vector<float> wo(37 * 9);
  for (int64_t j = 0; j < 9; ++j)
    for (int64_t i = 0; i < 37; ++i) wo[i + 37 * j] = hq[i] + hp[j];
  report("nd_sx_tag_outer_sum", O.rows() == 37 && O.cols() == 9 && same_bits(O.to_host(), wo));
}

// N-D front-end checks: a 4-D tree with a column expanded along three
// dimensions.
#include "common.hpp"

// 4-D (6, 5, 4, 3) with a column expanded along three dimensions
void nd_checks_4d() {
  const Ext4 sh{{6, 5, 4, 3}}, oc{{6, 1, 1, 1}};
  auto hd = uniform(sh.n(), -1.5f, 1.5f, 36), hv = uniform(6, -1.f, 1.f, 37);
  table<float> D(hd, nt2hip::extent(6, 5, 4, 3)), V(hv, 6, 1);
  auto e = nt2hip::sqr(D) - nt2hip::sx(V);
  std::vector<float> want(sh.n());
  for (int64_t i = 0; i < sh.n(); ++i) {
    float s = hd[i] * hd[i];
    want[i] = s - hv[expanded_index(i, sh, oc)];
  }
  table<float> Z;
  Z = e;
  report("nd_4d_tree_bit_exact", Z.dim(4) == 3 && same_bits(Z.to_host(), want));
  reduce_checks<float>("nd_folds_scans_of_tree_4d", e, want, sh);
}

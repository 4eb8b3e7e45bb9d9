// N-D front-end checks: a (1024, 12, 3) tree through the run kernels
// (dimension 1) and the tile kernels (dimensions 2 and 3); the default
// dimension of a reduction.
#include "common.hpp"

// (1024, 12, 3): the run kernels along 1 (1024 >= 512), the tile kernels along 2 and 3
void nd_checks_paths() {
  const Ext4 sh{{1024, 12, 3, 1}}, op{{1, 12, 3, 1}};
  auto hb = uniform(sh.n(), -3.f, 3.f, 38), hl = uniform(op.n(), 0.5f, 2.f, 39);
  table<float> B(hb, nt2hip::extent(1024, 12, 3)), L(hl, nt2hip::extent(1, 12, 3));
  auto e = nt2hip::abs(B) * nt2hip::sx(L) - 0.25f;
  std::vector<float> want(sh.n());
  for (int64_t i = 0; i < sh.n(); ++i) {
    float a = std::fabs(hb[i]) * hl[expanded_index(i, sh, op)];
    want[i] = a - 0.25f;
  }
  reduce_checks<float>("nd_folds_scans_of_tree_run_and_tile_paths", e, want, sh);
  // the defaults: first non-singleton dimension (3 for a 1 x 1 x p table)
  table<float> Pl(std::vector<float>(hl.begin(), hl.begin() + 3), nt2hip::extent(1, 1, 3));
  table<float> s = nt2hip::sum(Pl * 2.0f);
  const float w = (hl[0] * 2.0f + hl[1] * 2.0f) + hl[2] * 2.0f;
  report("nd_first_nonsingleton_default", s.numel() == 1 && std::fabs(s.to_host()[0] - w) <= 1e-6f * std::fabs(w));
}

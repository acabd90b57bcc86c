// Instantiations of the pipelined star-shape fused fragment kernel with 1 join: every count of
// HBM-bitmap joins (0..1) and aggregates (1..4).
#include "fragment_kernel.hpp"

namespace ghive {

namespace {
template <int NH>
FragKernel pick_na(int na) {
  switch (na) {
    case 1: return k_fragment_pipe<1, NH, 1>;
    case 2: return k_fragment_pipe<1, NH, 2>;
    case 3: return k_fragment_pipe<1, NH, 3>;
    default: return k_fragment_pipe<1, NH, 4>;
  }
}
}  // namespace

FragKernel pick_pipe_j1(int nh, int na) {
  switch (nh) {
    case 0: return pick_na<0>(na);
    case 1: return pick_na<1>(na);
    default: return pick_na<1>(na);
  }
}

}  // namespace ghive

// Instantiations of the pipelined star-shape fused fragment kernel with 4 joins: every count of
// HBM-bitmap joins (0..4) and aggregates (1..4).
#include "fragment_kernel.hpp"

namespace ghive {

namespace {
template <int NH>
FragKernel pick_na(int na) {
  switch (na) {
    case 1: return k_fragment_pipe<4, NH, 1>;
    case 2: return k_fragment_pipe<4, NH, 2>;
    case 3: return k_fragment_pipe<4, NH, 3>;
    default: return k_fragment_pipe<4, NH, 4>;
  }
}
}  // namespace

FragKernel pick_pipe_j4(int nh, int na) {
  switch (nh) {
    case 0: return pick_na<0>(na);
    case 1: return pick_na<1>(na);
    case 2: return pick_na<2>(na);
    case 3: return pick_na<3>(na);
    case 4: return pick_na<4>(na);
    default: return pick_na<4>(na);
  }
}

}  // namespace ghive

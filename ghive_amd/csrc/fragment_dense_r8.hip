// Instantiations of the dense-probe fused fragment kernel with 8 rows per thread.
#include "fragment_kernel.hpp"

namespace ghive {

template <int NJ>
static FragKernel pick_dense_na_r8(int na) {
  switch (na) {
    case 1: return k_fragment_dense<NJ, 1, 8>;
    case 2: return k_fragment_dense<NJ, 2, 8>;
    case 3: return k_fragment_dense<NJ, 3, 8>;
    default: return k_fragment_dense<NJ, 4, 8>;
  }
}

FragKernel pick_dense_r8(int nj, int na) {
  switch (nj) {
    case 1: return pick_dense_na_r8<1>(na);
    case 2: return pick_dense_na_r8<2>(na);
    case 3: return pick_dense_na_r8<3>(na);
    default: return pick_dense_na_r8<4>(na);
  }
}

}  // namespace ghive

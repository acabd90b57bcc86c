// Instantiations of the dense-probe fused fragment kernel with 4 rows per thread.
#include "fragment_kernel.hpp"

namespace ghive {

template <int NJ>
static FragKernel pick_dense_na_r4(int na) {
  switch (na) {
    case 1: return k_fragment_dense<NJ, 1, 4>;
    case 2: return k_fragment_dense<NJ, 2, 4>;
    case 3: return k_fragment_dense<NJ, 3, 4>;
    default: return k_fragment_dense<NJ, 4, 4>;
  }
}

FragKernel pick_dense_r4(int nj, int na) {
  switch (nj) {
    case 1: return pick_dense_na_r4<1>(na);
    case 2: return pick_dense_na_r4<2>(na);
    case 3: return pick_dense_na_r4<3>(na);
    default: return pick_dense_na_r4<4>(na);
  }
}

}  // namespace ghive

// Instantiation of the brick step for 2 bricks (wavefronts) per workgroup.
#include "k_bricks.hpp"

namespace cs {
hipError_t bricks_launch_w2(const cs_bricks_step_args& a, hipStream_t s, int xcd) {
    return launch_bricks<2>(a, s, xcd);
}
hipError_t bricks_launch_pair_w2(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd) {
    return launch_bricks_pair<2>(ab, ai, s, xcd);
}
}  // namespace cs

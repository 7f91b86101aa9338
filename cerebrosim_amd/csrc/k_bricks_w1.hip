// Instantiation of the brick step for 1 brick (wavefront) per workgroup.
#include "k_bricks.hpp"

namespace cs {
hipError_t bricks_launch_w1(const cs_bricks_step_args& a, hipStream_t s, int xcd) {
    return launch_bricks<1>(a, s, xcd);
}
hipError_t bricks_launch_pair_w1(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd) {
    return launch_bricks_pair<1>(ab, ai, s, xcd);
}
}  // namespace cs

/* ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of CerebroSim's sparse D3Q27 lattice-Boltzmann step (lbm.py) and the per-step reductions of
 * simulate.py, in FP32 and FP64, for the parity tests and as the timed CPU baseline. Only tests/, the smoke check
 * and the benchmark's CPU-baseline leg may load it; the product path never does.
 *
 * Lattice constants follow lbm.py:34-50: C = _C19 + _CORNERS (rest, 6 faces, 12 edges, 8 corners), weights
 * 8/27, 2/27, 1/54, 1/216, OPP by search, link codes WALL = -1, INLET = -2, OUTLET k = -3 - k.
 */
#include <math.h>
#include <stdint.h>

#define ORC_Q 27
#define ORC_WALL (-1)
#define ORC_INLET (-2)
#define ORC_OUTLET0 (-3)

static const int ORC_C[ORC_Q][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
    {1, 1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
static const double ORC_W[ORC_Q] = {
    8.0 / 27, 2.0 / 27, 2.0 / 27, 2.0 / 27, 2.0 / 27, 2.0 / 27, 2.0 / 27,
    1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54, 1.0 / 54,
    1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216};
static const int ORC_OPP[ORC_Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
                                   20, 19, 22, 21, 24, 23, 26, 25};

/* Exported so the Python side can check that its own tables agree with these. */
void orc_tables(int32_t* c_out, double* w_out, int32_t* opp_out) {
    for (int i = 0; i < ORC_Q; ++i) {
        for (int a = 0; a < 3; ++a) c_out[3 * i + a] = ORC_C[i][a];
        w_out[i] = ORC_W[i];
        opp_out[i] = ORC_OPP[i];
    }
}

#define REAL float
#define SUF _f32
#define ORC_COS cosf
#define ORC_SIN sinf
#define ORC_SQRT sqrtf
#include "step_impl.h"
#undef REAL
#undef SUF
#undef ORC_COS
#undef ORC_SIN
#undef ORC_SQRT

#define REAL double
#define SUF _f64
#define ORC_COS cos
#define ORC_SIN sin
#define ORC_SQRT sqrt
#include "step_impl.h"
#undef REAL
#undef SUF
#undef ORC_COS
#undef ORC_SIN
#undef ORC_SQRT

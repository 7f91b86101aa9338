// D3Q27 lattice tables for host and device code.
//
// The direction order is the reference's (lbm.py:34-40: rest, +-x, +-y, +-z, 12 edges, 8 corners) because the
// neighbour tables, fixtures and the finite-difference gradient (links 1..6) index by it. Weights are the FP64
// values rounded to float, as the reference's float32 constants (lbm.py:40, 57).
#pragma once
#include <cstdint>

namespace cs {

constexpr int Q = 27;
constexpr int WALL = -1;
constexpr int INLET = -2;
constexpr int OUTLET0 = -3;

struct Lat {
    static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
                                   20, 19, 22, 21, 24, 23, 26, 25};
    static constexpr float w[Q] = {
        float(8.0 / 27), float(2.0 / 27), float(2.0 / 27), float(2.0 / 27), float(2.0 / 27), float(2.0 / 27),
        float(2.0 / 27), float(1.0 / 54), float(1.0 / 54), float(1.0 / 54), float(1.0 / 54), float(1.0 / 54),
        float(1.0 / 54), float(1.0 / 54), float(1.0 / 54), float(1.0 / 54), float(1.0 / 54), float(1.0 / 54),
        float(1.0 / 54), float(1.0 / 216), float(1.0 / 216), float(1.0 / 216), float(1.0 / 216), float(1.0 / 216),
        float(1.0 / 216), float(1.0 / 216), float(1.0 / 216)};
};

}  // namespace cs

// Solver step over the reference's indexed layout: one thread per fluid cell, populations (Q, N) row-major,
// sources taken from the neighbour table. Semantics of lbm.py:140-257 (k_step); this path accepts any neighbour
// table, so it serves lattices that the brick layout cannot represent.
#include "collide.hpp"
#include "common.hpp"

namespace cs {
namespace {

constexpr int kBlock = 256;

// Non-uniform central difference of one velocity component along axis `ax` (links 2ax+1 from x - e, 2ax+2 from
// x + e), lbm.py:105-126. `comp` selects the component (the Galilean term needs only d u_a / d x_a).
__device__ __forceinline__ float ddx_component(const cs_step_args& a, int64_t n, int ax, int comp, float u0c) {
    const int64_t N = a.n_cells;
    const int im = 2 * ax + 1, ip = 2 * ax + 2;
    float h1 = 1.0f, v1 = u0c, h2 = 1.0f, v2 = u0c;
    const int s1 = a.nbr[im * N + n];
    if (s1 >= 0) {
        v1 = a.vel_old[3 * (int64_t)s1 + comp];
    } else if (s1 == WALL) {
        h1 = fmaxf(a.qw[im * N + n], 0.1f);
        v1 = 0.0f;
    }
    const int s2 = a.nbr[ip * N + n];
    if (s2 >= 0) {
        v2 = a.vel_old[3 * (int64_t)s2 + comp];
    } else if (s2 == WALL) {
        h2 = fmaxf(a.qw[ip * N + n], 0.1f);
        v2 = 0.0f;
    }
    return (h1 * h1 * (v2 - u0c) + h2 * h2 * (u0c - v1)) / (h1 * h2 * (h1 + h2));
}

template <int MODEL, bool GAL, bool HYB>
__global__ __launch_bounds__(kBlock) void k_step_indexed(cs_step_args a) {
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t N = a.n_cells;
    if (n >= N) return;
    const float om = a.om_cell[n];
    const float d = a.drive[0];
    const float up0 = a.vel_old[3 * n], up1 = a.vel_old[3 * n + 1], up2 = a.vel_old[3 * n + 2];
    const float upsq = up0 * up0 + up1 * up1 + up2 * up2;
    float f[Q];
    // pull streaming with boundary links (lbm.py:156-182)
    static_for<0, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int io = Lat::opp[i];
        const int src = a.nbr[i * N + n];
        float v;
        if (src >= 0) {
            v = a.f_in[i * N + src];
        } else if (src == WALL) {
            const float fo = a.f_in[io * N + n];
            v = fo;
            if (a.bouzidi) {
                const float q = a.qw[i * N + n];
                if (q < 0.5f) {
                    const int m = a.nbr[io * N + n];
                    if (m >= 0) v = 2.0f * q * fo + (1.0f - 2.0f * q) * a.f_in[io * N + m];
                } else {
                    v = fo / (2.0f * q) + (2.0f * q - 1.0f) / (2.0f * q) * a.f_in[i * N + n];
                }
            }
        } else if (src == INLET) {
            const float cu = cdot<i>(a.uin[3 * n] * d, a.uin[3 * n + 1] * d, a.uin[3 * n + 2] * d);
            v = a.f_in[io * N + n] + 6.0f * Lat::w[i] * cu;
        } else {
            const int k = OUTLET0 - src;
            const float cu = cdot<i>(up0, up1, up2);
            v = -a.f_in[io * N + n] + 2.0f * Lat::w[i] * (3.0f * a.p_out[k] + 4.5f * cu * cu - 1.5f * upsq);
        }
        f[i] = v;
    });
    CollideParams p;
    p.om = om;
    p.om_m = MODEL == 0 ? 1.0f / (0.5f + a.magic / (1.0f / om - 0.5f)) : 0.0f;
    p.F[0] = a.force[0] * d;
    p.F[1] = a.force[1] * d;
    p.F[2] = a.force[2] * d;
    p.sigma = a.sigma;
    p.t3 = GAL ? 1.0f : 0.0f;
    p.g[0] = p.g[1] = p.g[2] = 0.0f;
    if constexpr (GAL && !HYB) {
        p.g[0] = ddx_component(a, n, 0, 0, up0);
        p.g[1] = ddx_component(a, n, 1, 1, up1);
        p.g[2] = ddx_component(a, n, 2, 2, up2);
    }
    if constexpr (HYB) {
        const float upc[3] = {up0, up1, up2};
        for (int ax = 0; ax < 3; ++ax)
            for (int c = 0; c < 3; ++c) p.G[ax][c] = ddx_component(a, n, ax, c, upc[c]);
        if constexpr (GAL) {
            p.g[0] = p.G[0][0];
            p.g[1] = p.G[1][1];
            p.g[2] = p.G[2][2];
        }
    }
    if (a.inl[n] != 0) {
        // inlet cells: equilibrium at the prescribed velocity (lbm.py:205-216)
        const float dr = density_deviation(f);
        const float ut0 = a.uin[3 * n] * d, ut1 = a.uin[3 * n + 1] * d, ut2 = a.uin[3 * n + 2] * d;
        inlet_equilibrium(f, dr, ut0, ut1, ut2, p.t3);
        static_for<0, Q>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            a.f_out[i * N + n] = f[i];
        });
        a.vel[3 * n] = ut0;
        a.vel[3 * n + 1] = ut1;
        a.vel[3 * n + 2] = ut2;
        a.rho[n] = dr;
        return;
    }
    float dr, u[3];
    collide<MODEL, HYB>(f, p, dr, u);
    static_for<0, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        a.f_out[i * N + n] = f[i];
    });
    a.vel[3 * n] = u[0];
    a.vel[3 * n + 1] = u[1];
    a.vel[3 * n + 2] = u[2];
    a.rho[n] = dr;
}

template <int MODEL>
hipError_t launch_model(const cs_step_args& a, hipStream_t s, bool gal, bool hyb) {
    const unsigned grid = (unsigned)((a.n_cells + kBlock - 1) / kBlock);
    if (hyb) {
        if (gal) k_step_indexed<MODEL, true, true><<<grid, kBlock, 0, s>>>(a);
        else k_step_indexed<MODEL, false, true><<<grid, kBlock, 0, s>>>(a);
    } else {
        if (gal) k_step_indexed<MODEL, true, false><<<grid, kBlock, 0, s>>>(a);
        else k_step_indexed<MODEL, false, false><<<grid, kBlock, 0, s>>>(a);
    }
    return hipGetLastError();
}

}  // namespace
}  // namespace cs

extern "C" int cs_step_indexed(const cs_step_args* a, void* stream) {
    CS_REQUIRE(a != nullptr, "cs_step_indexed: null arguments");
    CS_REQUIRE(a->n_cells > 0 && a->n_cells < (int64_t)INT32_MAX, "cs_step_indexed: n_cells %lld out of range",
               (long long)a->n_cells);
    CS_REQUIRE(a->f_in && a->f_out && a->nbr && a->qw && a->uin && a->inl && a->p_out && a->drive && a->om_cell &&
                   a->vel_old && a->vel && a->rho,
               "cs_step_indexed: null device pointer");
    CS_REQUIRE(a->model == CS_MODEL_TRT || a->model == CS_MODEL_REGULARISED, "cs_step_indexed: unknown model %d",
               a->model);
    const bool hyb = a->model == CS_MODEL_REGULARISED && a->sigma < 1.0f;
    hipError_t e = a->model == CS_MODEL_REGULARISED
                       ? cs::launch_model<1>(*a, cs::as_stream(stream), a->galilean != 0, hyb)
                       : cs::launch_model<0>(*a, cs::as_stream(stream), a->galilean != 0, false);
    if (e != hipSuccess) {
        cs::set_error("cs_step_indexed launch failed: %s", hipGetErrorString(e));
        return CS_E_HIP;
    }
    return CS_OK;
}

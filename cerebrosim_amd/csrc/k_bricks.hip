// Solver step over the brick layout with AA-pattern in-place streaming (semantics of lbm.py:140-257).
//
// Storage: S[b][k][l] = f[(b * 27 + k) * 64 + l], l = x + 4 y + 16 z inside 4^3 brick b. One wavefront per brick,
// lane l = cell l; four bricks per workgroup.
//
// AA pattern. Let P_i(x) be the post-collision population of cell x in direction i.
//  * even step (parity 0) reads S[n][i] and writes P_i(n) to S[n][opp i]: everything stays in the cell's own slots;
//  * odd step (parity 1) reads S[n - c_i][opp i] (the even step's output of the upstream cell) and writes P_i(n) to
//    S[n + c_i][i], which is where the next even step of n + c_i reads it. Every slot is read and then written by one
//    thread only, so one copy of the populations suffices.
//  * A missing link i of n (wall, inlet, outlet: n - c_i not fluid) reads S[n][i]; the step that wrote it stored
//    there P_opp(n), the population that left n towards the boundary (odd steps redirect their write to a
//    non-fluid n + c_i into S[n][opp i]). That is the reference's f_o = f_in[opp i, n] in both parities.
//  * Bouzidi with q < 1/2 needs f_in[opp i, n + c_i]: the population n pulls for direction opp i, already in hand.
//    Bouzidi with q >= 1/2 needs P_i(n) of the previous step: when n + c_i is not fluid it sits in S[n][opp i]
//    (the missing link opp i's own raw value), otherwise it lives in a slot another thread rewrites during the
//    step, so each such link keeps a copy in the side buffer, written by n at the end of the previous step.
// Interior bricks (all cells fluid, all links fluid) run a separate instantiation without any boundary data.
#include "collide.hpp"
#include "common.hpp"

namespace cs {
namespace {

constexpr int BC = 64;
constexpr int WAVES = 4;
constexpr int WG = 64 * WAVES;
constexpr uint32_t kInletBit = 1u << 27;
constexpr uint32_t kAllLinks = (1u << Q) - 1;

// Brick-table index d and local index of the position (x, y, z) + (SX, SY, SZ), SX.. in {-1, 0, 1}.
template <int SX, int SY, int SZ>
__device__ __forceinline__ void shift_pos(int x, int y, int z, int& d, int& l) {
    int o = 13;
    if constexpr (SX < 0) o -= (x == 0) ? 1 : 0;
    if constexpr (SX > 0) o += (x == 3) ? 1 : 0;
    if constexpr (SY < 0) o -= (y == 0) ? 3 : 0;
    if constexpr (SY > 0) o += (y == 3) ? 3 : 0;
    if constexpr (SZ < 0) o -= (z == 0) ? 9 : 0;
    if constexpr (SZ > 0) o += (z == 3) ? 9 : 0;
    d = o;
    l = ((x + SX) & 3) + 4 * ((y + SY) & 3) + 16 * ((z + SZ) & 3);
}

__device__ __forceinline__ size_t brick_base(int b) { return (size_t)b * (Q * BC); }

// Value streamed into cell n through missing link I (wall / inlet / outlet), lbm.py:163-181.
//   fo        raw value of slot I: P_opp(n)
//   opp_fluid link opp(I) is fluid; f_opp its pulled population f_in[opp I, n + c_I]
//   p_self    P_I(n) of the previous step (for q >= 1/2)
template <int I>
__device__ __forceinline__ float link_value(float code, float fo, bool opp_fluid, float f_opp, float p_self, int bouzidi,
                                            const float (&uind)[3], const float (&up)[3], float upsq,
                                            const float* p_out) {
    if (code >= 0.0f) {
        float v = fo;
        if (bouzidi) {
            const float q = code;
            if (q < 0.5f) {
                if (opp_fluid) v = 2.0f * q * fo + (1.0f - 2.0f * q) * f_opp;
            } else {
                v = fo / (2.0f * q) + (2.0f * q - 1.0f) / (2.0f * q) * p_self;
            }
        }
        return v;
    }
    if (code > -1.5f) {  // inlet: bounce-back with the prescribed velocity
        return fo + 6.0f * Lat::w[I] * cdot<I>(uind[0], uind[1], uind[2]);
    }
    const int k = (int)(-2.0f - code);  // outlet k: anti-bounce-back at the prescribed pressure
    const float cu = cdot<I>(up[0], up[1], up[2]);
    return -fo + 2.0f * Lat::w[I] * (3.0f * p_out[k] + 4.5f * cu * cu - 1.5f * upsq);
}

template <int MODEL, bool GAL, bool BND, int PAR>
__global__ __launch_bounds__(WG) void k_bricks(cs_bricks_step_args a) {
    __shared__ int tbl[WAVES][32];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int L = threadIdx.x & 63;
    // XCD-aware order: the 8 XCDs take workgroups round-robin; give each XCD a contiguous run of bricks
    const int per = gridDim.x >> 3;
    const int wg = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int64_t li = (int64_t)wg * WAVES + wave;
    const bool valid = li < a.n_list;
    const int b = valid ? a.list[li] : 0;
    if (L < 27) tbl[wave][L] = valid ? a.nbr[(int64_t)b * 27 + L] : -1;
    __syncthreads();
    if (!valid) return;

    const int x = L & 3, y = (L >> 2) & 3, z = L >> 4;
    const size_t S = (size_t)a.n_slots;
    const size_t slot = (size_t)b * BC + L;
    const size_t base = brick_base(b);
    float* __restrict__ F = a.f;

    uint32_t mask = kAllLinks;
    int bb = 0;
    if constexpr (BND) {
        bb = a.bidx[b];
        mask = a.lmask[(size_t)bb * BC + L];
        if (mask == 0) return;  // solid slot
    }
    const float om = a.om_brick[b] ? a.om_cell[slot] : a.omega;
    const float d = a.drive[0];

    // ---- streaming: raw values ---------------------------------------------------------------------------
    float f[Q];
    static_for<0, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int io = Lat::opp[i];
        if constexpr (PAR == 0 || i == 0) {
            f[i] = F[base + i * BC + L];
        } else {
            int dd, l;
            shift_pos<-Lat::cx[i], -Lat::cy[i], -Lat::cz[i]>(x, y, z, dd, l);
            if (!BND || ((mask >> i) & 1u)) {
                f[i] = F[brick_base(tbl[wave][dd]) + io * BC + l];
            } else {
                f[i] = F[base + i * BC + L];
            }
        }
    });

    // ---- boundary links and the previous velocity ------------------------------------------------------------
    float up[3] = {0.0f, 0.0f, 0.0f};
    if constexpr (BND) {
        up[0] = a.vel_old[slot];
        up[1] = a.vel_old[S + slot];
        up[2] = a.vel_old[2 * S + slot];
        if ((mask & kAllLinks) != kAllLinks) {
            const float upsq = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
            float uind[3] = {0.0f, 0.0f, 0.0f};
            if (mask & kInletBit) {
                uind[0] = a.uin[((size_t)bb * 3 + 0) * BC + L] * d;
                uind[1] = a.uin[((size_t)bb * 3 + 1) * BC + L] * d;
                uind[2] = a.uin[((size_t)bb * 3 + 2) * BC + L] * d;
            }
            const uint32_t sm = a.bouzidi ? a.smask[(size_t)bb * BC + L] : 0u;
            const float* qc = a.qc + (size_t)bb * Q * BC + L;
            const float* side = a.side + (size_t)bb * Q * BC + L;
            static_for<1, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if constexpr (pair_head<i>()) {
                    constexpr int j = i + 1;
                    const bool fi = (mask >> i) & 1u, fj = (mask >> j) & 1u;
                    const float ri = f[i], rj = f[j];
                    float vi = ri, vj = rj;
                    if (!fi) {
                        const float ps = ((sm >> i) & 1u) ? side[i * BC] : rj;
                        vi = link_value<i>(qc[i * BC], ri, fj, rj, ps, a.bouzidi, uind, up, upsq, a.p_out);
                    }
                    if (!fj) {
                        const float ps = ((sm >> j) & 1u) ? side[j * BC] : ri;
                        vj = link_value<j>(qc[j * BC], rj, fi, ri, ps, a.bouzidi, uind, up, upsq, a.p_out);
                    }
                    f[i] = vi;
                    f[j] = vj;
                }
            });
        }
    }

    // ---- Galilean correction: d u_a / d x_a of the previous velocity (lbm.py:105-126, 196-202) ----------------
    CollideParams p;
    p.om = om;
    p.om_m = MODEL == 0 ? 1.0f / (0.5f + a.magic / (1.0f / om - 0.5f)) : 0.0f;
    p.F[0] = a.force[0] * d;
    p.F[1] = a.force[1] * d;
    p.F[2] = a.force[2] * d;
    p.t3 = GAL ? 1.0f : 0.0f;
    p.sigma = 1.0f;
    p.g[0] = p.g[1] = p.g[2] = 0.0f;
    if constexpr (GAL) {
        static_for<0, 3>([&](auto ac) {
            constexpr int ax = decltype(ac)::value;
            constexpr int i1 = 2 * ax + 1, i2 = 2 * ax + 2;
            int d1, l1, d2, l2;
            shift_pos<ax == 0 ? -1 : 0, ax == 1 ? -1 : 0, ax == 2 ? -1 : 0>(x, y, z, d1, l1);
            shift_pos<ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0>(x, y, z, d2, l2);
            const float* V = a.vel_old + ax * S;
            if constexpr (!BND) {
                const float v1 = V[(size_t)tbl[wave][d1] * BC + l1];
                const float v2 = V[(size_t)tbl[wave][d2] * BC + l2];
                p.g[ax] = (v2 - v1) * 0.5f;
            } else {
                float h1 = 1.0f, v1 = up[ax], h2 = 1.0f, v2 = up[ax];
                if ((mask >> i1) & 1u) {
                    v1 = V[(size_t)tbl[wave][d1] * BC + l1];
                } else {
                    const float c = a.qc[((size_t)bb * Q + i1) * BC + L];
                    if (c >= 0.0f) {
                        h1 = fmaxf(c, 0.1f);
                        v1 = 0.0f;
                    }
                }
                if ((mask >> i2) & 1u) {
                    v2 = V[(size_t)tbl[wave][d2] * BC + l2];
                } else {
                    const float c = a.qc[((size_t)bb * Q + i2) * BC + L];
                    if (c >= 0.0f) {
                        h2 = fmaxf(c, 0.1f);
                        v2 = 0.0f;
                    }
                }
                p.g[ax] = (h1 * h1 * (v2 - up[ax]) + h2 * h2 * (up[ax] - v1)) / (h1 * h2 * (h1 + h2));
            }
        });
    }

    // ---- collision ---------------------------------------------------------------------------------------------
    float dr, u[3];
    bool inlet_cell = false;
    if constexpr (BND) inlet_cell = (mask & kInletBit) != 0u;
    if (inlet_cell) {
        dr = density_deviation(f);
        u[0] = a.uin[((size_t)bb * 3 + 0) * BC + L] * d;
        u[1] = a.uin[((size_t)bb * 3 + 1) * BC + L] * d;
        u[2] = a.uin[((size_t)bb * 3 + 2) * BC + L] * d;
        inlet_equilibrium(f, dr, u[0], u[1], u[2], p.t3);
    } else {
        collide<MODEL, false>(f, p, dr, u);
    }

    // ---- outputs -----------------------------------------------------------------------------------------------
    a.vel[slot] = u[0];
    a.vel[S + slot] = u[1];
    a.vel[2 * S + slot] = u[2];
    a.rho[slot] = dr;
    if constexpr (BND) {
        const uint32_t sm = a.bouzidi ? a.smask[(size_t)bb * BC + L] : 0u;
        if (sm) {
            float* side = a.side + (size_t)bb * Q * BC + L;
            static_for<1, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if ((sm >> i) & 1u) side[i * BC] = f[i];
            });
        }
    }
    static_for<0, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int io = Lat::opp[i];
        if constexpr (PAR == 0 || i == 0) {
            F[base + io * BC + L] = f[i];
        } else {
            if (!BND || ((mask >> io) & 1u)) {
                int dd, l;
                shift_pos<Lat::cx[i], Lat::cy[i], Lat::cz[i]>(x, y, z, dd, l);
                F[brick_base(tbl[wave][dd]) + i * BC + l] = f[i];
            } else {
                F[base + io * BC + L] = f[i];
            }
        }
    });
}

// ---- conversions between the AA storage and the reference (Q, N) layout --------------------------------------
__device__ __forceinline__ uint32_t cell_mask(const cs_bricks_convert_args& a, int b, int L, int& bb) {
    bb = a.bidx[b];
    return bb < 0 ? kAllLinks : a.lmask[(size_t)bb * BC + L];
}

// Location of P_i(n) in the AA storage for the state read by a step of parity `par`.
template <int I>
__device__ __forceinline__ size_t location(const cs_bricks_convert_args& a, int b, int L, uint32_t mask, int par) {
    constexpr int io = Lat::opp[I];
    if (par == 1 || I == 0 || !((mask >> io) & 1u)) return brick_base(b) + io * BC + L;
    const int x = L & 3, y = (L >> 2) & 3, z = L >> 4;
    int dd, l;
    shift_pos<Lat::cx[I], Lat::cy[I], Lat::cz[I]>(x, y, z, dd, l);
    const int nbk = a.nbr[(size_t)b * 27 + dd];
    return brick_base(nbk) + I * BC + l;
}

template <bool SET>
__global__ __launch_bounds__(256) void k_convert(cs_bricks_convert_args a) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= a.n_cells) return;
    const int64_t s = a.slot_of_cell[n];
    const int b = (int)(s / BC), L = (int)(s % BC);
    int bb;
    const uint32_t mask = cell_mask(a, b, L, bb);
    const uint32_t sm = bb >= 0 ? a.smask[(size_t)bb * BC + L] : 0u;
    static_for<0, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const size_t loc = location<i>(a, b, L, mask, a.parity);
        float* ref = a.f_ref + (size_t)i * a.n_cells + n;
        if constexpr (SET) {
            a.f[loc] = *ref;
            if ((sm >> i) & 1u) a.side[((size_t)bb * Q + i) * BC + L] = *ref;
        } else {
            *ref = a.f[loc];
        }
    });
}

template <int MODEL, bool GAL, bool BND>
hipError_t launch_par(const cs_bricks_step_args& a, hipStream_t s) {
    const int64_t groups = (a.n_list + WAVES - 1) / WAVES;
    const unsigned grid = (unsigned)(((groups + 7) / 8) * 8);
    if (a.parity == 0) k_bricks<MODEL, GAL, BND, 0><<<grid, WG, 0, s>>>(a);
    else k_bricks<MODEL, GAL, BND, 1><<<grid, WG, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODEL>
hipError_t launch_model(const cs_bricks_step_args& a, hipStream_t s) {
    if (a.galilean) return a.boundary ? launch_par<MODEL, true, true>(a, s) : launch_par<MODEL, true, false>(a, s);
    return a.boundary ? launch_par<MODEL, false, true>(a, s) : launch_par<MODEL, false, false>(a, s);
}

}  // namespace
}  // namespace cs

extern "C" int cs_bricks_step(const cs_bricks_step_args* a, void* stream) {
    CS_REQUIRE(a && a->f && a->nbr && a->list && a->bidx && a->om_cell && a->om_brick && a->p_out && a->drive &&
                   a->vel_old && a->vel && a->rho,
               "cs_bricks_step: null argument");
    CS_REQUIRE(!a->boundary || (a->lmask && a->smask && a->qc && a->uin && a->side),
               "cs_bricks_step: boundary bricks need lmask/smask/qc/uin/side");
    CS_REQUIRE(a->model == CS_MODEL_TRT || a->model == CS_MODEL_REGULARISED, "cs_bricks_step: unknown model %d",
               a->model);
    CS_REQUIRE(a->parity == 0 || a->parity == 1, "cs_bricks_step: parity %d", a->parity);
    CS_REQUIRE(a->n_slots > 0 && a->n_slots % cs::BC == 0, "cs_bricks_step: n_slots %lld", (long long)a->n_slots);
    if (a->n_list == 0) return CS_OK;
    CS_REQUIRE(a->n_list > 0 && a->n_list <= a->n_slots / cs::BC, "cs_bricks_step: n_list %lld",
               (long long)a->n_list);
    hipError_t e = a->model == CS_MODEL_REGULARISED ? cs::launch_model<1>(*a, cs::as_stream(stream))
                                                     : cs::launch_model<0>(*a, cs::as_stream(stream));
    if (e != hipSuccess) {
        cs::set_error("cs_bricks_step launch failed: %s", hipGetErrorString(e));
        return CS_E_HIP;
    }
    return CS_OK;
}

static int convert(const cs_bricks_convert_args* a, void* stream, bool set) {
    CS_REQUIRE(a && a->f && a->nbr && a->bidx && a->slot_of_cell && a->f_ref && a->lmask && a->smask && a->side,
               "cs_bricks convert: null argument");
    CS_REQUIRE(a->parity == 0 || a->parity == 1, "cs_bricks convert: parity %d", a->parity);
    if (a->n_cells == 0) return CS_OK;
    const unsigned grid = (unsigned)((a->n_cells + 255) / 256);
    if (set) cs::k_convert<true><<<grid, 256, 0, cs::as_stream(stream)>>>(*a);
    else cs::k_convert<false><<<grid, 256, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_convert");
    return CS_OK;
}

extern "C" int cs_bricks_get_f(const cs_bricks_convert_args* a, void* stream) { return convert(a, stream, false); }
extern "C" int cs_bricks_set_f(const cs_bricks_convert_args* a, void* stream) { return convert(a, stream, true); }

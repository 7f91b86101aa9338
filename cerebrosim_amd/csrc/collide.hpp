// Collision of one cell, shared by every storage layout (device code).
//
// Implements the reference's regularised BGK with Guo forcing, the D3Q27 Galilean correction and the optional
// hybrid blend (lbm.py:183-240), its TRT branch (lbm.py:241-255) and the equilibrium of inlet cells
// (lbm.py:205-216). The arithmetic is reorganised for the GPU; the result equals the reference formula up to FP32
// rounding:
//  * directions are walked at compile time (static_for + if constexpr) so that lattice components equal to zero
//    produce no instructions (x * 0.0f cannot be folded away under IEEE rules);
//  * opposite directions i, OPP[i] are handled together: every term of the output is either even or odd in c_i,
//    so out_i = w_i (E + O) and out_opp = w_i (E - O);
//  * the second moment of the non-equilibrium part uses the D3Q27 isotropy identity
//    sum_i c_a c_b feq2_i = dr/3 delta_ab + u_a u_b, i.e. P = sum_i c_a c_b f_i - dr/3 delta - u u^T;
//  * the forcing coefficients are collected: -1.5 (1 - om) + 3 (1 - om/2) = 1.5 for c.F.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "d3q27.hpp"

namespace cs {

template <int I, int N, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// c_i . v with zero components removed at compile time (-0.0f is the additive identity that folds away).
template <int I>
__device__ __forceinline__ float cdot(float x, float y, float z) {
    constexpr int a = Lat::cx[I], b = Lat::cy[I], c = Lat::cz[I];
    float s = -0.0f;
    if constexpr (a == 1) s = s + x;
    if constexpr (a == -1) s = s - x;
    if constexpr (b == 1) s = s + y;
    if constexpr (b == -1) s = s - y;
    if constexpr (c == 1) s = s + z;
    if constexpr (c == -1) s = s - z;
    return s;
}

// Products of velocity components used by the third-order Hermite term (lbm.py:129-137).
struct Cubic {
    float xxy, xxz, xyy, yyz, xzz, yzz, xyz;
};

__device__ __forceinline__ Cubic cubic_terms(float u0, float u1, float u2) {
    Cubic k;
    float xx = u0 * u0, yy = u1 * u1, zz = u2 * u2;
    k.xxy = xx * u1;
    k.xxz = xx * u2;
    k.xyy = u0 * yy;
    k.yyz = yy * u2;
    k.xzz = u0 * zz;
    k.yzz = u1 * zz;
    k.xyz = u0 * u1 * u2;
    return k;
}

// _h3(c_i, u) for a compile-time direction: 13.5 sum (c_a^2 - 1/3)(...) + 27 cx cy cz ux uy uz.
template <int I>
__device__ __forceinline__ float h3(const Cubic& k) {
    constexpr int a = Lat::cx[I], b = Lat::cy[I], c = Lat::cz[I];
    constexpr float ka = 13.5f * (float(a * a) - 1.0f / 3.0f);
    constexpr float kb = 13.5f * (float(b * b) - 1.0f / 3.0f);
    constexpr float kc = 13.5f * (float(c * c) - 1.0f / 3.0f);
    float s = -0.0f;
    // (c_x^2 - 1/3)(c_y u_x^2 u_y + c_z u_x^2 u_z)
    if constexpr (b != 0) s = s + (ka * float(b)) * k.xxy;
    if constexpr (c != 0) s = s + (ka * float(c)) * k.xxz;
    // (c_y^2 - 1/3)(c_x u_x u_y^2 + c_z u_y^2 u_z)
    if constexpr (a != 0) s = s + (kb * float(a)) * k.xyy;
    if constexpr (c != 0) s = s + (kb * float(c)) * k.yyz;
    // (c_z^2 - 1/3)(c_x u_x u_z^2 + c_y u_y u_z^2)
    if constexpr (a != 0) s = s + (kc * float(a)) * k.xzz;
    if constexpr (b != 0) s = s + (kc * float(b)) * k.yzz;
    if constexpr (a != 0 && b != 0 && c != 0) s = s + (27.0f * float(a * b * c)) * k.xyz;
    return s;
}

// Directions with i odd (1, 3, ..., 25) are one member of each opposite pair; OPP[i] = i + 1.
template <int I>
constexpr bool pair_head() {
    return I % 2 == 1;
}

// Streamed populations -> density deviation, momentum j and the raw second moment Pi = sum c c^T f.
struct RawMoments {
    float dr, jx, jy, jz, pxx, pyy, pzz, pxy, pxz, pyz;
};

// GW: the inputs are g_i = f_i / w_i, and each sum takes the weight in a fused multiply-add (as many instructions
// as the plain sums, and the caller saves the 27 multiplications by w_i).
template <bool GW = false>
__device__ __forceinline__ RawMoments raw_moments(const float (&f)[Q]) {
    RawMoments m;
    float dr = GW ? Lat::w[0] * f[0] : f[0], jx = -0.0f, jy = -0.0f, jz = -0.0f;
    float pxx = -0.0f, pyy = -0.0f, pzz = -0.0f, pxy = -0.0f, pxz = -0.0f, pyz = -0.0f;
    static_for<1, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (pair_head<i>()) {
            constexpr int a = Lat::cx[i], b = Lat::cy[i], c = Lat::cz[i];
            float s = f[i] + f[i + 1];
            float d = f[i] - f[i + 1];
            // acc + sg * w * v (GW) or acc + sg * v
            auto acc = [](float& x, float sg, float v) {
                if constexpr (GW) x = __builtin_fmaf(sg * Lat::w[i], v, x);
                else x = sg > 0.0f ? x + v : x - v;
            };
            acc(dr, 1.0f, s);
            if constexpr (a != 0) acc(jx, float(a), d);
            if constexpr (b != 0) acc(jy, float(b), d);
            if constexpr (c != 0) acc(jz, float(c), d);
            if constexpr (a != 0) acc(pxx, 1.0f, s);
            if constexpr (b != 0) acc(pyy, 1.0f, s);
            if constexpr (c != 0) acc(pzz, 1.0f, s);
            if constexpr (a * b != 0) acc(pxy, float(a * b), s);
            if constexpr (a * c != 0) acc(pxz, float(a * c), s);
            if constexpr (b * c != 0) acc(pyz, float(b * c), s);
        }
    });
    m.dr = dr;
    m.jx = jx;
    m.jy = jy;
    m.jz = jz;
    m.pxx = pxx;
    m.pyy = pyy;
    m.pzz = pzz;
    m.pxy = pxy;
    m.pxz = pxz;
    m.pyz = pyz;
    return m;
}

// Inputs of a collision besides the populations.
struct CollideParams {
    float om;          // 1 / tau of this cell
    float om_m;        // TRT odd rate (unused by the regularised model)
    float F[3];        // force * drive
    float t3;          // 1 with the Galilean correction, else 0
    float g[3];        // d u_x/dx, d u_y/dy, d u_z/dz of the previous step (Galilean correction)
    float sigma;       // hybrid blend (regularised only)
    float G[3][3];     // full previous-step velocity gradient, rows d/dx, d/dy, d/dz (hybrid only)
};

// Regularised BGK (lbm.py:217-240) or TRT (lbm.py:241-255) of one cell, in place. Returns dr and u (= j + F/2).
// uset: use the velocity ut instead of j + F/2. With om = 1, F = 0 and g = 0 the regularised collision is then
// exactly the inlet equilibrium of lbm.py:205-216 (the same expressions, each added term an exact zero), so inlet
// cells share this code with their neighbours instead of diverging into a second path.
//
// GW and NEQ (regularised only) serve the float16 storage: with GW the inputs are g_i = f_i / w_i; with NEQ the
// outputs are the stored form nscale * (P_i - feq2_i) / w_i, where feq2_i / w_i = rho' + 3 c.u + 4.5 (c.u)^2 and
// rho' = dr - 1.5 |u|^2 with this cell's new dr and u. Subtracting feq2 from the expressions below cancels in closed
// form:
//   (P_i - feq2_i) / w_i = (A - rho') + k1 c.P.c + kc (c.u)(c.F) + (1.5 c.F + t3 h3_i) for the pair head i,
//   the same with the last bracket negated for OPP[i], and A - rho' for the rest direction, where
//   A - rho' = -k1 tr P - 3 (1 - om/2) u.F.
// So no equilibrium is formed and subtracted again, and nothing cancels in floating point.
template <int MODEL, bool HYB, bool GW = false, bool NEQ = false>
__device__ __forceinline__ void collide(float (&f)[Q], const CollideParams& p, float& dr_out, float (&u)[3],
                                        bool uset = false, float ut0 = 0.0f, float ut1 = 0.0f, float ut2 = 0.0f,
                                        float nscale = 1.0f) {
    static_assert(!(GW || NEQ) || (MODEL == 1 && !HYB), "float16 forms of the regularised collision only");
    RawMoments m = raw_moments<GW>(f);
    const float F0 = p.F[0], F1 = p.F[1], F2 = p.F[2];
    const float u0 = uset ? ut0 : m.jx + F0 * 0.5f, u1 = uset ? ut1 : m.jy + F1 * 0.5f,
                u2 = uset ? ut2 : m.jz + F2 * 0.5f;
    const float usq = u0 * u0 + u1 * u1 + u2 * u2;
    const float uF = u0 * F0 + u1 * F1 + u2 * F2;
    const float om = p.om;
    const float tp = 1.0f / om;
    // Galilean diagonal correction Ed_a = 3 tau u_a^2 d_a u_a (lbm.py:196-202); t3 == 0 disables it
    const float Ed0 = p.t3 * (3.0f * tp * u0 * u0 * p.g[0]);
    const float Ed1 = p.t3 * (3.0f * tp * u1 * u1 * p.g[1]);
    const float Ed2 = p.t3 * (3.0f * tp * u2 * u2 * p.g[2]);
    const Cubic k = cubic_terms(u0, u1, u2);
    const float dr = m.dr;
    if constexpr (MODEL == 1) {
        const float d3 = dr * (1.0f / 3.0f);
        float Pxx = m.pxx - d3 - u0 * u0, Pyy = m.pyy - d3 - u1 * u1, Pzz = m.pzz - d3 - u2 * u2;
        float Pxy = m.pxy - u0 * u1, Pxz = m.pxz - u0 * u2, Pyz = m.pyz - u1 * u2;
        if constexpr (HYB) {
            const float s = p.sigma, kk = (1.0f - s) / (3.0f * om);
            Pxx = s * Pxx - kk * (p.G[0][0] + p.G[0][0]);
            Pyy = s * Pyy - kk * (p.G[1][1] + p.G[1][1]);
            Pzz = s * Pzz - kk * (p.G[2][2] + p.G[2][2]);
            Pxy = s * Pxy - kk * (p.G[0][1] + p.G[1][0]);
            Pxz = s * Pxz - kk * (p.G[0][2] + p.G[2][0]);
            Pyz = s * Pyz - kk * (p.G[1][2] + p.G[2][1]);
        }
        Pxx -= Ed0;
        Pyy -= Ed1;
        Pzz -= Ed2;
        const float trP = (Pxx + Pyy + Pzz) * (1.0f / 3.0f);
        const float k1 = 4.5f * (1.0f - om);
        const float kc = 9.0f * (1.0f - 0.5f * om);
        const float Pxy2 = 2.0f * Pxy, Pxz2 = 2.0f * Pxz, Pyz2 = 2.0f * Pyz;
        const float sxy = Pxx + Pyy, sxz = Pxx + Pzz, syz = Pyy + Pzz, sxyz = Pxx + Pyy + Pzz;
        if constexpr (NEQ) {
            // the stored form, scaled by nscale (every coefficient carries the scale once)
            const float An = nscale * (-k1 * trP - 3.0f * (1.0f - 0.5f * om) * uF);
            const float k1n = nscale * k1, kcn = nscale * kc, fn = nscale * 1.5f, hn = nscale * p.t3;
            const float Fn0 = kcn * F0, Fn1 = kcn * F1, Fn2 = kcn * F2;
            const float Fo0 = fn * F0, Fo1 = fn * F1, Fo2 = fn * F2;
            f[0] = An;
            static_for<1, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if constexpr (pair_head<i>()) {
                    constexpr int a = Lat::cx[i], b = Lat::cy[i], c = Lat::cz[i];
                    constexpr int nz = (a != 0) + (b != 0) + (c != 0);
                    float cpc;
                    if constexpr (nz == 1) {
                        cpc = a != 0 ? Pxx : (b != 0 ? Pyy : Pzz);
                    } else if constexpr (nz == 2) {
                        if constexpr (c == 0) cpc = (a * b == 1) ? sxy + Pxy2 : sxy - Pxy2;
                        else if constexpr (b == 0) cpc = (a * c == 1) ? sxz + Pxz2 : sxz - Pxz2;
                        else cpc = (b * c == 1) ? syz + Pyz2 : syz - Pyz2;
                    } else {
                        float t = sxyz;
                        t = (a * b == 1) ? t + Pxy2 : t - Pxy2;
                        t = (a * c == 1) ? t + Pxz2 : t - Pxz2;
                        t = (b * c == 1) ? t + Pyz2 : t - Pyz2;
                        cpc = t;
                    }
                    const float cu = cdot<i>(u0, u1, u2);
                    const float even = An + k1n * cpc + cu * cdot<i>(Fn0, Fn1, Fn2);
                    const float odd = cdot<i>(Fo0, Fo1, Fo2) + hn * h3<i>(k);
                    f[i] = even + odd;
                    f[i + 1] = even - odd;
                }
            });
            dr_out = dr;
            u[0] = u0;
            u[1] = u1;
            u[2] = u2;
            return;
        }
        const float A = dr - 1.5f * usq - k1 * trP - 3.0f * (1.0f - 0.5f * om) * uF;
        f[0] = Lat::w[0] * A;
        static_for<1, Q>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if constexpr (pair_head<i>()) {
                constexpr int a = Lat::cx[i], b = Lat::cy[i], c = Lat::cz[i];
                constexpr int nz = (a != 0) + (b != 0) + (c != 0);
                const float cu = cdot<i>(u0, u1, u2);
                const float cF = cdot<i>(F0, F1, F2);
                // c . P . c
                float cpc;
                if constexpr (nz == 1) {
                    cpc = a != 0 ? Pxx : (b != 0 ? Pyy : Pzz);
                } else if constexpr (nz == 2) {
                    if constexpr (c == 0) cpc = (a * b == 1) ? sxy + Pxy2 : sxy - Pxy2;
                    else if constexpr (b == 0) cpc = (a * c == 1) ? sxz + Pxz2 : sxz - Pxz2;
                    else cpc = (b * c == 1) ? syz + Pyz2 : syz - Pyz2;
                } else {
                    float t = sxyz;
                    t = (a * b == 1) ? t + Pxy2 : t - Pxy2;
                    t = (a * c == 1) ? t + Pxz2 : t - Pxz2;
                    t = (b * c == 1) ? t + Pyz2 : t - Pyz2;
                    cpc = t;
                }
                const float even = A + cu * (4.5f * cu + kc * cF) + k1 * cpc;
                const float odd = 3.0f * cu + 1.5f * cF + p.t3 * h3<i>(k);
                f[i] = Lat::w[i] * (even + odd);
                f[i + 1] = Lat::w[i] * (even - odd);
            }
        });
    } else {
        // TRT with Guo forcing; psi carries the Galilean diagonal correction D = -(1 - om) Ed
        const float omp = om, omm = p.om_m;
        const float Dd0 = -(1.0f - omp) * Ed0, Dd1 = -(1.0f - omp) * Ed1, Dd2 = -(1.0f - omp) * Ed2;
        const float trD = (Dd0 + Dd1 + Dd2) * (1.0f / 3.0f);
        const float gp = 1.0f - 0.5f * omp, gm = 1.0f - 0.5f * omm;
        const float base = dr - 1.5f * usq;
        {
            // rest direction: f_0 is its own opposite, so fm = 0 and the odd parts vanish
            const float fp = f[0];
            const float eqp = Lat::w[0] * base;
            const float sp = Lat::w[0] * (-3.0f * uF);
            const float psi = Lat::w[0] * 4.5f * (-trD);
            f[0] = fp - omp * (fp - eqp) + gp * sp + psi;
        }
        static_for<1, Q>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if constexpr (pair_head<i>()) {
                constexpr int a = Lat::cx[i], b = Lat::cy[i], c = Lat::cz[i];
                const float w = Lat::w[i];
                const float cv = cdot<i>(u0, u1, u2);
                const float cF = cdot<i>(F0, F1, F2);
                const float fp = 0.5f * (f[i] + f[i + 1]);
                const float fm = 0.5f * (f[i] - f[i + 1]);
                const float eqp = w * (base + 4.5f * cv * cv);
                const float eqm = w * (3.0f * cv + p.t3 * h3<i>(k));
                float dsum = -0.0f;
                if constexpr (a != 0) dsum = dsum + Dd0;
                if constexpr (b != 0) dsum = dsum + Dd1;
                if constexpr (c != 0) dsum = dsum + Dd2;
                const float psi = w * 4.5f * (dsum - trD);
                const float sp = w * (9.0f * cv * cF - 3.0f * uF);
                const float sm = w * 3.0f * cF;
                const float even = fp - omp * (fp - eqp) + gp * sp + psi;
                const float odd = fm - omm * (fm - eqm) + gm * sm;
                f[i] = even + odd;
                f[i + 1] = even - odd;
            }
        });
    }
    dr_out = dr;
    u[0] = u0;
    u[1] = u1;
    u[2] = u2;
}

// Inlet cells: equilibrium at the prescribed velocity ut, with the density deviation of the streamed values
// (lbm.py:205-216). Writes the 27 outputs into f.
__device__ __forceinline__ void inlet_equilibrium(float (&f)[Q], float dr, float ut0, float ut1, float ut2, float t3) {
    const float utsq = ut0 * ut0 + ut1 * ut1 + ut2 * ut2;
    const Cubic k = cubic_terms(ut0, ut1, ut2);
    const float base = dr - 1.5f * utsq;
    f[0] = Lat::w[0] * base;
    static_for<1, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (pair_head<i>()) {
            const float cu = cdot<i>(ut0, ut1, ut2);
            const float even = base + 4.5f * cu * cu;
            const float odd = 3.0f * cu + t3 * h3<i>(k);
            f[i] = Lat::w[i] * (even + odd);
            f[i + 1] = Lat::w[i] * (even - odd);
        }
    });
}

// Sum of the streamed populations (density deviation), for inlet cells that skip the collision.
__device__ __forceinline__ float density_deviation(const float (&f)[Q]) {
    float dr = f[0];
    static_for<1, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (pair_head<i>()) dr = dr + (f[i] + f[i + 1]);
    });
    return dr;
}

}  // namespace cs

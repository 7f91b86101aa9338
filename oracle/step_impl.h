/* ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Scalar CPU restatement of CerebroSim's solver step in the reference's own (Q, N) indexed layout and loop order.
 * It exists to check the HIP product path and to serve as the timed CPU baseline; nothing shipped may call it.
 * This file is included twice by lbm_oracle.c, with REAL = float (suffix _f32) and REAL = double (suffix _f64).
 *
 * Reference followed (file:line of hisham-kadambot/CerebroSim):
 *   lattice ordering, weights, OPP ........ lbm.py:34-50
 *   k_clock ............................... lbm.py:90-102
 *   _ddx .................................. lbm.py:105-126
 *   _h3 ................................... lbm.py:129-137
 *   k_step (stream, moments, Galilean, inlet cells, regularised, hybrid, TRT, outputs) .. lbm.py:140-257
 *   k_flux / k_bc / k_gather .............. simulate.py:65-90
 *
 * Parity status of each function (how it is pinned) is recorded in the module header of oracle/__init__.py.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

/* lbm.py:129-137 — third-order Hermite term of the equilibrium divided by w_i. */
static inline REAL FN(orc_h3)(REAL c0, REAL c1, REAL c2, REAL u0, REAL u1, REAL u2) {
    const REAL t = (REAL)1.0 / (REAL)3.0;
    return (REAL)13.5 * ((c0 * c0 - t) * (c1 * u0 * u0 * u1 + c2 * u0 * u0 * u2)
                         + (c1 * c1 - t) * (c0 * u0 * u1 * u1 + c2 * u1 * u1 * u2)
                         + (c2 * c2 - t) * (c0 * u0 * u2 * u2 + c1 * u1 * u2 * u2))
           + (REAL)27.0 * c0 * c1 * c2 * u0 * u1 * u2;
}

/* lbm.py:105-126 — non-uniform central difference of the previous step's velocity along one axis.
 * i_minus pulls from x - e_a, i_plus from x + e_a. A wall neighbour has u = 0 at distance max(q, 0.1);
 * an opening (inlet/outlet code) counts as zero gradient (v = u0, h = 1). */
static inline void FN(orc_ddx)(const int32_t* nbr, const REAL* qw, const REAL* vel, int64_t N, int64_t n,
                               int i_minus, int i_plus, const REAL u0[3], REAL out[3]) {
    REAL h1 = (REAL)1.0, v1[3] = {u0[0], u0[1], u0[2]};
    int32_t s1 = nbr[(int64_t)i_minus * N + n];
    if (s1 >= 0) {
        v1[0] = vel[3 * (int64_t)s1]; v1[1] = vel[3 * (int64_t)s1 + 1]; v1[2] = vel[3 * (int64_t)s1 + 2];
    } else if (s1 == ORC_WALL) {
        REAL q = qw[(int64_t)i_minus * N + n];
        h1 = q > (REAL)0.1 ? q : (REAL)0.1;
        v1[0] = v1[1] = v1[2] = (REAL)0.0;
    }
    REAL h2 = (REAL)1.0, v2[3] = {u0[0], u0[1], u0[2]};
    int32_t s2 = nbr[(int64_t)i_plus * N + n];
    if (s2 >= 0) {
        v2[0] = vel[3 * (int64_t)s2]; v2[1] = vel[3 * (int64_t)s2 + 1]; v2[2] = vel[3 * (int64_t)s2 + 2];
    } else if (s2 == ORC_WALL) {
        REAL q = qw[(int64_t)i_plus * N + n];
        h2 = q > (REAL)0.1 ? q : (REAL)0.1;
        v2[0] = v2[1] = v2[2] = (REAL)0.0;
    }
    REAL den = h1 * h2 * (h1 + h2);
    for (int a = 0; a < 3; ++a) out[a] = (h1 * h1 * (v2[a] - u0[a]) + h2 * h2 * (u0[a] - v1[a])) / den;
}

/* lbm.py:90-102 — drive(t) Fourier series with the cosine start-up ramp; advances the tick. */
void FN(orc_clock)(int32_t* tick, const REAL* coef, int nh, int period, int ramp, REAL* drive) {
    int32_t t = tick[0];
    const REAL pi = (REAL)3.14159265358979323846;
    REAL ph = (REAL)2.0 * pi * (REAL)(t % period) / (REAL)period;
    REAL d = coef[0];
    for (int k = 1; k <= nh; ++k)
        d += coef[2 * k - 1] * ORC_COS((REAL)k * ph) + coef[2 * k] * ORC_SIN((REAL)k * ph);
    if (t < ramp) d *= (REAL)0.5 - (REAL)0.5 * ORC_COS(pi * (REAL)t / (REAL)ramp);
    drive[0] = d;
    tick[0] = t + 1;
}

/* lbm.py:140-257 — one solver step for cells [n0, n1). Arrays in the reference layout:
 * f_in/f_out (Q, N) row-major, nbr/qw (Q, N), uin/vel_old/vel (N, 3), inl/om_cell/rho (N). */
static void FN(orc_step_range)(const REAL* f_in, REAL* f_out, const int32_t* nbr, const REAL* qw, const REAL* uin,
                               const int32_t* inl, const REAL* p_out, REAL drive, const REAL* force,
                               const REAL* om_cell, REAL magic, int bouzidi, int model, REAL sigma, int gal,
                               const REAL* vel_old, REAL* vel, REAL* rho, int64_t N, int64_t n0, int64_t n1) {
    REAL Wr[ORC_Q];
    for (int i = 0; i < ORC_Q; ++i) Wr[i] = (REAL)ORC_W[i];
    for (int64_t n = n0; n < n1; ++n) {
        REAL om_p = om_cell[n];
        REAL om_m = (REAL)1.0 / ((REAL)0.5 + magic / ((REAL)1.0 / om_p - (REAL)0.5));
        REAL d = drive;
        REAL F[3] = {force[0] * d, force[1] * d, force[2] * d};
        REAL up[3] = {vel_old[3 * n], vel_old[3 * n + 1], vel_old[3 * n + 2]};
        REAL upsq = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
        REAL f[ORC_Q];
        /* pull streaming with boundary links (lbm.py:156-182) */
        for (int i = 0; i < ORC_Q; ++i) {
            int32_t src = nbr[(int64_t)i * N + n];
            int io = ORC_OPP[i];
            REAL cx = (REAL)ORC_C[i][0], cy = (REAL)ORC_C[i][1], cz = (REAL)ORC_C[i][2];
            REAL v = (REAL)0.0;
            if (src >= 0) {
                v = f_in[(int64_t)i * N + src];
            } else if (src == ORC_WALL) {
                REAL fo = f_in[(int64_t)io * N + n];
                v = fo;
                if (bouzidi != 0) {
                    REAL q = qw[(int64_t)i * N + n];
                    if (q < (REAL)0.5) {
                        int32_t m = nbr[(int64_t)io * N + n];
                        if (m >= 0) v = (REAL)2.0 * q * fo + ((REAL)1.0 - (REAL)2.0 * q) * f_in[(int64_t)io * N + m];
                    } else {
                        v = fo / ((REAL)2.0 * q) + ((REAL)2.0 * q - (REAL)1.0) / ((REAL)2.0 * q) * f_in[(int64_t)i * N + n];
                    }
                }
            } else if (src == ORC_INLET) {
                REAL uw0 = uin[3 * n] * d, uw1 = uin[3 * n + 1] * d, uw2 = uin[3 * n + 2] * d;
                REAL cu = cx * uw0 + cy * uw1 + cz * uw2;
                v = f_in[(int64_t)io * N + n] + (REAL)6.0 * Wr[i] * cu;
            } else {
                int k = ORC_OUTLET0 - src;
                REAL cu = cx * up[0] + cy * up[1] + cz * up[2];
                v = -f_in[(int64_t)io * N + n]
                    + (REAL)2.0 * Wr[i] * ((REAL)3.0 * p_out[k] + (REAL)4.5 * cu * cu - (REAL)1.5 * upsq);
            }
            f[i] = v;
        }
        /* moments, incompressible (lbm.py:183-191) */
        REAL dr = (REAL)0.0, j0 = (REAL)0.0, j1 = (REAL)0.0, j2 = (REAL)0.0;
        for (int i = 0; i < ORC_Q; ++i) {
            dr += f[i];
            j0 += (REAL)ORC_C[i][0] * f[i];
            j1 += (REAL)ORC_C[i][1] * f[i];
            j2 += (REAL)ORC_C[i][2] * f[i];
        }
        REAL u[3] = {j0 + F[0] * (REAL)0.5, j1 + F[1] * (REAL)0.5, j2 + F[2] * (REAL)0.5};
        REAL usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        REAL uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
        /* Galilean-invariance correction (lbm.py:192-204) */
        REAL t3 = (REAL)0.0, Ed[3] = {(REAL)0.0, (REAL)0.0, (REAL)0.0};
        REAL gx[3], gy[3], gz[3];
        if (gal != 0) {
            t3 = (REAL)1.0;
            FN(orc_ddx)(nbr, qw, vel_old, N, n, 1, 2, up, gx);
            FN(orc_ddx)(nbr, qw, vel_old, N, n, 3, 4, up, gy);
            FN(orc_ddx)(nbr, qw, vel_old, N, n, 5, 6, up, gz);
            REAL tp = (REAL)1.0 / om_p;
            Ed[0] = (REAL)3.0 * tp * u[0] * u[0] * gx[0];
            Ed[1] = (REAL)3.0 * tp * u[1] * u[1] * gy[1];
            Ed[2] = (REAL)3.0 * tp * u[2] * u[2] * gz[2];
        }
        REAL Dd[3] = {-((REAL)1.0 - om_p) * Ed[0], -((REAL)1.0 - om_p) * Ed[1], -((REAL)1.0 - om_p) * Ed[2]};
        REAL trD = (Dd[0] + Dd[1] + Dd[2]) / (REAL)3.0;
        /* inlet cells: equilibrium at the prescribed velocity (lbm.py:205-216) */
        if (inl[n] != 0) {
            REAL ut[3] = {uin[3 * n] * d, uin[3 * n + 1] * d, uin[3 * n + 2] * d};
            REAL utsq = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
            for (int i = 0; i < ORC_Q; ++i) {
                REAL cx = (REAL)ORC_C[i][0], cy = (REAL)ORC_C[i][1], cz = (REAL)ORC_C[i][2];
                REAL c0u = cx * ut[0] + cy * ut[1] + cz * ut[2];
                f_out[(int64_t)i * N + n] = Wr[i] * (dr + (REAL)3.0 * c0u + (REAL)4.5 * c0u * c0u - (REAL)1.5 * utsq
                                                     + t3 * FN(orc_h3)(cx, cy, cz, ut[0], ut[1], ut[2]));
            }
            vel[3 * n] = ut[0]; vel[3 * n + 1] = ut[1]; vel[3 * n + 2] = ut[2];
            rho[n] = dr;
            continue;
        }
        if (model == 1) {
            /* regularised BGK (lbm.py:217-240) */
            REAL P[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int i = 0; i < ORC_Q; ++i) {
                REAL c[3] = {(REAL)ORC_C[i][0], (REAL)ORC_C[i][1], (REAL)ORC_C[i][2]};
                REAL c1u = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
                REAL ne = f[i] - Wr[i] * (dr + (REAL)3.0 * c1u + (REAL)4.5 * c1u * c1u - (REAL)1.5 * usq);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) P[a][b] += c[a] * c[b] * ne;
            }
            if (sigma < (REAL)1.0) {
                /* hybrid blend with the finite-difference stress (lbm.py:224-229) */
                REAL G[3][3];
                FN(orc_ddx)(nbr, qw, vel_old, N, n, 1, 2, up, G[0]);
                FN(orc_ddx)(nbr, qw, vel_old, N, n, 3, 4, up, G[1]);
                FN(orc_ddx)(nbr, qw, vel_old, N, n, 5, 6, up, G[2]);
                REAL k = ((REAL)1.0 - sigma) / ((REAL)3.0 * om_p);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) P[a][b] = sigma * P[a][b] - k * (G[a][b] + G[b][a]);
            }
            P[0][0] -= Ed[0]; P[1][1] -= Ed[1]; P[2][2] -= Ed[2];
            REAL trP = (P[0][0] + P[1][1] + P[2][2]) / (REAL)3.0;
            for (int i = 0; i < ORC_Q; ++i) {
                REAL c[3] = {(REAL)ORC_C[i][0], (REAL)ORC_C[i][1], (REAL)ORC_C[i][2]};
                REAL c2u = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
                REAL c2F = c[0] * F[0] + c[1] * F[1] + c[2] * F[2];
                REAL feq = Wr[i] * (dr + (REAL)3.0 * c2u + (REAL)4.5 * c2u * c2u - (REAL)1.5 * usq
                                    + t3 * FN(orc_h3)(c[0], c[1], c[2], u[0], u[1], u[2]));
                REAL Pc[3];
                for (int a = 0; a < 3; ++a) Pc[a] = P[a][0] * c[0] + P[a][1] * c[1] + P[a][2] * c[2];
                REAL cPc = c[0] * Pc[0] + c[1] * Pc[1] + c[2] * Pc[2];
                REAL fneq = Wr[i] * ((REAL)4.5 * (cPc - trP) - (REAL)1.5 * c2F);
                REAL src2 = Wr[i] * ((REAL)3.0 * c2F + (REAL)9.0 * c2u * c2F - (REAL)3.0 * uF);
                f_out[(int64_t)i * N + n] = feq + ((REAL)1.0 - om_p) * fneq + ((REAL)1.0 - (REAL)0.5 * om_p) * src2;
            }
        } else {
            /* TRT with Guo forcing (lbm.py:241-255) */
            for (int i = 0; i < ORC_Q; ++i) {
                int jo = ORC_OPP[i];
                REAL c[3] = {(REAL)ORC_C[i][0], (REAL)ORC_C[i][1], (REAL)ORC_C[i][2]};
                REAL cv = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
                REAL cF = c[0] * F[0] + c[1] * F[1] + c[2] * F[2];
                REAL eqp = Wr[i] * (dr + (REAL)4.5 * cv * cv - (REAL)1.5 * usq);
                REAL eqm = Wr[i] * ((REAL)3.0 * cv + t3 * FN(orc_h3)(c[0], c[1], c[2], u[0], u[1], u[2]));
                REAL psi = Wr[i] * (REAL)4.5 * (c[0] * c[0] * Dd[0] + c[1] * c[1] * Dd[1] + c[2] * c[2] * Dd[2] - trD);
                REAL fp = (REAL)0.5 * (f[i] + f[jo]);
                REAL fm = (REAL)0.5 * (f[i] - f[jo]);
                REAL sp = Wr[i] * ((REAL)9.0 * cv * cF - (REAL)3.0 * uF);
                REAL sm = Wr[i] * (REAL)3.0 * cF;
                f_out[(int64_t)i * N + n] = f[i] - om_p * (fp - eqp) - om_m * (fm - eqm)
                                            + ((REAL)1.0 - (REAL)0.5 * om_p) * sp + ((REAL)1.0 - (REAL)0.5 * om_m) * sm + psi;
            }
        }
        vel[3 * n] = u[0]; vel[3 * n + 1] = u[1]; vel[3 * n + 2] = u[2];
        rho[n] = dr;
    }
}

void FN(orc_step)(const REAL* f_in, REAL* f_out, const int32_t* nbr, const REAL* qw, const REAL* uin,
                  const int32_t* inl, const REAL* p_out, REAL drive, const REAL* force, const REAL* om_cell,
                  REAL magic, int bouzidi, int model, REAL sigma, int gal, const REAL* vel_old, REAL* vel,
                  REAL* rho, int64_t N, int nthreads) {
    if (nthreads <= 1) {
        FN(orc_step_range)(f_in, f_out, nbr, qw, uin, inl, p_out, drive, force, om_cell, magic, bouzidi, model,
                           sigma, gal, vel_old, vel, rho, N, 0, N);
        return;
    }
    int64_t chunk = (N + 4 * (int64_t)nthreads - 1) / (4 * (int64_t)nthreads);
    if (chunk < 32) chunk = 32;
    if (chunk > 4096) chunk = 4096;
    int64_t nchunks = (N + chunk - 1) / chunk;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t c = 0; c < nchunks; ++c) {
        int64_t a = c * chunk, b = a + chunk < N ? a + chunk : N;
        FN(orc_step_range)(f_in, f_out, nbr, qw, uin, inl, p_out, drive, force, om_cell, magic, bouzidi, model,
                           sigma, gal, vel_old, vel, rho, N, a, b);
    }
}

/* simulate.py:65-82 — per-step flow through the slabs (k_flux), then the boundary update (k_bc).
 * Summation is sequential in list order (the reference's float atomics have no defined order). */
void FN(orc_flux_bc)(const int32_t* idx, const REAL* w, const int32_t* slot, int64_t M, const REAL* vel,
                     REAL* Qacc, const REAL* R, REAL pref, const REAL* drive, int n_open, REAL alpha, REAL* p_out,
                     REAL* q_last) {
    for (int64_t t = 0; t < M; ++t) {
        int64_t c = idx[t];
        Qacc[slot[t]] += vel[3 * c] * w[3 * t] + vel[3 * c + 1] * w[3 * t + 1] + vel[3 * c + 2] * w[3 * t + 2];
    }
    p_out[0] = drive[0];
    for (int k = 0; k < n_open; ++k) {
        q_last[k] = Qacc[k];
        if (k > 0) p_out[k] = p_out[k] + alpha * (R[k] * Qacc[k] - pref - p_out[k]);
        Qacc[k] = (REAL)0.0;
    }
}

/* simulate.py:85-90 — probe sums of rho/3 and the maximum speed over the probe cells. */
void FN(orc_gather)(const int32_t* idx, const int32_t* slot, int64_t M, const REAL* rho, const REAL* vel, REAL* out,
                    REAL* umax) {
    for (int64_t t = 0; t < M; ++t) {
        int64_t c = idx[t];
        out[slot[t]] += rho[c] / (REAL)3.0;
        REAL s = ORC_SQRT(vel[3 * c] * vel[3 * c] + vel[3 * c + 1] * vel[3 * c + 1] + vel[3 * c + 2] * vel[3 * c + 2]);
        if (s > umax[0]) umax[0] = s;
    }
}

/* lbm.py:422-448 semantics without graphs: `steps` x (k_clock then k_step), ping-pong of f and vel between the
 * two buffers, starting from buffer `cur`; returns the buffer index that holds the newest state. */
int FN(orc_run)(REAL* f0, REAL* f1, const int32_t* nbr, const REAL* qw, const REAL* uin, const int32_t* inl,
                const REAL* p_out, int32_t* tick, const REAL* coef, int nh, int period, int ramp, REAL* drive,
                const REAL* force, const REAL* om_cell, REAL magic, int bouzidi, int model, REAL sigma, int gal,
                REAL* v0, REAL* v1, REAL* rho, int64_t N, int cur, int64_t steps, int nthreads) {
    REAL* fb[2] = {f0, f1};
    REAL* vb[2] = {v0, v1};
    for (int64_t s = 0; s < steps; ++s) {
        FN(orc_clock)(tick, coef, nh, period, ramp, drive);
        FN(orc_step)(fb[cur], fb[1 - cur], nbr, qw, uin, inl, p_out, drive[0], force, om_cell, magic, bouzidi, model,
                     sigma, gal, vb[cur], vb[1 - cur], rho, N, nthreads);
        cur = 1 - cur;
    }
    return cur;
}

#undef FN
#undef CAT
#undef CAT2

// Host-side lattice builder: the sparse fluid-cell lattice of lbm.py:264-330 (build_lattice) from a signed distance
// sampled on grid nodes, in native code (threads over cells) so that lattices of tens of millions of cells build in
// seconds. The output is the reference's contract: coords in C order of the grid (argwhere), nbr (Q, N) source cell or
// link code (WALL -1, INLET -2, OUTLET k -3-k), q (Q, N) wall fraction (0.5 where unused), uin (N, 3) parabolic inlet
// profile (unit mean, pointing inward; the last inlet wins), and the link counts.
//
// Arithmetic follows the numpy builder of cerebrosim_amd/lattice.py step for step: opening tests in double
// (along >= -1e-6, lateral < radius + slack dx, inlets before outlets, first match wins); q = -phi_here /
// (phi_solid - phi_here) computed in FP64 from the FP32 distances (the reference's arithmetic is FP32; FP64 keeps q
// correctly rounded), clipped to [1e-3, 1], 0.5 when not finite (solid node off the grid), stored as FP32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <sched.h>
#include <thread>
#include <vector>

#include "common.hpp"
#include "d3q27.hpp"

namespace cs {
namespace {

struct LatticeHost {
    int64_t N = 0;
    std::vector<int32_t> coords;
    std::unique_ptr<int32_t[]> nbr;   // (Q, N), every entry written by the link pass
    std::unique_ptr<float[]> q;       // (Q, N), every entry written by the link pass
    std::vector<float> uin;
    std::vector<int64_t> counts;  // wall, inlet, outlet 0..n_out-1
};

// CPUs this process may run on (its affinity mask, which a container or a job's CPU share narrows), not the
// machine's count.
int cpus_available() {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
    return (int)std::max(1u, std::thread::hardware_concurrency());
}

template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
    if (threads <= 1 || n < 4096) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([&f, lo, hi, t] { f(lo, hi, t); });
    }
    for (auto& th : pool) th.join();
}

// Opening index the missing link belongs to (-1: wall), testing inlets then outlets in their given order.
int classify(const double p[3], const cs_opening* ops, int n_ops, double dx, double slack, int& kind_out,
             const std::vector<int>& outlet_rank) {
    for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k < n_ops; ++k) {
            const cs_opening& o = ops[k];
            if (o.kind != pass) continue;
            const double r0 = p[0] - o.centre[0], r1 = p[1] - o.centre[1], r2 = p[2] - o.centre[2];
            const double along = r0 * o.normal[0] + r1 * o.normal[1] + r2 * o.normal[2];
            const double l0 = r0 - along * o.normal[0], l1 = r1 - along * o.normal[1], l2 = r2 - along * o.normal[2];
            const double lateral = std::sqrt(l0 * l0 + l1 * l1 + l2 * l2);
            if (along >= -1e-6 && lateral < o.radius + slack * dx) {
                kind_out = pass;
                return pass == 0 ? k : outlet_rank[k];
            }
        }
    }
    kind_out = -1;
    return -1;
}

}  // namespace
}  // namespace cs

extern "C" int cs_lattice_build(const cs_grid_in* in, void** handle) {
    CS_REQUIRE(in && handle && in->phi, "cs_lattice_build: null argument");
    CS_REQUIRE(in->shape[0] > 0 && in->shape[1] > 0 && in->shape[2] > 0, "cs_lattice_build: shape");
    CS_REQUIRE(in->n_openings >= 0 && (in->n_openings == 0 || in->openings), "cs_lattice_build: openings");
    const int64_t nx = in->shape[0], ny = in->shape[1], nz = in->shape[2];
    const int64_t G = nx * ny * nz;
    int threads = in->threads > 0 ? in->threads : cs::cpus_available();
    auto* L = new (std::nothrow) cs::LatticeHost();
    CS_REQUIRE(L, "cs_lattice_build: out of memory");
    try {
        // outlet k = rank among the outlets, in the given order
        std::vector<int> outlet_rank(in->n_openings, -1);
        int n_out = 0;
        for (int k = 0; k < in->n_openings; ++k)
            if (in->openings[k].kind == 1) outlet_rank[k] = n_out++;
        // fluid nodes in C order (x slowest): count per x-slab, then fill
        std::vector<int64_t> slab(nx + 1, 0);
        cs::parallel_for(nx, threads, [&](int64_t lo, int64_t hi, int) {
            for (int64_t i = lo; i < hi; ++i) {
                int64_t c = 0;
                const float* p = in->phi + i * ny * nz;
                for (int64_t t = 0; t < ny * nz; ++t) c += p[t] < 0.0f;
                slab[i + 1] = c;
            }
        });
        for (int64_t i = 0; i < nx; ++i) slab[i + 1] += slab[i];
        const int64_t N = slab[nx];
        CS_REQUIRE(N < ((int64_t)1 << 31), "cs_lattice_build: %lld fluid cells exceed int32 ids", (long long)N);
        L->N = N;
        L->coords.resize(3 * N);
        std::vector<int32_t> ids(G, -1);
        cs::parallel_for(nx, threads, [&](int64_t lo, int64_t hi, int) {
            for (int64_t i = lo; i < hi; ++i) {
                int64_t n = slab[i];
                for (int64_t j = 0; j < ny; ++j)
                    for (int64_t k = 0; k < nz; ++k) {
                        const int64_t g = (i * ny + j) * nz + k;
                        if (in->phi[g] < 0.0f) {
                            ids[g] = (int32_t)n;
                            L->coords[3 * n] = (int32_t)i;
                            L->coords[3 * n + 1] = (int32_t)j;
                            L->coords[3 * n + 2] = (int32_t)k;
                            ++n;
                        }
                    }
            }
        });
        L->nbr.reset(new int32_t[(size_t)cs::Q * N]);
        L->q.reset(new float[(size_t)cs::Q * N]);
        L->uin.assign(3 * N, 0.0f);
        std::vector<std::vector<int64_t>> cnt(threads, std::vector<int64_t>(2 + n_out, 0));
        std::vector<uint8_t> has_inlet(N, 0);
        const int64_t shp[3] = {nx, ny, nz};
        const bool any_periodic = in->periodic[0] || in->periodic[1] || in->periodic[2];
        int64_t goff[cs::Q];  // grid-index offset of the source node x - c_i
        for (int i = 0; i < cs::Q; ++i) goff[i] = (cs::Lat::cx[i] * ny + cs::Lat::cy[i]) * nz + cs::Lat::cz[i];
        cs::parallel_for(N, threads, [&](int64_t lo, int64_t hi, int t) {
            std::vector<int64_t>& c = cnt[t];
            for (int64_t n = lo; n < hi; ++n) {
                const int64_t x[3] = {L->coords[3 * n], L->coords[3 * n + 1], L->coords[3 * n + 2]};
                const int64_t g0 = (x[0] * ny + x[1]) * nz + x[2];
                const float ph = in->phi[g0];
                // a node whose 27 sources all lie on the grid without wrapping: sources by a fixed offset
                const bool inner = !any_periodic && x[0] > 0 && x[0] < nx - 1 && x[1] > 0 && x[1] < ny - 1 &&
                                   x[2] > 0 && x[2] < nz - 1;
                for (int i = 0; i < cs::Q; ++i) {
                    int64_t s[3];
                    bool on_grid = true;
                    int64_t g;
                    const int ci[3] = {cs::Lat::cx[i], cs::Lat::cy[i], cs::Lat::cz[i]};
                    if (inner) {
                        g = g0 - goff[i];
                        const int32_t src = ids[g];
                        if (src >= 0) {
                            L->nbr[(size_t)i * N + n] = src;
                            L->q[(size_t)i * N + n] = 0.5f;
                            continue;
                        }
                        for (int a = 0; a < 3; ++a) s[a] = x[a] - ci[a];
                    } else {
                        for (int a = 0; a < 3; ++a) {
                            s[a] = x[a] - ci[a];
                            if (in->periodic[a]) s[a] = ((s[a] % shp[a]) + shp[a]) % shp[a];
                            else if (s[a] < 0 || s[a] >= shp[a]) on_grid = false;
                        }
                        g = on_grid ? (s[0] * ny + s[1]) * nz + s[2] : -1;
                        const int32_t src = on_grid ? ids[g] : -1;
                        if (src >= 0) {
                            L->nbr[(size_t)i * N + n] = src;
                            L->q[(size_t)i * N + n] = 0.5f;
                            continue;
                        }
                    }
                    const double p[3] = {in->origin[0] + (double)s[0] * in->dx, in->origin[1] + (double)s[1] * in->dx,
                                         in->origin[2] + (double)s[2] * in->dx};
                    int kind;
                    const int k = cs::classify(p, in->openings, in->n_openings, in->dx, in->slack, kind, outlet_rank);
                    L->q[(size_t)i * N + n] = 0.5f;
                    if (kind == 0) {
                        L->nbr[(size_t)i * N + n] = CS_INLET;
                        ++c[1];
                        has_inlet[n] = 1;
                    } else if (kind == 1) {
                        L->nbr[(size_t)i * N + n] = CS_OUTLET0 - k;
                        ++c[2 + k];
                    } else {
                        L->nbr[(size_t)i * N + n] = CS_WALL;
                        // the wall fraction in FP64 from the FP32 distances, clipped, then stored as FP32
                        const double phs = on_grid ? (double)in->phi[g] : NAN;
                        const double qv = -(double)ph / (phs - (double)ph);
                        L->q[(size_t)i * N + n] =
                            std::isfinite(qv) ? (float)std::min(std::max(qv, 1e-3), 1.0) : 0.5f;
                        ++c[0];
                    }
                }
            }
        });
        L->counts.assign(2 + n_out, 0);
        for (auto& c : cnt)
            for (int k = 0; k < 2 + n_out; ++k) L->counts[k] += c[k];
        // inlet profile: every inlet writes all inlet cells, in order (the last one wins)
        for (int k = 0; k < in->n_openings; ++k) {
            const cs_opening& o = in->openings[k];
            if (o.kind != 0) continue;
            cs::parallel_for(N, threads, [&](int64_t lo, int64_t hi, int) {
                for (int64_t n = lo; n < hi; ++n) {
                    if (!has_inlet[n]) continue;
                    double r[3];
                    for (int a = 0; a < 3; ++a)
                        r[a] = in->origin[a] + (double)L->coords[3 * n + a] * in->dx - o.centre[a];
                    const double along = r[0] * o.normal[0] + r[1] * o.normal[1] + r[2] * o.normal[2];
                    const double l0 = r[0] - along * o.normal[0], l1 = r[1] - along * o.normal[1],
                                 l2 = r[2] - along * o.normal[2];
                    const double lat = std::sqrt(l0 * l0 + l1 * l1 + l2 * l2);
                    const double t = lat / o.radius;
                    const double prof = std::max(2.0 * (1.0 - t * t), 0.0);
                    for (int a = 0; a < 3; ++a) L->uin[3 * n + a] = (float)(-prof * o.normal[a]);
                }
            });
        }
    } catch (const std::bad_alloc&) {
        delete L;
        cs::set_error("cs_lattice_build: out of memory");
        return CS_E_ARG;
    }
    *handle = L;
    return CS_OK;
}

extern "C" int cs_lattice_sizes(const void* handle, int64_t* n_cells, int64_t* counts, int32_t max_counts) {
    CS_REQUIRE(handle && n_cells, "cs_lattice_sizes: null argument");
    const auto* L = static_cast<const cs::LatticeHost*>(handle);
    *n_cells = L->N;
    if (counts)
        for (int k = 0; k < max_counts && k < (int)L->counts.size(); ++k) counts[k] = L->counts[k];
    return (int)L->counts.size();
}

extern "C" int cs_lattice_copy(const void* handle, const char* name, void* dst) {
    CS_REQUIRE(handle && name && dst, "cs_lattice_copy: null argument");
    const auto* L = static_cast<const cs::LatticeHost*>(handle);
    const std::string s(name);
    const size_t qn = (size_t)cs::Q * L->N;
    if (s == "coords") std::memcpy(dst, L->coords.data(), L->coords.size() * 4);
    else if (s == "nbr") std::memcpy(dst, L->nbr.get(), qn * 4);
    else if (s == "q") std::memcpy(dst, L->q.get(), qn * 4);
    else if (s == "uin") std::memcpy(dst, L->uin.data(), L->uin.size() * 4);
    else {
        cs::set_error("cs_lattice_copy: unknown array '%s'", name);
        return CS_E_ARG;
    }
    return CS_OK;
}

extern "C" int cs_lattice_view(const void* handle, const char* name, void** data) {
    CS_REQUIRE(handle && name && data, "cs_lattice_view: null argument");
    auto* L = static_cast<cs::LatticeHost*>(const_cast<void*>(handle));
    const std::string s(name);
    if (s == "coords") *data = L->coords.data();
    else if (s == "nbr") *data = L->nbr.get();
    else if (s == "q") *data = L->q.get();
    else if (s == "uin") *data = L->uin.data();
    else {
        cs::set_error("cs_lattice_view: unknown array '%s'", name);
        return CS_E_ARG;
    }
    return CS_OK;
}

extern "C" int cs_lattice_free(void* handle) {
    delete static_cast<cs::LatticeHost*>(handle);
    return CS_OK;
}

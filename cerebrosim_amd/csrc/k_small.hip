// Small per-step kernels: the drive clock (lbm.py:90-102), the slab-flow reduction with the opening-pressure
// update (simulate.py:65-82), the probe reduction (simulate.py:85-90) and the finite/mass reduction behind ok()
// (lbm.py:454-455). The reductions are deterministic: each thread sums a fixed strided subset in order, then a
// fixed-shape tree in LDS combines the partial sums (the reference's float atomics have no defined order).
#include <cmath>
#include <mutex>

#include "common.hpp"

namespace cs {
namespace {

// ---- k_clock ------------------------------------------------------------------------------------------------
__device__ __forceinline__ float clock_value(int t, const float* coef, int nh, int period, int ramp) {
    const float pi = 3.14159265358979323846f;
    const float ph = 2.0f * pi * (float)(t % period) / (float)period;
    float d = coef[0];
    for (int k = 1; k <= nh; ++k) d += coef[2 * k - 1] * cosf((float)k * ph) + coef[2 * k] * sinf((float)k * ph);
    if (t < ramp) d *= 0.5f - 0.5f * cosf(pi * (float)t / (float)ramp);
    return d;
}

__device__ __forceinline__ void clock_tick(int32_t* tick, const float* coef, int nh, int period, int ramp,
                                           float* drive) {
    const int t = tick[0];
    drive[0] = clock_value(t, coef, nh, period, ramp);
    tick[0] = t + 1;
}

// The drives of n_steps consecutive steps of one case per workgroup: drive[k * stride + c] = k_clock at tick + k,
// then tick += n_steps (read by every thread before the barrier, advanced after it).
__global__ __launch_bounds__(256) void k_clock_table(cs_clock_table_args a) {
    const int c = blockIdx.x;
    if (a.period[c] <= 0) return;  // workgroup-uniform
    const int t0 = a.tick[c];
    __syncthreads();
    const float* coef = a.coef + (int64_t)c * a.coef_stride;
    for (int k = threadIdx.x; k < a.n_steps; k += 256)
        a.drive[(int64_t)k * a.drive_stride + c] = clock_value(t0 + k, coef, a.nh[c], a.period[c], a.ramp[c]);
    if (threadIdx.x == 0) a.tick[c] = t0 + a.n_steps;
}

__global__ void k_clock(cs_clock_args a) {
    if (threadIdx.x != 0) return;
    clock_tick(a.tick, a.coef, a.nh, a.period, a.ramp, a.drive);
}

// one thread per case of a packed set of lattices
__global__ void k_clock_cases(cs_clock_cases_args a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_cases || a.period[c] <= 0) return;
    clock_tick(a.tick + c, a.coef + (int64_t)c * a.coef_stride, a.nh[c], a.period[c], a.ramp[c], a.drive + c);
}

// ---- bandwidth reference: 16-byte streaming copy -----------------------------------------------------------
// One 16-byte element per thread (a grid of n / 256 workgroups, capped at 2^24; a larger copy loops). On gfx950 this
// measured faster than grid-stride loops over fewer workgroups with several loads in flight per thread (6.2 against
// 4.6-5.5 TB/s for 1 GiB).
constexpr int64_t kCopyMaxGroups = (int64_t)1 << 24;

__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---- frame binning: one thread per bin, FP64 means in cell order, float16 outputs -----------------------------
__global__ __launch_bounds__(256) void k_bin_fields(cs_bin_args a) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= a.n_bins) return;
    const int32_t s0 = a.start[b], s1 = a.start[b + 1];
    double r = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
    for (int32_t k = s0; k < s1; ++k) {
        const int64_t s = a.slots[k];
        r += (double)a.rho[s * a.rho_elem];
        const float* v = a.vel + s * a.vel_elem;
        u0 += (double)v[0];
        u1 += (double)v[a.vel_comp];
        u2 += (double)v[2 * a.vel_comp];
    }
    const double inv = 1.0 / (double)(s1 - s0);
    reinterpret_cast<_Float16*>(a.p_out)[b] = (_Float16)(a.p_scale * (r * inv) + a.p_offset);
    _Float16* u = reinterpret_cast<_Float16*>(a.u_out) + 3 * b;
    u[0] = (_Float16)(a.u_scale * (u0 * inv));
    u[1] = (_Float16)(a.u_scale * (u1 * inv));
    u[2] = (_Float16)(a.u_scale * (u2 * inv));
}

// ---- block-wide sum in a fixed order (1024 threads = 16 waves of 64) ---------------------------------------
constexpr int kRedThreads = 1024;

template <typename T>
__device__ T block_sum(T v, T* lds) {
    // wave-level butterfly (fixed pattern), then the 16 wave totals summed in order by wave 0
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    T total = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kRedThreads / 64; ++k) total += lds[k];
    __syncthreads();
    return total;   // valid in thread 0
}

__device__ float block_max(float v, float* lds) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    float m = 0.0f;
    if (threadIdx.x == 0)
        for (int k = 0; k < kRedThreads / 64; ++k) m = fmaxf(m, lds[k]);
    __syncthreads();
    return m;
}

// ---- k_flux + k_bc: chunked partial sums, then one wave per lattice -------------------------------------------
constexpr int kMaxOpen = 64;
constexpr int kFluxGroup = 8;   // openings summed per pass over the slab cells

// Chunking of a lattice's m slab cells, a function of m alone (so a lattice gives the same bits alone or packed):
// each thread of a 1024-thread workgroup takes flux_per_thread(m) cells of the chunk, at most 256 chunks.
__device__ __forceinline__ int64_t flux_per_thread(int64_t m) {
    const int64_t per = (m + (int64_t)kRedThreads * 256 - 1) / ((int64_t)kRedThreads * 256);
    return per < 1 ? 1 : per;
}
__device__ __forceinline__ int64_t flux_chunks(int64_t m) {
    const int64_t ch = (int64_t)kRedThreads * flux_per_thread(m);
    return (m + ch - 1) / ch;
}

// Phase 1: chunk b of one lattice's slab cells; work[k * wc + b] = sum over the chunk's cells of opening k of
// vel . w (the product in FP32 from the FP32 inputs, accumulated in FP64: per thread in cell order, then the fixed
// block tree).
__device__ __forceinline__ void flux_partial(const int32_t* idx, const float* w, const int32_t* slot, int64_t m,
                                             const float* vel, int64_t vel_elem, int64_t vel_comp, int n_open,
                                             double* work, int wc, int64_t b, double* lds) {
    const int64_t per = flux_per_thread(m), chunk = (int64_t)kRedThreads * per;
    for (int k0 = 0; k0 < n_open; k0 += kFluxGroup) {
        double acc[kFluxGroup];
#pragma unroll
        for (int q = 0; q < kFluxGroup; ++q) acc[q] = 0.0;
        for (int64_t j = 0; j < per; ++j) {
            const int64_t t = b * chunk + j * kRedThreads + threadIdx.x;
            if (t >= m) break;
            const int kk = slot[t] - k0;
            if (kk < 0 || kk >= kFluxGroup) continue;
            const float* v = vel + (int64_t)idx[t] * vel_elem;
            const double val = (double)(v[0] * w[3 * t] + v[vel_comp] * w[3 * t + 1] + v[2 * vel_comp] * w[3 * t + 2]);
#pragma unroll
            for (int q = 0; q < kFluxGroup; ++q) acc[q] += (q == kk) ? val : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kFluxGroup; ++q) {
            if (k0 + q < n_open) {   // workgroup-uniform
                const double tot = block_sum(acc[q], lds);
                if (threadIdx.x == 0) work[(int64_t)(k0 + q) * wc + b] = tot;
            }
        }
    }
}

// Phase 2 (one wave): Q[k] = sum of the G chunk sums in FP64 (per lane in chunk order, then a fixed butterfly),
// then p_out[0] = drive and p_out[k] += alpha (R[k] Q[k] - pref - p_out[k]) for k >= 1 (in FP64, stored as float);
// q_last = Q (float). valid = false (a work buffer too small for the chunks) writes NaN, which the health check
// reports.
__device__ __forceinline__ void flux_final(const double* work, int wc, int64_t G, bool valid, int n_open,
                                           const float* R, float drive, float* p_out, float* q_last, float pref,
                                           float alpha) {
    const int lane = threadIdx.x;
    if (lane == 0) p_out[0] = drive;
    for (int k0 = 0; k0 < n_open; k0 += kFluxGroup) {
        double s[kFluxGroup];   // the group's loads are issued together; each sum keeps its order
#pragma unroll
        for (int q = 0; q < kFluxGroup; ++q) {
            s[q] = 0.0;
            if (valid && k0 + q < n_open)
                for (int64_t b = lane; b < G; b += 64) s[q] += work[(int64_t)(k0 + q) * wc + b];
        }
#pragma unroll
        for (int q = 0; q < kFluxGroup; ++q) {
            if (k0 + q >= n_open) break;   // wave-uniform
            double t = s[q];
            for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
            if (lane == 0) {
                const int k = k0 + q;
                const double qk = valid ? t : (double)__builtin_nanf("");
                q_last[k] = (float)qk;
                if (k > 0) {
                    const double p = (double)p_out[k];
                    p_out[k] = (float)(p + (double)alpha * ((double)R[k] * qk - (double)pref - p));
                }
            }
        }
    }
}

__global__ __launch_bounds__(kRedThreads) void k_flux_partial(cs_flux_args a) {
    __shared__ double lds[kRedThreads / 64];
    flux_partial(a.idx, a.w, a.slot, a.m, a.vel, a.vel_elem, a.vel_comp, a.n_open, a.work, a.work_chunks,
                 blockIdx.x, lds);
}

__global__ __launch_bounds__(64) void k_flux_final(cs_flux_args a) {
    flux_final(a.work, a.work_chunks, flux_chunks(a.m), true, a.n_open, a.R, a.drive[0], a.p_out, a.q_last, a.pref,
               a.alpha);
}

// packed lattices: grid (work_chunks, n_cases), case c's chunks b < flux_chunks(m_c)
__global__ __launch_bounds__(kRedThreads) void k_flux_partial_cases(cs_flux_cases_args a) {
    __shared__ double lds[kRedThreads / 64];
    const int c = blockIdx.y;
    const int n_open = a.n_open[c];
    if (n_open <= 0) return;  // workgroup-uniform
    const int64_t s0 = a.start[c], m = a.start[c + 1] - s0;
    if ((int64_t)blockIdx.x >= flux_chunks(m)) return;
    double* work = a.work + (int64_t)c * a.open_stride * a.work_chunks;
    flux_partial(a.idx + s0, a.w + 3 * s0, a.slot + s0, m, a.vel, a.vel_elem, a.vel_comp, n_open, work,
                 a.work_chunks, blockIdx.x, lds);
}

__global__ __launch_bounds__(64) void k_flux_final_cases(cs_flux_cases_args a) {
    const int c = blockIdx.x;
    const int n_open = a.n_open[c];
    if (n_open <= 0) return;  // workgroup-uniform
    const int64_t m = a.start[c + 1] - a.start[c];
    const int64_t G = flux_chunks(m);
    const int64_t o = (int64_t)c * a.open_stride;
    flux_final(a.work + o * a.work_chunks, a.work_chunks, G, G <= a.work_chunks, n_open, a.R + o, a.drive[c],
               a.p_out + o, a.q_last + o, a.pref[c], a.alpha[c]);
}

// ---- k_gather (one workgroup) ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kRedThreads) void k_gather(cs_gather_args a) {
    __shared__ float lds[kRedThreads / 64];
    float vmax = 0.0f;
    for (int64_t t = threadIdx.x; t < a.m; t += kRedThreads) {
        const int64_t s = a.idx[t];
        const float* v = a.vel + s * a.vel_elem;
        const float x = v[0], y = v[a.vel_comp], z = v[2 * a.vel_comp];
        vmax = fmaxf(vmax, sqrtf(x * x + y * y + z * z));
    }
    const float m = block_max(vmax, lds);
    if (threadIdx.x == 0) a.umax[0] = m;
    for (int k = 0; k < a.n_probe; ++k) {
        float acc = 0.0f;
        for (int64_t t = threadIdx.x; t < a.m; t += kRedThreads)
            if (a.slot[t] == k) acc += a.rho[(int64_t)a.idx[t] * (a.rho_elem > 1 ? a.rho_elem : 1)] / 3.0f;
        const float tot = block_sum(acc, lds);
        if (threadIdx.x == 0) a.out[k] = tot;
    }
}

// ---- finite check + FP64 sum ----------------------------------------------------------------------------------
__global__ __launch_bounds__(kRedThreads) void k_finite_partial(const float* x, int64_t n, double* work) {
    __shared__ double lds[kRedThreads / 64];
    double s = 0.0, bad = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kRedThreads;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride) {
        const float v = x[i];
        if (isfinite(v)) s += (double)v;
        else bad += 1.0;
    }
    const double ts = block_sum(s, lds);
    const double tb = block_sum(bad, lds);
    if (threadIdx.x == 0) {
        work[blockIdx.x] = ts;
        work[CS_REDUCE_BLOCKS + blockIdx.x] = tb;
    }
}

__global__ __launch_bounds__(64) void k_finite_final(const double* work, int nblocks, double* result) {
    if (threadIdx.x != 0) return;
    double s = 0.0, b = 0.0;
    for (int k = 0; k < nblocks; ++k) {
        s += work[k];
        b += work[CS_REDUCE_BLOCKS + k];
    }
    result[0] = s;
    result[1] = b;
}

// ---- wall shear stress: one thread per wall vertex, K cells summed in order ------------------------------------
__global__ __launch_bounds__(256) void k_wss(cs_wss_args a) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= a.n_vert) return;
    const float n0 = a.normal[3 * v], n1 = a.normal[3 * v + 1], n2 = a.normal[3 * v + 2];
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
    for (int j = 0; j < a.k; ++j) {
        const float cj = a.c[v * a.k + j];
        if (cj == 0.0f) continue;
        const float* u = a.vel + (int64_t)a.idx[v * a.k + j] * a.vel_elem;
        const float u0 = u[0], u1 = u[a.vel_comp], u2 = u[2 * a.vel_comp];
        const float un = u0 * n0 + u1 * n1 + u2 * n2;
        t0 += cj * (u0 - un * n0);
        t1 += cj * (u1 - un * n1);
        t2 += cj * (u2 - un * n2);
    }
    t0 *= a.scale;
    t1 *= a.scale;
    t2 *= a.scale;
    if (a.tau) {
        a.tau[3 * v] = t0;
        a.tau[3 * v + 1] = t1;
        a.tau[3 * v + 2] = t2;
    }
    if (a.acc_tau) {
        a.acc_tau[3 * v] += t0;
        a.acc_tau[3 * v + 1] += t1;
        a.acc_tau[3 * v + 2] += t2;
    }
    if (a.acc_mag) a.acc_mag[v] += sqrtf(t0 * t0 + t1 * t1 + t2 * t2);
}

}  // namespace
}  // namespace cs

extern "C" int cs_wss(const cs_wss_args* a, void* stream) {
    CS_REQUIRE(a && a->idx && a->c && a->normal && a->vel, "cs_wss: null argument");
    CS_REQUIRE(a->n_vert >= 0 && a->k > 0, "cs_wss: n_vert %lld, k %d", (long long)a->n_vert, a->k);
    if (a->n_vert == 0) return CS_OK;
    cs::k_wss<<<(unsigned)((a->n_vert + 255) / 256), 256, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_wss");
    return CS_OK;
}

extern "C" int cs_clock(const cs_clock_args* a, void* stream) {
    CS_REQUIRE(a && a->tick && a->coef && a->drive, "cs_clock: null argument");
    CS_REQUIRE(a->period > 0 && a->nh >= 0 && a->ramp >= 0, "cs_clock: period %d / nh %d / ramp %d invalid",
               a->period, a->nh, a->ramp);
    cs::k_clock<<<1, 64, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_clock");
    return CS_OK;
}

extern "C" int cs_flux_bc(const cs_flux_args* a, void* stream) {
    CS_REQUIRE(a && a->R && a->drive && a->p_out && a->q_last && a->work, "cs_flux_bc: null argument");
    CS_REQUIRE(a->n_open >= 1 && a->n_open <= cs::kMaxOpen, "cs_flux_bc: n_open %d not in [1, 64]", a->n_open);
    CS_REQUIRE(a->m >= 0 && (a->m == 0 || (a->idx && a->w && a->slot && a->vel)), "cs_flux_bc: null cell list");
    const int64_t per = (a->m + (int64_t)cs::kRedThreads * 256 - 1) / ((int64_t)cs::kRedThreads * 256);
    const int64_t chunk = (int64_t)cs::kRedThreads * (per < 1 ? 1 : per);
    const int64_t G = (a->m + chunk - 1) / chunk;
    CS_REQUIRE(a->work_chunks >= 1 && G <= a->work_chunks, "cs_flux_bc: %lld chunks need work_chunks >= %lld",
               (long long)G, (long long)(G < 1 ? 1 : G));
    if (G > 0) {
        cs::k_flux_partial<<<(unsigned)G, cs::kRedThreads, 0, cs::as_stream(stream)>>>(*a);
        CS_LAUNCHED("k_flux_partial");
    }
    cs::k_flux_final<<<1, 64, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_flux_final");
    return CS_OK;
}

// ---- cs_flux_bc_static: the same launches with the library's own chunk-sum scratch ---------------------------
// One array per device, for as many openings as cs_flux_bc accepts (64) x the most chunks a cell list can need (256),
// in FP64: 128 KiB, held by each device the library runs on until the process ends. In host code `scratch` names a
// host-side stand-in, not the device array: kernels get the device address from hipGetSymbolAddress.
constexpr int kStaticChunks = 256;
static __device__ __align__(256) double scratch[cs::kMaxOpen * kStaticChunks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device, looked up once per device
}  // namespace

extern "C" int cs_flux_bc_static(const cs_flux_args* a, void* stream) {
    CS_REQUIRE(a && a->R && a->drive && a->p_out && a->q_last, "cs_flux_bc_static: null argument");
    CS_REQUIRE(a->n_open >= 1 && a->n_open <= cs::kMaxOpen,
               "cs_flux_bc_static: n_open %d not in [1, %d] (the static scratch holds %d openings)", a->n_open,
               cs::kMaxOpen, cs::kMaxOpen);
    CS_REQUIRE(a->m >= 0 && (a->m == 0 || (a->idx && a->w && a->slot && a->vel)), "cs_flux_bc_static: null cell list");
    int dev = 0;
    CS_HIP(hipGetDevice(&dev));
    CS_REQUIRE(dev >= 0 && dev < kMaxDevices, "cs_flux_bc_static: device %d beyond %d", dev, kMaxDevices);
    double* work = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mutex);
        if (!scratch_ptr[dev]) {
            void* ptr = nullptr;
            CS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            scratch_ptr[dev] = static_cast<double*>(ptr);
        }
        work = scratch_ptr[dev];
    }
    cs_flux_args b = *a;
    b.work = work;
    b.work_chunks = kStaticChunks;
    const int64_t per = (b.m + (int64_t)cs::kRedThreads * 256 - 1) / ((int64_t)cs::kRedThreads * 256);
    const int64_t chunk = (int64_t)cs::kRedThreads * (per < 1 ? 1 : per);
    const int64_t G = (b.m + chunk - 1) / chunk;
    CS_REQUIRE(G <= kStaticChunks, "cs_flux_bc_static: %lld chunks exceed %d", (long long)G, kStaticChunks);
    if (G > 0) {
        cs::k_flux_partial<<<(unsigned)G, cs::kRedThreads, 0, cs::as_stream(stream)>>>(b);
        CS_LAUNCHED("k_flux_partial");
    }
    cs::k_flux_final<<<1, 64, 0, cs::as_stream(stream)>>>(b);
    CS_LAUNCHED("k_flux_final");
    return CS_OK;
}

extern "C" int cs_clock_table(const cs_clock_table_args* a, void* stream) {
    CS_REQUIRE(a && a->tick && a->coef && a->nh && a->period && a->ramp && a->drive, "cs_clock_table: null argument");
    CS_REQUIRE(a->n_cases >= 0 && a->n_cases <= CS_MAX_CASES && a->coef_stride >= 1 && a->n_steps >= 0 &&
                   a->drive_stride >= a->n_cases,
               "cs_clock_table: n_cases %d / coef_stride %d / n_steps %d / drive_stride %d", a->n_cases,
               a->coef_stride, a->n_steps, a->drive_stride);
    if (a->n_cases == 0 || a->n_steps == 0) return CS_OK;
    cs::k_clock_table<<<a->n_cases, 256, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_clock_table");
    return CS_OK;
}

extern "C" int cs_clock_cases(const cs_clock_cases_args* a, void* stream) {
    CS_REQUIRE(a && a->tick && a->coef && a->nh && a->period && a->ramp && a->drive, "cs_clock_cases: null argument");
    CS_REQUIRE(a->n_cases >= 0 && a->n_cases <= CS_MAX_CASES && a->coef_stride >= 1,
               "cs_clock_cases: n_cases %d / coef_stride %d", a->n_cases, a->coef_stride);
    if (a->n_cases == 0) return CS_OK;
    cs::k_clock_cases<<<(a->n_cases + 63) / 64, 64, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_clock_cases");
    return CS_OK;
}

extern "C" int cs_flux_cases(const cs_flux_cases_args* a, void* stream) {
    CS_REQUIRE(a && a->start && a->R && a->drive && a->p_out && a->q_last && a->n_open && a->pref && a->alpha &&
                   a->vel && a->work,
               "cs_flux_cases: null argument");
    CS_REQUIRE(a->n_cases >= 0 && a->n_cases <= CS_MAX_CASES && a->open_stride >= 1 &&
                   a->open_stride <= cs::kMaxOpen && a->work_chunks >= 1 && a->work_chunks <= 256,
               "cs_flux_cases: n_cases %d / open_stride %d / work_chunks %d", a->n_cases, a->open_stride,
               a->work_chunks);
    if (a->n_cases == 0) return CS_OK;
    cs::k_flux_partial_cases<<<dim3(a->work_chunks, a->n_cases), cs::kRedThreads, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_flux_partial_cases");
    cs::k_flux_final_cases<<<a->n_cases, 64, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_flux_final_cases");
    return CS_OK;
}

extern "C" int cs_gather(const cs_gather_args* a, void* stream) {
    CS_REQUIRE(a && a->out && a->umax, "cs_gather: null argument");
    CS_REQUIRE(a->n_probe >= 0 && a->n_probe <= cs::kMaxOpen, "cs_gather: n_probe %d not in [0, 64]", a->n_probe);
    CS_REQUIRE(a->m == 0 || (a->idx && a->slot && a->rho && a->vel), "cs_gather: null cell list");
    cs::k_gather<<<1, cs::kRedThreads, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_gather");
    return CS_OK;
}

extern "C" int cs_stream_copy(const void* src, void* dst, int64_t bytes, void* stream) {
    CS_REQUIRE(src && dst && bytes >= 0 && bytes % 16 == 0, "cs_stream_copy: null pointer or %lld bytes not a "
               "multiple of 16", (long long)bytes);
    if (bytes == 0) return CS_OK;
    const int64_t n = bytes / 16;
    const int64_t want = (n + 255) / 256;
    const unsigned groups = (unsigned)(want < cs::kCopyMaxGroups ? want : cs::kCopyMaxGroups);
    cs::k_stream_copy<<<groups, 256, 0, cs::as_stream(stream)>>>(reinterpret_cast<const uint4*>(src),
                                                                 reinterpret_cast<uint4*>(dst), n);
    CS_LAUNCHED("k_stream_copy");
    return CS_OK;
}

extern "C" int cs_bin_fields(const cs_bin_args* a, void* stream) {
    CS_REQUIRE(a && a->start && a->slots && a->rho && a->vel && a->p_out && a->u_out, "cs_bin_fields: null argument");
    CS_REQUIRE(a->n_bins >= 0 && a->rho_elem >= 1 && a->vel_elem >= 1 && a->vel_comp >= 1,
               "cs_bin_fields: n_bins %lld / strides", (long long)a->n_bins);
    if (a->n_bins == 0) return CS_OK;
    cs::k_bin_fields<<<(unsigned)((a->n_bins + 255) / 256), 256, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_bin_fields");
    return CS_OK;
}

extern "C" int cs_reduce_finite(const float* x, int64_t n, double* work, double* result, void* stream) {
    CS_REQUIRE(x && work && result && n >= 0, "cs_reduce_finite: invalid argument");
    int64_t want = (n + cs::kRedThreads - 1) / cs::kRedThreads;
    const int blocks = (int)(want < 1 ? 1 : (want > CS_REDUCE_BLOCKS ? CS_REDUCE_BLOCKS : want));
    cs::k_finite_partial<<<blocks, cs::kRedThreads, 0, cs::as_stream(stream)>>>(x, n, work);
    CS_LAUNCHED("k_finite_partial");
    cs::k_finite_final<<<1, 64, 0, cs::as_stream(stream)>>>(work, blocks, result);
    CS_LAUNCHED("k_finite_final");
    return CS_OK;
}

// Vessel signed distance (vessel.py:183-245): one thread per voxel centre, the branch stations read through
// wave-uniform addresses (scalar loads), FP32 like the reference's kernel.
#include <cmath>

#include "common.hpp"

namespace cs {
namespace {

constexpr int kMaxSdfBranches = 16;
constexpr int kSdfThreads = 256;

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 load3(const float* p) { return {p[0], p[1], p[2]}; }

// polynomial smooth minimum of width k (vessel.py:176-180)
__device__ __forceinline__ float smooth_min(float a, float b, float k) {
    const float h = fminf(fmaxf(0.5f + 0.5f * (b - a) / k, 0.0f), 1.0f);
    return b + (a - b) * h - k * h * (1.0f - h);
}

__global__ __launch_bounds__(kSdfThreads) void k_sdf_tubes(cs_sdf_tubes_args a) {
    const int64_t total = a.n[0] * a.n[1] * a.n[2];
    const int64_t v = (int64_t)blockIdx.x * kSdfThreads + threadIdx.x;
    if (v >= total) return;
    const int64_t l = v % a.n[2];
    const int64_t j = (v / a.n[2]) % a.n[1];
    const int64_t i = v / (a.n[2] * a.n[1]);
    const f3 p = {a.origin[0] + (float)i * a.dx, a.origin[1] + (float)j * a.dx, a.origin[2] + (float)l * a.dx};
    float d = 1.0e9f;
    for (int b = 0; b < a.n_branch; ++b) {
        float db = 1.0e9f;
        const int s0 = a.br_start[b], s1 = s0 + a.br_count[b] - 1;
        for (int s = s0; s < s1; ++s) {
            const float4 A = reinterpret_cast<const float4*>(a.pts)[s];
            const float4 B = reinterpret_cast<const float4*>(a.pts)[s + 1];
            const f3 ab = {B.x - A.x, B.y - A.y, B.z - A.z};
            const f3 ap = sub(p, {A.x, A.y, A.z});
            const float t = fminf(fmaxf(dot(ap, ab) / fmaxf(dot(ab, ab), 1.0e-12f), 0.0f), 1.0f);
            const float r = A.w + (B.w - A.w) * t;
            const f3 off = {p.x - (A.x + ab.x * t), p.y - (A.y + ab.y * t), p.z - (A.z + ab.z * t)};
            db = fminf(db, sqrtf(dot(off, off)) - r);
        }
        for (int e = 0; e < 2; ++e) {   // flat open end: a half-space, only near the cap
            const int c = 2 * b + e;
            const float cr = a.cap_r[c];
            if (cr > 0.0f) {
                const f3 q = sub(p, load3(a.cap_p + 3 * c));
                if (sqrtf(dot(q, q)) < 2.5f * cr) db = fmaxf(db, dot(q, load3(a.cap_n + 3 * c)));
            }
        }
        d = smooth_min(d, db, a.blend);
    }
    a.out[v] = d;
}

}  // namespace
}  // namespace cs

extern "C" int cs_sdf_tubes(const cs_sdf_tubes_args* a, void* stream) {
    CS_REQUIRE(a && a->pts && a->br_start && a->br_count && a->cap_p && a->cap_n && a->cap_r && a->out,
               "cs_sdf_tubes: null argument");
    CS_REQUIRE(a->n_branch >= 1 && a->n_branch <= cs::kMaxSdfBranches, "cs_sdf_tubes: n_branch %d not in [1, 16]",
               a->n_branch);
    CS_REQUIRE(a->blend > 0.0f && a->dx > 0.0f, "cs_sdf_tubes: blend and dx must be positive");
    CS_REQUIRE(a->n[0] > 0 && a->n[1] > 0 && a->n[2] > 0 && a->n[0] * a->n[1] * a->n[2] < ((int64_t)1 << 40),
               "cs_sdf_tubes: grid %lld x %lld x %lld out of range", (long long)a->n[0], (long long)a->n[1],
               (long long)a->n[2]);
    const int64_t total = a->n[0] * a->n[1] * a->n[2];
    const int64_t blocks = (total + cs::kSdfThreads - 1) / cs::kSdfThreads;
    CS_REQUIRE(blocks < ((int64_t)1 << 31), "cs_sdf_tubes: grid too large for one launch");
    cs::k_sdf_tubes<<<(unsigned)blocks, cs::kSdfThreads, 0, cs::as_stream(stream)>>>(*a);
    CS_LAUNCHED("k_sdf_tubes");
    return CS_OK;
}

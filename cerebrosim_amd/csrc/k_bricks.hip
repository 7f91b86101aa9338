// C entry points of the brick layout: the step (dispatch over the launch variants), the conversions between the
// AA storage and the reference (Q, N) layout, and the slot tables.
#include "k_bricks.hpp"

namespace cs {
namespace {

// Launch variants the autotuner chooses from: bricks (wavefronts) per workgroup and the XCD-aware workgroup order.
// Variant 0, the default, takes 4 bricks per workgroup in XCD-aware order: each XCD walks a contiguous Morton range,
// so a brick's neighbours (odd-step records, FP16 moment halo) are mostly in the same L2. In hardware order the
// interior launches read 8-55 % more from HBM (FETCH_SIZE); the step times of the variants are within about 2 %
// of each other, except 1 brick per workgroup.
struct Variant {
    int waves;
    int xcd;
};
constexpr Variant kVariant[] = {{4, 1}, {4, 0}, {2, 1}, {2, 0}, {1, 1}, {1, 0}};
constexpr int kVariants = (int)(sizeof(kVariant) / sizeof(kVariant[0]));

// ---- conversions between the AA storage and the reference (Q, N) layout --------------------------------------
__device__ __forceinline__ uint32_t cell_mask(const cs_bricks_convert_args& a, int b, int L, int& bb) {
    bb = a.bidx[b];
    return bb < 0 ? kAllLinks : a.lmask[(uint32_t)bb * BC + L];
}

// Location of P_i(n) in the AA storage for the state read by a step of parity `par`.
template <int I>
__device__ __forceinline__ uint32_t location(const cs_bricks_convert_args& a, int b, int L, uint32_t mask, int par) {
    constexpr int io = Lat::opp[I];
    if (par == 1 || I == 0 || !((mask >> io) & 1u)) return brick_base(b) + kSlots.p[io][L];
    const int x = L & 3, y = (L >> 2) & 3, z = L >> 4;
    int dd, l;
    shift_pos<Lat::cx[I], Lat::cy[I], Lat::cz[I]>(x, y, z, dd, l);
    const int nbk = a.nbr[(size_t)b * 27 + dd];
    return brick_base(nbk) + kSlots.p[I][l];
}

// Position of the side copy of link i of lane L of boundary brick bb in the compact side array: the lane's run
// (brick start + lane offset), then its side links below i (the order the step kernels use).
__device__ uint32_t side_index(const cs_bricks_convert_args& a, int bb, int L, int i) {
    const uint32_t s = a.smask[(uint32_t)bb * BC + L];
    return (uint32_t)a.sstart[bb] + (a.loff[(uint32_t)bb * BC + L] >> 16) + (uint32_t)__popc(s & ((1u << i) - 1u));
}

// P_i(n) is written by n itself, so the FP16 form is encoded and decoded with n's own newest moments.
template <bool SET, bool H16>
__global__ __launch_bounds__(256) void k_convert(cs_bricks_convert_args a) {
    using T = std::conditional_t<H16, _Float16, float>;
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= a.n_cells) return;
    const int64_t s = a.slot_of_cell[n];
    const int b = (int)(s / BC), L = (int)(s % BC);
    int bb;
    const uint32_t mask = cell_mask(a, b, L, bb);
    const uint32_t sm = bb >= 0 ? a.smask[(uint32_t)bb * BC + L] : 0u;
    T* F = reinterpret_cast<T*>(a.f);
    float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (H16) {
        m = reinterpret_cast<const float4*>(a.vel)[s];
        m.x = m.x - 1.5f * (m.y * m.y + m.z * m.z + m.w * m.w);
    }
    static_for<0, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const uint32_t loc = location<i>(a, b, L, mask, a.parity);
        float* ref = a.f_ref + (size_t)i * a.n_cells + n;
        if constexpr (SET) {
            if constexpr (H16) F[loc] = h16_encode<i>(*ref, m.x, m.y, m.z, m.w);
            else F[loc] = *ref;
            if ((sm >> i) & 1u) a.side[side_index(a, bb, L, i)] = *ref;
        } else {
            if constexpr (H16) *ref = h16_decode<i>(F[loc], m);
            else *ref = F[loc];
        }
    });
}

}  // namespace
}  // namespace cs

// Argument checks shared by the step entry points (the list may be empty).
static int check_step(const cs_bricks_step_args* a) {
    CS_REQUIRE(a && a->f && a->nbr && a->list && a->bidx && a->om_cell && a->om_brick && a->p_out && a->drive &&
                   a->vel_old && a->vel,
               "cs_bricks_step: null argument");
    CS_REQUIRE(!a->boundary || (a->lmask && a->smask && a->qlink && a->qstart && a->uin && a->side && a->sstart &&
                                a->loff && a->ldir),
               "cs_bricks_step: boundary bricks need lmask/smask/qlink/qstart/uin/side/sstart/loff/ldir");
    CS_REQUIRE(a->model == CS_MODEL_TRT || a->model == CS_MODEL_REGULARISED, "cs_bricks_step: unknown model %d",
               a->model);
    CS_REQUIRE(a->parity == 0 || a->parity == 1, "cs_bricks_step: parity %d", a->parity);
    CS_REQUIRE(a->storage == CS_STORAGE_FP32 || a->storage == CS_STORAGE_FP16, "cs_bricks_step: storage %d",
               a->storage);
    CS_REQUIRE(a->vel_old != a->vel, "cs_bricks_step: vel_old and vel must be distinct buffers");
    CS_REQUIRE(!a->hybrid || (a->model == CS_MODEL_REGULARISED && a->storage == CS_STORAGE_FP32 && a->sigma >= 0.0f &&
                              a->sigma < 1.0f),
               "cs_bricks_step: the hybrid blend needs the regularised model, FP32 storage and sigma in [0, 1)");
    CS_REQUIRE(a->storage == CS_STORAGE_FP16 || a->rho, "cs_bricks_step: FP32 storage needs rho");
    CS_REQUIRE(a->n_slots > 0 && a->n_slots % cs::BC == 0, "cs_bricks_step: n_slots %lld", (long long)a->n_slots);
    CS_REQUIRE(a->n_slots * cs::Q < ((int64_t)1 << 32), "cs_bricks_step: %lld slots exceed 32-bit population offsets",
               (long long)a->n_slots);
    CS_REQUIRE(a->n_list >= 0 && a->n_list < ((int64_t)1 << 31), "cs_bricks_step: n_list %lld",
               (long long)a->n_list);
    CS_REQUIRE(a->variant >= 0 && a->variant < cs::kVariants, "cs_bricks_step: variant %d not in [0, %d)", a->variant,
               cs::kVariants);
    return CS_OK;
}

extern "C" int cs_bricks_step(const cs_bricks_step_args* a, void* stream) {
    const int rc = check_step(a);
    if (rc != CS_OK) return rc;
    if (a->n_list == 0) return CS_OK;
    const cs::Variant v = cs::kVariant[a->variant];
    hipStream_t st = cs::as_stream(stream);
    hipError_t e = v.waves == 4 ? cs::bricks_launch_w4(*a, st, v.xcd)
                 : v.waves == 2 ? cs::bricks_launch_w2(*a, st, v.xcd) : cs::bricks_launch_w1(*a, st, v.xcd);
    if (e != hipSuccess) {
        cs::set_error("cs_bricks_step launch failed: %s", hipGetErrorString(e));
        return CS_E_HIP;
    }
    return CS_OK;
}

extern "C" int cs_bricks_step_pair(const cs_bricks_step_args* boundary, const cs_bricks_step_args* interior,
                                   void* stream) {
    CS_REQUIRE(boundary && interior, "cs_bricks_step_pair: null argument");
    CS_REQUIRE(boundary->boundary == 1 && interior->boundary == 0,
               "cs_bricks_step_pair: the first block must list boundary bricks, the second interior bricks");
    if (boundary->n_list == 0) return cs_bricks_step(interior, stream);
    if (interior->n_list == 0) return cs_bricks_step(boundary, stream);
    int rc = check_step(boundary);
    if (rc == CS_OK) rc = check_step(interior);
    if (rc != CS_OK) return rc;
    CS_REQUIRE(boundary->parity == interior->parity && boundary->storage == interior->storage &&
                   boundary->model == interior->model && boundary->galilean == interior->galilean &&
                   boundary->variant == interior->variant && boundary->hybrid == interior->hybrid &&
                   (!boundary->hybrid || boundary->sigma == interior->sigma),
               "cs_bricks_step_pair: the two blocks differ in parity, storage, model, galilean flag, variant or blend");
    const cs::Variant v = cs::kVariant[boundary->variant];
    hipStream_t st = cs::as_stream(stream);
    const hipError_t e = v.waves == 4 ? cs::bricks_launch_pair_w4(*boundary, *interior, st, v.xcd)
                       : v.waves == 2 ? cs::bricks_launch_pair_w2(*boundary, *interior, st, v.xcd)
                                      : cs::bricks_launch_pair_w1(*boundary, *interior, st, v.xcd);
    if (e != hipSuccess) {
        cs::set_error("cs_bricks_step_pair launch failed: %s", hipGetErrorString(e));
        return CS_E_HIP;
    }
    return CS_OK;
}

extern "C" int cs_bricks_run_pairs(const cs_bricks_step_args* args, int32_t n_pairs, int64_t n_steps, int32_t parity0,
                                   int64_t drive_stride, void* stream) {
    CS_REQUIRE(args && stream, "cs_bricks_run_pairs: null argument");
    CS_REQUIRE(n_pairs >= 0 && n_steps >= 0 && (parity0 == 0 || parity0 == 1) && drive_stride >= 0,
               "cs_bricks_run_pairs: n_pairs %d, n_steps %lld, parity0 %d, drive_stride %lld", n_pairs,
               (long long)n_steps, parity0, (long long)drive_stride);
    for (int k = 0; k < 4 * n_pairs; ++k)
        CS_REQUIRE(args[k].parity == (k < 2 * n_pairs ? 0 : 1), "cs_bricks_run_pairs: block %d has parity %d", k,
                   args[k].parity);
    for (int64_t s = 0; s < n_steps; ++s) {
        const cs_bricks_step_args* blk = args + (size_t)((parity0 + s) & 1) * 2 * n_pairs;
        for (int k = 0; k < n_pairs; ++k) {
            cs_bricks_step_args ab = blk[2 * k], ai = blk[2 * k + 1];
            ab.drive = blk[2 * k].drive + s * drive_stride;
            ai.drive = blk[2 * k + 1].drive + s * drive_stride;
            const int rc = cs_bricks_step_pair(&ab, &ai, stream);
            if (rc != CS_OK) return rc;
        }
    }
    return CS_OK;
}

static int convert(const cs_bricks_convert_args* a, void* stream, bool set) {
    CS_REQUIRE(a && a->f && a->nbr && a->bidx && a->slot_of_cell && a->f_ref && a->lmask && a->smask && a->side &&
                   a->sstart && a->loff,
               "cs_bricks convert: null argument");
    CS_REQUIRE(a->parity == 0 || a->parity == 1, "cs_bricks convert: parity %d", a->parity);
    CS_REQUIRE(a->storage == CS_STORAGE_FP32 || a->storage == CS_STORAGE_FP16, "cs_bricks convert: storage %d",
               a->storage);
    CS_REQUIRE(a->storage == CS_STORAGE_FP32 || (a->vel && a->n_slots > 0),
               "cs_bricks convert: FP16 storage needs the newest moments (vel) and n_slots");
    if (a->n_cells == 0) return CS_OK;
    const unsigned grid = (unsigned)((a->n_cells + 255) / 256);
    hipStream_t st = cs::as_stream(stream);
    if (a->storage == CS_STORAGE_FP16) {
        if (set) cs::k_convert<true, true><<<grid, 256, 0, st>>>(*a);
        else cs::k_convert<false, true><<<grid, 256, 0, st>>>(*a);
    } else {
        if (set) cs::k_convert<true, false><<<grid, 256, 0, st>>>(*a);
        else cs::k_convert<false, false><<<grid, 256, 0, st>>>(*a);
    }
    CS_LAUNCHED("k_convert");
    return CS_OK;
}

extern "C" int cs_bricks_slot_table(uint16_t* out) {
    CS_REQUIRE(out != nullptr, "cs_bricks_slot_table: null output");
    constexpr cs::SlotTable t = cs::make_slots();
    for (int k = 0; k < cs::Q; ++k)
        for (int l = 0; l < cs::BC; ++l) out[k * cs::BC + l] = t.p[k][l];
    return CS_OK;
}

extern "C" int cs_bricks_stage_table(uint16_t* out) {
    CS_REQUIRE(out != nullptr, "cs_bricks_stage_table: null output");
    constexpr cs::SlotTable t = cs::make_oddpos();
    for (int k = 0; k < cs::Q; ++k)
        for (int l = 0; l < cs::BC; ++l) out[k * cs::BC + l] = t.p[k][l];
    return CS_OK;
}

extern "C" int cs_bricks_record_table(uint16_t* out) {
    CS_REQUIRE(out != nullptr, "cs_bricks_record_table: null output");
    constexpr cs::RecordTable t = cs::make_records();
    for (int k = 0; k < cs::kRecs; ++k) out[k] = t.e[k];
    return CS_OK;
}

extern "C" int cs_bricks_get_f(const cs_bricks_convert_args* a, void* stream) { return convert(a, stream, false); }
extern "C" int cs_bricks_set_f(const cs_bricks_convert_args* a, void* stream) { return convert(a, stream, true); }

extern "C" int cs_bricks_variant_info(int32_t variant, int32_t* waves_per_group, int32_t* xcd_order) {
    if (variant < 0 || variant >= cs::kVariants) return cs::kVariants;
    CS_REQUIRE(waves_per_group && xcd_order, "cs_bricks_variant_info: null output");
    *waves_per_group = cs::kVariant[variant].waves;
    *xcd_order = cs::kVariant[variant].xcd;
    return CS_OK;
}

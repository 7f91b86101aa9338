// Brick-layout step kernel (semantics of lbm.py:140-257), shared by the translation units that instantiate it
// for each workgroup size (k_bricks_w*.hip) and by the conversions in k_bricks.hip.
//
// Storage: the 27 x 64 populations of 4^3 brick b occupy the block f[b * 1728 .. b * 1728 + 1728); population k of
// cell l (l = x + 4 y + 16 z) sits at offset slot[k][l] of the block. One wavefront per brick, lane l = cell l; four
// bricks per workgroup.
//
// Slot order. An even step touches exactly its brick's own block. An odd step reads and then writes slot (m, k) from
// the cell m - c_k, which lies in m's brick or in one of its 26 neighbours. slot[][] groups a block by that accessing
// brick: first the 1000 slots the brick's own wavefront reaches, then one contiguous record per neighbour offset (100
// slots for a face neighbour, 10 for an edge, 1 for a corner; plane by plane of the neighbour's z offset, faces,
// edges, corners inside a plane: group_at),
// population-major inside each group. In an odd step a
// wavefront therefore writes its own group and 26 records, each touched by no other wavefront, so cache lines fill
// from one wavefront instead of collecting partial writes from two.
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
#pragma once
#include <type_traits>

#include "collide.hpp"
#include "common.hpp"

namespace cs {
namespace {

constexpr int BC = 64;
constexpr int kLinkBatch = 8;  // boundary links evaluated per batch of loads (an even number: opposite pairs)
// LDS floats for a boundary brick's link run and side run, each with the one entry after it that lanes may read
// (the builder keeps the two runs within CS_BRICK_LINK_RUN_MAX); an FP16 stage holds exactly this many floats
constexpr int kLinkLds = CS_BRICK_LINK_RUN_MAX + 2;
static_assert(kLinkLds * 4 == 27 * 64 * 2, "the float16 even step's stage holds the link runs");
constexpr uint32_t kInletBit = 1u << 27;
constexpr uint32_t kAllLinks = (1u << Q) - 1;

struct SlotTable {
    uint16_t p[Q][BC];
};

// Slots of a block accessed from the brick at offset g ((ox+1) + 3 (oy+1) + 9 (oz+1)): 1000 from the brick itself,
// 100 from a face neighbour, 10 from an edge neighbour, 1 from a corner neighbour.
constexpr int group_size(int g) {
    int n = 0;
    for (int a = 0, v = g; a < 3; ++a, v /= 3) n += (v % 3 != 1);
    return n == 0 ? 1000 : (n == 1 ? 100 : (n == 2 ? 10 : 1));
}

// Group orders. In a block (group_at): the own group (13) first, then the records plane by plane of the accessing
// neighbour's z offset (oz = -1, 0, +1, as the offset index runs), inside a plane faces, then edges, then corners
// (each in offset order). Every plane holds an even number of slots and face and edge records hold an even number,
// so each face and edge record starts at an even slot: the float16 odd step moves them as 4-byte pairs. The planes
// stay together as in the plain offset order, which the direct accesses of boundary bricks walk. In an interior
// brick's staged odd-step footprint (stage_at, over neighbour offsets o): own group, faces, edges, corners, so that
// the pairs are contiguous in LDS.
struct GroupOrder {
    int g[27];
};

constexpr GroupOrder make_group_order(bool block) {
    GroupOrder t{};
    int k = 0;
    t.g[k++] = 13;
    const int sizes[3] = {100, 10, 1};
    if (block) {
        for (int gz = 0; gz < 3; ++gz)
            for (int sz : sizes)
                for (int g = 9 * gz; g < 9 * gz + 9; ++g)
                    if (g != 13 && group_size(g) == sz) t.g[k++] = g;
    } else {
        for (int sz : sizes)
            for (int g = 0; g < 27; ++g)
                if (g != 13 && group_size(g) == sz) t.g[k++] = g;
    }
    return t;
}

constexpr GroupOrder kBlockOrder = make_group_order(true);
constexpr GroupOrder kStageOrder = make_group_order(false);

constexpr int group_at(int gi) { return kBlockOrder.g[gi]; }
constexpr int stage_at(int oi) { return kStageOrder.g[oi]; }

constexpr SlotTable make_slots() {
    SlotTable t{};
    int pos = 0;
    // group g = brick offset (ox+1) + 3 (oy+1) + 9 (oz+1) of the accessing cell l - c_k, in group_at order
    for (int gi = 0; gi < 27; ++gi) {
        const int g = group_at(gi);
        for (int k = 0; k < Q; ++k)
            for (int l = 0; l < BC; ++l) {
                const int x = (l & 3) - Lat::cx[k], y = ((l >> 2) & 3) - Lat::cy[k], z = (l >> 4) - Lat::cz[k];
                const int o = (x < 0 ? 0 : (x > 3 ? 2 : 1)) + 3 * (y < 0 ? 0 : (y > 3 ? 2 : 1)) +
                              9 * (z < 0 ? 0 : (z > 3 ? 2 : 1));
                if (o == g) t.p[k][l] = (uint16_t)pos++;
            }
    }
    return t;
}

// 32-bit element offsets: the host guarantees n_bricks * 27 * 64 < 2^32 (one base pointer + a 32-bit VGPR offset
// per access keeps register use low).
__device__ __forceinline__ uint32_t brick_base(int b) { return (uint32_t)b * (uint32_t)(Q * BC); }

__constant__ SlotTable kSlots = make_slots();

// Brick offset index (ox+1) + 3 (oy+1) + 9 (oz+1) of the position (x, y, z) (components in [-1, 4]).
constexpr int offset_index(int x, int y, int z) {
    return (x < 0 ? 0 : (x > 3 ? 2 : 1)) + 3 * (y < 0 ? 0 : (y > 3 ? 2 : 1)) + 9 * (z < 0 ? 0 : (z > 3 ? 2 : 1));
}

// Start of group g inside a block (the order of make_slots: group_at).
constexpr int group_start(int g) {
    int pos = 0;
    for (int gi = 0; gi < 27; ++gi) {
        const int h = group_at(gi);
        if (h == g) return pos;
        pos += group_size(h);
    }
    return -1;
}

// The odd-step footprint of a brick, staged in LDS: its own group, then for every neighbour offset o != 13 the
// record of brick (b + o) that holds the slots accessed from b, i.e. group 26 - o of that brick. Staged order: own
// group, then the offsets o in stage_at order (faces, edges, corners), so that the face and edge records are
// contiguous 4-byte pairs; stage_start(o) is computed here.
constexpr int stage_start(int o) {
    int pos = 0;
    for (int oi = 0; oi < 27; ++oi) {
        const int h = stage_at(oi);
        if (h == o) return pos;
        pos += group_size(h == 13 ? 13 : 26 - h);
    }
    return -1;
}

// oddpos[i][l]: position in the staged odd-step footprint of the slot lane l reads for direction i,
// i.e. population opp(i) of the cell l - c_i (in this brick or a neighbour).
constexpr SlotTable make_oddpos() {
    SlotTable t{};
    constexpr SlotTable sl = make_slots();
    for (int i = 0; i < Q; ++i)
        for (int l = 0; l < BC; ++l) {
            const int x = (l & 3) - Lat::cx[i], y = ((l >> 2) & 3) - Lat::cy[i], z = (l >> 4) - Lat::cz[i];
            const int o = offset_index(x, y, z);
            const int ml = (x & 3) + 4 * (y & 3) + 16 * (z & 3);
            const int slot = sl.p[Lat::opp[i]][ml];
            const int g = o == 13 ? 13 : 26 - o;
            t.p[i][l] = (uint16_t)(stage_start(o) + slot - group_start(g));
        }
    return t;
}

// The odd-step footprint of an interior brick in staged order: the own group (block offsets [0, kOwn) of the own
// block), then the 26 neighbour records. Record table entry e[p] = (o << 11) | block offset, in the brick at
// neighbour offset o, of staged position kOwn + p (exported for tests; the kernels use the runs below).
constexpr int kOwn = 1000;
constexpr int kRecs = Q * BC - kOwn;
struct RecordTable {
    uint16_t e[kRecs];
};

constexpr RecordTable make_records() {
    RecordTable t{};
    int p = 0;
    for (int oi = 0; oi < 27; ++oi) {
        const int h = stage_at(oi);
        if (h == 13) continue;
        const int g = 26 - h;
        for (int k = 0; k < group_size(g); ++k) t.e[p++] = (uint16_t)((h << 11) | (group_start(g) + k));
    }
    return t;
}

// The same entries as runs: the 26 records in staged order, record r covering staged positions [start[r],
// start[r + 1]) with entry = base[r] + p (base[r] = (o << 11) + block offset of the record's first slot - start[r]),
// i.e. block offset off[r] + p in the brick at neighbour offset o = nbo[r]. The odd step computes its record
// addresses from these compile-time runs and the neighbour ids it holds in scalar registers.
struct RecordRuns {
    int start[26];
    int base[26];
    int nbo[26];
    int off[26];
};

constexpr RecordRuns make_runs() {
    RecordRuns t{};
    int p = 0, r = 0;
    for (int oi = 0; oi < 27; ++oi) {
        const int h = stage_at(oi);
        if (h == 13) continue;
        const int g = 26 - h;
        t.start[r] = p;
        t.base[r] = (h << 11) + group_start(g) - p;
        t.nbo[r] = h;
        t.off[r] = group_start(g) - p;
        p += group_size(g);
        ++r;
    }
    return t;
}

constexpr RecordRuns kRuns = make_runs();

constexpr bool runs_match_records() {
    constexpr RecordRuns t = make_runs();
    constexpr RecordTable rt = make_records();
    for (int p = 0; p < kRecs; ++p) {
        int r = 0;
        while (r + 1 < 26 && t.start[r + 1] <= p) ++r;
        if ((uint32_t)(t.base[r] + p) != rt.e[p]) return false;
    }
    return true;
}

static_assert(group_start(13) == 0 && group_size(13) == kOwn, "own group first");
static_assert(stage_start(stage_at(1)) == kOwn, "records after the own group");
static_assert(runs_match_records(), "record runs give the record table");

// Element address of staged record position p in [LO, HI] (the positions one load instruction covers): that of the
// last record starting at or before p, nb[o] * 1728 + block offset, as a select chain over the records the range
// touches. The neighbour ids nb are in scalar registers, so no table lookup stands between the brick's id and its
// record loads.
template <int LO, int HI>
__device__ __forceinline__ uint32_t record_addr(int p, const int (&nb)[27]) {
    uint32_t e = 0u;
    static_for<0, 26>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr int s = kRuns.start[r];
        constexpr int end = r + 1 < 26 ? kRuns.start[r + 1] : kRecs;
        if constexpr (s <= HI && end > LO) {
            const uint32_t v = brick_base(nb[kRuns.nbo[r]]) + (uint32_t)(kRuns.off[r] + p);
            if constexpr (s <= LO) e = v;
            else if (p >= s) e = v;
        }
    });
    return e;
}
static_assert([] {
    for (int r = 0; r < 26; ++r)
        if (kRuns.base[r] != (kRuns.nbo[r] << 11) + kRuns.off[r]) return false;
    return true;
}(), "record runs: base = (o << 11) + off");
__constant__ SlotTable kOddPos = make_oddpos();

// Odd steps of boundary bricks access global memory directly. pullsrc[i][l] = (d << 11) | block offset: where lane l
// pulls direction i from, population opp(i) of the cell l - c_i, in the brick at neighbour offset d (13: the brick
// itself). The same entry for direction opp(i) is where lane l pushes P_i (slot i of the cell l + c_i).
constexpr SlotTable make_pullsrc() {
    SlotTable t{};
    constexpr SlotTable sl = make_slots();
    for (int i = 0; i < Q; ++i)
        for (int l = 0; l < BC; ++l) {
            const int x = (l & 3) - Lat::cx[i], y = ((l >> 2) & 3) - Lat::cy[i], z = (l >> 4) - Lat::cz[i];
            const int ml = (x & 3) + 4 * (y & 3) + 16 * (z & 3);
            t.p[i][l] = (uint16_t)((offset_index(x, y, z) << 11) | sl.p[Lat::opp[i]][ml]);
        }
    return t;
}
static_assert(Q * BC <= 2048, "block offsets fit 11 bits");
__constant__ SlotTable kPullSrc = make_pullsrc();

// Float16 odd step: the face and edge records as 4-byte pairs (staged positions kOwn + 2p, 2p + 1, one record and
// an even block offset), then the 8 corner slots one by one (staged positions kOwn + 2 kNPairs + c).
constexpr int kNPairs = (6 * 100 + 12 * 10) / 2;
constexpr int kNCorners = 8;

constexpr bool pairs_valid() {
    constexpr RecordTable r = make_records();
    if (2 * kNPairs + kNCorners != kRecs) return false;
    for (int p = 0; p < kNPairs; ++p)
        if ((r.e[2 * p] & 1) != 0 || r.e[2 * p + 1] != r.e[2 * p] + 1) return false;  // same record, even offset
    return true;
}
static_assert(pairs_valid(), "face and edge records start at even offsets and pair up inside one record");
static_assert(kOwn % 8 == 0, "the staged records start 16-byte aligned in float16");

// c ? a : b as one v_cndmask on the wave mask of c. The compiler turns a select whose operand needs an LDS read into
// a divergent branch around that read; boundary bricks want both addresses computed and one access issued.
__device__ __forceinline__ uint32_t select_u32(bool c, uint32_t a, uint32_t b) {
    uint32_t r;
    const uint64_t m = __builtin_amdgcn_ballot_w64(c);
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
    return r;
}

// Cross-lane LDS hand-over inside one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Direct global -> LDS load of `size` bytes per lane (16 or 4): lane L's bytes land at lds + L * size. The LDS base
// is wave-uniform; the data arrive without passing through registers (wait for vmcnt before reading them).
template <int SIZE>
__device__ __forceinline__ void load_to_lds(const void* src, void* lds) {
    static_assert(SIZE == 16 || SIZE == 4, "16- or 4-byte lanes");
    auto* dst = (__attribute__((address_space(3))) void*)lds;
    if constexpr (SIZE == 16) __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
    else __builtin_amdgcn_global_load_lds(src, dst, 4, 0, 0);
}

// Copy a 27 x 64 uint16 table (216 x 16 bytes) into LDS with the workgroup's W waves, straight to LDS: nothing waits
// for it here (vmcnt and a workgroup barrier before the first read).
template <int W>
__device__ __forceinline__ void load_table(uint32_t* lds, const SlotTable& src, int wave, int L) {
    constexpr int NV = Q * BC * 2 / 16;
    static_for<0, (NV + 64 * W - 1) / (64 * W)>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int row = j * W + wave;  // 64 vectors per row, wave-uniform
        if (row * 64 < NV && row * 64 + L < NV)
            load_to_lds<16>(reinterpret_cast<const uint4*>(&src) + row * 64 + L,
                            reinterpret_cast<uint4*>(lds) + row * 64);
    });
}

// Brick id of the position (x, y, z) + e_a (one unit step, SX + SY + SZ = +-1): the brick itself or its face
// neighbour, from the neighbour ids in scalar registers.
template <int SX, int SY, int SZ>
__device__ __forceinline__ int face_brick(int x, int y, int z, const int (&nb)[27]) {
    static_assert((SX != 0) + (SY != 0) + (SZ != 0) == 1, "one axis");
    const bool out = SX < 0 ? x == 0 : SX > 0 ? x == 3 : SY < 0 ? y == 0 : SY > 0 ? y == 3 : SZ < 0 ? z == 0 : z == 3;
    return out ? nb[13 + SX + 3 * SY + 9 * SZ] : nb[13];
}

// Move the footprint of brick b between global memory and the wave's LDS stage with full-row accesses. DMA: loads
// go straight to LDS (interior bricks; boundary bricks keep the register copy, which times faster for them).
// Even steps: the whole own block. Odd steps (interior bricks): the own group as 16-byte vectors, then the 26
// neighbour records packed lane by lane (nb: the brick's neighbour ids). Loads are only issued here: the caller
// waits for them (vmcnt) before the stage is read.
// Record addresses per lane of an odd step's staged footprint, one per access instruction: FP32 12 rows of 4-byte
// slots; FP16 6 rows of 4-byte slot pairs and the row of corner slots. The load pass computes them (ra), the store
// pass reuses them.
template <typename T>
constexpr int kRecAddrs = sizeof(T) == 2 ? (kNPairs + BC - 1) / BC + 1 : (kRecs + BC - 1) / BC;

// PART: 0 everything, 1 the brick's own block (even) or own group (odd) only, 2 the odd step's records only (the
// parts that need the neighbour ids)
template <int PAR, bool STORE, bool DMA, int PART = 0, typename T>
__device__ __forceinline__ void stage_copy(T* __restrict__ F, T* stage, const int (&nb)[27], uint32_t base, int L,
                                           uint32_t (&ra)[kRecAddrs<T>]) {
    if constexpr (PAR == 0) {
        if constexpr (PART == 2) return;
        // the whole block as 16-byte vectors (a block is Q * BC * sizeof(T) bytes, a multiple of 16)
        constexpr int NV = Q * BC * (int)sizeof(T) / 16;
        uint4* g = reinterpret_cast<uint4*>(F + base);
        uint4* l = reinterpret_cast<uint4*>(stage);
        static_for<0, (NV + BC - 1) / BC>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int k = j * BC + L;
            if constexpr (!STORE && !DMA && NV % BC != 0 && (j + 1) * BC > NV) {
                // the partial last row goes straight to LDS, so that it is in flight with the full rows
                if (k < NV) load_to_lds<16>(g + k, l + j * BC);
            } else if (NV % BC == 0 || (j + 1) * BC <= NV || k < NV) {
                if constexpr (STORE) g[k] = l[k];
                else if constexpr (DMA) load_to_lds<16>(g + k, l + j * BC);
                else l[k] = g[k];
            }
        });
    } else {
        static_assert(kOwn * sizeof(T) % 16 == 0, "own group in 16-byte vectors");
        constexpr int NV = kOwn * (int)sizeof(T) / 16;
        uint4* g = reinterpret_cast<uint4*>(F + base);
        uint4* l = reinterpret_cast<uint4*>(stage);
        if constexpr (PART != 2)
            static_for<0, (NV + BC - 1) / BC>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const int k = j * BC + L;
                if (NV % BC == 0 || (j + 1) * BC <= NV || k < NV) {
                    if constexpr (STORE) g[k] = l[k];
                    else if constexpr (DMA) load_to_lds<16>(g + k, l + j * BC);
                    else l[k] = g[k];
                }
            });
        if constexpr (PART == 1) {
        } else if constexpr (sizeof(T) == 2) {
            // float16: face and edge records as 4-byte pairs (straight to LDS when DMA), corners one by one
            uint32_t* lp = reinterpret_cast<uint32_t*>(stage + kOwn);
            static_for<0, (kNPairs + BC - 1) / BC>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const int p = j * BC + L;
                if ((j + 1) * BC <= kNPairs || p < kNPairs) {
                    if constexpr (!STORE) ra[j] = record_addr<2 * j * BC, 2 * (j * BC + BC - 1)>(2 * p, nb);
                    uint32_t* g = reinterpret_cast<uint32_t*>(F + ra[j]);
                    if constexpr (STORE) *g = lp[p];
                    else if constexpr (DMA) load_to_lds<4>(g, lp + j * BC);
                    else lp[p] = *g;
                }
            });
            constexpr int jc = kRecAddrs<T> - 1;
            if (L < kNCorners) {
                if constexpr (!STORE) ra[jc] = record_addr<2 * kNPairs, 2 * kNPairs + kNCorners - 1>(2 * kNPairs + L, nb);
                const uint32_t at = ra[jc];
                if constexpr (STORE) F[at] = stage[kOwn + 2 * kNPairs + L];
                else stage[kOwn + 2 * kNPairs + L] = F[at];
            }
        } else {
            static_for<0, (kRecs + BC - 1) / BC>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const int p = j * BC + L;
                if ((j + 1) * BC <= kRecs || p < kRecs) {
                    if constexpr (!STORE) ra[j] = record_addr<j * BC, j * BC + BC - 1>(p, nb);
                    const uint32_t at = ra[j];
                    if constexpr (STORE) F[at] = stage[kOwn + p];
                    else if constexpr (DMA) load_to_lds<4>(F + at, stage + kOwn + j * BC);
                    else stage[kOwn + p] = F[at];
                }
            });
        }
    }
}

static_assert(Q * BC * 2 % 16 == 0, "table copy layout");
static_assert((Q * BC * 2) % 16 == 0, "a block of float16 populations is a whole number of 16-byte vectors");

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

// ---- FP16 storage -----------------------------------------------------------------------------------------------
// A stored population is h = (f~ - feq2(rho', u of the cell that wrote it)) / w * kH16Scale in float16, where
// feq2 = w (rho' + 3 c.u + 4.5 (c.u)^2) and rho' = drho - 1.5 |u|^2: the non-equilibrium part, which is small, keeps
// the float16's 11 significant bits. Every slot is written by the cell x that owns the population P_i(x); the reader
// decodes it with x's moments of the previous step (the velocity / density planes that step wrote). The writer of
// the value a cell n reads for direction i is n - c_i when link i is fluid and n itself otherwise (both parities).
constexpr float kH16Scale = 1024.0f;
constexpr int kHalo = 6 * 6 * 6;  // the brick and a one-cell shell

// Halo index of the position (x, y, z) + (SX, SY, SZ) in the 6^3 shell around the brick.
template <int SX, int SY, int SZ>
__device__ __forceinline__ int halo_index(int x, int y, int z) {
    return (x + 1 + SX) + 6 * (y + 1 + SY) + 36 * (z + 1 + SZ);
}

template <int I>
__device__ __forceinline__ _Float16 h16_encode(float fv, float rho_p, float u0, float u1, float u2) {
    const float cu = cdot<I>(u0, u1, u2);
    const float v = (fv * (1.0f / Lat::w[I]) - rho_p - cu * (3.0f + 4.5f * cu)) * kH16Scale;
    return (_Float16)v;
}

template <int I>
__device__ __forceinline__ float h16_decode(_Float16 h, const float4& m) {
    const float cu = cdot<I>(m.y, m.z, m.w);
    return Lat::w[I] * ((float)h * (1.0f / kH16Scale) + m.x + cu * (3.0f + 4.5f * cu));
}

// The same value divided by w_I (the input form of collide<1, false, true, ...>).
template <int I>
__device__ __forceinline__ float h16_decode_g(_Float16 h, const float4& m) {
    const float cu = cdot<I>(m.y, m.z, m.w);
    return __builtin_fmaf((float)h, 1.0f / kH16Scale, __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), m.x));
}

// Stage the previous-step moments {drho - 1.5 |u|^2, ux, uy, uz} of the brick and its one-cell shell in LDS
// (mom: float4 {drho, ux, uy, uz} per slot). Slots of absent neighbour bricks (boundary bricks only: every neighbour
// of an interior brick exists) read as zero. Two parts: halo_fetch issues the loads into registers (absent: bit k
// for chunk k), halo_put converts and stores them once they have arrived, so that other loads issue in between.
constexpr int kHaloChunks = (kHalo + BC - 1) / BC;
static_assert(kHaloChunks == 4, "four halo chunks: one 16-bit source word each, two per 32-bit word");

// Where halo entry k * 64 + L comes from, per lane L and chunk k: neighbour offset o (0..26) | cell l << 5. Lanes
// past the shell repeat its last entry (loaded, not stored). A compile-time table read with one 8-byte load per
// lane, instead of the base-6 digits of the entry computed in every wave.
struct HaloSrc {
    uint16_t v[BC][kHaloChunks];
};
constexpr HaloSrc make_halo_src() {
    HaloSrc t{};
    for (int L = 0; L < BC; ++L)
        for (int k = 0; k < kHaloChunks; ++k) {
            const int e = k * BC + L < kHalo ? k * BC + L : kHalo - 1;
            const int hx = e % 6, hy = (e / 6) % 6, hz = e / 36;
            const int o = (hx == 0 ? 0 : (hx == 5 ? 2 : 1)) + 3 * (hy == 0 ? 0 : (hy == 5 ? 2 : 1)) +
                          9 * (hz == 0 ? 0 : (hz == 5 ? 2 : 1));
            const int l = ((hx + 3) & 3) + 4 * ((hy + 3) & 3) + 16 * ((hz + 3) & 3);
            t.v[L][k] = (uint16_t)(o | l << 5);
        }
    return t;
}
__constant__ HaloSrc kHaloSrc = make_halo_src();

// src: the lane's row of kHaloSrc (chunks 0, 1 in .x, chunks 2, 3 in .y)
template <bool BND>
__device__ __forceinline__ void halo_fetch(float4 (&m)[kHaloChunks], uint32_t& absent, const int* tbl,
                                           const float4* __restrict__ mom, uint2 src) {
    static_for<0, kHaloChunks>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint32_t w = ((k < 2 ? src.x : src.y) >> (16 * (k & 1))) & 0xFFFFu;
        const int o = (int)(w & 31u), l = (int)(w >> 5);
        const int nb = tbl[o];
        m[k] = mom[(uint32_t)(BND && nb < 0 ? 0 : nb) * BC + l];
        if (BND && nb < 0) absent |= 1u << k;
    });
}

template <bool BND>
__device__ __forceinline__ void halo_put(float4* halo, const float4 (&m)[kHaloChunks], uint32_t absent, int L) {
    static_for<0, kHaloChunks>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int e = k * BC + L;
        if ((k + 1) * BC <= kHalo || e < kHalo) {
            float4 v = m[k];
            v.x = v.x - 1.5f * (v.y * v.y + v.z * v.z + v.w * v.w);
            if (BND && ((absent >> k) & 1u)) v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            halo[e] = v;
        }
    });
}


// Value streamed into cell n through missing link I (wall / inlet / outlet), lbm.py:163-181.
//   fo        raw value of slot I: P_opp(n)
//   opp_fluid link opp(I) is fluid; f_opp its pulled population f_in[opp I, n + c_I]
//   p_self    P_I(n) of the previous step (for q >= 1/2)
// GW: the populations are given divided by w_I (so are the results): the inlet and outlet terms drop the weight.
template <int I, bool GW = false>
__device__ __forceinline__ float link_value(float code, float fo, bool opp_fluid, float f_opp, float p_self, int bouzidi,
                                            const float (&uind)[3], const float (&up)[3], float upsq,
                                            const float* p_out) {
    constexpr float wI = GW ? 1.0f : Lat::w[I];
    if (code >= 0.0f) {
        float v = fo;
        if (bouzidi) {
            const float q = code;
            if (q < 0.5f) {
                if (opp_fluid) v = 2.0f * q * fo + (1.0f - 2.0f * q) * f_opp;
            } else {
                // fo / 2q + (2q - 1) / 2q * P_i: one hardware reciprocal (1 ulp) instead of two IEEE divisions
                const float t = 2.0f * q;
                v = (fo + (t - 1.0f) * p_self) * __builtin_amdgcn_rcpf(t);
            }
        }
        return v;
    }
    if (code > -1.5f) {  // inlet: bounce-back with the prescribed velocity
        return fo + 6.0f * wI * cdot<I>(uind[0], uind[1], uind[2]);
    }
    const int k = (int)(-2.0f - code);  // outlet k: anti-bounce-back at the prescribed pressure
    const float cu = cdot<I>(up[0], up[1], up[2]);
    return -fo + 2.0f * wI * (3.0f * p_out[k] + 4.5f * cu * cu - 1.5f * upsq);
}

// LDS of one workgroup of the step (byte offsets, 16-byte aligned): the slot table; in odd boundary steps the
// pull-source table; each wave's neighbour ids where lanes look them up; each wave's stage (staged steps), moment
// halo (FP16) and link-run buffer (odd boundary steps).
template <int W, bool BND, int PAR, bool H16>
struct BrickLds {
    static constexpr int kT = H16 ? 2 : 4;
    static constexpr bool STAGED = PAR == 0 || !BND;
    static constexpr bool TBL = H16 || (BND && PAR == 1);
    static constexpr int up16(int n) { return (n + 15) & ~15; }
    static constexpr int slots = 0;
    static constexpr int pull = slots + Q * BC * 2;
    static constexpr int tbl = pull + (BND && PAR == 1 ? Q * BC * 2 : 0);
    static constexpr int stage = up16(tbl + (TBL ? W * 32 * 4 : 0));
    static constexpr int halo = up16(stage + (STAGED ? W * Q * BC * kT : 0));
    static constexpr int lbuf = up16(halo + (H16 ? W * kHalo * 16 : 0));
    static constexpr int bytes = up16(lbuf + (BND && !STAGED ? W * kLinkLds * 4 : 0));
};

// One wavefront per brick, W bricks per workgroup; wg: the workgroup's index into the list (wg * W + wave is the
// wave's entry); smem: the workgroup's LDS (BrickLds layout).
template <int W, int MODEL, bool GAL, bool BND, int PAR, bool H16>
__device__ __forceinline__ void brick_step(const cs_bricks_step_args& a, int wg, unsigned char* smem) {
    constexpr int WAVES = W;
    using T = std::conditional_t<H16, _Float16, float>;
    using Lds = BrickLds<W, BND, PAR, H16>;
    // Even steps stage the brick's own block in LDS: vector loads and stores instead of 27 gathers that each touch
    // many lines of the block (in an even step no other wavefront touches the block). Odd steps of interior bricks
    // stage their footprint (own group + one record in each neighbour block: slots that no other wavefront
    // touches, since every cell and link around an interior brick is fluid). Odd steps of boundary bricks access
    // global memory directly: their records hold slots that neighbouring wavefronts write (redirected writes).
    // slt: in staged odd steps, the staged position of the value lane l reads for direction i (which is also where
    // it writes P_opp(i)); otherwise the block offset of population i of cell l.
    constexpr bool STAGED = PAR == 0 || !BND;
    constexpr bool ODD_STAGED = PAR == 1 && !BND;
    // Float16 storage, regularised model: the decoded inputs are g = f / w, the boundary links are evaluated in that
    // form (every formula is linear in the populations; the Bouzidi side copies stay actual populations), and the
    // collision writes the stored form directly (collide<1, false, true, true>).
    constexpr bool FUSED16 = H16 && MODEL == 1;
    // MODEL 2: the regularised collision with the hybrid blend (sigma < 1, lbm.py:224-229), FP32 storage only; it
    // needs the full velocity gradient of the previous step
    static_assert(MODEL != 2 || !H16, "the hybrid blend runs in FP32 storage");
    constexpr bool HYB = MODEL == 2;
    constexpr int CM = HYB ? 1 : MODEL;  // the collision model
    // per-lane lookups of the neighbour ids (odd boundary gathers, the float16 moment halo) read them from LDS
    constexpr bool LDS_TBL = H16 || (BND && PAR == 1);
    // FP32 + GAL: the neighbour velocities of the gradient stencil are loaded with the populations, except in odd
    // boundary steps (their direct gathers hold more registers: holding these too would lower the occupancy)
    constexpr bool HOIST_GRAD = !H16 && GAL && !HYB && (PAR == 0 || !BND);
    constexpr bool EARLY_OWN = !BND && !H16;
    static_assert(Lds::TBL == LDS_TBL && Lds::STAGED == STAGED, "LDS layout");
    // the odd-step record addresses are computed from compile-time runs, so that five workgroups of four waves fit
    // the LDS of a compute unit; odd boundary steps keep each wave's copy of its brick's link and side runs in lbuf
    // (even steps use the free stage)
    int (*tbl)[32] = reinterpret_cast<int (*)[32]>(smem + Lds::tbl);
    uint32_t* slots32 = reinterpret_cast<uint32_t*>(smem + Lds::slots);
    uint32_t* pull32 = reinterpret_cast<uint32_t*>(smem + Lds::pull);
    T* stage_all = reinterpret_cast<T*>(smem + Lds::stage);
    float4* halo_all = reinterpret_cast<float4*>(smem + Lds::halo);
    float* lbuf_all = reinterpret_cast<float*>(smem + Lds::lbuf);
    const uint16_t* slt = reinterpret_cast<const uint16_t*>(slots32);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int L = threadIdx.x & 63;
    const int64_t li = (int64_t)wg * WAVES + wave;
    // list entry = brick | case << CS_CASE_SHIFT (case 0 for a single lattice). A wave past the list or on a skip
    // entry only takes its part of the slot table copy (the workgroup barrier below waits for every wave).
    const uint32_t e0 = a.list[li < a.n_list ? li : a.n_list - 1];  // (launches have n_list >= 1)
    const uint32_t ent = li < a.n_list ? e0 : CS_LIST_SKIP;
    const bool valid = ent != CS_LIST_SKIP;
    const int b = valid ? (int)(ent & ((1u << CS_CASE_SHIFT) - 1u)) : 0;
    const int cs = (int)(ent >> CS_CASE_SHIFT);
    // The brick's 27 neighbour ids (13: the brick itself; -1: absent, around boundary bricks only), loaded as
    // scalars: the odd-step record addresses and the gradient stencil select among them, so that no table lookup
    // stands between the brick's id and its loads, and every load of the wave issues in one burst below.
    int nb[27];
    static_for<0, 27>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        nb[k] = a.nbr[(int64_t)b * 27 + k];
    });
    // the lanes' halo sources (FP16): a load that depends on nothing, issued first
    uint2 hsrc = make_uint2(0u, 0u);
    if constexpr (H16) hsrc = reinterpret_cast<const uint2*>(&kHaloSrc.v[0][0])[L];
    if constexpr (LDS_TBL) {
        // the same ids as a vector load of lanes 0..26 (the same cache lines): one load and one LDS write instead of
        // a select of each lane's id among the 27 scalars
        const int v = L < 27 ? a.nbr[(int64_t)b * 27 + L] : -1;
        if (L < 32) tbl[wave][L] = v;
        wave_sync();
    }

    const int x = L & 3, y = (L >> 2) & 3, z = L >> 4;
    const uint32_t S = (uint32_t)a.n_slots;
    const uint32_t slot = (uint32_t)b * BC + L;
    const uint32_t base = brick_base(b);
    T* __restrict__ F = reinterpret_cast<T*>(a.f);
    T* stage = stage_all + (STAGED ? wave * Q * BC : 0);
    float4* halo = halo_all + (H16 ? wave * kHalo : 0);

    uint32_t mask = kAllLinks;
    uint32_t sm = 0u, lo = 0u;  // boundary bricks: Bouzidi side bits and the lane's link-run offsets
    uint32_t wq = 0u, ws = 0u;  // boundary bricks: the brick's link unions (ldir)
    int bb = 0, qs = 0, ss = 0;  // boundary bricks: record, starts of its link and side runs
    uint32_t own_om = 0u;        // the brick has per-cell omegas (om_brick byte)
    // Inputs that do not depend on the populations are loaded before the populations arrive (the moment halo, the
    // boundary words, the cell's omega and previous velocity, the gradient stencil's neighbour velocities), so that
    // all round trips of the wave overlap. Every lane loads them (solid lanes too: all the addresses exist), so
    // that no divergent branch makes the wave wait for them; the omega of the cell is selected where it is used.
    float om_c = 0.0f, om_k = 0.0f, d = 0.0f, up[3] = {0.0f, 0.0f, 0.0f};
    // velocity ax at x - e_ax, x + e_ax (HOIST_GRAD)
    float gnb[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float4 hm[(kHalo + BC - 1) / BC];
    uint32_t habs = 0u;
    uint32_t ra[kRecAddrs<T>];  // odd interior steps: record addresses, computed for the loads, reused for the stores
    if (valid) {
        // the wave-uniform values first: scalar loads, issued before any direct-to-LDS copy (a store, after which
        // the compiler no longer takes a uniform load for a scalar one)
        if (a.case_omega) om_k = a.case_omega[cs];
        else om_k = a.omega;
        d = a.drive[cs];
        if constexpr (BND) {
            bb = a.brec ? a.brec[li] : (int)li;  // the entry's boundary record: bidx[brick] == bb
            wq = a.ldir[2 * bb];
            ws = a.bouzidi ? a.ldir[2 * bb + 1] : 0u;
            qs = a.qstart[bb];
            ss = a.sstart[bb];
        }
        // FP32 interior bricks: the brick's own part of the stage (the whole block in even steps, the own group in
        // odd ones) needs no neighbour id, so it is in flight before the wait for the ids
        if constexpr (EARLY_OWN) stage_copy<PAR, false, true, 1>(F, stage, nb, base, L, ra);
        if constexpr (H16) halo_fetch<BND>(hm, habs, tbl[wave], reinterpret_cast<const float4*>(a.vel_old), hsrc);
        own_om = a.om_brick[b];
        if constexpr (BND) {
            mask = a.lmask[(uint32_t)bb * BC + L];
            sm = a.bouzidi ? a.smask[(uint32_t)bb * BC + L] : 0u;
            lo = a.loff[(uint32_t)bb * BC + L];
            om_c = a.om_cell[slot];
            if constexpr (H16) {
                const float4 mo = reinterpret_cast<const float4*>(a.vel_old)[slot];
                up[0] = mo.y;
                up[1] = mo.z;
                up[2] = mo.w;
            } else {
                up[0] = a.vel_old[slot];
                up[1] = a.vel_old[S + slot];
                up[2] = a.vel_old[2 * S + slot];
            }
        }
        if constexpr (HOIST_GRAD)
            static_for<0, 3>([&](auto ac) {
                constexpr int ax = decltype(ac)::value;
                constexpr int sx = ax == 0, sy = ax == 1, sz = ax == 2;
                int d1, l1, d2, l2;
                shift_pos<-sx, -sy, -sz>(x, y, z, d1, l1);
                shift_pos<sx, sy, sz>(x, y, z, d2, l2);
                const float* V = a.vel_old + ax * S;
                // around boundary bricks an absent neighbour brick (-1) reads brick 0 instead: a valid address
                // whose value the link rules discard (the link is then not fluid)
                gnb[2 * ax] = V[(uint32_t)max(face_brick<-sx, -sy, -sz>(x, y, z, nb), 0) * BC + l1];
                gnb[2 * ax + 1] = V[(uint32_t)max(face_brick<sx, sy, sz>(x, y, z, nb), 0) * BC + l2];
            });
        if constexpr (EARLY_OWN) stage_copy<PAR, false, true, 2>(F, stage, nb, base, L, ra);
        else if constexpr (STAGED) stage_copy<PAR, false, !BND>(F, stage, nb, base, L, ra);
    }
    load_table<W>(slots32, ODD_STAGED ? kOddPos : kSlots, wave, L);
    if constexpr (BND && PAR == 1) load_table<W>(pull32, kPullSrc, wave, L);
    // the slot table (every wave's part) and the stage have landed in LDS
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!valid) return;  // wave-uniform
    // the byte is first used here: keeps its test (and the wait for it) after the barrier, not amid the loads
    asm volatile("" : "+v"(own_om));
    if constexpr (H16) {
        halo_put<BND>(halo, hm, habs, L);
        wave_sync();
    }
    const bool live = mask != 0u;  // solid slots of boundary bricks only join the staging copies

    // ---- streaming: raw values ---------------------------------------------------------------------------
    float f[Q];
    // value of direction i as stored -> float. FP16: a fluid link holds P_i(n - c_i), a non-fluid one the bounced
    // P_opp(i)(n); each is decoded with its writer's moments and its own direction.
    // Boundary bricks select without branching: the writer's moments (the upstream cell's, or the own cell's for a
    // missing link) and the sign of c.u (decoding with opp(i) flips c_i; w_opp(i) = w_i).
    auto value = [&](auto ic, T raw) -> float {
        constexpr int i = decltype(ic)::value;
        if constexpr (!H16) {
            return raw;
        } else if constexpr (!BND) {
            if constexpr (FUSED16) {
                if (!BND || ((mask >> i) & 1u))
                    return h16_decode_g<i>(raw, halo[halo_index<-Lat::cx[i], -Lat::cy[i], -Lat::cz[i]>(x, y, z)]);
                return h16_decode_g<Lat::opp[i]>(raw, halo[halo_index<0, 0, 0>(x, y, z)]);
            } else {
                if (!BND || ((mask >> i) & 1u))
                    return h16_decode<i>(raw, halo[halo_index<-Lat::cx[i], -Lat::cy[i], -Lat::cz[i]>(x, y, z)]);
                return h16_decode<Lat::opp[i]>(raw, halo[halo_index<0, 0, 0>(x, y, z)]);
            }
        } else {
            const bool fl = (mask >> i) & 1u;
            const float4 m = halo[select_u32(fl, halo_index<-Lat::cx[i], -Lat::cy[i], -Lat::cz[i]>(x, y, z),
                                             halo_index<0, 0, 0>(x, y, z))];
            const float c = cdot<i>(m.y, m.z, m.w);
            const float cu = fl ? c : -c;
            const float g = __builtin_fmaf((float)raw, 1.0f / kH16Scale,
                                           __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), m.x));
            if constexpr (FUSED16) return g;
            else return Lat::w[i] * g;
        }
    };
    if constexpr (STAGED) {
        if (live)
            static_for<0, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                f[i] = value(ic, stage[slt[i * BC + L]]);
            });
    } else {
        if (live) {
            // a fluid link pulls from its source slot (pullsrc), a missing one reads the cell's own slot i: the
            // address is selected, so each direction is one load without a divergent branch
            const uint16_t* pull = reinterpret_cast<const uint16_t*>(pull32);
            static_for<0, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if constexpr (i == 0) {
                    f[i] = value(ic, F[base + slt[L]]);
                } else {
                    const uint32_t e = pull[i * BC + L];
                    const uint32_t from = brick_base(tbl[wave][e >> 11]) + (e & 2047u);
                    f[i] = value(ic, F[select_u32((mask >> i) & 1u, from, base + slt[i * BC + L])]);
                }
            });
        }
    }
    // Boundary bricks: the brick's link run (qlink, with the entry after it that lanes may read ahead) and side run
    // are copied into LDS in one round trip (direct loads to LDS; the lengths from the last lane's offsets), so that
    // the per-direction link reads below are LDS reads. Even steps put them in the stage, free once the populations
    // are in registers (the wait makes sure of that). The builder keeps every brick's runs within the buffer (the
    // lengths are clamped to it all the same, so a bad layout cannot write past it).
    float* lb = nullptr;
    int nq = 0;
    if constexpr (BND) {
        const uint32_t cnt = mask ? (uint32_t)__popc(~mask & (kAllLinks - 1u)) : 0u;
        nq = min(__builtin_amdgcn_readlane((int)((lo & 0xFFFFu) + cnt), 63), kLinkLds - 2);
        const int ns = min(__builtin_amdgcn_readlane((int)((lo >> 16) + (uint32_t)__popc(sm & (kAllLinks - 1u))), 63),
                           kLinkLds - 2 - nq);
        if constexpr (STAGED) {
            lb = reinterpret_cast<float*>(stage);
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the stage reads above have returned
        } else {
            lb = lbuf_all + wave * kLinkLds;
        }
        for (int j = 0; j * BC < nq + 1; ++j)
            if (j * BC + L < nq + 1) load_to_lds<4>(a.qlink + qs + j * BC + L, lb + j * BC);
        for (int j = 0; j * BC < ns + 1; ++j)
            if (j * BC + L < ns + 1) load_to_lds<4>(a.side + ss + j * BC + L, lb + nq + 1 + j * BC);
        __builtin_amdgcn_s_waitcnt(0);
        wave_sync();
    }
    if (live) {
        // interior bricks read a per-cell omega only where the brick has them (sponge bricks: rare)
        const float om = own_om ? (BND ? om_c : a.om_cell[slot]) : om_k;
        const float* p_out = a.p_out + (int64_t)cs * a.p_out_stride;
        // ---- boundary links and the previous velocity --------------------------------------------------------
        float q16[6];  // BND + GAL or HYB: qc of links 1..6 (read by the gradient stencil), loaded with the link batch
        if constexpr (BND) {
            if ((mask & kAllLinks) != kAllLinks) {
                // compact link data: each lane's links in direction order from its offset in the brick's run, so
                // link i sits at the offset + the number of the lane's links below i. Directions that no cell of
                // the brick has (ldir, wave-uniform) are skipped; a lane that has not this link reads a neighbouring
                // entry (at most the end of its run: the arrays carry one element of padding) and ignores it.
                const uint32_t miss = ~mask & (kAllLinks - 1u);
                const uint32_t qo = lo & 0xFFFFu, so = lo >> 16;
                // held in a scalar register: otherwise the flag is reloaded from the kernel arguments (and waited
                // for, together with the outstanding LDS reads) at every link direction
                int bz = a.bouzidi;
                asm volatile("" : "+s"(bz));
                const float upsq = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
                float uind[3] = {0.0f, 0.0f, 0.0f};
                if (mask & kInletBit) {
                    uind[0] = a.uin[((uint32_t)bb * 3 + 0) * BC + L] * d;
                    uind[1] = a.uin[((uint32_t)bb * 3 + 1) * BC + L] * d;
                    uind[2] = a.uin[((uint32_t)bb * 3 + 2) * BC + L] * d;
                }
                const float* ql = lb;           // the brick's link run in LDS
                const float* sd = lb + nq + 1;  // its side run
                // links in batches of kLinkBatch (opposite pairs): first issue the batch's loads, then evaluate
                static_for<0, (Q - 1 + kLinkBatch - 1) / kLinkBatch>([&](auto bc) {
                    constexpr int h0 = 1 + kLinkBatch * decltype(bc)::value;
                    constexpr int nl = h0 + kLinkBatch <= Q ? kLinkBatch : Q - h0;
                    float qv[nl], pv[nl];
                    static_for<0, nl>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        constexpr uint32_t below = (1u << (h0 + k)) - 1u;
                        qv[k] = ((wq >> (h0 + k)) & 1u) ? ql[qo + (uint32_t)__popc(miss & below)] : 0.0f;
                        pv[k] = ((ws >> (h0 + k)) & 1u) ? sd[so + (uint32_t)__popc(sm & below)] : 0.0f;
                    });
                    if constexpr ((GAL || HYB) && h0 == 1)
                        static_for<0, 6>([&](auto kc) { q16[decltype(kc)::value] = qv[decltype(kc)::value]; });
                    static_for<0, nl / 2>([&](auto pc) {
                        constexpr int k = 2 * decltype(pc)::value;
                        constexpr int i = h0 + k, j = i + 1;
                        static_assert(pair_head<i>(), "batches start at pair heads");
                        const bool fi = (mask >> i) & 1u, fj = (mask >> j) & 1u;
                        const float ri = f[i], rj = f[j];
                        float vi = ri, vj = rj;
                        // side copies are actual populations: divided by w_i in the g form (w_i = w_j)
                        constexpr float sw = FUSED16 ? 1.0f / Lat::w[i] : 1.0f;
                        if (!fi) {
                            const float ps = ((sm >> i) & 1u) ? pv[k] * sw : rj;
                            vi = link_value<i, FUSED16>(qv[k], ri, fj, rj, ps, bz, uind, up, upsq, p_out);
                        }
                        if (!fj) {
                            const float ps = ((sm >> j) & 1u) ? pv[k + 1] * sw : ri;
                            vj = link_value<j, FUSED16>(qv[k + 1], rj, fi, ri, ps, bz, uind, up, upsq, p_out);
                        }
                        f[i] = vi;
                        f[j] = vj;
                    });
                });
            } else if constexpr (GAL || HYB) {
                static_for<0, 6>([&](auto kc) { q16[decltype(kc)::value] = 0.0f; });
            }
        }

        // ---- Galilean correction: d u_a / d x_a of the previous velocity (lbm.py:105-126, 196-202) ------------
        CollideParams p;
        p.om = om;
        p.om_m = MODEL == 0 ? 1.0f / (0.5f + a.magic / (1.0f / om - 0.5f)) : 0.0f;
        const float* force = a.case_force ? a.case_force + 3 * cs : a.force;
        p.F[0] = force[0] * d;
        p.F[1] = force[1] * d;
        p.F[2] = force[2] * d;
        p.t3 = GAL ? 1.0f : 0.0f;
        p.sigma = HYB ? a.sigma : 1.0f;
        p.g[0] = p.g[1] = p.g[2] = 0.0f;
        if constexpr (HYB) {
            // full gradient G[ax][c] = d u_c / d x_ax of the previous velocity (lbm.py:105-126 for every component):
            // the same one-sided rules as below, per component (a wall link at distance max(q, 0.1) with zero
            // velocity, an open link the cell's own)
            // (interior cells: both neighbours are fluid, h = 1, and the difference quotient needs no own velocity;
            // boundary cells have theirs in up)
            static_for<0, 3>([&](auto ac) {
                constexpr int ax = decltype(ac)::value;
                constexpr int sx = ax == 0, sy = ax == 1, sz = ax == 2;
                constexpr int i1 = 2 * ax + 1, i2 = 2 * ax + 2;
                int d1, l1, d2, l2;
                shift_pos<-sx, -sy, -sz>(x, y, z, d1, l1);
                shift_pos<sx, sy, sz>(x, y, z, d2, l2);
                const uint32_t s1 = (uint32_t)max(face_brick<-sx, -sy, -sz>(x, y, z, nb), 0) * BC + l1;
                const uint32_t s2 = (uint32_t)max(face_brick<sx, sy, sz>(x, y, z, nb), 0) * BC + l2;
                bool fl1 = true, fl2 = true, w1 = false, w2 = false;
                float h1 = 1.0f, h2 = 1.0f;
                if constexpr (BND) {
                    fl1 = (mask >> i1) & 1u;
                    fl2 = (mask >> i2) & 1u;
                    w1 = !fl1 && q16[i1 - 1] >= 0.0f;
                    w2 = !fl2 && q16[i2 - 1] >= 0.0f;
                    h1 = w1 ? fmaxf(q16[i1 - 1], 0.1f) : 1.0f;
                    h2 = w2 ? fmaxf(q16[i2 - 1], 0.1f) : 1.0f;
                }
                const float r = __builtin_amdgcn_rcpf(h1 * h2 * (h1 + h2));
                static_for<0, 3>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    const float n1 = a.vel_old[c * S + s1], n2 = a.vel_old[c * S + s2];
                    const float v1 = fl1 ? n1 : (w1 ? 0.0f : up[c]);
                    const float v2 = fl2 ? n2 : (w2 ? 0.0f : up[c]);
                    p.G[ax][c] = BND ? (h1 * h1 * (v2 - up[c]) + h2 * h2 * (up[c] - v1)) * r : (v2 - v1) * 0.5f;
                });
            });
            if constexpr (GAL) {
                p.g[0] = p.G[0][0];
                p.g[1] = p.G[1][1];
                p.g[2] = p.G[2][2];
            }
        } else if constexpr (GAL) {
            static_for<0, 3>([&](auto ac) {
                constexpr int ax = decltype(ac)::value;
                constexpr int i1 = 2 * ax + 1, i2 = 2 * ax + 2;
                int d1, l1, d2, l2;
                shift_pos<ax == 0 ? -1 : 0, ax == 1 ? -1 : 0, ax == 2 ? -1 : 0>(x, y, z, d1, l1);
                shift_pos<ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0>(x, y, z, d2, l2);
                const float* V = a.vel_old + ax * S;
                // neighbour velocity component ax of the previous step at x -+ e_ax: from the halo (FP16), loaded
                // with the populations (HOIST_GRAD) or here (odd boundary steps)
                auto nbv = [&](bool minus) -> float {
                    if constexpr (H16) {
                        const int e = minus ? halo_index<ax == 0 ? -1 : 0, ax == 1 ? -1 : 0, ax == 2 ? -1 : 0>(x, y, z)
                                            : halo_index<ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0>(x, y, z);
                        const float4 m = halo[e];
                        return ax == 0 ? m.y : (ax == 1 ? m.z : m.w);
                    } else if constexpr (HOIST_GRAD) {
                        return minus ? gnb[2 * ax] : gnb[2 * ax + 1];
                    } else {
                        return minus ? V[(uint32_t)max(tbl[wave][d1], 0) * BC + l1]
                                     : V[(uint32_t)max(tbl[wave][d2], 0) * BC + l2];
                    }
                };
                if constexpr (!BND) {
                    const float v1 = nbv(true);
                    const float v2 = nbv(false);
                    p.g[ax] = (v2 - v1) * 0.5f;
                } else {
                    // both neighbours are read whatever the links (the own slot stands in for a missing one: an
                    // address that always exists), so the loads issue together; then the wall and open-link rules
                    // select: a wall link at distance max(q, 0.1) with zero velocity, an open link the cell's own
                    const bool fl1 = (mask >> i1) & 1u, fl2 = (mask >> i2) & 1u;
                    const float n1 = nbv(true), n2 = nbv(false);
                    const float c1 = q16[i1 - 1], c2 = q16[i2 - 1];
                    const bool w1 = !fl1 && c1 >= 0.0f, w2 = !fl2 && c2 >= 0.0f;
                    const float h1 = w1 ? fmaxf(c1, 0.1f) : 1.0f, h2 = w2 ? fmaxf(c2, 0.1f) : 1.0f;
                    const float v1 = fl1 ? n1 : (w1 ? 0.0f : up[ax]);
                    const float v2 = fl2 ? n2 : (w2 ? 0.0f : up[ax]);
                    p.g[ax] = (h1 * h1 * (v2 - up[ax]) + h2 * h2 * (up[ax] - v1)) *
                              __builtin_amdgcn_rcpf(h1 * h2 * (h1 + h2));
                }
            });
        }

        // ---- collision -----------------------------------------------------------------------------------------
        float dr, u[3];
        bool inlet_cell = false;
        if constexpr (BND) inlet_cell = (mask & kInletBit) != 0u;
        if constexpr (CM == 1 && BND) {
            // inlet cells: the regularised collision at the prescribed velocity with om = 1, F = 0, g = 0
            float ut[3] = {0.0f, 0.0f, 0.0f};
            if (inlet_cell) {
                ut[0] = a.uin[((uint32_t)bb * 3 + 0) * BC + L] * d;
                ut[1] = a.uin[((uint32_t)bb * 3 + 1) * BC + L] * d;
                ut[2] = a.uin[((uint32_t)bb * 3 + 2) * BC + L] * d;
                p.om = 1.0f;
                p.F[0] = p.F[1] = p.F[2] = 0.0f;
                p.g[0] = p.g[1] = p.g[2] = 0.0f;
            }
            if constexpr (FUSED16) collide<1, false, true, true>(f, p, dr, u, inlet_cell, ut[0], ut[1], ut[2], kH16Scale);
            else collide<CM, HYB>(f, p, dr, u, inlet_cell, ut[0], ut[1], ut[2]);
        } else if (inlet_cell) {
            dr = density_deviation(f);
            u[0] = a.uin[((uint32_t)bb * 3 + 0) * BC + L] * d;
            u[1] = a.uin[((uint32_t)bb * 3 + 1) * BC + L] * d;
            u[2] = a.uin[((uint32_t)bb * 3 + 2) * BC + L] * d;
            inlet_equilibrium(f, dr, u[0], u[1], u[2], p.t3);
        } else if constexpr (FUSED16) {
            collide<1, false, true, true>(f, p, dr, u, false, 0.0f, 0.0f, 0.0f, kH16Scale);
        } else {
            collide<CM, HYB>(f, p, dr, u);
        }

        // ---- outputs -------------------------------------------------------------------------------------------
        if constexpr (H16) {
            reinterpret_cast<float4*>(a.vel)[slot] = make_float4(dr, u[0], u[1], u[2]);
        } else {
            a.vel[slot] = u[0];
            a.vel[S + slot] = u[1];
            a.vel[2 * S + slot] = u[2];
            a.rho[slot] = dr;
        }
        if constexpr (BND) {
            if (sm) {
                float* sd = a.side + ss;
                const uint32_t so = lo >> 16;
                // FUSED16: f holds the stored form; P_i = w_i (f_i / scale + rho' + c.u (3 + 4.5 c.u))
                const float rp = dr - 1.5f * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                static_for<1, Q>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    if ((ws >> i) & 1u)
                        if ((sm >> i) & 1u) {
                            float v = f[i];
                            if constexpr (FUSED16) {
                                const float cu = cdot<i>(u[0], u[1], u[2]);
                                v = Lat::w[i] * __builtin_fmaf(v, 1.0f / kH16Scale,
                                                               __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), rp));
                            }
                            sd[so + (uint32_t)__popc(sm & ((1u << i) - 1u))] = v;
                        }
                });
            }
        }
        if constexpr (H16 && !FUSED16) {
            // from here on f holds the stored form: the deviation from this cell's new equilibrium
            const float rp = dr - 1.5f * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            static_for<0, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                f[i] = (float)h16_encode<i>(f[i], rp, u[0], u[1], u[2]);
            });
        }
        if constexpr (!STAGED) {
            // odd: P_i(n) -> slot i of n + c_i (where the next even step of n + c_i reads it), or back into the own
            // slot opp(i) when n + c_i is not fluid
            const uint16_t* pull = reinterpret_cast<const uint16_t*>(pull32);
            static_for<0, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                constexpr int io = Lat::opp[i];
                if constexpr (i == 0) {
                    F[base + slt[L]] = (T)f[i];
                } else {
                    // slot i of n + c_i is where n + c_i pulls direction i from: pullsrc[opp i]
                    const uint32_t e = pull[io * BC + L];
                    const uint32_t to = brick_base(tbl[wave][e >> 11]) + (e & 2047u);
                    F[select_u32((mask >> io) & 1u, to, base + slt[io * BC + L])] = (T)f[i];
                }
            });
        }
    }
    if constexpr (STAGED) {
        // even: P_i(n) -> own slot opp(i); staged odd: P_i(n) -> S[n + c_i][i], the slot n reads for direction
        // opp(i) (the AA step reads and writes the same footprint); both at slt[opp(i)][l]
        wave_sync();
        if (live)
            static_for<0, Q>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                constexpr int io = Lat::opp[i];
                stage[slt[io * BC + L]] = (T)f[i];
            });
        wave_sync();
        stage_copy<PAR, true, !BND>(F, stage, nb, base, L, ra);
    }
}

// One list per launch. xcd: remap workgroup ids so that each XCD (which takes workgroups round-robin) processes a
// contiguous run of the list.
template <int W, int MODEL, bool GAL, bool BND, int PAR, bool H16>
__global__ __launch_bounds__(64 * W) void k_bricks(cs_bricks_step_args a, int xcd) {
    __shared__ __align__(16) unsigned char smem[BrickLds<W, BND, PAR, H16>::bytes];
    const int per = gridDim.x >> 3;
    const int wg = xcd ? (int)((blockIdx.x & 7) * per + (blockIdx.x >> 3)) : (int)blockIdx.x;
    brick_step<W, MODEL, GAL, BND, PAR, H16>(a, wg, smem);
}

// Both lists of a step in one launch (the two are independent within a step: every slot is read and then written by
// one thread of the step). gb workgroups take the boundary list, the rest the interior list; both parts hold a
// multiple of 8 workgroups, so each keeps its XCD-aware order. One launch per step instead of two on forked streams
// saves the fork and join between steps.
// Order, in rounds of 8 workgroups (one per XCD): the boundary rounds are spread evenly over the first kSpread %
// of the grid's rounds, the interior rounds fill the rest (0: every boundary round first, since their waves live
// longer). Spread, the boundary bricks' work overlaps the interior bricks' streaming instead of running as a phase
// of its own. Measured on the Womersley pipe and the story artery (paired launch, mean of both parities, HIP events):
// FP32 odd steps spread over 75 % (pipe 136.4 -> 134.9 us against spreading over the whole grid, 140.9 with the
// boundary first; artery 15.1 -> 15.7 k MLUPS), FP16 even steps over 50 % (102.2 -> 101.3 us); FP32 even and FP16
// odd steps are fastest boundary first.
template <int W, int MODEL, bool GAL, int PAR, bool H16>
__global__ __launch_bounds__(64 * W) void k_bricks_pair(cs_bricks_step_args ab, cs_bricks_step_args ai, int gb,
                                                        int xcd) {
    constexpr int kB = BrickLds<W, true, PAR, H16>::bytes, kI = BrickLds<W, false, PAR, H16>::bytes;
    __shared__ __align__(16) unsigned char smem[kB > kI ? kB : kI];
    const int bid = (int)blockIdx.x;
    constexpr int kSpread = H16 ? (PAR == 0 ? 50 : 0) : (PAR == 1 ? 75 : 0);
    if constexpr (kSpread > 0) {
        // round sb is a boundary round when floor(sb B8 / G8) steps up at sb: B8 of the first G8 rounds, evenly
        // spread; kb rounds before sb are boundary ones, sb - kb interior ones
        const int sb = bid >> 3, x = bid & 7;
        const int64_t B8 = gb >> 3;
        const int64_t G8 = max(B8, (int64_t)(gridDim.x >> 3) * kSpread / 100);
        const int kb = (int)min(B8, sb * B8 / G8), kb1 = (int)min(B8, (sb + 1) * B8 / G8);
        if (kb1 > kb) {
            brick_step<W, MODEL, GAL, true, PAR, H16>(ab, xcd ? x * (int)B8 + kb : kb * 8 + x, smem);
        } else {
            const int ki = sb - kb, I8 = (int)((int64_t)(gridDim.x >> 3) - B8);
            brick_step<W, MODEL, GAL, false, PAR, H16>(ai, xcd ? x * I8 + ki : ki * 8 + x, smem);
        }
        return;
    }
    if (bid < gb) {
        const int wg = xcd ? (bid & 7) * (gb >> 3) + (bid >> 3) : bid;
        brick_step<W, MODEL, GAL, true, PAR, H16>(ab, wg, smem);
    } else {
        const int g = bid - gb, gi = (int)gridDim.x - gb;
        const int wg = xcd ? (g & 7) * (gi >> 3) + (g >> 3) : g;
        brick_step<W, MODEL, GAL, false, PAR, H16>(ai, wg, smem);
    }
}

template <int W, int MODEL, bool GAL, bool BND>
hipError_t launch_par(const cs_bricks_step_args& a, hipStream_t s, int xcd) {
    const int64_t groups = (a.n_list + W - 1) / W;
    const unsigned grid = (unsigned)(((groups + 7) / 8) * 8);
    const bool h = a.storage == CS_STORAGE_FP16;
    if constexpr (MODEL == 2) {  // the hybrid blend: FP32 storage only (checked by the caller)
        if (a.parity == 0) k_bricks<W, MODEL, GAL, BND, 0, false><<<grid, 64 * W, 0, s>>>(a, xcd);
        else k_bricks<W, MODEL, GAL, BND, 1, false><<<grid, 64 * W, 0, s>>>(a, xcd);
    } else if (a.parity == 0) {
        if (h) k_bricks<W, MODEL, GAL, BND, 0, true><<<grid, 64 * W, 0, s>>>(a, xcd);
        else k_bricks<W, MODEL, GAL, BND, 0, false><<<grid, 64 * W, 0, s>>>(a, xcd);
    } else {
        if (h) k_bricks<W, MODEL, GAL, BND, 1, true><<<grid, 64 * W, 0, s>>>(a, xcd);
        else k_bricks<W, MODEL, GAL, BND, 1, false><<<grid, 64 * W, 0, s>>>(a, xcd);
    }
    return hipGetLastError();
}

template <int W, int MODEL>
hipError_t launch_model(const cs_bricks_step_args& a, hipStream_t s, int xcd) {
    if (a.galilean)
        return a.boundary ? launch_par<W, MODEL, true, true>(a, s, xcd) : launch_par<W, MODEL, true, false>(a, s, xcd);
    return a.boundary ? launch_par<W, MODEL, false, true>(a, s, xcd) : launch_par<W, MODEL, false, false>(a, s, xcd);
}

template <int W>
hipError_t launch_bricks(const cs_bricks_step_args& a, hipStream_t s, int xcd) {
    if (a.model != CS_MODEL_REGULARISED) return launch_model<W, 0>(a, s, xcd);
    return a.hybrid ? launch_model<W, 2>(a, s, xcd) : launch_model<W, 1>(a, s, xcd);
}

// The paired launch: both lists non-empty, same model, Galilean flag, parity and storage (checked by the caller).
template <int W, int MODEL, bool GAL>
hipError_t launch_pair_par(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd) {
    const int64_t gb = ((ab.n_list + W - 1) / W + 7) / 8 * 8, gi = ((ai.n_list + W - 1) / W + 7) / 8 * 8;
    const unsigned grid = (unsigned)(gb + gi);
    const bool h = ab.storage == CS_STORAGE_FP16;
    if constexpr (MODEL == 2) {  // the hybrid blend: FP32 storage only (checked by the caller)
        if (ab.parity == 0) k_bricks_pair<W, MODEL, GAL, 0, false><<<grid, 64 * W, 0, s>>>(ab, ai, (int)gb, xcd);
        else k_bricks_pair<W, MODEL, GAL, 1, false><<<grid, 64 * W, 0, s>>>(ab, ai, (int)gb, xcd);
    } else if (ab.parity == 0) {
        if (h) k_bricks_pair<W, MODEL, GAL, 0, true><<<grid, 64 * W, 0, s>>>(ab, ai, (int)gb, xcd);
        else k_bricks_pair<W, MODEL, GAL, 0, false><<<grid, 64 * W, 0, s>>>(ab, ai, (int)gb, xcd);
    } else {
        if (h) k_bricks_pair<W, MODEL, GAL, 1, true><<<grid, 64 * W, 0, s>>>(ab, ai, (int)gb, xcd);
        else k_bricks_pair<W, MODEL, GAL, 1, false><<<grid, 64 * W, 0, s>>>(ab, ai, (int)gb, xcd);
    }
    return hipGetLastError();
}

template <int W>
hipError_t launch_bricks_pair(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd) {
    if (ab.model == CS_MODEL_REGULARISED && ab.hybrid)
        return ab.galilean ? launch_pair_par<W, 2, true>(ab, ai, s, xcd) : launch_pair_par<W, 2, false>(ab, ai, s, xcd);
    if (ab.model == CS_MODEL_REGULARISED)
        return ab.galilean ? launch_pair_par<W, 1, true>(ab, ai, s, xcd) : launch_pair_par<W, 1, false>(ab, ai, s, xcd);
    return ab.galilean ? launch_pair_par<W, 0, true>(ab, ai, s, xcd) : launch_pair_par<W, 0, false>(ab, ai, s, xcd);
}

}  // namespace

// Launchers of the step for 1, 2 and 4 bricks (wavefronts) per workgroup (k_bricks_w1/2/4.hip).
hipError_t bricks_launch_w1(const cs_bricks_step_args& a, hipStream_t s, int xcd);
hipError_t bricks_launch_w2(const cs_bricks_step_args& a, hipStream_t s, int xcd);
hipError_t bricks_launch_w4(const cs_bricks_step_args& a, hipStream_t s, int xcd);
hipError_t bricks_launch_pair_w1(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd);
hipError_t bricks_launch_pair_w2(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd);
hipError_t bricks_launch_pair_w4(const cs_bricks_step_args& ab, const cs_bricks_step_args& ai, hipStream_t s, int xcd);
}  // namespace cs

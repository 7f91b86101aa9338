// Host-side builder of the brick layout (see cs_bricks_build in cerebro_lbm.h).
//
// Input is a lattice in the reference's indexed layout (lbm.py:264-330: nbr (Q, N) source cell or link code,
// q (Q, N), uin (N, 3), coords (N, 3)). Output: 4^3 bricks in Morton order, their 27-neighbour table (wrapped on
// periodic axes), the slot of every cell, and for bricks that touch a boundary the per-cell link masks, link codes
// with wall fractions, inlet profile and the mask of wall links that need a saved partner population.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"
#include "d3q27.hpp"

namespace cs {
namespace {

constexpr int B = 4;
constexpr int BC = B * B * B;
static_assert((B & (B - 1)) == 0 && BC == 64, "brick edge a power of two (bit masks), 64 cells (shift by 6)");
constexpr uint32_t kInletBit = 1u << 27;

struct BrickHost {
    int64_t N = 0, nb = 0, nbb = 0, nbi = 0;
    std::vector<int32_t> coords, nbr, bidx, interior, boundary, slot_of_cell, cell_of_slot;
    std::vector<uint32_t> lmask, smask;
    std::vector<float> qc, uin;
    // compact boundary link data (cs_bricks_copy "qlink" ... "ldir")
    int64_t n_links = 0, n_side = 0;
    std::vector<float> qlink;
    std::vector<int32_t> qstart, sstart;
    std::vector<uint32_t> loff, ldir;
};

// The per-step boundary data in compact form, only for links that exist: qlink holds qc of every missing link,
// brick by brick, lane by lane, each lane's links in direction order 1..26 (plus one element of padding, since a
// lane reads its entry for every direction of its brick's union, also where its own link is not missing);
// qstart / sstart are each brick's run start in qlink / in the Bouzidi side copies (links with an smask bit, same
// order); loff is each lane's run offset inside its brick's runs (qlink offset | side offset << 16), so link i of a
// lane is at its offset + the number of its links below i; ldir is the per-brick union of the missing links and of
// the side links.
int compact_links(BrickHost& h) {
    constexpr uint32_t kLinks = (1u << Q) - 2u;  // directions 1..26
    const int64_t nbb = h.nbb;
    h.qstart.assign(nbb + 1, 0);
    h.sstart.assign(nbb + 1, 0);
    h.loff.assign(nbb * BC, 0u);
    h.ldir.assign(nbb * 2, 0u);
    int64_t nq = 0;
    for (int64_t k = 0; k < nbb * BC; ++k) {
        const uint32_t m = h.lmask[k];
        nq += __builtin_popcount(m ? (~m & kLinks) : 0u);
    }
    h.qlink.assign(nq + 1, 0.0f);
    int64_t q = 0, s = 0;
    for (int64_t bb = 0; bb < nbb; ++bb) {
        CS_REQUIRE(q < (int64_t(1) << 31) && s < (int64_t(1) << 31), "cs_bricks_build: more than 2^31 boundary links");
        h.qstart[bb] = (int32_t)q;
        h.sstart[bb] = (int32_t)s;
        uint32_t wq = 0u, ws = 0u;
        for (int l = 0; l < BC; ++l) {
            const uint32_t m = h.lmask[bb * BC + l];
            const uint32_t miss = m ? (~m & kLinks) : 0u;
            const uint32_t sb = h.smask[bb * BC + l] & kLinks;
            h.loff[bb * BC + l] = (uint32_t)(q - h.qstart[bb]) | ((uint32_t)(s - h.sstart[bb]) << 16);
            for (int i = 1; i < Q; ++i)
                if ((miss >> i) & 1u) h.qlink[q++] = h.qc[(bb * Q + i) * BC + l];
            s += __builtin_popcount(sb);
            wq |= miss;
            ws |= sb;
        }
        h.ldir[2 * bb] = wq;
        h.ldir[2 * bb + 1] = ws;
        if ((q - h.qstart[bb]) + (s - h.sstart[bb]) > CS_BRICK_LINK_RUN_MAX) {
            set_error("cs_bricks_build: boundary brick %lld has %lld wall / open links and side copies (at most %d)",
                      (long long)bb, (long long)((q - h.qstart[bb]) + (s - h.sstart[bb])), CS_BRICK_LINK_RUN_MAX);
            return CS_E_LAYOUT;
        }
    }
    CS_REQUIRE(q < (int64_t(1) << 31) && s < (int64_t(1) << 31), "cs_bricks_build: more than 2^31 boundary links");
    h.qstart[nbb] = (int32_t)q;
    h.sstart[nbb] = (int32_t)s;
    h.n_links = q;
    h.n_side = s;
    return CS_OK;
}

uint64_t spread3(uint64_t v) {
    // interleave the low 21 bits of v with two zero bits between each
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

uint64_t morton(int x, int y, int z) { return spread3(x) | spread3(y) << 1 | spread3(z) << 2; }

int build(const cs_lattice_in* in, BrickHost& h) {
    const int64_t N = in->n_cells;
    h.N = N;
    int nbx[3];
    for (int a = 0; a < 3; ++a) {
        CS_REQUIRE(in->shape[a] > 0, "cs_bricks_build: shape[%d] = %d", a, in->shape[a]);
        if (in->periodic[a] && in->shape[a] % B != 0) {
            set_error("cs_bricks_build: periodic axis %d of length %d is not a multiple of the brick edge %d", a,
                      in->shape[a], B);
            return CS_E_LAYOUT;
        }
        nbx[a] = (in->shape[a] + B - 1) / B;
    }
    const int64_t ngrid = (int64_t)nbx[0] * nbx[1] * nbx[2];
    // brick occupancy on a dense brick grid
    std::vector<int32_t> grid(ngrid, -1);
    for (int64_t n = 0; n < N; ++n) {
        const int32_t* c = in->coords + 3 * n;
        for (int a = 0; a < 3; ++a)
            CS_REQUIRE(c[a] >= 0 && c[a] < in->shape[a], "cs_bricks_build: cell %lld outside the grid", (long long)n);
        grid[((int64_t)(c[0] / B) * nbx[1] + c[1] / B) * nbx[2] + c[2] / B] = 0;
    }
    std::vector<std::pair<uint64_t, int64_t>> keys;
    for (int64_t g = 0; g < ngrid; ++g)
        if (grid[g] == 0) {
            int bx = (int)(g / ((int64_t)nbx[1] * nbx[2])), by = (int)((g / nbx[2]) % nbx[1]), bz = (int)(g % nbx[2]);
            keys.push_back({morton(bx, by, bz), g});
        }
    std::sort(keys.begin(), keys.end());
    const int64_t nb = (int64_t)keys.size();
    if (nb * BC * Q >= ((int64_t)1 << 32)) {
        set_error("cs_bricks_build: %lld bricks exceed the 32-bit population offsets of one lattice", (long long)nb);
        return CS_E_LAYOUT;
    }
    h.nb = nb;
    h.coords.resize(3 * nb);
    for (int64_t b = 0; b < nb; ++b) {
        int64_t g = keys[b].second;
        grid[g] = (int32_t)b;
        h.coords[3 * b] = (int32_t)(g / ((int64_t)nbx[1] * nbx[2]));
        h.coords[3 * b + 1] = (int32_t)((g / nbx[2]) % nbx[1]);
        h.coords[3 * b + 2] = (int32_t)(g % nbx[2]);
    }
    // neighbour bricks, offset index d = (ox+1) + 3 (oy+1) + 9 (oz+1)
    h.nbr.assign(27 * nb, -1);
    for (int64_t b = 0; b < nb; ++b)
        for (int d = 0; d < 27; ++d) {
            int o[3] = {d % 3 - 1, (d / 3) % 3 - 1, d / 9 - 1};
            int p[3];
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                p[a] = h.coords[3 * b + a] + o[a];
                if (in->periodic[a]) p[a] = (p[a] + nbx[a]) % nbx[a];
                if (p[a] < 0 || p[a] >= nbx[a]) ok = false;
            }
            if (ok) h.nbr[27 * b + d] = grid[((int64_t)p[0] * nbx[1] + p[1]) * nbx[2] + p[2]];
        }
    // slots
    h.slot_of_cell.resize(N);
    h.cell_of_slot.assign(nb * BC, -1);
    for (int64_t n = 0; n < N; ++n) {
        const int32_t* c = in->coords + 3 * n;
        int64_t b = grid[((int64_t)(c[0] / B) * nbx[1] + c[1] / B) * nbx[2] + c[2] / B];
        int l = (c[0] % B) + B * (c[1] % B) + B * B * (c[2] % B);
        int64_t s = b * BC + l;
        CS_REQUIRE(s < INT32_MAX, "cs_bricks_build: lattice too large for 32-bit slots");
        h.slot_of_cell[n] = (int32_t)s;
        h.cell_of_slot[s] = (int32_t)n;
    }
    // verify that every fluid link is the implicit brick neighbour, symmetric, and classify bricks
    std::vector<uint8_t> is_boundary(nb, 0);
    for (int64_t s = 0; s < nb * BC; ++s)
        if (h.cell_of_slot[s] < 0) is_boundary[s / BC] = 1;
    // direction-major: row i of nbr is read in order, and its sources (n - c_i) and their rows opp(i) follow n
    // closely, so the checks stream through memory instead of jumping across the 27 rows for every cell
    for (int i = 0; i < Q; ++i) {
        const int ci[3] = {Lat::cx[i], Lat::cy[i], Lat::cz[i]};
        const int32_t* row = in->nbr + (int64_t)i * N;
        const int32_t* mirror = in->nbr + (int64_t)Lat::opp[i] * N;
        for (int64_t n = 0; n < N; ++n) {
            const int32_t src = row[n];
            const int64_t b = h.slot_of_cell[n] >> 6;  // slot / BC (slots are non-negative)
            if (src < 0) {
                CS_REQUIRE(src >= -3 - 64, "cs_bricks_build: link code %d out of range", src);
                is_boundary[b] = 1;
                continue;
            }
            CS_REQUIRE(src < N, "cs_bricks_build: nbr[%d, %lld] = %d out of range", i, (long long)n, src);
            // implicit source: position c - c_i inside brick b's neighbourhood (coordinates are non-negative, B = 4)
            const int32_t* c = in->coords + 3 * n;
            int o[3], l[3];
            for (int a = 0; a < 3; ++a) {
                const int local = (c[a] & (B - 1)) - ci[a];
                o[a] = local < 0 ? -1 : (local >= B ? 1 : 0);
                l[a] = local & (B - 1);
            }
            const int64_t nbk = h.nbr[27 * b + (o[0] + 1) + 3 * (o[1] + 1) + 9 * (o[2] + 1)];
            if (nbk < 0 || h.slot_of_cell[src] != nbk * BC + l[0] + B * l[1] + B * B * l[2]) {
                set_error("cs_bricks_build: link %d of cell %lld pulls from cell %d, which is not its grid neighbour "
                          "in the brick layout",
                          i, (long long)n, src);
                return CS_E_LAYOUT;
            }
            if (mirror[src] != (int32_t)n) {
                set_error("cs_bricks_build: link %d of cell %lld is not mirrored by link %d of cell %d", i,
                          (long long)n, Lat::opp[i], src);
                return CS_E_LAYOUT;
            }
        }
    }
    h.bidx.assign(nb, -1);
    for (int64_t b = 0; b < nb; ++b) {
        if (is_boundary[b]) {
            h.bidx[b] = (int32_t)h.boundary.size();
            h.boundary.push_back((int32_t)b);
        } else {
            h.interior.push_back((int32_t)b);
        }
    }
    h.nbb = (int64_t)h.boundary.size();
    h.nbi = (int64_t)h.interior.size();
    const int64_t nbb = h.nbb;
    h.lmask.assign(nbb * BC, 0u);
    h.smask.assign(nbb * BC, 0u);
    h.qc.assign(nbb * Q * BC, 0.0f);
    h.uin.assign(nbb * 3 * BC, 0.0f);
    for (int64_t n = 0; n < N; ++n) {
        const int64_t s = h.slot_of_cell[n];
        const int64_t bb = h.bidx[s / BC];
        if (bb < 0) continue;
        const int l = (int)(s % BC);
        uint32_t m = 0, sm = 0;
        bool inlet = false;
        for (int i = 0; i < Q; ++i) {
            const int32_t src = in->nbr[(int64_t)i * N + n];
            float code = 0.0f;
            if (src >= 0) {
                m |= 1u << i;
            } else if (src == WALL) {
                const float q = in->q[(int64_t)i * N + n];
                code = q;
                if (q >= 0.5f && in->nbr[(int64_t)Lat::opp[i] * N + n] >= 0) sm |= 1u << i;
            } else if (src == INLET) {
                code = -1.0f;
                inlet = true;
            } else {
                code = -2.0f - (float)(OUTLET0 - src);
            }
            h.qc[(bb * Q + i) * BC + l] = code;
        }
        if (inlet) m |= kInletBit;
        h.lmask[bb * BC + l] = m;
        h.smask[bb * BC + l] = sm;
        for (int a = 0; a < 3; ++a) h.uin[(bb * 3 + a) * BC + l] = in->uin[3 * n + a];
    }
    return CS_OK;
}

}  // namespace
}  // namespace cs

extern "C" int cs_bricks_build(const cs_lattice_in* in, void** handle) {
    CS_REQUIRE(in && handle && in->coords && in->nbr && in->q && in->uin, "cs_bricks_build: null argument");
    CS_REQUIRE(in->n_cells > 0, "cs_bricks_build: empty lattice");
    cs::BrickHost* h = new (std::nothrow) cs::BrickHost();
    if (!h) {
        cs::set_error("cs_bricks_build: out of host memory");
        return CS_E_NOMEM;
    }
    int rc;
    try {
        rc = cs::build(in, *h);
        if (rc == CS_OK) rc = cs::compact_links(*h);
    } catch (const std::bad_alloc&) {
        cs::set_error("cs_bricks_build: out of host memory");
        rc = CS_E_NOMEM;
    }
    if (rc != CS_OK) {
        delete h;
        return rc;
    }
    *handle = h;
    return CS_OK;
}

extern "C" int cs_bricks_sizes_of(const void* handle, cs_bricks_sizes* out) {
    CS_REQUIRE(handle && out, "cs_bricks_sizes_of: null argument");
    const cs::BrickHost* h = static_cast<const cs::BrickHost*>(handle);
    out->n_bricks = h->nb;
    out->n_boundary = h->nbb;
    out->n_interior = h->nbi;
    out->n_slots = h->nb * cs::BC;
    out->n_links = h->n_links;
    out->n_side = h->n_side;
    return CS_OK;
}

extern "C" int cs_bricks_copy(const void* handle, const char* name, void* dst) {
    CS_REQUIRE(handle && name && dst, "cs_bricks_copy: null argument");
    const cs::BrickHost* h = static_cast<const cs::BrickHost*>(handle);
    const std::string n(name);
    auto put = [&](const auto& v) {
        std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
        return CS_OK;
    };
    if (n == "coords") return put(h->coords);
    if (n == "nbr") return put(h->nbr);
    if (n == "bidx") return put(h->bidx);
    if (n == "interior") return put(h->interior);
    if (n == "boundary") return put(h->boundary);
    if (n == "slot_of_cell") return put(h->slot_of_cell);
    if (n == "cell_of_slot") return put(h->cell_of_slot);
    if (n == "lmask") return put(h->lmask);
    if (n == "smask") return put(h->smask);
    if (n == "qc") return put(h->qc);
    if (n == "uin") return put(h->uin);
    if (n == "qlink") return put(h->qlink);
    if (n == "qstart") return put(h->qstart);
    if (n == "sstart") return put(h->sstart);
    if (n == "loff") return put(h->loff);
    if (n == "ldir") return put(h->ldir);
    cs::set_error("cs_bricks_copy: unknown array '%s'", name);
    return CS_E_ARG;
}

extern "C" int cs_bricks_free(void* handle) {
    delete static_cast<cs::BrickHost*>(handle);
    return CS_OK;
}

// Checks dtb::group_by_owner (the grouped CSR that persistent delivery reads)
// on random delay-sorted CSRs, with g++ alone: every workgroup's part of
// every (row, delay) segment holds exactly that segment's synapses onto the
// workgroup's neurons, in CSR order, with the slice index, kind bit and
// rounded weight(s); the bounds are monotone and chain from one workgroup
// to the next; the arrivals that the persistent kernel's lookup
// (bptr + g * (n * D + 1) + row * D + delay - 1) accumulates equal the sums
// over the CSR.  Malformed input is refused.  Prints "ok <trials>".
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "dtb_group.hpp"

struct Csr {
    int64_t n, D;
    bool pair;
    uint32_t idx_bits, dbits, dmask, idx_mask;
    std::vector<int64_t> rp;
    std::vector<uint32_t> col;
    std::vector<float> w;
};

static Csr make_csr(std::mt19937_64& rng, int64_t n, int64_t D, bool pair, int max_deg)
{
    Csr c{n, D, pair, 1, 1, 0, 0, {}, {}, {}};
    while ((1ull << c.idx_bits) < (uint64_t)n) ++c.idx_bits;
    while ((1ll << c.dbits) < D) ++c.dbits;
    c.dmask = (1u << c.dbits) - 1u;
    c.idx_mask = (uint32_t)((1ull << c.idx_bits) - 1ull);
    c.rp.assign((size_t)n + 1, 0);
    for (int64_t j = 0; j < n; ++j) {
        const int deg = max_deg > 0 ? (int)(rng() % (uint64_t)(max_deg + 1)) : -max_deg;
        std::vector<std::pair<uint32_t, uint32_t>> row; // (delay, packed target)
        for (int e = 0; e < deg; ++e) {
            const uint32_t post = (uint32_t)(rng() % (uint64_t)n), d = 1u + (uint32_t)(rng() % (uint64_t)D);
            const uint32_t kind = (uint32_t)(rng() & 1u);
            row.push_back({d, post | ((d - 1u) << c.idx_bits) | (kind << 31)});
        }
        std::stable_sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (const auto& r : row) {
            c.col.push_back(r.second);
            // integers and halves, so the rounding (ties to even) is exercised
            for (int y = 0; y < (pair ? 2 : 1); ++y) c.w.push_back(0.5f * (float)((int)(rng() % 4001) - 2000));
        }
        c.rp[(size_t)j + 1] = (int64_t)c.col.size();
    }
    return c;
}

static int check(const Csr& c, int32_t G)
{
    const int64_t n = c.n, D = c.D, NS = n * D, per = (n + G - 1) / G;
    const int64_t nnz = (int64_t)c.col.size();
    const int R = c.pair ? 4 : 2, W = c.pair ? 2 : 1, C = c.pair ? 4 : 2;
    dtb::GroupedCsr g;
    const std::string err = dtb::group_by_owner(n, D, G, per, c.idx_bits, c.dmask, c.idx_mask, c.pair, c.rp.data(),
                                                c.col.data(), c.w.data(), g);
    if (!err.empty()) { std::printf("refused: %s\n", err.c_str()); return 1; }
    const size_t BS = (size_t)NS + 1;
    if (g.bptr.size() != (size_t)G * BS || g.rec.size() != (size_t)nnz * R) { std::printf("sizes\n"); return 1; }
    if (g.bptr[0] != 0 || g.bptr[(size_t)G * BS - 1] != nnz) { std::printf("ends\n"); return 1; }
    for (int32_t p = 0; p < G; ++p) {
        for (size_t q = 0; q < (size_t)NS; ++q)
            if (g.bptr[p * BS + q] > g.bptr[p * BS + q + 1]) { std::printf("not monotone\n"); return 1; }
        if (p + 1 < G && g.bptr[p * BS + NS] != g.bptr[(p + 1) * BS]) { std::printf("chain\n"); return 1; }
    }
    // expected parts, straight from the CSR: segment of synapse x = row * D + delay - 1
    std::vector<std::vector<int64_t>> want((size_t)G * NS);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t x = c.rp[(size_t)j]; x < c.rp[(size_t)j + 1]; ++x) {
            const int64_t d = ((c.col[(size_t)x] >> c.idx_bits) & c.dmask) + 1;
            const int64_t post = c.col[(size_t)x] & c.idx_mask;
            want[(size_t)((post / per) * NS + j * D + d - 1)].push_back(x);
        }
    std::vector<long long> sum_csr((size_t)n * C, 0), sum_grouped((size_t)n * C, 0);
    for (int64_t x = 0; x < nnz; ++x) {
        const uint32_t cc = c.col[(size_t)x];
        const int64_t post = cc & c.idx_mask;
        for (int y = 0; y < W; ++y)
            sum_csr[(size_t)(post * C + (cc >> 31) * (C / 2) + y)] += (long long)std::nearbyint(c.w[(size_t)x * W + y]);
    }
    for (int32_t p = 0; p < G; ++p)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t d = 1; d <= D; ++d) {
                const int32_t* bp = g.bptr.data() + (int64_t)p * (n * D + 1) + (j * D + (d - 1));
                const int32_t lo = bp[0], hi = bp[1];
                const std::vector<int64_t>& ws = want[(size_t)(p * NS + j * D + d - 1)];
                if ((size_t)(hi - lo) != ws.size()) { std::printf("part size g=%d row=%lld d=%lld\n", p, (long long)j, (long long)d); return 1; }
                for (int32_t x = lo; x < hi; ++x) {
                    const int64_t src = ws[(size_t)(x - lo)];
                    const uint32_t cc = c.col[(size_t)src];
                    const uint32_t r0 = g.rec[(size_t)x * R];
                    const int64_t li = r0 & 0x7fffffffu;
                    if (li != (int64_t)(cc & c.idx_mask) - p * per || (r0 >> 31) != (cc >> 31)) {
                        std::printf("record %d\n", x);
                        return 1;
                    }
                    for (int y = 0; y < W; ++y) {
                        const int32_t v = (int32_t)g.rec[(size_t)x * R + 1 + y];
                        if (v != (int32_t)std::nearbyint(c.w[(size_t)src * W + y])) { std::printf("weight %d\n", x); return 1; }
                        sum_grouped[(size_t)((p * per + li) * C + (r0 >> 31) * (C / 2) + y)] += v;
                    }
                    if (c.pair && g.rec[(size_t)x * R + 3] != 0u) { std::printf("pad %d\n", x); return 1; }
                }
            }
    if (sum_csr != sum_grouped) { std::printf("sums\n"); return 1; }
    return 0;
}

int main()
{
    std::mt19937_64 rng(12345);
    int trials = 0;
    // the plain block's shape (10,000 neurons, in-degree 100, delay 1, 32 workgroups)
    {
        Csr c = make_csr(rng, 10000, 1, false, -100);
        if (check(c, 32)) return 1;
        ++trials;
    }
    for (int t = 0; t < 30; ++t, ++trials) {
        const int64_t n = 1 + (int64_t)(rng() % 6000), D = 1 + (int64_t)(rng() % 40);
        const int32_t G = (int32_t)std::min<int64_t>(32, (n + 255) / 256);
        Csr c = make_csr(rng, n, D, t % 2 == 1, (int)(rng() % 120));
        if (check(c, G)) { std::printf("trial %d (n=%lld D=%lld G=%d)\n", t, (long long)n, (long long)D, G); return 1; }
    }
    // malformed input is refused, not grouped
    {
        Csr c = make_csr(rng, 300, 4, false, -8);
        // row 0 out of delay order: its first synapse at delay 4, its second at delay 1
        const uint32_t dfield = c.dmask << c.idx_bits;
        c.col[0] = (c.col[0] & ~dfield) | (3u << c.idx_bits);
        c.col[1] = (c.col[1] & ~dfield);
        dtb::GroupedCsr g;
        const int64_t per = (300 + 1) / 2;
        if (dtb::group_by_owner(300, 4, 2, per, c.idx_bits, c.dmask, c.idx_mask, false, c.rp.data(), c.col.data(),
                                c.w.data(), g).empty()) {
            std::printf("an unsorted row was accepted\n");
            return 1;
        }
        Csr e = make_csr(rng, 300, 1, false, -8);
        e.col[3] = (e.col[3] & ~e.idx_mask) | 300u; // a target past the last neuron
        if (dtb::group_by_owner(300, 1, 2, per, e.idx_bits, e.dmask, e.idx_mask, false, e.rp.data(), e.col.data(),
                                e.w.data(), g).empty()) {
            std::printf("a target past n was accepted\n");
            return 1;
        }
    }
    std::printf("ok %d\n", trials);
    return 0;
}

// Host-side grouping of a block's CSR for persistent delivery.  Plain C++
// (no HIP), so it also builds with g++ alone.
//
// The CSR rows are sorted by delay: segment q = j * D + d - 1 of row j and
// delay d is the run of row j whose delay field is d - 1.  Workgroup g of
// the persistent launch owns the neurons [g * per, (g + 1) * per).  The
// grouped copy stores every workgroup's synapses back to back, workgroup
// after workgroup, in segment order inside each (stable within a segment):
//   bptr[g * (n * D + 1) + q]  first record of workgroup g's part of segment q,
//   bptr[g * (n * D + 1) + n * D] = bptr[(g + 1) * (n * D + 1)]  (the end of g's run),
// so workgroup g reads segment q's records [bptr[g][q], bptr[g][q + 1]).
// A record is R = 2 words (compact weights) or 4 (pair weights):
//   word 0: post - g * per (the index in the workgroup's LDS slice) | kind << 31,
//   word 1 (and 2): the weight(s) rounded to the nearest int32, ties to even.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace dtb {

struct GroupedCsr {
    std::vector<int32_t> bptr; // [G][n * D + 1]
    std::vector<uint32_t> rec; // [nnz][R]
};

// rp: [n + 1] row starts; col: [nnz] packed targets (post in the low
// idx_bits, delay - 1 in the next dbits under dmask, kind in bit 31);
// w: [nnz] weights, or [nnz][2] with pair.  Returns an empty string, or what
// is wrong with the input.
inline std::string group_by_owner(int64_t n, int64_t D, int32_t G, int64_t per, uint32_t idx_bits, uint32_t dmask,
                                  uint32_t idx_mask, bool pair, const int64_t* rp, const uint32_t* col,
                                  const float* w, GroupedCsr& out)
{
    char msg[160];
    if (n <= 0 || D <= 0 || G <= 0 || per <= 0 || (int64_t)G * per < n) return "bad sizes";
    const int64_t nnz = rp[n];
    if (nnz < 0 || nnz >= (1ll << 31)) return "nnz out of range";
    const int64_t NS = n * D;
    std::vector<int64_t> sb((size_t)NS + 1); // CSR start of every segment
    for (int64_t j = 0; j < n; ++j) {
        int64_t e = rp[j];
        for (int64_t d = 1; d <= D; ++d) {
            sb[(size_t)(j * D + d - 1)] = e;
            while (e < rp[j + 1] && (int64_t)(((col[e] & 0x7fffffffu) >> idx_bits) & dmask) == d - 1) ++e;
        }
        if (e != rp[j + 1]) {
            std::snprintf(msg, sizeof msg, "row %lld is not sorted by delay", (long long)j);
            return msg;
        }
    }
    sb[(size_t)NS] = nnz;
    // pass 1: the size of every workgroup's part of every segment, into bptr[g][q + 1]
    const size_t BS = (size_t)NS + 1;
    out.bptr.assign((size_t)G * BS, 0);
    std::vector<int32_t>& bptr = out.bptr;
    for (int64_t x = 0; x < nnz; ++x) {
        const int64_t post = (int64_t)(col[x] & idx_mask);
        if (post >= n) {
            std::snprintf(msg, sizeof msg, "synapse %lld targets neuron %lld of %lld", (long long)x,
                          (long long)post, (long long)n);
            return msg;
        }
    }
    for (int64_t q = 0; q < NS; ++q)
        for (int64_t x = sb[(size_t)q]; x < sb[(size_t)q + 1]; ++x)
            ++bptr[(size_t)((int64_t)(col[x] & idx_mask) / per) * BS + (size_t)q + 1];
    int64_t run = 0;
    for (int32_t g = 0; g < G; ++g) {
        int32_t* bp = &bptr[(size_t)g * BS];
        bp[0] = (int32_t)run;
        for (size_t q = 0; q < (size_t)NS; ++q) bp[q + 1] += bp[q];
        run = bp[NS];
    }
    if (run != nnz) return "grouped synapse count differs from nnz";
    // pass 2: the records, each workgroup's part of a segment in CSR order
    const int R = pair ? 4 : 2;
    const int W = pair ? 2 : 1;
    out.rec.assign((size_t)nnz * R, 0u);
    std::vector<int32_t> pos((size_t)G);
    for (int64_t q = 0; q < NS; ++q) {
        for (int32_t g = 0; g < G; ++g) pos[(size_t)g] = bptr[(size_t)g * BS + (size_t)q];
        for (int64_t x = sb[(size_t)q]; x < sb[(size_t)q + 1]; ++x) {
            const uint32_t c = col[x];
            const int64_t post = (int64_t)(c & idx_mask);
            const int64_t g = post / per;
            const int32_t to = pos[(size_t)g]++;
            out.rec[(size_t)to * R] = (uint32_t)(post - g * per) | (c & 0x80000000u);
            for (int y = 0; y < W; ++y)
                out.rec[(size_t)to * R + 1 + y] = (uint32_t)(int32_t)std::nearbyint(w[x * W + y]);
        }
    }
    return std::string();
}

} // namespace dtb

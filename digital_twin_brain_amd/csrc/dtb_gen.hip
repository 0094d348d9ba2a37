// dtb_gen.hip -- seeded fixed in-degree networks built on the GPU, directly
// in the presynaptic-major packed CSR the step kernels read.
//
// The network is a list of members (one member for a single block; an
// uneven data-assimilation ensemble or the regions of a whole-brain share
// pack many): member m owns neurons [off_m, off_m + n_m), the first n_exc_m
// of them excitatory, and every one of its neurons receives Ki_m synapses
// from inside the member and K_m - Ki_m from other members chosen through a
// source table.  For post i of member m and slot k, one Philox4x32-10 call
// (ctr = (k, i_lo, i_hi, CONN), key = seed) gives r0..r3:
//   pre   = off_m + floor(r0 (n_m-1) / 2^32), shifted past i (no self-connection),
//           for k < Ki_m; for k >= Ki_m a second call (ctr = (k, i_lo, i_hi,
//           REGION)) picks the source s = smallest s with q0 < cdf[m][s] and
//           pre = off_s + floor(r0 n_s / 2^32)
//   w0    = fmaf(w_hi - w_lo, u24(r1), w_lo)
//   delay = 1 + floor(r2 D / 2^32)
//   w1    = fmaf(w_hi - w_lo, u24(r3), w_lo)      (pair weights)
//   kind  = pre - off_m >= n_exc_m
// Passes: (1) out-degree histogram, (2) exclusive scan -> rowptr,
// (3) scatter (delay, post, k) keys into their rows, (4) per-row LDS bitonic
// sort of the keys, then re-draw each synapse from its key and write the
// packed word, the weight and the row's delay-segment ends.  Rows come out
// ordered by (delay, post, k): independent of atomic ordering, and equal to
// the oracle's sequential generator after a stable per-row sort by delay.
#include <algorithm>
#include <vector>

#include "dtb_block.hpp"

namespace dtb {

struct GenMember {
    int64_t off, n, n_exc, syn_off; // syn_off: synapses of the members before this one
    int32_t K, Ki;                   // in-degree, of which the first Ki slots are intra-member
};

struct GenSpec {
    int64_t n;       // neurons in all members
    int64_t total;   // synapses in all members
    int32_t M, D;
    float w_lo, span;
    uint32_t k0, k1;
    const GenMember* __restrict__ mem;
    const uint64_t* __restrict__ cdf; // [R0][R0] source tables (scaled to 2^32), or null
    int32_t R0; // members per replica: member m = rep * R0 + r draws inter sources from row r, inside replica rep
};

constexpr uint32_t kStreamRegion = 0x52454749u; // ctr = (k, post_lo, post_hi, REGION)

__device__ inline u32x4 conn_draw(const GenSpec& s, int64_t i, int32_t k)
{
    return philox4x32_10(u32x4{(uint32_t)k, (uint32_t)i, (uint32_t)((uint64_t)i >> 32), kStreamConn}, s.k0, s.k1);
}

// member whose range [key_of(m), key_of(m+1)) holds v (members sorted)
__device__ inline int32_t member_of_syn(const GenSpec& s, int64_t x)
{
    int32_t lo = 0, hi = s.M - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (s.mem[mid].syn_off <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline int32_t member_of_neuron(const GenSpec& s, int64_t j)
{
    int32_t lo = 0, hi = s.M - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (s.mem[mid].off <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// synapse x -> (post i, slot k, member m)
__device__ inline void syn_coords(const GenSpec& s, int64_t x, int64_t& i, int32_t& k, int32_t& m)
{
    m = member_of_syn(s, x);
    const GenMember& g = s.mem[m];
    const int64_t local = x - g.syn_off;
    const int64_t il = local / g.K;
    k = (int32_t)(local - il * g.K);
    i = g.off + il;
}

__device__ inline uint32_t conn_pre(const GenMember& g, int64_t i, uint32_t r0)
{
    uint32_t pre = (uint32_t)(((uint64_t)r0 * (uint64_t)(g.n - 1)) >> 32);
    if ((int64_t)pre >= i - g.off) pre += 1;
    return (uint32_t)(g.off + pre);
}

__device__ inline uint32_t syn_pre(const GenSpec& s, int32_t m, int64_t i, int32_t k, uint32_t r0)
{
    const GenMember& g = s.mem[m];
    if (k < g.Ki) return conn_pre(g, i, r0);
    const u32x4 q =
        philox4x32_10(u32x4{(uint32_t)k, (uint32_t)i, (uint32_t)((uint64_t)i >> 32), kStreamRegion}, s.k0, s.k1);
    const int32_t rep = m / s.R0;
    const uint64_t* row = s.cdf + (int64_t)(m - rep * s.R0) * s.R0;
    int32_t lo = 0, hi = s.R0 - 1; // smallest source with q0 < row[source]; row[R0-1] = 2^32
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if ((uint64_t)q.x < row[mid]) hi = mid;
        else lo = mid + 1;
    }
    const GenMember& h = s.mem[rep * s.R0 + lo];
    return (uint32_t)(h.off + (int64_t)(((uint64_t)r0 * (uint64_t)h.n) >> 32));
}

__global__ void gen_count_kernel(GenSpec s, uint32_t* __restrict__ outdeg)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < s.total; x += stride) {
        int64_t i;
        int32_t k, m;
        syn_coords(s, x, i, k, m);
        const u32x4 r = conn_draw(s, i, k);
        atomicAdd(&outdeg[syn_pre(s, m, i, k, r.x)], 1u);
    }
}

// Exclusive scan of uint32 counts into int64 offsets: per-tile sums, a
// single-workgroup scan of the tile sums, then per-tile scans.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total)
{
    // Hillis-Steele over 256 lanes in LDS
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int64_t add = t >= off ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const int64_t incl = lds[t];
    *total = lds[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sums_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                       int64_t* __restrict__ tile_sums)
{
    __shared__ int64_t lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int64_t v = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t x = base + (int64_t)j * kScanThreads + threadIdx.x;
        if (x < n) v += in[x];
    }
    int64_t tot;
    block_exclusive_scan(v, lds, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_offsets_kernel(int64_t* __restrict__ tile_sums,
                                                                          int64_t ntiles)
{
    __shared__ int64_t lds[kScanThreads];
    int64_t carry = 0;
    for (int64_t base = 0; base < ntiles; base += kScanThreads) {
        const int64_t x = base + threadIdx.x;
        const int64_t v = x < ntiles ? tile_sums[x] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(v, lds, &tot);
        if (x < ntiles) tile_sums[x] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                   const int64_t* __restrict__ tile_off,
                                                                   int64_t* __restrict__ out)
{
    __shared__ int64_t lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    int64_t sum = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t x = base + j;
        v[j] = x < n ? in[x] : 0u;
        sum += v[j];
    }
    int64_t tot;
    int64_t run = tile_off[blockIdx.x] + block_exclusive_scan(sum, lds, &tot);
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t x = base + j;
        if (x < n) out[x] = run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_off[blockIdx.x] + tot;
}

__global__ void gen_fill_kernel(GenSpec s, unsigned long long* __restrict__ cursor, uint64_t* __restrict__ keys)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < s.total; x += stride) {
        int64_t i;
        int32_t k, m;
        syn_coords(s, x, i, k, m);
        const u32x4 r = conn_draw(s, i, k);
        const uint32_t pre = syn_pre(s, m, i, k, r.x);
        const uint64_t d = ((uint64_t)r.z * (uint64_t)s.D) >> 32; // delay - 1
        const unsigned long long pos = atomicAdd(&cursor[pre], 1ull);
        keys[pos] = (d << 58) | ((uint64_t)i << 32) | (uint32_t)k;
    }
}

// One workgroup per row: bitonic sort of the row's keys in LDS (rows up to
// kMaxRow synapses), then each synapse is re-drawn from its key.
constexpr int kSortThreads = 256;
constexpr int kMaxRow = 8192;

template <bool PAIR>
__global__ __launch_bounds__(kSortThreads) void gen_sort_finalize_kernel(GenSpec s, SynPack pk,
                                                                          const int64_t* __restrict__ rowptr,
                                                                          const uint64_t* __restrict__ keys,
                                                                          uint32_t* __restrict__ col,
                                                                          float* __restrict__ w,
                                                                          float2* __restrict__ w2,
                                                                          uint16_t* __restrict__ dend,
                                                                          uint32_t* __restrict__ too_long)
{
    __shared__ uint64_t lds[kMaxRow];
    for (int64_t j = blockIdx.x; j < s.n; j += gridDim.x) {
        const int64_t beg = rowptr[j];
        const int32_t len = (int32_t)(rowptr[j + 1] - beg);
        if (len > kMaxRow) {
            if (threadIdx.x == 0) atomicAdd(too_long, 1u);
            continue;
        }
        if (len == 0) {
            if (dend)
                for (int32_t dd = threadIdx.x; dd < s.D; dd += kSortThreads) dend[j * s.D + dd] = 0;
            continue;
        }
        int32_t P = 1;
        while (P < len) P <<= 1;
        for (int32_t x = threadIdx.x; x < P; x += kSortThreads) lds[x] = x < len ? keys[beg + x] : ~0ull;
        __syncthreads();
        for (int32_t size = 2; size <= P; size <<= 1) {
            for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (int32_t t = threadIdx.x; t < (P >> 1); t += kSortThreads) {
                    const int32_t lo = 2 * t - (t & (stride - 1));
                    const int32_t hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const uint64_t a = lds[lo], b = lds[hi];
                    if ((a > b) == up) {
                        lds[lo] = b;
                        lds[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        const GenMember& gm = s.mem[member_of_neuron(s, j)];
        const uint32_t kind = (j - gm.off >= gm.n_exc) ? 1u : 0u;
        if (dend) {
            // end of the delay-(dd+1) segment = first key with delay-1 > dd
            for (int32_t dd = threadIdx.x; dd < s.D; dd += kSortThreads) {
                int32_t lo = 0, hi = len;
                while (lo < hi) {
                    const int32_t mid = (lo + hi) >> 1;
                    if ((int32_t)(lds[mid] >> 58) <= dd) lo = mid + 1;
                    else hi = mid;
                }
                dend[j * s.D + dd] = (uint16_t)lo;
            }
        }
        for (int32_t x = threadIdx.x; x < len; x += kSortThreads) {
            const uint64_t key = lds[x];
            const int64_t i = (int64_t)((key >> 32) & 0x3FFFFFFull);
            const int32_t k = (int32_t)(uint32_t)key;
            const uint32_t d = (uint32_t)(key >> 58);
            const u32x4 r = conn_draw(s, i, k);
            col[beg + x] = (uint32_t)i | (d << pk.idx_bits) | (kind << 31);
            const float a0 = fmaf(s.span, u24(r.y), s.w_lo);
            if (PAIR)
                w2[beg + x] = make_float2(a0, fmaf(s.span, u24(r.w), s.w_lo));
            else
                w[beg + x] = a0;
        }
        __syncthreads();
    }
}

static inline unsigned grid_of(int64_t work, int bs, int64_t cap)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cap, (work + bs - 1) / bs));
}

static int gen_run(Block* b, const GenSpec& s, int32_t max_delay)
{
    const int64_t n = b->n;
    const int64_t nnz = b->nnz;
    void* p = nullptr;
    DTB_TRY(block_alloc(b, &p, sizeof(int64_t) * (size_t)(n + 1)));
    b->rowptr = (int64_t*)p;
    if (max_delay > 1) {
        DTB_TRY(block_alloc(b, &p, sizeof(uint16_t) * (size_t)n * (size_t)max_delay));
        b->dend = (uint16_t*)p;
    }
    DTB_TRY(block_alloc(b, &p, sizeof(uint32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    b->col = (uint32_t*)p;
    if (b->pair) {
        DTB_TRY(block_alloc(b, &p, sizeof(float2) * (size_t)std::max<int64_t>(nnz, 1)));
        b->w2 = (float2*)p;
    } else {
        DTB_TRY(block_alloc(b, &p, sizeof(float) * (size_t)std::max<int64_t>(nnz, 1)));
        b->w = (float*)p;
    }
    uint32_t* outdeg = nullptr;
    int64_t* tiles = nullptr;
    unsigned long long* cursor = nullptr;
    uint64_t* keys = nullptr;
    uint32_t* flag = nullptr;
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    int st = DTB_OK;
    auto fail = [&](hipError_t e, const char* what) {
        set_error("%s failed: %s", what, hipGetErrorString(e));
        st = e == hipErrorOutOfMemory ? DTB_OUT_OF_MEMORY : DTB_RUNTIME_ERROR;
    };
    hipError_t e;
    do {
        if ((e = hipMalloc(&outdeg, sizeof(uint32_t) * (size_t)n)) != hipSuccess) { fail(e, "hipMalloc"); break; }
        if ((e = hipMalloc(&tiles, sizeof(int64_t) * (size_t)std::max<int64_t>(ntiles, 1))) != hipSuccess) { fail(e, "hipMalloc"); break; }
        if ((e = hipMalloc(&cursor, sizeof(unsigned long long) * (size_t)n)) != hipSuccess) { fail(e, "hipMalloc"); break; }
        if ((e = hipMalloc(&keys, sizeof(uint64_t) * (size_t)std::max<int64_t>(nnz, 1))) != hipSuccess) { fail(e, "hipMalloc(keys)"); break; }
        if ((e = hipMalloc(&flag, sizeof(uint32_t))) != hipSuccess) { fail(e, "hipMalloc"); break; }
        hipMemsetAsync(outdeg, 0, sizeof(uint32_t) * (size_t)n, b->stream);
        hipMemsetAsync(flag, 0, sizeof(uint32_t), b->stream);
        const unsigned gg = grid_of(nnz, 256, 65536);
        hipLaunchKernelGGL(gen_count_kernel, dim3(gg), dim3(256), 0, b->stream, s, outdeg);
        hipLaunchKernelGGL(scan_tile_sums_kernel, dim3((unsigned)ntiles), dim3(kScanThreads), 0, b->stream, outdeg, n,
                           tiles);
        hipLaunchKernelGGL(scan_tile_offsets_kernel, dim3(1), dim3(kScanThreads), 0, b->stream, tiles, ntiles);
        hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)ntiles), dim3(kScanThreads), 0, b->stream, outdeg, n,
                           tiles, b->rowptr);
        hipMemcpyAsync(cursor, b->rowptr, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, b->stream);
        hipLaunchKernelGGL(gen_fill_kernel, dim3(gg), dim3(256), 0, b->stream, s, cursor, keys);
        const unsigned rg = grid_of(n, 1, 1 << 20);
        if (b->pair)
            hipLaunchKernelGGL(gen_sort_finalize_kernel<true>, dim3(rg), dim3(kSortThreads), 0, b->stream, s, b->pk,
                               b->rowptr, keys, b->col, b->w, b->w2, b->dend, flag);
        else
            hipLaunchKernelGGL(gen_sort_finalize_kernel<false>, dim3(rg), dim3(kSortThreads), 0, b->stream, s, b->pk,
                               b->rowptr, keys, b->col, b->w, b->w2, b->dend, flag);
        if ((e = hipGetLastError()) != hipSuccess) { fail(e, "generator launch"); break; }
        uint32_t nlong = 0;
        hipMemcpyAsync(&nlong, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream);
        if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) { fail(e, "generator"); break; }
        if (nlong) {
            set_error("%u presynaptic rows exceed %d synapses (generator row sort limit)", nlong, kMaxRow);
            st = DTB_UNSUPPORTED;
        }
    } while (0);
    if (outdeg) hipFree(outdeg);
    if (tiles) hipFree(tiles);
    if (cursor) hipFree(cursor);
    if (keys) hipFree(keys);
    if (flag) hipFree(flag);
    return st;
}

int gen_members(Block* b, const dtb_member_spec* mem, int32_t M, int32_t max_delay, float w_lo, float w_hi,
                uint64_t seed, const uint64_t* cdf, int32_t R0)
{
    if (R0 <= 0) R0 = M;
    DTB_REQUIRE(M % R0 == 0, "members must be whole replicas of the source table's %d members", R0);
    const int64_t n = b->n;
    const int64_t nnz = b->nnz;
    DTB_REQUIRE(n <= (1ll << 26), "generated networks are limited to 2^26 neurons");
    DTB_REQUIRE(max_delay <= 63, "generated networks are limited to delays <= 63");
    std::vector<GenMember> hm((size_t)M);
    int64_t off = 0, so = 0;
    bool inter = false;
    for (int32_t m = 0; m < M; ++m) {
        const int32_t K = mem[m].in_degree + mem[m].inter_degree;
        hm[(size_t)m] = GenMember{off, mem[m].n, mem[m].n_exc, so, K, mem[m].in_degree};
        inter = inter || mem[m].inter_degree > 0;
        off += mem[m].n;
        so += mem[m].n * (int64_t)K;
    }
    DTB_REQUIRE(!inter || cdf != nullptr, "inter-member synapses need a source table");
    GenMember* dmem = nullptr;
    uint64_t* dcdf = nullptr;
    DTB_TRY_HIP(hipMalloc(&dmem, sizeof(GenMember) * (size_t)M));
    DTB_TRY_HIP(hipMemcpy(dmem, hm.data(), sizeof(GenMember) * (size_t)M, hipMemcpyHostToDevice));
    if (inter) {
        std::vector<uint64_t> hc((size_t)R0 * (size_t)R0);
        hipError_t e = hipMemcpy(hc.data(), cdf, sizeof(uint64_t) * hc.size(), hipMemcpyDefault);
        int bad = e != hipSuccess;
        for (int32_t m = 0; m < R0 && !bad; ++m) {
            const uint64_t* row = hc.data() + (size_t)m * R0;
            for (int32_t j = 1; j < R0; ++j) bad |= row[j] < row[j - 1];
            bad |= row[R0 - 1] != (1ull << 32);
        }
        if (!bad) e = hipMalloc(&dcdf, sizeof(uint64_t) * hc.size());
        if (!bad && e == hipSuccess) e = hipMemcpy(dcdf, hc.data(), sizeof(uint64_t) * hc.size(), hipMemcpyHostToDevice);
        if (bad || e != hipSuccess) {
            hipFree(dmem);
            if (dcdf) hipFree(dcdf);
            if (bad) {
                set_error("source tables must be non-decreasing rows ending at 2^32");
                return DTB_INVALID_PARAMETER;
            }
            set_error("source table upload failed: %s", hipGetErrorString(e));
            return DTB_RUNTIME_ERROR;
        }
    }
    GenSpec s{n, nnz, M, max_delay, w_lo, w_hi - w_lo, (uint32_t)seed, (uint32_t)(seed >> 32), dmem, dcdf, R0};
    int st = gen_run(b, s, max_delay);
    hipFree(dmem);
    if (dcdf) hipFree(dcdf);
    return st;
}

int gen_fixed_indegree(Block* b, const dtb_fixed_indegree_spec* spec)
{
    dtb_member_spec m{spec->n, spec->n_exc, spec->in_degree, 0};
    return gen_members(b, &m, 1, spec->max_delay, spec->w_lo, spec->w_hi, spec->seed, nullptr, 1);
}

} // namespace dtb

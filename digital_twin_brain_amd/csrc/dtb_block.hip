// dtb_block.hip -- the per-timestep network step of the Digital Twin Brain
// on MI355X (gfx950), and its C ABI (include/dtb.h).
//
// One step n (t_n = t_{n-1} + dt, accumulated in double) is two kernels:
//
//   neuron_step_kernel  (one lane per neuron, coalesced SoA)
//     J_{n-1} = Jp_{n-1} + acc                     acc cleared
//     I_syn   = sum_k g_k (V_uk - V_{n-1}) J_k     models/block.py:83-87
//     V_n     = LIF Euler + refractory + clamp     models/block.py:89-111
//     s_n     = noise | threshold                  models/block.py:136-138
//     Jp_n    = J_{n-1} * exp(-dt/tau)             models/block.py:71
//     spiking ids appended to the step's spike list (ballot + LDS prefix,
//     one global atomic per workgroup); threshold counts and V sums per
//     sub-population (wavefront reductions)
//
//   deliver_kernel  (one wavefront per (spiking row, delay) segment)
//     for d = 1..D, every row j in the spike list of step n-d+1:
//       acc[post] += w over the row's delay-d segment   models/block.py:72-73
//     the event-driven form of W.s with per-synapse delays (d = 1 is the
//     reference's same-step delivery).
//
// Delays live on the spike side: the last D+1 spike lists are kept, and each
// CSR row is sorted by delay with per-row segment ends, so every arrival of
// a step lands in one [n] accumulator (8 B/neuron) that stays cache-resident
// instead of a [D][n] ring.  J_n = Jp_n + acc once the step's delivery ran.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <tuple>

#include "dtb_block.hpp"
#include "dtb_group.hpp"

namespace dtb {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int block_alloc(Block* b, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? DTB_OUT_OF_MEMORY : DTB_RUNTIME_ERROR;
    }
    b->allocs.push_back(*p);
    return DTB_OK;
}

template <typename T>
static int alloc_n(Block* b, T** p, int64_t count)
{
    return block_alloc(b, reinterpret_cast<void**>(p), sizeof(T) * (size_t)std::max<int64_t>(count, 1));
}

constexpr int kMaxPsets = 64; // parameter sets held in LDS by neuron_step_kernel

__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ---------------------------------------------------------------- kernels
// Spike lists.  The list of a step is kSubLists sub-lists: workgroup b of the
// neuron kernel appends to sub-list b % kSubLists (capacity spk_cap = the
// neurons of those workgroups, an exact bound), so each sub-list counter takes
// ~1/kSubLists of the workgroups' reservation atomics: returning atomics on
// one address serialise at ~11 ns each, which made one shared counter the
// neuron kernel's bottleneck.  Counters sit kCntStride words apart.
constexpr int kSubLists = 32;
constexpr int kCntStride = 32;
constexpr int kMaxEntries = DTB_MAX_DELAY * kSubLists; // (delay, sub-list) entries of a delivery
constexpr int kCursorStride = 32;    // one 128-byte line per bucket cursor / overflow flag / list count

// Segment lists (bucketed delivery with delays, D > 1).  A spike of step s
// is delivered over D steps, one (row, delay) segment per step, and reading a
// segment's bounds from rowptr and dend at each of those steps costs two
// random cache lines per segment: for short segments (in-degree / D of a
// dozen synapses or less) several times the synapse data itself.  Instead,
// right after the neuron kernel, segment_emit_kernel reads each spike's row
// start and segment ends once and writes its non-empty segments, packed as
// (row start + offset) << 24 | length, into the list of the step that
// delivers them: segment d of a spike of step s goes to step s + d - 1, list
// (s + d - 1) mod D, reserved by one returning atomic per wavefront and
// delay.  The stage kernel of step s then reads list s mod D with coalesced
// 8-byte loads, and the accumulation kernel clears its count.  A list over
// its capacity raises its overflow flag, and that step reads the spike lists
// as before; so do the first D - 1 steps after the lists were (re)started,
// whose lists lack the segments of spikes before the restart.
struct SegArgs {
    uint64_t* seg; // [D][cap] (NULL: segment lists off)
    uint32_t* cnt; // [D + 1][kCursorStride]: entries, overflow flag; entry D: first complete step (int64)
    int64_t cap;   // entries per list
};

__device__ inline int64_t seg_from(const SegArgs& g, int32_t D)
{
    return *reinterpret_cast<const int64_t*>(g.cnt + (int64_t)D * kCursorStride);
}


__host__ __device__ inline int64_t spk_cap_of(int64_t n)
{
    const int64_t nwg = (n + 255) / 256;
    return (nwg + kSubLists - 1) / kSubLists * 256;
}

struct NeuronArgs {
    int64_t n;
    int32_t D, S;
    const float4* __restrict__ g;
    const float4* __restrict__ vu;
    const float4* __restrict__ decay;
    const float* __restrict__ gL;
    const float* __restrict__ vL;
    const float* __restrict__ dtC;
    const float* __restrict__ vth;
    const float* __restrict__ vrst;
    const float* __restrict__ iext;
    const float* __restrict__ tref;
    const float* __restrict__ noisep;
    const int8_t* __restrict__ jexp; // [n][4] input exponents (read for the external drive)
    uint8_t* __restrict__ src;       // [n] forced-spike source flags (NULL: no sources)
    const int32_t* __restrict__ pop;
    const uint8_t* __restrict__ pset; // parameter set of each neuron (PSET kernels)
    const float* __restrict__ ptab;   // [npsets][16]: V_u[4], decay[4], g_L, V_L, dt/C, V_th, V_reset, I_ext
    int32_t npsets;
    float* __restrict__ V;
    float4* __restrict__ Jp;
    int32_t* __restrict__ ref_end;
    int32_t* __restrict__ last_step;
    double* __restrict__ t_last;
    uint8_t* __restrict__ flags;
    void* __restrict__ acc;
    int32_t* __restrict__ spk;     // [L][kSubLists][spk_cap]
    uint32_t* __restrict__ spk_cnt; // [L][kSubLists] (kCntStride apart)
    int64_t spk_cap;
    int32_t L;         // spike-list slots (D + 1)
    int32_t clear_acc; // direct delivery accumulates into acc, so it is cleared after reading
    const StepCtl* __restrict__ ctl;
    const double* __restrict__ ttab;
    double dt;
    uint64_t seed;
    int32_t noise_mode;
    float noise_rate;
    const float* __restrict__ uniforms;
    int32_t isolated;
    uint32_t* __restrict__ rec_freq;
    float* __restrict__ rec_vsum;
    ExtDrive ext;
};

// Steps until the accumulated time reaches t_spike + T_ref: the same double
// additions the host uses for t, so `step >= ref_end` equals block.py's
// `t >= t_last + T_ref` (models/block.py:100).
__device__ inline int32_t refractory_end(int64_t step, double t_spike, float tref, double dt)
{
    const double lim = t_spike + (double)tref;
    double tt = t_spike;
    int64_t c = 0;
    while (tt < lim && c < (1 << 22)) {
        tt += dt;
        ++c;
    }
    int64_t e = step + c;
    return (int32_t)(e > 0x7fffffff ? 0x7fffffff : e);
}

// Streaming loads.  Neuron state, synapse targets and weights and staged
// events are each read once per step from arrays far larger than L2, so they
// are loaded non-temporally: they then do not evict the lines that are
// reused (staging partitions being filled, spike lists, row pointers).  On
// the 5e6-neuron, 5e9-synapse block this shortens the step by ~10%.
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t nt_u2 __attribute__((ext_vector_type(2)));
typedef uint32_t nt_u4 __attribute__((ext_vector_type(4)));
__device__ inline float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }
__device__ inline uint32_t ld_stream(const uint32_t* p) { return __builtin_nontemporal_load(p); }
__device__ inline float2 ld_stream(const float2* p)
{
    const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(p));
    return make_float2(v.x, v.y);
}
__device__ inline float4 ld_stream(const float4* p)
{
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline uint2 ld_stream(const uint2* p)
{
    const nt_u2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u2*>(p));
    return make_uint2(v.x, v.y);
}
__device__ inline uint4 ld_stream(const uint4* p)
{
    const nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

// One neuron's step (models/block.py:89-114,136-138): J += this step's
// arrivals, I_syn from the stored V, the Euler step or the refractory hold,
// threshold, noise spike, J decayed for the next step.  GACC: the arrivals
// are loaded from a.acc (and cleared there for direct delivery); else they
// are `din`.  Every load of the neuron precedes its first store: the compiler
// makes a load that follows a store to another array wait for that store.
// PSET: the constants shared by groups of neurons (reversal potentials,
// synaptic decays, leak, threshold, reset, I_ext) come from a small table in
// LDS indexed by a 1-byte set id instead of 56 B of per-neuron arrays.
// One neuron's membrane step (models/block.py:83-111): I_syn from the
// conductances and J (arrivals added), Euler step of V, refractory hold,
// threshold and clamp.  Shared by every kernel that updates neurons, so all
// of them compute bit-identical values.
__device__ __forceinline__ void lif_core(float V, const float4& J, const float4& g, const float4& vu, float gL,
                                         float vL, float dtC, float vth, float vrst, float iext, bool not_ref,
                                         bool& act, float& vout)
{
    float isyn = g.x * (vu.x - V) * J.x;
    isyn += g.y * (vu.y - V) * J.y;
    isyn += g.z * (vu.z - V) * J.z;
    isyn += g.w * (vu.w - V) * J.w;
    const float main_part = -gL * (V - vL);
    const float vn = V + dtC * (main_part + isyn + iext);
    const float vi = not_ref ? vn : vrst;
    act = vi >= vth;
    vout = vi > vth ? vth : vi;
}

// J decayed to the next step (models/block.py:71)
__device__ __forceinline__ float4 lif_decay(const float4& J, const float4& dc)
{
    return make_float4(J.x * dc.x, J.y * dc.y, J.z * dc.z, J.w * dc.w);
}

template <bool PAIR, bool PSET, bool GACC>
__device__ __forceinline__ void neuron_update(const NeuronArgs& a, int64_t i, int64_t step, double t, int32_t k,
                                              float4 din, const float* s_tab, bool& act, bool& spike, float& vout)
{
    // the neuron kernel streams the state (read once per step from arrays far
    // larger than L2); the persistent launch re-reads its small block's state
    // every step, so it keeps it cached
    const float V = GACC ? ld_stream(a.V + i) : a.V[i];
    float4 J = GACC ? ld_stream(a.Jp + i) : a.Jp[i];
    float4 d = din; // this step's arrivals per receptor (compact: AMPA, NMDA share one sum, GABAs the other)
    if (GACC) {
        if (PAIR) {
            d = ld_stream(reinterpret_cast<const float4*>(a.acc) + i);
        } else {
            const float2 d2 = ld_stream(reinterpret_cast<const float2*>(a.acc) + i);
            d = make_float4(d2.x, d2.x, d2.y, d2.y);
        }
    }
    const float4 g = GACC ? ld_stream(a.g + i) : a.g[i];
    const int32_t rend = a.ref_end[i];
    const float pr = a.noise_mode == DTB_NOISE_PER_NEURON ? a.noisep[i] : a.noise_rate;
    const float un = a.noise_mode == DTB_NOISE_UNIFORMS ? a.uniforms[(int64_t)k * a.n + i] : 0.0f;
    const char2 ee = a.ext.n ? *reinterpret_cast<const char2*>(a.jexp + 4 * i) : make_char2(0, 0);
    const uint8_t ss = a.src ? a.src[i] : (uint8_t)0;
    J.x += d.x; J.y += d.y; J.z += d.z; J.w += d.w;
    float4 vu, dc;
    float gL, vL, dtC, vth, vrst, iext;
    if (PSET) {
        const float* T = s_tab + 16 * a.pset[i];
        vu = make_float4(T[0], T[1], T[2], T[3]);
        dc = make_float4(T[4], T[5], T[6], T[7]);
        gL = T[8]; vL = T[9]; dtC = T[10]; vth = T[11]; vrst = T[12]; iext = T[13];
    } else {
        vu = a.vu[i];
        dc = a.decay[i];
        gL = a.gL[i]; vL = a.vL[i]; dtC = a.dtC[i]; vth = a.vth[i]; vrst = a.vrst[i]; iext = a.iext[i];
    }
    lif_core(V, J, g, vu, gL, vL, dtC, vth, vrst, iext, step >= (int64_t)rend, act, vout);
    if (ss & 1) {
        // a source's threshold spike comes from the schedule (models/block.py:104-107);
        // a frozen one (an input block's neuron) also keeps its V
        act = (ss & 4) != 0;
        if (ss & 2) vout = V;
    }
    if (GACC && a.clear_acc) {
        if (PAIR) reinterpret_cast<float4*>(a.acc)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        else reinterpret_cast<float2*>(a.acc)[i] = make_float2(0.f, 0.f);
    }
    a.V[i] = vout;
    if (ss & 4) a.src[i] = ss & 3;
    if (act) {
        a.ref_end[i] = refractory_end(step, t, a.tref[i], a.dt);
        a.last_step[i] = (int32_t)step;
        a.t_last[i] = t;
    }
    const float u = a.noise_mode == DTB_NOISE_UNIFORMS ? un : noise_uniform(a.seed, (uint64_t)step, (uint64_t)i);
    spike = ((u < pr) && !(ss & 2)) || (act && !a.isolated);
    float4 jn = lif_decay(J, dc);
    if (a.ext.n) {
        // external arrivals of this step: part of J at the next step, like this step's deliveries
        const float x = (float)ext_count(a.ext, a.seed, (uint64_t)step, (uint64_t)i) * a.ext.w;
        jn.x += ldexpf(x, ee.x);
        jn.y += ldexpf(x, ee.y);
    }
    a.Jp[i] = jn;
    a.flags[i] = (uint8_t)((act ? 1 : 0) | (spike ? 2 : 0));
}

// Per sub-population statistics of a step (threshold spikes, V sum) for the
// 64 neurons of a wavefront: one atomic per wavefront when they share a
// population.
__device__ __forceinline__ void step_stats(const NeuronArgs& a, int32_t k, int32_t p, bool act, float vout)
{
    const uint32_t lane = __lane_id();
    const int32_t p0 = __shfl(p, 0);
    if (__all(p == p0)) {
        if (p0 >= 0) {
            if (a.rec_freq) {
                const uint32_t c = (uint32_t)__popcll(__ballot(act));
                if (lane == 0 && c) atomicAdd(&a.rec_freq[(int64_t)k * a.S + p0], c);
            }
            if (a.rec_vsum) {
                const float v = wave_sum(vout);
                if (lane == 0) unsafeAtomicAdd(&a.rec_vsum[(int64_t)k * a.S + p0], v);
            }
        }
    } else if (p >= 0) {
        if (a.rec_freq && act) atomicAdd(&a.rec_freq[(int64_t)k * a.S + p], 1u);
        if (a.rec_vsum) unsafeAtomicAdd(&a.rec_vsum[(int64_t)k * a.S + p], vout);
    }
}

template <bool PAIR, bool PSET>
__global__ __launch_bounds__(256) void neuron_step_kernel(NeuronArgs a, int32_t k)
{
    __shared__ uint32_t s_wcnt[4];
    __shared__ uint32_t s_base;
    __shared__ float s_tab[kMaxPsets * 16];
    if (PSET) {
        for (int x = threadIdx.x; x < a.npsets * 16; x += 256) s_tab[x] = a.ptab[x];
        __syncthreads();
    }
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t step = a.ctl->nbase + k + 1;
    const double t = a.ttab[k];
    const int32_t slot = (int32_t)(step % a.L);
    // the list of step+1 reuses the slot of step+1-L, whose last delivery was step
    if (blockIdx.x == 0 && threadIdx.x < kSubLists)
        a.spk_cnt[((int64_t)((step + 1) % a.L) * kSubLists + threadIdx.x) * kCntStride] = 0;

    bool act = false, spike = false;
    float vout = 0.0f;
    int32_t p = -1;
    if (i < a.n) {
        p = a.pop[i];
        neuron_update<PAIR, PSET, true>(a, i, step, t, k, make_float4(0.f, 0.f, 0.f, 0.f), s_tab, act, spike, vout);
    }

    // spike compaction: ballot + per-wave prefix, one global atomic per workgroup
    const uint32_t lane = __lane_id();
    const uint32_t wid = threadIdx.x >> 6;
    const uint64_t m = __ballot(spike);
    const uint32_t pre = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wcnt[wid] = (uint32_t)__popcll(m);
    __syncthreads();
    const int64_t sl = (int64_t)slot * kSubLists + blockIdx.x % kSubLists;
    if (threadIdx.x == 0) {
        const uint32_t tot = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
        s_base = tot ? atomicAdd(&a.spk_cnt[sl * kCntStride], tot) : 0u;
    }
    __syncthreads();
    if (spike) {
        uint32_t off = s_base;
        for (uint32_t w = 0; w < wid; ++w) off += s_wcnt[w];
        a.spk[sl * a.spk_cap + off + pre] = (int32_t)i;
    }

    // per sub-population statistics (threshold spikes, V sum)
    if (a.rec_freq || a.rec_vsum) step_stats(a, k, p, act, vout);
}

struct DeliverArgs {
    int64_t n;
    int32_t D, L;
    SynPack pk;
    const int64_t* __restrict__ rowptr;
    const uint16_t* __restrict__ dend; // [n][D] end of each delay segment, relative to the row start
    const uint32_t* __restrict__ col;
    const float* __restrict__ w;
    const float2* __restrict__ w2;
    void* __restrict__ acc;
    const int32_t* __restrict__ spk;
    const uint32_t* __restrict__ spk_cnt;
    int64_t spk_cap;
    const StepCtl* __restrict__ ctl;
    uint64_t* __restrict__ rec_events;
    uint32_t* __restrict__ rec_nspk;
};

// Inclusive scan over the 256 lanes of a workgroup (shuffles within each
// wavefront, then the four wavefront totals through LDS).
__device__ inline uint32_t block_inclusive_scan_256(uint32_t v, uint32_t* wtot)
{
    const uint32_t lane = __lane_id();
    const uint32_t wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(v, off);
        if (lane >= (uint32_t)off) v += y;
    }
    if (lane == kWave - 1) wtot[wid] = v;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wid; ++w) add += wtot[w];
    __syncthreads();
    return v + add;
}

// Inclusive scan over the W wavefronts of a workgroup (W * 64 lanes).
template <int W>
__device__ inline uint32_t block_inclusive_scan_n(uint32_t v, uint32_t* wtot)
{
    const uint32_t lane = __lane_id();
    const uint32_t wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(v, off);
        if (lane >= (uint32_t)off) v += y;
    }
    if (lane == kWave - 1) wtot[wid] = v;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wid; ++w) add += wtot[w];
    __syncthreads();
    return v + add;
}

// Spike-list slot of the step that spiked d - 1 steps before the step whose
// slot is s0 (d <= D < L, so one wrap at most).
__device__ inline int32_t delay_slot(int32_t s0, int32_t d, int32_t L)
{
    const int32_t s = s0 - (d - 1);
    return s < 0 ? s + L : s;
}

// where sub-list entry e = (d - 1) * kSubLists + q starts in spk, for the
// step whose slot is s0
__device__ inline int64_t entry_base(const DeliverArgs& a, int32_t s0, int32_t e)
{
    const int32_t d = e / kSubLists + 1, q = e % kSubLists;
    return ((int64_t)delay_slot(s0, d, a.L) * kSubLists + q) * a.spk_cap;
}

// Segments to deliver at batch step k: entry e = (d - 1) * kSubLists + q holds
// sub-list q of the step that spiked d - 1 steps ago; cum[e + 1] = segments in
// entries <= e.  Workgroups of 256 threads; returns the number of entries.
// Every counter load is issued before the first is waited for (entries past
// E read entry E - 1's counter and are then dropped).
__device__ inline int32_t delivery_entries(const DeliverArgs& a, int32_t k, uint32_t* cum, uint32_t* wtot)
{
    const int64_t step = a.ctl->nbase + k + 1;
    const int32_t s0 = (int32_t)(step % a.L);
    const int32_t E = a.D * kSubLists;
    constexpr int kPer = kMaxEntries / 256;
    uint32_t c[kPer];
    uint32_t sum = 0;
    const int32_t e0 = threadIdx.x * kPer;
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        const int32_t e = min(e0 + x, E - 1);
        const int32_t d = e / kSubLists + 1, q = e % kSubLists;
        c[x] = a.spk_cnt[((int64_t)delay_slot(s0, d, a.L) * kSubLists + q) * kCntStride];
    }
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        c[x] = e0 + x < E ? min(c[x], (uint32_t)a.spk_cap) : 0u;
        sum += c[x];
    }
    uint32_t run = block_inclusive_scan_256(sum, wtot) - sum;
    if (threadIdx.x == 0) cum[0] = 0;
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        run += c[x];
        if (e0 + x < E) cum[e0 + x + 1] = run;
    }
    __syncthreads();
    return E;
}

// entry holding segment g (cum[E] > g): the last e with cum[e] <= g
__device__ inline int32_t entry_of(const uint32_t* cum, int32_t E, uint32_t g)
{
    int32_t lo = 0, hi = E - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (cum[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// G lanes per (row, delay) segment (64, 32 or 16): a wavefront works on 64/G
// segments at once, each lane keeps two synapse loads in flight.
template <bool PAIR, int G>
__global__ __launch_bounds__(256) void deliver_kernel(DeliverArgs a, int32_t k)
{
    __shared__ uint32_t cum[kMaxEntries + 1];
    __shared__ uint32_t wtot[4];
    const int32_t s0 = (int32_t)((a.ctl->nbase + k + 1) % a.L);
    const int32_t E = delivery_entries(a, k, cum, wtot);
    const uint32_t total = cum[E];
    const uint32_t lane = __lane_id();
    const uint32_t sub = lane % G;
    constexpr uint32_t kGroups = kWave / G;
    const uint32_t gid = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * kGroups + lane / G;
    const uint32_t ngroups = ((gridDim.x * blockDim.x) >> 6) * kGroups;
    if (gid == 0 && sub == 0 && a.rec_nspk) a.rec_nspk[k] = cum[kSubLists];
    uint64_t ev = 0;
    int32_t e = 0;
    for (uint32_t r = gid; r < total; r += ngroups) {
        while (r >= cum[e + 1]) ++e; // r only grows, so e only moves forward
        const int32_t d = e / kSubLists + 1;
        const int32_t j = a.spk[entry_base(a, s0, e) + (r - cum[e])];
        const int64_t row = a.rowptr[j];
        int64_t beg, end;
        if (a.D == 1) {
            beg = row;
            end = a.rowptr[j + 1];
        } else {
            const uint16_t* de = a.dend + (int64_t)j * a.D;
            const uint32_t lo = de[d > 1 ? d - 2 : 0], hi = de[d - 1]; // both loads in flight
            beg = row + (d > 1 ? lo : 0u);
            end = row + hi;
        }
        ev += (uint64_t)(end - beg);
        for (int64_t e = beg + sub; e < end; e += 2 * G) {
            const int64_t e2 = e + G;
            const bool two = e2 < end;
            const uint32_t c0 = a.col[e];
            const uint32_t c1 = two ? a.col[e2] : 0u;
            if (PAIR) {
                const float2 w0 = a.w2[e];
                const float2 w1 = two ? a.w2[e2] : make_float2(0.f, 0.f);
                float* d0 = reinterpret_cast<float*>(a.acc) + ((int64_t)(c0 & a.pk.idx_mask) << 2) + ((c0 >> 31) << 1);
                unsafeAtomicAdd(d0, w0.x);
                unsafeAtomicAdd(d0 + 1, w0.y);
                if (two) {
                    float* d1 =
                        reinterpret_cast<float*>(a.acc) + ((int64_t)(c1 & a.pk.idx_mask) << 2) + ((c1 >> 31) << 1);
                    unsafeAtomicAdd(d1, w1.x);
                    unsafeAtomicAdd(d1 + 1, w1.y);
                }
            } else {
                const float w0 = a.w[e];
                const float w1 = two ? a.w[e2] : 0.0f;
                unsafeAtomicAdd(reinterpret_cast<float*>(a.acc) + ((int64_t)(c0 & a.pk.idx_mask) << 1) + (c0 >> 31), w0);
                if (two)
                    unsafeAtomicAdd(reinterpret_cast<float*>(a.acc) + ((int64_t)(c1 & a.pk.idx_mask) << 1) + (c1 >> 31),
                                    w1);
            }
        }
    }
    // per-group event counts -> one atomic per wavefront
#pragma unroll
    for (int off = G; off < kWave; off <<= 1) ev += __shfl_xor(ev, off);
    if (lane == 0 && ev && a.rec_events) atomicAdd((unsigned long long*)&a.rec_events[k], (unsigned long long)ev);
}

template <bool PAIR>
static void launch_deliver(const Block* b, const DeliverArgs& da, int32_t k)
{
    const dim3 g(b->deliver_grid), t(256);
    switch (b->deliver_lanes) {
    case 16: hipLaunchKernelGGL((deliver_kernel<PAIR, 16>), g, t, 0, b->stream, da, k); break;
    case 32: hipLaunchKernelGGL((deliver_kernel<PAIR, 32>), g, t, 0, b->stream, da, k); break;
    default: hipLaunchKernelGGL((deliver_kernel<PAIR, 64>), g, t, 0, b->stream, da, k); break;
    }
}

// ---------------------------------------------------------------------------
// Bucketed delivery.  Random float atomics on MI355X saturate near 2e10/s
// whatever the footprint, scope or XCD placement, so a large step's events
// are instead sorted by destination bucket (bsize consecutive neurons, one
// LDS slice of up to 160 KB; bsize is chosen at set-up so that the buckets
// fill the CUs in as few full waves as the LDS allows) and accumulated in LDS:
//   phase 1 (deliver_stage_kernel): each workgroup expands 256 (row, delay)
//     segments at a time into rounds of 256*kEPT events held in registers;
//     an LDS histogram ranks every event inside its bucket, one global atomic
//     per touched bucket reserves space, and events are written to that
//     bucket's staging array (bucket-major, contiguous per bucket).
//   phase 2 (deliver_accum_kernel): one workgroup per bucket sums its staged
//     events into the LDS slice and writes the bucket's slice of acc.
// A bucket whose staging array is full takes the rest of its events through
// global atomics into acc2, which phase 2 then folds in.
template <bool PAIR>
struct BucketCfg {
    static constexpr int kChan = PAIR ? 4 : 2; // floats per neuron in acc
};
constexpr int kMaxBucketBits = 15;   // neurons per bucket < 2^15 (local index in staged words)
constexpr int kStageParts = 8;       // staging partitions per bucket: workgroup w reserves in part
                                     // w % kStageParts, so each cursor takes 1/8 of the (serialising)
                                     // same-address reservation atomics
constexpr int kMaxBucketsLds = 8192; // n <= 2^26 neurons
constexpr int kRankBits = 13;        // rank of an event inside its bucket within a round (< 8192)
constexpr uint32_t kPartial = 0xffffffffu; // a bucket whose round reservation does not fit entirely
constexpr int kStageStaticLds = 4096; // static LDS of the stage kernel (segment table), bytes
constexpr int kAccumStaticLds = 256;  // static LDS of the accumulation kernel (partition counts), bytes

struct StageArgs {
    DeliverArgs d;
    int32_t nbk;      // buckets
    int32_t bsize;    // neurons per bucket (< 2^15)
    uint64_t bmagic;  // bucket = (post * bmagic) >> bshift == post / bsize for post < 2^26
    int32_t bshift;
    int64_t cap;      // staged events per bucket partition
    uint32_t* cursor; // [nbk][kStageParts] (kCursorStride apart)
    void* stage;      // [nbk][kStageParts][cap] uint2 (compact) / uint4 (pair)
    void* acc2;       // overflow accumulator [n]
    uint32_t* ovf;    // [nbk] overflow flags
    SegArgs sg;       // segment lists
    int32_t k;        // batch step (the accumulation kernel clears the step's segment list)
};


// Stage-kernel pieces.  Event e of a chunk lives in segment sg at offset
// e - segpre[sg]; loc packs (sg << 24) | offset, or ~0 past the chunk's end.
// sg is the last segment with segpre[sg] <= e (empty segments share their
// start with the next one), found by a fixed-depth bisection.  The step loop
// is outermost: each step issues the EPT independent LDS reads of a lane
// back to back and waits once, so a round costs log2(SEG) + 1 LDS latencies
// rather than EPT times that.
template <int SEG, int EPT>
__device__ inline void stage_locate(const uint32_t* segpre, uint32_t r0, uint32_t E, int t, uint32_t* loc)
{
    static_assert((SEG & (SEG - 1)) == 0, "SEG is a power of two");
    uint32_t sg[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) sg[m] = 0;
#pragma unroll
    for (int step = SEG / 2; step >= 1; step >>= 1) {
        uint32_t v[EPT];
#pragma unroll
        for (int m = 0; m < EPT; ++m) v[m] = segpre[sg[m] + step];
#pragma unroll
        for (int m = 0; m < EPT; ++m) sg[m] += v[m] <= r0 + (uint32_t)m * 256 + t ? step : 0;
    }
    uint32_t b[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) b[m] = segpre[sg[m]];
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const uint32_t e = r0 + (uint32_t)m * 256 + t;
        loc[m] = e < E ? ((sg[m] << 24) | (e - b[m])) : 0xffffffffu;
    }
}

// all of a lane's loads in flight together
template <bool PAIR, int EPT>
__device__ inline void stage_load(const DeliverArgs& a, const int64_t* segbeg, const uint32_t* loc, uint32_t* cw,
                                  float* w0, float* w1)
{
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        cw[m] = 0u;
        w0[m] = 0.0f;
        w1[m] = 0.0f;
        if (loc[m] != 0xffffffffu) {
            const int64_t idx = segbeg[loc[m] >> 24] + (loc[m] & 0xffffffu);
            cw[m] = ld_stream(a.col + idx);
            if (PAIR) {
                const float2 wv = ld_stream(a.w2 + idx);
                w0[m] = wv.x;
                w1[m] = wv.y;
            } else {
                w0[m] = ld_stream(a.w + idx);
            }
        }
    }
}

// staged event at position pos of bucket bk's partition, or through the
// overflow accumulator when the partition is full
template <bool PAIR>
__device__ inline void stage_put(const StageArgs& s, int32_t part, uint32_t bk, uint64_t pos, uint32_t pk, float w0,
                                 float w1)
{
    if (pos < (uint64_t)s.cap) {
        const int64_t at = ((int64_t)bk * kStageParts + part) * s.cap + (int64_t)pos;
        if (PAIR)
            reinterpret_cast<uint4*>(s.stage)[at] = make_uint4(pk, __float_as_uint(w0), __float_as_uint(w1), 0u);
        else
            reinterpret_cast<uint2*>(s.stage)[at] = make_uint2(pk, __float_as_uint(w0));
    } else {
        const int64_t post = (int64_t)bk * s.bsize + (pk & 0x7fffffffu);
        const uint32_t kind = pk >> 31;
        if (PAIR) {
            float* dst = reinterpret_cast<float*>(s.acc2) + (post << 2) + (kind << 1);
            unsafeAtomicAdd(dst, w0);
            unsafeAtomicAdd(dst + 1, w1);
        } else {
            unsafeAtomicAdd(reinterpret_cast<float*>(s.acc2) + (post << 1) + kind, w0);
        }
        s.ovf[(size_t)bk * kCursorStride] = 1u;
    }
}

// SEG (64, 128 or 256) segments per chunk: fewer give more, smaller chunks
// (parallelism when few neurons spike), more amortise the chunk prologue.
// EPT events per lane per round (16 or 32): fewer use fewer registers (more
// resident waves), more take fewer staging reservations.  The bucket
// histogram is dynamic LDS of nbk words, so small bucket counts leave room for
// more resident workgroups.  VAR selects how events reach the staging arrays:
//   0  one reservation per (round, bucket); each lane stores its own events;
//   1  one reservation per (chunk, bucket): a histogram pass over the chunk's
//      targets, then a second pass that stores (targets read twice);
//   2  one reservation per (round, bucket), the round's events first sorted
//      by bucket in LDS so that consecutive lanes store each bucket's run
//      contiguously (whole cache lines instead of scattered 8-byte stores).
// Four workgroups per CU at 16 events per lane: registers are capped at 128
// per lane and the dynamic LDS holds only what the block needs (entries of
// its delays, its buckets, the sorted round), so a launch of up to 4 x CUs
// workgroups is resident at once instead of leaving a tail wave.  (At 32
// events per lane that cap would spill; two per CU there.)
template <bool PAIR, int SEG, int EPT, int VAR>
__global__ __launch_bounds__(256, EPT <= 16 ? 4 : 2) void deliver_stage_kernel(StageArgs s, int32_t k)
{
    static_assert(SEG >= 64 && SEG <= 256, "segment index is packed in 8 bits");
    static_assert(256 * EPT <= (1 << 13), "rank inside a round must fit kRankBits");
    __shared__ uint32_t segpre[SEG + 1];
    __shared__ int64_t segbeg[SEG];
    extern __shared__ __align__(16) uint32_t dyn[];
    uint32_t* cum = dyn;                                // [D * kSubLists + 1]
    uint32_t* hist = dyn + s.d.D * kSubLists + 1;       // [nbk]
    __shared__ uint32_t wtot[4];
    const DeliverArgs& a = s.d;
    const int32_t s0 = (int32_t)((a.ctl->nbase + k + 1) % a.L);
    const int t = threadIdx.x;
    const int32_t part = (int32_t)(blockIdx.x % kStageParts);
    for (int b = t; b < s.nbk; b += 256) hist[b] = 0;
    // this step's segments: its segment list when complete, else the spike
    // lists of the last D steps (the same for every workgroup)
    bool use_seg = false;
    uint32_t total = 0;
    int64_t seg0 = 0;
    if (s.sg.seg) {
        const int64_t step = a.ctl->nbase + k + 1;
        const int32_t ls = (int32_t)(step % a.D);
        const uint32_t* c = s.sg.cnt + (int64_t)ls * kCursorStride;
        use_seg = step >= seg_from(s.sg, a.D) && c[1] == 0u;
        total = use_seg ? c[0] : 0u; // no overflow flag: every entry is in the list
        seg0 = (int64_t)ls * s.sg.cap;
    }
    int32_t NE = 0;
    if (!use_seg) {
        NE = delivery_entries(a, k, cum, wtot);
        total = cum[NE];
        if (blockIdx.x == 0 && t == 0 && a.rec_nspk) a.rec_nspk[k] = cum[kSubLists];
    } else if (blockIdx.x == 0 && t < 64 && a.rec_nspk) {
        // the step's spikes: the sum of its sub-list counts
        const int32_t slot = (int32_t)((a.ctl->nbase + k + 1) % a.L);
        uint32_t c = t < kSubLists ? min(a.spk_cnt[((int64_t)slot * kSubLists + t) * kCntStride], (uint32_t)a.spk_cap)
                                   : 0u;
        for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
        if (t == 0) a.rec_nspk[k] = c;
    }
    // The step's segments in kper * gridDim chunks of equal size (at most
    // SEG), dealt round-robin: every workgroup gets kper chunks, and each
    // gets chunks from all delays (segment lengths vary with the delay).
    const uint64_t G = gridDim.x;
    const uint64_t kper = ((uint64_t)total + G * SEG - 1) / (G * SEG);
    const uint64_t nch = G * kper;
    const uint64_t cs = nch ? ((uint64_t)total + nch - 1) / nch : 0;
    uint64_t ev_wg = 0;
    for (uint64_t c = blockIdx.x; c < nch && c * cs < total; c += G) {
        // segment lengths of this chunk -> LDS prefix
        const uint32_t c0 = (uint32_t)(c * cs);
        const uint32_t g_end = (uint32_t)min((uint64_t)total, c * cs + cs);
        const uint32_t g = c0 + t;
        uint32_t len = 0;
        int64_t beg = 0;
        if (t < SEG && g < g_end && use_seg) {
            const uint64_t v = s.sg.seg[seg0 + g];
            beg = (int64_t)(v >> 24);
            len = (uint32_t)(v & 0xffffffu);
        } else if (t < SEG && g < g_end) {
            const int32_t e = entry_of(cum, NE, g);
            const int32_t d = e / kSubLists + 1;
            const int32_t j = a.spk[entry_base(a, s0, e) + (g - cum[e])];
            const int64_t row = a.rowptr[j];
            int64_t end;
            if (a.D == 1) {
                beg = row;
                end = a.rowptr[j + 1];
            } else {
                const uint16_t* de = a.dend + (int64_t)j * a.D;
                const uint32_t lo = de[d > 1 ? d - 2 : 0], hi = de[d - 1]; // both loads in flight
                beg = row + (d > 1 ? lo : 0u);
                end = row + hi;
            }
            len = (uint32_t)(end - beg);
        }
        const uint32_t inc = block_inclusive_scan_256(len, wtot);
        if (t < SEG) {
            segbeg[t] = beg;
            segpre[t + 1] = inc;
        }
        if (t == 0) segpre[0] = 0;
        __syncthreads();
        const uint32_t E = segpre[SEG];
        ev_wg += E;
        if constexpr (VAR == 0) {
            for (uint32_t r0 = 0; r0 < E; r0 += 256 * EPT) {
                uint32_t br[EPT], pk_[EPT]; // br = bucket << kRankBits | rank
                float w0[EPT], w1[EPT];
                uint32_t loc[EPT];
                stage_locate<SEG, EPT>(segpre, r0, E, t, loc);
                uint32_t cw[EPT];
                stage_load<PAIR, EPT>(a, segbeg, loc, cw, w0, w1);
                // bucket and rank inside the bucket within this round
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    br[m] = 0xffffffffu;
                    if (loc[m] != 0xffffffffu) {
                        const uint32_t post = cw[m] & a.pk.idx_mask;
                        const uint32_t bk = (uint32_t)(((uint64_t)post * s.bmagic) >> s.bshift);
                        pk_[m] = (post - bk * (uint32_t)s.bsize) | (cw[m] & 0x80000000u);
                        br[m] = (bk << kRankBits) | atomicAdd(&hist[bk], 1u);
                    }
                }
                __syncthreads();
                // reserve space in each touched bucket; hist[b] becomes the base
                for (int b = t; b < s.nbk; b += 256) {
                    const uint32_t h = hist[b];
                    hist[b] = h ? atomicAdd(&s.cursor[((size_t)b * kStageParts + part) * kCursorStride], h) : 0u;
                }
                __syncthreads();
                uint32_t hb[EPT]; // all reads of the reserved bases before the stores
#pragma unroll
                for (int m = 0; m < EPT; ++m) hb[m] = hist[br[m] == 0xffffffffu ? 0u : br[m] >> kRankBits];
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    if (br[m] == 0xffffffffu) continue;
                    const uint32_t bk = br[m] >> kRankBits;
                    stage_put<PAIR>(s, part, bk, (uint64_t)hb[m] + (br[m] & ((1u << kRankBits) - 1)), pk_[m], w0[m],
                                    w1[m]);
                }
                __syncthreads();
                for (int b = t; b < s.nbk; b += 256) hist[b] = 0;
                __syncthreads();
            }
        } else if constexpr (VAR == 2) {
            uint32_t* pre = hist + s.nbk; // [nbk] first LDS slot of each bucket in the round
            // [nbk] staging index of the round's slot 0 for that bucket when the
            // bucket's whole reservation fits its partition, else kPartial
            uint32_t* off = pre + s.nbk;
            // sorted round (+ a spare slot): (packed target, weight 0) pairs, 8-byte aligned, then weight 1
            uint2* bev = reinterpret_cast<uint2*>(off + s.nbk + ((reinterpret_cast<uintptr_t>(off + s.nbk) >> 2) & 1));
            uint32_t* bw1 = reinterpret_cast<uint32_t*>(bev + 256 * EPT + 1);
            const int per = (s.nbk + 255) / 256;
            for (uint32_t r0 = 0; r0 < E; r0 += 256 * EPT) {
                uint32_t br[EPT], pk_[EPT];
                float w0[EPT], w1[EPT];
                uint32_t loc[EPT];
                stage_locate<SEG, EPT>(segpre, r0, E, t, loc);
                uint32_t cw[EPT];
                stage_load<PAIR, EPT>(a, segbeg, loc, cw, w0, w1);
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    br[m] = 0xffffffffu;
                    if (loc[m] != 0xffffffffu) {
                        const uint32_t post = cw[m] & a.pk.idx_mask;
                        const uint32_t bk = (uint32_t)(((uint64_t)post * s.bmagic) >> s.bshift);
                        pk_[m] = (post - bk * (uint32_t)s.bsize) | (cw[m] & 0x80000000u);
                        br[m] = (bk << kRankBits) | atomicAdd(&hist[bk], 1u);
                    }
                }
                __syncthreads();
                // bucket offsets inside the round (exclusive prefix) and staging reservations
                const int b0 = t * per;
                uint32_t sum = 0;
                for (int x = 0; x < per; ++x)
                    if (b0 + x < s.nbk) sum += hist[b0 + x];
                uint32_t run = block_inclusive_scan_256(sum, wtot) - sum;
                for (int x = 0; x < per; ++x) {
                    const int b = b0 + x;
                    if (b >= s.nbk) break;
                    const uint32_t h = hist[b];
                    pre[b] = run;
                    const uint32_t base =
                        h ? atomicAdd(&s.cursor[((size_t)b * kStageParts + part) * kCursorStride], h) : 0u;
                    // modulo 2^32, off + slot is the event's staging index (the
                    // staging array holds fewer than 2^32 events); a partly full
                    // reservation keeps the per-event check of stage_put
                    off[b] = (uint64_t)base + h <= (uint64_t)s.cap
                                 ? (uint32_t)(((uint64_t)b * kStageParts + part) * s.cap) + base - run
                                 : kPartial;
                    hist[b] = base - run; // position in the partition of the round's slot 0, for kPartial
                    run += h;
                }
                __syncthreads();
                // scatter into the sorted round: all reads of pre[] first, then
                // the writes (a lane without an event writes the spare slot)
                uint32_t jw[EPT];
#pragma unroll
                for (int m = 0; m < EPT; ++m) jw[m] = pre[br[m] == 0xffffffffu ? 0u : br[m] >> kRankBits];
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const bool ok = br[m] != 0xffffffffu;
                    const uint32_t bk = br[m] >> kRankBits;
                    const uint32_t j = ok ? jw[m] + (br[m] & ((1u << kRankBits) - 1)) : 256u * EPT;
                    // the bucket rides in bits 15..30 of the packed target
                    bev[j] = make_uint2(pk_[m] | (bk << kMaxBucketBits), __float_as_uint(w0[m]));
                    if (PAIR) bw1[j] = __float_as_uint(w1[m]);
                }
                __syncthreads();
                // consecutive lanes store consecutive events of a bucket; the
                // LDS reads of EPT events are issued before their stores
                const uint32_t cnt = min(E - r0, (uint32_t)(256 * EPT));
                uint32_t v[EPT], x0[EPT], x1[EPT], at[EPT];
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const uint32_t j = (uint32_t)m * 256 + t;
                    const uint2 e = bev[j];
                    v[m] = e.x;
                    x0[m] = e.y;
                    x1[m] = PAIR ? bw1[j] : 0u;
                }
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const uint32_t j = (uint32_t)m * 256 + t;
                    at[m] = off[j < cnt ? (v[m] >> kMaxBucketBits) & 0xffffu : 0u];
                }
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const uint32_t j = (uint32_t)m * 256 + t;
                    if (j >= cnt) break;
                    const uint32_t pk = v[m] & 0x80007fffu;
                    if (at[m] != kPartial) {
                        const uint32_t i = at[m] + j;
                        if (PAIR)
                            reinterpret_cast<uint4*>(s.stage)[i] = make_uint4(pk, x0[m], x1[m], 0u);
                        else
                            reinterpret_cast<uint2*>(s.stage)[i] = make_uint2(pk, x0[m]);
                    } else {
                        const uint32_t bk = (v[m] >> kMaxBucketBits) & 0xffffu;
                        stage_put<PAIR>(s, part, bk, (uint64_t)(hist[bk] + j), pk, __uint_as_float(x0[m]),
                                        __uint_as_float(x1[m]));
                    }
                }
                __syncthreads();
                for (int b = t; b < s.nbk; b += 256) hist[b] = 0;
                __syncthreads();
            }
        } else {
            // pass 1: the chunk's bucket histogram (targets only)
            for (uint32_t r0 = 0; r0 < E; r0 += 256 * EPT) {
                uint32_t loc[EPT];
                stage_locate<SEG, EPT>(segpre, r0, E, t, loc);
                uint32_t cw[EPT];
#pragma unroll
                for (int m = 0; m < EPT; ++m)
                    cw[m] = loc[m] != 0xffffffffu ? a.col[segbeg[loc[m] >> 24] + (loc[m] & 0xffffffu)] : 0u;
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    if (loc[m] == 0xffffffffu) continue;
                    const uint32_t post = cw[m] & a.pk.idx_mask;
                    atomicAdd(&hist[(uint32_t)(((uint64_t)post * s.bmagic) >> s.bshift)], 1u);
                }
            }
            __syncthreads();
            // one reservation per touched bucket for the whole chunk; hist[b] becomes the base
            for (int b = t; b < s.nbk; b += 256) {
                const uint32_t h = hist[b];
                hist[b] = h ? atomicAdd(&s.cursor[((size_t)b * kStageParts + part) * kCursorStride], h) : 0u;
            }
            __syncthreads();
            // pass 2: reload with the weights, take positions from the bases, write
            for (uint32_t r0 = 0; r0 < E; r0 += 256 * EPT) {
                float w0[EPT], w1[EPT];
                uint32_t loc[EPT];
                stage_locate<SEG, EPT>(segpre, r0, E, t, loc);
                uint32_t cw[EPT];
                stage_load<PAIR, EPT>(a, segbeg, loc, cw, w0, w1);
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    if (loc[m] == 0xffffffffu) continue;
                    const uint32_t post = cw[m] & a.pk.idx_mask;
                    const uint32_t bk = (uint32_t)(((uint64_t)post * s.bmagic) >> s.bshift);
                    const uint32_t pos = atomicAdd(&hist[bk], 1u);
                    stage_put<PAIR>(s, part, bk, pos, (post - bk * (uint32_t)s.bsize) | (cw[m] & 0x80000000u), w0[m],
                                    w1[m]);
                }
            }
            __syncthreads();
            for (int b = t; b < s.nbk; b += 256) hist[b] = 0;
            __syncthreads();
        }
    }
    if (t == 0 && ev_wg && a.rec_events) atomicAdd((unsigned long long*)&a.rec_events[k], (unsigned long long)ev_wg);
}

// Appends every spike of the step's non-empty segments to the segment lists
// of the steps that deliver them (D > 1), right after the neuron kernel.
// Wavefront-uniform loop: lane l of a wavefront takes spike w0 + l of the
// step's sub-lists in order.  Lane d - 1 keeps delay d's ballot of non-empty
// segments and reserves that delay's entries, so the D reservation atomics
// of a wavefront are in flight together; then every lane writes its entries.
__global__ __launch_bounds__(256) void segment_emit_kernel(DeliverArgs a, SegArgs sg, int32_t k)
{
    __shared__ uint32_t cum[kSubLists + 1];
    const int64_t step = a.ctl->nbase + k + 1;
    const int32_t slot = (int32_t)(step % a.L);
    const int t = threadIdx.x;
    const uint32_t lane = __lane_id();
    if (t < 64) {
        uint32_t c = t < kSubLists
                         ? min(a.spk_cnt[((int64_t)slot * kSubLists + t) * kCntStride], (uint32_t)a.spk_cap)
                         : 0u;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(c, o);
            if ((int)lane >= o) c += y;
        }
        if (t < kSubLists) cum[t + 1] = c;
        if (t == 0) cum[0] = 0u;
    }
    __syncthreads();
    const uint32_t total = cum[kSubLists];
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t w0 = blockIdx.x * 256u + ((uint32_t)t & ~63u); w0 < total; w0 += gridDim.x * 256u) {
        const uint32_t g = w0 + lane;
        const bool ok = g < total;
        int64_t row = 0;
        const uint16_t* de = a.dend;
        if (ok) {
            int q = 0; // the last sub-list starting at or before g
#pragma unroll
            for (int st = kSubLists / 2; st >= 1; st >>= 1) q += cum[q + st] <= g ? st : 0;
            const int32_t j = a.spk[((int64_t)slot * kSubLists + q) * a.spk_cap + (g - cum[q])];
            row = a.rowptr[j];
            de += (int64_t)j * a.D;
        }
        // pass 1: lane d - 1 keeps the ballot of delay d (segment ends read
        // eight at a time, the eight loads in flight together)
        constexpr int kU = 8;
        uint64_t mine = 0ull;
        uint32_t lo = 0u;
        for (int32_t d0 = 0; d0 < a.D; d0 += kU) {
            uint32_t hv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) hv[u] = ok && d0 + u < a.D ? (uint32_t)de[d0 + u] : 0u;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (d0 + u < a.D) { // uniform
                    const uint64_t m = __ballot(hv[u] > lo);
                    if ((int32_t)lane == d0 + u) mine = m;
                    lo = hv[u];
                }
            }
        }
        uint32_t base = 0u;
        if ((int32_t)lane < a.D && mine)
            base = atomicAdd(sg.cnt + ((step + lane) % a.D) * kCursorStride, (uint32_t)__popcll(mine));
        // pass 2: the entries (the segment ends again, from the cache)
        lo = 0u;
        for (int32_t d0 = 0; d0 < a.D; d0 += kU) {
            uint32_t hv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) hv[u] = ok && d0 + u < a.D ? (uint32_t)de[d0 + u] : 0u;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int32_t d = d0 + u + 1;
                if (d > a.D) continue; // uniform
                const uint32_t hi = hv[u];
                const uint32_t mlo = __shfl((uint32_t)mine, d - 1);
                const uint32_t mhi = __shfl((uint32_t)(mine >> 32), d - 1);
                const uint32_t bd = __shfl(base, d - 1);
                if (hi > lo) {
                    const uint64_t m = ((uint64_t)mhi << 32) | mlo;
                    const uint64_t pos = (uint64_t)bd + (uint64_t)__popcll(m & below);
                    const int64_t ls = (step + d - 1) % a.D;
                    if (pos < (uint64_t)sg.cap)
                        sg.seg[ls * sg.cap + (int64_t)pos] = ((uint64_t)(row + lo) << 24) | (uint64_t)(hi - lo);
                    else
                        sg.cnt[ls * kCursorStride + 1] = 1u;
                }
                lo = hi;
            }
        }
    }
}

// (Re)start the segment lists: empty, complete from step `from` on.
__global__ void segment_reset_kernel(uint32_t* cnt, int32_t D, int64_t from)
{
    for (int x = threadIdx.x; x < D; x += blockDim.x) {
        cnt[(int64_t)x * kCursorStride] = 0u;
        cnt[(int64_t)x * kCursorStride + 1] = 0u;
    }
    if (threadIdx.x == 0) *reinterpret_cast<int64_t*>(cnt + (int64_t)D * kCursorStride) = from;
}

// dynamic LDS of the stage kernel: the delivery entries and the bucket
// histogram, plus for the sorted variant the round's bucket offsets,
// reservations and sorted events
static size_t stage_lds(const Block* b, int ept, int var)
{
    size_t words = (size_t)b->D * kSubLists + 1 + (size_t)b->nbk;
    // + pre, off, one word of alignment, the round with a spare slot
    if (var == 2) words += 2 * (size_t)b->nbk + 1 + ((size_t)256 * ept + 1) * (b->pair ? 3 : 2);
    return sizeof(uint32_t) * words;
}

template <bool PAIR, int EPT, int VAR>
static void launch_stage_ept(const Block* b, const StageArgs& sa, int32_t k)
{
    const dim3 g(b->deliver_grid), t(256);
    const size_t lds = stage_lds(b, EPT, VAR);
    switch (b->stage_segs) {
    case 64: hipLaunchKernelGGL((deliver_stage_kernel<PAIR, 64, EPT, VAR>), g, t, lds, b->stream, sa, k); break;
    case 128: hipLaunchKernelGGL((deliver_stage_kernel<PAIR, 128, EPT, VAR>), g, t, lds, b->stream, sa, k); break;
    default: hipLaunchKernelGGL((deliver_stage_kernel<PAIR, 256, EPT, VAR>), g, t, lds, b->stream, sa, k); break;
    }
}

template <bool PAIR>
static void launch_stage(const Block* b, const StageArgs& sa, int32_t k)
{
    switch (b->stage_variant) {
    case 2: launch_stage_ept<PAIR, 16, 2>(b, sa, k); break; // sorted rounds use 16 events per lane
    case 1:
        if (b->stage_ept == 16) launch_stage_ept<PAIR, 16, 1>(b, sa, k);
        else launch_stage_ept<PAIR, 32, 1>(b, sa, k);
        break;
    default:
        if (b->stage_ept == 16) launch_stage_ept<PAIR, 16, 0>(b, sa, k);
        else launch_stage_ept<PAIR, 32, 0>(b, sa, k);
    }
}

template <bool PAIR, int U, bool FIXED>
__global__ void deliver_accum_kernel(StageArgs s);

template <bool PAIR, bool FIXED>
static void launch_accum_f(const Block* b, const StageArgs& sa)
{
    const size_t lds = (size_t)b->bsize * (PAIR ? 16 : 8);
    if (b->accum_unroll == 16)
        hipLaunchKernelGGL((deliver_accum_kernel<PAIR, 16, FIXED>), dim3(b->nbk), dim3(1024), lds, b->stream, sa);
    else
        hipLaunchKernelGGL((deliver_accum_kernel<PAIR, 8, FIXED>), dim3(b->nbk), dim3(1024), lds, b->stream, sa);
}

template <bool PAIR>
static void launch_accum(const Block* b, const StageArgs& sa)
{
    if (b->accum_fixed) launch_accum_f<PAIR, true>(b, sa);
    else launch_accum_f<PAIR, false>(b, sa);
}

// Input exponents (Block::jexp).  A neuron's arrivals on channel c in one
// step are bounded by B = the sum of |w| over its inputs on c (a synapse
// delivers at most once per step).  B is summed as an integer -- each |w|
// rounded up to a multiple of 2^-sg, 64-bit integer atomics -- so it does not
// depend on the order of the atomics, and the exponents a network gets are
// the same on every run.  e = 30 + sg - ceil(log2(B)) puts B x 2^e in
// (2^29, 2^30].

// nullable exponent table: NULL reads as 0
__device__ inline int jexp_at(const int8_t* j, int64_t k) { return j ? (int)j[k] : 0; }

template <bool PAIR>
__global__ void wmax_kernel(int64_t nnz, const uint32_t* __restrict__ col, const float* __restrict__ w,
                            const float2* __restrict__ w2, const int8_t* __restrict__ jexp, uint32_t idx_mask,
                            uint32_t* __restrict__ out)
{
    uint32_t m = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
        const uint32_t c = col[e];
        const int64_t k = (int64_t)(c & idx_mask) * 4 + ((c >> 31) << 1);
        if (PAIR) {
            const float2 v = w2[e];
            m = max(m, __float_as_uint(fabsf(ldexpf(v.x, -jexp_at(jexp, k)))));
            m = max(m, __float_as_uint(fabsf(ldexpf(v.y, -jexp_at(jexp, k + 1)))));
        } else {
            m = max(m, __float_as_uint(fabsf(ldexpf(w[e], -jexp_at(jexp, k)))));
        }
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) m = max(m, (uint32_t)__shfl_xor(m, off)); // non-negative floats order like their bits
    if (__lane_id() == 0) atomicMax(out, m);
}

template <bool PAIR>
__global__ void insum_kernel(int64_t nnz, const uint32_t* __restrict__ col, const float* __restrict__ w,
                             const float2* __restrict__ w2, const int8_t* __restrict__ jexp, uint32_t idx_mask,
                             int32_t sg, unsigned long long* __restrict__ sums)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
        const uint32_t c = col[e];
        const int64_t k = (int64_t)(c & idx_mask) * 4 + ((c >> 31) << 1);
        // compact weights: one sum per receptor pair, kept on its first channel
        const int nw = PAIR ? 2 : 1;
        for (int x = 0; x < nw; ++x) {
            const float v = PAIR ? (x ? w2[e].y : w2[e].x) : w[e];
            const double a = ldexp((double)fabsf(v), sg - jexp_at(jexp, k + x));
            const unsigned long long q = (unsigned long long)ceil(a);
            if (q) atomicAdd(&sums[k + x], q);
        }
    }
}

// out[0] = smallest exponent over channels with inputs (atomicMin, starts at
// 127), out[1] = 1 if some channel's bound overflows even at the lowest
// exponent
__global__ void jexp_kernel(int64_t n, int32_t pair, int32_t sg, const unsigned long long* __restrict__ sums,
                            int8_t* __restrict__ jnew, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 4; ++c) {
        const unsigned long long B = sums[i * 4 + (pair ? c : (c & 2))];
        int e = 0;
        if (B) {
            const int cl = 64 - __clzll(B - 1ull); // ceil(log2(B))
            e = 30 + sg - cl;
            if (e < -64) {
                e = -64;
                if (cl - sg + e > 30) out[1] = 1;
            }
            e = min(e, 64);
            atomicMin(&out[0], e);
        }
        jnew[i * 4 + c] = (int8_t)e;
    }
}

template <bool PAIR>
__global__ void rescale_w_kernel(int64_t nnz, const uint32_t* __restrict__ col, float* __restrict__ w,
                                 float2* __restrict__ w2, const int8_t* __restrict__ jold,
                                 const int8_t* __restrict__ jnew, uint32_t idx_mask)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
        const uint32_t c = col[e];
        const int64_t k = (int64_t)(c & idx_mask) * 4 + ((c >> 31) << 1);
        const int d0 = jexp_at(jnew, k) - jexp_at(jold, k);
        if (PAIR) {
            const int d1 = jexp_at(jnew, k + 1) - jexp_at(jold, k + 1);
            if (d0 | d1) {
                const float2 v = w2[e];
                w2[e] = make_float2(ldexpf(v.x, d0), ldexpf(v.y, d1));
            }
        } else if (d0) {
            w[e] = ldexpf(w[e], d0);
        }
    }
}

// Jp, acc (and the overflow accumulator) from exponents jold to jnew
template <bool PAIR>
__global__ void rescale_state_kernel(int64_t n, float4* __restrict__ Jp, void* __restrict__ acc,
                                     void* __restrict__ acc2, const int8_t* __restrict__ jold,
                                     const int8_t* __restrict__ jnew)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) d[c] = jexp_at(jnew, 4 * i + c) - jexp_at(jold, 4 * i + c);
    if (!(d[0] | d[1] | d[2] | d[3])) return;
    const float4 j = Jp[i];
    Jp[i] = make_float4(ldexpf(j.x, d[0]), ldexpf(j.y, d[1]), ldexpf(j.z, d[2]), ldexpf(j.w, d[3]));
    void* bufs[2] = {acc, acc2};
    for (int x = 0; x < 2; ++x) {
        if (!bufs[x]) continue;
        if (PAIR) {
            float4* a = reinterpret_cast<float4*>(bufs[x]) + i;
            const float4 v = *a;
            *a = make_float4(ldexpf(v.x, d[0]), ldexpf(v.y, d[1]), ldexpf(v.z, d[2]), ldexpf(v.w, d[3]));
        } else {
            float2* a = reinterpret_cast<float2*>(bufs[x]) + i;
            const float2 v = *a;
            *a = make_float2(ldexpf(v.x, d[0]), ldexpf(v.y, d[2]));
        }
    }
}

// unscaled weights of one pair member (field W0: which = 0, W1: which = 1)
template <bool PAIR>
__global__ void weight_out_kernel(int64_t nnz, const uint32_t* __restrict__ col, const float* __restrict__ w,
                                  const float2* __restrict__ w2, const int8_t* __restrict__ jexp, uint32_t idx_mask,
                                  int32_t which, float* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
        const uint32_t c = col[e];
        const int64_t k = (int64_t)(c & idx_mask) * 4 + ((c >> 31) << 1) + (PAIR ? which : 0);
        const float v = PAIR ? (which ? w2[e].y : w2[e].x) : w[e];
        out[e] = ldexpf(v, -jexp_at(jexp, k));
    }
}

static inline unsigned grid_syn(int64_t nnz);
static inline unsigned grid_for(int64_t n, int bs);

// (Re)compute the input exponents from the current weights and rescale the
// weights (and, with_state, J and the pending arrivals) to them.  Exact:
// every change is a multiplication by a power of two.
static int update_jexp(Block* b, bool with_state)
{
    const int64_t n = b->n, nnz = b->nnz;
    if (!b->jexp) {
        DTB_TRY(block_alloc(b, reinterpret_cast<void**>(&b->jexp), 4 * (size_t)n));
        DTB_TRY_HIP(hipMemsetAsync(b->jexp, 0, 4 * (size_t)n, b->stream));
    }
    int8_t* jnew = nullptr;
    unsigned long long* sums = nullptr;
    uint32_t* mx = nullptr;
    int32_t* mm = nullptr;
    int st = DTB_OK;
    hipError_t e = hipMalloc(&jnew, 4 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&sums, 8 * 4 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&mx, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&mm, 2 * sizeof(int32_t));
    do {
        if (e != hipSuccess) break;
        const int32_t mm0[2] = {127, 0};
        if ((e = hipMemsetAsync(mx, 0, sizeof(uint32_t), b->stream)) != hipSuccess) break;
        if ((e = hipMemsetAsync(sums, 0, 8 * 4 * (size_t)n, b->stream)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(mm, mm0, sizeof(mm0), hipMemcpyHostToDevice, b->stream)) != hipSuccess) break;
        float maxw = 0.0f;
        if (nnz > 0) {
            if (b->pair)
                hipLaunchKernelGGL(wmax_kernel<true>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col, b->w,
                                   b->w2, b->jexp, b->pk.idx_mask, mx);
            else
                hipLaunchKernelGGL(wmax_kernel<false>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col, b->w,
                                   b->w2, b->jexp, b->pk.idx_mask, mx);
            uint32_t bits = 0;
            if ((e = hipMemcpyAsync(&bits, mx, 4, hipMemcpyDeviceToHost, b->stream)) != hipSuccess) break;
            if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) break;
            std::memcpy(&maxw, &bits, 4);
        }
        if (!std::isfinite(maxw)) {
            set_error("synaptic weights must be finite");
            st = DTB_INVALID_PARAMETER;
            break;
        }
        // sums stay below 2^63: nnz * (maxw * 2^sg + 1) < 2^62
        int32_t sg = 0;
        if (maxw > 0.0f) {
            const int lw = (int)std::ceil(std::log2((double)maxw));
            const int ln = (int)std::ceil(std::log2((double)nnz + 1.0));
            sg = std::max(-200, std::min(200, 61 - ln - lw));
            if (nnz > 0) {
                if (b->pair)
                    hipLaunchKernelGGL(insum_kernel<true>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col,
                                       b->w, b->w2, b->jexp, b->pk.idx_mask, sg, sums);
                else
                    hipLaunchKernelGGL(insum_kernel<false>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col,
                                       b->w, b->w2, b->jexp, b->pk.idx_mask, sg, sums);
            }
        }
        hipLaunchKernelGGL(jexp_kernel, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, n, (int32_t)b->pair, sg,
                           sums, jnew, mm);
        if (nnz > 0) {
            if (b->pair)
                hipLaunchKernelGGL(rescale_w_kernel<true>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col,
                                   b->w, b->w2, b->jexp, jnew, b->pk.idx_mask);
            else
                hipLaunchKernelGGL(rescale_w_kernel<false>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col,
                                   b->w, b->w2, b->jexp, jnew, b->pk.idx_mask);
        }
        if (with_state) {
            if (b->pair)
                hipLaunchKernelGGL(rescale_state_kernel<true>, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, n,
                                   b->Jp, b->acc, b->acc2, b->jexp, jnew);
            else
                hipLaunchKernelGGL(rescale_state_kernel<false>, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, n,
                                   b->Jp, b->acc, b->acc2, b->jexp, jnew);
        }
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipMemcpyAsync(b->jexp, jnew, 4 * (size_t)n, hipMemcpyDeviceToDevice, b->stream)) != hipSuccess) break;
        int32_t hm[2] = {0, 0};
        if ((e = hipMemcpyAsync(hm, mm, sizeof(hm), hipMemcpyDeviceToHost, b->stream)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) break;
        b->acc_shift = hm[0] == 127 ? 0 : hm[0];
        b->fixed_ok = hm[1] == 0;
    } while (0);
    if (jnew) hipFree(jnew);
    if (sums) hipFree(sums);
    if (mx) hipFree(mx);
    if (mm) hipFree(mm);
    if (st) return st;
    DTB_TRY_HIP(e);
    if (!b->fixed_ok) b->accum_fixed = false;
    b->persist_dirty = true; // the grouped copy holds the rounded scaled weights
    if (!b->fixed_ok && b->deliver_mode == 2) b->deliver_mode = 1;
    return DTB_OK;
}

template <bool PAIR>
static int allow_sorted_stage_lds(int bytes)
{
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_stage_kernel<PAIR, 64, 16, 2>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_stage_kernel<PAIR, 128, 16, 2>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_stage_kernel<PAIR, 256, 16, 2>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return DTB_OK;
}

// U staged events in flight per lane (8 or 16).  FIXED: the slice holds int32
// sums of round(w * 2^S) (integer LDS atomics run ~10x faster than float ones
// on gfx950; S is chosen per block so that no neuron's sum can overflow, and
// integer sums do not depend on the order of the events); otherwise fp32.
template <bool PAIR, int U, bool FIXED>
__global__ __launch_bounds__(1024) void deliver_accum_kernel(StageArgs s)
{
    using Cfg = BucketCfg<PAIR>;
    constexpr int C = Cfg::kChan;
    extern __shared__ float slice[]; // [bsize][C]
    const DeliverArgs& a = s.d;
    const int32_t b = blockIdx.x;
    const int t = threadIdx.x;
    // the partition counts and the overflow flag, one lane each (loads in
    // parallel, overlapped with clearing the slice)
    __shared__ uint32_t s_part[kStageParts + 1];
    if (s.sg.seg && b == 0 && t == 64) {
        // the stage kernel of this step has read its segment list
        uint32_t* c = s.sg.cnt + ((a.ctl->nbase + s.k + 1) % a.D) * kCursorStride;
        c[0] = 0u;
        c[1] = 0u;
    }
    if (t < kStageParts)
        s_part[t] = (uint32_t)min((uint64_t)s.cursor[((size_t)b * kStageParts + t) * kCursorStride], (uint64_t)s.cap);
    else if (t == kStageParts)
        s_part[t] = s.ovf[(size_t)b * kCursorStride];
    for (int x = t; x < s.bsize * C; x += blockDim.x) slice[x] = 0.0f;
    __syncthreads();
    // the bucket's partitions back to back: pre[p] = events in partitions < p
    uint32_t pre[kStageParts + 1];
    pre[0] = 0;
#pragma unroll
    for (int p = 0; p < kStageParts; ++p) pre[p + 1] = pre[p] + s_part[p];
    const uint32_t cnt = pre[kStageParts];
    const bool ovf = s_part[kStageParts] != 0u;
    const int64_t sb = (int64_t)b * kStageParts * s.cap;
    auto at = [&](uint32_t e) -> int64_t {
        int p = 0;
#pragma unroll
        for (int q = 1; q < kStageParts; ++q) p += e >= pre[q] ? 1 : 0;
        return sb + (int64_t)p * s.cap + (e - pre[p]);
    };
    for (uint32_t base = t; base < cnt; base += blockDim.x * U) {
        if (PAIR) {
            uint4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t e = base + (uint32_t)u * blockDim.x;
                v[u] = e < cnt ? ld_stream(reinterpret_cast<const uint4*>(s.stage) + at(e)) : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (base + (uint32_t)u * blockDim.x >= cnt) break;
                const uint32_t o = (v[u].x & 0x7fffffffu) * 4 + ((v[u].x >> 31) << 1);
                if (FIXED) {
                    int* si = reinterpret_cast<int*>(slice);
                    atomicAdd(&si[o], __float2int_rn(__uint_as_float(v[u].y)));
                    atomicAdd(&si[o + 1], __float2int_rn(__uint_as_float(v[u].z)));
                } else {
                    atomicAdd(&slice[o], __uint_as_float(v[u].y));
                    atomicAdd(&slice[o + 1], __uint_as_float(v[u].z));
                }
            }
        } else {
            uint2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t e = base + (uint32_t)u * blockDim.x;
                v[u] = e < cnt ? ld_stream(reinterpret_cast<const uint2*>(s.stage) + at(e)) : make_uint2(0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (base + (uint32_t)u * blockDim.x >= cnt) break;
                const uint32_t o = (v[u].x & 0x7fffffffu) * 2 + (v[u].x >> 31);
                if (FIXED)
                    atomicAdd(&reinterpret_cast<int*>(slice)[o], __float2int_rn(__uint_as_float(v[u].y)));
                else
                    atomicAdd(&slice[o], __uint_as_float(v[u].y));
            }
        }
    }
    __syncthreads();
    const int64_t i0 = (int64_t)b * s.bsize;
    const int32_t nloc = (int32_t)min((int64_t)s.bsize, a.n - i0);
    if (PAIR) {
        float4* acc = reinterpret_cast<float4*>(a.acc) + i0;
        float4* acc2 = reinterpret_cast<float4*>(s.acc2) + i0;
        for (int x = t; x < nloc; x += blockDim.x) {
            float4 v;
            if (FIXED) {
                const int* si = reinterpret_cast<const int*>(slice) + 4 * x;
                v = make_float4((float)si[0], (float)si[1], (float)si[2], (float)si[3]);
            } else {
                v = make_float4(slice[4 * x], slice[4 * x + 1], slice[4 * x + 2], slice[4 * x + 3]);
            }
            if (ovf) {
                const float4 o = acc2[x];
                v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
                acc2[x] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            acc[x] = v;
        }
    } else {
        float2* acc = reinterpret_cast<float2*>(a.acc) + i0;
        float2* acc2 = reinterpret_cast<float2*>(s.acc2) + i0;
        for (int x = t; x < nloc; x += blockDim.x) {
            float2 v;
            if (FIXED) {
                const int* si = reinterpret_cast<const int*>(slice) + 2 * x;
                v = make_float2((float)si[0], (float)si[1]);
            } else {
                v = make_float2(slice[2 * x], slice[2 * x + 1]);
            }
            if (ovf) {
                const float2 o = acc2[x];
                v.x += o.x; v.y += o.y;
                acc2[x] = make_float2(0.f, 0.f);
            }
            acc[x] = v;
        }
    }
    __syncthreads();
    if (t < kStageParts) s.cursor[((size_t)b * kStageParts + t) * kCursorStride] = 0u;
    if (t == 0) s.ovf[(size_t)b * kCursorStride] = 0u;
}

// ---------------------------------------------------------------------------
// Persistent delivery (deliver_mode 2) for blocks small enough that every
// workgroup of one launch is resident: one launch runs all K steps of a
// batch, so a step costs one grid barrier instead of two or three kernel
// launches.  Workgroup p (p < G <= kSubLists) owns the neurons
// [p * per, (p + 1) * per), per = ceil(n / G), so every workgroup has the
// same share: it updates them, writes their spikes into spike sub-list p
// (per never exceeds a sub-list's capacity; its only writer, so in the standard list
// format without reservation atomics), and after the barrier accumulates
// every arrival for its own neurons into an int32 LDS slice.  The synapses
// are a copy of the CSR grouped by owning workgroup: each workgroup's
// synapses are contiguous, in (row, delay) segment order (one record per
// synapse: slice index and kind, then the weights pre-rounded to the same
// int32 fixed point as bucketed delivery, so one load fetches both), and so
// are its segment bounds.  The records and bounds that the workgroups of one
// XCD read then share no cache line with other workgroups', and that XCD's
// L2 holds only its own workgroups' share of the block (an eighth of it on
// 32 workgroups over 8 XCDs, instead of most of its lines).  The results
// equal bucketed delivery bit for bit.
constexpr int kPersistMaxLds = 64 * 1024; // LDS slice bytes per workgroup
constexpr uint32_t kBarrierSpinLimit = 1u << 22;
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima), gfx9 encoding
constexpr unsigned kWaitVmcnt0 = 0x0F70;
constexpr int kPubStride = 16;                 // words between publication words (one 64-byte line each)
constexpr int kPubAbort = kSubLists * kPubStride; // word index of the abort flag
constexpr int kPubWords = kPubAbort + 64;

struct PersistArgs {
    NeuronArgs na;
    const int32_t* __restrict__ bptr;  // [G][n*D+1]: start of workgroup g's synapses of each (row, delay) segment
    // grouped synapses, one record each: compact (local index | kind << 31,
    // round(scaled weight)) as 8 bytes, pair (index | kind, w0, w1, 0) as 16
    const uint32_t* __restrict__ pcw;
    int32_t G;                         // workgroups
    int32_t m;                         // blocks per workgroup (the LDS slice holds m*256 neurons)
    int32_t per;                       // neurons per workgroup: workgroup p owns [p * per, (p + 1) * per)
    int32_t K;                         // steps of the batch
    int32_t seg_loop;                  // 1: each thread walks its own (short) segment; 0: scan-balanced chunks
    uint32_t* pub;                     // [kSubLists] 64-bit words (step << 32 | spikes), kPubStride apart; abort flag
    uint64_t* __restrict__ rec_events;
    uint32_t* __restrict__ rec_nspk;
};

// Spike-list entries written by other workgroups of the same launch are read
// with agent-coherent loads (they bypass a stale copy in this XCD's L2);
// everything else the delivery reads is read-only during the launch or the
// workgroup's own, and stays cacheable.
// one grouped synapse: its slice index | kind bit, and its rounded weight(s)
template <bool PAIR>
__device__ __forceinline__ void persist_record(const uint32_t* __restrict__ pcw, int64_t x, uint32_t& c, int2& v)
{
    if (PAIR) {
        const uint4 r = reinterpret_cast<const uint4*>(pcw)[x];
        c = r.x;
        v = make_int2((int32_t)r.y, (int32_t)r.z);
    } else {
        const uint2 r = reinterpret_cast<const uint2*>(pcw)[x];
        c = r.x;
        v = make_int2((int32_t)r.y, 0);
    }
}

__device__ inline int32_t ld_coherent(const int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// MB blocks of 256 neurons per pass (workgroups of 256*MB threads).
// RES: every thread owns at most one neuron (the workgroup's blocks fit one
// pass), whose state (V, J, refractory end, flags) and constants stay in
// registers for the whole batch; it is read before the first step and written
// back after the last.  The per-step arithmetic is the neuron kernel's own.
template <bool PAIR, bool PSET, int MB, bool RES>
__global__ __launch_bounds__(256 * MB) void persist_step_kernel(PersistArgs P)
{
    constexpr int C = PAIR ? 4 : 2;
    constexpr int NT = 256 * MB;
    extern __shared__ int s_acc[]; // [m*256][C] arrivals of the workgroup's neurons, int32 fixed point
    __shared__ uint32_t s_wcnt[4 * MB];
    __shared__ float s_tab[kMaxPsets * 16];
    __shared__ uint32_t cum[kMaxEntries + 1];
    __shared__ uint32_t wtot[4 * MB];
    __shared__ uint32_t s_cnt[(DTB_MAX_DELAY + 1) * kSubLists]; // spike counts of every sub-list, per slot
    __shared__ int32_t s_lo[NT];       // delivery: start of each segment of the current chunk
    __shared__ uint32_t s_end[NT + 1]; // delivery: inclusive prefix of the chunk's segment lengths
    __shared__ int s_ok;
    const NeuronArgs& a = P.na;
    const int32_t p = blockIdx.x;
    const int32_t G = P.G;
    const uint32_t lane = __lane_id();
    const uint32_t wid = threadIdx.x >> 6;
    if (PSET)
        for (int x = threadIdx.x; x < a.npsets * 16; x += NT) s_tab[x] = a.ptab[x];
    for (int x = threadIdx.x; x < P.m * 256 * C; x += NT) s_acc[x] = 0;
    // counts of the earlier steps' lists (written before this launch)
    for (int x = threadIdx.x; x < a.L * kSubLists; x += NT) s_cnt[x] = a.spk_cnt[(int64_t)x * kCntStride];
    if (threadIdx.x == 0) s_ok = 1;
    __syncthreads();
    const int32_t E = a.D * kSubLists;
    // register-resident neuron (RES): block sb = threadIdx / 256 of this workgroup
    const int64_t ri = (int64_t)p * P.per + threadIdx.x;
    const bool rown = RES && (int32_t)threadIdx.x < P.per && ri < a.n;
    float rV = 0.0f, rgL = 0.0f, rvL = 0.0f, rdtC = 0.0f, rvth = 0.0f, rvrst = 0.0f, riext = 0.0f, rtref = 0.0f;
    float rpr = 0.0f;
    float4 rJ = make_float4(0.f, 0.f, 0.f, 0.f), rg = rJ, rvu = rJ, rdc = rJ;
    int32_t rrend = 0, rpop = -1;
    uint8_t rflags = 0;
    if (rown) {
        rV = a.V[ri];
        rJ = a.Jp[ri];
        rrend = a.ref_end[ri];
        rg = a.g[ri];
        rpop = a.pop[ri];
        rtref = a.tref[ri];
        rpr = a.noise_mode == DTB_NOISE_PER_NEURON ? a.noisep[ri] : a.noise_rate;
        if (PSET) {
            const float* T = s_tab + 16 * a.pset[ri];
            rvu = make_float4(T[0], T[1], T[2], T[3]);
            rdc = make_float4(T[4], T[5], T[6], T[7]);
            rgL = T[8]; rvL = T[9]; rdtC = T[10]; rvth = T[11]; rvrst = T[12]; riext = T[13];
        } else {
            rvu = a.vu[ri];
            rdc = a.decay[ri];
            rgL = a.gL[ri]; rvL = a.vL[ri]; rdtC = a.dtC[ri]; rvth = a.vth[ri]; rvrst = a.vrst[ri]; riext = a.iext[ri];
        }
    }
    for (int32_t k = 0; k < P.K; ++k) {
        const int64_t step = a.ctl->nbase + k + 1;
        const double t = a.ttab[k];
        const int32_t slot = (int32_t)(step % a.L);
        const int64_t sl = (int64_t)slot * kSubLists + p;
        uint32_t run = 0; // spikes written to sub-list p this step (uniform)
        for (int32_t jb = 0; jb < P.m; jb += MB) {
            const int li = jb * 256 + (int)threadIdx.x; // index in this workgroup's slice
            const int64_t i = (int64_t)p * P.per + li;
            bool act = false, spike = false;
            float vout = 0.0f;
            int32_t pp = -1;
            if (RES) {
                if (rown) {
                    pp = rpop;
                    float4 d;
                    if (k == 0) {
                        if (PAIR) {
                            d = reinterpret_cast<const float4*>(a.acc)[ri];
                        } else {
                            const float2 d2 = reinterpret_cast<const float2*>(a.acc)[ri];
                            d = make_float4(d2.x, d2.x, d2.y, d2.y);
                        }
                    } else if (PAIR) {
                        d = make_float4((float)s_acc[4 * li], (float)s_acc[4 * li + 1], (float)s_acc[4 * li + 2],
                                        (float)s_acc[4 * li + 3]);
                    } else {
                        const float e = (float)s_acc[2 * li], f = (float)s_acc[2 * li + 1];
                        d = make_float4(e, e, f, f);
                    }
#pragma unroll
                    for (int c = 0; c < C; ++c) s_acc[C * li + c] = 0;
                    float4 J = rJ;
                    J.x += d.x; J.y += d.y; J.z += d.z; J.w += d.w;
                    lif_core(rV, J, rg, rvu, rgL, rvL, rdtC, rvth, rvrst, riext, step >= (int64_t)rrend, act, vout);
                    rV = vout;
                    if (act) {
                        rrend = refractory_end(step, t, rtref, a.dt);
                        a.last_step[ri] = (int32_t)step;
                        a.t_last[ri] = t;
                    }
                    const float u = noise_uniform(a.seed, (uint64_t)step, (uint64_t)ri);
                    spike = (u < rpr) || (act && !a.isolated);
                    rJ = lif_decay(J, rdc);
                    rflags = (uint8_t)((act ? 1 : 0) | (spike ? 2 : 0));
                }
            } else if (li < P.per && i < a.n) {
                pp = a.pop[i];
                float4 d;
                if (k == 0) {
                    // the pending arrivals of the batch's first step, as the neuron kernel reads them
                    if (PAIR) {
                        d = reinterpret_cast<const float4*>(a.acc)[i];
                    } else {
                        const float2 d2 = reinterpret_cast<const float2*>(a.acc)[i];
                        d = make_float4(d2.x, d2.x, d2.y, d2.y);
                    }
                } else if (PAIR) {
                    d = make_float4((float)s_acc[4 * li], (float)s_acc[4 * li + 1], (float)s_acc[4 * li + 2],
                                    (float)s_acc[4 * li + 3]);
                } else {
                    const float e = (float)s_acc[2 * li], f = (float)s_acc[2 * li + 1];
                    d = make_float4(e, e, f, f);
                }
#pragma unroll
                for (int c = 0; c < C; ++c) s_acc[C * li + c] = 0;
                neuron_update<PAIR, PSET, false>(a, i, step, t, k, d, s_tab, act, spike, vout);
            }
            // compaction into sub-list p: the blocks in order, each after the earlier ones
            const uint64_t msk = __ballot(spike);
            const uint32_t pre = (uint32_t)__popcll(msk & ((1ull << lane) - 1ull));
            if (lane == 0) s_wcnt[wid] = (uint32_t)__popcll(msk);
            __syncthreads();
            uint32_t off = run, tot = 0;
            for (uint32_t w = 0; w < 4 * MB; ++w) {
                if (w < wid) off += s_wcnt[w];
                tot += s_wcnt[w];
            }
            if (spike) a.spk[sl * a.spk_cap + off + pre] = (int32_t)i;
            run += tot;
            if (a.rec_freq || a.rec_vsum) step_stats(a, k, pp, act, vout);
            __syncthreads();
        }
        // publish: every wave waits until its stores (list entries, state)
        // have reached L2, then one agent-scope release writes L2 back before
        // the (step, count) word that other workgroups poll
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __syncthreads();
        if (threadIdx.x == 0) {
            a.spk_cnt[sl * kCntStride] = run; // the standard counter, for launches after this one
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(reinterpret_cast<uint64_t*>(P.pub + kPubStride * p),
                               ((uint64_t)step << 32) | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // wait for every workgroup's word of this step (wave 0, one lane per
        // workgroup), then wave 0 builds the (delay, sub-list) entry prefix:
        // cum[e + 1] = spikes in entries <= e, e = (d - 1) * kSubLists + q
        if (wid == 0) {
            // a lane stops loading once its word shows this step, so the
            // waiting workgroups do not keep the few polled lines busy
            uint32_t spins = 0;
            bool ok = true;
            bool mine = (int32_t)lane >= G;
            uint64_t v = 0;
            for (;;) {
                if (!mine) {
                    v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(P.pub + kPubStride * lane),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    mine = (v >> 32) == (uint64_t)step;
                }
                if (__all(mine)) break;
                const bool ab = lane == 0 &&
                                __hip_atomic_load(P.pub + kPubAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                if (++spins > kBarrierSpinLimit || __any(ab)) {
                    if (lane == 0) __hip_atomic_store(P.pub + kPubAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (ok && (int32_t)lane < kSubLists) s_cnt[slot * kSubLists + lane] = (int32_t)lane < G ? (uint32_t)v : 0u;
            if (lane == 0) {
                s_ok = ok ? 1 : 0;
                cum[0] = 0;
            }
            uint32_t base = 0;
            for (int32_t e0 = 0; e0 < E; e0 += kWave) {
                const int32_t e = e0 + (int32_t)lane;
                uint32_t c = 0;
                if (e < E)
                    c = min(s_cnt[delay_slot(slot, e / kSubLists + 1, a.L) * kSubLists + e % kSubLists],
                            (uint32_t)a.spk_cap);
#pragma unroll
                for (int off = 1; off < kWave; off <<= 1) {
                    const uint32_t y = __shfl_up(c, off);
                    if (lane >= (uint32_t)off) c += y;
                }
                if (e < E) cum[e + 1] = base + c;
                base += __shfl(c, kWave - 1);
            }
        }
        __syncthreads();
        if (!s_ok) break;

        // delivery, NT spiking (row, delay) segments at a time: each thread
        // reads one segment's bounds for this workgroup, a scan spreads the
        // segments' synapses evenly over the threads, and every thread issues
        // its synapse loads together before its LDS atomics
        const uint32_t total = cum[E];
        uint64_t ev = 0;
        for (uint32_t c0 = 0; c0 < total; c0 += NT) {
            const uint32_t r = c0 + threadIdx.x;
            int32_t lo = 0, hi = 0;
            if (r < total) {
                const int32_t e = entry_of(cum, E, r);
                const int32_t d = e / kSubLists + 1;
                const int32_t j = ld_coherent(a.spk + ((int64_t)delay_slot(slot, d, a.L) * kSubLists + e % kSubLists) *
                                                          a.spk_cap + (r - cum[e]));
                if (j >= 0 && (int64_t)j < a.n) {
                    const int32_t* bp = P.bptr + (int64_t)p * (a.n * a.D + 1) + ((int64_t)j * a.D + (d - 1));
                    lo = bp[0];
                    hi = bp[1];
                }
            }
            const uint32_t len = (uint32_t)(hi - lo);
            ev += len; // every workgroup counts the events it receives
            if (P.seg_loop) {
                // short segments: each thread walks its own, eight synapse loads
                // in flight at a time (no scan, no barrier inside the chunk)
                for (int32_t x0 = lo; x0 < hi; x0 += 8) {
                    uint32_t cc[8];
                    int2 vv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (x0 + u >= hi) continue;
                        persist_record<PAIR>(P.pcw, x0 + u, cc[u], vv[u]);
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (x0 + u >= hi) continue;
                        const int32_t o = (int32_t)(cc[u] & 0x7fffffffu) * C + (int32_t)(cc[u] >> 31) * (C / 2);
                        atomicAdd(&s_acc[o], vv[u].x);
                        if (PAIR) atomicAdd(&s_acc[o + 1], vv[u].y);
                    }
                }
                continue;
            }
            const uint32_t incl = block_inclusive_scan_n<4 * MB>(len, wtot);
            s_lo[threadIdx.x] = lo;
            s_end[threadIdx.x + 1] = incl;
            if (threadIdx.x == 0) s_end[0] = 0;
            __syncthreads();
            const uint32_t tot = s_end[NT];
            for (uint32_t x0 = 0; x0 < tot; x0 += 4 * NT) {
                constexpr int U = 4;
                int32_t pos[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t x = x0 + (uint32_t)u * NT + threadIdx.x;
                    pos[u] = -1;
                    if (x < tot) {
                        int32_t l = 0, h = NT - 1; // the segment holding x: last t with s_end[t] <= x
                        while (l < h) {
                            const int32_t mid = (l + h + 1) >> 1;
                            if (s_end[mid] <= x) l = mid;
                            else h = mid - 1;
                        }
                        pos[u] = s_lo[l] + (int32_t)(x - s_end[l]);
                    }
                }
                uint32_t cc[U];
                int2 vv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (pos[u] < 0) continue;
                    persist_record<PAIR>(P.pcw, pos[u], cc[u], vv[u]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (pos[u] < 0) continue;
                    const int32_t o = (int32_t)(cc[u] & 0x7fffffffu) * C + (int32_t)(cc[u] >> 31) * (C / 2);
                    atomicAdd(&s_acc[o], vv[u].x);
                    if (PAIR) atomicAdd(&s_acc[o + 1], vv[u].y);
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) ev += __shfl_xor(ev, off);
        if (lane == 0 && ev) atomicAdd((unsigned long long*)&P.rec_events[k], (unsigned long long)ev);
        if (p == 0 && threadIdx.x == 0) P.rec_nspk[k] = cum[kSubLists];
        __syncthreads();
    }
    if (rown) {
        a.V[ri] = rV;
        a.Jp[ri] = rJ;
        a.ref_end[ri] = rrend;
        a.flags[ri] = rflags;
    }
    // the arrivals pending for the next step go to acc, as bucketed delivery leaves them
    for (int32_t li = threadIdx.x; li < P.per; li += NT) {
        const int64_t i = (int64_t)p * P.per + li;
        if (i >= a.n) continue;
        if (PAIR)
            reinterpret_cast<float4*>(a.acc)[i] = make_float4((float)s_acc[4 * li], (float)s_acc[4 * li + 1],
                                                              (float)s_acc[4 * li + 2], (float)s_acc[4 * li + 3]);
        else
            reinterpret_cast<float2*>(a.acc)[i] = make_float2((float)s_acc[2 * li], (float)s_acc[2 * li + 1]);
    }
}

// After a persistent batch: the spike-list counters of the slot the next step
// writes start at zero (the neuron kernel appends with atomics), as the
// neuron kernel of the batch's last step would have left them.
__global__ void persist_epilogue_kernel(const StepCtl* __restrict__ ctl, int32_t L, uint32_t* spk_cnt)
{
    const int64_t next = ctl->nsteps + 1;
    if (threadIdx.x < kSubLists) spk_cnt[((int64_t)(next % L) * kSubLists + threadIdx.x) * kCntStride] = 0;
}

// Forced spikes of batch step k: the sources listed for the step's 0-based
// index (nbase + k) get their fire mark, which the neuron kernel consumes.
__global__ void src_mark_kernel(const StepCtl* __restrict__ ctl, int32_t k, int64_t nts,
                                const uint32_t* __restrict__ ts, const int64_t* __restrict__ ptr,
                                const uint32_t* __restrict__ ids, uint8_t* __restrict__ src)
{
    const int64_t step0 = ctl->nbase + k;
    int64_t lo = 0, hi = nts; // first timestamp >= step0
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)ts[mid] < step0) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= nts || (int64_t)ts[lo] != step0) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t x = ptr[lo] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < ptr[lo + 1]; x += stride)
        src[ids[x]] |= (uint8_t)4;
}

// Batch prologue: times of the K steps, step counters, record clearing.
__global__ void prologue_kernel(StepCtl* ctl, double* ttab, int32_t K, double dt, uint32_t* rec_freq,
                                float* rec_vsum, float* rec_isum, int64_t nrec, uint64_t* rec_events,
                                uint32_t* rec_nspk, uint32_t* bar)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (tid == 0) {
        if (bar)
            for (int q = 0; q < kSubLists; ++q) *reinterpret_cast<uint64_t*>(bar + kPubStride * q) = 0ull;
        const int64_t n0 = ctl->nsteps;
        double t = ctl->t;
        for (int32_t k = 0; k < K; ++k) {
            t += dt;
            ttab[k] = t;
        }
        ctl->nbase = n0;
        ctl->nsteps = n0 + K;
        ctl->t = t;
    }
    for (int64_t x = tid; x < nrec; x += stride) {
        if (rec_freq) rec_freq[x] = 0u;
        if (rec_vsum) rec_vsum[x] = 0.0f;
        if (rec_isum) rec_isum[x] = 0.0f;
    }
    for (int64_t x = tid; x < K; x += stride) {
        rec_events[x] = 0ull;
        rec_nspk[x] = 0u;
    }
}

// Mean I_syn per sub-population at the end of a step (after delivery).
template <bool PAIR>
__global__ __launch_bounds__(256) void isum_kernel(NeuronArgs a, int32_t k, float* rec_isum)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float isyn = 0.0f;
    int32_t p = -1;
    if (i < a.n) {
        const float V = a.V[i];
        float4 J = a.Jp[i];
        if (PAIR) {
            const float4 d = reinterpret_cast<const float4*>(a.acc)[i];
            J.x += d.x; J.y += d.y; J.z += d.z; J.w += d.w;
        } else {
            const float2 d = reinterpret_cast<const float2*>(a.acc)[i];
            J.x += d.x; J.y += d.x; J.z += d.y; J.w += d.y;
        }
        const float4 g = a.g[i];
        const float4 vu = a.vu[i];
        isyn = g.x * (vu.x - V) * J.x;
        isyn += g.y * (vu.y - V) * J.y;
        isyn += g.z * (vu.z - V) * J.z;
        isyn += g.w * (vu.w - V) * J.w;
        p = a.pop[i];
    }
    const int32_t p0 = __shfl(p, 0);
    if (__all(p == p0)) {
        if (p0 >= 0) {
            const float v = wave_sum(isyn);
            if (__lane_id() == 0) unsafeAtomicAdd(&rec_isum[(int64_t)k * a.S + p0], v);
        }
    } else if (p >= 0) {
        unsafeAtomicAdd(&rec_isum[(int64_t)k * a.S + p], isyn);
    }
}

__global__ void sample_kernel(int64_t ns, const uint32_t* __restrict__ idx, const float* __restrict__ V,
                              const uint8_t* __restrict__ flags, int32_t k, uint8_t* __restrict__ sspk,
                              float* __restrict__ sv)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const uint32_t j = idx[s];
    sv[(int64_t)k * ns + s] = V[j];
    sspk[(int64_t)k * ns + s] = (uint8_t)((flags[j] >> 1) & 1);
}

// Attached Balloon-Windkessel step per voxel: input = threshold spikes of
// the voxel's sub-populations over the last sps steps / voxel size, floored
// at 1e-5 (simulation/simulation.py:274-279, the np.histogram merge of
// populations into voxels), then models/bold_model_pytorch.py:69-77.
__global__ void bold_counts_kernel(int32_t S, int32_t V, int32_t k, int32_t sps, const uint32_t* __restrict__ rec_freq,
                                   const int32_t* __restrict__ vptr, const int32_t* __restrict__ vpops,
                                   const double* __restrict__ size, const double* __restrict__ par,
                                   double* __restrict__ st, double* __restrict__ rec_bold)
{
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint64_t c = 0;
    for (int32_t x = vptr[v]; x < vptr[v + 1]; ++x) {
        const int32_t p = vpops[x];
        for (int32_t j = k - sps + 1; j <= k; ++j) c += rec_freq[(int64_t)j * S + p];
    }
    double u = (double)c / size[v];
    u = u > 1e-5 ? u : (u == u ? 1e-5 : u);
    const double eps = par[0], ts = par[1], tf = par[2], t0 = par[3], div_alpha = 1.0 / par[4], E0 = par[5],
                 V0 = par[6], dt = par[7];
    double* s = st;
    double* q = st + V;
    double* vv = st + 2 * V;
    double* f = st + 3 * V;
    const double si = s[v], qi = q[v], vi = vv[v], fi = f[v];
    const double d_s = eps * u - si / ts - (fi - 1.0) / tf;
    const double q_part = fi > 0.0 ? 1.0 - pow(1.0 - E0, 1.0 / fi) : 1.0;
    const double qn = qi + (fi * q_part / E0 - qi * pow(vi, div_alpha - 1.0)) / t0 * dt;
    const double vn = vi + (fi - pow(vi, div_alpha)) / t0 * dt;
    double fn = fi + si * dt;
    fn = fn > 0.0 ? fn : 0.0;
    s[v] = si + d_s * dt;
    q[v] = qn;
    vv[v] = vn;
    f[v] = fn;
    const double out = V0 * (7.0 * E0 * (1.0 - qn) + 2.0 * (1.0 - qn / vn) + (2.0 * E0 - 0.2) * (1.0 - vn));
    st[4 * V + v] = out;
    st[5 * V + v] = u;
    rec_bold[(int64_t)(k / sps) * V + v] = out;
}

__global__ void scale_kernel(float* x, const float* inv, int64_t K, int32_t S)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < K * S) x[t] *= inv[t % S];
}

// Derived per-neuron arrays from the [n][22] property rows.
__global__ void derive_kernel(int64_t n, const float* __restrict__ props, double dt, int32_t excl, float4* g,
                              float4* vu, float4* decay, float* gL, float* vL, float* dtC, float* vth, float* vrst,
                              float* iext, float* tref, float* noisep, int32_t* pop, int32_t* sub,
                              int32_t* ref_end, const int32_t* last_step, const double* t_last,
                              const int8_t* __restrict__ jexp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = props + i * DTB_NPROP;
    noisep[i] = p[0];
    iext[i] = p[2];
    const int32_t s = (int32_t)p[3];
    sub[i] = s;
    pop[i] = (excl && p[1] != 0.0f) ? -1 : s;
    dtC[i] = (float)(dt / (double)p[4]);
    tref[i] = p[5];
    gL[i] = p[6];
    vL[i] = p[7];
    vth[i] = p[8];
    vrst[i] = p[9];
    // J is held x 2^e per channel, so g x 2^-e gives the same products exactly
    g[i] = make_float4(ldexpf(p[10], -jexp_at(jexp, 4 * i)), ldexpf(p[11], -jexp_at(jexp, 4 * i + 1)),
                       ldexpf(p[12], -jexp_at(jexp, 4 * i + 2)), ldexpf(p[13], -jexp_at(jexp, 4 * i + 3)));
    vu[i] = make_float4(p[14], p[15], p[16], p[17]);
    decay[i] = make_float4((float)exp(-dt / (double)p[18]), (float)exp(-dt / (double)p[19]),
                           (float)exp(-dt / (double)p[20]), (float)exp(-dt / (double)p[21]));
    ref_end[i] = refractory_end(last_step[i], t_last[i], p[5], dt);
}

// Parameter-set table gathered from the derived arrays of one representative
// neuron per set, so both kernel paths read bit-identical constants.
__global__ void pset_table_kernel(int32_t m, const int64_t* __restrict__ rep, const float4* __restrict__ vu,
                                  const float4* __restrict__ decay, const float* __restrict__ gL,
                                  const float* __restrict__ vL, const float* __restrict__ dtC,
                                  const float* __restrict__ vth, const float* __restrict__ vrst,
                                  const float* __restrict__ iext, float* __restrict__ tab)
{
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int64_t r = rep[k];
    float* T = tab + 16 * k;
    const float4 u = vu[r], d = decay[r];
    T[0] = u.x; T[1] = u.y; T[2] = u.z; T[3] = u.w;
    T[4] = d.x; T[5] = d.y; T[6] = d.z; T[7] = d.w;
    T[8] = gL[r]; T[9] = vL[r]; T[10] = dtC[r]; T[11] = vth[r]; T[12] = vrst[r]; T[13] = iext[r];
    T[14] = 0.0f; T[15] = 0.0f;
}

__global__ void init_state_kernel(int64_t n, const float* __restrict__ props, float* V)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V[i] = (props[i * DTB_NPROP + 8] + props[i * DTB_NPROP + 9]) / 2.0f; // models/block.py:56
}

// CSR packing with validation; err[0] counts bad entries.
__global__ void pack_kernel(int64_t nnz, int64_t n, const int32_t* __restrict__ post, const uint8_t* __restrict__ kind,
                            const uint8_t* __restrict__ delay, SynPack pk, uint32_t* col, uint32_t* err)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = post[e];
        const uint32_t d = delay ? (uint32_t)delay[e] : 1u;
        const uint32_t kd = kind[e];
        if (q < 0 || q >= n || d < 1 || d > (pk.dmask + 1u) || kd > 1u) {
            atomicAdd(err, 1u);
            col[e] = 0;
            continue;
        }
        col[e] = (uint32_t)q | ((d - 1u) << pk.idx_bits) | (kd << 31);
    }
}

__global__ void interleave_kernel(int64_t nnz, const float* w0, const float* w1, float2* w2)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        w2[e] = make_float2(w0[e], w1[e]);
}

__global__ void unpack_kernel(int64_t nnz, const uint32_t* __restrict__ col, SynPack pk, int32_t field, void* out)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = col[e];
        if (field == DTB_FIELD_POST) reinterpret_cast<int32_t*>(out)[e] = (int32_t)(c & pk.idx_mask);
        if (field == DTB_FIELD_KIND) reinterpret_cast<uint8_t*>(out)[e] = (uint8_t)(c >> 31);
        if (field == DTB_FIELD_DELAY) reinterpret_cast<uint8_t*>(out)[e] = (uint8_t)(((c >> pk.idx_bits) & pk.dmask) + 1);
    }
}

template <bool PAIR>
__global__ void state_kernel(int64_t n, const float* __restrict__ V, const float4* __restrict__ Jp,
                             const void* __restrict__ acc, const float4* __restrict__ g,
                             const float4* __restrict__ vu, const int8_t* __restrict__ jexp, float* J_out,
                             float* I_out, float* Isyn_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 J = Jp[i];
    if (PAIR) {
        const float4 d = reinterpret_cast<const float4*>(acc)[i];
        J.x += d.x; J.y += d.y; J.z += d.z; J.w += d.w;
    } else {
        const float2 d = reinterpret_cast<const float2*>(acc)[i];
        J.x += d.x; J.y += d.x; J.z += d.y; J.w += d.y;
    }
    const float v = V[i];
    const float4 gg = g[i];
    const float4 uu = vu[i];
    const float i0 = gg.x * (uu.x - v) * J.x, i1 = gg.y * (uu.y - v) * J.y;
    const float i2 = gg.z * (uu.z - v) * J.z, i3 = gg.w * (uu.w - v) * J.w;
    if (J_out) {
        J_out[i] = ldexpf(J.x, -jexp_at(jexp, 4 * i));
        J_out[n + i] = ldexpf(J.y, -jexp_at(jexp, 4 * i + 1));
        J_out[2 * n + i] = ldexpf(J.z, -jexp_at(jexp, 4 * i + 2));
        J_out[3 * n + i] = ldexpf(J.w, -jexp_at(jexp, 4 * i + 3));
    }
    if (I_out) {
        I_out[i] = i0; I_out[n + i] = i1; I_out[2 * n + i] = i2; I_out[3 * n + i] = i3;
    }
    if (Isyn_out) Isyn_out[i] = ((i0 + i1) + i2) + i3;
}

__global__ void flag_bit_kernel(int64_t n, const uint8_t* __restrict__ flags, int bit, uint8_t* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint8_t)((flags[i] >> bit) & 1);
}

// Property edits.  Requests are grouped by column into per-sub-population
// tables (NaN: no request) so one pass over the neurons applies them all.
__global__ void mul_props_kernel(int64_t n, int32_t nsub, const int32_t* __restrict__ sub,
                                 const float* __restrict__ table, const int32_t* __restrict__ cols, int32_t ncols,
                                 float* props)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t s = sub[i];
    if (s < 0 || s >= nsub) return;
    for (int32_t c = 0; c < ncols; ++c) {
        const float f = table[(int64_t)c * nsub + s];
        if (f == f) props[i * DTB_NPROP + cols[c]] *= f;
    }
}

// Marsaglia & Tsang (2000), "A simple method for generating gamma
// variables", ACM TOMS 26(3); shape < 1 boosted by U^(1/shape).  Normals by
// Box-Muller.  All draws come from Philox4x32-10 keyed by (seed) with
// counter (neuron, draw, GAMMA), 4 words per call.
struct PhiloxStream {
    uint32_t k0, k1;
    uint64_t idx;
    uint32_t ctr;
    u32x4 buf;
    int pos;
    __device__ uint32_t next()
    {
        if (pos == 4) {
            buf = philox4x32_10(u32x4{(uint32_t)idx, (uint32_t)(idx >> 32), ctr++, kStreamGamma}, k0, k1);
            pos = 0;
        }
        const uint32_t v = pos == 0 ? buf.x : pos == 1 ? buf.y : pos == 2 ? buf.z : buf.w;
        ++pos;
        return v;
    }
    __device__ float uniform_open() { return ((float)(next() >> 8) + 0.5f) * (1.0f / 16777216.0f); } // (0,1)
    __device__ float normal()
    {
        const float u1 = uniform_open(), u2 = uniform_open();
        return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    }
};

__device__ float gamma_draw(float alpha, PhiloxStream& rs)
{
    if (!(alpha > 0.0f)) return 0.0f;
    float scale = 1.0f;
    if (alpha < 1.0f) {
        scale = powf(rs.uniform_open(), 1.0f / alpha);
        alpha += 1.0f;
    }
    const float d = alpha - 1.0f / 3.0f;
    const float c = 1.0f / sqrtf(9.0f * d);
    for (int it = 0; it < 1000; ++it) {
        float x = 0.0f, v = 0.0f;
        for (int tries = 0; tries < 64; ++tries) {
            x = rs.normal();
            v = 1.0f + c * x;
            if (v > 0.0f) break;
        }
        if (!(v > 0.0f)) continue;
        v = v * v * v;
        const float u = rs.uniform_open();
        const float x2 = x * x;
        if (u < 1.0f - 0.0331f * x2 * x2) return scale * d * v;
        if (logf(u) < 0.5f * x2 + d * (1.0f - v + logf(v))) return scale * d * v;
    }
    return scale * d;
}

__global__ void gamma_props_kernel(int64_t n, int32_t nsub, const int32_t* __restrict__ sub,
                                   const float2* __restrict__ table, const int32_t* __restrict__ cols, int32_t ncols,
                                   uint64_t seed, float* props)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t s = sub[i];
    if (s < 0 || s >= nsub) return;
    for (int32_t c = 0; c < ncols; ++c) {
        const float2 ab = table[(int64_t)c * nsub + s];
        if (!(ab.x == ab.x)) continue;
        PhiloxStream rs{(uint32_t)seed, (uint32_t)(seed >> 32), (uint64_t)i, (uint32_t)cols[c] << 24, {}, 4};
        props[i * DTB_NPROP + cols[c]] = gamma_draw(ab.x, rs) / ab.y;
    }
}

__global__ void set_props_kernel(int64_t m, const uint32_t* __restrict__ neuron, const uint32_t* __restrict__ col,
                                 const float* __restrict__ val, float* props)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) props[(int64_t)neuron[r] * DTB_NPROP + col[r]] = val[r];
}

// ------------------------------------------------------------- host side
static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }
// Per-synapse kernels loop over their range: a dispatch holds at most 2^32-1
// work-items, fewer than the synapses of a large block.
static inline unsigned grid_syn(int64_t nnz) { return (unsigned)std::min<int64_t>(grid_for(nnz, 256), 1 << 16); }

static int upload(Block* b, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return DTB_OK;
    DTB_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, b->stream));
    return DTB_OK;
}

// Columns whose values enter a parameter set: I_ext, C, g_L, V_L, V_th,
// V_reset, V_ui[4], tau_ui[4].
static bool pset_column(uint32_t c) { return c == 2 || c == 4 || (c >= 6 && c <= 9) || (c >= 14 && c <= 21); }

static int build_psets(Block* b)
{
    if (b->pset_off) {
        b->npsets = 0;
        return DTB_OK;
    }
    std::vector<float> props((size_t)b->n * DTB_NPROP);
    DTB_TRY_HIP(hipMemcpyAsync(props.data(), b->props, sizeof(float) * props.size(), hipMemcpyDeviceToHost, b->stream));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    static const int cols[14] = {2, 4, 6, 7, 8, 9, 14, 15, 16, 17, 18, 19, 20, 21};
    std::map<std::vector<uint32_t>, int32_t> sets;
    std::vector<uint8_t> idx((size_t)b->n);
    std::vector<int64_t> rep;
    std::vector<uint32_t> key(14);
    for (int64_t i = 0; i < b->n; ++i) {
        for (int c = 0; c < 14; ++c) std::memcpy(&key[(size_t)c], &props[(size_t)i * DTB_NPROP + cols[c]], 4);
        auto it = sets.find(key);
        if (it == sets.end()) {
            if ((int32_t)sets.size() == kMaxPsets) {
                b->npsets = 0; // too many distinct sets: per-neuron arrays
                return DTB_OK;
            }
            it = sets.emplace(key, (int32_t)sets.size()).first;
            rep.push_back(i);
        }
        idx[(size_t)i] = (uint8_t)it->second;
    }
    if (!b->pset) DTB_TRY(alloc_n(b, &b->pset, b->n));
    if (!b->ptab) DTB_TRY(alloc_n(b, &b->ptab, kMaxPsets * 16));
    int64_t* drep = nullptr;
    DTB_TRY_HIP(hipMalloc(&drep, sizeof(int64_t) * rep.size()));
    int st = upload(b, drep, rep.data(), sizeof(int64_t) * rep.size());
    if (!st) st = upload(b, b->pset, idx.data(), (size_t)b->n);
    if (!st) {
        hipLaunchKernelGGL(pset_table_kernel, dim3(1), dim3(kMaxPsets), 0, b->stream, (int32_t)rep.size(), drep,
                           b->vu, b->decay, b->gL, b->vL, b->dtC, b->vth, b->vrst, b->iext, b->ptab);
        if (hipStreamSynchronize(b->stream) != hipSuccess) st = DTB_RUNTIME_ERROR;
    }
    hipFree(drep);
    if (st) return st;
    b->npsets = (int32_t)rep.size();
    return DTB_OK;
}

int derive_props(Block* b, bool psets_dirty)
{
    hipLaunchKernelGGL(derive_kernel, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n, b->props, b->dt,
                       b->stats_excl, b->g, b->vu, b->decay, b->gL, b->vL, b->dtC, b->vth, b->vrst, b->iext, b->tref,
                       b->noisep, b->pop, b->sub, b->ref_end, b->last_step, b->t_last, b->jexp);
    DTB_TRY_HIP(hipGetLastError());
    // statistics populations and their sizes (host copy of sub_idx / pop)
    std::vector<int32_t> pop((size_t)b->n);
    DTB_TRY_HIP(hipMemcpyAsync(pop.data(), b->pop, sizeof(int32_t) * (size_t)b->n, hipMemcpyDeviceToHost, b->stream));
    std::vector<int32_t> sub((size_t)b->n);
    DTB_TRY_HIP(hipMemcpyAsync(sub.data(), b->sub, sizeof(int32_t) * (size_t)b->n, hipMemcpyDeviceToHost, b->stream));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    int32_t S = 0;
    for (int32_t s : sub) {
        DTB_REQUIRE(s >= 0, "property column 3 (sub_idx) must be >= 0");
        S = std::max(S, s + 1);
    }
    DTB_REQUIRE(b->S == 0 || S <= b->S, "sub_idx range grew from %d to %d after creation", b->S, S);
    if (b->S == 0) b->S = S;
    b->pop_size.assign((size_t)b->S, 0);
    for (int32_t p : pop)
        if (p >= 0) b->pop_size[(size_t)p]++;
    std::vector<float> inv((size_t)b->S);
    for (int32_t s = 0; s < b->S; ++s) inv[(size_t)s] = b->pop_size[(size_t)s] ? 1.0f / (float)b->pop_size[(size_t)s] : NAN;
    if (!b->pop_inv) DTB_TRY(alloc_n(b, &b->pop_inv, b->S));
    if (!b->pop_sized) DTB_TRY(alloc_n(b, &b->pop_sized, b->S));
    std::vector<double> szd((size_t)b->S);
    for (int32_t s = 0; s < b->S; ++s) szd[(size_t)s] = (double)b->pop_size[(size_t)s];
    DTB_TRY_HIP(hipMemcpyAsync(b->pop_inv, inv.data(), sizeof(float) * (size_t)b->S, hipMemcpyHostToDevice, b->stream));
    DTB_TRY_HIP(hipMemcpyAsync(b->pop_sized, szd.data(), sizeof(double) * (size_t)b->S, hipMemcpyHostToDevice,
                               b->stream));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    if (psets_dirty) {
        DTB_TRY(build_psets(b));
        for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second); // the kernel variant may change
        b->graphs.clear();
    }
    return DTB_OK;
}

static int alloc_state(Block* b)
{
    const int64_t n = b->n;
    DTB_TRY(alloc_n(b, &b->props, n * DTB_NPROP));
    DTB_TRY(alloc_n(b, &b->g, n));
    DTB_TRY(alloc_n(b, &b->vu, n));
    DTB_TRY(alloc_n(b, &b->decay, n));
    DTB_TRY(alloc_n(b, &b->gL, n));
    DTB_TRY(alloc_n(b, &b->vL, n));
    DTB_TRY(alloc_n(b, &b->dtC, n));
    DTB_TRY(alloc_n(b, &b->vth, n));
    DTB_TRY(alloc_n(b, &b->vrst, n));
    DTB_TRY(alloc_n(b, &b->iext, n));
    DTB_TRY(alloc_n(b, &b->tref, n));
    DTB_TRY(alloc_n(b, &b->noisep, n));
    DTB_TRY(alloc_n(b, &b->pop, n));
    DTB_TRY(alloc_n(b, &b->sub, n));
    DTB_TRY(alloc_n(b, &b->V, n));
    DTB_TRY(alloc_n(b, &b->Jp, n));
    DTB_TRY(alloc_n(b, &b->ref_end, n));
    DTB_TRY(alloc_n(b, &b->last_step, n));
    DTB_TRY(alloc_n(b, &b->t_last, n));
    DTB_TRY(alloc_n(b, &b->flags, n));
    b->spk_cap = spk_cap_of(n);
    DTB_TRY(alloc_n(b, &b->spk, (int64_t)(b->D + 1) * kSubLists * b->spk_cap));
    DTB_TRY(alloc_n(b, &b->spk_cnt, (int64_t)(b->D + 1) * kSubLists * kCntStride));
    DTB_TRY(alloc_n(b, &b->ctl, 1));
    b->kmax = 1000;
    DTB_TRY(alloc_n(b, &b->ttab, b->kmax));
    DTB_TRY(alloc_n(b, &b->rec_events, b->kmax));
    DTB_TRY(alloc_n(b, &b->rec_nspk, b->kmax));
    DTB_TRY_HIP(hipMemsetAsync(b->Jp, 0, sizeof(float4) * (size_t)n, b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->last_step, 0, sizeof(int32_t) * (size_t)n, b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->t_last, 0, sizeof(double) * (size_t)n, b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->flags, 0, (size_t)n, b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->spk_cnt, 0, (size_t)(b->D + 1) * kSubLists * kCntStride * sizeof(uint32_t),
                               b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->ctl, 0, sizeof(StepCtl), b->stream));
    return DTB_OK;
}

static int ensure_bucketed(Block* b);

static int alloc_ring_and_records(Block* b)
{
    const size_t esz = b->pair ? sizeof(float4) : sizeof(float2);
    DTB_TRY(block_alloc(b, &b->acc, esz * (size_t)b->n));
    DTB_TRY_HIP(hipMemsetAsync(b->acc, 0, esz * (size_t)b->n, b->stream));
    DTB_TRY(alloc_n(b, &b->rec_freq, b->kmax * b->S));
    DTB_TRY(alloc_n(b, &b->rec_vsum, b->kmax * b->S));
    DTB_TRY(alloc_n(b, &b->rec_isum, b->kmax * b->S));
    // delivery launch shape: about one segment per lane group at ~20 Hz of
    // spikes (an autotuner may replace it, dtb_block_set_launch)
    const double segs = std::max(64.0, (double)b->n * 2e-3 * b->D);
    const int32_t avg_seg = (int32_t)(b->n ? b->nnz / std::max<int64_t>(1, b->n * b->D) : 64);
    b->deliver_lanes = avg_seg >= 48 ? 64 : avg_seg >= 24 ? 32 : 16;
    const double groups_per_wg = 4.0 * (64 / b->deliver_lanes);
    b->deliver_grid = (int32_t)std::min<double>(16384.0, std::max(4.0, std::ceil(segs / groups_per_wg)));
    b->stage_ept = b->pair ? 16 : 32;
    if (b->n >= (1 << 17) && b->n <= (1ll << 26)) {
        b->deliver_mode = 1;
        b->deliver_grid = 1024;
        DTB_TRY(ensure_bucketed(b));
    }
    return DTB_OK;
}

// Staging buffers of the bucketed delivery, allocated on first use.
// Bucket geometry for `occ` accumulation workgroups per CU: the largest LDS
// slice the device allows (at most 160 KB) divided by occ, then as few full
// waves of occ buckets per CU as that permits.  The bucket of a target is
// (post * bmagic) >> bshift, an exact division for post < 2^26.
struct BucketGeom {
    int32_t bsize, bshift, nbk;
    uint64_t bmagic;
    int64_t cap;
};

static BucketGeom bucket_geom(const Block* b, int occ)
{
    const int64_t ncu = std::max(b->ncu, 1);
    const int64_t bpn = b->pair ? 16 : 8; // LDS bytes per neuron
    const int64_t max_bs =
        std::min<int64_t>((1 << kMaxBucketBits) - 64, (b->lds_max - kAccumStaticLds) / occ / bpn / 64 * 64);
    const int64_t per_wave = ncu * occ * max_bs;
    const int64_t waves = (b->n + per_wave - 1) / per_wave;
    const int64_t per = (b->n + waves * ncu * occ - 1) / (waves * ncu * occ);
    BucketGeom g;
    g.bsize = (int32_t)std::min<int64_t>(max_bs, std::max<int64_t>(64, (per + 63) / 64 * 64));
    g.bshift = 26;
    while ((1ll << (g.bshift - 26)) < g.bsize) ++g.bshift;
    g.bmagic = ((1ull << g.bshift) + (uint64_t)g.bsize - 1) / (uint64_t)g.bsize;
    g.nbk = (int32_t)((b->n + g.bsize - 1) / g.bsize);
    // room for 1% of all synapses per step (a mean rate of 100 Hz at 0.1 ms),
    // at least 1024 events per bucket partition; fuller partitions fall back to atomics
    g.cap = std::max<int64_t>(1024, (b->nnz / 100 + (int64_t)g.nbk * kStageParts - 1) / ((int64_t)g.nbk * kStageParts));
    return g;
}

static int use_bucket_geom(Block* b, int occ)
{
    const BucketGeom g = bucket_geom(b, occ);
    DTB_REQUIRE(g.nbk <= kMaxBucketsLds, "too many delivery buckets");
    b->acc_occ = occ;
    b->bsize = g.bsize;
    b->bshift = g.bshift;
    b->bmagic = g.bmagic;
    b->nbk = g.nbk;
    b->stage_cap = g.cap;
    // the sorted staging variant needs its round in LDS next to the static segment table
    b->sorted_ok = stage_lds(b, 16, 2) + kStageStaticLds <= (size_t)b->lds_max;
    return DTB_OK;
}

// Staging buffers of the bucketed delivery, allocated on first use for the
// largest geometry (1 or 2 accumulation workgroups per CU) so that the
// launch option can switch between them.
// Segment lists (allocated on first use, switched on): room for the segments
// of a mean rate of 100 Hz at 0.1 ms (1% of the neurons spiking in each of
// the last D steps); a fuller list sends its step back to the spike lists.
static int ensure_segments(Block* b)
{
    DTB_REQUIRE(b->D > 1 && b->nnz < (1ll << 40),
                "segment lists need delays (max delay > 1) and fewer than 2^40 synapses");
    if (!b->seg) {
        b->seg_alloc = std::max<int64_t>(65536, (int64_t)b->D * (b->n / 100 + 1));
        DTB_TRY(alloc_n(b, &b->seg, (int64_t)b->D * b->seg_alloc));
        DTB_TRY(alloc_n(b, &b->seg_cnt, (int64_t)(b->D + 1) * kCursorStride));
        DTB_TRY_HIP(hipMemsetAsync(b->seg_cnt, 0, 4 * (size_t)(b->D + 1) * kCursorStride, b->stream));
        DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    }
    b->seg_cap = b->seg_alloc;
    b->seg_dirty = true;
    return DTB_OK;
}

static int ensure_bucketed(Block* b)
{
    if (b->stage) return DTB_OK;
    DTB_REQUIRE(b->n <= (1ll << 26), "bucketed delivery supports n <= 2^26");
    DTB_REQUIRE(b->max_row < (1ll << 24), "bucketed delivery supports rows of fewer than 2^24 synapses");
    int lds = 0;
    DTB_TRY_HIP(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device));
    DTB_TRY_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device));
    b->lds_max = std::min(std::max(lds, 131072), 163840);
    // the attribute is per kernel and process-wide: allow the largest slice any block may use
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<true, 8, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<true, 8, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<true, 16, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<true, 16, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<false, 8, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<false, 8, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<false, 16, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY_HIP(hipFuncSetAttribute((const void*)deliver_accum_kernel<false, 16, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, b->lds_max - kAccumStaticLds));
    DTB_TRY(allow_sorted_stage_lds<true>(b->lds_max - kStageStaticLds));
    DTB_TRY(allow_sorted_stage_lds<false>(b->lds_max - kStageStaticLds));
    int64_t nbk_max = 0, events_max = 0;
    for (int occ = 1; occ <= 2; ++occ) {
        const BucketGeom g = bucket_geom(b, occ);
        nbk_max = std::max<int64_t>(nbk_max, g.nbk);
        events_max = std::max<int64_t>(events_max, (int64_t)g.nbk * kStageParts * g.cap);
    }
    DTB_TRY(use_bucket_geom(b, 1));
    // the sorted staging variant indexes the staging array in 32 bits
    DTB_REQUIRE(events_max < (int64_t)kPartial, "staging array of %lld events exceeds 2^32 - 1",
                (long long)events_max);
    const size_t esz = b->pair ? 16 : 8;
    DTB_TRY(alloc_n(b, &b->cursor, nbk_max * kStageParts * kCursorStride));
    DTB_TRY(alloc_n(b, &b->ovf, nbk_max * kCursorStride));
    DTB_TRY(block_alloc(b, &b->stage, esz * (size_t)events_max));
    DTB_TRY(block_alloc(b, &b->acc2, (b->pair ? 16 : 8) * (size_t)b->n));
    DTB_TRY_HIP(hipMemsetAsync(b->cursor, 0, 4 * (size_t)nbk_max * kStageParts * kCursorStride, b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->ovf, 0, 4 * (size_t)nbk_max * kCursorStride, b->stream));
    DTB_TRY_HIP(hipMemsetAsync(b->acc2, 0, (b->pair ? 16 : 8) * (size_t)b->n, b->stream));
    // segment lists by default where their stage-kernel saving (two random
    // cache lines per delivered segment) outweighs the emission kernel's
    // fixed cost: many (neuron, delay) pairs and short segments
    const double pairs = (double)b->n * (double)b->D;
    if (b->D > 1 && pairs >= 4e8 && (double)b->nnz <= 16.0 * pairs) DTB_TRY(ensure_segments(b));
    b->accum_fixed = b->fixed_ok;
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    return DTB_OK;
}

static int finish_create(Block* b, const float* props)
{
    DTB_TRY(upload(b, b->props, props, sizeof(float) * (size_t)b->n * DTB_NPROP));
    hipLaunchKernelGGL(init_state_kernel, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n, b->props, b->V);
    DTB_TRY_HIP(hipGetLastError());
    DTB_TRY(update_jexp(b, false));
    DTB_TRY(derive_props(b));
    DTB_TRY(alloc_ring_and_records(b));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    return DTB_OK;
}

static void destroy(Block* b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    for (hipEvent_t e : b->events) hipEventDestroy(e);
    for (void* p : b->allocs) hipFree(p);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

static int begin(Block** out, int device, int64_t n)
{
    DTB_REQUIRE(out != nullptr, "out handle pointer is NULL");
    *out = nullptr;
    DTB_REQUIRE(n >= 1, "n must be >= 1");
    DTB_REQUIRE(n < (1ll << 30), "n = %lld exceeds the 2^30 neurons of one block", (long long)n);
    int count = 0;
    DTB_TRY_HIP(hipGetDeviceCount(&count));
    DTB_REQUIRE(device >= 0 && device < count, "device %d not present (%d devices)", device, count);
    DTB_TRY_HIP(hipSetDevice(device));
    Block* b = new Block();
    b->device = device;
    b->n = n;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
        return DTB_RUNTIME_ERROR;
    }
    *out = b;
    return DTB_OK;
}

static int make_pack(Block* b, int32_t D)
{
    b->D = D;
    b->pk.idx_bits = std::max<uint32_t>(1, bits_for((uint64_t)b->n));
    b->pk.dbits = bits_for((uint64_t)D);
    DTB_REQUIRE(b->pk.idx_bits + b->pk.dbits <= 31, "n=%lld with max delay %d does not fit the 32-bit synapse word",
                (long long)b->n, D);
    b->pk.idx_mask = (uint32_t)((1ull << b->pk.idx_bits) - 1);
    b->pk.dmask = (uint32_t)((1ull << b->pk.dbits) - 1);
    return DTB_OK;
}

static int create_from_csr(Block* b, const dtb_block_desc* d)
{
    const int64_t n = d->n, nnz = d->nnz;
    DTB_REQUIRE(nnz >= 0, "nnz must be >= 0");
    DTB_REQUIRE(d->props && d->rowptr, "props and rowptr are required");
    DTB_REQUIRE(nnz == 0 || (d->post && d->kind && d->w0), "post, kind and w0 are required when nnz > 0");
    DTB_REQUIRE(d->dt > 0.0, "dt must be > 0");
    b->nnz = nnz;
    b->dt = d->dt;
    b->seed = d->noise_seed;
    b->stats_excl = d->stats_exclude_blocked;
    b->pair = d->w1 != nullptr;
    // rowptr check and max delay on the host
    std::vector<int64_t> rp((size_t)n + 1);
    DTB_TRY_HIP(hipMemcpy(rp.data(), d->rowptr, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyDefault));
    DTB_REQUIRE(rp[0] == 0 && rp[(size_t)n] == nnz, "rowptr must start at 0 and end at nnz");
    for (int64_t j = 0; j < n; ++j) {
        DTB_REQUIRE(rp[(size_t)j + 1] >= rp[(size_t)j], "rowptr must be non-decreasing (row %lld)", (long long)j);
        b->max_row = std::max<int64_t>(b->max_row, rp[(size_t)j + 1] - rp[(size_t)j]);
    }
    int32_t D = 1;
    if (d->delay && nnz > 0) {
        std::vector<uint8_t> dl((size_t)nnz);
        DTB_TRY_HIP(hipMemcpy(dl.data(), d->delay, (size_t)nnz, hipMemcpyDefault));
        for (uint8_t x : dl) D = std::max<int32_t>(D, x);
        DTB_REQUIRE(D <= DTB_MAX_DELAY, "delay %d exceeds DTB_MAX_DELAY", D);
    }
    DTB_TRY(make_pack(b, D));
    DTB_TRY(alloc_state(b));
    DTB_TRY(alloc_n(b, &b->rowptr, n + 1));
    DTB_TRY(upload(b, b->rowptr, rp.data(), sizeof(int64_t) * (size_t)(n + 1)));
    DTB_TRY(alloc_n(b, &b->col, nnz));
    // with delays, each row is reordered by delay (stable) and the segment
    // ends recorded, so the delivery of a (row, delay) pair is contiguous
    std::vector<int32_t> h_post;
    std::vector<uint8_t> h_kind, h_delay;
    std::vector<float> h_w0, h_w1;
    dtb_block_desc dd = *d;
    if (D > 1 && nnz > 0) {
        h_post.resize((size_t)nnz);
        h_kind.resize((size_t)nnz);
        h_delay.resize((size_t)nnz);
        h_w0.resize((size_t)nnz);
        DTB_TRY_HIP(hipMemcpy(h_post.data(), d->post, 4 * (size_t)nnz, hipMemcpyDefault));
        DTB_TRY_HIP(hipMemcpy(h_kind.data(), d->kind, (size_t)nnz, hipMemcpyDefault));
        DTB_TRY_HIP(hipMemcpy(h_delay.data(), d->delay, (size_t)nnz, hipMemcpyDefault));
        DTB_TRY_HIP(hipMemcpy(h_w0.data(), d->w0, 4 * (size_t)nnz, hipMemcpyDefault));
        if (d->w1) {
            h_w1.resize((size_t)nnz);
            DTB_TRY_HIP(hipMemcpy(h_w1.data(), d->w1, 4 * (size_t)nnz, hipMemcpyDefault));
        }
        std::vector<uint16_t> dend((size_t)n * D);
        std::vector<int64_t> perm((size_t)nnz);
        std::vector<int64_t> cnt((size_t)D + 1);
        for (int64_t j = 0; j < n; ++j) {
            const int64_t b0 = rp[(size_t)j], b1 = rp[(size_t)j + 1];
            DTB_REQUIRE(b1 - b0 <= 65535, "row %lld has %lld synapses; rows with delays are limited to 65535",
                        (long long)j, (long long)(b1 - b0));
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int64_t e = b0; e < b1; ++e) {
                const int dl = h_delay[(size_t)e];
                DTB_REQUIRE(dl >= 1 && dl <= D, "delay %d outside 1..%d", dl, D);
                cnt[(size_t)dl]++;
            }
            for (int x = 1; x <= D; ++x) cnt[(size_t)x] += cnt[(size_t)x - 1];
            for (int x = 1; x <= D; ++x) dend[(size_t)j * D + x - 1] = (uint16_t)cnt[(size_t)x];
            for (int64_t e = b0; e < b1; ++e) perm[(size_t)(b0 + cnt[(size_t)h_delay[(size_t)e] - 1]++)] = e;
        }
        auto permute = [&](auto& v) {
            auto tmp = v;
            for (int64_t e = 0; e < nnz; ++e) v[(size_t)e] = tmp[(size_t)perm[(size_t)e]];
        };
        permute(h_post);
        permute(h_kind);
        permute(h_delay);
        permute(h_w0);
        if (d->w1) permute(h_w1);
        dd.post = h_post.data();
        dd.kind = h_kind.data();
        dd.delay = h_delay.data();
        dd.w0 = h_w0.data();
        dd.w1 = d->w1 ? h_w1.data() : nullptr;
        DTB_TRY(alloc_n(b, &b->dend, n * (int64_t)D));
        DTB_TRY(upload(b, b->dend, dend.data(), sizeof(uint16_t) * dend.size()));
        d = &dd;
    }
    if (nnz > 0) {
        // stage the raw arrays on the device, pack and validate there
        void *dpost, *dkind, *ddelay = nullptr, *derr;
        DTB_TRY_HIP(hipMalloc(&dpost, sizeof(int32_t) * (size_t)nnz));
        DTB_TRY_HIP(hipMalloc(&dkind, (size_t)nnz));
        if (d->delay) DTB_TRY_HIP(hipMalloc(&ddelay, (size_t)nnz));
        DTB_TRY_HIP(hipMalloc(&derr, sizeof(uint32_t)));
        int st = DTB_OK;
        uint32_t nerr = 0;
        do {
            if ((st = upload(b, dpost, d->post, sizeof(int32_t) * (size_t)nnz))) break;
            if ((st = upload(b, dkind, d->kind, (size_t)nnz))) break;
            if (d->delay && (st = upload(b, ddelay, d->delay, (size_t)nnz))) break;
            if (hipMemsetAsync(derr, 0, sizeof(uint32_t), b->stream) != hipSuccess) { st = DTB_RUNTIME_ERROR; break; }
            hipLaunchKernelGGL(pack_kernel, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, n,
                               (const int32_t*)dpost, (const uint8_t*)dkind, (const uint8_t*)ddelay, b->pk, b->col,
                               (uint32_t*)derr);
            if (hipMemcpyAsync(&nerr, derr, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
                hipStreamSynchronize(b->stream) != hipSuccess) {
                st = DTB_RUNTIME_ERROR;
                set_error("CSR packing failed");
            }
        } while (0);
        hipFree(dpost);
        hipFree(dkind);
        if (ddelay) hipFree(ddelay);
        hipFree(derr);
        if (st) return st;
        DTB_REQUIRE(nerr == 0, "%u synapses have post outside [0,n), kind not in {0,1} or delay outside 1..64", nerr);
        if (b->pair) {
            DTB_TRY(alloc_n(b, &b->w2, nnz));
            float *d0, *d1;
            DTB_TRY_HIP(hipMalloc(&d0, sizeof(float) * (size_t)nnz));
            DTB_TRY_HIP(hipMalloc(&d1, sizeof(float) * (size_t)nnz));
            int s2 = upload(b, d0, d->w0, sizeof(float) * (size_t)nnz);
            if (!s2) s2 = upload(b, d1, d->w1, sizeof(float) * (size_t)nnz);
            if (!s2) {
                hipLaunchKernelGGL(interleave_kernel, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, d0, d1,
                                   b->w2);
                if (hipStreamSynchronize(b->stream) != hipSuccess) s2 = DTB_RUNTIME_ERROR;
            }
            hipFree(d0);
            hipFree(d1);
            if (s2) return s2;
        } else {
            DTB_TRY(alloc_n(b, &b->w, nnz));
            DTB_TRY(upload(b, b->w, d->w0, sizeof(float) * (size_t)nnz));
        }
    } else {
        if (b->pair) DTB_TRY(alloc_n(b, &b->w2, 1));
        else DTB_TRY(alloc_n(b, &b->w, 1));
    }
    return finish_create(b, d->props);
}

// ------------------------------------------------------------- the run loop
static int ensure_persist(Block* b);
static bool persist_fits(const Block* b, int32_t* G = nullptr, int32_t* m = nullptr, int32_t* per = nullptr);

static SegArgs seg_args(const Block* b)
{
    SegArgs g;
    const bool on = b->seg && b->seg_cap > 0 && b->D > 1;
    g.seg = on ? b->seg : nullptr;
    g.cnt = b->seg_cnt;
    g.cap = b->seg_cap;
    return g;
}

static NeuronArgs neuron_args(Block* b, const dtb_run_args* a, const float* uniforms, bool freq, bool vsum)
{
    NeuronArgs x;
    x.n = b->n; x.D = b->D; x.S = b->S;
    x.g = b->g; x.vu = b->vu; x.decay = b->decay; x.gL = b->gL; x.vL = b->vL; x.dtC = b->dtC; x.vth = b->vth;
    x.vrst = b->vrst; x.iext = b->iext; x.tref = b->tref; x.noisep = b->noisep; x.jexp = b->jexp; x.pop = b->pop;
    x.src = b->src;
    x.pset = b->pset; x.ptab = b->ptab; x.npsets = b->npsets;
    x.V = b->V; x.Jp = b->Jp; x.ref_end = b->ref_end; x.last_step = b->last_step; x.t_last = b->t_last;
    x.flags = b->flags; x.acc = b->acc; x.spk = b->spk; x.spk_cnt = b->spk_cnt; x.spk_cap = b->spk_cap;
    x.L = b->D + 1; x.ctl = b->ctl;
    x.clear_acc = b->deliver_mode == 0;
    x.ttab = b->ttab;
    x.dt = b->dt; x.seed = b->seed;
    x.noise_mode = a->noise_mode; x.noise_rate = a->noise_rate; x.uniforms = uniforms; x.isolated = a->isolated;
    x.rec_freq = freq ? b->rec_freq : nullptr;
    x.rec_vsum = vsum ? b->rec_vsum : nullptr;
    x.ext = b->ext;
    return x;
}

static DeliverArgs deliver_args(Block* b)
{
    DeliverArgs x;
    x.n = b->n; x.D = b->D; x.L = b->D + 1; x.pk = b->pk; x.rowptr = b->rowptr; x.dend = b->dend; x.col = b->col;
    x.w = b->w; x.w2 = b->w2;
    x.acc = b->acc; x.spk = b->spk; x.spk_cnt = b->spk_cnt; x.spk_cap = b->spk_cap; x.ctl = b->ctl;
    x.rec_events = b->rec_events;
    x.rec_nspk = b->rec_nspk;
    return x;
}

struct Want {
    bool freq, vsum, isum, sample, bold, timing;
};


template <bool PAIR, bool PSET>
static void launch_persist_t(const Block* b, const PersistArgs& pa, int mb, size_t lds)
{
    // every thread holds at most one neuron when the workgroup's blocks fit one pass
    const bool res = b->pm <= mb;
    if (mb == 4) {
        if (res) hipLaunchKernelGGL((persist_step_kernel<PAIR, PSET, 4, true>), dim3(b->pG), dim3(1024), lds, b->stream, pa);
        else hipLaunchKernelGGL((persist_step_kernel<PAIR, PSET, 4, false>), dim3(b->pG), dim3(1024), lds, b->stream, pa);
    } else if (mb == 2) {
        if (res) hipLaunchKernelGGL((persist_step_kernel<PAIR, PSET, 2, true>), dim3(b->pG), dim3(512), lds, b->stream, pa);
        else hipLaunchKernelGGL((persist_step_kernel<PAIR, PSET, 2, false>), dim3(b->pG), dim3(512), lds, b->stream, pa);
    } else {
        if (res) hipLaunchKernelGGL((persist_step_kernel<PAIR, PSET, 1, true>), dim3(b->pG), dim3(256), lds, b->stream, pa);
        else hipLaunchKernelGGL((persist_step_kernel<PAIR, PSET, 1, false>), dim3(b->pG), dim3(256), lds, b->stream, pa);
    }
}

static void launch_persist(const Block* b, const PersistArgs& pa, int mb, size_t lds)
{
    if (b->pair) {
        if (b->npsets) launch_persist_t<true, true>(b, pa, mb, lds);
        else launch_persist_t<true, false>(b, pa, mb, lds);
    } else {
        if (b->npsets) launch_persist_t<false, true>(b, pa, mb, lds);
        else launch_persist_t<false, false>(b, pa, mb, lds);
    }
}

// A batch runs as one persistent launch when the block is set to deliver_mode
// 2 and the batch needs nothing that launch does not produce (injected
// uniforms, sources, the external drive, I_syn means, samples, per-kernel
// times); otherwise it runs on the bucketed kernels, whose fixed-point sums
// give the same results.
static bool persist_batch(const Block* b, const dtb_run_args* a, const Want& w)
{
    return b->deliver_mode == 2 && b->pcol && !b->persist_dirty && a->noise_mode != DTB_NOISE_UNIFORMS &&
           b->src_nts == 0 && !b->src && b->ext.n == 0 && !w.isum && !w.sample && !w.timing;
}

static int enqueue_batch(Block* b, int32_t K, const dtb_run_args* a, const float* uniforms, Want w)
{
    if (w.timing && (int64_t)b->events.size() < 3ll * K) {
        for (size_t e = b->events.size(); e < (size_t)(3 * K); ++e) {
            hipEvent_t ev;
            DTB_TRY_HIP(hipEventCreate(&ev));
            b->events.push_back(ev);
        }
    }
    const int64_t nrec = (int64_t)K * b->S;
    hipLaunchKernelGGL(prologue_kernel, dim3(grid_for(std::max<int64_t>(nrec, K), 256 * 4)), dim3(256), 0,
                       b->stream, b->ctl, b->ttab, K, b->dt, w.freq ? b->rec_freq : nullptr,
                       w.vsum ? b->rec_vsum : nullptr, w.isum ? b->rec_isum : nullptr, nrec, b->rec_events,
                       b->rec_nspk, b->pbar);
    const NeuronArgs na = neuron_args(b, a, uniforms, w.freq, w.vsum);
    const DeliverArgs da = deliver_args(b);
    const unsigned ng = grid_for(b->n, 256);
    if (persist_batch(b, a, w)) {
        PersistArgs pa;
        pa.na = na;
        pa.bptr = b->pbptr; pa.pcw = b->pcol;
        pa.G = b->pG; pa.m = b->pm; pa.K = K; pa.per = b->pper;
        // a thread per (row, delay, workgroup) segment when they average at most 8 synapses
        pa.seg_loop = b->nnz <= 8 * b->n * (int64_t)b->D * b->pG ? 1 : 0;
        pa.pub = b->pbar;
        pa.rec_events = b->rec_events; pa.rec_nspk = b->rec_nspk;
        const size_t lds = sizeof(int) * (size_t)b->pm * 256 * (b->pair ? 4 : 2);
        const int mb = b->pm >= 4 ? 4 : b->pm >= 2 ? 2 : 1;
        launch_persist(b, pa, mb, lds);
        hipLaunchKernelGGL(persist_epilogue_kernel, dim3(1), dim3(64), 0, b->stream, b->ctl, b->D + 1, b->spk_cnt);
        if (w.bold)
            for (int32_t k = b->bold_sps - 1; k < K; k += b->bold_sps)
                hipLaunchKernelGGL(bold_counts_kernel, dim3(grid_for(b->nvox, 256)), dim3(256), 0, b->stream, b->S,
                                   b->nvox, k, b->bold_sps, b->rec_freq, b->vox_ptr, b->vox_pops, b->vox_size,
                                   b->bold_state + 6 * (int64_t)b->nvox, b->bold_state, b->rec_bold);
        if (w.vsum)
            hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nrec, 256)), dim3(256), 0, b->stream, b->rec_vsum,
                               b->pop_inv, (int64_t)K, b->S);
        DTB_TRY_HIP(hipGetLastError());
        return DTB_OK;
    }
    for (int32_t k = 0; k < K; ++k) {
        if (b->src_nts > 0)
            hipLaunchKernelGGL(src_mark_kernel, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (b->src_max + 255) / 256))),
                               dim3(256), 0, b->stream, b->ctl, k, b->src_nts, b->src_ts, b->src_ptr, b->src_ids,
                               b->src);
        if (w.timing) DTB_TRY_HIP(hipEventRecord(b->events[(size_t)(3 * k)], b->stream));
        if (b->pair) {
            if (b->npsets)
                hipLaunchKernelGGL((neuron_step_kernel<true, true>), dim3(ng), dim3(256), 0, b->stream, na, k);
            else
                hipLaunchKernelGGL((neuron_step_kernel<true, false>), dim3(ng), dim3(256), 0, b->stream, na, k);
        } else {
            if (b->npsets)
                hipLaunchKernelGGL((neuron_step_kernel<false, true>), dim3(ng), dim3(256), 0, b->stream, na, k);
            else
                hipLaunchKernelGGL((neuron_step_kernel<false, false>), dim3(ng), dim3(256), 0, b->stream, na, k);
        }
        if (w.timing) DTB_TRY_HIP(hipEventRecord(b->events[(size_t)(3 * k + 1)], b->stream));
        if (b->deliver_mode >= 1) {
            const SegArgs sg = seg_args(b);
            StageArgs sa{da, b->nbk, b->bsize, b->bmagic, b->bshift, b->stage_cap, b->cursor, b->stage, b->acc2,
                         b->ovf, sg, k};
            if (sg.seg)
                hipLaunchKernelGGL(segment_emit_kernel, dim3((unsigned)std::min<int64_t>(1024, (b->n + 255) / 256)),
                                   dim3(256), 0, b->stream, da, sg, k);
            if (b->pair) {
                launch_stage<true>(b, sa, k);
                launch_accum<true>(b, sa);
            } else {
                launch_stage<false>(b, sa, k);
                launch_accum<false>(b, sa);
            }
        } else if (b->pair) {
            launch_deliver<true>(b, da, k);
        } else {
            launch_deliver<false>(b, da, k);
        }
        if (w.timing) DTB_TRY_HIP(hipEventRecord(b->events[(size_t)(3 * k + 2)], b->stream));
        if (w.isum) {
            if (b->pair)
                hipLaunchKernelGGL(isum_kernel<true>, dim3(ng), dim3(256), 0, b->stream, na, k, b->rec_isum);
            else
                hipLaunchKernelGGL(isum_kernel<false>, dim3(ng), dim3(256), 0, b->stream, na, k, b->rec_isum);
        }
        if (w.bold && (k + 1) % b->bold_sps == 0)
            hipLaunchKernelGGL(bold_counts_kernel, dim3(grid_for(b->nvox, 256)), dim3(256), 0, b->stream, b->S,
                               b->nvox, k, b->bold_sps, b->rec_freq, b->vox_ptr, b->vox_pops, b->vox_size,
                               b->bold_state + 6 * (int64_t)b->nvox, b->bold_state, b->rec_bold);
        if (w.sample)
            hipLaunchKernelGGL(sample_kernel, dim3(grid_for(b->ns, 256)), dim3(256), 0, b->stream, b->ns, b->samples,
                               b->V, b->flags, k, b->rec_sspk, b->rec_sv);
    }
    if (w.vsum)
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nrec, 256)), dim3(256), 0, b->stream, b->rec_vsum, b->pop_inv,
                           (int64_t)K, b->S);
    if (w.isum)
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nrec, 256)), dim3(256), 0, b->stream, b->rec_isum, b->pop_inv,
                           (int64_t)K, b->S);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

static int copy_out(Block* b, void* dst, const void* src, size_t bytes)
{
    if (!dst || bytes == 0) return DTB_OK;
    DTB_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, b->stream));
    return DTB_OK;
}

static bool is_device_ptr(const void* p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice;
}

static int run_impl(Block* b, int64_t nsteps, const dtb_run_args* a)
{
    DTB_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    DTB_REQUIRE(a != nullptr, "run args are NULL");
    DTB_REQUIRE(a->noise_mode >= 0 && a->noise_mode <= 2, "noise_mode must be 0, 1 or 2");
    DTB_REQUIRE(a->noise_mode != DTB_NOISE_UNIFORMS || a->uniforms, "noise_mode 2 needs uniforms");
    DTB_REQUIRE(b->steps_done + nsteps < 0x7fff0000ll, "step counter limit (2^31) reached");
    const bool bold = b->bold_sps > 0;
    const Want w{a->freqs != nullptr || bold, a->vmeans != nullptr, a->imeans != nullptr,
                 a->sample_spike != nullptr || a->sample_v != nullptr, bold, a->kernel_ms != nullptr};
    DTB_REQUIRE(!bold || nsteps % b->bold_sps == 0, "with BOLD attached, nsteps must be a multiple of %d",
                b->bold_sps);
    DTB_REQUIRE(bold || !a->bold, "bold output requested but no BOLD integrator is attached");
    if (w.sample && b->ns == 0) {
        set_error("sample outputs requested but no samples are set");
        return DTB_UNSET_SAMPLE_FAILURE;
    }
    DTB_TRY_HIP(hipSetDevice(b->device));
    if (b->deliver_mode == 2 && b->persist_dirty) {
        // weights or layout changed since the grouped copy was built
        if (persist_fits(b)) DTB_TRY(ensure_persist(b));
        else b->deliver_mode = 1;
    }
    // segment lists: written only by bucketed batches; restarted after any
    // batch that did not write them (or a state load)
    const bool seg_batches = b->deliver_mode >= 1 && seg_args(b).seg != nullptr && !persist_batch(b, a, w);
    if (seg_batches && b->seg_dirty && nsteps > 0) {
        hipLaunchKernelGGL(segment_reset_kernel, dim3(1), dim3(64), 0, b->stream, b->seg_cnt, b->D,
                           (int64_t)(b->steps_done + b->D));
        DTB_TRY_HIP(hipGetLastError());
        b->seg_dirty = false;
    }
    if (!seg_batches && nsteps > 0) b->seg_dirty = true;
    float* ustage = nullptr;
    const bool uhost = a->noise_mode == DTB_NOISE_UNIFORMS && !is_device_ptr(a->uniforms);
    if (uhost) DTB_TRY_HIP(hipMalloc(&ustage, sizeof(float) * (size_t)b->n * (size_t)std::min<int64_t>(nsteps, b->kmax)));
    int st = DTB_OK;
    int64_t done = 0;
    while (done < nsteps && st == DTB_OK) {
        int64_t kcap = b->kmax;
        if (bold) kcap -= kcap % b->bold_sps;
        const int32_t K = (int32_t)std::min<int64_t>(nsteps - done, kcap);
        const float* u = nullptr;
        if (a->noise_mode == DTB_NOISE_UNIFORMS) {
            const float* src = a->uniforms + (size_t)done * (size_t)b->n;
            if (uhost) {
                if ((st = upload(b, ustage, src, sizeof(float) * (size_t)b->n * (size_t)K))) break;
                u = ustage;
            } else {
                u = src;
            }
        }
        const bool graph = a->use_graph && a->noise_mode != DTB_NOISE_UNIFORMS && !w.timing;
        if (graph) {
            uint32_t nr;
            std::memcpy(&nr, &a->noise_rate, 4);
            std::vector<uint64_t> key{(uint64_t)K, (uint64_t)a->noise_mode, (uint64_t)a->isolated, nr,
                                      (uint64_t)w.freq, (uint64_t)w.vsum, (uint64_t)w.isum, (uint64_t)w.sample,
                                      (uint64_t)b->ns, (uint64_t)b->bold_sps};
            auto it = b->graphs.find(key);
            if (it == b->graphs.end()) {
                hipGraph_t g;
                if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                    set_error("hipStreamBeginCapture failed");
                    st = DTB_RUNTIME_ERROR;
                    break;
                }
                int s2 = enqueue_batch(b, K, a, nullptr, w);
                hipError_t e2 = hipStreamEndCapture(b->stream, &g);
                if (s2 || e2 != hipSuccess) {
                    if (!s2) set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e2));
                    st = s2 ? s2 : DTB_RUNTIME_ERROR;
                    break;
                }
                hipGraphExec_t ge;
                e2 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (e2 != hipSuccess) {
                    set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e2));
                    st = DTB_RUNTIME_ERROR;
                    break;
                }
                it = b->graphs.emplace(key, ge).first;
            }
            if (hipGraphLaunch(it->second, b->stream) != hipSuccess) {
                set_error("hipGraphLaunch failed");
                st = DTB_RUNTIME_ERROR;
                break;
            }
        } else {
            if ((st = enqueue_batch(b, K, a, u, w))) break;
        }
        const size_t S = (size_t)b->S;
        if ((st = copy_out(b, a->freqs ? a->freqs + done * S : nullptr, b->rec_freq, sizeof(uint32_t) * S * K))) break;
        if ((st = copy_out(b, a->vmeans ? a->vmeans + done * S : nullptr, b->rec_vsum, sizeof(float) * S * K))) break;
        if ((st = copy_out(b, a->imeans ? a->imeans + done * S : nullptr, b->rec_isum, sizeof(float) * S * K))) break;
        if ((st = copy_out(b, a->sample_spike ? a->sample_spike + done * b->ns : nullptr, b->rec_sspk,
                           (size_t)b->ns * K)))
            break;
        if ((st = copy_out(b, a->sample_v ? a->sample_v + done * b->ns : nullptr, b->rec_sv,
                           sizeof(float) * (size_t)b->ns * K)))
            break;
        if ((st = copy_out(b, a->events ? a->events + done : nullptr, b->rec_events, sizeof(uint64_t) * K))) break;
        if ((st = copy_out(b, a->spikes ? a->spikes + done : nullptr, b->rec_nspk, sizeof(uint32_t) * K))) break;
        if (bold && (st = copy_out(b, a->bold ? a->bold + (done / b->bold_sps) * b->nvox : nullptr, b->rec_bold,
                                   sizeof(double) * (size_t)b->nvox * (size_t)(K / b->bold_sps))))
            break;
        if (w.timing) {
            if (hipStreamSynchronize(b->stream) != hipSuccess) {
                set_error("step batch failed: %s", hipGetErrorString(hipGetLastError()));
                st = DTB_RUNTIME_ERROR;
                break;
            }
            for (int32_t k = 0; k < K; ++k) {
                float t0 = 0.f, t1 = 0.f;
                hipEventElapsedTime(&t0, b->events[(size_t)(3 * k)], b->events[(size_t)(3 * k + 1)]);
                hipEventElapsedTime(&t1, b->events[(size_t)(3 * k + 1)], b->events[(size_t)(3 * k + 2)]);
                float tt[2] = {t0, t1};
                if ((st = copy_out(b, a->kernel_ms + 2 * (done + k), tt, sizeof(tt)))) break;
            }
            if (st) break;
        }
        if (uhost || K < nsteps - done) {
            // the staging buffer and records are reused by the next batch
            if (hipStreamSynchronize(b->stream) != hipSuccess) {
                set_error("step batch failed: %s", hipGetErrorString(hipGetLastError()));
                st = DTB_RUNTIME_ERROR;
                break;
            }
        }
        for (int32_t k = 0; k < K; ++k) b->t_host += b->dt;
        b->steps_done += K;
        done += K;
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    if (ustage) hipFree(ustage);
    if (st) return st;
    if (e != hipSuccess) {
        set_error("step batch failed: %s", hipGetErrorString(e));
        return DTB_RUNTIME_ERROR;
    }
    if (b->pbar) {
        uint32_t ab = 0;
        DTB_TRY_HIP(hipMemcpy(&ab, b->pbar + kPubAbort, 4, hipMemcpyDeviceToHost));
        if (ab) {
            DTB_TRY_HIP(hipMemset(b->pbar + kPubAbort, 0, 4));
            set_error("persistent step launch: a workgroup waited past the barrier limit (not all workgroups "
                      "resident); the block state is undefined");
            return DTB_RUNTIME_ERROR;
        }
    }
    return DTB_OK;
}

static int get_impl(Block* b, int32_t field, void* dst, int64_t bytes, int64_t* bytes_out)
{
    DTB_REQUIRE(dst != nullptr || field == DTB_FIELD_SPIKE_LIST, "dst is NULL");
    DTB_TRY_HIP(hipSetDevice(b->device));
    const int64_t n = b->n, nnz = b->nnz;
    auto need = [&](int64_t want) -> int {
        if (bytes_out) *bytes_out = want;
        if (bytes < want) {
            set_error("field %d needs %lld bytes, got %lld", field, (long long)want, (long long)bytes);
            return DTB_INVALID_PARAMETER;
        }
        return DTB_OK;
    };
    void* tmp = nullptr;
    int st = DTB_OK;
    switch (field) {
    case DTB_FIELD_V:
        DTB_TRY(need(4 * n));
        DTB_TRY(copy_out(b, dst, b->V, 4 * (size_t)n));
        break;
    case DTB_FIELD_J:
    case DTB_FIELD_I_UI:
    case DTB_FIELD_I_SYN: {
        const int64_t sz = field == DTB_FIELD_I_SYN ? 4 * n : 16 * n;
        DTB_TRY(need(sz));
        DTB_TRY_HIP(hipMalloc(&tmp, (size_t)sz));
        float* J = field == DTB_FIELD_J ? (float*)tmp : nullptr;
        float* I = field == DTB_FIELD_I_UI ? (float*)tmp : nullptr;
        float* Is = field == DTB_FIELD_I_SYN ? (float*)tmp : nullptr;
        if (b->pair)
            hipLaunchKernelGGL(state_kernel<true>, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, n, b->V, b->Jp,
                               b->acc, b->g, b->vu, b->jexp, J, I, Is);
        else
            hipLaunchKernelGGL(state_kernel<false>, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, n, b->V, b->Jp,
                               b->acc, b->g, b->vu, b->jexp, J, I, Is);
        st = copy_out(b, dst, tmp, (size_t)sz);
        break;
    }
    case DTB_FIELD_ACTIVE:
    case DTB_FIELD_SPIKE:
        DTB_TRY(need(n));
        DTB_TRY_HIP(hipMalloc(&tmp, (size_t)n));
        hipLaunchKernelGGL(flag_bit_kernel, dim3(grid_for(n, 256)), dim3(256), 0, b->stream, n, b->flags,
                           field == DTB_FIELD_ACTIVE ? 0 : 1, (uint8_t*)tmp);
        st = copy_out(b, dst, tmp, (size_t)n);
        break;
    case DTB_FIELD_SPIKE_LIST: {
        // the sub-lists of the last step, concatenated and sorted
        const int64_t slot = b->steps_done % (b->D + 1);
        std::vector<uint32_t> c((size_t)kSubLists * kCntStride, 0u);
        if (b->steps_done > 0) {
            DTB_TRY_HIP(hipMemcpyAsync(c.data(), b->spk_cnt + slot * kSubLists * kCntStride, 4 * c.size(),
                                       hipMemcpyDeviceToHost, b->stream));
            DTB_TRY_HIP(hipStreamSynchronize(b->stream));
        }
        int64_t cnt = 0;
        for (int q = 0; q < kSubLists; ++q) cnt += std::min<int64_t>(c[(size_t)q * kCntStride], b->spk_cap);
        DTB_TRY(need(4 * cnt));
        if (cnt) {
            std::vector<int32_t> l((size_t)cnt);
            int64_t at = 0;
            for (int q = 0; q < kSubLists; ++q) {
                const int64_t m = std::min<int64_t>(c[(size_t)q * kCntStride], b->spk_cap);
                if (m)
                    DTB_TRY_HIP(hipMemcpyAsync(l.data() + at, b->spk + (slot * kSubLists + q) * b->spk_cap,
                                               4 * (size_t)m, hipMemcpyDeviceToHost, b->stream));
                at += m;
            }
            DTB_TRY_HIP(hipStreamSynchronize(b->stream));
            std::sort(l.begin(), l.end());
            DTB_TRY_HIP(hipMemcpy(dst, l.data(), 4 * (size_t)cnt, hipMemcpyDefault));
        }
        break;
    }
    case DTB_FIELD_PROPS:
        DTB_TRY(need(4 * n * DTB_NPROP));
        DTB_TRY(copy_out(b, dst, b->props, 4 * (size_t)n * DTB_NPROP));
        break;
    case DTB_FIELD_ROWPTR:
        DTB_TRY(need(8 * (n + 1)));
        DTB_TRY(copy_out(b, dst, b->rowptr, 8 * (size_t)(n + 1)));
        break;
    case DTB_FIELD_POST:
    case DTB_FIELD_KIND:
    case DTB_FIELD_DELAY: {
        const int64_t sz = field == DTB_FIELD_POST ? 4 * nnz : nnz;
        DTB_TRY(need(sz));
        if (nnz == 0) break;
        DTB_TRY_HIP(hipMalloc(&tmp, (size_t)sz));
        hipLaunchKernelGGL(unpack_kernel, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col, b->pk, field,
                           tmp);
        if (hipError_t le = hipGetLastError(); le != hipSuccess) {
            set_error("unpack launch failed: %s", hipGetErrorString(le));
            st = DTB_RUNTIME_ERROR;
            break;
        }
        st = copy_out(b, dst, tmp, (size_t)sz);
        break;
    }
    case DTB_FIELD_W0:
    case DTB_FIELD_W1:
        DTB_TRY(need(4 * nnz));
        if (nnz == 0) break;
        // the stored weights are scaled by the input exponents of their targets
        DTB_TRY_HIP(hipMalloc(&tmp, 4 * (size_t)nnz));
        if (b->pair)
            hipLaunchKernelGGL(weight_out_kernel<true>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col,
                               b->w, b->w2, b->jexp, b->pk.idx_mask, (int32_t)(field == DTB_FIELD_W1), (float*)tmp);
        else
            hipLaunchKernelGGL(weight_out_kernel<false>, dim3(grid_syn(nnz)), dim3(256), 0, b->stream, nnz, b->col,
                               b->w, b->w2, b->jexp, b->pk.idx_mask, (int32_t)(field == DTB_FIELD_W1), (float*)tmp);
        st = copy_out(b, dst, tmp, 4 * (size_t)nnz);
        break;
    case DTB_FIELD_T_LAST:
        DTB_TRY(need(8 * n));
        DTB_TRY(copy_out(b, dst, b->t_last, 8 * (size_t)n));
        break;
    case DTB_FIELD_POP_SIZES:
        DTB_TRY(need(8 * (int64_t)b->S));
        DTB_TRY_HIP(hipMemcpy(dst, b->pop_size.data(), 8 * (size_t)b->S, hipMemcpyDefault));
        break;
    default:
        set_error("unknown field %d", field);
        return DTB_INVALID_PARAMETER;
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    if (tmp) hipFree(tmp);
    if (st) return st;
    DTB_TRY_HIP(e);
    return DTB_OK;
}

// State snapshot: a 256-byte header, then the device sections below, each
// starting on a 256-byte boundary.  The sections are everything the next
// steps read that a step writes; properties, connectivity, the launch shape
// and the per-batch records are not part of it.  Jp and acc are stored in the
// saving block's input scaling, so the input exponents are the last section;
// a loading block re-expresses them in its own (a power-of-two rescale).
namespace {
constexpr uint64_t kSnapMagic = 0x31504e5342544455ull; // "UDTBSNP1"
constexpr uint32_t kSnapVersion = 2;

struct SnapHeader {
    uint64_t magic;
    uint32_t version;
    uint32_t pair;
    int64_t n;
    int32_t D;
    int32_t nvox;
    int32_t bold_sps;
    int32_t pad0;
    int64_t spk_cap;
    int64_t steps_done;
    double t_host;
    uint64_t seed;
    int64_t bytes;
};
static_assert(sizeof(SnapHeader) <= 256, "snapshot header exceeds its slot");

struct SnapSection {
    void* p;
    size_t bytes;
};
} // namespace

static std::vector<SnapSection> snap_sections(Block* b)
{
    const size_t n = (size_t)b->n, L = (size_t)b->D + 1;
    std::vector<SnapSection> v{
        {b->ctl, sizeof(StepCtl)},
        {b->V, 4 * n},
        {b->Jp, 16 * n},
        {b->acc, (b->pair ? 16 : 8) * n},
        {b->ref_end, 4 * n},
        {b->last_step, 4 * n},
        {b->t_last, 8 * n},
        {b->flags, n},
        {b->spk, 4 * L * kSubLists * (size_t)b->spk_cap},
        {b->spk_cnt, 4 * L * kSubLists * kCntStride},
    };
    if (b->nvox > 0) v.push_back({b->bold_state, 6 * sizeof(double) * (size_t)b->nvox});
    v.push_back({b->jexp, 4 * n}); // last: loaded into a scratch copy, not over the block's own
    return v;
}

static int64_t snap_bytes(Block* b)
{
    int64_t off = 256;
    for (const SnapSection& x : snap_sections(b)) off += ((int64_t)x.bytes + 255) / 256 * 256;
    return off;
}

static int save_state_impl(Block* b, void* dst, int64_t bytes)
{
    const int64_t need = snap_bytes(b);
    DTB_REQUIRE(dst != nullptr, "dst is NULL");
    DTB_REQUIRE(bytes >= need, "a state snapshot takes %lld bytes, got %lld", (long long)need, (long long)bytes);
    DTB_TRY_HIP(hipSetDevice(b->device));
    SnapHeader h{};
    h.magic = kSnapMagic;
    h.version = kSnapVersion;
    h.pair = b->pair ? 1u : 0u;
    h.n = b->n;
    h.D = b->D;
    h.nvox = b->nvox;
    h.bold_sps = b->bold_sps;
    h.spk_cap = b->spk_cap;
    h.steps_done = b->steps_done;
    h.t_host = b->t_host;
    h.seed = b->seed;
    h.bytes = need;
    char* out = static_cast<char*>(dst);
    int st = copy_out(b, out, &h, sizeof(h));
    int64_t off = 256;
    for (const SnapSection& x : snap_sections(b)) {
        if (st) break;
        st = copy_out(b, out + off, x.p, x.bytes);
        off += ((int64_t)x.bytes + 255) / 256 * 256;
    }
    // the header is read from this frame: wait for the copies before returning
    const hipError_t e = hipStreamSynchronize(b->stream);
    if (st) return st;
    DTB_TRY_HIP(e);
    return DTB_OK;
}

static int load_state_impl(Block* b, const void* src, int64_t bytes)
{
    DTB_REQUIRE(src != nullptr, "src is NULL");
    DTB_REQUIRE(bytes >= 256, "a state snapshot is at least 256 bytes, got %lld", (long long)bytes);
    DTB_TRY_HIP(hipSetDevice(b->device));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    SnapHeader h{};
    DTB_TRY_HIP(hipMemcpy(&h, src, sizeof(h), hipMemcpyDefault));
    DTB_REQUIRE(h.magic == kSnapMagic && h.version == kSnapVersion, "not a state snapshot (version %u)", kSnapVersion);
    const int64_t need = snap_bytes(b);
    DTB_REQUIRE(h.n == b->n && h.D == b->D && h.pair == (b->pair ? 1u : 0u) && h.spk_cap == b->spk_cap,
                "snapshot of a block with n=%lld, max delay %d, pair %u; this block has n=%lld, max delay %d, pair %d",
                (long long)h.n, h.D, h.pair, (long long)b->n, b->D, b->pair ? 1 : 0);
    DTB_REQUIRE(h.nvox == b->nvox && h.bold_sps == b->bold_sps,
                "snapshot taken with %d BOLD voxels every %d steps; this block has %d every %d", h.nvox, h.bold_sps,
                b->nvox, b->bold_sps);
    DTB_REQUIRE(h.bytes == need && bytes >= need, "snapshot size %lld (buffer %lld), expected %lld",
                (long long)h.bytes, (long long)bytes, (long long)need);
    DTB_REQUIRE(h.steps_done >= 0 && h.steps_done < 0x7fff0000ll, "snapshot step counter out of range");
    const char* in = static_cast<const char*>(src);
    int64_t off = 256;
    int st = DTB_OK;
    b->seg_dirty = true; // the segment lists are not part of the snapshot: rebuilt from the spike lists
    int8_t* jsnap = nullptr;
    DTB_TRY_HIP(hipMalloc(&jsnap, 4 * (size_t)b->n));
    const std::vector<SnapSection> secs = snap_sections(b);
    for (size_t k = 0; k < secs.size(); ++k) {
        const SnapSection& x = secs[k];
        void* to = k + 1 == secs.size() ? (void*)jsnap : x.p;
        if (hipMemcpyAsync(to, in + off, x.bytes, hipMemcpyDefault, b->stream) != hipSuccess) {
            set_error("state snapshot copy failed");
            st = DTB_RUNTIME_ERROR;
            break;
        }
        off += ((int64_t)x.bytes + 255) / 256 * 256;
    }
    if (!st) {
        if (b->pair)
            hipLaunchKernelGGL(rescale_state_kernel<true>, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n,
                               b->Jp, b->acc, nullptr, jsnap, b->jexp);
        else
            hipLaunchKernelGGL(rescale_state_kernel<false>, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n,
                               b->Jp, b->acc, nullptr, jsnap, b->jexp);
    }
    const hipError_t e = hipStreamSynchronize(b->stream);
    hipFree(jsnap);
    if (st) return st;
    DTB_TRY_HIP(e);
    if (h.seed != b->seed) {
        // captured batches hold the seed as a kernel argument
        for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->seed = h.seed;
    }
    b->steps_done = h.steps_done;
    b->t_host = h.t_host;
    return DTB_OK;
}

// Property requests grouped by column into [ncols][S] tables.
template <typename T>
static int group_requests(Block* b, int64_t m, const uint32_t* pop, const uint32_t* col, const T* val, T none,
                          std::vector<int32_t>& cols, std::vector<T>& table)
{
    std::vector<uint32_t> hp((size_t)m), hc((size_t)m);
    std::vector<T> hv((size_t)m);
    DTB_TRY_HIP(hipMemcpy(hp.data(), pop, 4 * (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hc.data(), col, 4 * (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hv.data(), val, sizeof(T) * (size_t)m, hipMemcpyDefault));
    int32_t colpos[DTB_NPROP];
    std::fill(colpos, colpos + DTB_NPROP, -1);
    cols.clear();
    for (int64_t r = 0; r < m; ++r) {
        DTB_REQUIRE(hc[(size_t)r] < DTB_NPROP, "property column %u out of range", hc[(size_t)r]);
        DTB_REQUIRE(hc[(size_t)r] != 3, "property column 3 (sub_idx) cannot be changed");
        DTB_REQUIRE(hp[(size_t)r] < (uint32_t)b->S, "sub-population %u out of range [0,%d)", hp[(size_t)r], b->S);
        if (colpos[hc[(size_t)r]] < 0) {
            colpos[hc[(size_t)r]] = (int32_t)cols.size();
            cols.push_back((int32_t)hc[(size_t)r]);
        }
    }
    table.assign(cols.size() * (size_t)b->S, none);
    for (int64_t r = 0; r < m; ++r) {
        T& slot = table[(size_t)colpos[hc[(size_t)r]] * b->S + hp[(size_t)r]];
        slot = hv[(size_t)r];
    }
    return DTB_OK;
}

} // namespace dtb

using namespace dtb;

// ------------------------------------------------------- weight updates
// Requests name synapses by (pre, post, channel); the rows of a presynaptic
// neuron are short (at most a few thousand synapses), so each request scans
// its row.  Pass 1 counts the matches; pass 2 writes the first match and
// zeroes the others, so several synapses joining one pair on one receptor
// group sum to the requested value, as one dense w_uij entry does.
__global__ void find_syn_kernel(int64_t m, const uint32_t* __restrict__ pre, const uint32_t* __restrict__ post,
                                const uint8_t* __restrict__ chan, const int64_t* __restrict__ rowptr,
                                const uint32_t* __restrict__ col, uint32_t idx_mask, int64_t* __restrict__ first)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const uint32_t want = post[r] | ((uint32_t)(chan[r] >> 1) << 31);
    int64_t f = -1;
    for (int64_t e = rowptr[pre[r]]; e < rowptr[pre[r] + 1]; ++e) {
        const uint32_t c = col[e];
        if (((c & idx_mask) | (c & 0x80000000u)) == want) {
            f = e;
            break;
        }
    }
    first[r] = f;
}

__global__ void write_syn_kernel(int64_t m, const uint32_t* __restrict__ pre, const uint32_t* __restrict__ post,
                                 const uint8_t* __restrict__ chan, const float* __restrict__ val,
                                 const int64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                 uint32_t idx_mask, const int8_t* __restrict__ jexp, int32_t pair, float* w)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const uint32_t ch = chan[r];
    const uint32_t want = post[r] | ((ch >> 1) << 31);
    // stored in the target's input scaling (exact)
    const float x = ldexpf(val[r], jexp[(int64_t)post[r] * 4 + ch]);
    bool first = true;
    for (int64_t e = rowptr[pre[r]]; e < rowptr[pre[r] + 1]; ++e) {
        const uint32_t c = col[e];
        if (((c & idx_mask) | (c & 0x80000000u)) != want) continue;
        w[pair ? 2 * e + (ch & 1) : e] = first ? x : 0.0f;
        first = false;
    }
}

__global__ void dup_pair_kernel(int64_t nnz, const float* __restrict__ w, float2* __restrict__ w2)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) w2[e] = make_float2(w[e], w[e]);
}

// compact arrivals (E, I) -> per channel (E, E, I, I)
__global__ void widen_acc_kernel(int64_t n, const float2* __restrict__ a2, float4* __restrict__ a4)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 v = a2[i];
        a4[i] = make_float4(v.x, v.x, v.y, v.y);
    }
}

static void block_free(Block* b, void* p)
{
    if (!p) return;
    auto it = std::find(b->allocs.begin(), b->allocs.end(), p);
    if (it != b->allocs.end()) b->allocs.erase(it);
    hipFree(p);
}

// Compact weights (one per synapse, shared by the receptor pair) to pair
// weights (one per channel): weights, arrivals and the bucketed staging take
// the pair layout; results continue unchanged.
static int make_pair(Block* b)
{
    if (b->pair) return DTB_OK;
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    float2* w2 = nullptr;
    DTB_TRY(alloc_n(b, &w2, std::max<int64_t>(b->nnz, 1)));
    if (b->nnz > 0)
        hipLaunchKernelGGL(dup_pair_kernel, dim3(grid_syn(b->nnz)), dim3(256), 0, b->stream, b->nnz, b->w, w2);
    void* acc4 = nullptr;
    DTB_TRY(block_alloc(b, &acc4, sizeof(float4) * (size_t)b->n));
    hipLaunchKernelGGL(widen_acc_kernel, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n,
                       (const float2*)b->acc, (float4*)acc4);
    DTB_TRY_HIP(hipGetLastError());
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    block_free(b, b->w);
    block_free(b, b->acc);
    b->w = nullptr;
    b->w2 = w2;
    b->acc = acc4;
    // staging holds 16-byte events for pair weights: rebuilt on demand
    const bool had = b->stage != nullptr;
    block_free(b, b->stage);
    block_free(b, b->cursor);
    block_free(b, b->ovf);
    block_free(b, b->acc2);
    b->stage = nullptr;
    b->cursor = nullptr;
    b->ovf = nullptr;
    b->acc2 = nullptr;
    b->pair = true;
    b->stage_ept = 16;
    b->persist_dirty = true;
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    if (had || b->deliver_mode == 1) {
        const bool fixed = b->accum_fixed;
        DTB_TRY(ensure_bucketed(b));
        b->accum_fixed = fixed && b->fixed_ok;
        if (b->stage_variant == 2 && !b->sorted_ok) b->stage_variant = 0;
    }
    return DTB_OK;
}

namespace dtb {

// Persistent delivery fits when at most kSubLists workgroups of 256 threads
// cover the neurons with an LDS slice of at most kPersistMaxLds bytes, the
// synapse positions fit int32 and the fixed-point sums cannot overflow.
static bool persist_fits(const Block* b, int32_t* G, int32_t* m, int32_t* per)
{
    // G workgroups own consecutive, equal ranges of ceil(n / G) neurons (at most
    // a sub-list's capacity), each held in an LDS slice of m 256-neuron blocks
    const int64_t nblk = (b->n + 255) / 256;
    const int64_t g = std::min<int64_t>(kSubLists, nblk);
    const int64_t pp = g > 0 ? (b->n + g - 1) / g : 0;
    const int64_t mm = (pp + 255) / 256;
    const int64_t lds = mm * 256 * (b->pair ? 4 : 2) * 4;
    if (G) *G = (int32_t)g;
    if (m) *m = (int32_t)mm;
    if (per) *per = (int32_t)pp;
    return b->n > 0 && b->fixed_ok && lds <= kPersistMaxLds && b->nnz < (1ll << 31) &&
           b->n * b->D * (std::min<int64_t>(kSubLists, nblk) + 1) < (1ll << 28);
}

// (Re)build the grouped copy of the CSR (group_by_owner): the synapses of
// each owning workgroup (post / per) back to back, workgroup after
// workgroup, in (row, delay) segment order inside each, each stored as its
// index in that workgroup's LDS slice (post - owner * per) with the kind bit,
// and its weight rounded to the int32 the bucketed accumulation adds
// (__float2int_rn of the scaled weight).
static int ensure_persist(Block* b)
{
    if (!b->persist_dirty && b->pcol) return DTB_OK;
    int32_t G = 0, m = 0, per = 0;
    DTB_REQUIRE(persist_fits(b, &G, &m, &per),
                "persistent delivery needs at most %d x 256 x %d neurons, fixed-point sums and fewer than 2^31 "
                "synapses", kSubLists, kPersistMaxLds / (256 * 4 * (b->pair ? 4 : 2)), kSubLists);
    DTB_TRY_HIP(hipSetDevice(b->device));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    const int64_t n = b->n, nnz = b->nnz, D = b->D;
    std::vector<int64_t> rp((size_t)n + 1);
    std::vector<uint32_t> col((size_t)nnz);
    std::vector<float> w((size_t)nnz * (b->pair ? 2 : 1));
    DTB_TRY_HIP(hipMemcpy(rp.data(), b->rowptr, 8 * (size_t)(n + 1), hipMemcpyDeviceToHost));
    if (nnz > 0) {
        DTB_TRY_HIP(hipMemcpy(col.data(), b->col, 4 * (size_t)nnz, hipMemcpyDeviceToHost));
        DTB_TRY_HIP(hipMemcpy(w.data(), b->pair ? (const void*)b->w2 : (const void*)b->w, 4 * w.size(),
                              hipMemcpyDeviceToHost));
    }
    DTB_REQUIRE(rp[(size_t)n] == nnz, "the CSR row pointer does not end at nnz");
    GroupedCsr gc;
    const std::string err = group_by_owner(n, D, G, per, b->pk.idx_bits, b->pk.dmask, b->pk.idx_mask, b->pair,
                                           rp.data(), col.data(), w.data(), gc);
    DTB_REQUIRE(err.empty(), "persistent delivery: %s", err.c_str());
    const std::vector<int32_t>& bptr = gc.bptr;
    const std::vector<uint32_t>& rec = gc.rec;
    const int R = b->pair ? 4 : 2; // 32-bit words per record
    block_free(b, b->pbptr);
    block_free(b, b->pcol);
    b->pbptr = nullptr;
    b->pcol = nullptr;
    DTB_TRY(alloc_n(b, &b->pbptr, (int64_t)bptr.size()));
    DTB_TRY(alloc_n(b, &b->pcol, std::max<int64_t>(nnz * R, R)));
    if (!b->pbar) {
        DTB_TRY(alloc_n(b, &b->pbar, kPubWords));
        DTB_TRY_HIP(hipMemsetAsync(b->pbar, 0, 4 * kPubWords, b->stream));
    }
    DTB_TRY(upload(b, b->pbptr, bptr.data(), 4 * bptr.size()));
    if (nnz > 0) {
        DTB_TRY(upload(b, b->pcol, rec.data(), 4 * rec.size()));
    }
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    b->pG = G;
    b->pm = m;
    b->pper = per;
    b->persist_dirty = false;
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    return DTB_OK;
}

} // namespace dtb

static int set_weights_impl(Block* b, int64_t m, const uint32_t* pre, const uint32_t* post, const uint8_t* chan,
                            const float* val)
{
    DTB_REQUIRE(m >= 0, "m must be >= 0");
    if (m == 0) return DTB_OK;
    DTB_REQUIRE(pre && post && chan && val, "pre, post, channel and w are required");
    std::vector<uint32_t> hp((size_t)m), hq((size_t)m);
    std::vector<uint8_t> hc((size_t)m);
    std::vector<float> hv((size_t)m);
    DTB_TRY_HIP(hipMemcpy(hp.data(), pre, 4 * (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hq.data(), post, 4 * (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hc.data(), chan, (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hv.data(), val, 4 * (size_t)m, hipMemcpyDefault));
    // validate; the last request for a (pre, post, channel) wins
    std::map<std::tuple<uint32_t, uint32_t, uint8_t>, float> req;
    for (int64_t r = 0; r < m; ++r) {
        DTB_REQUIRE(hp[(size_t)r] < (uint64_t)b->n && hq[(size_t)r] < (uint64_t)b->n,
                    "request %lld: neuron index out of range", (long long)r);
        DTB_REQUIRE(hc[(size_t)r] <= 3, "request %lld: channel must be 0..3", (long long)r);
        DTB_REQUIRE(std::isfinite(hv[(size_t)r]), "request %lld: weight is not finite", (long long)r);
        req[std::make_tuple(hp[(size_t)r], hq[(size_t)r], hc[(size_t)r])] = hv[(size_t)r];
    }
    // compact weights hold one value per receptor pair: a request that sets
    // the two channels of a pair differently (or only one of them) needs pair weights
    bool split = false;
    if (!b->pair) {
        for (const auto& kv : req) {
            const auto& [p, q, c] = kv.first;
            auto it = req.find(std::make_tuple(p, q, (uint8_t)(c ^ 1)));
            if (it == req.end() || it->second != kv.second) {
                split = true;
                break;
            }
        }
    }
    const int64_t u = (int64_t)req.size();
    std::vector<uint32_t> up, uq;
    std::vector<uint8_t> uc;
    std::vector<float> uv;
    for (const auto& kv : req) {
        up.push_back(std::get<0>(kv.first));
        uq.push_back(std::get<1>(kv.first));
        uc.push_back(std::get<2>(kv.first));
        uv.push_back(kv.second);
    }
    DTB_TRY_HIP(hipSetDevice(b->device));
    char* dbuf = nullptr;
    const size_t off_q = 4 * (size_t)u, off_c = 8 * (size_t)u, off_v = (9 * (size_t)u + 7) & ~(size_t)7,
                 off_f = (off_v + 4 * (size_t)u + 7) & ~(size_t)7;
    DTB_TRY_HIP(hipMalloc(&dbuf, off_f + 8 * (size_t)u));
    int st = DTB_OK;
    do {
        if ((st = upload(b, dbuf, up.data(), 4 * (size_t)u))) break;
        if ((st = upload(b, dbuf + off_q, uq.data(), 4 * (size_t)u))) break;
        if ((st = upload(b, dbuf + off_c, uc.data(), (size_t)u))) break;
        if ((st = upload(b, dbuf + off_v, uv.data(), 4 * (size_t)u))) break;
        const uint32_t* dp = (const uint32_t*)dbuf;
        const uint32_t* dq = (const uint32_t*)(dbuf + off_q);
        const uint8_t* dc = (const uint8_t*)(dbuf + off_c);
        const float* dv = (const float*)(dbuf + off_v);
        int64_t* df = (int64_t*)(dbuf + off_f);
        hipLaunchKernelGGL(find_syn_kernel, dim3(grid_for(u, 256)), dim3(256), 0, b->stream, u, dp, dq, dc, b->rowptr,
                           b->col, b->pk.idx_mask, df);
        std::vector<int64_t> hf((size_t)u);
        if (hipMemcpyAsync(hf.data(), df, 8 * (size_t)u, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess) {
            set_error("synapse search failed");
            st = DTB_RUNTIME_ERROR;
            break;
        }
        for (int64_t r = 0; r < u; ++r) {
            if (hf[(size_t)r] < 0) {
                set_error("no synapse from neuron %u to neuron %u on channel %u: only existing synapses can be "
                          "changed", up[(size_t)r], uq[(size_t)r], (unsigned)uc[(size_t)r]);
                st = DTB_INVALID_PARAMETER;
                break;
            }
        }
        if (st) break;
        if (split && (st = make_pair(b))) break;
        float* wp = b->pair ? reinterpret_cast<float*>(b->w2) : b->w;
        hipLaunchKernelGGL(write_syn_kernel, dim3(grid_for(u, 256)), dim3(256), 0, b->stream, u, dp, dq, dc, dv,
                           b->rowptr, b->col, b->pk.idx_mask, b->jexp, (int32_t)b->pair, wp);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
            set_error("weight update failed");
            st = DTB_RUNTIME_ERROR;
        }
    } while (0);
    hipFree(dbuf);
    if (st) return st;
    // the input bounds changed: new exponents, weights and state rescaled, g re-derived
    DTB_TRY(update_jexp(b, true));
    return derive_props(b, false);
}

// ------------------------------------------------------- forced spikes
static int set_sources_impl(Block* b, int64_t n_src, const uint32_t* sources, const uint8_t* frozen,
                            int64_t n_events, const uint32_t* step, const uint32_t* neuron)
{
    DTB_REQUIRE(n_src >= 0 && n_events >= 0, "counts must be >= 0");
    DTB_REQUIRE(n_src == 0 || sources, "sources is NULL");
    DTB_REQUIRE(n_events == 0 || (step && neuron), "step and neuron are required with events");
    DTB_REQUIRE(n_src > 0 || n_events == 0, "events need sources");
    const int64_t n = b->n;
    std::vector<uint32_t> hs((size_t)n_src), ht((size_t)n_events), hn((size_t)n_events);
    std::vector<uint8_t> hf((size_t)n_src, 0);
    if (n_src) DTB_TRY_HIP(hipMemcpy(hs.data(), sources, 4 * (size_t)n_src, hipMemcpyDefault));
    if (n_src && frozen) DTB_TRY_HIP(hipMemcpy(hf.data(), frozen, (size_t)n_src, hipMemcpyDefault));
    if (n_events) {
        DTB_TRY_HIP(hipMemcpy(ht.data(), step, 4 * (size_t)n_events, hipMemcpyDefault));
        DTB_TRY_HIP(hipMemcpy(hn.data(), neuron, 4 * (size_t)n_events, hipMemcpyDefault));
    }
    std::vector<uint8_t> flag((size_t)n, 0);
    for (int64_t k = 0; k < n_src; ++k) {
        DTB_REQUIRE(hs[(size_t)k] < (uint64_t)n, "source %lld: neuron %u out of range", (long long)k, hs[(size_t)k]);
        flag[hs[(size_t)k]] = (uint8_t)(1 | (hf[(size_t)k] ? 2 : 0));
    }
    std::vector<int64_t> order((size_t)n_events);
    for (int64_t k = 0; k < n_events; ++k) {
        DTB_REQUIRE(hn[(size_t)k] < (uint64_t)n && (flag[hn[(size_t)k]] & 1),
                    "event %lld: neuron %u is not a source", (long long)k, hn[(size_t)k]);
        DTB_REQUIRE(ht[(size_t)k] < 0x7fff0000u, "event %lld: step index out of range", (long long)k);
        order[(size_t)k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return ht[(size_t)x] < ht[(size_t)y]; });
    std::vector<uint32_t> ts, ids((size_t)n_events);
    std::vector<int64_t> ptr{0};
    int64_t mx = 0;
    for (int64_t k = 0; k < n_events; ++k) {
        const uint32_t t = ht[(size_t)order[(size_t)k]];
        if (ts.empty() || ts.back() != t) {
            if (!ts.empty()) ptr.push_back(k);
            ts.push_back(t);
        }
        ids[(size_t)k] = hn[(size_t)order[(size_t)k]];
    }
    if (!ts.empty()) ptr.push_back(n_events);
    for (size_t k = 0; k + 1 < ptr.size(); ++k) mx = std::max<int64_t>(mx, ptr[k + 1] - ptr[k]);
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    block_free(b, b->src);
    block_free(b, b->src_ts);
    block_free(b, b->src_ptr);
    block_free(b, b->src_ids);
    b->src = nullptr;
    b->src_ts = nullptr;
    b->src_ptr = nullptr;
    b->src_ids = nullptr;
    b->src_nts = 0;
    b->src_max = 0;
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    if (n_src == 0) return DTB_OK;
    DTB_TRY(alloc_n(b, &b->src, n));
    DTB_TRY(upload(b, b->src, flag.data(), (size_t)n));
    if (!ts.empty()) {
        DTB_TRY(alloc_n(b, &b->src_ts, (int64_t)ts.size()));
        DTB_TRY(alloc_n(b, &b->src_ptr, (int64_t)ptr.size()));
        DTB_TRY(alloc_n(b, &b->src_ids, n_events));
        DTB_TRY(upload(b, b->src_ts, ts.data(), 4 * ts.size()));
        DTB_TRY(upload(b, b->src_ptr, ptr.data(), 8 * ptr.size()));
        DTB_TRY(upload(b, b->src_ids, ids.data(), 4 * ids.size()));
    }
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    b->src_nts = (int64_t)ts.size();
    b->src_max = mx;
    return DTB_OK;
}

struct dtb_block : public Block {};

extern "C" {

const char* dtb_last_error(void) { return g_err; }

int32_t dtb_abi_version(void) { return 1; }

int dtb_block_create(dtb_block** out, const dtb_block_desc* desc)
{
    DTB_REQUIRE(desc != nullptr, "desc is NULL");
    Block* b = nullptr;
    DTB_TRY(begin(&b, desc->device, desc->n));
    int st = create_from_csr(b, desc);
    if (st) {
        destroy(b);
        return st;
    }
    *out = static_cast<dtb_block*>(b);
    return DTB_OK;
}

int dtb_block_create_fixed_indegree(dtb_block** out, const dtb_fixed_indegree_spec* spec, const float* props,
                                    double dt, uint64_t noise_seed, int32_t stats_exclude_blocked, int32_t device)
{
    DTB_REQUIRE(spec != nullptr && props != nullptr, "spec and props are required");
    DTB_REQUIRE(spec->n >= 2, "fixed in-degree networks need n >= 2");
    DTB_REQUIRE(spec->in_degree >= 0, "in_degree must be >= 0");
    DTB_REQUIRE(spec->max_delay >= 1 && spec->max_delay <= DTB_MAX_DELAY, "max_delay must be in 1..64");
    DTB_REQUIRE(spec->n_exc >= 0 && spec->n_exc <= spec->n, "n_exc must be in [0, n]");
    DTB_REQUIRE(dt > 0.0, "dt must be > 0");
    Block* b = nullptr;
    DTB_TRY(begin(&b, device, spec->n));
    int st = DTB_OK;
    do {
        b->dt = dt;
        b->seed = noise_seed;
        b->stats_excl = stats_exclude_blocked;
        b->pair = spec->pair_weights != 0;
        b->nnz = spec->n * (int64_t)spec->in_degree;
        if ((st = make_pack(b, spec->max_delay))) break;
        b->max_row = 8192; // the generator's row limit
        if ((st = alloc_state(b))) break;
        if ((st = gen_fixed_indegree(b, spec))) break;
        st = finish_create(b, props);
    } while (0);
    if (st) {
        destroy(b);
        return st;
    }
    *out = static_cast<dtb_block*>(b);
    return DTB_OK;
}

static int create_members_impl(dtb_block** out, const dtb_member_spec* members, int32_t n_members,
                               const uint64_t* source_cdf, int32_t max_delay, float w_lo, float w_hi,
                               int32_t pair_weights, uint64_t seed, const float* props, double dt,
                               uint64_t noise_seed, int32_t stats_exclude_blocked, int32_t device, bool regions,
                               int32_t replicas = 1)
{
    DTB_REQUIRE(out != nullptr, "out handle pointer is NULL");
    DTB_REQUIRE(members != nullptr && props != nullptr && n_members >= 1, "members and props are required");
    DTB_REQUIRE(max_delay >= 1 && max_delay <= DTB_MAX_DELAY, "max_delay must be in 1..64");
    DTB_REQUIRE(dt > 0.0, "dt must be > 0");
    DTB_REQUIRE(!regions || source_cdf != nullptr, "source_cdf is required");
    DTB_REQUIRE(replicas >= 1 && (int64_t)n_members * replicas < (1ll << 31), "replicas must be >= 1");
    const int32_t R0 = n_members;
    std::vector<dtb_member_spec> hm((size_t)n_members);
    DTB_TRY_HIP(hipMemcpy(hm.data(), members, sizeof(dtb_member_spec) * (size_t)n_members, hipMemcpyDefault));
    for (int32_t r = 1; r < replicas; ++r) hm.insert(hm.end(), hm.begin(), hm.begin() + R0);
    n_members *= replicas;
    int64_t n = 0, nnz = 0;
    for (const auto& m : hm) {
        DTB_REQUIRE(m.n >= 1 && m.n_exc >= 0 && m.n_exc <= m.n && m.in_degree >= 0 && m.inter_degree >= 0,
                    "bad member spec");
        DTB_REQUIRE(regions || m.inter_degree == 0, "inter_degree must be 0 (use dtb_block_create_regions)");
        DTB_REQUIRE(m.in_degree == 0 || m.n >= 2, "a member with synapses needs at least 2 neurons");
        n += m.n;
        nnz += m.n * (int64_t)(m.in_degree + m.inter_degree);
    }
    Block* b = nullptr;
    DTB_TRY(begin(&b, device, n));
    int st = DTB_OK;
    do {
        b->dt = dt;
        b->seed = noise_seed;
        b->stats_excl = stats_exclude_blocked;
        b->pair = pair_weights != 0;
        b->nnz = nnz;
        if ((st = make_pack(b, max_delay))) break;
        b->max_row = 8192;
        if ((st = alloc_state(b))) break;
        if ((st = gen_members(b, hm.data(), n_members, max_delay, w_lo, w_hi, seed, regions ? source_cdf : nullptr,
                              R0)))
            break;
        st = finish_create(b, props);
    } while (0);
    if (st) {
        destroy(b);
        return st;
    }
    *out = static_cast<dtb_block*>(b);
    return DTB_OK;
}

int dtb_block_create_members(dtb_block** out, const dtb_member_spec* members, int32_t n_members,
                             int32_t max_delay, float w_lo, float w_hi, int32_t pair_weights, uint64_t seed,
                             const float* props, double dt, uint64_t noise_seed, int32_t stats_exclude_blocked,
                             int32_t device)
{
    return create_members_impl(out, members, n_members, nullptr, max_delay, w_lo, w_hi, pair_weights, seed, props,
                               dt, noise_seed, stats_exclude_blocked, device, false);
}

int dtb_block_create_regions(dtb_block** out, const dtb_member_spec* members, int32_t n_members,
                             const uint64_t* source_cdf, int32_t max_delay, float w_lo, float w_hi,
                             int32_t pair_weights, uint64_t seed, const float* props, double dt,
                             uint64_t noise_seed, int32_t stats_exclude_blocked, int32_t device)
{
    return create_members_impl(out, members, n_members, source_cdf, max_delay, w_lo, w_hi, pair_weights, seed, props,
                               dt, noise_seed, stats_exclude_blocked, device, true);
}

int dtb_block_create_region_replicas(dtb_block** out, const dtb_member_spec* regions, int32_t n_regions,
                                     int32_t replicas, const uint64_t* source_cdf, int32_t max_delay, float w_lo,
                                     float w_hi, int32_t pair_weights, uint64_t seed, const float* props, double dt,
                                     uint64_t noise_seed, int32_t stats_exclude_blocked, int32_t device)
{
    return create_members_impl(out, regions, n_regions, source_cdf, max_delay, w_lo, w_hi, pair_weights, seed, props,
                               dt, noise_seed, stats_exclude_blocked, device, true, replicas);
}

int dtb_block_destroy(dtb_block* b)
{
    destroy(b);
    return DTB_OK;
}

int dtb_block_run(dtb_block* b, int64_t nsteps, const dtb_run_args* args)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    return run_impl(b, nsteps, args);
}

int dtb_block_get(dtb_block* b, int32_t field, void* dst, int64_t bytes, int64_t* bytes_out)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    return get_impl(b, field, dst, bytes, bytes_out);
}

int dtb_block_state_bytes(dtb_block* b, int64_t* bytes)
{
    DTB_REQUIRE(b != nullptr && bytes != nullptr, "block or bytes is NULL");
    *bytes = snap_bytes(b);
    return DTB_OK;
}

int dtb_block_save_state(dtb_block* b, void* dst, int64_t bytes)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    return save_state_impl(b, dst, bytes);
}

int dtb_block_load_state(dtb_block* b, const void* src, int64_t bytes)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    return load_state_impl(b, src, bytes);
}

int dtb_block_info(dtb_block* b, int64_t* info)
{
    DTB_REQUIRE(b != nullptr && info != nullptr, "block or info is NULL");
    info[0] = b->n;
    info[1] = b->nnz;
    info[2] = b->S;
    info[3] = b->D;
    info[4] = b->steps_done;
    info[5] = b->ns;
    info[6] = b->pair ? 1 : 0;
    info[7] = b->device;
    info[8] = b->nvox;
    return DTB_OK;
}

int dtb_block_time(dtb_block* b, double* t)
{
    DTB_REQUIRE(b != nullptr && t != nullptr, "block or t is NULL");
    *t = b->t_host;
    return DTB_OK;
}

int dtb_block_set_props(dtb_block* b, int64_t m, const uint32_t* neuron, const uint32_t* col, const float* val)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_REQUIRE(m >= 0, "m must be >= 0");
    if (m == 0) return DTB_OK;
    DTB_REQUIRE(neuron && col && val, "neuron, col and val are required");
    DTB_TRY_HIP(hipSetDevice(b->device));
    std::vector<uint32_t> hn((size_t)m), hc((size_t)m);
    DTB_TRY_HIP(hipMemcpy(hn.data(), neuron, 4 * (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hc.data(), col, 4 * (size_t)m, hipMemcpyDefault));
    for (int64_t r = 0; r < m; ++r) {
        DTB_REQUIRE(hn[(size_t)r] < (uint64_t)b->n, "neuron %u out of range", hn[(size_t)r]);
        DTB_REQUIRE(hc[(size_t)r] < DTB_NPROP && hc[(size_t)r] != 3, "property column %u not settable", hc[(size_t)r]);
    }
    void *dn, *dc, *dv;
    DTB_TRY_HIP(hipMalloc(&dn, 4 * (size_t)m));
    DTB_TRY_HIP(hipMalloc(&dc, 4 * (size_t)m));
    DTB_TRY_HIP(hipMalloc(&dv, 4 * (size_t)m));
    int st = upload(b, dn, hn.data(), 4 * (size_t)m);
    if (!st) st = upload(b, dc, hc.data(), 4 * (size_t)m);
    if (!st) st = upload(b, dv, val, 4 * (size_t)m);
    if (!st) {
        hipLaunchKernelGGL(set_props_kernel, dim3(grid_for(m, 256)), dim3(256), 0, b->stream, m, (const uint32_t*)dn,
                           (const uint32_t*)dc, (const float*)dv, b->props);
        st = derive_props(b, std::any_of(hc.begin(), hc.end(), pset_column));
    }
    hipStreamSynchronize(b->stream);
    hipFree(dn);
    hipFree(dc);
    hipFree(dv);
    return st;
}

int dtb_block_mul_props_by_pop(dtb_block* b, int64_t m, const uint32_t* pop, const uint32_t* col, const float* val)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_REQUIRE(m >= 0, "m must be >= 0");
    if (m == 0) return DTB_OK;
    DTB_REQUIRE(pop && col && val, "pop, col and val are required");
    DTB_TRY_HIP(hipSetDevice(b->device));
    std::vector<int32_t> cols;
    std::vector<float> table;
    DTB_TRY(group_requests<float>(b, m, pop, col, val, NAN, cols, table));
    void *dt, *dc;
    DTB_TRY_HIP(hipMalloc(&dt, sizeof(float) * table.size()));
    DTB_TRY_HIP(hipMalloc(&dc, sizeof(int32_t) * cols.size()));
    int st = upload(b, dt, table.data(), sizeof(float) * table.size());
    if (!st) st = upload(b, dc, cols.data(), sizeof(int32_t) * cols.size());
    if (!st) {
        hipLaunchKernelGGL(mul_props_kernel, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n, b->S, b->sub,
                           (const float*)dt, (const int32_t*)dc, (int32_t)cols.size(), b->props);
        st = derive_props(b, std::any_of(cols.begin(), cols.end(), [](int32_t c) { return pset_column((uint32_t)c); }));
    }
    hipStreamSynchronize(b->stream);
    hipFree(dt);
    hipFree(dc);
    return st;
}

int dtb_block_gamma_props_by_pop(dtb_block* b, int64_t m, const uint32_t* pop, const uint32_t* col,
                                 const float* alpha, const float* beta, uint64_t seed)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_REQUIRE(m >= 0, "m must be >= 0");
    if (m == 0) return DTB_OK;
    DTB_REQUIRE(pop && col && alpha && beta, "pop, col, alpha and beta are required");
    DTB_TRY_HIP(hipSetDevice(b->device));
    std::vector<float> ha((size_t)m), hb((size_t)m);
    DTB_TRY_HIP(hipMemcpy(ha.data(), alpha, 4 * (size_t)m, hipMemcpyDefault));
    DTB_TRY_HIP(hipMemcpy(hb.data(), beta, 4 * (size_t)m, hipMemcpyDefault));
    std::vector<float2> ab((size_t)m);
    for (int64_t r = 0; r < m; ++r) {
        DTB_REQUIRE(ha[(size_t)r] >= 0.0f && hb[(size_t)r] > 0.0f, "gamma needs alpha >= 0 and beta > 0");
        ab[(size_t)r] = make_float2(ha[(size_t)r], hb[(size_t)r]);
    }
    std::vector<int32_t> cols;
    std::vector<float2> table;
    DTB_TRY(group_requests<float2>(b, m, pop, col, ab.data(), make_float2(NAN, NAN), cols, table));
    void *dt, *dc;
    DTB_TRY_HIP(hipMalloc(&dt, sizeof(float2) * table.size()));
    DTB_TRY_HIP(hipMalloc(&dc, sizeof(int32_t) * cols.size()));
    int st = upload(b, dt, table.data(), sizeof(float2) * table.size());
    if (!st) st = upload(b, dc, cols.data(), sizeof(int32_t) * cols.size());
    if (!st) {
        hipLaunchKernelGGL(gamma_props_kernel, dim3(grid_for(b->n, 256)), dim3(256), 0, b->stream, b->n, b->S, b->sub,
                           (const float2*)dt, (const int32_t*)dc, (int32_t)cols.size(), seed, b->props);
        st = derive_props(b, std::any_of(cols.begin(), cols.end(), [](int32_t c) { return pset_column((uint32_t)c); }));
    }
    hipStreamSynchronize(b->stream);
    hipFree(dt);
    hipFree(dc);
    return st;
}

int dtb_block_set_weights(dtb_block* b, int64_t m, const uint32_t* pre, const uint32_t* post,
                          const uint8_t* channel, const float* w)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_TRY_HIP(hipSetDevice(b->device));
    return set_weights_impl(b, m, pre, post, channel, w);
}

int dtb_block_set_sources(dtb_block* b, int64_t n_src, const uint32_t* sources, const uint8_t* frozen,
                          int64_t n_events, const uint32_t* step, const uint32_t* neuron)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_TRY_HIP(hipSetDevice(b->device));
    return set_sources_impl(b, n_src, sources, frozen, n_events, step, neuron);
}

int dtb_block_replace_props(dtb_block* b, const float* props)
{
    DTB_REQUIRE(b != nullptr && props != nullptr, "block or props is NULL");
    DTB_TRY_HIP(hipSetDevice(b->device));
    DTB_TRY(upload(b, b->props, props, sizeof(float) * (size_t)b->n * DTB_NPROP));
    return derive_props(b);
}

int dtb_block_attach_bold(dtb_block* b, const double* par, int32_t steps_per_sample, int32_t nvox,
                          const int32_t* voxel_of_pop, const double* voxel_size)
{
    DTB_REQUIRE(b != nullptr && par != nullptr, "block or par is NULL");
    DTB_REQUIRE(steps_per_sample >= 1 && steps_per_sample <= b->kmax, "steps_per_sample must be in 1..%lld",
                (long long)b->kmax);
    DTB_REQUIRE(b->steps_done % steps_per_sample == 0, "attach the integrator at a multiple of steps_per_sample");
    DTB_TRY_HIP(hipSetDevice(b->device));
    const int32_t S = b->S;
    if (!voxel_of_pop) nvox = S;
    DTB_REQUIRE(nvox >= 1, "nvox must be >= 1");
    std::vector<int32_t> vop((size_t)S);
    for (int32_t p = 0; p < S; ++p) vop[(size_t)p] = p;
    if (voxel_of_pop) DTB_TRY_HIP(hipMemcpy(vop.data(), voxel_of_pop, 4 * (size_t)S, hipMemcpyDefault));
    std::vector<int32_t> vptr((size_t)nvox + 1, 0), vpops;
    for (int32_t p = 0; p < S; ++p) {
        DTB_REQUIRE(vop[(size_t)p] >= -1 && vop[(size_t)p] < nvox, "voxel of sub-population %d out of range", p);
        if (vop[(size_t)p] >= 0) vptr[(size_t)vop[(size_t)p] + 1]++;
    }
    for (int32_t v = 0; v < nvox; ++v) vptr[(size_t)v + 1] += vptr[(size_t)v];
    vpops.assign((size_t)std::max(vptr[(size_t)nvox], 1), 0);
    std::vector<int32_t> cur(vptr.begin(), vptr.end() - 1);
    for (int32_t p = 0; p < S; ++p)
        if (vop[(size_t)p] >= 0) vpops[(size_t)cur[(size_t)vop[(size_t)p]]++] = p;
    std::vector<double> vsz((size_t)nvox, 0.0);
    if (voxel_size) {
        DTB_TRY_HIP(hipMemcpy(vsz.data(), voxel_size, 8 * (size_t)nvox, hipMemcpyDefault));
    } else {
        for (int32_t p = 0; p < S; ++p)
            if (vop[(size_t)p] >= 0) vsz[(size_t)vop[(size_t)p]] += (double)b->pop_size[(size_t)p];
    }
    if (nvox > b->nvox_cap) {
        DTB_TRY(alloc_n(b, &b->bold_state, 6 * (int64_t)nvox + 8));
        DTB_TRY(alloc_n(b, &b->rec_bold, b->kmax * (int64_t)nvox));
        DTB_TRY(alloc_n(b, &b->vox_ptr, (int64_t)nvox + 1));
        DTB_TRY(alloc_n(b, &b->vox_size, nvox));
        b->nvox_cap = nvox;
    }
    if (!b->vox_pops) DTB_TRY(alloc_n(b, &b->vox_pops, S));
    std::vector<double> init((size_t)(6 * nvox + 8), 0.0);
    for (int32_t v = 0; v < nvox; ++v) {
        init[(size_t)(2 * nvox + v)] = 1.0; // v
        init[(size_t)(3 * nvox + v)] = 1.0; // f_in
    }
    for (int j = 0; j < 8; ++j) init[(size_t)(6 * nvox + j)] = par[j];
    std::memcpy(b->bold_par, par, sizeof(b->bold_par));
    DTB_TRY(upload(b, b->bold_state, init.data(), sizeof(double) * init.size()));
    DTB_TRY(upload(b, b->vox_ptr, vptr.data(), sizeof(int32_t) * vptr.size()));
    DTB_TRY(upload(b, b->vox_pops, vpops.data(), sizeof(int32_t) * std::min<size_t>(vpops.size(), (size_t)S)));
    DTB_TRY(upload(b, b->vox_size, vsz.data(), sizeof(double) * vsz.size()));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    b->nvox = nvox;
    b->bold_sps = steps_per_sample;
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    return DTB_OK;
}

int dtb_block_bold_state(dtb_block* b, int32_t which, int32_t dir, double* data)
{
    DTB_REQUIRE(b != nullptr && data != nullptr, "block or data is NULL");
    DTB_REQUIRE(b->bold_sps > 0, "no BOLD integrator attached");
    DTB_REQUIRE(which >= 0 && which < 6 && (dir == 0 || dir == 1), "which must be 0..5 and dir 0 or 1");
    DTB_TRY_HIP(hipSetDevice(b->device));
    double* src = b->bold_state + (int64_t)which * b->nvox;
    if (dir == 0)
        DTB_TRY_HIP(hipMemcpyAsync(data, src, sizeof(double) * (size_t)b->nvox, hipMemcpyDefault, b->stream));
    else
        DTB_TRY_HIP(hipMemcpyAsync(src, data, sizeof(double) * (size_t)b->nvox, hipMemcpyDefault, b->stream));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    return DTB_OK;
}

int dtb_block_set_external(dtb_block* b, int32_t n_thresholds, const uint32_t* thresholds, float weight)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_REQUIRE(n_thresholds >= 0 && n_thresholds <= kMaxExtThr, "at most 16 thresholds");
    DTB_REQUIRE(n_thresholds == 0 || thresholds != nullptr, "thresholds are required");
    DTB_REQUIRE(std::isfinite(weight), "weight must be finite");
    ExtDrive e{};
    e.n = n_thresholds;
    e.w = weight;
    for (int32_t j = 0; j < n_thresholds; ++j) {
        e.thr[j] = thresholds[j];
        DTB_REQUIRE(j == 0 || e.thr[j] >= e.thr[j - 1], "thresholds must be non-decreasing");
    }
    b->ext = e;
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    return DTB_OK;
}

int dtb_block_set_launch(dtb_block* b, int32_t what, int32_t value)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    if (what == DTB_LAUNCH_DELIVER_GRID) {
        DTB_REQUIRE(value >= 1 && value <= 65536, "delivery grid must be in 1..65536 workgroups");
        b->deliver_grid = value;
    } else if (what == DTB_LAUNCH_DELIVER_LANES) {
        DTB_REQUIRE(value == 16 || value == 32 || value == 64, "lanes per segment must be 16, 32 or 64");
        b->deliver_lanes = value;
    } else if (what == DTB_LAUNCH_NEURON_PSETS) {
        DTB_REQUIRE(value == 0 || value == 1, "parameter sets must be 0 (off) or 1 (auto)");
        b->pset_off = value == 0;
        DTB_TRY(build_psets(b));
    } else if (what == DTB_LAUNCH_STAGE_SEGS) {
        DTB_REQUIRE(value == 64 || value == 128 || value == 256, "segments per chunk must be 64, 128 or 256");
        b->stage_segs = value;
    } else if (what == DTB_LAUNCH_STAGE_EPT) {
        DTB_REQUIRE(value == 16 || value == 32, "events per lane per round must be 16 or 32");
        b->stage_ept = value;
    } else if (what == DTB_LAUNCH_ACCUM_FIXED) {
        DTB_REQUIRE(value == 0 || value == 1, "fixed-point accumulation must be 0 or 1");
        DTB_TRY(ensure_bucketed(b));
        DTB_REQUIRE(value == 0 || b->fixed_ok,
                    "fixed-point accumulation needs 2^S resolution with S >= 16 (largest per-step input sum < 16384)");
        DTB_REQUIRE(value == 1 || b->deliver_mode != 2, "persistent delivery sums in fixed point");
        b->accum_fixed = value != 0;
    } else if (what == DTB_LAUNCH_ACCUM_UNROLL) {
        DTB_REQUIRE(value == 8 || value == 16, "accumulation unroll must be 8 or 16");
        b->accum_unroll = value;
    } else if (what == DTB_LAUNCH_ACCUM_OCC) {
        DTB_REQUIRE(value == 1 || value == 2, "accumulation workgroups per CU must be 1 or 2");
        DTB_TRY(ensure_bucketed(b));
        DTB_TRY_HIP(hipStreamSynchronize(b->stream)); // cursors are all zero between steps
        DTB_TRY(use_bucket_geom(b, value));
        if (b->stage_variant == 2 && !b->sorted_ok) b->stage_variant = 0;
    } else if (what == DTB_LAUNCH_STAGE_VARIANT) {
        DTB_REQUIRE(value >= 0 && value <= 2, "staging variant must be 0, 1 or 2");
        if (value == 2) {
            DTB_TRY(ensure_bucketed(b));
            DTB_REQUIRE(b->sorted_ok, "too many delivery buckets for the sorted staging variant");
        }
        b->stage_variant = value;
    } else if (what == DTB_LAUNCH_SEGMENT_LISTS) {
        DTB_REQUIRE(value >= 0, "segment list capacity must be >= 0");
        DTB_TRY(ensure_bucketed(b));
        if (value > 0) DTB_TRY(ensure_segments(b));
        b->seg_cap = value == 0 ? 0 : value == 1 ? b->seg_alloc : std::min<int64_t>(value, b->seg_alloc);
        b->seg_dirty = true;
    } else if (what == DTB_LAUNCH_DELIVER_MODE) {
        DTB_REQUIRE(value >= 0 && value <= 2,
                    "delivery mode must be 0 (atomics), 1 (bucketed) or 2 (persistent, small blocks)");
        if (value >= 1) DTB_TRY(ensure_bucketed(b));
        if (value == 2) {
            DTB_TRY(ensure_persist(b));
            b->accum_fixed = true; // batches that fall back to bucketed kernels sum as the persistent launch does
        }
        b->deliver_mode = value;
    } else {
        set_error("unknown launch parameter %d", what);
        return DTB_INVALID_PARAMETER;
    }
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    return DTB_OK;
}

int dtb_block_get_launch(dtb_block* b, int32_t what, int32_t* value)
{
    DTB_REQUIRE(b != nullptr && value != nullptr, "block or value is NULL");
    if (what == DTB_LAUNCH_DELIVER_GRID) *value = b->deliver_grid;
    else if (what == DTB_LAUNCH_DELIVER_LANES) *value = b->deliver_lanes;
    else if (what == DTB_LAUNCH_DELIVER_MODE) *value = b->deliver_mode;
    else if (what == DTB_LAUNCH_NEURON_PSETS) *value = b->npsets;
    else if (what == DTB_LAUNCH_STAGE_SEGS) *value = b->stage_segs;
    else if (what == DTB_LAUNCH_STAGE_EPT) *value = b->stage_ept;
    else if (what == DTB_LAUNCH_STAGE_VARIANT) *value = b->stage_variant;
    else if (what == DTB_LAUNCH_ACCUM_OCC) *value = b->acc_occ;
    else if (what == DTB_LAUNCH_ACCUM_UNROLL) *value = b->accum_unroll;
    else if (what == DTB_LAUNCH_ACCUM_FIXED) *value = b->accum_fixed ? b->acc_shift : 0;
    else if (what == DTB_LAUNCH_SEGMENT_LISTS) *value = (int32_t)std::min<int64_t>(b->seg ? b->seg_cap : 0, INT32_MAX);
    else {
        set_error("unknown launch parameter %d", what);
        return DTB_INVALID_PARAMETER;
    }
    return DTB_OK;
}

int dtb_block_set_samples(dtb_block* b, int64_t m, const uint32_t* idx)
{
    DTB_REQUIRE(b != nullptr, "block is NULL");
    DTB_REQUIRE(m >= 0 && (m == 0 || idx), "bad sample list");
    DTB_TRY_HIP(hipSetDevice(b->device));
    std::vector<uint32_t> h((size_t)m);
    if (m) DTB_TRY_HIP(hipMemcpy(h.data(), idx, 4 * (size_t)m, hipMemcpyDefault));
    for (uint32_t x : h) DTB_REQUIRE(x < (uint64_t)b->n, "sample index %u out of range", x);
    if (m > b->ns) {
        // new buffers; the old ones stay owned by the block until destroy
        DTB_TRY(alloc_n(b, &b->samples, m));
        DTB_TRY(alloc_n(b, &b->rec_sspk, m * b->kmax));
        DTB_TRY(alloc_n(b, &b->rec_sv, m * b->kmax));
        for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
    }
    b->ns = m;
    if (m) DTB_TRY(upload(b, b->samples, h.data(), 4 * (size_t)m));
    DTB_TRY_HIP(hipStreamSynchronize(b->stream));
    return DTB_OK;
}

} // extern "C"

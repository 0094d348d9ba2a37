// Block state shared by the engine (dtb_block.hip) and the network
// generator (dtb_gen.hip).
#pragma once

#include <map>
#include <vector>

#include "dtb_internal.hpp"

namespace dtb {

// Step bookkeeping kept on the device so captured graphs replay without
// host arguments: nbase = steps done before the current batch.
struct StepCtl {
    int64_t nsteps;
    int64_t nbase;
    double t;
    double pad;
};

struct Block {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0, nnz = 0;
    int64_t max_row = 0; // longest presynaptic row
    int32_t S = 0;    // statistics sub-populations: max(sub_idx)+1
    int32_t D = 1;    // ring depth = largest delay
    bool pair = false; // float2 weights (else one weight per synapse)
    SynPack pk{};
    double dt = 0.1;
    uint64_t seed = 0;
    int stats_excl = 0;

    // properties: master copy [n][22] and the derived per-neuron arrays
    float* props = nullptr;
    float4 *g = nullptr, *vu = nullptr, *decay = nullptr;
    float *gL = nullptr, *vL = nullptr, *dtC = nullptr, *vth = nullptr, *vrst = nullptr, *iext = nullptr;
    float *tref = nullptr, *noisep = nullptr;
    int32_t* pop = nullptr; // statistics sub-population (-1: excluded)
    uint8_t* pset = nullptr; // parameter set of each neuron (when npsets > 0)
    float* ptab = nullptr;   // [64][16] parameter-set table
    int32_t npsets = 0;
    bool pset_off = false;
    int32_t* sub = nullptr; // raw sub_idx (property column 3)

    // dynamic state
    float* V = nullptr;
    float4* Jp = nullptr;          // J decayed to the current step, arrivals of the current step not added
    int32_t* ref_end = nullptr;    // first step that is not refractory
    int32_t* last_step = nullptr;  // step of the last threshold spike (0: none)
    double* t_last = nullptr;      // time of the last threshold spike
    uint8_t* flags = nullptr;      // bit0 threshold spike, bit1 delivered spike (last step)
    void* acc = nullptr;           // [n] arrivals of the current step: float2 (E,I) or float4 (AMPA..GABAb)

    // presynaptic-major CSR, each row sorted by delay
    int64_t* rowptr = nullptr;
    uint16_t* dend = nullptr; // [n][D] end of each delay segment relative to the row start (D > 1)
    uint32_t* col = nullptr; // packed (post, delay-1, kind)
    float* w = nullptr;      // compact weights
    float2* w2 = nullptr;    // pair weights

    // spike lists of the last D+1 steps (slot = step mod (D+1))
    int32_t* spk = nullptr;      // [D+1][sub-lists][spk_cap]
    uint32_t* spk_cnt = nullptr; // [D+1][sub-lists], padded
    int64_t spk_cap = 0;         // entries per sub-list

    StepCtl* ctl = nullptr;
    double* ttab = nullptr; // [kmax] time of each step of the batch
    int64_t kmax = 0;

    // per-step records of the current batch
    uint32_t* rec_freq = nullptr; // [kmax][S]
    float* rec_vsum = nullptr;    // [kmax][S]
    float* rec_isum = nullptr;    // [kmax][S]
    uint64_t* rec_events = nullptr; // [kmax]
    uint32_t* rec_nspk = nullptr;   // [kmax]
    uint8_t* rec_sspk = nullptr;    // [kmax][ns]
    float* rec_sv = nullptr;        // [kmax][ns]
    float* pop_inv = nullptr;       // [S] 1/size (NaN for empty)

    // attached BOLD integrator (per sub-population), FP64
    int32_t bold_sps = 0;
    double bold_par[8] = {0};
    int32_t nvox = 0, nvox_cap = 0;
    double* bold_state = nullptr; // [6][nvox]: s, q, v, f_in, bold, u; then the 8 parameters
    double* rec_bold = nullptr;   // [kmax][nvox]
    int32_t* vox_ptr = nullptr;   // [nvox+1] sub-populations of each voxel (CSR)
    int32_t* vox_pops = nullptr;  // [S]
    double* vox_size = nullptr;   // [nvox] neurons per voxel (input divisor)
    double* pop_sized = nullptr;  // [S] sizes as double
    std::vector<hipEvent_t> events;

    uint32_t* samples = nullptr;
    int64_t ns = 0;

    // forced-spike sources: src[i] bit0 = source (threshold spikes from the
    // schedule), bit1 = frozen (no membrane dynamics, no noise), bit2 = fires
    // at the current step (set by the mark kernel, cleared by the neuron
    // kernel); schedule: nts ascending step indices src_ts, neurons
    // src_ids[src_ptr[k] .. src_ptr[k+1]) fire at the step whose 0-based index is src_ts[k]
    uint8_t* src = nullptr;
    uint32_t* src_ts = nullptr;
    int64_t* src_ptr = nullptr;
    uint32_t* src_ids = nullptr;
    int64_t src_nts = 0;
    int64_t src_max = 0; // most neurons of one timestamp

    std::vector<int64_t> pop_size;
    int64_t steps_done = 0;
    double t_host = 0.0;
    int32_t deliver_grid = 64;  // workgroups of the delivery kernel
    int32_t deliver_lanes = 64; // lanes per (row, delay) segment
    int32_t deliver_mode = 0;   // 0: global atomics, 1: bucketed (staged + LDS accumulation)
    int32_t stage_segs = 256;   // bucketed: (row, delay) segments per staging chunk
    int32_t stage_ept = 32;     // bucketed: events per lane per staging round
    int32_t stage_variant = 0;  // bucketed: 0 scattered stores, 1 two passes per chunk, 2 LDS-sorted rounds
    bool sorted_ok = false;     // the sorted variant's LDS fits
    int32_t acc_occ = 1;        // bucketed: accumulation workgroups per CU (bucket size follows)
    int32_t accum_unroll = 8;   // bucketed: staged events in flight per accumulation lane
    bool accum_fixed = false;   // bucketed: int32 fixed-point accumulation of the scaled weights
    bool fixed_ok = false;      // no neuron's largest per-step sum can overflow int32
    int32_t acc_shift = 0;      // smallest input exponent over neurons with inputs (coarsest quantum 2^-acc_shift)
    // Input exponents: for neuron i and receptor channel c, e = jexp[4i + c] is
    // chosen so that the sum of |w| over i's inputs on c, times 2^e, lies in
    // [2^29, 2^30].  CSR weights, Jp and acc hold values x 2^e and g holds
    // g x 2^-e (exact power-of-two scalings), so an int32 sum of rounded
    // scaled weights resolves every neuron relative to its own input scale.
    int8_t* jexp = nullptr;
    int32_t ncu = 0;            // compute units of the device
    int32_t lds_max = 0;        // LDS bytes a workgroup may use
    ExtDrive ext{};             // external Poisson drive (n = 0: off)
    int32_t bsize = 0;          // bucketed: neurons per destination bucket
    uint64_t bmagic = 0;        // bucket = (post * bmagic) >> bshift
    int32_t bshift = 0;
    int32_t nbk = 0;            // destination buckets
    int64_t stage_cap = 0;      // staged events per bucket
    uint32_t* cursor = nullptr; // [nbk]
    uint32_t* ovf = nullptr;    // [nbk]
    void* stage = nullptr;      // [nbk][stage_cap]
    void* acc2 = nullptr;       // [n] overflow accumulator

    // segment lists (bucketed delivery, D > 1): every spike's non-empty
    // (row, delay) segments, written once after the neuron kernel into the
    // list of the step that delivers them
    uint64_t* seg = nullptr;      // [D][seg_alloc]: (row start + offset) << 24 | length
    uint32_t* seg_cnt = nullptr;  // [D + 1][32 words]: entries, overflow flag; entry D: first complete step (int64)
    int64_t seg_alloc = 0;        // entries allocated per list
    int64_t seg_cap = 0;          // entries used per list (0: segment lists off)
    bool seg_dirty = true;        // the lists must be restarted before the next bucketed batch

    // persistent delivery (deliver_mode 2): the CSR grouped by owning
    // workgroup (each workgroup's synapses contiguous, in (row, delay) segment
    // order), rebuilt when the weights change
    int32_t* pbptr = nullptr;   // [pG][n*D+1] start of workgroup g's synapses of each segment
    uint32_t* pcol = nullptr;   // [nnz] records: local index | kind << 31, round(scaled weight) (x2 pair, + pad)
    uint32_t* pbar = nullptr;   // [0] barrier counter, [64] abort flag
    int32_t pG = 0, pm = 0;     // workgroups, 256-neuron blocks of LDS slice per workgroup
    int32_t pper = 0;           // neurons per workgroup (consecutive ranges)
    bool persist_dirty = true;

    std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
    std::vector<void*> allocs;
};

int block_alloc(Block* b, void** p, size_t bytes);
int derive_props(Block* b, bool psets_dirty = true);
int gen_fixed_indegree(Block* b, const dtb_fixed_indegree_spec* spec);
int gen_members(Block* b, const dtb_member_spec* mem, int32_t M, int32_t max_delay, float w_lo, float w_hi,
                uint64_t seed, const uint64_t* cdf, int32_t R0);

} // namespace dtb

/*
 * dtb.h -- C ABI of the MI355X Digital Twin Brain network-step library
 * (libdtb.so, built from the .hip sources in digital_twin_brain_amd/csrc for
 * gfx950).
 *
 * Plain C types only: pointers, sizes and POD structs.  Every buffer
 * argument may be host memory or device memory (the library copies with
 * hipMemcpyDefault); device pointers avoid the PCIe copy.  Calls on one
 * block handle are not re-entrant; separate handles are independent.
 * No entry point exits the process: each returns a status code and leaves a
 * thread-local message for dtb_last_error().
 */
#ifndef DTB_H_
#define DTB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  0..3 keep the values of SnnStatus
 * (cuda/protos/snn.proto:36-41); 4 and 5 are new. */
#define DTB_OK 0
#define DTB_INVALID_PARAMETER 1
#define DTB_UNSET_SAMPLE_FAILURE 2
#define DTB_UNSUPPORTED 3
#define DTB_RUNTIME_ERROR 4
#define DTB_OUT_OF_MEMORY 5

/* Number of property columns per neuron: noise_rate, blocked_in_stat, I_ext,
 * sub_idx, C, T_ref, g_L, V_L, V_th, V_reset, g_ui[4], V_ui[4], tau_ui[4]
 * (generation/make_block.py:394-434, models/block.py:156). */
#define DTB_NPROP 22

/* Largest synaptic delay in steps. */
#define DTB_MAX_DELAY 64

/* Thread-local message of the last failing call (replaces the reference's
 * exit() on error, cuda/src/common.cpp:199-209). */
const char* dtb_last_error(void);
int32_t dtb_abi_version(void);

typedef struct dtb_block dtb_block;

/* A network given as presynaptic-major CSR: row j lists the synapses of
 * presynaptic neuron j (cuda/src/configuration.cpp:747-828).  kind 0 feeds
 * the AMPA/NMDA pair with (w0, w1), kind 1 the GABAa/GABAb pair
 * (generation/read_block.py:68-74).  w1 == NULL stores one weight per
 * synapse shared by the pair ("compact", 8 B/synapse). */
typedef struct {
    int64_t n;                     /* neurons */
    int64_t nnz;                   /* synapses */
    const float* props;            /* [n][22] */
    const int64_t* rowptr;         /* [n+1] */
    const int32_t* post;           /* [nnz] postsynaptic neuron */
    const uint8_t* kind;           /* [nnz] 0 or 1 */
    const uint8_t* delay;          /* [nnz] steps in 1..DTB_MAX_DELAY, NULL = all 1 */
    const float* w0;               /* [nnz] */
    const float* w1;               /* [nnz] or NULL */
    double dt;                     /* ms */
    uint64_t noise_seed;           /* Philox key of the background noise */
    int32_t stats_exclude_blocked; /* 1: property column 1 != 0 excludes a neuron from statistics */
    int32_t device;
} dtb_block_desc;

/* block(node_property, w_uij, delta_t) -- models/block.py:30-60, with the
 * connectivity already converted to CSR; also the engine init of
 * cuda/src/block.cpp:37-193 (parse_conn_table + parse_params). */
int dtb_block_create(dtb_block** out, const dtb_block_desc* desc);

/* Seeded fixed in-degree network generated on the GPU (the synthetic
 * counterpart of generation/make_block.py:206-323): for post i and slot k
 * one Philox4x32-10 call keyed by seed gives the presynaptic neuron
 * (uniform, no self-connection), w0, the delay (uniform 1..max_delay) and
 * w1.  kind = (pre >= n_exc).  Rows are sorted by (post, k). */
typedef struct {
    int64_t n;
    int64_t n_exc;
    int32_t in_degree;
    int32_t max_delay;
    float w_lo, w_hi;
    int32_t pair_weights; /* 0: compact (one weight per synapse) */
    uint64_t seed;
} dtb_fixed_indegree_spec;

int dtb_block_create_fixed_indegree(dtb_block** out, const dtb_fixed_indegree_spec* spec, const float* props,
                                    double dt, uint64_t noise_seed, int32_t stats_exclude_blocked,
                                    int32_t device);

/* A network of members packed in one block (a data-assimilation ensemble of
 * uneven sizes, or the regions of one GPU's share of a whole-brain model):
 * member m owns the next n neurons, the first n_exc excitatory, each
 * receiving in_degree synapses from inside the member with the draws of
 * dtb_block_create_fixed_indegree, and inter_degree more from other members
 * (dtb_block_create_regions).  No reference counterpart for the ensemble:
 * the reference replicates equal members inside one network
 * (data_assimilation/DataAssimilation.py:97); regions follow the
 * population/region layout of generation/make_block.py:206-434. */
typedef struct {
    int64_t n;
    int64_t n_exc;
    int32_t in_degree;
    int32_t inter_degree; /* 0 for dtb_block_create_members */
} dtb_member_spec;

int dtb_block_create_members(dtb_block** out, const dtb_member_spec* members, int32_t n_members,
                             int32_t max_delay, float w_lo, float w_hi, int32_t pair_weights, uint64_t seed,
                             const float* props, double dt, uint64_t noise_seed, int32_t stats_exclude_blocked,
                             int32_t device);

/* Members with inter-member synapses: inter slot k of a neuron of member m
 * draws its source member s as the smallest s with q < source_cdf[m*M + s]
 * (q a 32-bit Philox draw, ctr = (k, post_lo, post_hi, 0x52454749)), then a
 * presynaptic neuron uniformly inside s.  source_cdf: [M][M] cumulative
 * source probabilities scaled to 2^32, rows non-decreasing and ending at
 * exactly 2^32 (host or device memory).  The reference draws inter-region
 * partners from a connectome in generation/make_block.py:206-330. */
int dtb_block_create_regions(dtb_block** out, const dtb_member_spec* members, int32_t n_members,
                             const uint64_t* source_cdf, int32_t max_delay, float w_lo, float w_hi,
                             int32_t pair_weights, uint64_t seed, const float* props, double dt,
                             uint64_t noise_seed, int32_t stats_exclude_blocked, int32_t device);

/* `replicas` copies of one region network back to back (replica r holds
 * regions r*n_regions .. (r+1)*n_regions - 1): every replica draws its
 * inter-region sources through the same source_cdf [n_regions][n_regions],
 * inside itself.  The same network as dtb_block_create_regions with the
 * block-diagonal source table of all replicas, without that table's
 * (replicas*n_regions)^2 entries.  For ensembles of whole-brain members
 * (data_assimilation/DataAssimilation.py:97 replicates equal members); props
 * holds the rows of all replicas. */
int dtb_block_create_region_replicas(dtb_block** out, const dtb_member_spec* regions, int32_t n_regions,
                                     int32_t replicas, const uint64_t* source_cdf, int32_t max_delay, float w_lo,
                                     float w_hi, int32_t pair_weights, uint64_t seed, const float* props, double dt,
                                     uint64_t noise_seed, int32_t stats_exclude_blocked, int32_t device);

int dtb_block_destroy(dtb_block* b);

/* Noise modes of dtb_block_run. */
#define DTB_NOISE_SCALAR 0    /* Bernoulli(noise_rate) per neuron (block.run, models/block.py:136) */
#define DTB_NOISE_PER_NEURON 1 /* Bernoulli(property column 0) (cuda/src/stage/spike.cpp:156-157) */
#define DTB_NOISE_UNIFORMS 2  /* spike iff uniforms[step][i] < noise_rate (injected noise) */

typedef struct {
    int32_t noise_mode;
    float noise_rate;
    const float* uniforms; /* DTB_NOISE_UNIFORMS: [nsteps][n] */
    int32_t isolated;      /* 1: only noise spikes reach synapses (models/block.py:137-138) */
    int32_t use_graph;     /* 1: replay captured hipGraphs of whole step batches */
    /* per-step outputs, NULL = not wanted (host or device memory) */
    uint32_t* freqs;       /* [nsteps][S] threshold spikes per sub-population */
    float* vmeans;         /* [nsteps][S] mean V per sub-population */
    float* imeans;         /* [nsteps][S] mean I_syn per sub-population */
    uint8_t* sample_spike; /* [nsteps][n_samples] delivered spike of sampled neurons */
    float* sample_v;       /* [nsteps][n_samples] V of sampled neurons */
    uint64_t* events;      /* [nsteps] synaptic events delivered (sum of out-degrees of spiking rows) */
    uint32_t* spikes;      /* [nsteps] spiking rows delivered */
    double* bold;          /* [nsteps/steps_per_sample][nvox] BOLD of the attached integrator */
    float* kernel_ms;      /* [nsteps][2] device time of the neuron and delivery kernels of each step,
                              measured with hipEvents on the block's stream (direct launches) */
} dtb_run_args;

/* nsteps of block.run (models/block.py:116-150) / BrainBlock's step
 * sequence (cuda/tools/dist_simulator.cpp:3822-4114); returns when the
 * outputs are written. */
int dtb_block_run(dtb_block* b, int64_t nsteps, const dtb_run_args* args);

/* Attach a Balloon-Windkessel integrator over nvox voxels, each the union
 * of the sub-populations p with voxel_of_pop[p] == voxel (-1: none; NULL:
 * one voxel per sub-population).  Every steps_per_sample steps,
 * u = max(threshold spikes of the voxel in the window / voxel_size, 1e-5)
 * drives one Euler step of par[7] seconds, in FP64 -- the per-step merge and
 * BOLD call of simulation/simulation.py:269-279 with
 * models/bold_model_pytorch.py:43-78.  voxel_size NULL: neurons counted in
 * statistics.  The state starts at s=0, q=0, v=1, f_in=1
 * (models/bold_model_pytorch.py:60-67).  Runs must then advance a multiple
 * of steps_per_sample steps. */
int dtb_block_attach_bold(dtb_block* b, const double* par, int32_t steps_per_sample, int32_t nvox,
                          const int32_t* voxel_of_pop, const double* voxel_size);

/* Read (dir=0) or write (dir=1) the attached integrator's state:
 * which = 0 s, 1 q, 2 v, 3 f_in, 4 last BOLD, 5 last input u; [nvox] doubles
 * (BOLD.state_update, models/bold_model_pytorch.py:34-41, writes 0..3). */
int dtb_block_bold_state(dtb_block* b, int32_t which, int32_t dir, double* data);

/* State fields for dtb_block_get. */
#define DTB_FIELD_V 0          /* float[n] */
#define DTB_FIELD_J 1          /* float[4][n] J_ui */
#define DTB_FIELD_I_UI 2       /* float[4][n] */
#define DTB_FIELD_I_SYN 3      /* float[n] */
#define DTB_FIELD_ACTIVE 4     /* uint8[n] threshold spikes of the last step */
#define DTB_FIELD_SPIKE 5      /* uint8[n] delivered spikes of the last step */
#define DTB_FIELD_SPIKE_LIST 6 /* int32[count] delivered spike list of the last step, ascending */
#define DTB_FIELD_PROPS 7      /* float[n][22] */
#define DTB_FIELD_ROWPTR 8     /* int64[n+1] */
#define DTB_FIELD_POST 9       /* int32[nnz] */
#define DTB_FIELD_KIND 10      /* uint8[nnz] */
#define DTB_FIELD_DELAY 11     /* uint8[nnz] */
#define DTB_FIELD_W0 12        /* float[nnz] */
#define DTB_FIELD_W1 13        /* float[nnz] (== W0 for compact weights) */
#define DTB_FIELD_T_LAST 14    /* double[n] time of the last threshold spike */
#define DTB_FIELD_POP_SIZES 15 /* int64[S] neurons per sub-population counted in statistics */

/* Copies a field (block.V_i, .J_ui, .I_ui, .I_syn, .active, models/block.py)
 * into dst; bytes must be at least the field size.  For SPIKE_LIST the
 * byte count of the list is returned through *bytes_out. */
int dtb_block_get(dtb_block* b, int32_t field, void* dst, int64_t bytes, int64_t* bytes_out);

/* info[0..8] = n, nnz, S (sub-populations), max delay, steps done,
 * n_samples, pair weights (0/1), device, BOLD voxels. */
int dtb_block_info(dtb_block* b, int64_t* info);
/* Simulated time t (models/block.py:58,133), accumulated in double. */
int dtb_block_time(dtb_block* b, double* t);

/* State checkpoint and restore.  No counterpart in the reference, which keeps
 * no V, J or t_last across runs (its only offset, RunRequest.iter_offset in
 * cuda/tools/dist_simulator.cpp:3811-3820, re-aligns input spike times); for
 * restartable data-assimilation windows (data_assimilation/DataAssimilation.py:402-453).
 * A snapshot holds everything the next steps read: V, J, the arrivals of the
 * current step, refractory ends, t_last, spike flags, the spike lists of the
 * last D+1 steps (the synaptic delays in flight), the step counter and
 * simulated time, the noise seed (noise is counter-based: seed and step fix
 * it) and, when attached, the BOLD integrator's state.  Properties,
 * connectivity, the external drive's settings and the launch shape are not
 * part of it.  dst and src may be
 * host or device memory.  Loading checks that the snapshot comes from a
 * block of the same size, maximum delay, weight form and BOLD attachment;
 * after it the block continues exactly as the saved one did. */
int dtb_block_state_bytes(dtb_block* b, int64_t* bytes);
int dtb_block_save_state(dtb_block* b, void* dst, int64_t bytes);
int dtb_block_load_state(dtb_block* b, const void* src, int64_t bytes);

/* Point-wise property assignment, cuda/src/stage/property.cpp:17-120:
 * props[neuron[i]][col[i]] = val[i] (col is a property column 0..21). */
int dtb_block_set_props(dtb_block* b, int64_t m, const uint32_t* neuron, const uint32_t* col, const float* val);

/* Multiply a property column over every neuron of a sub-population,
 * cuda/src/stage/property.cpp:145-256 (mul_property_by_subblk). */
int dtb_block_mul_props_by_pop(dtb_block* b, int64_t m, const uint32_t* pop, const uint32_t* col,
                               const float* val);

/* Draw a property column from Gamma(alpha, beta) (shape, rate) for every
 * neuron of a sub-population, cuda/src/stage/property.cpp:347-471
 * (Marsaglia-Tsang on Philox4x32-10 keyed by (seed, neuron)). */
int dtb_block_gamma_props_by_pop(dtb_block* b, int64_t m, const uint32_t* pop, const uint32_t* col,
                                 const float* alpha, const float* beta, uint64_t seed);

/* Forced-spike sources: the threshold spikes of the listed neurons come from
 * a schedule instead of their membrane potential -- block(src=...)
 * (models/block.py:38-50,104-109) and the input blocks of the block npz
 * format (src_timestamp, src_neuron_idx: cuda/src/configuration.cpp:753-762,
 * timed injection in cuda/src/block.cpp:195-221).  Event k fires neuron[k]
 * at the step whose 0-based index since the block's creation is step[k]
 * (the first step run has index 0, as the reference's iter and timestamps
 * count); at every other step a source does not fire.  A source keeps its
 * membrane dynamics and background noise (block.py); a frozen source
 * (frozen[k] != 0, an input block's neuron) keeps its V and draws no noise.
 * Refractoriness, t_last and the statistics follow the forced spikes.  Each
 * call replaces the previous sources and schedule; n_src = 0 removes them.
 * frozen may be NULL (none frozen). */
int dtb_block_set_sources(dtb_block* b, int64_t n_src, const uint32_t* sources, const uint8_t* frozen,
                          int64_t n_events, const uint32_t* step, const uint32_t* neuron);

/* Change existing synaptic weights: w_uij[channel][post][pre] = w for each
 * request (block.update_conn_weight, models/block.py:172-177, with the
 * request as the (channel, post, pre) index of w_uij).  Channels 0/1 are
 * the AMPA/NMDA weights of a kind-0 synapse, 2/3 the GABAa/GABAb weights of
 * a kind-1 synapse.  Where several synapses join one pair on one receptor
 * group, the first takes w and the others 0, so they sum to w as the dense
 * w_uij entry does; the last of several requests for the same entry wins.
 * A block with compact weights switches to pair weights when a call sets
 * the two channels of a pair differently (or only one of them).  A request
 * naming no existing synapse fails with DTB_INVALID_PARAMETER and changes
 * nothing.  The dynamic state continues unchanged. */
int dtb_block_set_weights(dtb_block* b, int64_t m, const uint32_t* pre, const uint32_t* post, const uint8_t* channel,
                          const float* w);

/* Replace every property (block.update_property, models/block.py:152-170). */
int dtb_block_replace_props(dtb_block* b, const float* props);

/* Sampled neurons for sample_spike / sample_v (BlockWrapper.set_samples,
 * cuda/python/dist_blockwrapper_pytorch.py:505-528). */
int dtb_block_set_samples(dtb_block* b, int64_t m, const uint32_t* idx);

/* External Poisson drive (spikes from regions outside the block, no
 * reference counterpart in block.py): at every step each neuron receives
 * count * weight on its excitatory channels (AMPA and NMDA), as part of J from
 * the next step on, like that step's synaptic arrivals.  count =
 * #{j : r >= thresholds[j]}, r a 32-bit Philox draw keyed by the noise seed
 * (ctr = (i_lo, i_hi, step, 0x45585420)); thresholds[j] = 2^32 P(X <= j) of
 * the wanted count distribution (a Poisson CDF for a Poisson stream),
 * non-decreasing, at most 16.  n_thresholds = 0 turns the drive off. */
int dtb_block_set_external(dtb_block* b, int32_t n_thresholds, const uint32_t* thresholds, float weight);

/* Launch shape of the delivery kernel (no counterpart in the reference,
 * whose launch shapes are fixed, cuda/include/common.hpp:70-71): the
 * workgroup count, the lanes that share one (row, delay) segment and the
 * delivery mode.  Used by the autotuner; results do not depend on them
 * beyond floating-point summation order. */
#define DTB_LAUNCH_DELIVER_GRID 0  /* 1..65536 */
#define DTB_LAUNCH_DELIVER_LANES 1 /* 16, 32 or 64 */
#define DTB_LAUNCH_DELIVER_MODE 2  /* 0: global atomics, 1: bucketed LDS accumulation, 2: persistent (small
                                      blocks: one launch per step batch with a grid barrier per step, at most
                                      32 workgroups; fixed-point sums, results equal to mode 1; batches that
                                      need injected uniforms, sources, the external drive, I_syn means, samples
                                      or kernel times run on the mode-1 kernels) */
#define DTB_LAUNCH_NEURON_PSETS 3  /* set: 0 per-neuron constants, 1 shared parameter sets when <= 64
                                      distinct; get: number of sets in use (0 = per-neuron) */
#define DTB_LAUNCH_STAGE_SEGS 4    /* bucketed delivery: at most 64, 128 or 256 (row, delay) segments per chunk (the
                                      step's segments are split into equal chunks, the same count per workgroup) */
#define DTB_LAUNCH_STAGE_EPT 5     /* bucketed delivery: 16 or 32 events per lane per staging round */
#define DTB_LAUNCH_STAGE_VARIANT 6 /* bucketed delivery staging: 0 a reservation per round, lanes store their own
                                      events; 1 a reservation per chunk (two passes); 2 rounds sorted by
                                      bucket in LDS, stored as contiguous runs (16 events per lane) */
#define DTB_LAUNCH_ACCUM_OCC 7     /* bucketed delivery: 1 or 2 accumulation workgroups per CU (the
                                      destination bucket is the LDS slice divided by this) */
#define DTB_LAUNCH_ACCUM_UNROLL 8  /* bucketed delivery: 8 or 16 staged events in flight per accumulation lane */
#define DTB_LAUNCH_ACCUM_FIXED 9   /* bucketed delivery: set 1 to sum arrivals as int32 round(w * 2^S), 0 for fp32
                                      atomics; S is the largest shift keeping every neuron's largest possible
                                      per-step sum (sum of |w| over its inputs) below 2^30, allowed when S >= 16
                                      (the default then); get returns S when on, else 0 */
#define DTB_LAUNCH_SEGMENT_LISTS 10 /* bucketed delivery with delays (max delay > 1): 1 (the default) writes each
                                       spike's (row, delay) segments once, after the neuron kernel, into the list
                                       of the step that delivers them, so the stage kernel reads them coalesced
                                       instead of the CSR's row pointer and segment ends at every delay; 0 turns
                                       this off; a value > 1 caps the entries per list (a step whose list is full
                                       reads the spike lists instead); get returns the cap in use, 0 when off */
int dtb_block_set_launch(dtb_block* b, int32_t what, int32_t value);
int dtb_block_get_launch(dtb_block* b, int32_t what, int32_t* value);

/* Balloon-Windkessel forward-Euler step on device arrays of n voxels,
 * models/bold_model_pytorch.py:43-78.  par = {epsilon, tao_s, tao_f, tao_0,
 * alpha, E_0, V_0, delta_t}.  stream: hipStream_t or NULL. */
int dtb_bold_step_f32(int64_t n, const double* par, float* s, float* q, float* v, float* f, const float* u,
                      float* out, void* stream);
int dtb_bold_step_f64(int64_t n, const double* par, double* s, double* q, double* v, double* f,
                      const double* u, double* out, void* stream);
/* `steps` calls of dtb_bold_step_f64 with the same input u in one launch (the
 * state kept in registers; out = the last step's BOLD).  No counterpart in the
 * reference: the whole-brain DA's rate surrogate integrates a window's
 * hemodynamics from a predicted window-mean input with it. */
int dtb_bold_run_f64(int64_t n, int32_t steps, const double* par, double* s, double* q, double* v, double* f,
                     const double* u, double* out, void* stream);

/* Diffusion EnKF analysis in FP64 on device arrays,
 * data_assimilation/DataAssimilation.py:40-60.  w: [E][V][S] updated in
 * place (S = hyper-parameters + 6 hemodynamic states, last = BOLD);
 * y: [V] observation; xi: [E][V] standard-normal draws. */
int dtb_enkf_analysis_f64(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y, double rho,
                          const double* xi, void* stream);

/* The same analysis without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The scratch is a fixed per-GPU device array
 * of 2^19 doubles (4 MiB), held until the process ends; sizes whose scratch
 * (V*S + E*V*S + 2*E*V + V*S + E*E doubles) exceeds it return
 * DTB_INVALID_PARAMETER.  Asynchronous on `stream`; every call on one GPU
 * shares that GPU's scratch, so calls that overlap on one GPU (other streams,
 * other threads) race. */
int dtb_enkf_analysis_f64_noalloc(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y,
                                  double rho, const double* xi, void* stream);

/* Learned surrogate kernels (no counterpart in the reference: the surrogate
 * is new in this library).  bf16 operands are raw uint16 bit patterns on the
 * device.  All calls are asynchronous on `stream` (hipStream_t or NULL). */

/* C[M,N] = act(A[M,K] * B[N,K]^T + bias[N]) with bf16 A, B and FP32
 * accumulation on the matrix cores.  M and N multiples of 32, K a multiple
 * of 16, lda/ldb multiples of 8 (elements), A and B 16-byte aligned.
 * act: 0 identity, 1 ReLU.  bias may be NULL.  Writes C (FP32, ldc) and/or
 * Cb (bf16, ldcb); either may be NULL but not both. */
int dtb_mlp_gemm_bf16_nt(int32_t M, int32_t N, int32_t K, const void* A, int64_t lda, const void* B, int64_t ldb,
                         const float* bias, int32_t act, float* C, int64_t ldc, void* Cb, int64_t ldcb, void* stream);
/* dZ = dY where Z > 0, else 0 (FP32 and/or bf16 output) */
int dtb_mlp_relu_backward(int64_t n, const float* dY, const float* Z, float* dZ, void* dZb, void* stream);
/* db[j] = sum_i dZ[i*N + j] */
int dtb_mlp_bias_grad(int32_t M, int32_t N, const float* dZ, float* db, void* stream);
/* grad = scale * mask * (out - target); lsum[b] = sum of (mask*(out-target))^2
 * over the 256 elements of block b (ceil(n/256) entries; may be NULL). */
int dtb_mlp_mse_grad(int64_t n, const float* out, const float* target, const float* mask, float scale, float* grad,
                     void* gradb, float* lsum, void* stream);
/* one Adam step (t >= 1) on FP32 weights w with moments m, v; refreshes the
 * bf16 copy wb when not NULL */
int dtb_mlp_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                 int32_t t, void* wb, void* stream);
/* out[N,M] = in[M,N]^T (bf16) */
int dtb_mlp_transpose_bf16(int32_t M, int32_t N, const void* in, void* out, void* stream);
/* out[r][c] = in[r][c]*scale[c] + shift[c] for c < cin and 0 for cin <= c < cout
 * (normalise and zero-pad features); FP32 and/or bf16 output with row
 * stride ldo; scale/shift may be NULL (1 and 0). */
int dtb_mlp_affine(int32_t rows, int32_t cin, int32_t cout, const float* in, int64_t ldi, const float* scale,
                   const float* shift, float* out, void* outb, int64_t ldo, void* stream);
/* GRU cell, gate order (r, z, n) as torch.nn.GRU.  gi = x W_ih^T + b_ih and
 * gh = h W_hh^T + b_hh are [B][3H] FP32 (from dtb_mlp_gemm_bf16_nt);
 * forward: hn = (1-z) n + z h with r = sigmoid(gi_r+gh_r),
 * z = sigmoid(gi_z+gh_z), n = tanh(gi_n + r gh_n); FP32 and optional bf16. */
int dtb_gru_forward(int32_t B, int32_t H, const float* gi, const float* gh, const float* h, float* hn, void* hnb,
                    void* stream);
/* backward of one step, gates recomputed: with dh = dh_a + dh_b + dh_c
 * (dh_b, dh_c may be NULL) writes the gradients of gi and gh (FP32, optional
 * bf16) and dh_z = dh * z; the full gradient of h is dh_z + dgh W_hh. */
int dtb_gru_backward(int32_t B, int32_t H, const float* gi, const float* gh, const float* h, const float* dh_a,
                     const float* dh_b, const float* dh_c, float* dgi, void* dgib, float* dgh, void* dghb, float* dh_z,
                     void* stream);
/* round-to-nearest-even FP32 -> bf16 */
int dtb_mlp_f32_to_bf16(int64_t n, const float* in, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DTB_H_ */

/* tissuemc_hip.h — C ABI of the MI355X photon-transport engine.
 *
 * Plain C types only (no C++ or torch types), so that any FFI can bind it (ctypes, cgo, JNI,
 * N-API). The library owns all device memory and streams (a run's stream and timing events come
 * from a per-device pool and go back to it, so later calls skip their creation; the pool lives
 * until the process ends); the caller owns every host buffer it passes in, and no pointer is
 * retained after a call returns. Calls are synchronous and
 * blocking; they are safe to make one after another from one thread. Errors never cross the ABI
 * as exceptions: every entry point returns a TMC_* status and tmc_last_error() gives the message
 * (per calling thread).
 *
 * Reference interface replaced (Raghava-Dwivedi/tissuemc):
 *   tmc_hip_available    <- bool cuda_available()                      include/tissuemc/cuda_engine.hpp:9
 *   tmc_hip_run          <- RunResult run_cuda(const RunSpec&,
 *                                               const EngineOptions&)  include/tissuemc/cuda_engine.hpp:14
 *   tmc_spec             <- struct RunSpec / LayerSpec                 include/tissuemc/config.hpp:14-29
 *   tmc_opts             <- struct EngineOptions                       include/tissuemc/cpu_engine.hpp:18-30
 *   tmc_out              <- struct RunResult / BatchResult             include/tissuemc/result.hpp:15-46
 *   tmc_parse_mci_file   <- parse_mci_file(const std::string&)         include/tissuemc/mci.hpp:21
 *   tmc_validate_spec    <- validate(const RunSpec&)                   include/tissuemc/config.hpp:38
 *   TMC_ERR_CONFIG       <- ConfigError                                include/tissuemc/config.hpp:31
 *   TMC_ERR_DEVICE/RUNTIME <- std::runtime_error (no device, API error) src/cuda/engine.cu:18-26,351
 *
 * Tallies are integers in units of 2^-32 of one launched photon's weight (tmc_fixed_point_unit).
 * Integer sums are exact and associative, so a given (input, seed, photons, batches) gives
 * bit-identical tallies on every run, whatever the launch shape.
 */
#ifndef TISSUEMC_HIP_H
#define TISSUEMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMC_MAX_LAYERS 16

enum {
    TMC_OK = 0,
    TMC_ERR_CONFIG = 1,  /* invalid input or options (the reference's ConfigError) */
    TMC_ERR_DEVICE = 2,  /* no usable GPU */
    TMC_ERR_RUNTIME = 3  /* a HIP API call failed */
};

/* Option flags (tmc_opts.flags). */
#define TMC_FLAG_ARZ 1u           /* also tally A(r, z) into tmc_out.arz */
#define TMC_FLAG_NO_PROFILES 2u   /* diagnostic: skip the A_z/R_r/T_r/A_rz profiles (zeros) */
#define TMC_FLAG_NO_GRAPH 4u      /* batched runs: launch directly instead of replaying a graph */
#define TMC_FLAG_ANGLES 8u        /* also tally the exit-angle profiles into tmc_out.ra/ta/rra/tra */
#define TMC_FLAG_AGGREGATE 16u    /* combine a wave's equal-bin A_z and A(r,z) adds (ballot + DPP
                                   * sum) before the LDS atomics; results are identical */
#define TMC_FLAG_TIME 32u         /* also tally R_d(t), T_t(t) into tmc_out.rt/tt_t (tmc_opts.nt,
                                   * dt) */

typedef struct tmc_layer {
    double n;   /* refractive index */
    double mua; /* absorption coefficient [1/cm] */
    double mus; /* scattering coefficient [1/cm] */
    double g;   /* anisotropy */
    double d;   /* thickness [cm] */
} tmc_layer;

/* One MCML run, as read from an .mci file. */
typedef struct tmc_spec {
    uint64_t photons;
    double dz, dr;          /* grid bin widths [cm] */
    int32_t nz, nr, na;     /* grid bin counts (na: exit-angle bins, used by TMC_FLAG_ANGLES) */
    double n_above, n_below;
    int32_t num_layers;
    tmc_layer layers[TMC_MAX_LAYERS];
    char output_name[128];
} tmc_spec;

typedef struct tmc_opts {
    uint64_t seed;           /* 64-bit Philox key */
    int32_t batches;         /* >= 1 and <= photons */
    int32_t device;          /* HIP device ordinal */
    uint64_t photons;        /* nonzero overrides spec->photons */
    uint32_t stream;         /* Philox counter word 3; 0 gives the reference's streams */
    uint32_t flags;          /* TMC_FLAG_* */
    int32_t waves_per_group; /* 0 = automatic */
    int32_t groups;          /* persistent workgroups; 0 = one per compute unit */
    int32_t chunk;           /* photon ids claimed per wave at a time; 0 = automatic (about
                                1/64 of a wave's share of the photons, 64 to 1024) */
    int32_t arz_tile_z;      /* depth bins of A(r,z) kept in LDS; -1 = automatic */
    int32_t arz_tile_r;      /* radial bins of A(r,z) kept in LDS; -1 = automatic */
    uint64_t step_limit;     /* steps after which a photon is ended as an anomaly; 0 = 1e9 */
    /* With TMC_FLAG_TIME (no reference counterpart): time-of-flight bins of the escaping weight,
     * nt bins of dt ns from launch (the last bin takes later photons); time = optical path / c,
     * the optical path summing n of the layer times each hop's length. */
    double dt;               /* bin width [ns] */
    int32_t nt;              /* bin count */
} tmc_opts;

/* Results. Array pointers are caller-owned host buffers sized by B = batches, L = num_layers,
 * nz and nr; all values are in units of tmc_fixed_point_unit() except the counters. */
typedef struct tmc_out {
    uint64_t *rd;             /* [B]    diffuse reflectance weight per batch */
    uint64_t *tt;             /* [B]    transmitted weight per batch */
    uint64_t *absorbed;       /* [B*L]  absorbed weight per batch and layer */
    uint64_t *events;         /* [B]    absorption events */
    uint64_t *anomalies;      /* [B]    non-finite weights, non-unit directions, step limit */
    uint64_t *boundary_steps; /* [B]    steps that ended on a layer boundary */
    uint64_t *az;             /* [B*nz] absorbed weight per batch and depth bin */
    uint64_t *rr;             /* [B*nr] diffuse reflectance per batch and radial bin */
    uint64_t *tr;             /* [B*nr] transmittance per batch and radial bin */
    uint64_t *arz;            /* [nz*nr] absorbed weight per (z, r) bin, all batches; or NULL */
    double specular;          /* specular reflectance of the launched beam */
    double kernel_seconds;    /* transport kernel time (HIP events) */
    double wall_seconds;      /* whole call */
    char device[96];          /* device name and architecture */
    int32_t groups, waves_per_group, lds_bytes, chunk, arz_tile_z, arz_tile_r;
    /* With TMC_FLAG_ANGLES (no reference counterpart; MCML's Rd_a, Tt_a, Rd_ra, Tt_ra): weight
     * escaping at an angle from the normal in na bins of pi / (2 na), all batches. */
    uint64_t *ra, *ta;        /* [na]    diffuse reflectance, transmittance per angle bin */
    uint64_t *rra, *tra;      /* [nr*na] per (radius, angle) bin, row-major in radius */
    /* With TMC_FLAG_TIME: escaping weight per time-of-flight bin, all batches. */
    uint64_t *rt, *tt_t;      /* [nt]    diffuse reflectance, transmittance per time bin */
} tmc_out;

/* ---------------------------------------------------------------- batched multi-stack runs
 *
 * No reference counterpart (the reference simulates one stack per process). Many independent
 * layer stacks, each with its own photon budget, run in one persistent launch; every stack is cut
 * into batches_per_stack batches (the reference's BatchPlan over its budget) for standard errors.
 * Stack s uses the Philox streams (seed, photon id, word3 = s): its results equal, bit for bit,
 * a single run of the same stack with tmc_opts.stream = s and the same seed, batches and grid. */

typedef struct tmc_stack {
    int32_t num_layers;
    double n_above, n_below;
    tmc_layer layers[TMC_MAX_LAYERS];
    uint64_t photons; /* this stack's photon budget */
} tmc_stack;

typedef struct tmc_batched_opts {
    uint64_t seed;
    int32_t batches_per_stack; /* >= 1 and <= every stack's budget */
    int32_t device;
    double dz, dr;             /* one grid for every stack */
    int32_t nz, nr;
    uint32_t flags;            /* TMC_FLAG_NO_PROFILES, TMC_FLAG_NO_GRAPH, TMC_FLAG_AGGREGATE */
    int32_t waves_per_group, groups, chunk; /* 0 = automatic */
    uint64_t step_limit;       /* 0 = 1e9 */
    uint64_t photons_hint;     /* expected photons per execute (sizes the claim chunk); 0 = auto */
} tmc_batched_opts;

/* Results of a batched run; S = stacks, Bs = batches_per_stack. Caller-owned buffers. */
typedef struct tmc_batched_out {
    uint64_t *rd, *tt;        /* [S*Bs]   row s*Bs + b */
    uint64_t *absorbed;       /* [S*Bs*TMC_MAX_LAYERS] */
    uint64_t *events, *anomalies, *boundary_steps; /* [S*Bs] */
    uint64_t *az;             /* [S*nz] all batches of a stack */
    uint64_t *rr, *tr;        /* [S*nr] */
    double kernel_seconds;    /* clear + transport (HIP events) */
    double wall_seconds;
    char device[96];
    int32_t groups, waves_per_group, lds_bytes, chunk;
} tmc_batched_out;

typedef struct tmc_batched_plan tmc_batched_plan; /* opaque */

/* A resident plan for repeated batched runs of n_stacks stacks: device buffers are allocated
 * once and the clear + transport launches are captured in a HIP graph on the first execute and
 * replayed afterwards. */
int tmc_batched_create(int32_t n_stacks, const tmc_batched_opts *opts, tmc_batched_plan **plan);
int tmc_batched_execute(tmc_batched_plan *plan, const tmc_stack *stacks, int32_t n_stacks,
                        uint64_t seed, tmc_batched_out *out);
void tmc_batched_destroy(tmc_batched_plan *plan);

/* One-shot form: create, execute once, destroy. */
int tmc_hip_run_batched(const tmc_stack *stacks, int32_t n_stacks, const tmc_batched_opts *opts,
                        tmc_batched_out *out);

/* ---------------------------------------------------------------- launch-shape autotuning
 *
 * No reference counterpart (the reference's CUDA engine launches one fixed shape,
 * src/cuda/engine.cu). Tallies do not depend on the launch shape, so the shape is picked by
 * timing pilot runs of the caller's own workload (every budget scaled by pilot_fraction, at
 * least one photon per batch) at each candidate: waves per workgroup and workgroups per compute
 * unit, which together decide how the LDS of a compute unit is split between the waves' tally
 * slots and, for single runs, the workgroup's A(r,z) tile. Each candidate is timed three times;
 * its fastest launch counts. */

#define TMC_TUNE_MAX 16

typedef struct tmc_tune_entry {
    int32_t waves_per_group, groups, lds_bytes, arz_tile_z, arz_tile_r;
    double kernel_seconds;  /* fastest pilot launch */
    double confirm_seconds; /* faster of two launches at confirm_fraction; 0 if not re-timed */
} tmc_tune_entry;

typedef struct tmc_tune_report {
    double pilot_fraction;
    double confirm_fraction; /* 0 if no shape came close enough to the default to re-time */
    int32_t n_tried, best;   /* tried[best] is the shape chosen */
    tmc_tune_entry tried[TMC_TUNE_MAX];
} tmc_tune_report;

/* Single runs: copies *opts into *tuned with waves_per_group, groups and (with TMC_FLAG_ARZ) the
 * A(r,z) tile set to the fastest candidate; the launch-shape fields of *opts are not used.
 * Every candidate is first timed on a pilot at pilot_fraction (fastest of three launches; the
 * default shape is listed first). Short pilots are dominated by the last, longest photon walks,
 * so their ranking is then confirmed: the default and up to two shapes whose pilot lies within
 * 10% of the fastest are re-timed at confirm_fraction = min(1, 8 * pilot_fraction), and a shape
 * replaces the default only if it is more than 3% faster there. pilot_fraction is in (0, 1].
 * report may be NULL. */
int tmc_autotune(const tmc_spec *spec, const tmc_opts *opts, double pilot_fraction,
                 tmc_opts *tuned, tmc_tune_report *report);

/* Batched plans: times pilots of these stacks (with the same confirmation step) and rebuilds the
 * plan's device state (and its graph) with the fastest shape, which later executes use. report
 * may be NULL. */
int tmc_batched_autotune(tmc_batched_plan *plan, const tmc_stack *stacks, int32_t n_stacks,
                         uint64_t seed, double pilot_fraction, tmc_tune_report *report);

/* Distribution of random layer stacks. Log-uniform: mua, mus, d, photons; uniform: g, n; the
 * layer count is uniform in [min_layers, max_layers]; the ambient index is fixed. */
typedef struct tmc_stack_dist {
    int32_t min_layers, max_layers;
    double mua_lo, mua_hi;         /* [1/cm] */
    double mus_lo, mus_hi;         /* [1/cm] */
    double g_lo, g_hi;
    double n_lo, n_hi;
    double d_lo, d_hi;             /* [cm] */
    double photons_lo, photons_hi; /* photon budget per stack */
    double n_ambient;
} tmc_stack_dist;

/* The distribution of the batched workload: 1-5 layers, mua 0.01-10 /cm, mus 10-500 /cm,
 * g 0.6-0.95, n 1.33-1.5, d 0.005-0.5 cm, budgets 1e4-1e6 photons, ambient 1.0. */
void tmc_default_stack_dist(tmc_stack_dist *dist);

/* Stacks first_id .. first_id + count - 1 of the sequence keyed by seed. Stack k depends only on
 * (seed, k): field blocks of Philox4x32-10 with counter {k_lo, k_hi, block, 0x53544B31} and key
 * seed, each word mapped to (x + 0.5) * 2^-32. Block 0 gives the layer count (word 0) and budget
 * (word 1); block 1 + i gives layer i's mua, mus, g, n (words 0-3); block 17 + i its thickness. */
int tmc_generate_stacks(uint64_t seed, uint64_t first_id, int32_t count, const tmc_stack_dist *dist,
                        tmc_stack *out);

/* ---------------------------------------------------------------- learned surrogate
 *
 * No reference counterpart. An ensemble of `heads` independent MLPs (inputs -> depth hidden
 * layers of width `hidden` with GELU (tanh form) -> outputs), trained with Adam on a mean squared
 * error, every matrix product in BF16 MFMA with FP32 accumulation. Parameters per head, in order
 * per layer k: W_k [out_k x in_k] row-major, then b_k [out_k] (float32). */

typedef struct tmc_mlp_config {
    int32_t heads, inputs, hidden, depth, outputs;
    int32_t batch;                  /* training batch per step and head (multiple of 64) */
    float lr, beta1, beta2, eps, weight_decay;
    uint64_t seed;                  /* initialisation and batch sampling */
    int32_t device;
} tmc_mlp_config;

typedef struct tmc_mlp tmc_mlp; /* opaque */

/* 4 heads, 30 inputs, 4 x 512 hidden, 202 outputs, batch 4096, Adam lr 1e-3 (0.9, 0.999, 1e-8). */
void tmc_mlp_default_config(tmc_mlp_config *config);
int tmc_mlp_create(const tmc_mlp_config *config, tmc_mlp **model);
void tmc_mlp_destroy(tmc_mlp *model);
/* Parameters per head (logical, unpadded). */
int64_t tmc_mlp_param_count(const tmc_mlp *model);
/* Upload the training set: x [n x inputs], y [n x outputs] (float32), loss weight per output
 * (NULL = all 1). The set stays in device memory. */
int tmc_mlp_set_data(tmc_mlp *model, const float *x, const float *y, int64_t n,
                     const float *loss_weights);
/* `steps` Adam steps on batches sampled from the set (each head its own samples); the step is a
 * captured HIP graph replayed per step. losses: [steps x heads] or NULL; seconds: device time. */
int tmc_mlp_train(tmc_mlp *model, int64_t steps, float *losses, double *seconds);
/* Gradients of the loss on the given rows of the set (batch rows, no update): grads
 * [heads x param_count], loss [heads]. */
int tmc_mlp_gradients(tmc_mlp *model, const int64_t *rows, float *grads, float *loss);
/* Ensemble prediction for x [n x inputs]: mean [n x outputs] over heads, spread [n x 2] (standard
 * deviation over heads of outputs 0 and 1), per_head [heads x n x outputs] (any may be NULL).
 * seconds: device time of the forward passes and reductions (transfers excluded). */
int tmc_mlp_predict(tmc_mlp *model, const float *x, int64_t n, float *mean, float *spread,
                    float *per_head, double *seconds);
int tmc_mlp_get_params(tmc_mlp *model, float *params /* [heads x param_count] */);
int tmc_mlp_set_params(tmc_mlp *model, const float *params);
/* Learning rate of the following steps (kept in device memory; the captured step reads it). */
int tmc_mlp_set_lr(tmc_mlp *model, float lr);
/* Algorithmic matrix-product FLOPs of one training step and of one query (all heads). */
int tmc_mlp_flops(const tmc_mlp *model, uint64_t *per_step, uint64_t *per_query);

/* Surrogate answers for layer stacks, end to end on the device (no reference counterpart).
 * The network must have 30 inputs (5 layer slots x (log10 mua, log10 mus, g, n, log10 d), each
 * mapped to [-1, 1] over gate->box, 0 in empty slots, then 5 slot masks) and >= 2 outputs, the
 * first two being R_d and T_t; output 0 is log10(R_d + floor_rd), output 1 log10(T_t + floor_tt)
 * and every other output log10(value + floor), each standardized by y_mean and y_std. For each stack the call encodes the inputs on the device (in double precision, the
 * host's arithmetic), checks the domain (layer count within the box's range, every present
 * layer's five encoded parameters within the box, to 1e-9), runs the ensemble, and routes the
 * stack to Monte Carlo (route = 1) if it is outside the domain, if the ensemble's spread of R_d
 * relative to max(R_d, 1e-3) exceeds gate->tau_rd, or if that of T_t relative to max(T_t, 1e-2)
 * exceeds gate->tau_tt. Stacks go through in chunks (128 x the CU count), the host packing the
 * next chunk (on up to 8 threads of its own, started and joined inside the call, when n is above
 * half a chunk) while that chunk's copy and the current chunk's kernels run; only the answers
 * come back. */
typedef struct tmc_surrogate_gate {
    tmc_stack_dist box;  /* the training distribution: encoding box and domain */
    const float *y_mean; /* [outputs] */
    const float *y_std;  /* [outputs] */
    double floor;        /* outputs 2.. are log10(value + floor): 1e-7 in the pipeline */
    double floor_rd;     /* output 0 is log10(R_d + floor_rd): 3e-4 in the pipeline */
    double floor_tt;     /* output 1 is log10(T_t + floor_tt): 3e-3 in the pipeline */
    double tau_rd;       /* spread gate of R_d */
    double tau_tt;       /* spread gate of T_t */
} tmc_surrogate_gate;

typedef struct tmc_surrogate_answers {
    double *rd, *tt;               /* [n] ensemble R_d, T_t (>= 0) */
    double *rd_spread, *tt_spread; /* [n] or NULL: ensemble standard deviation of R_d, T_t */
    uint8_t *route;                /* [n] 1: answer with Monte Carlo */
    uint8_t *in_domain;            /* [n] or NULL */
    float *profiles;               /* [n x (outputs - 2)] or NULL: the other outputs' values */
    double device_seconds;         /* the kernels (encode, network, gate), HIP events */
    double wall_seconds;           /* the whole call: packing, transfers and kernels */
} tmc_surrogate_answers;

int tmc_mlp_answer(tmc_mlp *model, const tmc_surrogate_gate *gate, const tmc_stack *stacks,
                   int64_t n, tmc_surrogate_answers *out);

/* 1 if a HIP device is present and usable, else 0. */
int tmc_hip_available(void);

/* Simulate one run on the GPU (blocking). Batch b covers photon ids
 * [b*base + min(b, extra), ...) exactly as the reference's BatchPlan, and photon i uses the
 * Philox stream keyed by (seed, i), so per-batch results are comparable photon for photon with
 * the reference's CPU engine. */
int tmc_hip_run(const tmc_spec *spec, const tmc_opts *opts, tmc_out *out);

/* tmc_hip_run for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. Same
 * arguments, results and errors as tmc_hip_run, bit for bit, but it calls no hipMalloc or
 * hipFree and creates no stream or event: every device buffer of the run (unit and layer tables,
 * seed, tallies) comes from one static per-device scratch array of tmc_static_scratch_bytes()
 * (16 MiB; every GPU the library runs on holds it until the process ends), and the work goes to
 * the null stream. A run needing more scratch (about 8 bytes per tally bin and per-batch
 * counter) fails with TMC_ERR_CONFIG before any device work. kernel_seconds is measured on the
 * host from the transport launch to its completion. All calls on one GPU share its scratch
 * array: calls that overlap on one GPU (from several host threads) race and must not be made. */
int tmc_hip_run_static(const tmc_spec *spec, const tmc_opts *opts, tmc_out *out);

/* Bytes of the static per-device scratch array used by tmc_hip_run_static. */
uint64_t tmc_static_scratch_bytes(void);

/* Message for the last failed call on this thread ("" if none). */
const char *tmc_last_error(void);

/* 2^-32: the weight of one tally unit. */
double tmc_fixed_point_unit(void);

/* Fill opts with the defaults (seed 1, 10 batches, device 0, automatic launch shape). */
void tmc_default_opts(tmc_opts *opts);

/* Check a run description like the reference's validate(); TMC_ERR_CONFIG names the field. */
int tmc_validate_spec(const tmc_spec *spec);

/* Parse an MCML .mci file into up to max_runs runs; *n_runs receives the count. Errors name the
 * file and line (TMC_ERR_CONFIG). */
int tmc_parse_mci_file(const char *path, tmc_spec *runs, int32_t max_runs, int32_t *n_runs);

/* The same from the text of an .mci file (<- parse_mci(std::istream&), include/tissuemc/mci.hpp:18). */
int tmc_parse_mci_text(const char *text, tmc_spec *runs, int32_t max_runs, int32_t *n_runs);

/* The first n 32-bit words of the photon stream (seed, photon_id, word3), generated by the
 * device's Philox4x32-10 code path. A parity hook: the stream must equal the reference's
 * PhotonRng word for word (include/tissuemc/philox.hpp:62-83). */
int tmc_hip_philox_stream(uint64_t seed, uint64_t photon_id, uint32_t word3, int64_t n,
                          uint32_t *out, int32_t device);

#ifdef __cplusplus
}
#endif

#endif /* TISSUEMC_HIP_H */

/*
 * cpm_engine.h — C ABI of the MI355X copy-attempt engine.
 *
 * This is the drop-in boundary for cpm-core's hot path.  Each entry point
 * names the reference interface it replaces (paths relative to the
 * reference repository).  Plain pointers and sizes only; every call blocks
 * until its device work is complete; every entry point returns a
 * cpm_status and never aborts: CPM_ERR_CONFIG and CPM_ERR_INIT carry the
 * same messages the reference's ConfigError / InitError / BatchError do
 * (crates/cpm-py/src/errors.rs:16-33 maps all of them to ValueError), and
 * cpm_last_error() returns the message of the calling thread's last
 * failure.
 *
 * Threading: an engine is not re-entrant (one caller at a time); separate
 * engines, models and results are independent.
 */
#ifndef CPM_ENGINE_H
#define CPM_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CPM_OK = 0,
    CPM_ERR_CONFIG = 1, /* ConfigError / CellTypeError / theta errors   */
    CPM_ERR_INIT = 2,   /* InitError                                     */
    CPM_ERR_DEVICE = 3, /* HIP runtime failure or no device              */
    CPM_ERR_OOM = 4,    /* device or host allocation failure             */
    CPM_ERR_ARG = 5     /* invalid argument to this ABI (null pointer...) */
} cpm_status;

/* Boundary (crates/cpm-core/src/lattice.rs:11-21) */
enum { CPM_BOUNDARY_FIXED = 0, CPM_BOUNDARY_PERIODIC = 1 };
/* NeighborhoodKind (crates/cpm-core/src/neighborhood.rs:147-167) */
enum {
    CPM_NB_DEFAULT = -1, /* resolve() default for energy (Moore / Eighteen) */
    CPM_NB_VON_NEUMANN = 0,
    CPM_NB_MOORE = 1,
    CPM_NB_EIGHTEEN = 2,
    CPM_NB_TWENTY_SIX = 3,
    CPM_NB_TWENTY_SIX_WEIGHTED = 4
};
/* AcceptanceMode / ProposalMode (crates/cpm-core/src/config.rs:42-66) */
enum { CPM_ACCEPT_METROPOLIS = 0, CPM_ACCEPT_METROPOLIS_HASTINGS = 1 };
enum { CPM_PROPOSAL_UNIFORM = 0, CPM_PROPOSAL_EDGE_LIST = 1 };
/* Device execution modes of cpm_run_mcs */
enum {
    /* reproduces the reference's serial attempt order bit-for-bit */
    CPM_MODE_EXACT_REPLAY = 0,
    /* halo-disjoint checkerboard sublattices updated in parallel; same
       energy function, different update order (statistically validated) */
    CPM_MODE_SUBLATTICE = 1
};

/*
 * The user-facing configuration: UserConfig<D> plus the builder's
 * active_terms (crates/cpm-core/src/config.rs:141-257;
 * crates/cpm-py/src/builder.rs:117-195).  cpm_model_create resolves the
 * documented defaults and validates exactly like ResolvedConfig::validate
 * (config.rs:287-413).  The adhesion matrix is passed row by row so that a
 * ragged matrix can be reported with the reference's message.
 */
typedef struct {
    uint32_t dim;                  /* 2 or 3 (grid length)                   */
    uint64_t grid[3];
    int32_t boundary;              /* CPM_BOUNDARY_*                         */
    int32_t copy_neighborhood;     /* CPM_NB_*                               */
    int32_t energy_neighborhood;   /* CPM_NB_* or CPM_NB_DEFAULT             */
    int32_t connectivity_neighborhood;
    int32_t acceptance;            /* CPM_ACCEPT_*                           */
    int32_t proposal;              /* CPM_PROPOSAL_*                         */
    int32_t act_active;            /* active_terms.act                       */
    uint32_t n_types;
    const char* const* type_names;
    const uint32_t* target_volume;
    const uint32_t* target_interface;
    const double* lambda_volume;
    const double* lambda_interface;
    const double* lambda_act;
    const uint32_t* max_act;
    const uint64_t* cell_counts;   /* per type (add_cells)                   */
    uint32_t adhesion_rows;        /* 0 = adhesion not set                   */
    const uint32_t* adhesion_row_len;
    const double* adhesion;        /* rows concatenated                      */
    int32_t init_explicit;         /* InitializationSpec::Explicit           */
    const uint32_t* init_lattice;  /* row-major, unpadded                    */
    uint64_t init_lattice_len;
    const uint64_t* init_cell_type_of;
    uint64_t init_n_cells;
    int32_t warn_on_default_init;
    uint64_t burn_in_mcs;
    uint64_t readout_mcs;
    uint64_t sampling_interval_mcs;
    uint64_t seed;
} cpm_config;

/* Per-launch event counters: A attempts, N non-null attempts (sigma_t !=
   sigma_s), Nt attempts that read the simple-point table, E energy
   evaluations, C accepted copies; exp_near_ties counts Metropolis draws
   within 2^-40 (relative) of exp(-dH), decided with the correctly rounded
   exp instead of the device libm's. */
typedef struct {
    uint64_t attempts, non_null, topology, energy, accepted, exp_near_ties;
    double kernel_ms; /* device time of the launches, HIP events */
    /* the exact-replay LDS kernel alone: its launches and their summed
       durations, from HIP events recorded around each launch on the
       stream it runs on */
    uint64_t replay_launches;
    double replay_launch_ms;
    /* exact-replay replicas whose long chains, with the lattice left in
       HBM by the LDS budget, ran on the pipelined kernel (summed over
       launches) */
    uint64_t piped_stragglers;
} cpm_counters;

typedef struct cpm_engine cpm_engine;
typedef struct cpm_model cpm_model;
typedef struct cpm_result cpm_result;

const char* cpm_last_error(void);
const char* cpm_version(void);

/* One engine per device: owns the device buffers and the simple-point
   tables (2D: 256 bits; 3D: 2^26 bits generated on the device).
   Replaces the process-wide OnceLock tables of
   crates/cpm-core/src/connectivity.rs:20-24 and simple_point.rs:361-366. */
cpm_status cpm_engine_create(int device, cpm_engine** out);
void cpm_engine_destroy(cpm_engine* engine);

/* UserConfig::resolve + CPM::new (crates/cpm-core/src/config.rs:200-257,
   crates/cpm-core/src/model.rs:44-76): resolve defaults, validate, build an
   empty model seeded with rng_from_seed(seed). */
cpm_status cpm_model_create(const cpm_config* config, cpm_model** out);
void cpm_model_destroy(cpm_model* model);
/* initialization::initialize (crates/cpm-core/src/initialization.rs:39-55);
   *used_default = 1 iff scatter-and-grow ran. Host-side, as in the
   reference. */
cpm_status cpm_model_initialize(cpm_model* model, int32_t* used_default);
/* initialization::initialize for n models at once, as run_batch's workers
   do (crates/cpm-core/src/batch.rs:126-145, initialization.rs:39-55): the
   default scatter_and_grow (initialization.rs:202-330) runs on the engine's
   device, one wave per model of at most 4096 cells, every draw in the
   reference's order; explicit placements and larger models initialise on
   the host threads (all of them on the host threads when
   cpm_engine_set_device_init(engine, 0) was called).  Each model ends in
   the state cpm_model_initialize gives it.  On failure returns the error
   (and message) of the lowest failing index on the device path, of some
   failing index on the host path; used_default (n entries) may be NULL. */
cpm_status cpm_model_initialize_batch(cpm_engine* engine, cpm_model* const* models, size_t n,
                                      int32_t* used_default);

/* run_mcs (crates/cpm-core/src/monte_carlo.rs:327-357) for n models at
   once: every model advances by n_mcs MCS on the device.  In
   CPM_MODE_EXACT_REPLAY the result is bit-identical to calling the
   reference's run_mcs on each model in turn.  counters may be NULL. */
cpm_status cpm_run_mcs(cpm_engine* engine, cpm_model* const* models, size_t n,
                       uint64_t n_mcs, int32_t mode, cpm_counters* counters);
/* Same, with exactly one model per kernel launch and one launch per MCS:
   the "plain" one-replica-per-launch baseline the batched path is
   compared against. */
cpm_status cpm_run_mcs_plain(cpm_engine* engine, cpm_model* model, uint64_t n_mcs,
                             cpm_counters* counters);

/* No counterpart in the reference: cpm_run_mcs in CPM_MODE_EXACT_REPLAY
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  It
   allocates nothing on the device and creates no HIP events: every model
   must already be resident on this engine's device (advanced by an earlier
   cpm_run_mcs or cpm_run on it; 3D models also need the engine's 3D table,
   built by that call), and the call's replica descriptors and launch order
   live in a fixed device array of the library, at most
   CPM_NOALLOC_MAX_REPLICAS models per call (larger n returns CPM_ERR_ARG;
   the array takes about 2.5 MiB of each device the library runs on, until
   the process ends).  The results are bit-identical to cpm_run_mcs.  Every
   call on one device shares that array: calls that overlap on one device
   race and must not be made.  counters->replay_launch_ms stays 0 (no
   per-launch events). */
#define CPM_NOALLOC_MAX_REPLICAS 4096
cpm_status cpm_run_mcs_noalloc(cpm_engine* engine, cpm_model* const* models, size_t n,
                               uint64_t n_mcs, cpm_counters* counters);

/* State access (the fields of State<D>, crates/cpm-core/src/state.rs:168-203,
   plus the RNG state and mcs of CPM<D>, model.rs:15-38).  Sizes: lattice
   n_sites (row-major, unpadded), volume/interface/type_index n_cells,
   sum_position n_cells*dim, sum_outer n_cells*dim*dim, contact
   (n_cells+1)^2 upper-triangular.  Any pointer may be NULL. */
typedef struct {
    uint64_t n_sites, n_cells, mcs;
    double conservative_energy;
    uint64_t rng_state[4];
} cpm_model_info;
cpm_status cpm_model_get_info(cpm_model* model, cpm_model_info* out);
cpm_status cpm_model_get_state(cpm_model* model, uint32_t* lattice, uint32_t* volume,
                               uint32_t* interface, uint32_t* type_index,
                               int64_t* sum_position, int64_t* sum_outer,
                               uint32_t* contact);
/* convention_hash / model_hash (crates/cpm-core/src/config.rs:488-556) of
   the model's current state: 64 lowercase hex digits + NUL each. */
cpm_status cpm_model_hashes(cpm_model* model, char convention_hash[65], char model_hash[65]);
/* global_energy (crates/cpm-core/src/energy.rs:561-565) from scratch */
cpm_status cpm_model_global_energy(cpm_model* model, double* out);

/*
 * State import: replaces the whole simulation state of a created model by
 * the caller's, the fields of CPM<D> (rng, mcs; crates/cpm-core/src/
 * model.rs:15-38) and State<D> (lattice, cells, volume, interface,
 * position sums, contact, conservative_energy; state.rs:168-203).  This is
 * what a cpm-core shim needs to hand an existing `&mut CPM<D>` to the
 * device inside `run_mcs` (monte_carlo.rs:327) and continue its trajectory
 * exactly: after the import the model is the caller's model, at the
 * caller's RNG position and MCS count.
 *
 * lattice (n_sites, row-major, unpadded) and type_index (n_cells) are
 * required; any other per-cell pointer may be NULL, and is then derived
 * from the lattice as initialization does (volume by counting;
 * interface, sum_position, sum_outer, contact by
 * recompute_interface_and_positions, initialization.rs:396-436).
 * sum_position / sum_outer are the unwrapped sums the reference carries
 * (a recompute from a periodic lattice gives the wrapped coordinates).
 * has_energy = 0 recomputes conservative_energy with global_energy
 * (energy.rs:561-565).  rng_state NULL keeps the model's current stream.
 * The values are taken as given — no connectivity or consistency check,
 * as in the reference, whose fields are public — except what device
 * memory safety needs: every lattice id <= n_cells, every type_index <
 * n_types, 1 <= n_cells < 2^32 - 1, and a non-zero RNG state (Xoshiro's
 * fixed point); violations return CPM_ERR_ARG and leave the model as it
 * was.  Any device copy of the previous state is dropped.
 */
typedef struct {
    const uint32_t* lattice;      /* n_sites, row-major, unpadded     */
    uint64_t n_cells;             /* cell ids are 1..n_cells          */
    const uint32_t* type_index;   /* n_cells                          */
    const uint32_t* volume;       /* n_cells, or NULL                 */
    const uint32_t* interface;    /* n_cells, or NULL                 */
    const int64_t* sum_position;  /* n_cells * dim, or NULL           */
    const int64_t* sum_outer;     /* n_cells * dim * dim, or NULL     */
    const uint32_t* contact;      /* (n_cells + 1)^2, or NULL         */
    int32_t has_energy;
    double conservative_energy;
    const uint64_t* rng_state;    /* Xoshiro256++ s[0..4], or NULL    */
    uint64_t mcs;
    const uint32_t* birth_mcs;    /* n_sites, row-major: the Act term's
                                     per-site birth MCS (energy.rs:
                                     337-344), or NULL (all 0); ignored
                                     when the term is off             */
    const uint64_t* edge_sites;   /* EdgeList proposal: the edge list in
                                     its order (edge_list.rs:42-45), as
                                     interior row-major site indices,
                                     each at most once; NULL rebuilds
                                     it from the lattice (EdgeList::
                                     build); ignored under Uniform    */
    uint64_t n_edge_sites;
} cpm_model_state;
cpm_status cpm_model_set_state(cpm_model* model, const cpm_model_state* state);

/* The Act term's per-site birth MCS (ActTerm::birth_mcs, energy.rs:
   337-344; written by ActTerm::commit, energy.rs:459-464) for the interior
   sites, row-major, into out[n_sites]; all 0 when the term is off. */
cpm_status cpm_model_birth_mcs(cpm_model* model, uint32_t* out, size_t n_sites);

/* The EdgeList proposal's edge list (edge_list.rs:42-45) in its order, as
   interior row-major site indices: writes up to cap entries and sets *n to
   its length (0 under the uniform proposal). */
cpm_status cpm_model_edge_sites(cpm_model* model, uint64_t* out, size_t cap, size_t* n);

/* take_sample (crates/cpm-core/src/output.rs:289-337, with
   label_components, labelling.rs:70-154) of the model's current state, as
   a result holding that one sample (include_lattice as in cpm_run),
   computed on the device when the model is resident there (it has run)
   and on the host otherwise, with the same results.  Its
   status applies run_with's rules (output.rs:470-509) to this one sample
   with an empty burn-in trace: Fragmented if a cell is in pieces, CellLost
   if a cell has volume <= 1, Degenerate if the contact graph joins every
   cell, else Ok. */
cpm_status cpm_model_sample(cpm_model* model, int32_t include_lattice, cpm_result** out);

/* not_equilibrated (crates/cpm-core/src/output.rs:388-420, private in the
   reference, where run_with applies it to the burn-in energy trace):
   *out = 1 if the least-squares drift over the trace exceeds twice the
   residual scatter. */
cpm_status cpm_not_equilibrated(const double* trace, size_t n, int32_t* out);

/* output::run_with (crates/cpm-core/src/output.rs:434-529): burn-in,
   equilibration check, sampled readout window, statuses, metadata.  The
   MC steps run on the device in `mode`. */
cpm_status cpm_run(cpm_engine* engine, cpm_model* model, int32_t include_lattice,
                   int32_t mode, cpm_result** out);

/* batch::run_batch (crates/cpm-core/src/batch.rs:110-153): one replica per
   theta (theta::inject, theta.rs:93-147) with seed derive_seed_u64(master,
   index), all replicas resident on the device and advanced together.
   out must hold n_thetas result pointers. */
cpm_status cpm_run_batch(cpm_engine* engine, const cpm_config* base, const double* thetas,
                         size_t n_thetas, size_t theta_len, uint64_t master_seed,
                         int32_t include_lattice, int32_t mode, cpm_result** out,
                         int32_t* any_used_default);
/* The same batched driver over fully independent configurations (uneven
   grids, per-replica V*, I* and seeds): per-replica CPM::new + initialize +
   run_with (batch.rs:130-145) without the shared-grid restriction. */
cpm_status cpm_run_batch_configs(cpm_engine* engine, const cpm_config* const* configs,
                                 size_t n, int32_t include_lattice, int32_t mode,
                                 cpm_result** out, int32_t* any_used_default,
                                 cpm_counters* counters);

/* RunResult / Sample views (crates/cpm-core/src/output.rs:245-287;
   crates/cpm-py/src/convert.rs:27-115).  Pointers stay valid until
   cpm_result_destroy. */
typedef struct {
    const char* cpm_version;
    uint32_t convention_version;
    const char* convention_hash;
    const char* model_hash;
    uint64_t seed;
    const char* seed_policy;
    double derived_phi;
    uint32_t theta_layout_version;
    const char* status_kind; /* "Ok", "CellLost", "Fragmented", ... */
    int32_t status_has_cell;
    uint32_t status_cell_id;
    uint64_t status_at_mcs;
    uint64_t n_samples;
    uint64_t n_cells;
    uint32_t dim;
    const uint32_t* cell_type_index;
    int32_t used_default;
} cpm_result_info;
typedef struct {
    uint64_t mcs;
    double conservative_energy;
    uint64_t n_cells;
    const uint32_t* ids;
    const uint32_t* volume;
    const uint32_t* interface;
    const double* centroid;     /* n_cells x dim */
    const double* eigenvalues;  /* n_cells x dim */
    const double* radius_of_gyration_sq;
    const double* kappa_sq;
    const uint32_t* contact;    /* (n_cells+1)^2 */
    uint64_t n_fragmented;
    const uint32_t* fragmented_cells;
    uint64_t n_medium_components;
    const uint32_t* medium_component_volumes;
    const uint8_t* medium_component_touches_boundary;
    const uint32_t* lattice;    /* NULL unless include_lattice */
} cpm_sample_view;
/* Every sample of a result in one call, field by field (the same data as
   cpm_result_get_sample, for bindings that convert many samples at once):
   cpm_result_get_samples_shape gives the sizes, the caller allocates one
   buffer per field and cpm_result_copy_samples fills them.  Sample k's
   values of a per-cell field start at k * n_cells (k * n_cells * dim for
   centroid and eigenvalues, k * (n_cells+1)^2 for contact, k *
   lattice_sites for lattice); the variable-length lists of sample k are
   [offsets[k], offsets[k+1]) (offsets hold n_samples + 1 entries).  Sample
   ids are 1..n_cells.  No counterpart in the reference (its PyO3 layer
   converts in Rust, crates/cpm-py/src/convert.rs:53-115). */
typedef struct {
    uint64_t n_samples;
    uint64_t n_cells;
    uint32_t dim;
    int32_t has_lattice;
    uint64_t lattice_sites;
    uint64_t n_fragmented_total;
    uint64_t n_medium_total;
} cpm_samples_shape;
typedef struct {
    uint64_t* mcs;                          /* n_samples */
    double* conservative_energy;            /* n_samples */
    uint32_t* volume;                       /* n_samples x n_cells */
    uint32_t* interface;
    double* centroid;                       /* n_samples x n_cells x dim */
    double* eigenvalues;
    double* radius_of_gyration_sq;          /* n_samples x n_cells */
    double* kappa_sq;
    uint32_t* contact;                      /* n_samples x (n_cells+1)^2 */
    uint64_t* fragmented_offsets;           /* n_samples + 1 */
    uint32_t* fragmented_cells;             /* n_fragmented_total */
    uint64_t* medium_offsets;               /* n_samples + 1 */
    uint32_t* medium_component_volumes;     /* n_medium_total */
    uint8_t* medium_component_touches_boundary;
    uint32_t* lattice;                      /* n_samples x lattice_sites, or NULL */
} cpm_samples_buffers;
cpm_status cpm_result_get_info(const cpm_result* result, cpm_result_info* out);
cpm_status cpm_result_get_sample(const cpm_result* result, uint64_t index,
                                 cpm_sample_view* out);
cpm_status cpm_result_get_samples_shape(const cpm_result* result, cpm_samples_shape* out);
cpm_status cpm_result_copy_samples(const cpm_result* result, const cpm_samples_buffers* out);
void cpm_result_destroy(cpm_result* result);

/* theta (crates/cpm-core/src/theta.rs:43-178) */
size_t cpm_theta_len(uint32_t n_types);
cpm_status cpm_theta_extract(const cpm_config* config, double* out, size_t cap);
/* writes the NUL-terminated name of entry `index` into buf */
cpm_status cpm_theta_name(const cpm_config* config, size_t index, char* buf, size_t cap);

/* derive_seed_u64 (crates/cpm-core/src/rng.rs:37-40) */
uint64_t cpm_derive_seed_u64(uint64_t master_seed, uint64_t index);

/* Copies the engine's simple-point tables to the host (2D: 256 entries as
   bytes; 3D: 2^20 little-endian 64-bit words), for verification. */
cpm_status cpm_engine_sp_tables(cpm_engine* engine, uint8_t* table2d, uint64_t* table3d_words);

/* The Metropolis test of exact replay (crates/cpm-core/src/
   monte_carlo.rs:147-158: accept iff u < exp(-delta_h)), for verification
   against the reference's libm.  For each i: exp_libm[i] = exp(-delta_h[i])
   from the device libm, exp_cr[i] = exp(-delta_h[i]) correctly rounded
   (double-double evaluation; results >= 2^-1022), and, when u is not
   NULL, accept[i] = the kernels' decision for u[i]: the libm value decides
   unless u[i] lies within 2^-40 (relative) of it, and there the correctly
   rounded one does.  engine NULL computes the same on the host (host
   libm).  Any output may be NULL only together with u for accept. */
cpm_status cpm_metropolis_check(cpm_engine* engine, const double* delta_h, const double* u,
                                size_t n, double* exp_libm, double* exp_cr, uint8_t* accept);

/* The Act term's geometric-mean root (energy.rs:443: product.powf(1.0 /
   count as f64)) as the exact-replay kernels compute it: out[i] =
   product[i]^(1/count[i]), correctly rounded, for products of non-negative
   integer activities (count >= 1).  engine NULL computes the same on the
   host. */
cpm_status cpm_act_root_check(cpm_engine* engine, const double* product, const uint32_t* count,
                              size_t n, double* out);

/* The Metropolis-Hastings decision (monte_carlo.rs:168-196) as the
   exact-replay kernels take it: accept[i] = u[i] < min((n_a[i] / n_b[i])
   exp(-delta_h[i]), 1), with the ratio recomputed from the correctly
   rounded exp when u[i] lies within 2^-40 (relative) of the libm one.
   n_a[i] >= 1, n_b[i] >= 1.  engine NULL computes the same on the host. */
cpm_status cpm_mh_check(cpm_engine* engine, const uint32_t* n_a, const uint32_t* n_b,
                        const double* delta_h, const double* u, size_t n, uint8_t* accept);

/* The EdgeList proposal's per-MCS draw k ~ Binomial(n, p) (monte_carlo.rs:
   359-366, rand_distr 0.4.3 Binomial::sample) as the exact-replay kernels
   take it: for each i, one sample from the Xoshiro256++ stream whose state
   is rng_state[4i..4i+4], written to k[i], and the advanced state written
   back to rng_state.  0 <= p[i] <= 1.  engine NULL computes the same on
   the host. */
cpm_status cpm_binomial_check(cpm_engine* engine, uint64_t* rng_state, const uint64_t* n,
                              const double* p, size_t count, uint64_t* k);

/* Exact-replay sigma placement (no counterpart in the reference, whose
   lattices live in host memory).  A batch keeps each lattice's sigma in
   LDS or in HBM; by default an autotuner picks per batch signature.
   order 1 fixes "smallest lattices first", 2 "longest chains first", each
   filling at most `share` of the chip's LDS (a tiny share puts every
   sigma in HBM); order 0 returns the choice to the autotuner.  Results do
   not depend on the placement. */
enum { CPM_PLACEMENT_AUTO = 0, CPM_PLACEMENT_SMALLEST = 1, CPM_PLACEMENT_LONGEST = 2 };
cpm_status cpm_engine_set_sigma_placement(cpm_engine* engine, int32_t order, double share);
/* The placement the engine's last exact-replay call used (the caller's,
   or the autotuner's choice for that batch): order and share as above,
   order 0 before any such call. */
cpm_status cpm_engine_sigma_placement(cpm_engine* engine, int32_t* order, double* share);

/* Exact-replay pipelining (no counterpart in the reference): by default
   the replicas with the longest chains run on a two-wave kernel (one wave
   decodes and evaluates ahead while the other walks the chain) while the
   batch leaves SIMDs free, and so do the long chains whose lattices the
   LDS budget leaves in HBM (counters.piped_stragglers); on = 0 runs every
   replica on one wave.  Results do not depend on it. */
cpm_status cpm_engine_set_pipelining(cpm_engine* engine, int32_t on);

/* Where batches initialise (no counterpart in the reference; the result
   is the same either way): by default cpm_run_batch and
   cpm_run_batch_configs run scatter_and_grow (initialization.rs:202-330)
   on the device, one wave per model of at most 4096 cells, drawing from
   each model's stream in the reference's order, and so does
   cpm_model_initialize_batch; on = 0 initialises every model on the host
   threads, as cpm_model_initialize does. */
cpm_status cpm_engine_set_device_init(cpm_engine* engine, int32_t on);

/* Copy attempts per second of the model's own chain in its last
   exact-replay call on the wave path, from the device clock around the
   chain (0 before any such call).  No counterpart in the reference: a
   diagnostic for schedulers and load balance. */
cpm_status cpm_model_chain_rate(cpm_model* model, double* attempts_per_s);

/* How often each exact-replay wave-kernel variant has been launched by
   this engine: writes up to cap (key, launches) pairs and sets *n to the
   number of variants launched so far.  Key = variant * 2000 + base, where
   variant is 0 (plain), 1 (pipelined: two waves per replica), 2
   (time-limited) or 3 (pipelined and time-limited), and base is D * 100 +
   the energy-stencil size (204, 208, 306, 318, 326), plus 1000 when sigma
   is read from HBM instead of LDS. */
cpm_status cpm_engine_kernel_launches(cpm_engine* engine, int32_t* keys, uint64_t* launches,
                                      size_t cap, size_t* n);

/*
 * Serving.  No counterpart in the reference, which runs one configuration
 * per call; each request here is one output::run_with
 * (crates/cpm-core/src/output.rs:434-529) and returns the same RunResult.
 * A server owns a worker thread that admits queued requests in arrival
 * order while their device footprint fits the HBM budget, and advances all
 * admitted requests together in rounds (continuous batching): in each
 * round every exact-replay chain runs for round_ms on the device's clock
 * (stopping mid-MCS if need be; the next round resumes it) and a
 * sublattice-mode request for quantum_attempts attempts, never past its
 * next sample, so requests that arrive while others run join at the next
 * round and leave when done.  A request's result is the same alone and in
 * any mix: exact replay follows the reference's stream whatever the
 * chunking, and sublattice randomness is indexed by absolute MCS.
 */
typedef struct cpm_server cpm_server;
typedef struct {
    uint64_t hbm_budget_bytes; /* 0: 80 % of the device's free memory      */
    uint32_t max_active;       /* requests on the device at once; 0: 4096  */
    uint64_t quantum_attempts; /* per sublattice-mode request per round;    */
                               /* 0: 2^18                                   */
    double round_ms;           /* device time of an exact-replay chain per  */
                               /* round (at most 40 s); 0: 30 ms            */
} cpm_server_options;
enum { CPM_REQ_QUEUED = 0, CPM_REQ_ACTIVE = 1, CPM_REQ_DONE = 2, CPM_REQ_FAILED = 3 };
typedef struct {
    int32_t state;                    /* CPM_REQ_*                          */
    double submit_s, admit_s, done_s; /* seconds since the server started   */
    uint64_t footprint_bytes, mcs_done, mcs_total;
} cpm_request_info;
typedef struct {
    uint64_t submitted, completed, failed, queued, active;
    uint64_t hbm_budget_bytes, hbm_in_use_bytes;
    uint64_t rounds, attempts;
    double busy_s;   /* time in rounds                                       */
    double device_s; /* of which in the batched engine calls                 */
    double uptime_s;
} cpm_server_stats;
cpm_status cpm_serve_create(cpm_engine* engine, const cpm_server_options* options,
                            cpm_server** out);
/* Stops the worker after its current round; unfinished requests are
   dropped. */
void cpm_serve_destroy(cpm_server* server);
/* Validates and initialises the configuration in the caller's thread
   (errors as cpm_model_create / cpm_model_initialize), then queues it. */
cpm_status cpm_serve_submit(cpm_server* server, const cpm_config* config,
                            int32_t include_lattice, int32_t mode, uint64_t* request_id,
                            int32_t* used_default);
/* Waits up to timeout_s (0: returns at once).  *result is set (and the
   request forgotten) once it is done, else NULL; a failed request returns
   its error (and is forgotten). */
cpm_status cpm_serve_poll(cpm_server* server, uint64_t request_id, double timeout_s,
                          cpm_request_info* info, cpm_result** result);
cpm_status cpm_serve_stats(cpm_server* server, cpm_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif

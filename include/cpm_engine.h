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
   that fell within 4 ulp of exp(-dH) (decided with the device exp). */
typedef struct {
    uint64_t attempts, non_null, topology, energy, accepted, exp_near_ties;
    double kernel_ms; /* device time of the launches, HIP events */
    /* the exact-replay LDS kernel alone: its launches and their summed
       durations, from HIP events recorded around each launch on the
       stream it runs on */
    uint64_t replay_launches;
    double replay_launch_ms;
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

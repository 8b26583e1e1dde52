/*
 * CPU oracle for the Metropolis copy-attempt engine — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's serial step loop
 * (crates/cpm-core/src/{monte_carlo,energy,connectivity,simple_point,
 * dynamics,lattice,state,initialization,rng}.rs) and of the third-party
 * arithmetic it depends on (rand_xoshiro 0.6.0 seeding and Xoshiro256++,
 * rand 0.8 integer/float sampling, blake3 1.8.7 hashing).  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load
 * this library; the product library never links or calls it.
 *
 * Pins (every function below has at least one check that does not just
 * restate its formula):
 *   - RNG, sampling, FP64 energy order, commit order, halo refresh: the
 *     reference's frozen trajectory (crates/cpm-core/tests/
 *     dimension_consistency.rs:113-131) — volume, interface, energy bits
 *     and model_hash of the final lattice.
 *   - BLAKE3: published test vectors plus derive_seed_u64 known answers
 *     (notebooks/batch_execution_demo.ipynb:179-183) and the frozen
 *     model_hash (multi-chunk input).
 *   - Simple-point tables: slow flood-fill definition vs bitmask checker,
 *     the reference's hand-built cases, the 2D ring-walk cross-check and
 *     cube-symmetry invariance.
 *   - scatter_and_grow: structural invariants (connected cells, volumes
 *     <= V*, contact/interface recompute identities) and derived_phi of the
 *     README quickstart (notebooks/performance_demo.ipynb:41).  The exact
 *     draw order of scatter_and_grow has no reference-held golden value:
 *     parity unpinned beyond those invariants.
 */
#ifndef CPM_ORACLE_H
#define CPM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OC_FIXED = 0, OC_PERIODIC = 1 };
enum {
    OC_NB_VON_NEUMANN = 0,
    OC_NB_MOORE = 1,
    OC_NB_EIGHTEEN = 2,
    OC_NB_TWENTY_SIX = 3
};
enum { OC_HALO_FULL_RESCAN = 0, OC_HALO_MIRROR = 1 };
enum { OC_METROPOLIS = 0, OC_METROPOLIS_HASTINGS = 1 };

typedef struct ocpm ocpm;

/* ---- RNG (rng.rs:7-40; rand_xoshiro 0.6.0; rand 0.8) ---- */
void oc_rng_seed(uint64_t seed, uint64_t out_state[4]);
uint64_t oc_rng_next(uint64_t state[4]);
uint64_t oc_rng_gen_range(uint64_t state[4], uint64_t n);      /* 0..n   */
uint64_t oc_rng_gen_range_incl(uint64_t state[4], uint64_t hi); /* 0..=hi */
double oc_rng_gen_f64(uint64_t state[4]);

/* ---- BLAKE3 (blake3 1.8.7, rng.rs:20-40, config.rs:510,554) ---- */
void oc_blake3(const uint8_t* data, size_t len, uint8_t out[32]);
uint64_t oc_derive_seed_u64(uint64_t master_seed, uint64_t index);

/* ---- simple-point tables (simple_point.rs:35-366) ---- */
int oc_sp_slow(int dim, uint32_t pattern);      /* flood-fill definition */
int oc_sp_fast(int dim, uint32_t pattern);      /* bitmask checker      */
void oc_sp_table2d(uint8_t out[256]);
/* 2^26 bits into out_words[2^20] (bit p of word p/64) */
void oc_sp_table3d(uint64_t* out_words);
/* the canonical full-neighbourhood offset order (neighborhood.rs:75-95) */
int oc_full_offsets(int dim, int32_t* out /* (3^dim-1)*dim */);

/* ---- model (model.rs, state.rs, lattice.rs) ---- */
ocpm* oc_model_new(int dim, const uint64_t* dims, int boundary, int copy_nb,
                   int energy_nb, int n_types, const double* lambda_v,
                   const double* lambda_i, const uint32_t* target_v,
                   const uint32_t* target_i, const double* adhesion,
                   uint64_t seed);
void oc_model_free(ocpm* m);
void oc_set_halo_mode(ocpm* m, int mode);
/* acceptance rule and Act term (config.rs:84-103, 140-200); Act parameters
   per type, ignored when act == 0 */
void oc_set_modes(ocpm* m, int acceptance, int act, const double* lambda_act,
                  const uint32_t* max_act);
/* the connectivity check of each attempt (default on; the exact-Boltzmann
   checks switch it off as exact_boltzmann.rs does) */
void oc_set_connectivity_check(ocpm* m, int on);
/* ProposalMode::EdgeList (monte_carlo.rs:336-351; edge_list.rs): call
   before initialisation */
void oc_set_proposal(ocpm* m, int edge_list);
uint64_t oc_edge_sites(const ocpm* m, uint64_t* out, uint64_t cap);
void oc_set_edge_sites(ocpm* m, const uint64_t* in, uint64_t n);
/* rand_distr 0.4.3 Binomial(n, p).sample on the given stream */
uint64_t oc_binomial(uint64_t state[4], uint64_t n, double p);
/* per-site birth MCS of the Act term, interior row-major */
void oc_birth_mcs(const ocpm* m, uint32_t* out);
void oc_set_birth_mcs(ocpm* m, const uint32_t* in);
double oc_act_delta(const ocpm* m, const uint64_t* target_coord, int copy_k, uint64_t mcs);
void oc_root(const double* x, const uint32_t* n, uint64_t len, double* out);
/* initialization.rs:61-137 (explicit) — returns 0 or a nonzero error code */
int oc_init_explicit(ocpm* m, const uint32_t* lattice, uint64_t n_cells,
                     const uint64_t* cell_type_of);
/* initialization.rs:202-330 (scatter_and_grow) — returns 0 or nonzero */
int oc_init_scatter(ocpm* m, const uint64_t* cell_counts);
/* arbitrary lattice, no checks, derived state recomputed (checker view) */
void oc_load_state(ocpm* m, const uint32_t* lattice, uint64_t n_cells,
                   const uint64_t* cell_type_of);
/* monte_carlo.rs:327-357 */
void oc_run_mcs(ocpm* m, uint64_t n);
/* energy.rs:561-565 */
double oc_global_energy(const ocpm* m);

/* ---- state access ---- */
uint64_t oc_n_sites(const ocpm* m);
uint64_t oc_n_cells(const ocpm* m);
uint64_t oc_mcs(const ocpm* m);
double oc_energy(const ocpm* m);
void oc_rng_state(const ocpm* m, uint64_t out[4]);
void oc_lattice(const ocpm* m, uint32_t* out /* n_sites, row-major */);
void oc_padded_lattice(const ocpm* m, uint32_t* out /* padded */);
void oc_volume(const ocpm* m, uint32_t* out);
void oc_interface(const ocpm* m, uint32_t* out);
void oc_type_index(const ocpm* m, uint32_t* out);
void oc_sum_position(const ocpm* m, int64_t* out /* n*D */);
void oc_sum_outer(const ocpm* m, int64_t* out /* n*D*D */);
void oc_contact(const ocpm* m, uint32_t* out /* (n+1)^2 */);
/* out = {A: attempts, N: non-null (sigma_t != sigma_s), Nt: reached the
   simple-point table, E: energy evaluations, C: accepted} */
void oc_counters(const ocpm* m, uint64_t out[5]);

/* recompute from scratch (initialization.rs:396-436; state.rs:127-157) */
void oc_recompute(const ocpm* m, uint32_t* interface, int64_t* sum_pos,
                  int64_t* sum_outer, uint32_t* contact);

/* labelling.rs:70-154: returns number of fragmented cells written; medium
   component volumes/touches written up to max_components; the total
   number of medium components goes to *n_medium */
uint64_t oc_label(const ocpm* m, uint32_t* fragmented, uint64_t max_frag,
                  uint32_t* medium_volume, uint8_t* medium_touch,
                  uint64_t max_components, uint64_t* n_medium);

/* Records (dH, u) of every uphill Metropolis test (monte_carlo.rs:158)
   into buf, up to cap pairs; returns the tests seen since the last call. */
uint64_t oc_metropolis_log(ocpm* m, double* buf, uint64_t cap);
/* exp(-x[i]) with the platform libm (glibc here, as in the reference's CI). */
void oc_exp_neg(const double* x, uint64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif

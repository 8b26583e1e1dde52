// Interface between the host library and the device engine (engine.hip).
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/cpm_engine.h"

namespace cpm {

struct Model;
struct Sample;
struct Engine;

Engine* engine_create(int device);
void engine_destroy(Engine* e);

// Advance every model by n_mcs MCS on the device (uploads on first use).
void run_mcs(Engine* e, Model* const* models, size_t n, uint64_t n_mcs, int mode,
             cpm_counters* counters);
// Advance models by per-model MCS counts in one launch (todo[i] may be 0).
// With budget_s > 0 (exact replay), each wave-path chain also stops once it
// has run budget_s seconds, possibly mid-MCS (Model::attempt_in_mcs), and a
// lane-path model (more than 4095 cells) runs at most one MCS; todo counts
// an MCS in progress as one.  The results do not depend on where calls stop.
void run_mcs_each(Engine* e, Model* const* models, size_t n, const uint64_t* todo, int mode,
                  cpm_counters* counters, double budget_s = 0.0);
// run_mcs in exact replay without allocating device memory or events:
// every model must already be resident on this engine's device; the
// call's descriptors live in the library's fixed scratch array (at most
// CPM_NOALLOC_MAX_REPLICAS models per call).
void run_mcs_noalloc(Engine* e, Model* const* models, size_t n, uint64_t n_mcs,
                     cpm_counters* counters);
// One model per launch, one launch per MCS.
void run_mcs_plain(Engine* e, Model* m, uint64_t n_mcs, cpm_counters* counters);

// take_sample (output.rs:289-337) of each model's device-resident state,
// readouts computed on the device (at_mcs / at_energy name the state).
void device_samples(Engine* e, Model* const* models, size_t n, bool include_lattice,
                    const uint64_t* at_mcs, const double* at_energy, Sample* out);
// The engine a model is resident on (nullptr: host only).
Engine* device_engine(const Model* m);
// Bring the host copy of the state up to date (no-op when not stale).
void sync_to_host(Model* m);
// Scalars only: energy, rng, mcs.
void sync_scalars(Model* m);
// Drop the device copy of a model.
void device_release(Model* m);

void engine_sp_tables(Engine* e, uint8_t* t2, uint64_t* t3);
// Fix the sigma placement of exact-replay calls (order 1: smallest
// lattices first, 2: longest chains first, filling `share` of the chip's
// LDS) or hand it back to the autotuner (order 0).
void engine_set_placement(Engine* e, int order, double share);
// The placement the last exact-replay call used (order 0: none yet).
void engine_last_placement(Engine* e, int* order, double* share);
// exp(-dh[i]) with the device libm and correctly rounded, and (when u is
// given) the exact-replay kernels' Metropolis decision for u[i].
void engine_metropolis_check(Engine* e, const double* dh, const double* u, size_t n,
                             double* p_libm, double* p_cr, uint8_t* acc);
// the Act root and the Metropolis-Hastings decision as the lane path
// computes them (device copies of the inputs, one launch)
void engine_act_root_check(Engine* e, const double* product, const uint32_t* count, size_t n,
                           double* out);
void engine_mh_check(Engine* e, const uint32_t* n_a, const uint32_t* n_b, const double* dh,
                     const double* u, size_t n, uint8_t* accept);
void engine_binomial_check(Engine* e, uint64_t* rng_state, const uint64_t* n, const double* p,
                           size_t count, uint64_t* k);
// Allow (default) or forbid the pipelined two-wave exact-replay kernel.
void engine_set_pipelining(Engine* e, bool on);
// (instantiation key, launches) of the exact-replay wave kernels so far.
std::vector<std::pair<int, uint64_t>> engine_launch_counts(Engine* e);

// initialization::initialize of n models (initialization.rs:39-55): the
// default scatter_and_grow runs on the device, one wave per model, for
// models of at most 4096 cells; the others initialise on the host.
// Throws the lowest failing index's error.
void device_initialize(Engine* e, Model* const* models, size_t n);

// Device bytes a model occupies once uploaded.
size_t device_footprint(Engine* e, const Model& m);
// Whether sublattice mode can run the model (LDS-resident lattice, even
// periodic sides).
bool sublattice_supported(Engine* e, const Model& m);
// Free device memory now.
size_t device_free_bytes(Engine* e);

}  // namespace cpm

// Host-side model of one replica: the resolved configuration, the
// halo-padded lattice and per-cell bookkeeping in the layout the device
// kernels use, plus the host-only steps around the hot loop (validation,
// initial placement, readouts, hashes).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cpm_engine.h"
#include "core.hpp"

namespace cpm {

constexpr int kMaxTypes = 15;  // J is (T+1)^2 <= 256 entries on the device

struct Error : std::runtime_error {
    cpm_status code;
    Error(cpm_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct CellType {
    std::string name;
    uint32_t target_volume = 0, target_interface = 0;
    double lambda_volume = 0, lambda_interface = 0, lambda_act = 0;
    uint32_t max_act = 0;
};

// ResolvedConfig<D> (crates/cpm-core/src/config.rs:262-283)
struct Config {
    uint32_t dim = 2;
    uint64_t grid[3] = {0, 0, 0};
    int boundary = CPM_BOUNDARY_PERIODIC;
    int copy_nb = NB_VON_NEUMANN, energy_nb = NB_MOORE, conn_nb = NB_VON_NEUMANN;
    int acceptance = CPM_ACCEPT_METROPOLIS, proposal = CPM_PROPOSAL_UNIFORM;
    bool act = false;
    std::vector<CellType> types;
    std::vector<uint64_t> counts;
    std::vector<std::vector<double>> adhesion;
    bool init_explicit = false;
    std::vector<uint32_t> init_lattice;
    std::vector<uint64_t> init_cell_type_of;
    bool warn_on_default_init = true;
    uint64_t burn_in = 0, readout = 0, interval = 0, seed = 0;
};

Config resolve(const cpm_config& c);  // UserConfig::resolve, throws Error
void validate(const Config& c);       // ResolvedConfig::validate, throws Error

std::string rust_f64(double v);       // f64 Display
std::string grid_debug(const Config& c);

struct Device;  // device residency of one model (engine.hip)

// scatter_and_grow's inputs derived from the configuration: the type of
// each cell in id order, the minimum seed separation and the smallest
// integer sum of squares whose square root reaches it.
struct ScatterPlan {
    std::vector<uint32_t> type_of;
    double min_sep = 0.0;
    uint64_t sep_sq = 0;
};

// CPM<D> + State<D> (crates/cpm-core/src/model.rs:15-38, state.rs:168-203)
struct Model {
    Config cfg;
    int D = 2;
    uint64_t dims[3] = {1, 1, 1}, pdims[3] = {1, 1, 1}, strides[3] = {1, 1, 1};
    uint64_t n_sites = 0, n_padded = 0;
    int n_copy = 0, n_energy = 0, n_topo = 0, n_conn = 0;
    int64_t copy_off[26], energy_off[26], topo_off[26], conn_off[26];
    std::vector<uint32_t> sigma;  // padded, halo mirrors the opposite face
    uint64_t n_cells = 0;
    std::vector<uint32_t> type_index, volume, interface, contact;
    std::vector<int64_t> sum_pos, sum_outer;
    // Act term (crates/cpm-core/src/energy.rs:329-344): the MCS at which
    // each site was last taken by its owner, padded-lattice indexed (halo
    // entries stay 0); empty when the term is off
    std::vector<uint32_t> birth;
    // EdgeList proposal (crates/cpm-core/src/edge_list.rs:42-186): the edge
    // sites (padded flat indices) in list order, the first n_edge entries
    // of edge_sites, and each padded site's position in the list
    // (kNotEdge when absent); empty under the uniform proposal
    std::vector<uint32_t> edge_sites, edge_pos;
    uint64_t n_edge = 0;
    static constexpr uint32_t kNotEdge = 0xffffffffu;
    double energy = 0.0;
    Xoshiro rng{};
    uint64_t mcs = 0;
    bool initialized = false, used_default = false;
    // sublattice mode: key of its counter-based randomness, drawn once from
    // `rng` on first use; draws are indexed by absolute MCS, so a run gives
    // the same result however its MCS are split into calls
    uint64_t sublattice_key = 0;
    bool has_sublattice_key = false;
    // attempts per second of this model's exact-replay chain in its last
    // wave-path call (0 until measured)
    double attempt_rate = 0.0;
    // attempts done in the MCS in progress: non-zero only between the
    // time-limited calls of a serving loop (a chain out of time stops
    // mid-MCS; the next call resumes the same MCS)
    uint32_t attempt_in_mcs = 0;

    // the device copy is authoritative while `device_dirty` (host arrays
    // stale) — see engine.hip
    Device* dev = nullptr;
    bool device_dirty = false;

    explicit Model(const Config& c);
    ~Model();

    uint64_t flat(const uint64_t* c) const;
    void interior_coord(uint64_t flat, uint64_t* c) const;
    bool is_interior(uint64_t flat) const;
    uint64_t wrapped(uint64_t flat) const;
    void refresh_halo();
    // EdgeList::build (edge_list.rs:121-131): interior sites in row-major
    // order whose copy neighbourhood holds another owner
    void build_edge_list();
    void initialize();  // initialization::initialize, throws Error(INIT)
    // scatter_and_grow in parts, for a device that places and grows the
    // cells: the plan (throws Error(INIT) for zero cells), the error for a
    // seed that finds no room, and the steps after growth (sigma and
    // volume hold the grown cells, rng the state after the last draw)
    ScatterPlan scatter_plan() const;
    [[noreturn]] void crowded_error(const ScatterPlan& plan) const;
    void finish_scatter(const ScatterPlan& plan);
    // Replaces the state by a caller's (cpm_model_set_state); throws
    // Error(ARG) and leaves the model unchanged on invalid input.
    void set_state(const cpm_model_state& st);
    void recompute(std::vector<uint32_t>& itf, std::vector<int64_t>& sp,
                   std::vector<int64_t>& so, std::vector<uint32_t>& ct) const;
    double global_energy() const;
    std::vector<uint32_t> unpadded_lattice() const;
    // the interior of a padded lattice, row-major
    std::vector<uint32_t> unpadded(const uint32_t* padded) const;
    size_t ckey(uint32_t a, uint32_t b) const {
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        return (size_t)lo * (n_cells + 1) + hi;
    }
};

// --- readouts (crates/cpm-core/src/output.rs, labelling.rs) ---
struct Sample {
    uint64_t mcs = 0;
    double energy = 0;
    std::vector<uint32_t> ids, volume, interface, contact, fragmented, medium_volumes;
    std::vector<uint8_t> medium_touch;
    std::vector<double> centroid, eigen, rg2, kappa;
    std::vector<uint32_t> lattice;
    bool has_lattice = false;
};

// A readout of the model's host copy; mcs and energy name the state it was
// synced at (the two scalars change with every device call, the host copy
// only when it is synced again).
Sample take_sample(const Model& m, bool include_lattice, uint64_t mcs, double energy);
inline Sample take_sample(const Model& m, bool include_lattice) {
    return take_sample(m, include_lattice, m.mcs, m.energy);
}
bool not_equilibrated(const std::vector<double>& trace);
// shape_descriptors (output.rs:147-187) from a cell's gyration tensor g
// (D x D): eigenvalues in descending order, R_g^2 and kappa^2.
void shape_from_tensor(int D, const double* g, double* eigen, double* rg2, double* kappa);

struct Result {
    std::string cpm_version, convention_hash, model_hash, seed_policy, status_kind;
    uint64_t seed = 0;
    double derived_phi = 0;
    bool status_has_cell = false;
    uint32_t status_cell = 0;
    uint64_t status_mcs = 0;
    std::vector<Sample> samples;
    std::vector<uint32_t> cell_type_index;
    uint32_t dim = 2;
    uint64_t n_cells = 0;
    bool used_default = false;
};

// Status rules and metadata after the readout window (output.rs:470-529).
void finish_result(const Model& m, const std::vector<double>& trace, Result& r);

std::string convention_hash(const Config& c);
std::string model_hash(const Model& m);

// theta (crates/cpm-core/src/theta.rs)
size_t theta_len(size_t n_types);
std::vector<double> theta_extract(const Config& c);
void theta_inject(Config& c, const double* theta, size_t len);
std::string theta_name(const Config& c, size_t index);

// BLAKE3 (blake3 1.8.7 as used by rng.rs:20-40 and config.rs:510,554)
void blake3(const uint8_t* data, size_t len, uint8_t out[32]);
uint64_t derive_seed_u64(uint64_t master, uint64_t index);

// simple-point tables, generated from the definition
void sp_table2d(uint8_t out[256]);
bool sp_is_simple(int dim, uint32_t pattern);

}  // namespace cpm

// Host-side steps of the engine: config resolution and validation,
// initial placement, from-scratch recomputation, readouts, run records and
// theta handling.  Each function cites the reference code whose behaviour
// it keeps (paths relative to the reference repository).
#include "host.hpp"

#include <algorithm>
#include <array>
#include <charconv>
#include <cmath>
#include <cstring>
#include <deque>

#include "device.hpp"

namespace cpm {

// ----------------------------------------------------------------------
// formatting helpers matching Rust's Display / Debug output
// ----------------------------------------------------------------------

std::string rust_f64(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[400];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::fixed);
    return std::string(buf, r.ptr);
}

std::string grid_debug(const Config& c) {
    std::string s = "[";
    for (uint32_t d = 0; d < c.dim; ++d) {
        if (d) s += ", ";
        s += std::to_string(c.grid[d]);
    }
    return s + "]";
}

static const char* nb_debug(int k) {
    switch (k) {
        case NB_VON_NEUMANN: return "VonNeumann";
        case NB_MOORE: return "Moore";
        case NB_EIGHTEEN: return "Eighteen";
        case NB_TWENTY_SIX: return "TwentySix";
        default: return "TwentySixWeighted";
    }
}

// ----------------------------------------------------------------------
// UserConfig::resolve (crates/cpm-core/src/config.rs:200-257) plus the
// builder's active_terms (crates/cpm-py/src/builder.rs:420-425)
// ----------------------------------------------------------------------

Config resolve(const cpm_config& u) {
    Config c;
    if (u.dim != 2 && u.dim != 3)
        throw Error(CPM_ERR_CONFIG,
                    "grid must have length 2 or 3 (dimensionality is inferred from it, "
                    "not from an explicit `dimension` argument); got length " +
                        std::to_string(u.dim));
    c.dim = u.dim;
    for (uint32_t d = 0; d < u.dim; ++d) c.grid[d] = u.grid[d];
    if (u.adhesion_rows == 0) throw Error(CPM_ERR_CONFIG, "adhesion matrix is required");
    c.boundary = u.boundary;
    c.copy_nb = u.copy_neighborhood;
    c.energy_nb = u.energy_neighborhood == CPM_NB_DEFAULT
                      ? (u.dim == 2 ? NB_MOORE : NB_EIGHTEEN)
                      : u.energy_neighborhood;
    c.conn_nb = u.connectivity_neighborhood;
    c.acceptance = u.acceptance;
    c.proposal = u.proposal;
    c.act = u.act_active != 0;
    for (uint32_t t = 0; t < u.n_types; ++t) {
        CellType ct;
        ct.name = u.type_names && u.type_names[t] ? u.type_names[t] : "";
        ct.target_volume = u.target_volume[t];
        ct.target_interface = u.target_interface[t];
        ct.lambda_volume = u.lambda_volume[t];
        ct.lambda_interface = u.lambda_interface[t];
        ct.lambda_act = u.lambda_act ? u.lambda_act[t] : 0.0;
        ct.max_act = u.max_act ? u.max_act[t] : 0;
        c.types.push_back(ct);
        c.counts.push_back(u.cell_counts ? u.cell_counts[t] : 0);
    }
    size_t pos = 0;
    for (uint32_t r = 0; r < u.adhesion_rows; ++r) {
        c.adhesion.emplace_back(u.adhesion + pos, u.adhesion + pos + u.adhesion_row_len[r]);
        pos += u.adhesion_row_len[r];
    }
    c.init_explicit = u.init_explicit != 0;
    if (c.init_explicit) {
        c.init_lattice.assign(u.init_lattice, u.init_lattice + u.init_lattice_len);
        c.init_cell_type_of.assign(u.init_cell_type_of, u.init_cell_type_of + u.init_n_cells);
    }
    c.warn_on_default_init = u.warn_on_default_init != 0;
    c.burn_in = u.burn_in_mcs;
    c.readout = u.readout_mcs;
    c.interval = u.sampling_interval_mcs;
    c.seed = u.seed;
    return c;
}

// CellType::validate (crates/cpm-core/src/cell.rs:40-66)
static void validate_type(const CellType& t) {
    if (t.name.empty()) throw Error(CPM_ERR_CONFIG, "cell type name must not be empty");
    const std::pair<const char*, double> ls[3] = {{"lambda_volume", t.lambda_volume},
                                                   {"lambda_interface", t.lambda_interface},
                                                   {"lambda_act", t.lambda_act}};
    for (auto& [label, v] : ls)
        if (!std::isfinite(v) || v < 0.0)
            throw Error(CPM_ERR_CONFIG, std::string(label) +
                                            " must be finite and non-negative, got " +
                                            rust_f64(v));
    if (t.lambda_act > 0.0 && t.max_act == 0)
        throw Error(CPM_ERR_CONFIG, "max_act must be positive when lambda_act is positive");
}

// NeighborhoodKind::to_stencil's dimension checks (neighborhood.rs:179-263)
static void stencil_check(int kind, uint32_t dim) {
    const std::string D = std::to_string(dim);
    if (kind == NB_MOORE && dim != 2)
        throw Error(CPM_ERR_CONFIG, "Moore neighbourhood is 2D only, got D = " + D);
    if (kind == NB_EIGHTEEN && dim != 3)
        throw Error(CPM_ERR_CONFIG, "Eighteen-neighbour is 3D only, got D = " + D);
    if (kind == NB_TWENTY_SIX && dim != 3)
        throw Error(CPM_ERR_CONFIG, "Twenty-six-neighbour is 3D only, got D = " + D);
    if (kind == NB_TWENTY_SIX_WEIGHTED && dim != 3)
        throw Error(CPM_ERR_CONFIG, "Weighted twenty-six-neighbour is 3D only, got D = " + D);
}

// ResolvedConfig::validate (crates/cpm-core/src/config.rs:287-413)
void validate(const Config& c) {
    for (uint32_t d = 0; d < c.dim; ++d)
        if (c.grid[d] == 0)
            throw Error(CPM_ERR_CONFIG,
                        "every grid dimension must be nonzero, got " + grid_debug(c));
    if (c.conn_nb != NB_VON_NEUMANN)
        throw Error(CPM_ERR_CONFIG,
                    std::string("connectivity_neighborhood must be VonNeumann in this version "
                                "(the simple-point tables only support the (4,8)/(6,26) "
                                "pairing), got ") +
                        nb_debug(c.conn_nb));
    stencil_check(c.copy_nb, c.dim);
    if (c.copy_nb != NB_VON_NEUMANN)  // every other stencil has a diagonal offset
        throw Error(CPM_ERR_CONFIG, std::string("copy_neighborhood ") + nb_debug(c.copy_nb) +
                                        " must be a subset of connectivity_neighborhood " +
                                        nb_debug(c.conn_nb));
    stencil_check(c.energy_nb, c.dim);
    if (c.energy_nb == NB_TWENTY_SIX_WEIGHTED)
        throw Error(CPM_ERR_CONFIG,
                    "weighted energy neighbourhoods are not supported yet: interface "
                    "bookkeeping is integer-typed; use an unweighted energy_neighborhood "
                    "(Moore, Eighteen, or TwentySix)");
    if (c.types.empty()) throw Error(CPM_ERR_CONFIG, "at least one cell type is required");
    for (auto& t : c.types) validate_type(t);
    if (c.counts.size() != c.types.size())
        throw Error(CPM_ERR_CONFIG, "cell_counts has " + std::to_string(c.counts.size()) +
                                        " entries, expected one per cell type (" +
                                        std::to_string(c.types.size()) + ")");
    const size_t n1 = c.types.size() + 1;
    bool bad_shape = c.adhesion.size() != n1;
    for (auto& r : c.adhesion) bad_shape |= r.size() != n1;
    if (bad_shape)
        throw Error(CPM_ERR_CONFIG,
                    "adhesion matrix must be " + std::to_string(n1) + "x" + std::to_string(n1) +
                        " (n_types + 1 for medium), got " + std::to_string(c.adhesion.size()) +
                        "x" + std::to_string(c.adhesion.empty() ? 0 : c.adhesion[0].size()));
    for (size_t i = 0; i < n1; ++i)
        for (size_t j = 0; j < n1; ++j) {
            double a = c.adhesion[i][j], b = c.adhesion[j][i];
            if (!std::isfinite(a))
                throw Error(CPM_ERR_CONFIG, "adhesion[" + std::to_string(i) + "][" +
                                                std::to_string(j) + "] = " + rust_f64(a) +
                                                " is not finite");
            if (std::fabs(a - b) > 0.0)
                throw Error(CPM_ERR_CONFIG,
                            "adhesion matrix must be symmetric in v1: [" + std::to_string(i) +
                                "][" + std::to_string(j) + "] = " + rust_f64(a) + " but [" +
                                std::to_string(j) + "][" + std::to_string(i) +
                                "] = " + rust_f64(b));
        }
    if (c.readout == 0) throw Error(CPM_ERR_CONFIG, "readout_mcs must be positive");
    if (c.interval == 0) throw Error(CPM_ERR_CONFIG, "sampling_interval_mcs must be positive");
    if (c.proposal == CPM_PROPOSAL_EDGE_LIST && c.acceptance == CPM_ACCEPT_METROPOLIS_HASTINGS)
        throw Error(CPM_ERR_CONFIG,
                    "proposal = EdgeList is not valid with acceptance = MetropolisHastings: "
                    "the exact-Boltzmann validation's reversibility argument depends on the "
                    "uniform proposal's forward/reverse symmetry, which EdgeList does not "
                    "establish");
    if (c.proposal != CPM_PROPOSAL_UNIFORM && c.proposal != CPM_PROPOSAL_EDGE_LIST)
        throw Error(CPM_ERR_CONFIG, "unknown proposal mode " + std::to_string(c.proposal));
    if (c.acceptance != CPM_ACCEPT_METROPOLIS && c.acceptance != CPM_ACCEPT_METROPOLIS_HASTINGS)
        throw Error(CPM_ERR_CONFIG, "unknown acceptance mode " + std::to_string(c.acceptance));
    if (c.types.size() > (size_t)kMaxTypes)
        throw Error(CPM_ERR_CONFIG, "this engine supports at most " + std::to_string(kMaxTypes) +
                                        " cell types");
}

// ----------------------------------------------------------------------
// Model (crates/cpm-core/src/model.rs:44-76; lattice.rs:24-254)
// ----------------------------------------------------------------------

Model::Model(const Config& c) : cfg(c) {
    validate(cfg);
    D = (int)cfg.dim;
    n_sites = n_padded = 1;
    for (int d = 0; d < D; ++d) {
        dims[d] = cfg.grid[d];
        pdims[d] = dims[d] + 2;
        n_sites *= dims[d];
        n_padded *= pdims[d];
    }
    strides[D - 1] = 1;
    for (int d = D - 2; d >= 0; --d) strides[d] = strides[d + 1] * pdims[d + 1];
    int32_t o[26 * 3];
    auto deltas = [&](int kind, int64_t* out) {
        int n = stencil(D, kind, o);
        for (int i = 0; i < n; ++i) {
            int64_t dl = 0;
            for (int d = 0; d < D; ++d) dl += (int64_t)o[i * D + d] * (int64_t)strides[d];
            out[i] = dl;
        }
        return n;
    };
    n_copy = deltas(cfg.copy_nb, copy_off);
    n_energy = deltas(cfg.energy_nb, energy_off);
    n_conn = deltas(NB_VON_NEUMANN, conn_off);
    n_topo = deltas(D == 2 ? NB_MOORE : NB_TWENTY_SIX, topo_off);
    sigma.assign(n_padded, 0);
    if (cfg.act) birth.assign(n_padded, 0);
    if (cfg.proposal == CPM_PROPOSAL_EDGE_LIST) {
        edge_sites.assign(n_sites, 0);
        edge_pos.assign(n_padded, kNotEdge);
    }
    rng = xo_seed(cfg.seed);
}

uint64_t Model::flat(const uint64_t* c) const {
    uint64_t f = 0;
    for (int d = 0; d < D; ++d) f += (c[d] + 1) * strides[d];
    return f;
}

void Model::interior_coord(uint64_t f, uint64_t* c) const {
    for (int d = 0; d < D; ++d) {
        c[d] = f / strides[d] - 1;
        f %= strides[d];
    }
}

bool Model::is_interior(uint64_t f) const {
    for (int d = 0; d < D; ++d) {
        uint64_t p = f / strides[d];
        f %= strides[d];
        if (p == 0 || p == pdims[d] - 1) return false;
    }
    return true;
}

uint64_t Model::wrapped(uint64_t f) const {
    uint64_t c[3];
    for (int d = 0; d < D; ++d) {
        uint64_t p = f / strides[d];
        f %= strides[d];
        c[d] = p == 0 ? dims[d] - 1 : (p == pdims[d] - 1 ? 0 : p - 1);
    }
    return flat(c);
}

void Model::refresh_halo() {
    if (cfg.boundary != CPM_BOUNDARY_PERIODIC) return;
    for (uint64_t f = 0; f < n_padded; ++f)
        if (!is_interior(f)) sigma[f] = sigma[wrapped(f)];
}

std::vector<uint32_t> Model::unpadded_lattice() const { return unpadded(sigma.data()); }

std::vector<uint32_t> Model::unpadded(const uint32_t* padded) const {
    std::vector<uint32_t> out(n_sites);
    uint64_t c[3] = {0, 0, 0};
    for (uint64_t i = 0; i < n_sites; ++i) {
        out[i] = padded[flat(c)];
        for (int d = D - 1; d >= 0; --d) {
            if (++c[d] < dims[d]) break;
            c[d] = 0;
        }
    }
    return out;
}

// Row-major interior iteration helper.
template <class F>
static void each_interior(const Model& m, F&& f) {
    uint64_t c[3] = {0, 0, 0};
    for (uint64_t i = 0; i < m.n_sites; ++i) {
        f(c);
        for (int d = m.D - 1; d >= 0; --d) {
            if (++c[d] < m.dims[d]) break;
            c[d] = 0;
        }
    }
}

// initialization.rs:396-436 and state.rs:127-157
void Model::recompute(std::vector<uint32_t>& itf, std::vector<int64_t>& sp,
                      std::vector<int64_t>& so, std::vector<uint32_t>& ct) const {
    itf.assign(n_cells, 0);
    sp.assign(n_cells * D, 0);
    so.assign(n_cells * D * D, 0);
    ct.assign((n_cells + 1) * (n_cells + 1), 0);
    const bool periodic = cfg.boundary == CPM_BOUNDARY_PERIODIC;
    each_interior(*this, [&](const uint64_t* c) {
        uint64_t fi = flat(c);
        uint32_t ci = sigma[fi];
        if (ci) {
            size_t i = ci - 1;
            for (int d = 0; d < D; ++d) {
                sp[i * D + d] += (int64_t)c[d];
                for (int e = 0; e < D; ++e) so[(i * D + d) * D + e] += (int64_t)(c[d] * c[e]);
            }
            for (int k = 0; k < n_energy; ++k)
                if (sigma[fi + energy_off[k]] != ci) itf[i] += 1;
        }
        for (int k = 0; k < n_energy; ++k) {
            uint64_t fj = fi + energy_off[k];
            bool keyed = true;
            uint64_t key = fj;
            if (!is_interior(fj)) {
                if (periodic) key = wrapped(fj);
                else keyed = false;
            }
            if (keyed && key <= fi) continue;
            uint32_t cj = sigma[fj];
            if (cj != ci) ct[ckey(ci, cj)] += 1;
        }
    });
}

// energy.rs:106-114, 205-213, 269-308, 521-565
double Model::global_energy() const {
    auto term = [](double lam, uint32_t tgt, double x) {
        double d = x - (double)tgt;
        return lam * (d * d);
    };
    double vol = 0.0, itf = 0.0, adh = 0.0;
    for (size_t i = 0; i < n_cells; ++i) {
        auto& t = cfg.types[type_index[i]];
        vol += term(t.lambda_volume, t.target_volume, (double)volume[i]);
    }
    for (size_t i = 0; i < n_cells; ++i) {
        auto& t = cfg.types[type_index[i]];
        itf += term(t.lambda_interface, t.target_interface, (double)interface[i]);
    }
    const bool periodic = cfg.boundary == CPM_BOUNDARY_PERIODIC;
    auto aidx = [&](uint32_t id) { return id == 0 ? 0u : type_index[id - 1] + 1; };
    each_interior(*this, [&](const uint64_t* c) {
        uint64_t fi = flat(c);
        uint32_t ci = sigma[fi];
        for (int k = 0; k < n_energy; ++k) {
            uint64_t fj = fi + energy_off[k];
            bool keyed = true;
            uint64_t key = fj;
            if (!is_interior(fj)) {
                if (periodic) key = wrapped(fj);
                else keyed = false;
            }
            if (keyed && key <= fi) continue;
            uint32_t cj = sigma[fj];
            if (cj == ci) continue;
            adh += 1.0 * cfg.adhesion[aidx(ci)][aidx(cj)];
        }
    });
    return vol + itf + adh;
}

// ----------------------------------------------------------------------
// initialization.rs
// ----------------------------------------------------------------------

[[noreturn]] static void init_error(const std::string& m) { throw Error(CPM_ERR_INIT, m); }

void Model::initialize() {
    if (cfg.init_explicit) {
        // place_explicit (initialization.rs:61-137)
        const auto& lat = cfg.init_lattice;
        const auto& cto = cfg.init_cell_type_of;
        if (lat.size() != n_sites)
            init_error("explicit lattice has " + std::to_string(lat.size()) +
                       " entries, expected N_sites = " + std::to_string(n_sites));
        for (size_t i = 0; i < cto.size(); ++i)
            if (cto[i] >= cfg.types.size())
                init_error("cell " + std::to_string(i + 1) + " has type_index " +
                           std::to_string(cto[i]) + ", but only " +
                           std::to_string(cfg.types.size()) + " types are declared");
        for (uint32_t id : lat)
            if (id > cto.size())
                init_error("lattice references cell id " + std::to_string(id) +
                           ", but cell_type_of only covers " + std::to_string(cto.size()) +
                           " cells");
        n_cells = cto.size();
        std::fill(sigma.begin(), sigma.end(), 0u);
        volume.assign(n_cells, 0);
        size_t k = 0;
        each_interior(*this, [&](const uint64_t* c) {
            uint32_t v = lat[k++];
            sigma[flat(c)] = v;
            if (v) volume[v - 1] += 1;
        });
        refresh_halo();
        for (size_t i = 0; i < n_cells; ++i)
            if (!volume[i])
                init_error("cell " + std::to_string(i + 1) +
                           " has zero volume in the supplied lattice");
        type_index.resize(n_cells);
        for (size_t i = 0; i < n_cells; ++i) type_index[i] = (uint32_t)cto[i];
        // BFS connectivity under von Neumann (initialization.rs:145-178)
        std::vector<uint8_t> seen(n_padded);
        std::vector<uint64_t> stack;
        std::vector<uint64_t> first(n_cells + 1, UINT64_MAX);
        each_interior(*this, [&](const uint64_t* c) {
            uint64_t f = flat(c);
            uint32_t v = sigma[f];
            if (v && first[v] == UINT64_MAX) first[v] = f;
        });
        for (uint32_t id = 1; id <= n_cells; ++id) {
            uint64_t count = 0;
            stack.assign(1, first[id]);
            seen[first[id]] = 1;
            while (!stack.empty()) {
                uint64_t f = stack.back();
                stack.pop_back();
                ++count;
                for (int q = 0; q < n_conn; ++q) {
                    uint64_t nb = f + conn_off[q];
                    if (!is_interior(nb) && cfg.boundary == CPM_BOUNDARY_PERIODIC) nb = wrapped(nb);
                    if (sigma[nb] == id && !seen[nb]) {
                        seen[nb] = 1;
                        stack.push_back(nb);
                    }
                }
            }
            if (count != volume[id - 1])
                init_error("cell " + std::to_string(id) +
                           " is not connected under connectivity_neighborhood \xE2\x80\x94 "
                           "every explicitly-placed cell must already be connected");
        }
        used_default = false;
    } else {
        // scatter_and_grow (initialization.rs:202-330)
        const ScatterPlan plan = scatter_plan();
        const std::vector<uint32_t>& type_of = plan.type_of;
        const size_t n = type_of.size();
        const double min_sep = plan.min_sep;
        const bool periodic = cfg.boundary == CPM_BOUNDARY_PERIODIC;
        std::vector<uint64_t> seeds;
        seeds.reserve(3 * n);
        for (size_t i = 0; i < n; ++i) {
            bool placed = false;
            for (int a = 0; a < 20000 && !placed; ++a) {
                uint64_t cand[3];
                for (int d = 0; d < D; ++d) cand[d] = gen_range(rng, dims[d]);
                bool far = true;
                for (size_t j = 0; j < i && far; ++j) {
                    double sq = 0.0;
                    for (int d = 0; d < D; ++d) {
                        double raw = std::fabs((double)seeds[3 * j + d] - (double)cand[d]);
                        double delta = periodic ? std::fmin(raw, (double)dims[d] - raw) : raw;
                        sq += delta * delta;
                    }
                    far = std::sqrt(sq) >= min_sep;
                }
                if (far) {
                    for (int d = 0; d < 3; ++d) seeds.push_back(d < D ? cand[d] : 0);
                    placed = true;
                }
            }
            if (!placed) crowded_error(plan);
        }
        n_cells = n;
        std::fill(sigma.begin(), sigma.end(), 0u);
        volume.assign(n, 0);
        type_index = type_of;
        for (size_t i = 0; i < n; ++i) {
            sigma[flat(&seeds[3 * i])] = (uint32_t)(i + 1);
            volume[i] += 1;
        }
        // canonical_growable_neighbour (initialization.rs:308-327)
        auto growable = [&](uint64_t f, int64_t off, uint64_t* out) {
            uint64_t raw = f + off;
            if (is_interior(raw)) {
                *out = raw;
                return true;
            }
            if (periodic) {
                *out = wrapped(raw);
                return true;
            }
            return false;
        };
        std::vector<std::vector<uint64_t>> frontier(n);
        for (size_t i = 0; i < n; ++i) {
            uint64_t sf = flat(&seeds[3 * i]), nb;
            for (int k = 0; k < n_copy; ++k)
                if (growable(sf, copy_off[k], &nb) && sigma[nb] == 0) frontier[i].push_back(nb);
        }
        std::vector<size_t> order(n);
        for (size_t i = 0; i < n; ++i) order[i] = i;
        for (;;) {
            bool grew = false;
            for (size_t i : order) {
                if (volume[i] >= cfg.types[type_of[i]].target_volume) continue;
                auto& fr = frontier[i];
                fr.erase(std::remove_if(fr.begin(), fr.end(),
                                        [&](uint64_t f) { return sigma[f] != 0; }),
                         fr.end());
                if (fr.empty()) continue;
                size_t pick = (size_t)gen_range(rng, fr.size());
                uint64_t claimed = fr[pick];
                fr[pick] = fr.back();
                fr.pop_back();
                sigma[claimed] = (uint32_t)(i + 1);
                volume[i] += 1;
                uint64_t nb;
                for (int k = 0; k < n_copy; ++k)
                    if (growable(claimed, copy_off[k], &nb) && sigma[nb] == 0) fr.push_back(nb);
                grew = true;
            }
            if (!grew) break;
            for (size_t i = n - 1; i >= 1; --i) {
                size_t j = (size_t)gen_range_incl(rng, i);
                std::swap(order[i], order[j]);
            }
        }
        finish_scatter(plan);
        return;
    }
    recompute(interface, sum_pos, sum_outer, contact);
    energy = global_energy();
    build_edge_list();  // initialization.rs:427-435
    initialized = true;
}

// What scatter_and_grow derives from the configuration before its first
// draw (initialization.rs:203-224): the type of each cell in id order and
// the minimum seed separation.  Seed coordinates are integers, so every
// separation the placement tests is the square root of an exact integer
// sum of squares (periodic_distance, initialization.rs:369-386); since
// sqrt is monotone, `sqrt(s) >= min_sep` holds exactly for the integers
// s >= sep_sq, the smallest such integer.
ScatterPlan Model::scatter_plan() const {
    ScatterPlan p;
    for (size_t t = 0; t < cfg.counts.size(); ++t)
        p.type_of.insert(p.type_of.end(), cfg.counts[t], (uint32_t)t);
    if (p.type_of.empty()) init_error("cell_counts sums to zero cells \xE2\x80\x94 nothing to place");
    double avg = 0.0;
    for (auto& t : cfg.types) avg += (double)t.target_volume;
    avg /= (double)cfg.types.size();
    p.min_sep = std::sqrt(avg / 3.141592653589793);
    const double sq = p.min_sep * p.min_sep;
    uint64_t s = sq > 2.0 ? (uint64_t)sq - 2 : 0;
    while (std::sqrt((double)s) < p.min_sep) ++s;
    p.sep_sq = s;
    return p;
}

void Model::crowded_error(const ScatterPlan& plan) const {
    char sep[64];
    snprintf(sep, sizeof sep, "%.2f", plan.min_sep);
    init_error("could not place all " + std::to_string(plan.type_of.size()) +
               " seed points with minimum separation " + sep + " on a grid of " +
               grid_debug(cfg) + " \xE2\x80\x94 too crowded for this configuration");
}

// The end of scatter_and_grow once sigma and the volumes hold the grown
// cells (initialization.rs:296-305, 396-436): halo, derived state, energy.
void Model::finish_scatter(const ScatterPlan& plan) {
    n_cells = plan.type_of.size();
    type_index = plan.type_of;
    refresh_halo();
    used_default = true;
    recompute(interface, sum_pos, sum_outer, contact);
    energy = global_energy();
    build_edge_list();  // initialization.rs:427-435
    initialized = true;
}

void Model::build_edge_list() {
    if (cfg.proposal != CPM_PROPOSAL_EDGE_LIST) return;
    std::fill(edge_pos.begin(), edge_pos.end(), kNotEdge);
    n_edge = 0;
    each_interior(*this, [&](const uint64_t* c) {
        const uint64_t f = flat(c);
        bool edge = false;
        for (int k = 0; k < n_copy && !edge; ++k)
            edge = sigma[(uint64_t)((int64_t)f + copy_off[k])] != sigma[f];
        if (edge) {
            edge_pos[f] = (uint32_t)n_edge;
            edge_sites[n_edge++] = (uint32_t)f;
        }
    });
}

// The fields of CPM<D> and State<D> (model.rs:15-38, state.rs:168-203)
// taken from the caller; derived fields that are not given are computed
// from the lattice as initialization does (initialization.rs:396-436,
// energy.rs:561-565).  Only what device memory safety needs is checked.
void Model::set_state(const cpm_model_state& st) {
    auto bad = [](const std::string& m) { throw Error(CPM_ERR_ARG, "invalid state: " + m); };
    if (!st.lattice || !st.type_index) bad("lattice and type_index are required");
    if (st.n_cells < 1 || st.n_cells > n_sites || st.n_cells >= 0xffffffffull)
        bad("n_cells must be in 1..=N_sites (" + std::to_string(n_sites) + "), got " +
            std::to_string(st.n_cells));
    const size_t n = st.n_cells;
    for (size_t i = 0; i < n; ++i)
        if (st.type_index[i] >= cfg.types.size())
            bad("cell " + std::to_string(i + 1) + " has type_index " +
                std::to_string(st.type_index[i]) + ", but only " +
                std::to_string(cfg.types.size()) + " types are declared");
    for (uint64_t k = 0; k < n_sites; ++k)
        if (st.lattice[k] > n)
            bad("lattice references cell id " + std::to_string(st.lattice[k]) + ", but only " +
                std::to_string(n) + " cells are given");
    if (st.rng_state && !(st.rng_state[0] | st.rng_state[1] | st.rng_state[2] | st.rng_state[3]))
        bad("the all-zero Xoshiro256++ state is a fixed point");
    n_cells = n;
    std::fill(sigma.begin(), sigma.end(), 0u);
    size_t k = 0;
    each_interior(*this, [&](const uint64_t* c) { sigma[flat(c)] = st.lattice[k++]; });
    refresh_halo();
    type_index.assign(st.type_index, st.type_index + n);
    if (st.volume) {
        volume.assign(st.volume, st.volume + n);
    } else {
        volume.assign(n, 0);
        for (uint64_t q = 0; q < n_sites; ++q)
            if (st.lattice[q]) volume[st.lattice[q] - 1] += 1;
    }
    std::vector<uint32_t> itf, ct;
    std::vector<int64_t> sp, so;
    recompute(itf, sp, so, ct);
    interface = st.interface ? std::vector<uint32_t>(st.interface, st.interface + n) : itf;
    sum_pos = st.sum_position ? std::vector<int64_t>(st.sum_position, st.sum_position + n * D)
                              : sp;
    sum_outer = st.sum_outer
                    ? std::vector<int64_t>(st.sum_outer, st.sum_outer + n * D * D)
                    : so;
    contact = st.contact ? std::vector<uint32_t>(st.contact, st.contact + (n + 1) * (n + 1))
                         : ct;
    energy = st.has_energy ? st.conservative_energy : global_energy();
    if (cfg.act) {
        std::fill(birth.begin(), birth.end(), 0u);
        k = 0;
        if (st.birth_mcs)
            each_interior(*this, [&](const uint64_t* c) { birth[flat(c)] = st.birth_mcs[k++]; });
    }
    if (cfg.proposal == CPM_PROPOSAL_EDGE_LIST) {
        if (!st.edge_sites) {
            build_edge_list();
        } else {
            std::fill(edge_pos.begin(), edge_pos.end(), kNotEdge);
            n_edge = 0;
            for (uint64_t i = 0; i < st.n_edge_sites; ++i) {
                uint64_t r = st.edge_sites[i], c[3] = {0, 0, 0};
                if (r >= n_sites) bad("edge site " + std::to_string(r) + " is not a site");
                for (int d = D - 1; d >= 0; --d) {
                    c[d] = r % dims[d];
                    r /= dims[d];
                }
                const uint64_t f = flat(c);
                if (edge_pos[f] != kNotEdge)
                    bad("edge site " + std::to_string(st.edge_sites[i]) + " is listed twice");
                edge_pos[f] = (uint32_t)n_edge;
                edge_sites[n_edge++] = (uint32_t)f;
            }
        }
    }
    if (st.rng_state)
        for (int q = 0; q < 4; ++q) rng.s[q] = st.rng_state[q];
    mcs = st.mcs;
    attempt_in_mcs = 0;
    has_sublattice_key = false;
    attempt_rate = 0.0;
    initialized = true;
    used_default = false;
}

Model::~Model() { device_release(this); }

// ----------------------------------------------------------------------
// Readouts (crates/cpm-core/src/output.rs:147-337; labelling.rs:70-154)
// ----------------------------------------------------------------------

static void sym_eigen(int D, const double* g, double* out) {
    if (D == 2) {
        double a = g[0], b = g[1], d = g[3];
        double tr = a + d;
        double disc = std::sqrt(std::max((a - d) * (a - d) + 4.0 * b * b, 0.0));
        out[0] = (tr + disc) / 2.0;
        out[1] = (tr - disc) / 2.0;
        return;
    }
    auto G = [&](int i, int j) { return g[i * 3 + j]; };
    double p1 = G(0, 1) * G(0, 1) + G(0, 2) * G(0, 2) + G(1, 2) * G(1, 2);
    if (p1 == 0.0) {
        out[0] = G(0, 0);
        out[1] = G(1, 1);
        out[2] = G(2, 2);
        return;
    }
    double q = (G(0, 0) + G(1, 1) + G(2, 2)) / 3.0;
    double a0 = G(0, 0) - q, a1 = G(1, 1) - q, a2 = G(2, 2) - q;
    double p2 = a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * p1;
    double p = std::sqrt(p2 / 6.0);
    auto B = [&](int i, int j) { return (G(i, j) - (i == j ? q : 0.0)) / p; };
    double det = B(0, 0) * (B(1, 1) * B(2, 2) - B(1, 2) * B(2, 1)) -
                 B(0, 1) * (B(1, 0) * B(2, 2) - B(1, 2) * B(2, 0)) +
                 B(0, 2) * (B(1, 0) * B(2, 1) - B(1, 1) * B(2, 0));
    double r = std::min(std::max(det / 2.0, -1.0), 1.0);
    double phi = std::acos(r) / 3.0;
    double e1 = q + 2.0 * p * std::cos(phi);
    double e3 = q + 2.0 * p * std::cos(phi + 2.0 * 3.141592653589793 / 3.0);
    out[0] = e1;
    out[1] = 3.0 * q - e1 - e3;
    out[2] = e3;
}

void shape_from_tensor(int D, const double* g, double* eigen, double* rg2, double* kappa) {
    double ev[3];
    sym_eigen(D, g, ev);
    std::sort(ev, ev + D, [](double x, double y) { return x > y; });
    double sum = 0.0, ssq = 0.0;
    for (int d = 0; d < D; ++d) sum += ev[d];
    for (int d = 0; d < D; ++d) ssq += ev[d] * ev[d];
    double k = 0.0;
    if (sum > 0.0) {
        double dd = (double)D;
        k = (dd / (dd - 1.0)) * (ssq / (sum * sum)) - 1.0 / (dd - 1.0);
    }
    for (int d = 0; d < D; ++d) eigen[d] = ev[d];
    *rg2 = sum;
    *kappa = k;
}

static void label(const Model& m, Sample& s) {
    const bool periodic = m.cfg.boundary == CPM_BOUNDARY_PERIODIC;
    std::vector<uint8_t> visited(m.n_padded);
    std::vector<uint32_t> comps(m.n_cells + 1);
    std::vector<uint64_t> queue;
    queue.reserve(m.n_sites);
    each_interior(m, [&](const uint64_t* c) {
        uint64_t seed = m.flat(c);
        if (visited[seed]) return;
        uint32_t owner = m.sigma[seed];
        const int64_t* offs = owner == 0 ? m.topo_off : m.conn_off;
        int noff = owner == 0 ? m.n_topo : m.n_conn;
        queue.clear();
        queue.push_back(seed);
        visited[seed] = 1;
        uint32_t volume = 0;
        bool touches = false;
        for (size_t h = 0; h < queue.size(); ++h) {
            uint64_t site = queue[h];
            volume += 1;
            for (int k = 0; k < noff; ++k) {
                uint64_t nb = site + offs[k];
                if (!m.is_interior(nb)) {
                    if (owner == 0 && !periodic) touches = true;
                    if (!periodic) continue;
                    nb = m.wrapped(nb);
                }
                if (visited[nb] || m.sigma[nb] != owner) continue;
                visited[nb] = 1;
                queue.push_back(nb);
            }
        }
        if (owner == 0) {
            s.medium_volumes.push_back(volume);
            s.medium_touch.push_back(touches ? 1 : 0);
        } else {
            comps[owner] += 1;
        }
    });
    for (uint32_t id = 1; id <= m.n_cells; ++id)
        if (comps[id] > 1) s.fragmented.push_back(id);
}

Sample take_sample(const Model& m, bool include_lattice, uint64_t mcs, double energy) {
    Sample s;
    const int D = m.D;
    const size_t n = m.n_cells;
    s.mcs = mcs;
    s.energy = energy;
    s.ids.resize(n);
    s.volume = m.volume;
    s.interface = m.interface;
    s.centroid.resize(n * D);
    s.eigen.resize(n * D);
    s.rg2.resize(n);
    s.kappa.resize(n);
    for (size_t i = 0; i < n; ++i) {
        s.ids[i] = (uint32_t)(i + 1);
        double v = (double)m.volume[i];
        double cen[3], g[9];
        for (int d = 0; d < D; ++d) cen[d] = (double)m.sum_pos[i * D + d] / v;
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
                g[a * D + b] = (double)m.sum_outer[(i * D + a) * D + b] / v - cen[a] * cen[b];
        for (int d = 0; d < D; ++d) s.centroid[i * D + d] = cen[d];
        shape_from_tensor(D, g, &s.eigen[i * D], &s.rg2[i], &s.kappa[i]);
    }
    s.contact = m.contact;
    label(m, s);
    if (include_lattice) {
        s.lattice = m.unpadded_lattice();
        s.has_lattice = true;
    }
    return s;
}

// output.rs:388-420
bool not_equilibrated(const std::vector<double>& y) {
    const size_t n = y.size();
    if (n < 3) return false;
    double xm = 0.0, ym = 0.0;
    for (size_t i = 0; i < n; ++i) xm += (double)i;
    xm /= (double)n;
    for (size_t i = 0; i < n; ++i) ym += y[i];
    ym /= (double)n;
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double dx = (double)i - xm;
        num += dx * (y[i] - ym);
        den += dx * dx;
    }
    if (den == 0.0) return false;
    double slope = num / den, icpt = ym - slope * xm, rv = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double r = y[i] - (icpt + slope * (double)i);
        rv += r * r;
    }
    rv /= (double)n;
    double rs = std::sqrt(rv);
    double drift = std::fabs(slope) * (double)(n - 1);
    if (rs == 0.0) return drift > 0.0;
    return drift > 2.0 * rs;
}

// output.rs:531-559
static bool fully_connected(const std::vector<uint32_t>& contact, size_t n_cells) {
    if (n_cells < 2) return false;
    const size_t n1 = n_cells + 1;
    std::vector<std::vector<uint32_t>> adj(n1);
    bool any = false;
    for (size_t a = 1; a < n1; ++a)
        for (size_t b = 1; b < n1; ++b)
            if (contact[a * n1 + b]) {
                adj[a].push_back((uint32_t)b);
                adj[b].push_back((uint32_t)a);
                any = true;
            }
    if (!any) return false;
    size_t start = 1;
    while (adj[start].empty()) ++start;
    std::vector<uint8_t> seen(n1);
    std::vector<uint32_t> st{(uint32_t)start};
    seen[start] = 1;
    size_t count = 1;
    while (!st.empty()) {
        uint32_t x = st.back();
        st.pop_back();
        for (uint32_t y : adj[x])
            if (!seen[y]) {
                seen[y] = 1;
                ++count;
                st.push_back(y);
            }
    }
    return count == n_cells;
}

void finish_result(const Model& m, const std::vector<double>& trace, Result& r) {
    r.status_kind = not_equilibrated(trace) ? "NotEquilibrated" : "Ok";
    if (r.status_kind == "Ok")
        for (auto& s : r.samples)
            if (!s.fragmented.empty()) {
                r.status_kind = "Fragmented";
                r.status_has_cell = true;
                r.status_cell = s.fragmented[0];
                r.status_mcs = s.mcs;
                break;
            }
    if (r.status_kind == "Ok" && !r.samples.empty()) {
        size_t ns = r.samples.size();
        size_t window = (size_t)std::ceil((double)ns * 0.5);
        window = std::min(std::max(window, (size_t)1), ns);
        for (size_t i = 0; i < m.n_cells; ++i) {
            bool all = true;
            for (size_t k = ns - window; k < ns && all; ++k) all = r.samples[k].volume[i] <= 1;
            if (all) {
                r.status_kind = "CellLost";
                r.status_has_cell = true;
                r.status_cell = (uint32_t)(i + 1);
                r.status_mcs = r.samples[ns - window].mcs;
                break;
            }
        }
    }
    if (r.status_kind == "Ok" && !r.samples.empty() &&
        fully_connected(r.samples.back().contact, m.n_cells))
        r.status_kind = "Degenerate";
    r.cpm_version = "0.1.0";
    r.convention_hash = convention_hash(m.cfg);
    r.model_hash = model_hash(m);
    r.seed = m.cfg.seed;
    r.seed_policy = "independent";
    uint64_t occ = 0;
    for (uint32_t v : m.volume) occ += v;
    r.derived_phi = (double)(uint32_t)occ / (double)m.n_sites;
    r.cell_type_index = m.type_index;
    r.dim = m.cfg.dim;
    r.n_cells = m.n_cells;
    r.used_default = m.used_default;
}

// ----------------------------------------------------------------------
// Run records (crates/cpm-core/src/config.rs:423-556): serde_json bytes of
// the hashed views, BLAKE3 hex
// ----------------------------------------------------------------------

static void json_str(std::string& o, const std::string& s) {
    static const char* hex = "0123456789abcdef";
    o += '"';
    for (unsigned char ch : s) {
        switch (ch) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            default:
                if (ch < 0x20) {
                    o += "\\u00";
                    o += hex[ch >> 4];
                    o += hex[ch & 15];
                } else {
                    o += (char)ch;
                }
        }
    }
    o += '"';
}

template <class T>
static void json_uints(std::string& o, const std::vector<T>& v) {
    o += '[';
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) o += ',';
        o += std::to_string(v[i]);
    }
    o += ']';
}

static std::string hex32(const uint8_t h[32]) {
    static const char* hx = "0123456789abcdef";
    std::string s;
    for (int i = 0; i < 32; ++i) {
        s += hx[h[i] >> 4];
        s += hx[h[i] & 15];
    }
    return s;
}

std::string convention_hash(const Config& c) {
    std::string o = "{\"dimensionality\":" + std::to_string(c.dim) + ",\"grid\":[";
    for (uint32_t d = 0; d < c.dim; ++d) {
        if (d) o += ',';
        o += std::to_string(c.grid[d]);
    }
    o += "],\"copy_neighborhood\":{\"kind\":\"" + std::string(nb_debug(c.copy_nb)) + "\"}";
    o += ",\"energy_neighborhood\":{\"kind\":\"" + std::string(nb_debug(c.energy_nb)) + "\"}";
    o += ",\"connectivity_neighborhood\":{\"kind\":\"" + std::string(nb_debug(c.conn_nb)) + "\"}";
    o += ",\"boundary\":";
    o += c.boundary == CPM_BOUNDARY_PERIODIC ? "\"Periodic\"" : "\"Fixed\"";
    o += ",\"acceptance\":";
    o += c.acceptance == CPM_ACCEPT_METROPOLIS ? "\"Metropolis\"" : "\"MetropolisHastings\"";
    o += ",\"proposal\":";
    o += c.proposal == CPM_PROPOSAL_UNIFORM ? "\"Uniform\"" : "\"EdgeList\"";
    o += ",\"cell_types\":[";
    for (size_t t = 0; t < c.types.size(); ++t) {
        if (t) o += ',';
        o += "{\"name\":";
        json_str(o, c.types[t].name);
        o += ",\"target_volume\":" + std::to_string(c.types[t].target_volume);
        o += ",\"target_interface\":" + std::to_string(c.types[t].target_interface) + "}";
    }
    o += "],\"active_terms\":{\"volume\":true,\"interface\":true,\"adhesion\":true,\"act\":";
    o += c.act ? "true}" : "false}";
    o += ",\"initialization\":";
    if (c.init_explicit) {
        o += "{\"Explicit\":{\"lattice\":";
        json_uints(o, c.init_lattice);
        o += ",\"cell_type_of\":";
        json_uints(o, c.init_cell_type_of);
        o += "}}";
    } else {
        o += "{\"Default\":{\"warn_on_default_init\":";
        o += c.warn_on_default_init ? "true}}" : "false}}";
    }
    o += ",\"burn_in_mcs\":" + std::to_string(c.burn_in);
    o += ",\"readout_mcs\":" + std::to_string(c.readout);
    o += ",\"sampling_interval_mcs\":" + std::to_string(c.interval);
    o += ",\"min_cell_volume\":1,\"seed_policy\":\"Independent\",\"requested_readouts\":[]";
    o += ",\"convention_version\":1}";
    uint8_t h[32];
    blake3((const uint8_t*)o.data(), o.size(), h);
    return hex32(h);
}

std::string model_hash(const Model& m) {
    std::string o = "{\"convention_hash\":\"" + convention_hash(m.cfg) + "\",\"lattice\":";
    json_uints(o, m.unpadded_lattice());
    o += ",\"cell_type_of\":";
    json_uints(o, m.type_index);
    o += "}";
    uint8_t h[32];
    blake3((const uint8_t*)o.data(), o.size(), h);
    return hex32(h);
}

// ----------------------------------------------------------------------
// theta (crates/cpm-core/src/theta.rs:43-178)
// ----------------------------------------------------------------------

size_t theta_len(size_t n) { return 4 * n + (n + 1) * (n + 2) / 2; }

std::vector<double> theta_extract(const Config& c) {
    const size_t n = c.types.size();
    std::vector<double> th;
    for (auto& t : c.types) th.push_back(t.lambda_volume);
    for (auto& t : c.types) th.push_back(t.lambda_interface);
    for (size_t i = 0; i <= n; ++i)
        for (size_t j = i; j <= n; ++j) th.push_back(c.adhesion[i][j]);
    for (auto& t : c.types) th.push_back(t.lambda_act);
    for (auto& t : c.types) th.push_back((double)t.max_act);
    return th;
}

void theta_inject(Config& c, const double* th, size_t len) {
    const size_t n = c.types.size();
    if (len != theta_len(n))
        throw Error(CPM_ERR_CONFIG, "theta has " + std::to_string(len) + " entries, expected " +
                                        std::to_string(theta_len(n)) + " for " +
                                        std::to_string(n) + " cell types");
    const size_t ma = len - n;
    for (size_t i = 0; i < n; ++i) {
        double v = th[ma + i], r = std::round(v);
        if (!std::isfinite(r) || r < 0.0 || r > 4294967295.0)
            throw Error(CPM_ERR_CONFIG,
                        "max_act[" + c.types[i].name + "] = " + rust_f64(v) +
                            " does not round to a value representable as u32 (must be finite "
                            "and within [0, 4294967295])");
    }
    size_t k = 0;
    for (auto& t : c.types) t.lambda_volume = th[k++];
    for (auto& t : c.types) t.lambda_interface = th[k++];
    for (size_t i = 0; i <= n; ++i)
        for (size_t j = i; j <= n; ++j) {
            c.adhesion[i][j] = th[k];
            c.adhesion[j][i] = th[k];
            ++k;
        }
    for (auto& t : c.types) t.lambda_act = th[k++];
    for (auto& t : c.types) t.max_act = (uint32_t)std::round(th[k++]);
}

std::string theta_name(const Config& c, size_t index) {
    const size_t n = c.types.size();
    auto lbl = [&](size_t i) { return i == 0 ? std::string("medium") : c.types[i - 1].name; };
    if (index < n) return "lambda_volume[" + c.types[index].name + "]";
    index -= n;
    if (index < n) return "lambda_interface[" + c.types[index].name + "]";
    index -= n;
    for (size_t i = 0; i <= n; ++i)
        for (size_t j = i; j <= n; ++j) {
            if (index == 0) return "J[" + lbl(i) + "][" + lbl(j) + "]";
            --index;
        }
    if (index < n) return "lambda_act[" + c.types[index].name + "]";
    index -= n;
    if (index < n) return "max_act[" + c.types[index].name + "]";
    throw Error(CPM_ERR_ARG, "theta index out of range");
}

// ----------------------------------------------------------------------
// BLAKE3: streaming compressor with a chaining-value stack
// ----------------------------------------------------------------------

namespace {
constexpr uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                             0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr int kSchedule[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void mix(uint32_t v[16], int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] += v[b] + x;
    v[d] = ror(v[d] ^ v[a], 16);
    v[c] += v[d];
    v[b] = ror(v[b] ^ v[c], 12);
    v[a] += v[b] + y;
    v[d] = ror(v[d] ^ v[a], 8);
    v[c] += v[d];
    v[b] = ror(v[b] ^ v[c], 7);
}

void compress8(const uint32_t cv[8], const uint8_t block[64], uint32_t len, uint64_t ctr,
               uint32_t flags, uint32_t out[8]) {
    uint32_t m[16];
    for (int i = 0; i < 16; ++i)
        m[i] = (uint32_t)block[4 * i] | (uint32_t)block[4 * i + 1] << 8 |
               (uint32_t)block[4 * i + 2] << 16 | (uint32_t)block[4 * i + 3] << 24;
    uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      kIV[0], kIV[1], kIV[2], kIV[3], (uint32_t)ctr, (uint32_t)(ctr >> 32),
                      len, flags};
    for (int round = 0; round < 7; ++round) {
        mix(v, 0, 4, 8, 12, m[0], m[1]);
        mix(v, 1, 5, 9, 13, m[2], m[3]);
        mix(v, 2, 6, 10, 14, m[4], m[5]);
        mix(v, 3, 7, 11, 15, m[6], m[7]);
        mix(v, 0, 5, 10, 15, m[8], m[9]);
        mix(v, 1, 6, 11, 12, m[10], m[11]);
        mix(v, 2, 7, 8, 13, m[12], m[13]);
        mix(v, 3, 4, 9, 14, m[14], m[15]);
        uint32_t t[16];
        for (int i = 0; i < 16; ++i) t[i] = m[kSchedule[i]];
        std::memcpy(m, t, sizeof m);
    }
    for (int i = 0; i < 8; ++i) out[i] = v[i] ^ v[i + 8];
}
}  // namespace

void blake3(const uint8_t* data, size_t len, uint8_t out[32]) {
    // Each completed chunk's chaining value is pushed; after chunk k+1
    // completes, merge while the count of completed chunks has trailing
    // zero bits (the spec's lazy merge).  The last chunk stays unmerged
    // until finalisation.
    std::vector<std::array<uint32_t, 8>> stack;
    const size_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
    uint32_t last_cv[8];
    uint8_t blk[64];
    for (size_t ci = 0; ci < nchunks; ++ci) {
        const uint8_t* cp = data + ci * 1024;
        size_t clen = std::min<size_t>(1024, len - ci * 1024);
        if (len == 0) clen = 0;
        uint32_t cv[8];
        std::memcpy(cv, kIV, sizeof cv);
        size_t nb = clen == 0 ? 1 : (clen + 63) / 64;
        const bool root_chunk = nchunks == 1;
        for (size_t b = 0; b < nb; ++b) {
            size_t bl = std::min<size_t>(64, clen - b * 64);
            if (clen == 0) bl = 0;
            std::memset(blk, 0, 64);
            if (bl) std::memcpy(blk, cp + b * 64, bl);
            uint32_t fl = (b == 0 ? 1u : 0u) | (b + 1 == nb ? 2u : 0u);
            if (b + 1 == nb && root_chunk) fl |= 8u;
            compress8(cv, blk, (uint32_t)bl, ci, fl, cv);
        }
        if (ci + 1 == nchunks) {
            std::memcpy(last_cv, cv, sizeof cv);
            break;
        }
        std::array<uint32_t, 8> a;
        std::memcpy(a.data(), cv, 32);
        stack.push_back(a);
        uint64_t total = ci + 1;
        while ((total & 1) == 0) {
            auto r = stack.back();
            stack.pop_back();
            auto l = stack.back();
            stack.pop_back();
            std::memcpy(blk, l.data(), 32);
            std::memcpy(blk + 32, r.data(), 32);
            std::array<uint32_t, 8> p;
            compress8(kIV, blk, 64, 0, 4u, p.data());
            stack.push_back(p);
            total >>= 1;
        }
    }
    uint32_t cv[8];
    std::memcpy(cv, last_cv, sizeof cv);
    while (!stack.empty()) {
        auto l = stack.back();
        stack.pop_back();
        std::memcpy(blk, l.data(), 32);
        std::memcpy(blk + 32, cv, 32);
        compress8(kIV, blk, 64, 0, 4u | (stack.empty() ? 8u : 0u), cv);
    }
    for (int i = 0; i < 8; ++i)
        for (int b = 0; b < 4; ++b) out[4 * i + b] = (uint8_t)(cv[i] >> (8 * b));
}

uint64_t derive_seed_u64(uint64_t master, uint64_t index) {
    uint8_t in[16], h[32];
    for (int i = 0; i < 8; ++i) {
        in[i] = (uint8_t)(master >> (8 * i));
        in[8 + i] = (uint8_t)(index >> (8 * i));
    }
    blake3(in, 16, h);
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)h[i] << (8 * i);
    return v;
}

// ----------------------------------------------------------------------
// Simple-point definition (crates/cpm-core/src/simple_point.rs:97-117,
// 242-276): one background component under full adjacency, and exactly one
// face-adjacency object component that touches the origin face-wise.
// ----------------------------------------------------------------------

struct SpAdj {
    int n;
    uint32_t face[26], full[26], origin_face, all;
    explicit SpAdj(int dim) {
        int32_t o[26 * 3];
        n = stencil(dim, dim == 2 ? NB_MOORE : NB_TWENTY_SIX, o);
        origin_face = 0;
        for (int i = 0; i < n; ++i) {
            face[i] = full[i] = 0;
            int manh0 = 0;
            for (int d = 0; d < dim; ++d) manh0 += std::abs(o[i * dim + d]);
            if (manh0 == 1) origin_face |= 1u << i;
            for (int j = 0; j < n; ++j) {
                if (i == j) continue;
                int manh = 0, cheb = 0;
                for (int d = 0; d < dim; ++d) {
                    int df = std::abs(o[i * dim + d] - o[j * dim + d]);
                    manh += df;
                    cheb = std::max(cheb, df);
                }
                if (manh == 1) face[i] |= 1u << j;
                if (cheb == 1) full[i] |= 1u << j;
            }
        }
        all = (uint32_t)((1ull << n) - 1);
    }
    static uint32_t grow(uint32_t seed, uint32_t uni, const uint32_t* adj) {
        uint32_t comp = seed;
        for (;;) {
            uint32_t nxt = comp;
            for (uint32_t f = comp; f; f &= f - 1) nxt |= adj[__builtin_ctz(f)];
            nxt &= uni;
            if (nxt == comp) return comp;
            comp = nxt;
        }
    }
    bool simple(uint32_t obj) const {
        uint32_t bg = ~obj & all;
        if (!bg || grow(bg & -bg, bg, full) != bg) return false;
        int touching = 0;
        for (uint32_t rem = obj; rem;) {
            uint32_t part = grow(rem & -rem, obj, face);
            rem &= ~part;
            touching += (part & origin_face) != 0;
        }
        return touching == 1;
    }
};

bool sp_is_simple(int dim, uint32_t pattern) {
    static const SpAdj a2(2), a3(3);
    return (dim == 2 ? a2 : a3).simple(pattern);
}

void sp_table2d(uint8_t out[256]) {
    for (uint32_t p = 0; p < 256; ++p) out[p] = sp_is_simple(2, p) ? 1 : 0;
}

}  // namespace cpm

// extern "C" boundary (include/cpm_engine.h) and the run drivers that sit
// on top of the device engine: output::run_with (crates/cpm-core/src/
// output.rs:434-529) and batch::run_batch (crates/cpm-core/src/
// batch.rs:110-153).
#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <future>
#include <vector>

#include "../../include/cpm_engine.h"
#include "device.hpp"
#include "driver.hpp"
#include "host.hpp"
#include "binomial.hpp"
#include "metropolis.hpp"
#include "serve.hpp"

using namespace cpm;

struct cpm_engine {
    Engine* e;
    // batches place and grow their default-initialised models on the device
    bool device_init = true;
};
struct cpm_model {
    Model* m;
};
struct cpm_result {
    Result r;
};
struct cpm_server {
    Server* s;
};

namespace {
thread_local std::string g_last_error;

template <class F>
cpm_status guarded(F&& f) {
    try {
        f();
        return CPM_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return CPM_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CPM_ERR_DEVICE;
    }
}

void need(bool ok, const char* what) {
    if (!ok) throw Error(CPM_ERR_ARG, std::string("invalid argument: ") + what);
}

// Drives many initialised models through their run_with schedules in
// lockstep: each launch advances every replica by its own next chunk.
void run_lockstep(Engine* e, std::vector<Model*>& ms, bool include_lattice, int mode,
                  std::vector<Result>& out, cpm_counters* counters) {
    const size_t n = ms.size();
    std::vector<std::vector<Step>> sch(n);
    size_t steps = 0;
    for (size_t i = 0; i < n; ++i) {
        sch[i] = schedule(ms[i]->cfg);
        steps = std::max(steps, sch[i].size());
    }
    std::vector<std::vector<double>> traces(n);
    out.assign(n, Result());
    std::vector<uint64_t> todo(n);
    // the readouts of step k are made on the device right after step k
    // (before step k + 1 changes the state); a second host thread appends
    // them to the results while the device runs step k + 1
    std::future<void> readouts;
    for (size_t k = 0; k < steps; ++k) {
        for (size_t i = 0; i < n; ++i) todo[i] = k < sch[i].size() ? sch[i][k].mcs : 0;
        run_mcs_each(e, ms.data(), n, todo.data(), mode, counters);
        if (readouts.valid()) readouts.get();
        std::vector<size_t> sampled;
        std::vector<uint64_t> at_mcs;
        std::vector<double> at_energy;
        for (size_t i = 0; i < n; ++i) {
            if (k >= sch[i].size()) continue;
            if (sch[i][k].trace) traces[i].push_back(ms[i]->energy);
            if (sch[i][k].sample) {
                sampled.push_back(i);
                at_mcs.push_back(ms[i]->mcs);
                at_energy.push_back(ms[i]->energy);
            }
        }
        // readouts on the device (labelling, shape); the samples are
        // appended on a second host thread while the device runs step k + 1
        std::vector<Model*> sm(sampled.size());
        for (size_t j = 0; j < sampled.size(); ++j) sm[j] = ms[sampled[j]];
        std::vector<Sample> smp(sampled.size());
        device_samples(e, sm.data(), sm.size(), include_lattice, at_mcs.data(), at_energy.data(),
                       smp.data());
        readouts = std::async(std::launch::async, [&out, sampled = std::move(sampled),
                                                   smp = std::move(smp)]() mutable {
            for (size_t j = 0; j < sampled.size(); ++j)
                out[sampled[j]].samples.push_back(std::move(smp[j]));
        });
    }
    if (readouts.valid()) readouts.get();
    parallel_for(n, [&](size_t i) {
        sync_to_host(ms[i]);
        finish_result(*ms[i], traces[i], out[i]);
    });
}
}  // namespace

extern "C" {

const char* cpm_last_error(void) { return g_last_error.c_str(); }
const char* cpm_version(void) { return "0.1.0"; }

cpm_status cpm_engine_create(int device, cpm_engine** out) {
    return guarded([&] {
        need(out, "out");
        *out = new cpm_engine{engine_create(device)};
    });
}

void cpm_engine_destroy(cpm_engine* engine) {
    if (!engine) return;
    engine_destroy(engine->e);
    delete engine;
}

cpm_status cpm_model_create(const cpm_config* config, cpm_model** out) {
    return guarded([&] {
        need(config && out, "config/out");
        Config c = resolve(*config);
        *out = new cpm_model{new Model(c)};
    });
}

void cpm_model_destroy(cpm_model* model) {
    if (!model) return;
    delete model->m;
    delete model;
}

cpm_status cpm_model_initialize(cpm_model* model, int32_t* used_default) {
    return guarded([&] {
        need(model, "model");
        device_release(model->m);
        model->m->initialize();
        if (used_default) *used_default = model->m->used_default ? 1 : 0;
    });
}

cpm_status cpm_model_initialize_batch(cpm_engine* engine, cpm_model* const* models, size_t n,
                                      int32_t* used_default) {
    return guarded([&] {
        need(engine && (models || n == 0), "engine/models");
        std::vector<Model*> ms(n);
        for (size_t i = 0; i < n; ++i) {
            need(models[i], "models[i]");
            ms[i] = models[i]->m;
            device_release(ms[i]);
        }
        if (engine->device_init)
            device_initialize(engine->e, ms.data(), n);
        else
            parallel_for(n, [&](size_t i) { ms[i]->initialize(); });
        if (used_default)
            for (size_t i = 0; i < n; ++i) used_default[i] = ms[i]->used_default ? 1 : 0;
    });
}

cpm_status cpm_run_mcs(cpm_engine* engine, cpm_model* const* models, size_t n, uint64_t n_mcs,
                       int32_t mode, cpm_counters* counters) {
    return guarded([&] {
        need(engine && (models || n == 0), "engine/models");
        std::vector<Model*> ms(n);
        for (size_t i = 0; i < n; ++i) {
            need(models[i], "models[i]");
            ms[i] = models[i]->m;
        }
        if (counters) *counters = cpm_counters{};
        if (n) run_mcs(engine->e, ms.data(), n, n_mcs, mode, counters);
    });
}

cpm_status cpm_run_mcs_noalloc(cpm_engine* engine, cpm_model* const* models, size_t n,
                               uint64_t n_mcs, cpm_counters* counters) {
    return guarded([&] {
        need(engine && (models || n == 0), "engine/models");
        if (n > CPM_NOALLOC_MAX_REPLICAS)
            throw Error(CPM_ERR_ARG, "cpm_run_mcs_noalloc: " + std::to_string(n) +
                                         " models exceed the maximum of " +
                                         std::to_string(CPM_NOALLOC_MAX_REPLICAS) + " per call");
        Model* stack_models[64];
        std::vector<Model*> heap_models;
        Model** ms = stack_models;
        if (n > 64) {
            heap_models.resize(n);
            ms = heap_models.data();
        }
        for (size_t i = 0; i < n; ++i) {
            need(models[i], "models[i]");
            ms[i] = models[i]->m;
        }
        if (counters) *counters = cpm_counters{};
        if (n) run_mcs_noalloc(engine->e, ms, n, n_mcs, counters);
    });
}

cpm_status cpm_run_mcs_plain(cpm_engine* engine, cpm_model* model, uint64_t n_mcs,
                             cpm_counters* counters) {
    return guarded([&] {
        need(engine && model, "engine/model");
        if (counters) *counters = cpm_counters{};
        run_mcs_plain(engine->e, model->m, n_mcs, counters);
    });
}

cpm_status cpm_model_get_info(cpm_model* model, cpm_model_info* out) {
    return guarded([&] {
        need(model && out, "model/out");
        Model* m = model->m;
        out->n_sites = m->n_sites;
        out->n_cells = m->n_cells;
        out->mcs = m->mcs;
        out->conservative_energy = m->energy;
        for (int k = 0; k < 4; ++k) out->rng_state[k] = m->rng.s[k];
    });
}

cpm_status cpm_model_get_state(cpm_model* model, uint32_t* lattice, uint32_t* volume,
                               uint32_t* interface, uint32_t* type_index,
                               int64_t* sum_position, int64_t* sum_outer, uint32_t* contact) {
    return guarded([&] {
        need(model, "model");
        Model* m = model->m;
        sync_to_host(m);
        const size_t n = m->n_cells, D = m->D;
        if (lattice) {
            auto l = m->unpadded_lattice();
            std::memcpy(lattice, l.data(), l.size() * 4);
        }
        if (volume) std::memcpy(volume, m->volume.data(), n * 4);
        if (interface) std::memcpy(interface, m->interface.data(), n * 4);
        if (type_index) std::memcpy(type_index, m->type_index.data(), n * 4);
        if (sum_position) std::memcpy(sum_position, m->sum_pos.data(), n * D * 8);
        if (sum_outer) std::memcpy(sum_outer, m->sum_outer.data(), n * D * D * 8);
        if (contact) std::memcpy(contact, m->contact.data(), (n + 1) * (n + 1) * 4);
    });
}

cpm_status cpm_model_birth_mcs(cpm_model* model, uint32_t* out, size_t n_sites) {
    return guarded([&] {
        need(model && out, "model/out");
        Model* m = model->m;
        need(n_sites == m->n_sites, "n_sites");
        sync_to_host(m);
        if (!m->cfg.act) {
            std::memset(out, 0, n_sites * 4);
            return;
        }
        auto b = m->unpadded(m->birth.data());
        std::memcpy(out, b.data(), b.size() * 4);
    });
}

cpm_status cpm_model_edge_sites(cpm_model* model, uint64_t* out, size_t cap, size_t* n) {
    return guarded([&] {
        need(model && n && (cap == 0 || out), "model/out/n");
        Model* m = model->m;
        sync_to_host(m);
        *n = (size_t)m->n_edge;
        for (size_t i = 0; i < m->n_edge && i < cap; ++i) {
            uint64_t c[3];
            m->interior_coord(m->edge_sites[i], c);
            uint64_t r = 0;
            for (int d = 0; d < m->D; ++d) r = r * m->dims[d] + c[d];
            out[i] = r;
        }
    });
}

cpm_status cpm_model_hashes(cpm_model* model, char convention_hex[65], char model_hex[65]) {
    return guarded([&] {
        need(model && convention_hex && model_hex, "model/out");
        sync_to_host(model->m);
        std::string c = convention_hash(model->m->cfg), m = model_hash(*model->m);
        std::memcpy(convention_hex, c.c_str(), 65);
        std::memcpy(model_hex, m.c_str(), 65);
    });
}

cpm_status cpm_model_global_energy(cpm_model* model, double* out) {
    return guarded([&] {
        need(model && out, "model/out");
        sync_to_host(model->m);
        *out = model->m->global_energy();
    });
}

cpm_status cpm_model_set_state(cpm_model* model, const cpm_model_state* state) {
    return guarded([&] {
        need(model && state, "model/state");
        // validated before anything changes; the device copy (if any) holds
        // the previous state and is dropped only once the new one is in
        Model* m = model->m;
        Model next(m->cfg);
        next.set_state(*state);
        device_release(m);
        m->n_cells = next.n_cells;
        m->sigma.swap(next.sigma);
        m->type_index.swap(next.type_index);
        m->volume.swap(next.volume);
        m->interface.swap(next.interface);
        m->sum_pos.swap(next.sum_pos);
        m->sum_outer.swap(next.sum_outer);
        m->contact.swap(next.contact);
        m->birth.swap(next.birth);
        m->edge_sites.swap(next.edge_sites);
        m->edge_pos.swap(next.edge_pos);
        m->n_edge = next.n_edge;
        m->energy = next.energy;
        m->rng = state->rng_state ? next.rng : m->rng;
        m->mcs = next.mcs;
        m->attempt_in_mcs = 0;
        m->has_sublattice_key = false;
        m->attempt_rate = 0.0;
        m->initialized = true;
        m->used_default = false;
        m->device_dirty = false;
    });
}

cpm_status cpm_model_sample(cpm_model* model, int32_t include_lattice, cpm_result** out) {
    return guarded([&] {
        need(model && out, "model/out");
        Model* m = model->m;
        need(m->initialized, "model must be initialized before cpm_model_sample");
        Result r;
        r.samples.emplace_back();
        if (Engine* e = device_engine(m)) {  // resident: readouts on the device
            const uint64_t mcs = m->mcs;
            const double energy = m->energy;
            device_samples(e, &m, 1, include_lattice != 0, &mcs, &energy, &r.samples[0]);
        } else {
            r.samples[0] = take_sample(*m, include_lattice != 0);
        }
        sync_to_host(m);
        finish_result(*m, {}, r);
        *out = new cpm_result{std::move(r)};
    });
}

cpm_status cpm_not_equilibrated(const double* trace, size_t n, int32_t* out) {
    return guarded([&] {
        need(out && (trace || n == 0), "trace/out");
        *out = not_equilibrated(std::vector<double>(trace, trace + n)) ? 1 : 0;
    });
}

cpm_status cpm_run(cpm_engine* engine, cpm_model* model, int32_t include_lattice, int32_t mode,
                   cpm_result** out) {
    return guarded([&] {
        need(engine && model && out, "engine/model/out");
        need(model->m->initialized, "model must be initialized before cpm_run");
        std::vector<Model*> ms{model->m};
        std::vector<Result> rs;
        run_lockstep(engine->e, ms, include_lattice != 0, mode, rs, nullptr);
        *out = new cpm_result{std::move(rs[0])};
    });
}

static void batch_models(cpm_engine* engine, std::vector<Config>& cfgs, int32_t include_lattice,
                         int32_t mode, cpm_result** out, int32_t* any_used_default,
                         cpm_counters* counters) {
    const size_t n = cfgs.size();
    std::vector<std::unique_ptr<Model>> owned(n);
    std::vector<Model*> ms(n);
    // CPM::new + initialize per replica (batch.rs:126-145): construction
    // over the host cores, scatter_and_grow on the device (one wave per
    // model) unless it is switched off
    parallel_for(n, [&](size_t i) {
        owned[i] = std::make_unique<Model>(cfgs[i]);
        ms[i] = owned[i].get();
        if (!engine->device_init) owned[i]->initialize();
    });
    if (engine->device_init) device_initialize(engine->e, ms.data(), n);
    bool any = false;
    for (size_t i = 0; i < n; ++i) any |= ms[i]->used_default;
    if (counters) *counters = cpm_counters{};
    std::vector<Result> rs;
    run_lockstep(engine->e, ms, include_lattice != 0, mode, rs, counters);
    for (size_t i = 0; i < n; ++i) out[i] = new cpm_result{std::move(rs[i])};
    if (any_used_default) *any_used_default = any ? 1 : 0;
}

cpm_status cpm_run_batch(cpm_engine* engine, const cpm_config* base, const double* thetas,
                         size_t n_thetas, size_t theta_len_, uint64_t master_seed,
                         int32_t include_lattice, int32_t mode, cpm_result** out,
                         int32_t* any_used_default) {
    return guarded([&] {
        need(engine && base && out && (thetas || n_thetas == 0), "engine/base/out/thetas");
        Config c = resolve(*base);
        const size_t expected = theta_len(c.types.size());
        // batch.rs:115-124 — every theta length checked before any work
        if (theta_len_ != expected)
            throw Error(CPM_ERR_CONFIG, "theta at index 0 has length " + std::to_string(theta_len_) +
                                            ", expected " + std::to_string(expected));
        std::vector<Config> cfgs(n_thetas, c);
        for (size_t i = 0; i < n_thetas; ++i) {
            theta_inject(cfgs[i], thetas + i * theta_len_, theta_len_);
            cfgs[i].seed = derive_seed_u64(master_seed, i);
            validate(cfgs[i]);
        }
        batch_models(engine, cfgs, include_lattice, mode, out, any_used_default, nullptr);
    });
}

cpm_status cpm_run_batch_configs(cpm_engine* engine, const cpm_config* const* configs, size_t n,
                                 int32_t include_lattice, int32_t mode, cpm_result** out,
                                 int32_t* any_used_default, cpm_counters* counters) {
    return guarded([&] {
        need(engine && out && (configs || n == 0), "engine/configs/out");
        std::vector<Config> cfgs;
        cfgs.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            need(configs[i], "configs[i]");
            cfgs.push_back(resolve(*configs[i]));
            validate(cfgs.back());
        }
        batch_models(engine, cfgs, include_lattice, mode, out, any_used_default, counters);
    });
}

cpm_status cpm_result_get_info(const cpm_result* result, cpm_result_info* o) {
    return guarded([&] {
        need(result && o, "result/out");
        const Result& r = result->r;
        o->cpm_version = r.cpm_version.c_str();
        o->convention_version = 1;
        o->convention_hash = r.convention_hash.c_str();
        o->model_hash = r.model_hash.c_str();
        o->seed = r.seed;
        o->seed_policy = r.seed_policy.c_str();
        o->derived_phi = r.derived_phi;
        o->theta_layout_version = 1;
        o->status_kind = r.status_kind.c_str();
        o->status_has_cell = r.status_has_cell ? 1 : 0;
        o->status_cell_id = r.status_cell;
        o->status_at_mcs = r.status_mcs;
        o->n_samples = r.samples.size();
        o->n_cells = r.n_cells;
        o->dim = r.dim;
        o->cell_type_index = r.cell_type_index.data();
        o->used_default = r.used_default ? 1 : 0;
    });
}

cpm_status cpm_result_get_sample(const cpm_result* result, uint64_t index, cpm_sample_view* o) {
    return guarded([&] {
        need(result && o, "result/out");
        need(index < result->r.samples.size(), "sample index");
        const Sample& s = result->r.samples[index];
        o->mcs = s.mcs;
        o->conservative_energy = s.energy;
        o->n_cells = s.ids.size();
        o->ids = s.ids.data();
        o->volume = s.volume.data();
        o->interface = s.interface.data();
        o->centroid = s.centroid.data();
        o->eigenvalues = s.eigen.data();
        o->radius_of_gyration_sq = s.rg2.data();
        o->kappa_sq = s.kappa.data();
        o->contact = s.contact.data();
        o->n_fragmented = s.fragmented.size();
        o->fragmented_cells = s.fragmented.data();
        o->n_medium_components = s.medium_volumes.size();
        o->medium_component_volumes = s.medium_volumes.data();
        o->medium_component_touches_boundary = s.medium_touch.data();
        o->lattice = s.has_lattice ? s.lattice.data() : nullptr;
    });
}

cpm_status cpm_result_get_samples_shape(const cpm_result* result, cpm_samples_shape* o) {
    return guarded([&] {
        need(result && o, "result/out");
        const Result& r = result->r;
        *o = cpm_samples_shape{};
        o->n_samples = r.samples.size();
        o->n_cells = r.samples.empty() ? r.n_cells : r.samples[0].ids.size();
        o->dim = r.dim;
        const bool lat = !r.samples.empty() && r.samples[0].has_lattice;
        o->has_lattice = lat ? 1 : 0;
        o->lattice_sites = lat ? r.samples[0].lattice.size() : 0;
        for (const Sample& s : r.samples) {
            o->n_fragmented_total += s.fragmented.size();
            o->n_medium_total += s.medium_volumes.size();
        }
    });
}

cpm_status cpm_result_copy_samples(const cpm_result* result, const cpm_samples_buffers* o) {
    return guarded([&] {
        need(result && o, "result/out");
        const Result& r = result->r;
        const size_t ns = r.samples.size();
        if (!ns) {
            if (o->fragmented_offsets) o->fragmented_offsets[0] = 0;
            if (o->medium_offsets) o->medium_offsets[0] = 0;
            return;
        }
        const size_t n = r.samples[0].ids.size(), D = r.dim, n1 = n + 1;
        const bool lat = r.samples[0].has_lattice;
        const size_t sites = lat ? r.samples[0].lattice.size() : 0;
        need(o->mcs && o->conservative_energy && o->volume && o->interface && o->centroid &&
                 o->eigenvalues && o->radius_of_gyration_sq && o->kappa_sq && o->contact &&
                 o->fragmented_offsets && o->medium_offsets && (!lat || o->lattice),
             "sample buffers");
        uint64_t fo = 0, mo = 0;
        for (size_t k = 0; k < ns; ++k) {
            const Sample& s = r.samples[k];
            need(s.ids.size() == n && s.lattice.size() == sites, "samples of one shape");
            o->mcs[k] = s.mcs;
            o->conservative_energy[k] = s.energy;
            o->fragmented_offsets[k] = fo;
            o->medium_offsets[k] = mo;
            fo += s.fragmented.size();
            mo += s.medium_volumes.size();
        }
        o->fragmented_offsets[ns] = fo;
        o->medium_offsets[ns] = mo;
        need((fo == 0 || o->fragmented_cells) &&
                 (mo == 0 || (o->medium_component_volumes && o->medium_component_touches_boundary)),
             "list buffers");
        // the per-sample arrays, on several host cores when there is much to
        // copy (a sample holds a dense contact matrix)
        auto copy_one = [&](size_t k) {
            const Sample& s = r.samples[k];
            std::copy(s.volume.begin(), s.volume.end(), o->volume + k * n);
            std::copy(s.interface.begin(), s.interface.end(), o->interface + k * n);
            std::copy(s.centroid.begin(), s.centroid.end(), o->centroid + k * n * D);
            std::copy(s.eigen.begin(), s.eigen.end(), o->eigenvalues + k * n * D);
            std::copy(s.rg2.begin(), s.rg2.end(), o->radius_of_gyration_sq + k * n);
            std::copy(s.kappa.begin(), s.kappa.end(), o->kappa_sq + k * n);
            std::copy(s.contact.begin(), s.contact.end(), o->contact + k * n1 * n1);
            std::copy(s.fragmented.begin(), s.fragmented.end(),
                      o->fragmented_cells + o->fragmented_offsets[k]);
            std::copy(s.medium_volumes.begin(), s.medium_volumes.end(),
                      o->medium_component_volumes + o->medium_offsets[k]);
            std::copy(s.medium_touch.begin(), s.medium_touch.end(),
                      o->medium_component_touches_boundary + o->medium_offsets[k]);
            if (lat) std::copy(s.lattice.begin(), s.lattice.end(), o->lattice + k * sites);
        };
        if ((double)ns * (double)(n1 * n1 * 4 + sites * 4) > (double)(16u << 20))
            parallel_for(ns, copy_one);
        else
            for (size_t k = 0; k < ns; ++k) copy_one(k);
    });
}

void cpm_result_destroy(cpm_result* result) { delete result; }

size_t cpm_theta_len(uint32_t n_types) { return theta_len(n_types); }

cpm_status cpm_theta_extract(const cpm_config* config, double* out, size_t cap) {
    return guarded([&] {
        need(config && out, "config/out");
        Config c = resolve(*config);
        validate(c);
        auto th = theta_extract(c);
        need(cap >= th.size(), "capacity");
        std::copy(th.begin(), th.end(), out);
    });
}

cpm_status cpm_theta_name(const cpm_config* config, size_t index, char* buf, size_t cap) {
    return guarded([&] {
        need(config && buf && cap, "config/buf");
        Config c = resolve(*config);
        validate(c);
        std::string s = theta_name(c, index);
        need(s.size() < cap, "capacity");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

uint64_t cpm_derive_seed_u64(uint64_t master_seed, uint64_t index) {
    return derive_seed_u64(master_seed, index);
}

cpm_status cpm_engine_sp_tables(cpm_engine* engine, uint8_t* table2d, uint64_t* table3d_words) {
    return guarded([&] {
        need(engine, "engine");
        engine_sp_tables(engine->e, table2d, table3d_words);
    });
}

cpm_status cpm_engine_set_sigma_placement(cpm_engine* engine, int32_t order, double share) {
    return guarded([&] {
        need(engine, "engine");
        engine_set_placement(engine->e, order, share);
    });
}

cpm_status cpm_engine_sigma_placement(cpm_engine* engine, int32_t* order, double* share) {
    return guarded([&] {
        need(engine && order && share, "engine/order/share");
        int o = 0;
        engine_last_placement(engine->e, &o, share);
        *order = o;
    });
}

cpm_status cpm_metropolis_check(cpm_engine* engine, const double* delta_h, const double* u,
                                size_t n, double* exp_libm, double* exp_cr, uint8_t* accept) {
    return guarded([&] {
        need((delta_h && exp_libm && exp_cr && (!u || accept)) || n == 0,
             "delta_h/exp_libm/exp_cr/accept");
        if (engine) {
            engine_metropolis_check(engine->e, delta_h, u, n, exp_libm, exp_cr, accept);
            return;
        }
        for (size_t i = 0; i < n; ++i) {
            exp_libm[i] = std::exp(-delta_h[i]);
            exp_cr[i] = cpm::exp_cr(-delta_h[i]);
            if (u) {
                bool tie = false;
                accept[i] = metropolis_accept(u[i], delta_h[i], &tie) ? 1 : 0;
            }
        }
    });
}

cpm_status cpm_act_root_check(cpm_engine* engine, const double* product, const uint32_t* count,
                              size_t n, double* out) {
    return guarded([&] {
        need((product && count && out) || n == 0, "product/count/out");
        for (size_t i = 0; i < n; ++i) need(count[i] >= 1, "count >= 1");
        if (engine) {
            engine_act_root_check(engine->e, product, count, n, out);
            return;
        }
        for (size_t i = 0; i < n; ++i) out[i] = act_root(product[i], count[i]);
    });
}

cpm_status cpm_mh_check(cpm_engine* engine, const uint32_t* n_a, const uint32_t* n_b,
                        const double* delta_h, const double* u, size_t n, uint8_t* accept) {
    return guarded([&] {
        need((n_a && n_b && delta_h && u && accept) || n == 0, "n_a/n_b/delta_h/u/accept");
        for (size_t i = 0; i < n; ++i) need(n_a[i] >= 1 && n_b[i] >= 1, "n_a, n_b >= 1");
        if (engine) {
            engine_mh_check(engine->e, n_a, n_b, delta_h, u, n, accept);
            return;
        }
        for (size_t i = 0; i < n; ++i) {
            bool tie = false;
            accept[i] = mh_accept(u[i], n_a[i], n_b[i], delta_h[i], &tie) ? 1 : 0;
        }
    });
}

cpm_status cpm_binomial_check(cpm_engine* engine, uint64_t* rng_state, const uint64_t* n,
                              const double* p, size_t count, uint64_t* k) {
    return guarded([&] {
        need((rng_state && n && p && k) || count == 0, "rng_state/n/p/k");
        for (size_t i = 0; i < count; ++i) need(p[i] >= 0.0 && p[i] <= 1.0, "0 <= p <= 1");
        if (engine) {
            engine_binomial_check(engine->e, rng_state, n, p, count, k);
            return;
        }
        for (size_t i = 0; i < count; ++i) {
            Xoshiro r;
            for (int j = 0; j < 4; ++j) r.s[j] = rng_state[4 * i + j];
            if (!binomial_sample(r, n[i], p[i], &k[i]))
                throw Error(CPM_ERR_DEVICE, "the binomial sampler ran away");
            for (int j = 0; j < 4; ++j) rng_state[4 * i + j] = r.s[j];
        }
    });
}

cpm_status cpm_engine_set_pipelining(cpm_engine* engine, int32_t on) {
    return guarded([&] {
        need(engine, "engine");
        engine_set_pipelining(engine->e, on != 0);
    });
}

cpm_status cpm_engine_set_device_init(cpm_engine* engine, int32_t on) {
    return guarded([&] {
        need(engine, "engine");
        engine->device_init = on != 0;
    });
}

cpm_status cpm_model_chain_rate(cpm_model* model, double* attempts_per_s) {
    return guarded([&] {
        need(model && attempts_per_s, "model/attempts_per_s");
        *attempts_per_s = model->m->attempt_rate;
    });
}

cpm_status cpm_engine_kernel_launches(cpm_engine* engine, int32_t* keys, uint64_t* launches,
                                      size_t cap, size_t* n) {
    return guarded([&] {
        need(engine && n && (cap == 0 || (keys && launches)), "engine/keys/launches/n");
        auto v = engine_launch_counts(engine->e);
        *n = v.size();
        for (size_t i = 0; i < v.size() && i < cap; ++i) {
            keys[i] = v[i].first;
            launches[i] = v[i].second;
        }
    });
}


cpm_status cpm_serve_create(cpm_engine* engine, const cpm_server_options* options,
                            cpm_server** out) {
    return guarded([&] {
        need(engine && out, "engine/out");
        ServeOptions o;
        if (options) {
            o.hbm_budget = options->hbm_budget_bytes;
            o.max_active = options->max_active;
            o.quantum = options->quantum_attempts;
            o.round_s = options->round_ms * 1e-3;
        }
        *out = new cpm_server{serve_create(engine->e, o)};
    });
}

void cpm_serve_destroy(cpm_server* server) {
    if (!server) return;
    serve_destroy(server->s);
    delete server;
}

cpm_status cpm_serve_submit(cpm_server* server, const cpm_config* config,
                            int32_t include_lattice, int32_t mode, uint64_t* request_id,
                            int32_t* used_default) {
    return guarded([&] {
        need(server && config && request_id, "server/config/request_id");
        auto m = std::make_unique<Model>(resolve(*config));
        m->initialize();
        if (used_default) *used_default = m->used_default ? 1 : 0;
        *request_id = serve_submit(server->s, m.release(), include_lattice != 0, mode);
    });
}

cpm_status cpm_serve_poll(cpm_server* server, uint64_t request_id, double timeout_s,
                          cpm_request_info* info, cpm_result** result) {
    return guarded([&] {
        need(server, "server");
        if (result) *result = nullptr;
        RequestInfo ri;
        Result r;
        const bool done = serve_poll(server->s, request_id, timeout_s, ri, &r);
        if (info) {
            info->state = ri.state;
            info->submit_s = ri.submit_s;
            info->admit_s = ri.admit_s;
            info->done_s = ri.done_s;
            info->footprint_bytes = ri.footprint;
            info->mcs_done = ri.mcs_done;
            info->mcs_total = ri.mcs_total;
        }
        if (done && result) *result = new cpm_result{std::move(r)};
    });
}

cpm_status cpm_serve_stats(cpm_server* server, cpm_server_stats* out) {
    return guarded([&] {
        need(server && out, "server/out");
        const ServerStats st = serve_stats(server->s);
        out->submitted = st.submitted;
        out->completed = st.completed;
        out->failed = st.failed;
        out->queued = st.queued;
        out->active = st.active;
        out->hbm_budget_bytes = st.hbm_budget;
        out->hbm_in_use_bytes = st.hbm_in_use;
        out->rounds = st.rounds;
        out->attempts = st.attempts;
        out->busy_s = st.busy_s;
        out->device_s = st.device_s;
        out->uptime_s = st.uptime_s;
    });
}
}  // extern "C"

// Device engine: HIP kernels for gfx950 and their host-side launch logic.
//
// Exact replay (cpm_run_mcs, CPM_MODE_EXACT_REPLAY) keeps every replica's
// state resident in HBM and advances one replica per lane: the lane runs
// the reference's serial attempt loop (crates/cpm-core/src/
// monte_carlo.rs:37-145, 327-357) against its own haloed lattice, per-cell
// bookkeeping and Xoshiro256++ stream, so its results are bit-identical to
// the reference's serial loop.  All FP64 energy arithmetic is compiled
// with -ffp-contract=off and follows the reference's operation order
// (crates/cpm-core/src/energy.rs:92-103, 167-202, 242-264, 521-523).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "core.hpp"
#include "device.hpp"
#include "host.hpp"

namespace cpm {

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            throw Error(_e == hipErrorOutOfMemory ? CPM_ERR_OOM : CPM_ERR_DEVICE,      \
                        std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

// ---------------------------------------------------------------------------
// Replica descriptor: geometry, stencils, parameters, state pointers and the
// scalar state a launch reads and writes back.
// ---------------------------------------------------------------------------
struct Rep {
    uint32_t D, boundary, n_types, n_cells;
    uint32_t dims[3], strides[3];
    uint32_t n_sites, n_copy, n_energy, n_topo;
    int32_t copy_off[6], energy_off[26], topo_off[26];
    uint32_t* sigma;
    uint32_t* type_index;
    uint32_t* volume;
    uint32_t* interface;
    int64_t* sum_pos;
    int64_t* sum_outer;
    uint32_t* contact;
    const double* params;     // lambda_v[T], lambda_i[T], J[(T+1)^2]
    const uint32_t* targets;  // V*[T], I*[T]
    uint64_t rng[4];
    double energy;
    uint64_t mcs;
    uint64_t todo;            // MCS to run in this launch
    unsigned long long cnt[6];
};

__constant__ uint32_t c_sp2d[8];  // 256-bit 2D simple-point table

// ---------------------------------------------------------------------------
// One replica's serial chain.  D is a template parameter so coordinates,
// strides and offsets live in registers.
// ---------------------------------------------------------------------------
template <int D>
__device__ void replica_chain(Rep& R, const uint64_t* __restrict__ sp3d) {
    Xoshiro rng;
    for (int i = 0; i < 4; ++i) rng.s[i] = R.rng[i];
    double energy = R.energy;
    unsigned long long cnt[6] = {0, 0, 0, 0, 0, 0};
    uint32_t dims[D], strides[D];
    uint64_t zone[D];
    for (int d = 0; d < D; ++d) {
        dims[d] = R.dims[d];
        strides[d] = R.strides[d];
        zone[d] = range_zone(dims[d]);
    }
    const uint32_t n_copy = R.n_copy, n_energy = R.n_energy, n_topo = R.n_topo;
    const uint64_t zone_k = range_zone(n_copy);
    uint32_t* __restrict__ sig = R.sigma;
    uint32_t* __restrict__ vol = R.volume;
    uint32_t* __restrict__ itf = R.interface;
    const uint32_t* __restrict__ typ = R.type_index;
    const uint32_t T = R.n_types, T1 = T + 1;
    const double* __restrict__ lv = R.params;
    const double* __restrict__ li = R.params + T;
    const double* __restrict__ J = R.params + 2 * T;
    const uint32_t* __restrict__ tv = R.targets;
    const uint32_t* __restrict__ ti = R.targets + T;
    const uint64_t ncell1 = (uint64_t)R.n_cells + 1;
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;

    for (uint64_t m = 0; m < R.todo; ++m) {
        for (uint32_t a = 0; a < R.n_sites; ++a) {
            // monte_carlo.rs:37-42 — target coordinates, axis 0 first
            uint32_t c[D];
            uint32_t t = 0;
            for (int d = 0; d < D; ++d) {
                uint64_t v;
                while (!range_trial(xo_next(rng), dims[d], zone[d], &v)) {
                }
                c[d] = (uint32_t)v;
                t += (c[d] + 1) * strides[d];
            }
            // monte_carlo.rs:57 — copy source
            uint64_t k;
            while (!range_trial(xo_next(rng), n_copy, zone_k, &k)) {
            }
            const uint32_t s = (uint32_t)((int32_t)t + R.copy_off[k]);
            cnt[0]++;
            const uint32_t L = sig[t], G = sig[s];
            if (L == G) continue;  // early-out
            cnt[1]++;
            if (L != 0) {
                // connectivity.rs:42-65 — losing-cell pattern over the full shell
                uint32_t pat = 0;
                for (uint32_t i = 0; i < n_topo; ++i)
                    pat |= (uint32_t)(sig[(int32_t)t + R.topo_off[i]] == L) << i;
                cnt[2]++;
                const bool simple = D == 2 ? ((c_sp2d[pat >> 5] >> (pat & 31)) & 1u)
                                           : ((sp3d[pat >> 6] >> (pat & 63)) & 1ull);
                if (!simple) continue;
                if (vol[L - 1] <= 1) continue;  // min_cell_volume
            }
            cnt[3]++;
            // energy.rs:521-523 — (volume + interface) + adhesion
            const uint32_t tl = L ? typ[L - 1] : 0, tg = G ? typ[G - 1] : 0;
            double dvol = 0.0;
            if (L) {
                const double v = (double)vol[L - 1], T0 = (double)tv[tl];
                const double x1 = (v - 1.0) - T0, x0 = v - T0;
                dvol += lv[tl] * (x1 * x1) - lv[tl] * (x0 * x0);
            }
            if (G) {
                const double v = (double)vol[G - 1], T0 = (double)tv[tg];
                const double x1 = (v + 1.0) - T0, x0 = v - T0;
                dvol += lv[tg] * (x1 * x1) - lv[tg] * (x0 * x0);
            }
            double dl = 0.0, dg = 0.0, dadh = 0.0;
            const uint32_t li_idx = L ? tl + 1 : 0, gi_idx = G ? tg + 1 : 0;
            for (uint32_t q = 0; q < n_energy; ++q) {
                const uint32_t nb = sig[(int32_t)t + R.energy_off[q]];
                dl += nb == L ? 1.0 : -1.0;
                dg += nb == G ? -1.0 : 1.0;
                const uint32_t ni = nb ? typ[nb - 1] + 1 : 0;
                const double before = nb != L ? J[li_idx * T1 + ni] : 0.0;
                const double after = nb != G ? J[gi_idx * T1 + ni] : 0.0;
                dadh += 1.0 * (after - before);
            }
            double ditf = 0.0;
            if (L) {
                const double i0 = (double)itf[L - 1], T0 = (double)ti[tl];
                const double x1 = (i0 + dl) - T0, x0 = i0 - T0;
                ditf += li[tl] * (x1 * x1) - li[tl] * (x0 * x0);
            }
            if (G) {
                const double i0 = (double)itf[G - 1], T0 = (double)ti[tg];
                const double x1 = (i0 + dg) - T0, x0 = i0 - T0;
                ditf += li[tg] * (x1 * x1) - li[tg] * (x0 * x0);
            }
            const double dh = dvol + ditf + dadh;
            // monte_carlo.rs:158 — Metropolis at T = 1
            bool accept = dh <= 0.0;
            if (!accept) {
                const double u = gen_f64(rng);
                const double p = exp(-dh);
                accept = u < p;
                if (fabs(u - p) <= p * 0x1p-50) cnt[5]++;
            }
            if (!accept) continue;
            cnt[4]++;
            // dynamics.rs:73-98 — contact counts on the pre-move lattice
            for (uint32_t q = 0; q < n_energy; ++q) {
                const uint32_t nb = sig[(int32_t)t + R.energy_off[q]];
                if (nb != L) {
                    const uint64_t key = (uint64_t)min(L, nb) * ncell1 + max(L, nb);
                    const uint32_t cur = R.contact[key];
                    R.contact[key] = cur ? cur - 1 : 0;
                }
                if (nb != G) {
                    const uint64_t key = (uint64_t)min(G, nb) * ncell1 + max(G, nb);
                    R.contact[key] += 1;
                }
            }
            sig[t] = G;
            // dynamics.rs:44-53, 106-110
            if (L) {
                vol[L - 1] -= 1;
                itf[L - 1] = (uint32_t)((int64_t)itf[L - 1] + (int64_t)round(dl));
            }
            if (G) {
                vol[G - 1] += 1;
                itf[G - 1] = (uint32_t)((int64_t)itf[G - 1] + (int64_t)round(dg));
            }
            // dynamics.rs:125-183 — unwrapped position sums
            if (L) {
                int64_t* sp = R.sum_pos + (uint64_t)(L - 1) * D;
                int64_t* so = R.sum_outer + (uint64_t)(L - 1) * D * D;
                for (int d = 0; d < D; ++d) sp[d] -= (int64_t)c[d];
                for (int d = 0; d < D; ++d)
                    for (int e = 0; e < D; ++e) so[d * D + e] -= (int64_t)c[d] * (int64_t)c[e];
            }
            if (G) {
                int64_t* sp = R.sum_pos + (uint64_t)(G - 1) * D;
                int64_t* so = R.sum_outer + (uint64_t)(G - 1) * D * D;
                const uint32_t v = vol[G - 1];
                const uint32_t vbg = max(v > 0 ? v - 1 : 0u, 1u);
                double cen[D];
                for (int d = 0; d < D; ++d) cen[d] = (double)sp[d] / (double)vbg;
                int64_t img[D];
                for (int d = 0; d < D; ++d) {
                    int64_t best = (int64_t)c[d];
                    if (periodic) {
                        double bd = fabs((double)best - cen[d]);
                        const int64_t lo = (int64_t)c[d] - (int64_t)dims[d];
                        const int64_t hi = (int64_t)c[d] + (int64_t)dims[d];
                        const double dlo = fabs((double)lo - cen[d]);
                        if (dlo < bd) {
                            best = lo;
                            bd = dlo;
                        }
                        const double dhi = fabs((double)hi - cen[d]);
                        if (dhi < bd) best = hi;
                    }
                    img[d] = best;
                    sp[d] += best;
                }
                for (int d = 0; d < D; ++d)
                    for (int e = 0; e < D; ++e) so[d * D + e] += img[d] * img[e];
            }
            // dynamics.rs:206-224 — periodic halo: rewrite t's mirrors only
            if (periodic) {
                uint32_t opt[D][3], nopt[D];
                bool edge = false;
                for (int d = 0; d < D; ++d) {
                    nopt[d] = 1;
                    opt[d][0] = c[d] + 1;
                    if (c[d] == 0) opt[d][nopt[d]++] = dims[d] + 1;
                    if (c[d] == dims[d] - 1) opt[d][nopt[d]++] = 0;
                    edge |= nopt[d] > 1;
                }
                if (edge) {
                    uint32_t total = 1;
                    for (int d = 0; d < D; ++d) total *= nopt[d];
                    for (uint32_t q = 1; q < total; ++q) {
                        uint32_t r = q, f = 0;
                        for (int d = D - 1; d >= 0; --d) {
                            f += opt[d][r % nopt[d]] * strides[d];
                            r /= nopt[d];
                        }
                        sig[f] = G;
                    }
                }
            }
            energy += dh;  // monte_carlo.rs:270
        }
    }
    for (int i = 0; i < 4; ++i) R.rng[i] = rng.s[i];
    R.energy = energy;
    R.mcs += R.todo;
    for (int i = 0; i < 6; ++i) R.cnt[i] = cnt[i];
}

__global__ void __launch_bounds__(64) k_exact_replay(Rep* __restrict__ reps, uint32_t n,
                                                     const uint64_t* __restrict__ sp3d) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    Rep& R = reps[r];
    if (R.D == 2) replica_chain<2>(R, sp3d);
    else replica_chain<3>(R, sp3d);
}

// ---------------------------------------------------------------------------
// 3D simple-point table, generated from the definition on the device: one
// thread per 64 patterns (simple_point.rs:242-276, 340-354).
// ---------------------------------------------------------------------------
struct SpMasks {
    uint32_t face[26], full[26], origin_face, all;
};

__device__ __forceinline__ uint32_t sp_grow(uint32_t seed, uint32_t uni, const uint32_t* adj) {
    uint32_t comp = seed;
    for (;;) {
        uint32_t nxt = comp;
        for (uint32_t f = comp; f; f &= f - 1) nxt |= adj[__builtin_ctz(f)];
        nxt &= uni;
        if (nxt == comp) return comp;
        comp = nxt;
    }
}

__global__ void __launch_bounds__(256) k_sp3d(SpMasks M, uint64_t* __restrict__ out) {
    __shared__ uint32_t face[26], full[26];
    if (threadIdx.x < 26) {
        face[threadIdx.x] = M.face[threadIdx.x];
        full[threadIdx.x] = M.full[threadIdx.x];
    }
    __syncthreads();
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (1u << 20)) return;
    uint64_t bits = 0;
    for (uint32_t b = 0; b < 64; ++b) {
        const uint32_t obj = w * 64 + b;
        const uint32_t bg = ~obj & M.all;
        bool simple = false;
        if (bg && sp_grow(bg & (0u - bg), bg, full) == bg) {
            int touching = 0;
            for (uint32_t rem = obj; rem;) {
                const uint32_t part = sp_grow(rem & (0u - rem), obj, face);
                rem &= ~part;
                touching += (part & M.origin_face) != 0;
            }
            simple = touching == 1;
        }
        bits |= (uint64_t)simple << b;
    }
    out[w] = bits;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct Device {
    char* base = nullptr;
    Rep rep{};  // device pointers + static fields
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t* sp3d = nullptr;
    Rep* reps_dev = nullptr;
    size_t reps_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;
};

Engine* engine_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device)
        throw Error(CPM_ERR_DEVICE, "no HIP device " + std::to_string(device) + " available");
    auto* e = new Engine();
    e->device = device;
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&e->ev0));
    HIP_OK(hipEventCreate(&e->ev1));
    uint8_t t2[256];
    sp_table2d(t2);
    uint32_t w[8] = {0};
    for (int p = 0; p < 256; ++p)
        if (t2[p]) w[p >> 5] |= 1u << (p & 31);
    HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_sp2d), w, sizeof w));
    return e;
}

static void ensure_sp3d(Engine* e) {
    if (e->sp3d) return;
    HIP_OK(hipMalloc(&e->sp3d, sizeof(uint64_t) << 20));
    SpMasks M{};
    int32_t o[26 * 3];
    int n = stencil(3, NB_TWENTY_SIX, o);
    for (int i = 0; i < n; ++i) {
        int manh0 = 0;
        for (int d = 0; d < 3; ++d) manh0 += std::abs(o[i * 3 + d]);
        if (manh0 == 1) M.origin_face |= 1u << i;
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            int manh = 0, cheb = 0;
            for (int d = 0; d < 3; ++d) {
                int df = std::abs(o[i * 3 + d] - o[j * 3 + d]);
                manh += df;
                cheb = std::max(cheb, df);
            }
            if (manh == 1) M.face[i] |= 1u << j;
            if (cheb == 1) M.full[i] |= 1u << j;
        }
    }
    M.all = (1u << 26) - 1;
    hipLaunchKernelGGL(k_sp3d, dim3((1u << 20) / 256), dim3(256), 0, e->stream, M, e->sp3d);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(e->stream));
}

void engine_destroy(Engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->sp3d) (void)hipFree(e->sp3d);
    if (e->reps_dev) (void)hipFree(e->reps_dev);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

void engine_sp_tables(Engine* e, uint8_t* t2, uint64_t* t3) {
    HIP_OK(hipSetDevice(e->device));
    if (t2) {
        uint32_t w[8];
        HIP_OK(hipMemcpyFromSymbol(w, HIP_SYMBOL(c_sp2d), sizeof w));
        for (int p = 0; p < 256; ++p) t2[p] = (w[p >> 5] >> (p & 31)) & 1u;
    }
    if (t3) {
        ensure_sp3d(e);
        HIP_OK(hipMemcpy(t3, e->sp3d, sizeof(uint64_t) << 20, hipMemcpyDeviceToHost));
    }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Layout of one model's device block.
struct Layout {
    size_t sigma, type_index, volume, interface, sum_pos, sum_outer, contact, params, targets,
        total;
};

static Layout layout_of(const Model& m) {
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off += align256(bytes ? bytes : 4);
        return at;
    };
    const size_t n = m.n_cells, T = m.cfg.types.size(), D = m.D;
    L.sigma = take(m.n_padded * 4);
    L.type_index = take(n * 4);
    L.volume = take(n * 4);
    L.interface = take(n * 4);
    L.sum_pos = take(n * D * 8);
    L.sum_outer = take(n * D * D * 8);
    L.contact = take((n + 1) * (n + 1) * 4);
    L.params = take((2 * T + (T + 1) * (T + 1)) * 8);
    L.targets = take(2 * T * 4);
    L.total = off;
    return L;
}

static void upload(Engine* e, Model* m) {
    if (m->dev) return;
    if (!m->initialized) throw Error(CPM_ERR_ARG, "model is not initialized");
    const Layout L = layout_of(*m);
    auto* d = new Device();
    hipError_t err = hipMalloc(&d->base, L.total);
    if (err != hipSuccess) {
        delete d;
        throw Error(CPM_ERR_OOM, std::string("device allocation failed: ") + hipGetErrorString(err));
    }
    m->dev = d;
    Rep& R = d->rep;
    R.D = m->D;
    R.boundary = m->cfg.boundary;
    R.n_types = (uint32_t)m->cfg.types.size();
    R.n_cells = (uint32_t)m->n_cells;
    for (int k = 0; k < 3; ++k) {
        R.dims[k] = k < m->D ? (uint32_t)m->dims[k] : 1;
        R.strides[k] = k < m->D ? (uint32_t)m->strides[k] : 0;
    }
    R.n_sites = (uint32_t)m->n_sites;
    R.n_copy = m->n_copy;
    R.n_energy = m->n_energy;
    R.n_topo = m->n_topo;
    for (int i = 0; i < m->n_copy; ++i) R.copy_off[i] = (int32_t)m->copy_off[i];
    for (int i = 0; i < m->n_energy; ++i) R.energy_off[i] = (int32_t)m->energy_off[i];
    for (int i = 0; i < m->n_topo; ++i) R.topo_off[i] = (int32_t)m->topo_off[i];
    R.sigma = (uint32_t*)(d->base + L.sigma);
    R.type_index = (uint32_t*)(d->base + L.type_index);
    R.volume = (uint32_t*)(d->base + L.volume);
    R.interface = (uint32_t*)(d->base + L.interface);
    R.sum_pos = (int64_t*)(d->base + L.sum_pos);
    R.sum_outer = (int64_t*)(d->base + L.sum_outer);
    R.contact = (uint32_t*)(d->base + L.contact);
    R.params = (const double*)(d->base + L.params);
    R.targets = (const uint32_t*)(d->base + L.targets);
    const size_t T = R.n_types;
    std::vector<double> params(2 * T + (T + 1) * (T + 1));
    std::vector<uint32_t> targets(2 * T);
    for (size_t t = 0; t < T; ++t) {
        params[t] = m->cfg.types[t].lambda_volume;
        params[T + t] = m->cfg.types[t].lambda_interface;
        targets[t] = m->cfg.types[t].target_volume;
        targets[T + t] = m->cfg.types[t].target_interface;
    }
    for (size_t i = 0; i <= T; ++i)
        for (size_t j = 0; j <= T; ++j) params[2 * T + i * (T + 1) + j] = m->cfg.adhesion[i][j];
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) HIP_OK(hipMemcpyAsync(d->base + off, src, bytes, hipMemcpyHostToDevice, e->stream));
    };
    const size_t n = m->n_cells, D = m->D;
    put(L.sigma, m->sigma.data(), m->n_padded * 4);
    put(L.type_index, m->type_index.data(), n * 4);
    put(L.volume, m->volume.data(), n * 4);
    put(L.interface, m->interface.data(), n * 4);
    put(L.sum_pos, m->sum_pos.data(), n * D * 8);
    put(L.sum_outer, m->sum_outer.data(), n * D * D * 8);
    put(L.contact, m->contact.data(), (n + 1) * (n + 1) * 4);
    put(L.params, params.data(), params.size() * 8);
    put(L.targets, targets.data(), targets.size() * 4);
    HIP_OK(hipStreamSynchronize(e->stream));
    m->device_dirty = false;
}

void device_release(Model* m) {
    if (!m || !m->dev) return;
    (void)hipFree(m->dev->base);
    delete m->dev;
    m->dev = nullptr;
}

void sync_to_host(Model* m) {
    if (!m->dev || !m->device_dirty) return;
    const Layout L = layout_of(*m);
    const Rep& R = m->dev->rep;
    (void)R;
    const size_t n = m->n_cells, D = m->D;
    auto get = [&](void* dst, size_t off, size_t bytes) {
        if (bytes) HIP_OK(hipMemcpy(dst, m->dev->base + off, bytes, hipMemcpyDeviceToHost));
    };
    get(m->sigma.data(), L.sigma, m->n_padded * 4);
    get(m->volume.data(), L.volume, n * 4);
    get(m->interface.data(), L.interface, n * 4);
    get(m->sum_pos.data(), L.sum_pos, n * D * 8);
    get(m->sum_outer.data(), L.sum_outer, n * D * D * 8);
    get(m->contact.data(), L.contact, (n + 1) * (n + 1) * 4);
    m->device_dirty = false;
}

void sync_scalars(Model*) {}  // scalars are written back after every launch

static void ensure_reps(Engine* e, size_t n) {
    if (n <= e->reps_cap) return;
    if (e->reps_dev) HIP_OK(hipFree(e->reps_dev));
    e->reps_dev = nullptr;
    size_t cap = std::max(n, (size_t)64);
    HIP_OK(hipMalloc(&e->reps_dev, cap * sizeof(Rep)));
    e->reps_cap = cap;
}

static void add_counters(cpm_counters* c, const Rep& R) {
    if (!c) return;
    c->attempts += R.cnt[0];
    c->non_null += R.cnt[1];
    c->topology += R.cnt[2];
    c->energy += R.cnt[3];
    c->accepted += R.cnt[4];
    c->exp_near_ties += R.cnt[5];
}

void run_mcs_each(Engine* e, Model* const* models, size_t n, const uint64_t* todo, int mode,
                  cpm_counters* counters) {
    if (mode != CPM_MODE_EXACT_REPLAY)
        throw Error(CPM_ERR_ARG, "sublattice mode is not available in this build");
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    bool any3d = false;
    for (size_t i = 0; i < n; ++i) {
        upload(e, models[i]);
        any3d |= models[i]->D == 3;
    }
    if (any3d) ensure_sp3d(e);
    ensure_reps(e, n);
    std::vector<Rep> reps(n);
    for (size_t i = 0; i < n; ++i) {
        Model* m = models[i];
        Rep R = m->dev->rep;
        for (int k = 0; k < 4; ++k) R.rng[k] = m->rng.s[k];
        R.energy = m->energy;
        R.mcs = m->mcs;
        R.todo = todo[i];
        reps[i] = R;
    }
    HIP_OK(hipMemcpyAsync(e->reps_dev, reps.data(), n * sizeof(Rep), hipMemcpyHostToDevice,
                          e->stream));
    HIP_OK(hipEventRecord(e->ev0, e->stream));
    const uint32_t block = 64;
    hipLaunchKernelGGL(k_exact_replay, dim3((uint32_t)((n + block - 1) / block)), dim3(block), 0,
                       e->stream, e->reps_dev, (uint32_t)n, e->sp3d);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(e->ev1, e->stream));
    HIP_OK(hipMemcpyAsync(reps.data(), e->reps_dev, n * sizeof(Rep), hipMemcpyDeviceToHost,
                          e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    if (counters) counters->kernel_ms += ms;
    for (size_t i = 0; i < n; ++i) {
        Model* m = models[i];
        for (int k = 0; k < 4; ++k) m->rng.s[k] = reps[i].rng[k];
        m->energy = reps[i].energy;
        m->mcs = reps[i].mcs;
        if (todo[i]) m->device_dirty = true;
        add_counters(counters, reps[i]);
    }
}

void run_mcs(Engine* e, Model* const* models, size_t n, uint64_t n_mcs, int mode,
             cpm_counters* counters) {
    std::vector<uint64_t> todo(n, n_mcs);
    run_mcs_each(e, models, n, todo.data(), mode, counters);
}

void run_mcs_plain(Engine* e, Model* m, uint64_t n_mcs, cpm_counters* counters) {
    for (uint64_t i = 0; i < n_mcs; ++i) {
        uint64_t one = 1;
        run_mcs_each(e, &m, 1, &one, CPM_MODE_EXACT_REPLAY, counters);
    }
}

}  // namespace cpm

// Device engine: HIP kernels for gfx950 and their host-side launch logic.
//
// Exact replay reproduces the reference's serial attempt loop
// (crates/cpm-core/src/monte_carlo.rs:37-145, 327-357) bit for bit.  Two
// execution paths share the same arithmetic:
//
//  * LDS path (lattices whose haloed u16 copy fits in LDS): one wave per
//    replica, one launch per call.  The wave writes its Xoshiro256++
//    stream ahead into a 2 MiB HBM buffer — 64 segments of 4096 draws, lane
//    g jumping to segment g with a GF(2) jump polynomial — and refills it
//    whenever its cursor reaches the end.  k_lds_engine keeps sigma (haloed,
//    packed u16), V and I resident in LDS, holds a 128-draw window of the
//    stream in registers, evaluates the attempt starting at every window
//    position lane-parallel, and follows the real chain of attempts by
//    pointer doubling.
//  * Global path (everything else, e.g. 64^3 lattices): one lane per
//    replica running the serial loop against HBM-resident state.
//
// FP64 energy arithmetic is compiled with -ffp-contract=off and keeps the
// reference's operation order (crates/cpm-core/src/energy.rs:92-103,
// 167-202, 242-264, 521-523); sums of +-1.0 are formed from popcounts,
// which is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <vector>

#include "core.hpp"
#include "device.hpp"
#include "driver.hpp"
#include "host.hpp"
#include "binomial.hpp"
#include "metropolis.hpp"
#include "rng_jump.hpp"

namespace cpm {

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            throw Error(_e == hipErrorOutOfMemory ? CPM_ERR_OOM : CPM_ERR_DEVICE,      \
                        std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

constexpr uint32_t kSegLen = 4096;             // draws per stream segment (one lane's share)
constexpr uint32_t kSegs = 64;                 // segments per fill (one per lane)
constexpr uint32_t kFill = kSegLen * kSegs;    // draws buffered per replica

// ---------------------------------------------------------------------------
// Replica descriptor: geometry, stencils, parameters, state pointers and the
// scalar state a launch reads and writes back.
// ---------------------------------------------------------------------------
struct Rep {
    uint32_t D, boundary, n_types, n_cells;
    uint32_t dims[3], strides[3];
    uint32_t n_sites, n_padded, n_copy, n_energy, n_topo;
    int32_t copy_off[6], energy_off[26], topo_off[26];
    uint32_t* sigma;
    uint32_t* type_index;
    uint32_t* volume;
    uint32_t* interface;
    int64_t* sum_pos;
    int64_t* sum_outer;
    uint32_t* contact;
    const double* params;     // lambda_v[T], lambda_i[T], J[(T+1)^2], lambda_act[T]
    const uint32_t* targets;  // V*[T], I*[T], max_act[T]
    uint32_t acceptance;      // CPM_ACCEPT_* (the lane path takes both)
    uint32_t act;             // Act term on (lane path only)
    uint32_t* birth;          // Act: per padded site, the MCS it was last taken
    uint32_t proposal;        // CPM_PROPOSAL_* (EdgeList: lane path only)
    uint32_t* edge_sites;     // EdgeList: edge sites (padded indices) in list order
    uint32_t* edge_pos;       // EdgeList: each padded site's list position, or ~0
    uint64_t n_edge;          // EdgeList: list length at the start / end of a call
    uint64_t* draws;          // kFill buffered draws + kSegs segment states (wave path)
    uint16_t* sig16;          // packed sigma in HBM (wave path with sigma outside LDS)
    uint32_t* labels;         // readouts: union-find parent per padded site
    uint32_t* lcount;         // readouts: per-root counts (medium volume | touch << 31)
    unsigned char* readout;   // readouts: the sample's device part (ReadoutLayout)
    uint64_t rng[4];          // consumer state at the start / end of a call
    double energy;
    uint64_t mcs;
    uint64_t todo;            // MCS left to run in this call
    uint32_t attempt_in_mcs;  // attempts done within the current MCS
    uint32_t done;
    uint32_t error;           // an attempt needed more draws than a window holds
    uint64_t fill_state[4];   // sublattice mode: [0] keys its randomness
    unsigned long long cnt[6];
    uint64_t t_run;           // wave path: 100 MHz ticks the replica's chain ran in this call
    uint64_t budget;          // wave path: 100 MHz ticks the chain may run (0: no limit, at
                              // most 40 s); a chain out of time stops after a walked window,
                              // mid-MCS if so
};

__constant__ uint32_t c_sp2d[8];  // 256-bit 2D simple-point table

// Scratch of the entry point that allocates nothing (cpm_run_mcs_noalloc):
// one call's replica descriptors, then its launch order.  Each device holds
// its own copy, whose address the engine takes once at creation; every
// call on a device overwrites the whole used range before its kernels read
// it, so nothing needs clearing, and overlapping calls on one device race.
constexpr size_t kNoallocMaxReplicas = CPM_NOALLOC_MAX_REPLICAS;
constexpr size_t kNoallocRepBytes = (kNoallocMaxReplicas * sizeof(Rep) + 255) / 256 * 256;
constexpr size_t kNoallocScratchBytes = kNoallocRepBytes + kNoallocMaxReplicas * sizeof(uint32_t);
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

// ===========================================================================
// Global path: one lane per replica (lattices too large for LDS, and every
// Metropolis-Hastings or Act model)
// ===========================================================================

// The Act term's geometric mean (crates/cpm-core/src/energy.rs:403-444)
// around the interior site at coordinate cc (0-based): activity
// max(0, max_act - (mcs - birth)) over the site and its copy neighbours
// owned by `rel`, multiplied in that order, then the count-th root.  A
// neighbour across a periodic face is read at the interior site it
// mirrors; across a fixed face it is the halo (medium, birth 0).
template <int D>
__device__ double act_mean(const Rep& R, const uint32_t* cc, const int* axis, const int* sgn,
                           uint32_t rel, uint32_t max_act, uint64_t mcs) {
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;
    double product = 1.0;
    uint32_t count = 0;
    for (int k = -1; k < (int)R.n_copy; ++k) {
        uint32_t f = 0;
        for (int d = 0; d < D; ++d) {
            int64_t x = (int64_t)cc[d] + 1;  // padded coordinate
            if (k >= 0 && axis[k] == d) {
                x += sgn[k];
                if (periodic) x = x == 0 ? R.dims[d] : (x == (int64_t)R.dims[d] + 1 ? 1 : x);
            }
            f += (uint32_t)x * R.strides[d];
        }
        if (R.sigma[f] != rel) continue;
        const uint64_t birth = R.birth[f];
        const uint64_t elapsed = mcs > birth ? mcs - birth : 0;
        const uint64_t activity = (uint64_t)max_act > elapsed ? (uint64_t)max_act - elapsed : 0;
        product *= (double)activity;
        count += 1;
    }
    return act_root(product, count);
}

// delta_H_Act (energy.rs:446-457) of copying the source (copy neighbour k
// of the target at coordinate c, owned by G) onto the target (owned by L).
template <int D>
__device__ double act_delta(const Rep& R, const uint32_t* c, uint32_t k, const int* axis,
                            const int* sgn, uint32_t L, uint32_t G, uint64_t mcs) {
    if (G == 0) return 0.0;  // medium gaining: no type to take lambda and max_act from
    const uint32_t T = R.n_types;
    const uint32_t ty = R.type_index[G - 1];
    const double lambda = R.params[2 * T + (T + 1) * (T + 1) + ty];
    const uint32_t max_act = R.targets[2 * T + ty];
    uint32_t cs[D];
    for (int d = 0; d < D; ++d) cs[d] = c[d];
    // the source's interior coordinate (a gaining cell is never a fixed halo)
    const int64_t x = (int64_t)cs[axis[k]] + sgn[k];
    cs[axis[k]] = x < 0 ? R.dims[axis[k]] - 1 : (x == (int64_t)R.dims[axis[k]] ? 0 : (uint32_t)x);
    const double gm_src = act_mean<D>(R, cs, axis, sgn, G, max_act, mcs);
    const double gm_tgt = act_mean<D>(R, c, axis, sgn, L, max_act, mcs);
    return -(lambda / (double)max_act) * (gm_src - gm_tgt);
}

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
    const bool mh = R.acceptance == CPM_ACCEPT_METROPOLIS_HASTINGS;
    // the axis and direction of each copy offset (von Neumann: +-stride)
    int axis[6], sgn[6];
    for (uint32_t k = 0; k < n_copy; ++k) {
        const int32_t o = R.copy_off[k];
        for (int d = 0; d < D; ++d)
            if ((uint32_t)(o < 0 ? -o : o) == strides[d]) {
                axis[k] = d;
                sgn[k] = o < 0 ? -1 : 1;
            }
    }

    // EdgeList (edge_list.rs:42-186): the list lives in HBM; its length here
    const bool edges = R.proposal == CPM_PROPOSAL_EDGE_LIST;
    uint64_t n_edge = R.n_edge;
    uint32_t* __restrict__ esites = R.edge_sites;
    uint32_t* __restrict__ epos = R.edge_pos;
    // EdgeList::update_around (edge_list.rs:150-174): the target and its
    // copy neighbours (a periodic halo site canonicalised, a fixed one
    // skipped), each inserted or swap-removed by its new edge status
    auto update_edges = [&](uint32_t t, const uint32_t* c) {
        for (int k = -1; k < (int)n_copy; ++k) {
            uint32_t site = t;
            if (k >= 0) {
                const int ax = axis[k];
                const int64_t x = (int64_t)c[ax] + sgn[k];
                if (x < 0 || x >= (int64_t)dims[ax]) {
                    if (!periodic) continue;
                    const uint32_t w = x < 0 ? dims[ax] - 1 : 0;
                    site = (uint32_t)((int64_t)t + ((int64_t)w - (int64_t)c[ax]) * strides[ax]);
                } else {
                    site = (uint32_t)((int32_t)t + R.copy_off[k]);
                }
            }
            const uint32_t own = sig[site];
            bool edge = false;
            for (uint32_t q = 0; q < n_copy; ++q) edge |= sig[(int32_t)site + R.copy_off[q]] != own;
            const uint32_t at = epos[site];
            if (edge) {
                if (at == 0xffffffffu) {
                    epos[site] = (uint32_t)n_edge;
                    esites[n_edge++] = site;
                }
            } else if (at != 0xffffffffu) {
                const uint32_t last = (uint32_t)(n_edge - 1);
                if (at != last) {
                    const uint32_t moved = esites[last];
                    esites[at] = moved;
                    epos[moved] = at;
                }
                n_edge -= 1;
                epos[site] = 0xffffffffu;
            }
        }
    };

    for (uint64_t m = 0; m < R.todo; ++m) {
        const uint64_t mcs = R.mcs + m;
        // one attempt at target t (coordinate c): monte_carlo.rs:52-145;
        // `return` ends the attempt
        auto attempt_at = [&](const uint32_t t, const uint32_t* c) {
            uint64_t k;
            while (!range_trial(xo_next(rng), n_copy, zone_k, &k)) {
            }
            const uint32_t s = (uint32_t)((int32_t)t + R.copy_off[k]);
            cnt[0]++;
            const uint32_t L = sig[t], G = sig[s];
            if (L == G) return;
            cnt[1]++;
            if (L != 0) {
                uint32_t pat = 0;
                for (uint32_t i = 0; i < n_topo; ++i)
                    pat |= (uint32_t)(sig[(int32_t)t + R.topo_off[i]] == L) << i;
                cnt[2]++;
                const bool simple = D == 2 ? ((c_sp2d[pat >> 5] >> (pat & 31)) & 1u)
                                           : ((sp3d[pat >> 6] >> (pat & 63)) & 1ull);
                if (!simple) return;
                if (vol[L - 1] <= 1) return;
            }
            cnt[3]++;
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
            // monte_carlo.rs:80-97: the Act term enters the test, never the energy
            const double dh_acc =
                R.act ? dh + act_delta<D>(R, c, (uint32_t)k, axis, sgn, L, G, mcs) : dh;
            bool accept;
            if (mh) {
                // monte_carlo.rs:168-222: no downhill shortcut; n_A == 0
                // rejects before the draw
                uint32_t na = 0, nb = 0;
                for (uint32_t q = 0; q < n_copy; ++q) {
                    const uint32_t v = sig[(int32_t)t + R.copy_off[q]];
                    na += v == L;
                    nb += v == G;
                }
                if (na == 0) return;
                const double u = gen_f64(rng);
                bool tie = false;
                accept = mh_accept(u, na, nb, dh_acc, &tie);
                if (tie) cnt[5]++;
            } else {
                accept = dh_acc <= 0.0;
                if (!accept) {
                    const double u = gen_f64(rng);
                    bool tie = false;
                    accept = metropolis_accept(u, dh_acc, &tie);
                    if (tie) cnt[5]++;
                }
            }
            if (!accept) return;
            cnt[4]++;
            if (R.act) R.birth[t] = (uint32_t)mcs;  // energy.rs:459-464
            for (uint32_t q = 0; q < n_energy; ++q) {
                const uint32_t nb = sig[(int32_t)t + R.energy_off[q]];
                if (nb != L) {
                    const uint64_t key = (uint64_t)min(L, nb) * ncell1 + max(L, nb);
                    const uint32_t cur = R.contact[key];
                    R.contact[key] = cur ? cur - 1 : 0;
                }
                if (nb != G) R.contact[(uint64_t)min(G, nb) * ncell1 + max(G, nb)] += 1;
            }
            sig[t] = G;
            if (L) {
                vol[L - 1] -= 1;
                itf[L - 1] = (uint32_t)((int64_t)itf[L - 1] + (int64_t)round(dl));
            }
            if (G) {
                vol[G - 1] += 1;
                itf[G - 1] = (uint32_t)((int64_t)itf[G - 1] + (int64_t)round(dg));
            }
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
                        if (fabs((double)hi - cen[d]) < bd) best = hi;
                    }
                    img[d] = best;
                    sp[d] += best;
                }
                for (int d = 0; d < D; ++d)
                    for (int e = 0; e < D; ++e) so[d * D + e] += img[d] * img[e];
            }
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
            if (edges) update_edges(t, c);  // monte_carlo.rs:257-264
            energy += dh;
        };
        if (!edges) {
            for (uint32_t a = 0; a < R.n_sites; ++a) {
                uint32_t c[D];
                uint32_t t = 0;
                for (int d = 0; d < D; ++d) {
                    uint64_t v;
                    while (!range_trial(xo_next(rng), dims[d], zone[d], &v)) {
                    }
                    c[d] = (uint32_t)v;
                    t += (c[d] + 1) * strides[d];
                }
                attempt_at(t, c);
            }
        } else {
            // monte_carlo.rs:336-351: k ~ Binomial(N_sites, |E| / N_sites),
            // |E| read once; k targets drawn from the list as it changes
            uint64_t todo_k = 0;
            if (n_edge && !binomial_sample(rng, R.n_sites, (double)n_edge / (double)R.n_sites,
                                           &todo_k))
                R.error = 2;
            for (uint64_t a = 0; a < todo_k; ++a) {
                if (n_edge == 0) break;
                const uint32_t t = esites[gen_range(rng, n_edge)];
                uint32_t c[D], f = t;
                for (int d = 0; d < D; ++d) {
                    c[d] = f / strides[d] - 1;
                    f %= strides[d];
                }
                attempt_at(t, c);
            }
        }
    }
    for (int i = 0; i < 4; ++i) R.rng[i] = rng.s[i];
    R.energy = energy;
    R.mcs += R.todo;
    R.todo = 0;
    R.done = 1;
    R.n_edge = n_edge;
    for (int i = 0; i < 6; ++i) R.cnt[i] += cnt[i];
}

__global__ void __launch_bounds__(64) k_exact_replay(Rep* __restrict__ reps,
                                                     const uint32_t* __restrict__ idx, uint32_t n,
                                                     const uint64_t* __restrict__ sp3d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rep& R = reps[idx[i]];
    if (R.D == 2) replica_chain<2>(R, sp3d);
    else replica_chain<3>(R, sp3d);
}

// ===========================================================================
// LDS path
// ===========================================================================

// x^(g * kSegLen) mod p(x): the jump from a fill's start to segment g
__constant__ uint64_t c_jump[kSegs][4];

// Orders one wave's LDS and global accesses around a point where its lanes
// exchange data (a workgroup barrier minus the rendezvous, so that it is
// safe inside code only one wave of a workgroup runs).
// Orders this thread's earlier global-memory stores before its later
// accesses for every thread of its workgroup (all of them on one CU, one
// L1): the exact-replay kernels' lattice words in HBM are written and read
// by the replica's own wave(s) only, with plain loads and stores.
__device__ __forceinline__ void wg_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// The wave writes kFill consecutive draws of the stream whose state at the
// buffer start is s0: lane g jumps to draw g * kSegLen (sum of T^i s0 over
// the polynomial's terms) and steps through its segment.  Segment start
// states go to segs[g] for later rebasing.
__device__ void wave_fill(uint64_t* __restrict__ draws, uint64_t* __restrict__ segs,
                          const uint64_t* s0, uint32_t lane) {
    Xoshiro s{{s0[0], s0[1], s0[2], s0[3]}};
    if (lane) {
        Xoshiro acc{{0, 0, 0, 0}};
        for (int i = 0; i < 256; ++i) {
            if ((c_jump[lane][i >> 6] >> (i & 63)) & 1)
                for (int k = 0; k < 4; ++k) acc.s[k] ^= s.s[k];
            xo_next(s);
        }
        s = acc;
    }
    for (int k = 0; k < 4; ++k) segs[lane * 4 + k] = s.s[k];
    uint64_t* __restrict__ out = draws + (size_t)lane * kSegLen;
    for (uint32_t i = 0; i < kSegLen; ++i) out[i] = xo_next(s);
    wg_fence();  // the wave reads other lanes' segments next
    wsync();
}

// Whether a chain with a time budget (100 MHz ticks, 0: none) has used it.
__device__ __forceinline__ bool out_of_time(const uint32_t& budget_lds, uint64_t t_begin) {
    const uint32_t b = __builtin_amdgcn_readfirstlane(*(const volatile uint32_t*)&budget_lds);
    return b && (uint32_t)(__builtin_amdgcn_s_memrealtime() - t_begin) >= b;
}

__device__ __forceinline__ uint32_t rl32(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, uint32_t lane) {
    return ((uint64_t)rl32((uint32_t)(v >> 32), lane) << 32) | rl32((uint32_t)v, lane);
}
__device__ __forceinline__ double rlf64(double v, uint32_t lane) {
    return __longlong_as_double((long long)rl64((uint64_t)__double_as_longlong(v), lane));
}
__device__ __forceinline__ uint32_t uni(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// First set bit at or after p (< 128) in the 128-bit mask {m0, m1}; 128 if none.
__device__ __forceinline__ uint32_t next_bit(uint64_t m0, uint64_t m1, uint32_t p) {
    if (p < 64) {
        const uint64_t m = m0 & (~0ull << p);
        if (m) return (uint32_t)__builtin_ctzll(m);
        p = 64;
    }
    if (p >= 128) return 128;
    const uint64_t m = m1 & (~0ull << (p - 64));
    return m ? 64 + (uint32_t)__builtin_ctzll(m) : 128;
}

// LDS layout of one replica (host computes the same offsets).
struct LdsLayout {
    uint32_t sigma, type, vol, itf, J, lv, li, tv, ti, total;
};

__host__ __device__ inline LdsLayout lds_layout(uint32_t n_padded, uint32_t n_cells, uint32_t T,
                                                bool sigma_in_lds = true) {
    LdsLayout L;
    uint32_t off = 0;
    auto take = [&](uint32_t bytes, uint32_t align) {
        off = (off + align - 1) / align * align;
        uint32_t at = off;
        off += bytes;
        return at;
    };
    L.J = take((T + 1) * (T + 1) * 8, 8);
    L.lv = take(T * 8, 8);
    L.li = take(T * 8, 8);
    L.vol = take(n_cells * 4 + 4, 4);
    L.itf = take(n_cells * 4 + 4, 4);
    L.tv = take(T * 4, 4);
    L.ti = take(T * 4, 4);
    L.sigma = take(sigma_in_lds ? n_padded * 2 : 0, 4);
    L.type = take(n_cells + 1, 4);
    L.total = (off + 15) / 16 * 16;
    return L;
}

// One consumer's per-lane sample for one 64-draw half of the window.
struct Half {
    uint64_t w;        // raw draw
    uint32_t hi[4];    // sample for consumers 0..D-1 (axes) and D (copy index)
};

template <int D>
__device__ __forceinline__ void prep_half(Half& h, uint64_t w, const uint32_t* range,
                                          const uint64_t* zone, uint64_t* mask) {
    h.w = w;
#pragma unroll
    for (int c = 0; c <= D; ++c) {
        uint64_t v;
        const bool ok = range_trial(w, range[c], zone[c], &v);
        h.hi[c] = (uint32_t)v;
        mask[c] = __ballot(ok);
    }
}

// First set bit at or after the per-lane position p (0..129) of the 128-bit
// mask {m0, m1}; 128 if there is none.
__device__ __forceinline__ uint32_t lane_next_bit(uint64_t m0, uint64_t m1, uint32_t p) {
    const uint64_t lo = p < 64 ? (m0 & (~0ull << (p & 63))) : 0ull;
    const uint32_t ph = p < 64 ? 0u : p - 64;
    const uint64_t hi = ph < 64 ? (m1 & (~0ull << (ph & 63))) : 0ull;
    return lo ? (uint32_t)__builtin_ctzll(lo) : (hi ? 64u + (uint32_t)__builtin_ctzll(hi) : 128u);
}

__device__ __forceinline__ uint32_t bperm(uint32_t v, uint32_t src_lane) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src_lane << 2), (int)v);
}

__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}

// LDS words of the exact-replay engine: cell id in the low 12 bits, the
// cell's adhesion-matrix index (type + 1; 0 for medium) in the high 4, so
// one read yields both.
constexpr uint32_t kIdMask = 0xFFF;
constexpr uint32_t kMaxPackedCells = 4095;

// Index, in the full-shell (topology) stencil order, of entry q of the
// energy stencil with NE entries.  Full shell: code 0..3^D-1 with
// component 0 fastest, origin skipped; von Neumann: axis 0 +1, axis 0 -1,
// axis 1 +1, ...; eighteen: the full shell without its corners
// (crates/cpm-core/src/neighborhood.rs:45-115).
__host__ __device__ constexpr int shell_index(int D, int NE, int q) {
    const int total = D == 2 ? 9 : 27, origin = total / 2;
    const bool vn = NE == 2 * D;
    if (vn) {
        const int axis = q / 2, sgn = q % 2 == 0 ? 1 : -1;
        int code = 0, pw = 1;
        for (int d = 0; d < D; ++d) {
            code += ((d == axis ? sgn : 0) + 1) * pw;
            pw *= 3;
        }
        return code > origin ? code - 1 : code;
    }
    int n = 0;
    for (int code = 0; code < total; ++code) {
        if (code == origin) continue;
        int c = code, nz = 0;
        for (int d = 0; d < D; ++d) {
            nz += (c % 3) != 1;
            c /= 3;
        }
        if (NE == 18 && nz > 2) continue;
        if (n == q) return code > origin ? code - 1 : code;
        ++n;
    }
    return -1;
}

// Meta word of a decoded attempt, per slot and lane of the wave path: the
// window position after its draws, and what its evaluation found.
constexpr uint32_t kMe = 0xFF;      // position after the attempt's draws (window-relative)
constexpr uint32_t kMv = 1u << 8;   // every draw inside the 128-draw window
constexpr uint32_t kMn = 1u << 9;   // non-null: sigma(t) != sigma(s)
constexpr uint32_t kMu = 1u << 10;  // dH > 0 after the checks: u is drawn
constexpr uint32_t kMa = 1u << 11;  // accepted
constexpr uint32_t kMp = 1u << 12;  // topology checked (t not medium)
constexpr uint32_t kMg = 1u << 13;  // passed the checks: energy evaluated
constexpr uint32_t kMt = 1u << 14;  // u near exp(-dH): decided by the correctly rounded exp
// LDS the wave path adds after the replica's tables: one evaluation pass's
// 64 staged attempts (target, source, u draw).
constexpr uint32_t kStageBytes = 64 * 16;
// LDS the pipelined kernel adds to the single-wave one: the second wave's
// stage and two block buffers of KB slots (16 bytes per lane).
__host__ __device__ constexpr uint32_t pipe_extra_bytes(int D) {
    return kStageBytes + 2u * (D == 2 ? 3u : 4u) * 64u * 16u;
}

// Packed sigma words, in LDS or (G) in HBM.  The HBM copy is read through
// global-address-space loads (flat loads would also count against the LDS
// wait counter).
template <bool G>
struct SigmaRef {
    uint16_t* p;
    __device__ uint32_t ld(int64_t i) const { return p[i]; }
    __device__ void st(int64_t i, uint32_t v) const { p[i] = (uint16_t)v; }
};
template <>
struct SigmaRef<true> {
    __attribute__((address_space(1))) uint16_t* p;
    __device__ uint32_t ld(int64_t i) const { return p[i]; }
    __device__ void st(int64_t i, uint32_t v) const { p[i] = (uint16_t)v; }
};

// Exact replay on the wave path: one wave per replica.  The stream is cut
// into 64-position windows; lane i of a window decodes the attempt that
// would start at window position i (the first accepted draw of each
// consumer, found in the window's 128 draws by mask search).  A block of KB
// windows is decoded at once; its attempts that are not null on the current
// lattice (about one in six) are packed onto consecutive lanes through LDS
// and evaluated completely, 64 per pass.  The wave then walks the actual
// chain of attempts through each window (pointer doubling) up to the first
// accepted copy, evaluates that attempt again on every lane, commits it,
// re-evaluates the rest of the block and goes on.  NE is the energy stencil
// size; NT = 3^D - 1 the full shell.
template <int D, int NE, bool SG, bool PW, bool TB>
__global__ void __launch_bounds__(PW ? 128 : 64) __attribute__((amdgpu_waves_per_eu(D == 2 ? (PW ? 3 : 4) : 1))) k_lds_engine(Rep* __restrict__ reps,
                                                   const uint32_t* __restrict__ idx,
                                                   const uint64_t* __restrict__ sp3d) {
    constexpr int NT = D == 2 ? 8 : 26;
    constexpr uint32_t NTH = PW ? 128 : 64;
    extern __shared__ __align__(16) unsigned char smem[];
    Rep& R = reps[idx[blockIdx.x]];
    if (R.done) return;
    const uint64_t t_begin = __builtin_amdgcn_s_memrealtime();
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = PW ? threadIdx.x >> 6 : 0u;  // pipelined: 0 walks, 1 decodes/evaluates
    const uint32_t T = R.n_types, T1 = T + 1, n_cells = R.n_cells;
    const uint32_t n_padded = R.n_padded, n_sites = R.n_sites;
    const LdsLayout Ly = lds_layout(n_padded, n_cells, T, !SG);
    SigmaRef<SG> sig;
    if constexpr (SG)
        sig.p = (__attribute__((address_space(1))) uint16_t*)R.sig16;
    else
        sig.p = (uint16_t*)(smem + Ly.sigma);
    uint32_t* vol = (uint32_t*)(smem + Ly.vol);  // vol[id - 1]
    uint32_t* itf = (uint32_t*)(smem + Ly.itf);
    double* J = (double*)(smem + Ly.J);
    double* lv = (double*)(smem + Ly.lv);
    double* li = (double*)(smem + Ly.li);
    uint32_t* tv = (uint32_t*)(smem + Ly.tv);
    uint32_t* ti = (uint32_t*)(smem + Ly.ti);

    // ---- stage the replica into LDS
    const uint32_t* __restrict__ gtyp = R.type_index;
    for (uint32_t i = threadIdx.x; i < n_padded; i += NTH) {
        const uint32_t id = R.sigma[i];
        sig.st(i, id ? id | ((gtyp[id - 1] + 1) << 12) : 0u);
    }
    if constexpr (SG) wg_fence();
    for (uint32_t i = threadIdx.x; i < n_cells; i += NTH) {
        vol[i] = R.volume[i];
        itf[i] = R.interface[i];
    }
    for (uint32_t i = threadIdx.x; i < T1 * T1; i += NTH) J[i] = R.params[2 * T + i];
    for (uint32_t i = threadIdx.x; i < T; i += NTH) {
        lv[i] = R.params[i];
        li[i] = R.params[T + i];
        tv[i] = R.targets[i];
        ti[i] = R.targets[T + i];
    }
    __syncthreads();

    // ---- uniform geometry
    const uint32_t n_copy = R.n_copy;
    const int32_t eoff = lane < NE ? R.energy_off[lane] : 0;  // for the contact updates
    const uint32_t coff = lane < n_copy ? (uint32_t)R.copy_off[lane] : 0u;
    uint32_t range[D + 1], stride[D], dims[D];
    uint64_t zone[D + 1];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        dims[d] = R.dims[d];
        range[d] = dims[d];
        stride[d] = R.strides[d];
        zone[d] = range_zone(range[d]);
    }
    range[D] = n_copy;
    zone[D] = range_zone(n_copy);
    uint32_t t_safe = 0;  // an interior site, for lanes whose decode is void
#pragma unroll
    for (int d = 0; d < D; ++d) t_safe += stride[d];
    const uint64_t sp2a = (uint64_t)c_sp2d[0] | (uint64_t)c_sp2d[1] << 32;
    const uint64_t sp2b = (uint64_t)c_sp2d[2] | (uint64_t)c_sp2d[3] << 32;
    const uint64_t sp2c = (uint64_t)c_sp2d[4] | (uint64_t)c_sp2d[5] << 32;
    const uint64_t sp2d = (uint64_t)c_sp2d[6] | (uint64_t)c_sp2d[7] << 32;
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;
    const uint64_t ncell1 = (uint64_t)n_cells + 1;
    // global address space: a flat load would also count against the LDS
    // wait counter, so every LDS wait would stall on the draw prefetch
    const __attribute__((address_space(1))) uint64_t* __restrict__ draws =
        (const __attribute__((address_space(1))) uint64_t*)R.draws;

    uint64_t todo = R.todo;
    uint32_t ain = R.attempt_in_mcs;
    uint64_t mcs = R.mcs;
    // TB (a time-limited call): the time budget in 100 MHz ticks, at most
    // 40 s, kept in LDS and read once per walked window, so that it holds no
    // scalar register.  Calls without a budget run the TB = false kernels,
    // whose loop carries no check at all.
    __shared__ uint32_t s_budget;
    if (TB && threadIdx.x == 0) s_budget = (uint32_t)min(R.budget, (uint64_t)4000000000u);
    double energy = R.energy;
    uint32_t c_att = 0, c_nn = 0, c_top = 0, c_en = 0, c_acc = 0, c_tie = 0;
    bool dirty = false, error = false;

    // ---- the stream buffer, written by this wave from the call's RNG state
    uint64_t* const gdraws = R.draws;
    uint64_t* const segs = R.draws + kFill;  // start state of each segment
    if (wave == (PW ? 1u : 0u)) {
        const uint64_t s0[4] = {R.rng[0], R.rng[1], R.rng[2], R.rng[3]};
        wave_fill(gdraws, segs, s0, lane);
    }

    // ---- blocks of KB consecutive 64-position windows.  `base` is the
    // buffer position where slot 0's window starts, p the chain position in
    // it; the decode of the window in slot k reads the 128 draws from its
    // start, halves hr[k] and hr[k + 1] of the block.
    // (3D: four windows amortise the block's dependent global-memory round
    // trips — null check, neighbourhood, 2^26-bit table — over more attempts)
    constexpr int KB = D == 2 ? 3 : 4;
    static_assert(pipe_extra_bytes(D) == kStageBytes + 2u * KB * 64u * 16u, "block buffers");
    uint4* const stage = (uint4*)(smem + Ly.total + wave * kStageBytes);  // one pass's attempts
    uint32_t base = 0, p = 0;
    uint64_t hr[KB + 1];
    uint32_t st_t[KB], st_m[KB];  // per slot and lane: target, meta (copy index in bits 16-20)
    uint64_t st_u[KB];            // the draw after the attempt (its u when dH > 0)
#pragma unroll
    for (int j = 0; j <= KB; ++j) hr[j] = 0;
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        st_t[k] = t_safe;
        st_m[k] = 0;
        st_u[k] = 0;
    }
    uint32_t nvalid = 0, commit_p = 0;
    bool need_block = true, need_eval = false, hr_ready = false, pending = false;
    bool walk_hit = false;
    // Copies committed since the block's evaluation (padded coordinates and
    // the two cells), so that only attempts they can have changed are
    // evaluated again; clog_all: evaluate everything (a copy on a periodic
    // boundary, whose halo mirrors also changed, or a full log).
    constexpr uint32_t kClog = 8;
    __shared__ uint32_t s_clog[kClog][4];
    uint32_t clog_n = 0;
    bool clog_all = false, eval_sel = false;

    // The buffer is restarted at the segment holding position `at` (its
    // draws regenerated from the segment's start state).
#define CPM_REFILL(at_)                                                                      \
    do {                                                                                     \
        const uint32_t at = (at_), g = min(at / kSegLen, kSegs - 1);                          \
        /* a vector load: the segment states were written by vector stores */                \
        const uint64_t mine = segs[g * 4 + (lane & 3)];                                       \
        const uint64_t s0[4] = {rl64(mine, 0), rl64(mine, 1), rl64(mine, 2), rl64(mine, 3)}; \
        wsync();                                                                             \
        wave_fill(gdraws, segs, s0, lane);                                                   \
        base = at - g * kSegLen;                                                             \
        p = 0;                                                                               \
        hr_ready = false;                                                                    \
        need_block = true;                                                                   \
    } while (0)

    // Decode of slot k: lane i takes the attempt starting at window
    // position i — the first accepted draw of each consumer in turn, found
    // by mask search — and the draw right after it.
#define CPM_DECODE(k)                                                                        \
    do {                                                                                     \
        uint32_t pos = lane, t = 0, kv = 0;                                                  \
        bool valid = true;                                                                   \
        _Pragma("unroll") for (int c = 0; c <= D; ++c) {                                     \
            const uint32_t q = lane_next_bit(m0[c], m1[c], pos);                             \
            valid = valid && q < 128;                                                        \
            const uint32_t v0 = bperm(h0.hi[c], q & 63), v1 = bperm(h1.hi[c], q & 63);       \
            const uint32_t v = q < 64 ? v0 : v1;                                             \
            if (c < D)                                                                       \
                t += (v + 1) * stride[c];                                                    \
            else                                                                             \
                kv = v;                                                                      \
            pos = q + 1;                                                                     \
        }                                                                                    \
        st_t[k] = valid ? t : t_safe;                                                        \
        st_m[k] = min(pos, (uint32_t)kMe) | (valid ? kMv | kv << 16 : 0u);                   \
        const uint32_t ue = min(pos, 127u), ul = ue & 63;                                    \
        const uint32_t lo0 = bperm((uint32_t)h0.w, ul), hi0 = bperm((uint32_t)(h0.w >> 32), ul); \
        const uint32_t lo1 = bperm((uint32_t)h1.w, ul), hi1 = bperm((uint32_t)(h1.w >> 32), ul); \
        st_u[k] = ue < 64 ? ((uint64_t)hi0 << 32 | lo0) : ((uint64_t)hi1 << 32 | lo1);      \
    } while (0)

    // Full evaluation of one attempt per lane on the current lattice
    // (dynamics.rs:22-69, energy.rs:521-523): early-out, simple-point and
    // volume checks, the local dH in FP64 in the reference's order, and the
    // Metropolis test with u = gen::<f64>() of the draw after the attempt.
#define CPM_EVAL(IT_, IS_, IU_, IV_)                                                         \
    do {                                                                                     \
        const uint32_t et_ = (IT_), es_ = (IS_);                                                \
        const uint64_t eu_ = (IU_);                                                           \
        ev_Lw = sig.ld(et_);                                                                   \
        ev_Gw = sig.ld(es_);                                                                  \
        const uint32_t L = ev_Lw & kIdMask, G = ev_Gw & kIdMask;                             \
        const uint32_t lidx = ev_Lw >> 12, gidx = ev_Gw >> 12;                               \
        const bool nn = (IV_) && ev_Lw != ev_Gw;                                             \
        uint32_t nb[NT];                                                                     \
        _Pragma("unroll") for (int m = 0; m < NT; ++m) {                                     \
            const int code = m < NT / 2 ? m : m + 1;                                         \
            const int o0 = code % 3 - 1, o1 = (code / 3) % 3 - 1, o2 = code / 9 - 1;         \
            const int32_t a = D == 2 ? o0 * (int32_t)stride[0] + o1                          \
                                     : o0 * (int32_t)stride[0] + o1 * (int32_t)stride[1 % D] + o2; \
            nb[m] = sig.ld((int32_t)et_ + a);                                                  \
        }                                                                                    \
        ev_vL = vol[L ? L - 1 : 0];                                                          \
        ev_vG = vol[G ? G - 1 : 0];                                                          \
        ev_iL = itf[L ? L - 1 : 0];                                                          \
        ev_iG = itf[G ? G - 1 : 0];                                                          \
        uint32_t pat = 0;                                                                    \
        _Pragma("unroll") for (int m = 0; m < NT; ++m)                                       \
            pat |= (uint32_t)((nb[m] & kIdMask) == L) << m;                                  \
        const bool top = nn && L != 0;                                                       \
        bool simple;                                                                         \
        if (D == 2) {                                                                        \
            const uint64_t w = pat < 128 ? (pat < 64 ? sp2a : sp2b)                          \
                                         : (pat < 192 ? sp2c : sp2d);                        \
            simple = (w >> (pat & 63)) & 1ull;                                               \
        } else {                                                                             \
            simple = top ? ((sp3d[pat >> 6] >> (pat & 63)) & 1ull) : false;                  \
        }                                                                                    \
        const bool go = nn && (L == 0 || (simple && ev_vL > 1));                             \
        double dadh = 0.0; /* summed in stencil order */                                     \
        int cl = 0, cg = 0;                                                                  \
        _Pragma("unroll") for (int q = 0; q < NE; ++q) {                                     \
            const uint32_t w = nb[shell_index(D, NE, q)];                                    \
            const uint32_t nbe = w & kIdMask, ni = w >> 12;                                  \
            const double jl = J[lidx * T1 + ni], jg = J[gidx * T1 + ni];                     \
            dadh += 1.0 * ((nbe != G ? jg : 0.0) - (nbe != L ? jl : 0.0));                   \
            cl += nbe == L;                                                                  \
            cg += nbe == G;                                                                  \
        }                                                                                    \
        ev_dl = 2 * cl - NE;                                                                 \
        ev_dg = NE - 2 * cg;                                                                 \
        const uint32_t tl = lidx ? lidx - 1 : 0u, tg = gidx ? gidx - 1 : 0u;                 \
        double dvol = 0.0, ditf = 0.0;                                                       \
        if (L) {                                                                             \
            const double v = (double)ev_vL, T0 = (double)tv[tl];                             \
            const double x1 = (v - 1.0) - T0, x0 = v - T0;                                   \
            dvol += lv[tl] * (x1 * x1) - lv[tl] * (x0 * x0);                                 \
            const double i0 = (double)ev_iL, I0 = (double)ti[tl];                            \
            const double y1 = (i0 + (double)ev_dl) - I0, y0 = i0 - I0;                       \
            ditf += li[tl] * (y1 * y1) - li[tl] * (y0 * y0);                                 \
        }                                                                                    \
        if (G) {                                                                             \
            const double v = (double)ev_vG, T0 = (double)tv[tg];                             \
            const double x1 = (v + 1.0) - T0, x0 = v - T0;                                   \
            dvol += lv[tg] * (x1 * x1) - lv[tg] * (x0 * x0);                                 \
            const double i0 = (double)ev_iG, I0 = (double)ti[tg];                            \
            const double y1 = (i0 + (double)ev_dg) - I0, y0 = i0 - I0;                       \
            ditf += li[tg] * (y1 * y1) - li[tg] * (y0 * y0);                                 \
        }                                                                                    \
        ev_dh = dvol + ditf + dadh;                                                          \
        const bool needu = go && !(ev_dh <= 0.0);                                            \
        bool acc = go, tie = false;                                                          \
        if (needu) {                                                                         \
            const double u = (double)(eu_ >> 11) * (1.0 / 9007199254740992.0);                \
            acc = metropolis_accept(u, ev_dh, &tie);                                         \
        }                                                                                    \
        ev_f = (nn ? kMn : 0u) | (needu ? kMu : 0u) | (acc ? kMa : 0u) | (top ? kMp : 0u) |  \
               (go ? kMg : 0u) | (tie ? kMt : 0u);                                           \
    } while (0)

    // ---- the steps of the loop, shared by the single-wave and the
    // pipelined kernels.
    // index within the pass starting at q0 of the lane's attempt in
    // slot k (>= 64: not in it)
#define CPM_CIDX(k, q0)                                                                      \
    (((nnm[k] >> lane) & 1) &&                                                               \
             pre[k] + __builtin_amdgcn_mbcnt_hi((uint32_t)(nnm[k] >> 32),                    \
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)nnm[k], 0u)) >= (q0) \
         ? pre[k] + __builtin_amdgcn_mbcnt_hi((uint32_t)(nnm[k] >> 32),                      \
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)nnm[k], 0u)) - (q0) \
         : 64u)
    // An accepted attempt (slot 0, position commit_p): evaluated again on
    // every lane, then committed (dynamics.rs:73-224).
#define CPM_DO_COMMIT()                                                                         \
    do {                                                                                        \
        const uint32_t t = rl32(st_t[0], commit_p);                                             \
        uint32_t ev_Lw, ev_Gw, ev_vL, ev_vG, ev_iL, ev_iG, ev_f;                                \
        int32_t ev_dl, ev_dg;                                                                   \
        double ev_dh;                                                                           \
        const uint32_t s = t + rl32(coff, (rl32(st_m[0], commit_p) >> 16) & 31);                \
        CPM_EVAL(t, s, rl64(st_u[0], commit_p), true);                                          \
        (void)ev_f;                                                                             \
        const uint32_t Lw = uni(ev_Lw), Gw = uni(ev_Gw);                                        \
        const uint32_t L = Lw & kIdMask, G = Gw & kIdMask;                                      \
        const int32_t dl_i = (int32_t)uni((uint32_t)ev_dl);                                     \
        const int32_t dg_i = (int32_t)uni((uint32_t)ev_dg);                                     \
        const uint32_t vL = uni(ev_vL), vG = uni(ev_vG);                                        \
        const uint32_t iL = uni(ev_iL), iG = uni(ev_iG);                                        \
        const double dh = rlf64(ev_dh, 0);                                                      \
        uint32_t cval[D];                                                                       \
        {                                                                                       \
            uint32_t f = t;                                                                     \
        _Pragma("unroll")                                                                       \
            for (int d = 0; d < D; ++d) {                                                       \
                cval[d] = f / stride[d] - 1;                                                    \
                f %= stride[d];                                                                 \
            }                                                                                   \
        }                                                                                       \
        /* log the copy for the selective re-evaluation */                                      \
        {                                                                                       \
            bool edge = false;                                                                  \
        _Pragma("unroll")                                                                       \
            for (int d = 0; d < D; ++d) edge = edge || cval[d] == 0 || cval[d] + 1 == dims[d];  \
            if ((periodic && edge) || clog_n == kClog) {                                        \
                clog_all = true;                                                                \
            } else {                                                                            \
                if (lane == 0) {                                                                \
                    s_clog[clog_n][0] = cval[0] + 1;                                            \
                    s_clog[clog_n][1] = cval[1 % D] + 1;                                        \
                    s_clog[clog_n][2] = D == 3 ? cval[D - 1] + 1 : 0u;                          \
                    s_clog[clog_n][3] = L | G << 16;                                            \
                }                                                                               \
                ++clog_n;                                                                       \
            }                                                                                   \
        }                                                                                       \
        /* dynamics.rs:73-98 — contact counts, pre-move lattice */                              \
        if (lane < NE) {                                                                        \
            const uint32_t nbe = sig.ld((int32_t)t + eoff) & kIdMask;                           \
            if (nbe != L)                                                                       \
                atomicSub(&R.contact[(uint64_t)min(L, nbe) * ncell1 + max(L, nbe)], 1u);        \
            if (nbe != G)                                                                       \
                atomicAdd(&R.contact[(uint64_t)min(G, nbe) * ncell1 + max(G, nbe)], 1u);        \
        }                                                                                       \
        wsync();  /* every lane has read t's neighbourhood */                                   \
        if (lane == 0) {                                                                        \
            sig.st(t, Gw);                                                                      \
            if (L) {                                                                            \
                vol[L - 1] = vL - 1;                                                            \
                itf[L - 1] = (uint32_t)((int32_t)iL + dl_i);                                    \
            }                                                                                   \
            if (G) {                                                                            \
                vol[G - 1] = vG + 1;                                                            \
                itf[G - 1] = (uint32_t)((int32_t)iG + dg_i);                                    \
            }                                                                                   \
            /* dynamics.rs:125-183 — unwrapped position sums */                                 \
            if (L) {                                                                            \
                int64_t* sp = R.sum_pos + (uint64_t)(L - 1) * D;                                \
                int64_t* so = R.sum_outer + (uint64_t)(L - 1) * D * D;                          \
                for (int d = 0; d < D; ++d) sp[d] -= (int64_t)cval[d];                          \
                for (int d = 0; d < D; ++d)                                                     \
                    for (int e2 = 0; e2 < D; ++e2)                                              \
                        so[d * D + e2] -= (int64_t)cval[d] * (int64_t)cval[e2];                 \
            }                                                                                   \
            if (G) {                                                                            \
                int64_t* sp = R.sum_pos + (uint64_t)(G - 1) * D;                                \
                int64_t* so = R.sum_outer + (uint64_t)(G - 1) * D * D;                          \
                const uint32_t v = vG + 1;                                                      \
                const uint32_t vbg = max(v > 0 ? v - 1 : 0u, 1u);                               \
                double cen[D];                                                                  \
                for (int d = 0; d < D; ++d) cen[d] = (double)sp[d] / (double)vbg;               \
                int64_t img[D];                                                                 \
                for (int d = 0; d < D; ++d) {                                                   \
                    int64_t best = (int64_t)cval[d];                                            \
                    if (periodic) {                                                             \
                        double bd = fabs((double)best - cen[d]);                                \
                        const int64_t lo = (int64_t)cval[d] - (int64_t)dims[d];                 \
                        const int64_t hi = (int64_t)cval[d] + (int64_t)dims[d];                 \
                        const double dlo = fabs((double)lo - cen[d]);                           \
                        if (dlo < bd) {                                                         \
                            best = lo;                                                          \
                            bd = dlo;                                                           \
                        }                                                                       \
                        if (fabs((double)hi - cen[d]) < bd) best = hi;                          \
                    }                                                                           \
                    img[d] = best;                                                              \
                    sp[d] += best;                                                              \
                }                                                                               \
                for (int d = 0; d < D; ++d)                                                     \
                    for (int e2 = 0; e2 < D; ++e2) so[d * D + e2] += img[d] * img[e2];          \
            }                                                                                   \
            /* dynamics.rs:206-224 — rewrite t's halo mirrors */                                \
            if (periodic) {                                                                     \
                uint32_t opt[D][3], nopt[D], total = 1;                                         \
                for (int d = 0; d < D; ++d) {                                                   \
                    nopt[d] = 1;                                                                \
                    opt[d][0] = cval[d] + 1;                                                    \
                    if (cval[d] == 0) opt[d][nopt[d]++] = dims[d] + 1;                          \
                    if (cval[d] == dims[d] - 1) opt[d][nopt[d]++] = 0;                          \
                    total *= nopt[d];                                                           \
                }                                                                               \
                for (uint32_t q = 1; q < total; ++q) {                                          \
                    uint32_t r = q, f = 0;                                                      \
                    for (int d = D - 1; d >= 0; --d) {                                          \
                        f += opt[d][r % nopt[d]] * stride[d];                                   \
                        r /= nopt[d];                                                           \
                    }                                                                           \
                    sig.st(f, Gw);                                                              \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        if constexpr (SG) wg_fence();  /* the next evaluation reads sigma from HBM */          \
        wsync();                                                                                \
        energy += dh;                                                                           \
        /* the lattice changed: every later attempt is evaluated again */                       \
        need_eval = true;                                                                       \
    } while (0)
    // Decode a block of nk windows from `base` (the lattice plays no part).
#define CPM_DO_BLOCK(nk_)                                                                       \
    do {                                                                                        \
        const uint32_t nkb = (nk_);                                                                 \
        if (!hr_ready) {                                                                        \
        _Pragma("unroll")                                                                       \
            for (int j = 0; j <= KB; ++j)                                                       \
                hr[j] = base + 64 * j + 64 <= kFill ? draws[base + 64 * j + lane] : 0ull;       \
        }                                                                                       \
        Half h0, h1;                                                                            \
        uint64_t m0[D + 1], m1[D + 1];                                                          \
        prep_half<D>(h0, hr[0], range, zone, m0);                                               \
        _Pragma("unroll")                                                                       \
        for (int k = 0; k < KB; ++k) {                                                          \
            if (k < (int)nkb) {                                                                  \
                prep_half<D>(h1, hr[k + 1], range, zone, m1);                                   \
                CPM_DECODE(k);                                                                  \
                h0 = h1;                                                                        \
        _Pragma("unroll")                                                                       \
                for (int c = 0; c <= D; ++c) m0[c] = m1[c];                                     \
            }                                                                                   \
        }                                                                                       \
        nvalid = nkb;                                                                           \
        /* the next block's draws, in flight while this block is walked */                      \
        hr_ready = nkb == (uint32_t)KB;                                                          \
        if (hr_ready) {                                                                         \
            hr[0] = hr[KB];                                                                     \
        _Pragma("unroll")                                                                       \
            for (int j = 1; j <= KB; ++j) {                                                     \
                const uint32_t at = base + 64 * (KB + j);                                       \
                hr[j] = at + 64 <= kFill ? draws[at + lane] : 0ull;                             \
            }                                                                                   \
        }                                                                                       \
        need_block = false;                                                                     \
        need_eval = true;                                                                       \
    } while (0)
    // Evaluate the block's non-null attempts, 64 per pass, packed onto
    // consecutive lanes through the LDS stage.
#define CPM_DO_EVAL()                                                                           \
    do {                                                                                        \
        uint64_t nnm[KB];                                                                       \
        uint32_t pre[KB + 1], src[KB];                                                          \
        pre[0] = 0;                                                                             \
        _Pragma("unroll")                                                                       \
        for (int k = 0; k < KB; ++k) {                                                          \
            bool nn = false;                                                                    \
            /* source site: target + offset of the copy index (a void */                        \
            /* decode keeps its interior stand-in as both) */                                   \
            src[k] = st_t[k] + ((st_m[k] & kMv) ? bperm(coff, (st_m[k] >> 16) & 31) : 0u);      \
            bool want = false;                                                                  \
            if (k < (int)nvalid) {                                                              \
                const uint32_t Lw = sig.ld(st_t[k]), Gw = sig.ld(src[k]);                       \
                nn = (st_m[k] & kMv) && Lw != Gw && (k > 0 || lane >= p);                       \
                want = nn;                                                                      \
                if (eval_sel && nn && (st_m[k] & kMn) && !clog_all) {                           \
                    /* still non-null and evaluated before: again only if a */                  \
                    /* logged copy lies in its 3^D box or changed its cells */                  \
                    const uint32_t Li = Lw & kIdMask, Gi = Gw & kIdMask;                        \
                    const uint32_t c0 = st_t[k] / stride[0], rem = st_t[k] - c0 * stride[0];    \
                    const uint32_t c1 = D == 3 ? rem / stride[1 % D] : rem;                     \
                    const uint32_t c2 = D == 3 ? rem - c1 * stride[1 % D] : 0u;                 \
                    bool aff = false;                                                           \
                    for (uint32_t j = 0; j < clog_n; ++j) {                                     \
                        const uint32_t cells = s_clog[j][3];                                    \
                        const uint32_t cl = cells & 0xFFFF, cg = cells >> 16;                   \
                        const bool box = c0 + 1 - s_clog[j][0] <= 2u &&                         \
                                         c1 + 1 - s_clog[j][1] <= 2u &&                         \
                                         (D == 2 || c2 + 1 - s_clog[j][2] <= 2u);               \
                        aff = aff || box || (Li && (Li == cl || Li == cg)) ||                   \
                              (Gi && (Gi == cl || Gi == cg));                                   \
                    }                                                                           \
                    want = aff;                                                                 \
                }                                                                               \
            }                                                                                   \
            if (want || !nn)                                                                    \
                st_m[k] = (st_m[k] & (kMe | kMv | 0x1F0000u)) | (nn ? kMn : 0u);                \
            nnm[k] = uni64(__ballot(want));                                                     \
            pre[k + 1] = pre[k] + (uint32_t)__popcll(nnm[k]);                                   \
        }                                                                                       \
        const uint32_t total = pre[KB];                                                         \
        for (uint32_t q0 = 0; q0 < total; q0 += 64) {                                           \
        _Pragma("unroll")                                                                       \
            for (int k = 0; k < KB; ++k) {                                                      \
                const uint32_t c = CPM_CIDX(k, q0);                                             \
                if (pre[k + 1] > q0 && pre[k] < q0 + 64 && c < 64)                              \
                    stage[c] = make_uint4(st_t[k], src[k], (uint32_t)st_u[k],                   \
                                          (uint32_t)(st_u[k] >> 32));                           \
            }                                                                                   \
            wsync();                                                                            \
            const uint32_t n_here = min(total - q0, 64u);                                       \
            uint4 in = make_uint4(t_safe, t_safe, 0u, 0u);                                      \
            if (lane < n_here) in = stage[lane];                                                \
            wsync();  /* the stage is rewritten by the next pass */                             \
            uint32_t ev_Lw, ev_Gw, ev_vL, ev_vG, ev_iL, ev_iG, ev_f;                            \
            int32_t ev_dl, ev_dg;                                                               \
            double ev_dh;                                                                       \
            CPM_EVAL(in.x, in.y, (uint64_t)in.w << 32 | in.z, lane < n_here);                   \
        _Pragma("unroll")                                                                       \
            for (int k = 0; k < KB; ++k) {                                                      \
                if (pre[k + 1] > q0 && pre[k] < q0 + 64) {                                      \
                    const uint32_t c = CPM_CIDX(k, q0);                                         \
                    const uint32_t f = bperm(ev_f, c & 63);                                     \
                    if (c < 64) st_m[k] |= f;                                                   \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
    } while (0)
    // Walk the chain through slot 0 up to the first accepted copy, the
    // window's end or the MCS target, in one step: lane k finds the k-th
    // attempt position x = next^k(p) by pointer doubling.  An attempt takes
    // at least D + 1 >= 3 draws, so at most 22 attempts start in a 64-draw
    // window and 5 doublings suffice.  walk_hit: it stopped at an accepted
    // copy (now pending, with p past it).
#define CPM_DO_WALK()                                                                           \
    do {                                                                                        \
        const uint32_t m_own = st_m[0];                                                         \
        {                                                                                       \
            const uint32_t e = m_own & kMe, u1 = (m_own & kMu) ? 1u : 0u;                       \
            const bool okb = (m_own & kMv) && (!u1 || e < 128);                                 \
            uint32_t jb = okb ? e + u1 : 127u;  /* next^(2^b); >= 64 = leaves */                \
            uint32_t x = p;                                                                     \
        _Pragma("unroll")                                                                       \
            for (int b = 0; b < 5; ++b) {                                                       \
                const uint32_t y = bperm(jb, x & 63);                                           \
                if (((lane >> b) & 1) && x < 64) x = y;                                         \
                if (b < 4) {                                                                    \
                    const uint32_t z = bperm(jb, jb & 63);                                      \
                    if (jb < 64) jb = z;                                                        \
                }                                                                               \
            }                                                                                   \
            const uint32_t fx = bperm(m_own, x & 63);  /* meta of the lane's attempt */         \
            const bool okx = (fx & kMv) && (!(fx & kMu) || (fx & kMe) < 128);                   \
            const bool vk = lane < 32 && x < 64 && okx;                                         \
            const uint32_t k_inv = (uint32_t)__builtin_ctzll(uni64(__ballot(!vk)));             \
            const uint64_t acc_k = uni64(__ballot(vk && (fx & kMa)));                           \
            const uint32_t k_acc = acc_k ? (uint32_t)__builtin_ctzll(acc_k) : 64u;              \
            /* attempts left before the MCS target: (todo - 1) * n_sites + (n_sites - ain) */   \
            const uint64_t left = (todo - 1) * (uint64_t)n_sites + (n_sites - ain);             \
            uint32_t n_att = min(k_inv, left < 64 ? (uint32_t)left : 64u);                      \
            const bool hit = k_acc < n_att;                                                     \
            if (hit) n_att = k_acc + 1;                                                         \
            {                                                                                   \
                const bool w = lane < n_att;                                                    \
                c_att += n_att;                                                                 \
                c_nn += __popcll(uni64(__ballot(w && (fx & kMn))));                             \
                c_top += __popcll(uni64(__ballot(w && (fx & kMp))));                            \
                c_en += __popcll(uni64(__ballot(w && (fx & kMg))));                             \
                c_acc += hit ? 1u : 0u;                                                         \
                c_tie += __popcll(uni64(__ballot(w && (fx & kMt))));                            \
            }                                                                                   \
            /* monte_carlo.rs:330-356 — MCS accounting */                                       \
            ain += n_att;                                                                       \
            while (ain >= n_sites) {                                                            \
                ain -= n_sites;                                                                 \
                ++mcs;                                                                          \
                --todo;                                                                         \
            }                                                                                   \
            const uint32_t np = rl32(x, n_att);  /* the position after the walked attempts */   \
            p = np;                                                                             \
            walk_hit = hit;                                                                     \
            if (hit) {                                                                          \
                commit_p = rl32(x, k_acc);                                                      \
                pending = true;                                                                 \
            }                                                                                   \
        }                                                                                       \
    } while (0)
    // Slot 0 is walked: the next window becomes slot 0.
#define CPM_SHIFT_SLOTS()                                                                       \
    do {                                                                                        \
        _Pragma("unroll")                                                                       \
        for (int k = 0; k + 1 < KB; ++k) {                                                      \
            st_t[k] = st_t[k + 1];                                                              \
            st_m[k] = st_m[k + 1];                                                              \
            st_u[k] = st_u[k + 1];                                                              \
        }                                                                                       \
        base += 64;                                                                             \
        p -= 64;                                                                                \
        --nvalid;                                                                               \
    } while (0)
    // Pipelined producer: decode and evaluate the block at pb into block
    // buffer b (none when the stream buffer must first be refilled and the
    // walker may still need it).
#define CPM_PRODUCE(b_, idle_)                                                                   \
    do {                                                                                         \
        const uint32_t bb = (b_);                                                                  \
        const bool walker_idle = (idle_);                                                        \
        uint32_t room = kFill - pb;                                                              \
        uint32_t nk = room >= 128 ? min((uint32_t)KB, (room - 64) / 64) : 0u;                    \
        if (nk == 0 && walker_idle) {                                                            \
            CPM_REFILL(pb);                                                                      \
            pb = base;                                                                           \
            room = kFill - pb;                                                                   \
            nk = min((uint32_t)KB, (room - 64) / 64);                                            \
        }                                                                                        \
        if (nk) {                                                                                \
            base = pb;                                                                           \
            p = 0;                                                                               \
            CPM_DO_BLOCK(nk);                                                                    \
            eval_sel = false;                                                                    \
            CPM_DO_EVAL();                                                                       \
            need_eval = false;                                                                   \
        _Pragma("unroll")                                                                        \
            for (int k = 0; k < KB; ++k)                                                         \
                blk[(bb * KB + k) * 64 + lane] =                                                  \
                    make_uint4(st_t[k], st_m[k], (uint32_t)st_u[k], (uint32_t)(st_u[k] >> 32));  \
            pb += 64 * nk;                                                                       \
        }                                                                                        \
        if (lane == 0) {                                                                         \
            s_nv[bb] = nk;                                                                        \
            s_base[bb] = base;                                                                    \
        }                                                                                        \
    } while (0)

    if constexpr (!PW) {
        for (;;) {
            if (pending) {
                pending = false;
                dirty = true;
                CPM_DO_COMMIT();
                eval_sel = true;
            }
            if (todo == 0) break;
            if (need_block) {
                const uint32_t room = kFill - base;
                const uint32_t nk = room >= 128 ? min((uint32_t)KB, (room - 64) / 64) : 0u;
                if (nk == 0) {
                    CPM_REFILL(base + p);
                    continue;
                }
                CPM_DO_BLOCK(nk);
                clog_n = 0;
                clog_all = false;
                eval_sel = false;
            }
            if (need_eval) {
                need_eval = false;
                CPM_DO_EVAL();
            }
            CPM_DO_WALK();
            if (walk_hit) continue;
            if (todo == 0) break;
            if constexpr (TB) {
                if (out_of_time(s_budget, t_begin)) break;
            }
            if (p >= 64) {
                CPM_SHIFT_SLOTS();
                if (nvalid == 0) need_block = true;
                continue;
            }
            // The attempt at p does not fit in its window: restart the block
            // at it.  An attempt needing more than 128 draws even from a
            // window of its own is reported, never looped on.
            if (p == 0) {
                error = true;
                break;
            }
            base += p;
            p = 0;
            hr_ready = false;
            need_block = true;
        }
    } else {
        // Pipelined: wave 1 decodes and evaluates block i + 1 while wave 0
        // walks block i, one workgroup barrier per block.  The walker
        // commits as above and evaluates again the attempts its own commits
        // can have changed (in the rest of its block, and in the next block
        // for the copies it made while that one was being evaluated).  Shared words, by iteration
        // parity: the walker's command for the next iteration (go on,
        // restart at s_pos, done) and whether it committed; each block
        // buffer's start and window count.  The stream buffer is refilled
        // only while the walker has no block, so the walker's final position
        // always lies in the current buffer.
        __shared__ uint32_t s_cmd[2], s_pos[2], s_cmt[2], s_nv[2], s_base[2];
        constexpr uint32_t kGo = 0, kRestart = 1, kDone = 2;
        uint4* const blk = (uint4*)(smem + Ly.total + 2 * kStageBytes);  // [2][KB][64]
        uint32_t pb = 0;  // producer: where its next block starts
        if (wave == 1) {
            CPM_PRODUCE(0u, true);
        } else if (lane == 0) {
            s_cmd[0] = kGo;
            s_cmt[0] = 0;
        }
        __syncthreads();
        for (uint32_t it = 0;; ++it) {
            const uint32_t cur = it & 1, nxt = cur ^ 1;
            const uint32_t cmd = s_cmd[cur];
            if (cmd == kDone) break;
            if (wave == 1) {
                if (cmd == kRestart) {
                    pb = s_pos[cur];
                    hr_ready = false;
                }
                if constexpr (SG) {
                    if (s_cmt[cur]) wg_fence();  // sigma words the walker rewrote
                }
                CPM_PRODUCE(nxt, cmd == kRestart || s_nv[cur] == 0);
            } else {
                uint32_t next_cmd = kGo, next_pos = 0;
                bool committed = false, timed_out = false;
                if (cmd != kRestart && s_nv[cur] > 0) {
                    nvalid = s_nv[cur];
                    base = s_base[cur];
#pragma unroll
                    for (int k = 0; k < KB; ++k) {
                        const uint4 v = blk[(cur * KB + k) * 64 + lane];
                        st_t[k] = v.x;
                        st_m[k] = v.y;
                        st_u[k] = (uint64_t)v.w << 32 | v.z;
                    }
                    // the producer evaluated this block while the walker made
                    // the logged copies: evaluate again what they touched
                    eval_sel = true;
                    if (clog_n > 0 || clog_all) CPM_DO_EVAL();
                    clog_n = 0;
                    clog_all = false;
                    for (;;) {
                        if (pending) {
                            pending = false;
                            dirty = true;
                            committed = true;
                            CPM_DO_COMMIT();
                        }
                        if (todo == 0) break;
                        if (need_eval) {
                            need_eval = false;
                            CPM_DO_EVAL();
                        }
                        CPM_DO_WALK();
                        if (walk_hit) continue;
                        if (todo == 0) break;
                        if constexpr (TB) {
                            if (out_of_time(s_budget, t_begin)) {
                                timed_out = true;
                                break;
                            }
                        }
                        if (p >= 64) {
                            CPM_SHIFT_SLOTS();
                            if (nvalid == 0) break;
                            continue;
                        }
                        if (p == 0) {
                            error = true;
                            break;
                        }
                        next_cmd = kRestart;
                        next_pos = base + p;
                        p = 0;
                        break;
                    }
                } else {
                    clog_n = 0;  // nothing changed while this iteration's block was made
                    clog_all = false;
                }
                if (todo == 0 || error || timed_out) next_cmd = kDone;
                if (lane == 0) {
                    s_cmd[nxt] = next_cmd;
                    s_pos[nxt] = next_pos;
                    s_cmt[nxt] = committed ? 1u : 0u;
                }
            }
            __syncthreads();
        }
        if (wave == 1) return;
        wg_fence();  // the producer's last refill of the segment states
    }
    base += p;  // buffer position of the next unconsumed draw
#undef CPM_DECODE
#undef CPM_EVAL
#undef CPM_REFILL
#undef CPM_CIDX
#undef CPM_DO_COMMIT
#undef CPM_DO_BLOCK
#undef CPM_DO_EVAL
#undef CPM_DO_WALK
#undef CPM_SHIFT_SLOTS
#undef CPM_PRODUCE
    wsync();
    // the consumer's RNG state: the segment's start state (a vector load:
    // it was written by vector stores) stepped to the cursor
    const uint32_t g_end = min(base / kSegLen, kSegs - 1);
    const uint64_t seg_end = segs[g_end * 4 + (lane & 3)];
    const uint64_t st_end[4] = {rl64(seg_end, 0), rl64(seg_end, 1), rl64(seg_end, 2),
                                rl64(seg_end, 3)};
    if (dirty) {
        for (uint32_t i = lane; i < n_padded; i += 64) R.sigma[i] = sig.ld(i) & kIdMask;
        for (uint32_t i = lane; i < n_cells; i += 64) {
            R.volume[i] = vol[i];
            R.interface[i] = itf[i];
        }
    }
    if (lane == 0) {
        R.todo = todo;
        R.attempt_in_mcs = ain;
        R.mcs = mcs;
        R.energy = energy;
        R.cnt[0] += c_att;
        R.cnt[1] += c_nn;
        R.cnt[2] += c_top;
        R.cnt[3] += c_en;
        R.cnt[4] += c_acc;
        R.cnt[5] += c_tie;
        R.t_run = __builtin_amdgcn_s_memrealtime() - t_begin;
        R.done = 1;
        if (error) {
            R.error = 1;
        } else {
            Xoshiro st{{st_end[0], st_end[1], st_end[2], st_end[3]}};
            for (uint32_t i = g_end * kSegLen; i < base; ++i) xo_next(st);
            for (int k = 0; k < 4; ++k) R.rng[k] = st.s[k];
        }
    }
}

// key = 4000 * (time-limited) + 2000 * (pipelined) + 1000 * (sigma in HBM) + 100 * D +
// energy stencil size
#define CPM_WAVE_KERNELS_PW(X, pw, tb, k0)                                             \
    X(k0 + 204, 2, 4, false, pw, tb) X(k0 + 208, 2, 8, false, pw, tb)                  \
    X(k0 + 306, 3, 6, false, pw, tb) X(k0 + 318, 3, 18, false, pw, tb)                 \
    X(k0 + 326, 3, 26, false, pw, tb) X(k0 + 1204, 2, 4, true, pw, tb)                 \
    X(k0 + 1208, 2, 8, true, pw, tb) X(k0 + 1306, 3, 6, true, pw, tb)                  \
    X(k0 + 1318, 3, 18, true, pw, tb) X(k0 + 1326, 3, 26, true, pw, tb)
#define CPM_WAVE_KERNELS(X)                                                              \
    CPM_WAVE_KERNELS_PW(X, false, false, 0) CPM_WAVE_KERNELS_PW(X, true, false, 2000)  \
    CPM_WAVE_KERNELS_PW(X, false, true, 4000) CPM_WAVE_KERNELS_PW(X, true, true, 6000)

// ===========================================================================
// Sublattice-parallel mode: halo-disjoint checkerboard colours updated in
// parallel.  An attempt at t reads only t's 3^D neighbourhood and writes
// only sigma[t]; sites of one colour (c_d mod 2 on every axis) are at
// Chebyshev distance >= 2, so within a colour phase no site reads what
// another writes (periodic sides are even, so the wrap keeps the
// colouring), and topology checks stay exact.
//
// Schedule: an MCS is kSlSub sub-sweeps; a sub-sweep visits the 2^D
// colours in a random order, and each site of a colour is the target of an
// attempt with probability 1/kSlSub, so a site is attempted Binomial(kSlSub,
// 1/kSlSub) times per MCS (the reference's uniform target choice gives
// about Poisson(1)), n_sites attempts per MCS on average.
//
// Per-cell V and I are exact: attempts of a colour phase run in batches of
// one site per thread; within a batch, the attempts that reach the volume
// check (non-null, simple point) commit in descending order of a random
// priority among those sharing a cell — each round, an attempt proceeds if
// it holds the highest priority on both of its cells (an LDS atomicMax per
// cell), so every decision sees V and I exactly as a sequential pass in
// that order would.  Outcomes depend only on the counter-based random
// numbers (SplitMix64 of replica key, MCS, sub-sweep, colour, site and
// draw), not on thread timing.
// ===========================================================================
__device__ __forceinline__ uint64_t splitmix_at(uint64_t key, uint64_t ctr) {
    uint64_t z = key + ctr * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

constexpr uint32_t kSlThreads = 256;
// Lattices with at least kSlBigSites sites (or whose sigma does not fit in
// LDS) run on workgroups of kSlThreadsBig threads, one per replica, smaller
// ones on workgroups of kSlThreads; a batch is one site per thread.  The
// choice depends on the replica alone, so its trajectory does not depend on
// what else is in the batch.
constexpr uint32_t kSlThreadsBig = 512;
constexpr uint32_t kSlSitesPerThread = 16;  // sites per thread per batch, big workgroups
constexpr uint64_t kSlBigSites = 16384;
// the sublattice kernels: (key, D, sigma in HBM, threads)
#define CPM_SL_KERNELS(X)                                                                    \
    X(200, 2, false, kSlThreads) X(300, 3, false, kSlThreads)                                \
    X(210, 2, false, kSlThreadsBig) X(211, 2, true, kSlThreadsBig)                           \
    X(310, 3, false, kSlThreadsBig) X(311, 3, true, kSlThreadsBig)
constexpr uint32_t kSlSub = 4;  // sub-sweeps per MCS (a power of two)

// Dynamic LDS of the sublattice kernel: the exact-replay layout plus one
// 64-bit priority word per cell.
__host__ __device__ inline uint32_t sl_prio_offset(const LdsLayout& Ly) { return Ly.total; }
__host__ __device__ inline uint32_t sl_lds_bytes(const LdsLayout& Ly, uint32_t n_cells) {
    return Ly.total + (n_cells * 8 + 15) / 16 * 16;
}

// NT threads per workgroup; SG: sigma in HBM (the packed copy) with the
// per-cell tables and priorities still in LDS.
template <int D, bool SG, uint32_t NT>
__global__ void __launch_bounds__(NT) k_sublattice(Rep* __restrict__ reps,
                                                   const uint32_t* __restrict__ idx,
                                                   const uint64_t* __restrict__ sp3d) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned long long s_cnt[6];
    Rep& R = reps[idx[blockIdx.x]];
    const uint32_t tid = threadIdx.x;
    const uint32_t T = R.n_types, T1 = T + 1, n_cells = R.n_cells;
    const uint32_t n_padded = R.n_padded;
    const LdsLayout Ly = lds_layout(n_padded, n_cells, T, !SG);
    SigmaRef<SG> sig;  // packed: id | (type + 1) << 12
    if constexpr (SG)
        sig.p = (__attribute__((address_space(1))) uint16_t*)R.sig16;
    else
        sig.p = (uint16_t*)(smem + Ly.sigma);
    uint32_t* vol = (uint32_t*)(smem + Ly.vol);
    uint32_t* itf = (uint32_t*)(smem + Ly.itf);
    double* J = (double*)(smem + Ly.J);
    double* lv = (double*)(smem + Ly.lv);
    double* li = (double*)(smem + Ly.li);
    uint32_t* tv = (uint32_t*)(smem + Ly.tv);
    uint32_t* ti = (uint32_t*)(smem + Ly.ti);
    unsigned long long* prio = (unsigned long long*)(smem + sl_prio_offset(Ly));
    const uint32_t* __restrict__ gtyp = R.type_index;
    for (uint32_t i = tid; i < n_padded; i += NT) {
        const uint32_t id = R.sigma[i];
        sig.st(i, id ? id | ((gtyp[id - 1] + 1) << 12) : 0u);
    }
    for (uint32_t i = tid; i < n_cells; i += NT) {
        vol[i] = R.volume[i];
        itf[i] = R.interface[i];
        prio[i] = 0;
    }
    for (uint32_t i = tid; i < T1 * T1; i += NT) J[i] = R.params[2 * T + i];
    for (uint32_t i = tid; i < T; i += NT) {
        lv[i] = R.params[i];
        li[i] = R.params[T + i];
        tv[i] = R.targets[i];
        ti[i] = R.targets[T + i];
    }
    if (tid < 6) s_cnt[tid] = 0;
    if constexpr (SG) wg_fence();  // the staged sigma words, HBM
    __syncthreads();

    uint32_t dims[D], stride[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        dims[d] = R.dims[d];
        stride[d] = R.strides[d];
    }
    const uint32_t n_copy = R.n_copy, n_energy = R.n_energy;
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;
    const uint64_t key = R.fill_state[0];
    unsigned long long c_att = 0, c_nn = 0, c_top = 0, c_en = 0, c_acc = 0;
    constexpr uint32_t NC = 1u << D;
    // sites per thread in a batch: several on the big workgroups, which
    // amortise each round's barriers and commit latency over more attempts
    constexpr uint32_t KS = NT == kSlThreadsBig ? kSlSitesPerThread : 1u;

    for (uint64_t m = 0; m < R.todo; ++m) {
        const uint64_t mcs = R.mcs + m;
        for (uint32_t sub = 0; sub < kSlSub; ++sub) {
            const uint64_t phase0 = (mcs * kSlSub + sub) * NC;
            // colour order of this sub-sweep: Fisher-Yates from one hash
            uint32_t order[NC];
            for (uint32_t c = 0; c < NC; ++c) order[c] = c;
            uint64_t hperm = splitmix_at(key, ~phase0);
            for (uint32_t c = NC - 1; c > 0; --c) {
                const uint32_t j = (uint32_t)((hperm & 0xffff) % (c + 1));
                hperm >>= 16;
                const uint32_t tmp = order[c];
                order[c] = order[j];
                order[j] = tmp;
            }
            for (uint32_t ci = 0; ci < NC; ++ci) {
                const uint32_t col = order[ci];
                const uint64_t phase = phase0 + col;
                uint32_t half[D], nsite = 1;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const uint32_t cd = (col >> d) & 1u;
                    half[d] = dims[d] > cd ? (dims[d] - cd + 1) / 2 : 0;
                    nsite *= half[d];
                }
                if (nsite == 0) continue;
                // batches start at a random rotation of the colour's sites
                const uint32_t rot = (uint32_t)(splitmix_at(key ^ 0x5bd1e995ull, phase) % nsite);
                // a colour site's coordinates and padded flat index
                auto site_of = [&](uint32_t j, uint32_t* c, uint32_t& t) {
                    uint32_t rem = j;
                    t = 0;
#pragma unroll
                    for (int d = D - 1; d >= 0; --d) {
                        c[d] = 2 * (rem % half[d]) + ((col >> d) & 1u);
                        rem /= half[d];
                        t += (c[d] + 1) * stride[d];
                    }
                };
                // a winning attempt: volume check, dH, Metropolis, commit
                // (dynamics.rs:23-224); no other winner of its round shares
                // a cell with it or a site of its neighbourhood
                auto commit_win = [&](uint32_t j, uint32_t Lw, uint32_t Gw) {
                    const uint32_t L = Lw & kIdMask, G = Gw & kIdMask;
                    uint32_t c[D], t;
                    site_of(j, c, t);
                    const uint64_t ctr = (phase * (uint64_t)R.n_sites + j) * 4;
                    const uint32_t vL = L ? vol[L - 1] : 0u;
                    if (L == 0 || vL > 1) {
                        c_en++;
                        const uint32_t lidx = Lw >> 12, gidx = Gw >> 12;
                        const uint32_t tl = lidx ? lidx - 1 : 0u, tg = gidx ? gidx - 1 : 0u;
                        int nl = 0, ng = 0;
                        double dadh = 0.0;  // energy.rs:242-264, stencil order
                        // the stencil's words, all reads in flight together
                        uint32_t ew[26];
#pragma unroll
                        for (uint32_t q = 0; q < 26; ++q)
                            ew[q] = q < n_energy ? sig.ld((int32_t)t + R.energy_off[q]) : 0u;
#pragma unroll
                        for (uint32_t q = 0; q < 26; ++q) {
                            if (q >= n_energy) break;
                            const uint32_t w = ew[q];
                            const uint32_t nb = w & kIdMask, ni = w >> 12;
                            nl += nb == L;
                            ng += nb == G;
                            const double before = nb != L ? J[lidx * T1 + ni] : 0.0;
                            const double after = nb != G ? J[gidx * T1 + ni] : 0.0;
                            dadh += 1.0 * (after - before);
                        }
                        const int dl_i = 2 * nl - (int)n_energy, dg_i = (int)n_energy - 2 * ng;
                        double dvol = 0.0, ditf = 0.0;
                        if (L) {
                            const double v = (double)vL;
                            const double x1 = (v - 1.0) - (double)tv[tl], x0 = v - (double)tv[tl];
                            dvol += lv[tl] * (x1 * x1) - lv[tl] * (x0 * x0);
                            const double i0 = (double)itf[L - 1];
                            const double y1 = (i0 + (double)dl_i) - (double)ti[tl], y0 = i0 - (double)ti[tl];
                            ditf += li[tl] * (y1 * y1) - li[tl] * (y0 * y0);
                        }
                        if (G) {
                            const double v = (double)vol[G - 1];
                            const double x1 = (v + 1.0) - (double)tv[tg], x0 = v - (double)tv[tg];
                            dvol += lv[tg] * (x1 * x1) - lv[tg] * (x0 * x0);
                            const double i0 = (double)itf[G - 1];
                            const double y1 = (i0 + (double)dg_i) - (double)ti[tg], y0 = i0 - (double)ti[tg];
                            ditf += li[tg] * (y1 * y1) - li[tg] * (y0 * y0);
                        }
                        const double dh = dvol + ditf + dadh;
                        bool accept = dh <= 0.0;
                        if (!accept) {
                            const uint64_t hu = splitmix_at(key, ctr + 2);
                            const double u = (double)(hu >> 11) * (1.0 / 9007199254740992.0);
                            accept = u < exp(-dh);
                        }
                        if (accept) {
                            c_acc++;
                            // (the contact counts are recounted from the lattice
                            // after the call: k_contact_count)
                            sig.st(t, Gw);
                            // this round owns L's and G's V, I and position sums
                            if (L) {
                                vol[L - 1] = vL - 1;
                                itf[L - 1] = (uint32_t)((int32_t)itf[L - 1] + dl_i);
                                // no other winner of this round touches cell L
                                // or G: the sums are updated without waiting
                                unsigned long long* sp =
                                    (unsigned long long*)(R.sum_pos + (uint64_t)(L - 1) * D);
                                unsigned long long* so =
                                    (unsigned long long*)(R.sum_outer + (uint64_t)(L - 1) * D * D);
                                for (int d = 0; d < D; ++d)
                                    atomicAdd(&sp[d], (unsigned long long)(-(int64_t)c[d]));
                                for (int d = 0; d < D; ++d)
                                    for (int e2 = 0; e2 < D; ++e2)
                                        atomicAdd(&so[d * D + e2],
                                                  (unsigned long long)(-((int64_t)c[d] * (int64_t)c[e2])));
                            }
                            if (G) {
                                const uint32_t vg_new = vol[G - 1] + 1;
                                vol[G - 1] = vg_new;
                                itf[G - 1] = (uint32_t)((int32_t)itf[G - 1] + dg_i);
                                int64_t* sp = R.sum_pos + (uint64_t)(G - 1) * D;
                                unsigned long long* so =
                                    (unsigned long long*)(R.sum_outer + (uint64_t)(G - 1) * D * D);
                                const uint32_t vbg = max(vg_new - 1, 1u);
                                int64_t img[D];
                                for (int d = 0; d < D; ++d) {
                                    // an atomic (device-coherent) read: the sums
                                    // are updated with atomics by other threads
                                    const double cen =
                                        (double)__hip_atomic_load(&sp[d], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT) /
                                        (double)vbg;
                                    int64_t best = (int64_t)c[d];
                                    if (periodic) {
                                        double bd = fabs((double)best - cen);
                                        const int64_t lo = (int64_t)c[d] - (int64_t)dims[d];
                                        const int64_t hi = (int64_t)c[d] + (int64_t)dims[d];
                                        const double dlo = fabs((double)lo - cen);
                                        if (dlo < bd) {
                                            best = lo;
                                            bd = dlo;
                                        }
                                        if (fabs((double)hi - cen) < bd) best = hi;
                                    }
                                    img[d] = best;
                                }
                                for (int d = 0; d < D; ++d)
                                    atomicAdd((unsigned long long*)&sp[d], (unsigned long long)img[d]);
                                for (int d = 0; d < D; ++d)
                                    for (int e2 = 0; e2 < D; ++e2)
                                        atomicAdd(&so[d * D + e2], (unsigned long long)(img[d] * img[e2]));
                            }
                            if (periodic) {  // t's halo mirrors
                                uint32_t opt[D][3], nopt[D], total = 1;
                                for (int d = 0; d < D; ++d) {
                                    nopt[d] = 1;
                                    opt[d][0] = c[d] + 1;
                                    if (c[d] == 0) opt[d][nopt[d]++] = dims[d] + 1;
                                    if (c[d] == dims[d] - 1) opt[d][nopt[d]++] = 0;
                                    total *= nopt[d];
                                }
                                for (uint32_t q = 1; q < total; ++q) {
                                    uint32_t r = q, f = 0;
                                    for (int d = D - 1; d >= 0; --d) {
                                        f += opt[d][r % nopt[d]] * stride[d];
                                        r /= nopt[d];
                                    }
                                    sig.st(f, Gw);
                                }
                            }
                        }
                    }
                };
                // batches of KS sites per thread (NT * KS sites)
                for (uint32_t b0 = 0; b0 < nsite; b0 += NT * KS) {
                    uint32_t pj[KS], pw[KS];  // site index; packed words L | G << 16
                    unsigned long long ppr[KS];
                    uint32_t pmask = 0;  // bit q: slot q holds a pending attempt
                    // the thread's attempts of this batch: targets and sources
                    // first, then every lattice read of the null checks at
                    // once (one memory round trip, not one per attempt)
                    uint32_t tq[KS], sq[KS], lq[KS], gq[KS];
                    uint32_t amask = 0;  // bit q: slot q is attempted
#pragma unroll
                    for (uint32_t q = 0; q < KS; ++q) {
                        pj[q] = 0;
                        pw[q] = 0;
                        ppr[q] = 0;
                        tq[q] = 0;
                        sq[q] = 0;
                        const uint32_t jb = b0 + q * NT + tid;
                        if (jb >= nsite) continue;
                        uint32_t j = jb + rot;
                        if (j >= nsite) j -= nsite;
                        uint32_t c[D], t;
                        site_of(j, c, t);
                        const uint64_t ctr = (phase * (uint64_t)R.n_sites + j) * 4;
                        const uint64_t h0 = splitmix_at(key, ctr);
                        if ((h0 & (kSlSub - 1)) != 0) continue;  // this site is not attempted
                        const uint32_t k = (uint32_t)(((h0 >> 32) * (uint64_t)n_copy) >> 32);
                        amask |= 1u << q;
                        pj[q] = j;
                        tq[q] = t;
                        sq[q] = (uint32_t)((int32_t)t + R.copy_off[k]);
                    }
#pragma unroll
                    for (uint32_t q = 0; q < KS; ++q) {
                        lq[q] = (amask >> q) & 1u ? sig.ld(tq[q]) : 0u;
                        gq[q] = (amask >> q) & 1u ? sig.ld(sq[q]) : 0u;
                    }
#pragma unroll
                    for (uint32_t q = 0; q < KS; ++q) {
                        if (!((amask >> q) & 1u)) continue;
                        const uint32_t j = pj[q], t = tq[q];
                        const uint64_t ctr = (phase * (uint64_t)R.n_sites + j) * 4;
                        c_att++;
                        const uint32_t Lw = lq[q], Gw = gq[q];
                        if (Lw == Gw) continue;
                        c_nn++;
                        const uint32_t L = Lw & kIdMask;
                        bool simple = true;
                        if (L != 0) {
                            // the full shell (3^D - 1 sites, as the packed-word
                            // upload checks), all reads in flight together
                            constexpr uint32_t NTOPO = D == 2 ? 8u : 26u;
                            uint32_t nbw[NTOPO];
#pragma unroll
                            for (uint32_t i = 0; i < NTOPO; ++i)
                                nbw[i] = sig.ld((int32_t)t + R.topo_off[i]);
                            uint32_t pat = 0;
#pragma unroll
                            for (uint32_t i = 0; i < NTOPO; ++i)
                                pat |= (uint32_t)((nbw[i] & kIdMask) == L) << i;
                            c_top++;
                            simple = D == 2 ? ((c_sp2d[pat >> 5] >> (pat & 31)) & 1u)
                                            : ((sp3d[pat >> 6] >> (pat & 63)) & 1ull);
                        }
                        if (!simple) continue;
                        pmask |= 1u << q;
                        pw[q] = Lw | Gw << 16;
                        // unique within the phase: site index in the low 24 bits
                        ppr[q] = 0x8000000000000000ull |
                                 (splitmix_at(key, ctr + 1) & 0x7fffffffff000000ull) | j;
                    }
                    // commit rounds: the highest priority on both cells goes first
                    for (;;) {
                        if (!__syncthreads_or(pmask != 0)) break;
#pragma unroll
                        for (uint32_t q = 0; q < KS; ++q)
                            if ((pmask >> q) & 1u) {
                                const uint32_t L = pw[q] & kIdMask, G = (pw[q] >> 16) & kIdMask;
                                if (L) atomicMax(&prio[L - 1], ppr[q]);
                                if (G) atomicMax(&prio[G - 1], ppr[q]);
                            }
                        __syncthreads();
                        uint32_t wmask = 0;
#pragma unroll
                        for (uint32_t q = 0; q < KS; ++q)
                            if ((pmask >> q) & 1u) {
                                const uint32_t L = pw[q] & kIdMask, G = (pw[q] >> 16) & kIdMask;
                                if ((!L || prio[L - 1] == ppr[q]) && (!G || prio[G - 1] == ppr[q]))
                                    wmask |= 1u << q;
                            }
                        __syncthreads();
#pragma unroll
                        for (uint32_t q = 0; q < KS; ++q)
                            if ((pmask >> q) & 1u) {
                                const uint32_t L = pw[q] & kIdMask, G = (pw[q] >> 16) & kIdMask;
                                if (L) prio[L - 1] = 0;
                                if (G) prio[G - 1] = 0;
                            }
                        // one commit per thread and round: a thread's other
                        // winners bid again and win again next round (nothing
                        // that shares a cell with them can commit first)
                        if (wmask) {
                            const uint32_t qw = (uint32_t)__builtin_ctz(wmask);
                            uint32_t jw = 0, ww = 0;
#pragma unroll
                            for (uint32_t q = 0; q < KS; ++q)
                                if (q == qw) {
                                    jw = pj[q];
                                    ww = pw[q];
                                }
                            pmask &= ~(1u << qw);
                            commit_win(jw, ww & 0xffffu, ww >> 16);
                        }
                        if constexpr (SG) wg_fence();  // sigma writes, HBM
                    }
                }
                __syncthreads();
            }
        }
    }
    atomicAdd(&s_cnt[0], c_att);
    atomicAdd(&s_cnt[1], c_nn);
    atomicAdd(&s_cnt[2], c_top);
    atomicAdd(&s_cnt[3], c_en);
    atomicAdd(&s_cnt[4], c_acc);
    __syncthreads();
    for (uint32_t i = tid; i < n_padded; i += NT) R.sigma[i] = sig.ld(i) & kIdMask;
    for (uint32_t i = tid; i < n_cells; i += NT) {
        R.volume[i] = vol[i];
        R.interface[i] = itf[i];
    }
    if (tid == 0) {
        R.mcs += R.todo;
        R.todo = 0;
        R.done = 1;
        for (int i = 0; i < 5; ++i) R.cnt[i] += s_cnt[i];
    }
}

// ===========================================================================
// Readouts on the device (output.rs:147-337, labelling.rs:70-154): the
// connected components of the lattice and the per-cell shape, written into
// the replica's readout area and copied to the host once per sample.
//
// Components: a lock-free union-find over the interior sites (each link
// points from the larger root to the smaller, so a component's root is its
// first site in row-major order — the order in which the reference's scan
// meets components), cells joined under von Neumann connectivity, medium
// under the full shell, across the periodic wrap where there is one.  A cell
// with more than one root is fragmented; each medium root gets its
// component's volume and whether it touches a fixed boundary.  Shape: the
// centroid and gyration tensor of every cell from its position sums, and
// in 2D the eigenvalues, R_g^2 and kappa^2 with the reference's closed
// form (+, -, *, /, sqrt only, so the same bits as the host's); the 3D
// closed form needs acos and cos and is finished on the host with the same
// libm as the reference.
// ===========================================================================
struct ReadoutLayout {
    uint32_t cells;    // per cell: centroid[D], tensor[D*D], eigen[D], rg2, kappa (doubles)
    uint32_t header;   // n_fragmented, n_medium (u32)
    uint32_t frag;     // fragmented ids (n_cells u32)
    uint32_t medium;   // medium components (n_sites u32: volume | touches << 31)
    uint32_t ncomp;    // components per cell id (n_cells + 1 u32; not copied back)
    uint32_t total;
};
__host__ __device__ inline uint32_t readout_doubles_per_cell(uint32_t D) {
    return D + D * D + D + 2;
}
__host__ __device__ inline ReadoutLayout readout_layout(uint32_t D, uint32_t n_cells,
                                                        uint32_t n_sites) {
    ReadoutLayout L;
    L.cells = 0;
    L.header = n_cells * readout_doubles_per_cell(D) * 8;
    L.frag = L.header + 8;
    L.medium = L.frag + n_cells * 4;
    L.ncomp = L.medium + n_sites * 4;
    L.total = L.ncomp + (n_cells + 1) * 4;
    return L;
}

__device__ __forceinline__ uint32_t uf_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ uint32_t uf_find(uint32_t* lab, uint32_t x) {
    for (uint32_t p = uf_load(&lab[x]); p != x; p = uf_load(&lab[x])) x = p;
    return x;
}
__device__ void uf_unite(uint32_t* lab, uint32_t a, uint32_t b) {
    a = uf_find(lab, a);
    b = uf_find(lab, b);
    while (a != b) {
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        // link the larger root a under b if a is still a root
        const uint32_t old = atomicCAS(&lab[a], a, b);
        if (old == a) return;
        a = uf_find(lab, old);
        b = uf_find(lab, b);
    }
}

// Interior site q (row-major, last axis fastest) -> coordinates and padded flat index.
__device__ __forceinline__ uint32_t interior_flat(const Rep& R, uint32_t q, uint32_t* c) {
    uint32_t f = 0;
    for (int d = (int)R.D - 1; d >= 0; --d) {
        c[d] = q % R.dims[d];
        q /= R.dims[d];
        f += (c[d] + 1) * R.strides[d];
    }
    return f;
}
// The neighbour of interior site (c, flat x) along flat offset `off`
// (components -1, 0, 1): its padded flat index, or its interior mirror
// under a periodic boundary; false if it lies outside a fixed box.
__device__ __forceinline__ bool neighbour_site(const Rep& R, const uint32_t* c, int64_t off,
                                               uint32_t* out) {
    int64_t rem = off, f = 0;
    bool inside = true;
    for (int d = 0; d < (int)R.D; ++d) {
        const int64_t st = R.strides[d];
        const int64_t od = rem >= 0 ? (rem + st / 2) / st : -((-rem + st / 2) / st);
        rem -= od * st;
        int64_t p = (int64_t)c[d] + od;  // interior coordinate
        if (p < 0 || p >= (int64_t)R.dims[d]) {
            inside = false;
            p = p < 0 ? (int64_t)R.dims[d] - 1 : 0;
        }
        f += (p + 1) * st;
    }
    *out = (uint32_t)f;
    return inside || R.boundary == CPM_BOUNDARY_PERIODIC;
}

constexpr uint32_t kRoThreads = 1024;
__global__ void __launch_bounds__(kRoThreads) k_readout(Rep* __restrict__ reps,
                                                        const uint32_t* __restrict__ idx) {
    __shared__ uint32_t s_wave[kRoThreads / 64];
    __shared__ uint32_t s_base;
    Rep& R = reps[idx[blockIdx.x]];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t D = R.D, n_sites = R.n_sites, n_cells = R.n_cells, n_padded = R.n_padded;
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;
    const uint32_t* __restrict__ sig = R.sigma;
    uint32_t* lab = R.labels;
    uint32_t* cnt = R.lcount;
    const ReadoutLayout Lo = readout_layout(D, n_cells, n_sites);
    uint32_t* ncomp = (uint32_t*)(R.readout + Lo.ncomp);  // [n_cells + 1]
    for (uint32_t x = tid; x < n_padded; x += kRoThreads) {
        lab[x] = x;
        cnt[x] = 0;
    }
    for (uint32_t i = tid; i <= n_cells; i += kRoThreads) ncomp[i] = 0;
    __threadfence();
    __syncthreads();
    // union over each edge once: forward offsets (positive flat delta)
    for (uint32_t q = tid; q < n_sites; q += kRoThreads) {
        uint32_t c[3];
        const uint32_t x = interior_flat(R, q, c), o = sig[x];
        if (o == 0) {
            for (uint32_t k = 0; k < R.n_topo; ++k) {
                const int32_t off = R.topo_off[k];
                uint32_t y;
                if (off > 0 && neighbour_site(R, c, off, &y) && sig[y] == 0) uf_unite(lab, x, y);
            }
        } else {
            for (uint32_t d = 0; d < D; ++d) {
                uint32_t y;
                if (neighbour_site(R, c, (int64_t)R.strides[d], &y) && sig[y] == o)
                    uf_unite(lab, x, y);
            }
        }
    }
    __threadfence();
    __syncthreads();
    for (uint32_t q = tid; q < n_sites; q += kRoThreads) {
        uint32_t c[3];
        const uint32_t x = interior_flat(R, q, c);
        lab[x] = uf_find(lab, x);
    }
    __threadfence();
    __syncthreads();
    for (uint32_t q = tid; q < n_sites; q += kRoThreads) {
        uint32_t c[3];
        const uint32_t x = interior_flat(R, q, c), o = sig[x], r = uf_load(&lab[x]);
        if (o == 0) {
            atomicAdd(&cnt[r], 1u);
            bool edge = false;
            for (uint32_t d = 0; d < D; ++d) edge |= c[d] == 0 || c[d] == R.dims[d] - 1;
            if (!periodic && edge) atomicOr(&cnt[r], 0x80000000u);
        } else if (r == x) {
            atomicAdd(&ncomp[o], 1u);
        }
    }
    __threadfence();
    __syncthreads();
    // ordered compaction: medium roots by site, then fragmented cells by id
    uint32_t* med = (uint32_t*)(R.readout + Lo.medium);
    uint32_t* frag = (uint32_t*)(R.readout + Lo.frag);
    uint32_t* header = (uint32_t*)(R.readout + Lo.header);
    for (int pass = 0; pass < 2; ++pass) {
        const uint32_t n_items = pass == 0 ? n_sites : n_cells;
        if (tid == 0) s_base = 0;
        __syncthreads();
        for (uint32_t b0 = 0; b0 < n_items; b0 += kRoThreads) {
            const uint32_t q = b0 + tid;
            bool keep = false;
            uint32_t val = 0;
            if (q < n_items) {
                if (pass == 0) {
                    uint32_t c[3];
                    const uint32_t x = interior_flat(R, q, c);
                    keep = sig[x] == 0 && uf_load(&lab[x]) == x;
                    if (keep) val = uf_load(&cnt[x]);
                } else {
                    keep = uf_load(&ncomp[q + 1]) > 1;
                    val = q + 1;
                }
            }
            const uint64_t bal = __ballot(keep);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (lane == 0) s_wave[wave] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = s_base;
            for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
            if (keep) (pass == 0 ? med : frag)[before + rank] = val;
            __syncthreads();
            if (tid == 0)
                for (uint32_t w = 0; w < kRoThreads / 64; ++w) s_base += s_wave[w];
            __syncthreads();
        }
        if (tid == 0) header[pass == 0 ? 1 : 0] = s_base;
        __syncthreads();
    }
    // per-cell shape (output.rs:147-231, 289-302)
    const uint32_t per = readout_doubles_per_cell(D);
    for (uint32_t i = tid; i < n_cells; i += kRoThreads) {
        double* out = (double*)(R.readout + Lo.cells) + (size_t)i * per;
        const double v = (double)R.volume[i];
        double cen[3], g[9];
        for (uint32_t d = 0; d < D; ++d) cen[d] = (double)R.sum_pos[i * D + d] / v;
        for (uint32_t a = 0; a < D; ++a)
            for (uint32_t b = 0; b < D; ++b)
                g[a * D + b] = (double)R.sum_outer[(i * D + a) * D + b] / v - cen[a] * cen[b];
        for (uint32_t d = 0; d < D; ++d) out[d] = cen[d];
        for (uint32_t k = 0; k < D * D; ++k) out[D + k] = g[k];
        if (D == 2) {
            const double a = g[0], b = g[1], dd = g[3];
            const double tr = a + dd;
            const double t2 = (a - dd) * (a - dd) + 4.0 * b * b;
            const double disc = sqrt(t2 < 0.0 ? 0.0 : t2);  // std::max(t2, 0.0)
            double e0 = (tr + disc) / 2.0, e1 = (tr - disc) / 2.0;
            if (e1 > e0) {
                const double t = e0;
                e0 = e1;
                e1 = t;
            }
            double sum = 0.0, ssq = 0.0;
            sum += e0;
            sum += e1;
            ssq += e0 * e0;
            ssq += e1 * e1;
            double kappa = 0.0;
            if (sum > 0.0) kappa = (2.0 / (2.0 - 1.0)) * (ssq / (sum * sum)) - 1.0 / (2.0 - 1.0);
            out[D + 4] = e0;
            out[D + 5] = e1;
            out[D + 6] = sum;
            out[D + 7] = kappa;
        }
    }
}

// ===========================================================================
// Contact counts recounted from the lattice (sublattice mode keeps no
// running counts): the pairs of sites that differ, each unordered pair of
// interior sites once (a halo neighbour counts as the interior site it
// mirrors), exactly as initialization's recompute counts them
// (initialization.rs:396-436).  One replica per blockIdx.y; integer counts,
// so the result does not depend on the order of the atomics.
// ===========================================================================
__global__ void __launch_bounds__(256) k_contact_clear(Rep* __restrict__ reps,
                                                       const uint32_t* __restrict__ idx) {
    Rep& R = reps[idx[blockIdx.y]];
    const uint64_t n1 = (uint64_t)R.n_cells + 1, total = n1 * n1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x)
        R.contact[i] = 0;
}

__global__ void __launch_bounds__(256) k_contact_count(Rep* __restrict__ reps,
                                                       const uint32_t* __restrict__ idx) {
    __shared__ int32_t s_off[26][3];
    Rep& R = reps[idx[blockIdx.y]];
    const uint32_t D = R.D;
    const uint64_t n1 = (uint64_t)R.n_cells + 1;
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;
    // the offsets' components (each -1, 0 or 1: the flat offset rounded per axis)
    if (threadIdx.x < R.n_energy) {
        int64_t off = R.energy_off[threadIdx.x];
        for (int d = 0; d < 3; ++d) {
            int64_t od = 0;
            if (d < (int)D) {
                const int64_t st = R.strides[d];
                od = off >= 0 ? (off + st / 2) / st : -((-off + st / 2) / st);
                off -= od * st;
            }
            s_off[threadIdx.x][d] = (int32_t)od;
        }
    }
    __syncthreads();
    const uint32_t dimv[3] = {R.dims[0], R.dims[1], R.dims[2]};
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < R.n_sites;
         q += gridDim.x * blockDim.x) {
        uint32_t c[3] = {0, 0, 0};
        uint32_t rem = q;
        uint64_t fi = 0;
        for (int d = (int)D - 1; d >= 0; --d) {
            c[d] = rem % dimv[d];
            rem /= dimv[d];
            fi += (uint64_t)(c[d] + 1) * R.strides[d];
        }
        const uint32_t ci = R.sigma[fi];
        for (uint32_t k = 0; k < R.n_energy; ++k) {
            // the neighbour's padded coordinates and the interior site it
            // is or mirrors
            uint64_t fj = (uint64_t)((int64_t)fi + R.energy_off[k]), key = 0;
            bool interior = true;
            for (int d = 0; d < (int)D; ++d) {
                const int32_t p = (int32_t)c[d] + 1 + s_off[k][d];  // padded coordinate
                uint32_t w = (uint32_t)p;
                if (p == 0 || p == (int32_t)dimv[d] + 1) {
                    interior = false;
                    w = p == 0 ? dimv[d] : 1u;  // the mirrored interior coordinate
                }
                key += (uint64_t)w * R.strides[d];
            }
            if (interior) key = fj;
            else if (!periodic) key = ~0ull;  // Fixed: a distinct, permanently medium site
            if (key != ~0ull && key <= fi) continue;
            const uint32_t cj = R.sigma[fj];
            if (cj != ci) {
                const uint32_t lo = min(ci, cj), hi = max(ci, cj);
                atomicAdd(&R.contact[(uint64_t)lo * n1 + hi], 1u);
            }
        }
    }
}

// ===========================================================================
// Sublattice mode's energy after a call.
// ===========================================================================
// global_energy (energy.rs:561-565) of each listed replica from its device
// state, one wave per replica, in the host's summation order: the volume
// terms cell by cell, the interface terms cell by cell, then the adhesion
// over interior sites in row-major order and, per site, over the energy
// stencil in order (each unordered pair once, keyed by the interior site it
// is or mirrors; a fixed halo is a distinct medium site).  The lanes compute
// 64 cells' or sites' terms at a time and stage them in LDS in that order;
// every lane then adds the staged terms one after another, so the FP64
// rounding is the host's.
constexpr uint32_t kEnergyStage = 64 * 26;  // terms staged per pass (doubles)

__device__ __forceinline__ uint32_t lane_prefix(uint32_t v, uint32_t lane) {
    // exclusive prefix sum over the wave
    uint32_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (lane >= d) x += y;
    }
    return x - v;
}

__global__ void __launch_bounds__(64) k_energy(Rep* __restrict__ reps,
                                               const uint32_t* __restrict__ idx) {
    __shared__ double stage[kEnergyStage];
    __shared__ int32_t s_off[26][3];  // the energy stencil, per-axis offsets
    Rep& R = reps[idx[blockIdx.x]];
    const uint32_t lane = threadIdx.x, T = R.n_types, T1 = T + 1, D = R.D;
    if (lane < R.n_energy) {
        // each component is -1, 0 or 1: the flat offset rounded per axis
        int64_t off = R.energy_off[lane];
        for (int d = 0; d < 3; ++d) {
            int64_t od = 0;
            if (d < (int)D) {
                const int64_t st = R.strides[d];
                od = off >= 0 ? (off + st / 2) / st : -((-off + st / 2) / st);
                off -= od * st;
            }
            s_off[lane][d] = (int32_t)od;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const uint32_t dims0 = R.dims[0], dims1 = R.dims[1], dims2 = R.dims[2];
    const uint32_t dimv[3] = {dims0, dims1, dims2};
    const double* lv = R.params;
    const double* li = R.params + T;
    const double* J = R.params + 2 * T;
    const uint32_t* tv = R.targets;
    const uint32_t* ti = R.targets + T;
    const bool periodic = R.boundary == CPM_BOUNDARY_PERIODIC;
    double sums[2] = {0.0, 0.0};
    // volume, then interface: lambda * (x - target)^2 per cell
    for (int which = 0; which < 2; ++which) {
        const uint32_t* x = which ? R.interface : R.volume;
        for (uint32_t c0 = 0; c0 < R.n_cells; c0 += 64) {
            const uint32_t c = c0 + lane;
            if (c < R.n_cells) {
                const uint32_t t = R.type_index[c];
                const double d = (double)x[c] - (double)(which ? ti[t] : tv[t]);
                stage[lane] = (which ? li[t] : lv[t]) * (d * d);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            const uint32_t n = min(64u, R.n_cells - c0);
            double acc = sums[which];
            for (uint32_t k = 0; k < n; ++k) acc += stage[k];
            sums[which] = acc;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
        }
    }
    // adhesion: sum over unlike keyed pairs of J[type(ci)][type(cj)]; each
    // lane counts its site's terms, then writes them at its prefix offset
    double adh = 0.0;
    for (uint32_t q0 = 0; q0 < R.n_sites; q0 += 64) {
        const uint32_t q = q0 + lane;
        const bool site = q < R.n_sites;
        uint32_t c[3] = {0, 0, 0};
        uint64_t fi = 0;
        if (site) {
            uint32_t rem = q;
            for (int d = (int)D - 1; d >= 0; --d) {
                c[d] = rem % dimv[d];
                rem /= dimv[d];
                fi += (uint64_t)(c[d] + 1) * R.strides[d];
            }
        }
        const uint32_t ci = site ? R.sigma[fi] : 0u;
        const uint32_t ai = ci ? R.type_index[ci - 1] + 1 : 0u;
        // the pair (site, stencil entry k) counts: the neighbour's adhesion index in *aj
        auto pair = [&](uint32_t k, uint32_t* aj) {
            const uint64_t fj = (uint64_t)((int64_t)fi + R.energy_off[k]);
            uint64_t key = 0;
            bool interior = true;
            for (int d = 0; d < (int)D; ++d) {
                const int32_t p = (int32_t)c[d] + 1 + s_off[k][d];  // padded coordinate
                uint32_t w = (uint32_t)p;
                if (p == 0 || p == (int32_t)dimv[d] + 1) {
                    interior = false;
                    w = p == 0 ? dimv[d] : 1u;  // the mirrored interior coordinate
                }
                key += (uint64_t)w * R.strides[d];
            }
            if (interior) key = fj;
            else if (!periodic) key = ~0ull;
            if (key != ~0ull && key <= fi) return false;
            const uint32_t cj = R.sigma[fj];
            if (cj == ci) return false;
            *aj = cj ? R.type_index[cj - 1] + 1 : 0u;
            return true;
        };
        uint32_t cnt = 0, aj = 0;
        if (site)
            for (uint32_t k = 0; k < R.n_energy; ++k) cnt += pair(k, &aj) ? 1u : 0u;
        const uint32_t at = lane_prefix(cnt, lane);
        if (site)
            for (uint32_t k = 0, r = 0; k < R.n_energy; ++k)
                if (pair(k, &aj)) stage[at + r++] = J[ai * T1 + aj];
        const uint32_t total = (uint32_t)__shfl((int)(at + cnt), 63, 64);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        // eight staged terms loaded at a time, added one after another
        double acc = adh;
        uint32_t k = 0;
        for (; k + 8 <= total; k += 8) {
            double v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = stage[k + j];
#pragma unroll
            for (int j = 0; j < 8; ++j) acc += 1.0 * v[j];
        }
        for (; k < total; ++k) acc += 1.0 * stage[k];
        adh = acc;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) R.energy = sums[0] + sums[1] + adh;
}

// ===========================================================================
// 3D simple-point table, generated from the definition on the device: one
// thread per 64 patterns (simple_point.rs:242-276, 340-354).
// ===========================================================================
struct SpMasks {
    uint32_t face[26], full[26], origin_face, all;
};

__device__ __forceinline__ uint32_t sp_grow(uint32_t seed, uint32_t uni_, const uint32_t* adj) {
    uint32_t comp = seed;
    for (;;) {
        uint32_t nxt = comp;
        for (uint32_t f = comp; f; f &= f - 1) nxt |= adj[__builtin_ctz(f)];
        nxt &= uni_;
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

// ===========================================================================
// scatter_and_grow on the device (initialization.rs:202-330): one wave per
// model, every draw in the reference's order.  The wave places the seeds
// (lanes test the separation from 64 earlier seeds at a time), then grows
// the cells round-robin in the shuffled order: each claim prunes the cell's
// frontier of claimed sites keeping its order (lane-parallel compaction),
// draws the pick, swap-removes it and appends the new site's unclaimed
// neighbours in copy-stencil order.  Draws and the uniform steps run on
// every lane alike, so each lane's reads of the lattice follow its own
// writes; the frontier moves between lanes and is fenced.
// ===========================================================================
struct InitJob {
    uint32_t* sigma;           // padded lattice (the model's own device copy)
    uint32_t* frontier;        // cell i's frontier at fr_start[i]
    uint32_t* seeds;           // 3 coordinates per cell
    const uint32_t* fr_start;  // frontier offset of each cell
    const uint32_t* target;    // target volume of each cell
    uint32_t* volume;          // out: volume of each cell
    uint64_t rng[4];           // in: the model's state; out: after the last draw
    uint64_t sep_sq;           // seeds closer than sqrt(sep_sq) are refused
    uint32_t dims[3], strides[3];
    int32_t copy_vec[26][3];   // copy stencil, as per-axis offsets
    uint32_t n_cells, n_copy, D, periodic;
    uint32_t status;           // out: 0 grown; 1 + i: seed i found no room
    uint32_t n_padded;         // sigma's length (cleared first)
};
// cells per model the device places (order, volume and frontier length in LDS)
constexpr uint32_t kInitMaxCells = 4096;

__device__ __forceinline__ void init_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(64) k_scatter_grow(InitJob* __restrict__ jobs) {
    extern __shared__ __align__(16) unsigned char smem[];
    InitJob& J = jobs[blockIdx.x];
    const uint32_t lane = threadIdx.x;
    const uint32_t n = J.n_cells, D = J.D, n_copy = J.n_copy;
    const bool periodic = J.periodic != 0;
    uint32_t dims[3], strides[3];
    for (int d = 0; d < 3; ++d) {
        dims[d] = J.dims[d];
        strides[d] = J.strides[d];
    }
    uint32_t* const order = (uint32_t*)smem;
    uint32_t* const vol = order + n;
    uint32_t* const flen = vol + n;
    uint32_t* const seeds = J.seeds;
    uint32_t* const sigma = J.sigma;
    uint32_t* const fr = J.frontier;
    Xoshiro r{{J.rng[0], J.rng[1], J.rng[2], J.rng[3]}};
    for (uint32_t i = lane; i < J.n_padded; i += 64) sigma[i] = 0;
    init_fence();
    auto flat_of = [&](const uint32_t* c) {
        uint32_t f = 0;
        for (uint32_t d = 0; d < D; ++d) f += (c[d] + 1) * strides[d];
        return f;
    };
    // canonical_growable_neighbour (initialization.rs:183-196): the
    // neighbour of interior site c along copy offset k, wrapped on a
    // periodic box; none past a fixed boundary
    auto growable = [&](const uint32_t* c, uint32_t k, uint32_t* out) {
        uint32_t q[3];
        for (uint32_t d = 0; d < D; ++d) {
            const int64_t x = (int64_t)c[d] + J.copy_vec[k][d];
            if (x < 0 || x >= (int64_t)dims[d]) {
                if (!periodic) return false;
                q[d] = x < 0 ? dims[d] - 1 : 0u;
            } else {
                q[d] = (uint32_t)x;
            }
        }
        *out = flat_of(q);
        return true;
    };
    auto coords_of = [&](uint32_t f, uint32_t* c) {
        for (uint32_t d = 0; d < D; ++d) {
            c[d] = f / strides[d] - 1;
            f %= strides[d];
        }
    };

    // ---- seeds (initialization.rs:226-249)
    for (uint32_t i = 0; i < n; ++i) {
        bool placed = false;
        for (uint32_t a = 0; a < 20000 && !placed; ++a) {
            uint32_t cand[3] = {0, 0, 0};
            for (uint32_t d = 0; d < D; ++d) cand[d] = (uint32_t)gen_range(r, dims[d]);
            bool far = true;
            for (uint32_t j0 = 0; j0 < i && far; j0 += 64) {
                const uint32_t j = j0 + lane;
                bool near = false;
                if (j < i) {
                    uint64_t sq = 0;
                    for (uint32_t d = 0; d < D; ++d) {
                        const uint32_t sd = seeds[3 * j + d];
                        const uint32_t raw = sd > cand[d] ? sd - cand[d] : cand[d] - sd;
                        const uint32_t del = periodic ? min(raw, dims[d] - raw) : raw;
                        sq += (uint64_t)del * del;
                    }
                    near = sq < J.sep_sq;
                }
                far = __ballot(near) == 0;
            }
            if (far) {
                for (uint32_t d = 0; d < 3; ++d) seeds[3 * i + d] = cand[d];
                placed = true;
            }
        }
        if (!placed) {
            if (lane == 0) J.status = 1 + i;
            return;
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (uint32_t d = 0; d < 3; ++d) c[d] = seeds[3 * i + d];
        sigma[flat_of(c)] = i + 1;
    }
    for (uint32_t i = lane; i < n; i += 64) {
        order[i] = i;
        vol[i] = 1;
    }
    init_fence();
    // frontier_of each seed once every seed is placed (initialization.rs:332-343)
    for (uint32_t i = lane; i < n; i += 64) {
        uint32_t c[3], len = 0;
        for (uint32_t d = 0; d < 3; ++d) c[d] = seeds[3 * i + d];
        const uint32_t base = J.fr_start[i];
        for (uint32_t k = 0; k < n_copy; ++k) {
            uint32_t nb;
            if (growable(c, k, &nb) && sigma[nb] == 0) fr[base + len++] = nb;
        }
        flen[i] = len;
    }
    init_fence();

    // ---- growth rounds (initialization.rs:262-294)
    for (;;) {
        bool grew = false;
        for (uint32_t idx = 0; idx < n; ++idx) {
            const uint32_t i = order[idx];
            if (vol[i] >= J.target[i]) continue;
            const uint32_t base = J.fr_start[i];
            // retain the unclaimed sites, in order
            const uint32_t len0 = flen[i];
            uint32_t len = 0;
            for (uint32_t c0 = 0; c0 < len0; c0 += 64) {
                const bool in = c0 + lane < len0;
                const uint32_t v = in ? fr[base + c0 + lane] : 0u;
                const bool keep = in && sigma[v] == 0;
                const uint64_t m = __ballot(keep);
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (keep) fr[base + len + rank] = v;
                len += (uint32_t)__popcll(m);
            }
            if (len == 0) {
                flen[i] = 0;
                continue;
            }
            init_fence();
            const uint32_t pick = (uint32_t)gen_range(r, len);
            const uint32_t claimed = fr[base + pick];
            const uint32_t last = fr[base + len - 1];
            init_fence();
            if (lane == 0) fr[base + pick] = last;
            --len;
            sigma[claimed] = i + 1;
            vol[i] += 1;
            uint32_t c[3];
            coords_of(claimed, c);
            uint32_t nb = 0;
            const bool ok = lane < n_copy && growable(c, lane, &nb) && sigma[nb] == 0;
            const uint64_t m = __ballot(ok);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (ok) fr[base + len + rank] = nb;
            flen[i] = len + (uint32_t)__popcll(m);
            init_fence();
            grew = true;
        }
        if (!grew) break;
        // the per-round shuffle (initialization.rs:289-293)
        for (uint32_t i = n - 1; i >= 1; --i) {
            const uint32_t j = (uint32_t)gen_range_incl(r, i);
            const uint32_t a = order[i], b = order[j];
            order[i] = b;
            order[j] = a;
        }
    }
    for (uint32_t i = lane; i < n; i += 64) J.volume[i] = vol[i];
    if (lane == 0) {
        for (int k = 0; k < 4; ++k) J.rng[k] = r.s[k];
        J.status = 0;
    }
}

// What initialization.rs:296-305 derives once the cells are grown, for
// models whose state the device wrote: the periodic halo, then (after the
// cells' tables are cleared) interfaces and position sums per site, with
// integer atomics, so the order of the updates does not matter.  Contact
// counts follow with k_contact_count, the energy with k_energy.
__global__ void __launch_bounds__(256) k_halo(Rep* __restrict__ reps) {
    Rep& R = reps[blockIdx.y];
    if (R.boundary != CPM_BOUNDARY_PERIODIC) return;
    const uint32_t D = R.D;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < R.n_padded;
         f += gridDim.x * blockDim.x) {
        uint32_t rem = f, src = 0;
        bool halo = false;
        for (uint32_t d = 0; d < D; ++d) {
            const uint32_t p = rem / R.strides[d];
            rem -= p * R.strides[d];
            uint32_t w = p;
            if (p == 0) {
                w = R.dims[d];
                halo = true;
            } else if (p == R.dims[d] + 1) {
                w = 1;
                halo = true;
            }
            src += w * R.strides[d];
        }
        if (halo) R.sigma[f] = R.sigma[src];
    }
}

__global__ void __launch_bounds__(256) k_state_clear(Rep* __restrict__ reps) {
    Rep& R = reps[blockIdx.y];
    const uint64_t n = R.n_cells, D = R.D, n1 = n + 1;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t0; i < n; i += step) R.interface[i] = 0;
    for (uint64_t i = t0; i < n * D; i += step) R.sum_pos[i] = 0;
    for (uint64_t i = t0; i < n * D * D; i += step) R.sum_outer[i] = 0;
    for (uint64_t i = t0; i < n1 * n1; i += step) R.contact[i] = 0;
}

__global__ void __launch_bounds__(256) k_recompute_cells(Rep* __restrict__ reps) {
    Rep& R = reps[blockIdx.y];
    const uint32_t D = R.D;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < R.n_sites;
         q += gridDim.x * blockDim.x) {
        uint32_t c[3] = {0, 0, 0}, rem = q;
        uint64_t fi = 0;
        for (int d = (int)D - 1; d >= 0; --d) {
            c[d] = rem % R.dims[d];
            rem /= R.dims[d];
            fi += (uint64_t)(c[d] + 1) * R.strides[d];
        }
        const uint32_t ci = R.sigma[fi];
        if (!ci) continue;
        const uint64_t i = ci - 1;
        uint32_t unlike = 0;
        for (uint32_t k = 0; k < R.n_energy; ++k)
            unlike += R.sigma[(uint64_t)((int64_t)fi + R.energy_off[k])] != ci ? 1u : 0u;
        if (unlike) atomicAdd(&R.interface[i], unlike);
        unsigned long long* sp = (unsigned long long*)(R.sum_pos + i * D);
        unsigned long long* so = (unsigned long long*)(R.sum_outer + i * D * D);
        for (uint32_t d = 0; d < D; ++d) {
            atomicAdd(&sp[d], (unsigned long long)c[d]);
            for (uint32_t e2 = 0; e2 < D; ++e2)
                atomicAdd(&so[d * D + e2], (unsigned long long)((uint64_t)c[d] * c[e2]));
        }
    }
}

// ===========================================================================
// Host side
// ===========================================================================
struct Device {
    char* base = nullptr;
    Rep rep{};  // device pointers + static fields
    bool lds = false;     // packed sigma words fit (LDS or HBM copy): wave and sublattice paths
    bool wave = false;    // exact replay on the wave path (else the lane path)
    bool gsig = false;    // sigma does not fit in LDS: always in HBM
    uint32_t lds_bytes = 0;       // LDS with sigma in LDS (when it fits)
    uint32_t lds_bytes_hbm = 0;   // LDS with sigma in HBM
    int device = 0;               // HIP device holding `base`
    Engine* eng = nullptr;        // the engine whose launches write `base`
};

// LDS kept free of the dynamic allocation for the kernels' static LDS
// (the sublattice kernel's counters and the compiler's own use)
constexpr uint32_t kLdsStaticReserve = 1024;
// Replicas whose chain (MCS x sites) is at least this share of the batch's
// longest run on the pipelined kernel.
constexpr double kPipeShare = 0.75;
// A sigma placement: the share of the chip's LDS it plans to fill
// (workgroups of mixed sizes do not pack a CU's LDS perfectly; 0: sigma in
// HBM for every replica, 8: in LDS wherever it fits) and which replicas get
// LDS first — the smallest lattices (most replicas resident in LDS) or the
// longest chains (MCS x sites: the replicas that set the call's time).
struct Placement {
    double fill;
    bool longest_first;
};
constexpr Placement kDefaultPlacement = {0.75, false};
// the candidates the autotuner tries
constexpr Placement kPlacements[] = {{0.0, false}, {0.5, false}, {0.75, false},
                                     {1.0, false}, {8.0, false}, {0.6, true},
                                     {0.75, true}, {0.85, true}, {1.0, true}};
constexpr int kNumPlacements = sizeof(kPlacements) / sizeof(kPlacements[0]);
constexpr uint64_t kTuneMinSliceMcs = 4;  // tune only calls of at least 4 MCS per slice

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t* sp3d = nullptr;
    Rep* reps_dev = nullptr;
    size_t reps_cap = 0;
    uint32_t* idx_dev = nullptr;
    size_t idx_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // replicas of different LDS footprints run as separate launches on
    // these streams, concurrently, after each round's stream fill
    static constexpr int kClassStreams = 4;
    hipStream_t cls_stream[kClassStreams] = {};
    hipEvent_t cls_done[kClassStreams] = {};
    uint32_t max_lds = 0;
    uint32_t n_cu = 0;
    std::vector<hipEvent_t> launch_ev;  // start/stop pairs around replay launches
    // autotuned sigma placement: batch signature -> best candidate
    std::map<uint64_t, int> tuned_placement;
    // Streams for state copies to the host, shared by the host threads that
    // sync replicas in parallel: on the null stream their copies would run
    // one after another, and a stream of each thread's own would be created
    // and destroyed with each short-lived worker thread.  Every copy first
    // waits for `work_done`, recorded on `stream` after each call's device
    // work, so a copy never reads state a kernel is still writing.
    static constexpr int kCopyStreams = 16;
    hipStream_t copy_stream[kCopyStreams] = {};
    hipEvent_t work_done = nullptr;
    // a sigma placement fixed by the caller (fill 0: the autotuner decides)
    Placement forced{0.0, false};
    // the placement of the last exact-replay call (fill 0: none yet)
    Placement last_placement{0.0, false};
    // false: no replica runs on the pipelined (two-wave) kernel
    bool pipelining = true;
    // this device's copy of `scratch` (hipGetSymbolAddress at creation)
    unsigned char* noalloc_scratch = nullptr;
    // launches of the exact-replay wave kernels per instantiation key
    std::map<int, uint64_t> launches_by_key;
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
    for (int i = 0; i < Engine::kClassStreams; ++i) {
        HIP_OK(hipStreamCreateWithFlags(&e->cls_stream[i], hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&e->cls_done[i], hipEventDisableTiming));
    }
    for (int i = 0; i < Engine::kCopyStreams; ++i)
        HIP_OK(hipStreamCreateWithFlags(&e->copy_stream[i], hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&e->work_done, hipEventDisableTiming));
    HIP_OK(hipEventRecord(e->work_done, e->stream));
    // the device address of this device's `scratch` (the array's name is a
    // host-side stand-in, not a device address)
    HIP_OK(hipGetSymbolAddress((void**)&e->noalloc_scratch, HIP_SYMBOL(scratch)));
    if (!e->noalloc_scratch) throw Error(CPM_ERR_DEVICE, "no device address for the scratch array");
    uint8_t t2[256];
    sp_table2d(t2);
    uint32_t w[8] = {0};
    for (int p = 0; p < 256; ++p)
        if (t2[p]) w[p >> 5] |= 1u << (p & 31);
    HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_sp2d), w, sizeof w));
    xo_jump_selfcheck();
    {
        uint64_t jump[kSegs][4];
        for (uint32_t g = 0; g < kSegs; ++g) {
            JumpPoly q = xo_jump_poly((uint64_t)g * kSegLen);
            std::memcpy(jump[g], q.w, 32);
        }
        HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_jump), jump, sizeof jump));
    }
    int lds = 0;
    HIP_OK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    e->max_lds = (uint32_t)lds;
    int n_cu = 0;
    HIP_OK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    e->n_cu = (uint32_t)n_cu;
#define CPM_FN(k, d, ne, gs, pw, tb) (const void*)k_lds_engine<d, ne, gs, pw, tb>,
#define CPM_SLFN(k, d, sg, nt) (const void*)k_sublattice<d, sg, nt>,
    for (auto fn : {CPM_WAVE_KERNELS(CPM_FN) CPM_SL_KERNELS(CPM_SLFN)})
#undef CPM_SLFN
#undef CPM_FN
        HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   lds - kLdsStaticReserve));
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
    if (e->idx_dev) (void)hipFree(e->idx_dev);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    for (int i = 0; i < Engine::kClassStreams; ++i) {
        if (e->cls_stream[i]) (void)hipStreamDestroy(e->cls_stream[i]);
        if (e->cls_done[i]) (void)hipEventDestroy(e->cls_done[i]);
    }
    for (hipEvent_t ev : e->launch_ev) (void)hipEventDestroy(ev);
    for (int i = 0; i < Engine::kCopyStreams; ++i)
        if (e->copy_stream[i]) (void)hipStreamDestroy(e->copy_stream[i]);
    if (e->work_done) (void)hipEventDestroy(e->work_done);
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

struct Layout {
    size_t sigma, type_index, volume, interface, sum_pos, sum_outer, contact, params, targets,
        birth, edge_sites, edge_pos, draws, sig16, labels, lcount, readout, total;
};

static Layout layout_of(const Model& m, bool with_draws) {
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
    L.params = take((3 * T + (T + 1) * (T + 1)) * 8);
    L.targets = take(3 * T * 4);
    L.birth = take(m.cfg.act ? m.n_padded * 4 : 0);
    const bool edges = m.cfg.proposal == CPM_PROPOSAL_EDGE_LIST;
    L.edge_sites = take(edges ? m.n_sites * 4 : 0);
    L.edge_pos = take(edges ? m.n_padded * 4 : 0);
    L.draws = with_draws ? take((size_t)kFill * 8 + (size_t)kSegs * 32) : 0;
    L.sig16 = with_draws ? take(m.n_padded * 2) : 0;
    L.labels = take(m.n_padded * 4);
    L.lcount = take(m.n_padded * 4);
    L.readout = take(readout_layout((uint32_t)D, (uint32_t)n, (uint32_t)m.n_sites).total);
    L.total = off;
    return L;
}

// Whether the model runs on the wave path (one wave per replica), and with
// sigma in LDS or in HBM (*gsig): packed 16-bit words need at most 4095
// cells and 15 types; sigma stays in LDS when it fits beside the tables.
static bool lds_eligible(const Engine* e, const Model& m, uint32_t* bytes, bool* gsig = nullptr) {
    if (m.n_cells > kMaxPackedCells || m.n_padded > (1u << 30)) return false;
    // Metropolis-Hastings and the Act term run on the lane path only
    if (m.cfg.acceptance != CPM_ACCEPT_METROPOLIS || m.cfg.act) return false;
    if (m.cfg.types.size() > 15) return false;
    const uint32_t T = (uint32_t)m.cfg.types.size();
    const LdsLayout in = lds_layout((uint32_t)m.n_padded, (uint32_t)m.n_cells, T, true);
    const LdsLayout out = lds_layout((uint32_t)m.n_padded, (uint32_t)m.n_cells, T, false);
    const bool fits = in.total + kStageBytes + kLdsStaticReserve <= e->max_lds;
    if (gsig) *gsig = !fits;
    *bytes = (fits ? in.total : out.total) + kStageBytes;
    if (!gsig) return fits;  // callers that need sigma in LDS
    return out.total + kStageBytes + kLdsStaticReserve <= e->max_lds;
}

// LDS bytes a launch reserves for replicas needing `bytes` (grouping key).
static uint32_t lds_class(const Engine* e, uint32_t bytes) {
    return std::min((bytes + 1023) / 1024 * 1024, e->max_lds - kLdsStaticReserve);
}

// Allocates the model's device state and copies its parameters; with
// `with_state` also its lattice and per-cell state (the device-side
// initialisation writes those itself).
static void upload(Engine* e, Model* m, bool with_state = true) {
    if (m->dev) return;
    if (with_state && !m->initialized) throw Error(CPM_ERR_ARG, "model is not initialized");
    uint32_t lds_bytes = 0;
    bool gsig = false;
    const bool lds = lds_eligible(e, *m, &lds_bytes, &gsig);
    const Layout L = layout_of(*m, lds);
    auto* d = new Device();
    hipError_t err = hipMalloc(&d->base, L.total);
    if (err != hipSuccess) {
        delete d;
        throw Error(CPM_ERR_OOM, std::string("device allocation failed: ") + hipGetErrorString(err));
    }
    m->dev = d;
    d->device = e->device;
    d->eng = e;
    d->lds = lds;
    // the EdgeList proposal runs exact replay on the lane path (the
    // sublattice mode takes it on the packed words)
    d->wave = lds && m->cfg.proposal == CPM_PROPOSAL_UNIFORM;
    d->gsig = lds && gsig;
    d->lds_bytes = lds_bytes;
    d->lds_bytes_hbm = lds_layout((uint32_t)m->n_padded, (uint32_t)m->n_cells,
                                  (uint32_t)m->cfg.types.size(), false)
                           .total +
                       kStageBytes;
    if (lds) {
        // the LDS kernel addresses the full shell from the strides and maps
        // the energy stencil into it at compile time
        const int NT = m->D == 2 ? 8 : 26;
        bool same = m->n_topo == NT;
        for (int k = 0; same && k < NT; ++k) {
            const int code = k < NT / 2 ? k : k + 1;
            const int64_t o0 = code % 3 - 1, o1 = (code / 3) % 3 - 1, o2 = code / 9 - 1;
            const int64_t a = m->D == 2 ? o0 * (int64_t)m->strides[0] + o1
                                        : o0 * (int64_t)m->strides[0] +
                                              o1 * (int64_t)m->strides[1] + o2;
            same = a == m->topo_off[k];
        }
        for (int q = 0; same && q < m->n_energy; ++q)
            same = m->energy_off[q] == m->topo_off[shell_index(m->D, m->n_energy, q)];
        if (!same) throw Error(CPM_ERR_DEVICE, "stencil order differs from the LDS kernel's");
    }
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
    R.n_padded = (uint32_t)m->n_padded;
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
    R.acceptance = (uint32_t)m->cfg.acceptance;
    R.act = m->cfg.act ? 1u : 0u;
    R.birth = m->cfg.act ? (uint32_t*)(d->base + L.birth) : nullptr;
    const bool edges = m->cfg.proposal == CPM_PROPOSAL_EDGE_LIST;
    R.proposal = (uint32_t)m->cfg.proposal;
    R.edge_sites = edges ? (uint32_t*)(d->base + L.edge_sites) : nullptr;
    R.edge_pos = edges ? (uint32_t*)(d->base + L.edge_pos) : nullptr;
    R.draws = lds ? (uint64_t*)(d->base + L.draws) : nullptr;
    R.sig16 = lds ? (uint16_t*)(d->base + L.sig16) : nullptr;
    R.labels = (uint32_t*)(d->base + L.labels);
    R.lcount = (uint32_t*)(d->base + L.lcount);
    R.readout = (unsigned char*)(d->base + L.readout);
    const size_t T = R.n_types;
    std::vector<double> params(3 * T + (T + 1) * (T + 1));
    std::vector<uint32_t> targets(3 * T);
    for (size_t t = 0; t < T; ++t) {
        params[t] = m->cfg.types[t].lambda_volume;
        params[T + t] = m->cfg.types[t].lambda_interface;
        params[2 * T + (T + 1) * (T + 1) + t] = m->cfg.types[t].lambda_act;
        targets[t] = m->cfg.types[t].target_volume;
        targets[T + t] = m->cfg.types[t].target_interface;
        targets[2 * T + t] = m->cfg.types[t].max_act;
    }
    for (size_t i = 0; i <= T; ++i)
        for (size_t j = 0; j <= T; ++j) params[2 * T + i * (T + 1) + j] = m->cfg.adhesion[i][j];
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) HIP_OK(hipMemcpyAsync(d->base + off, src, bytes, hipMemcpyHostToDevice, e->stream));
    };
    const size_t n = m->n_cells, D = m->D;
    put(L.type_index, m->type_index.data(), n * 4);
    put(L.params, params.data(), params.size() * 8);
    put(L.targets, targets.data(), targets.size() * 4);
    if (!with_state) return;  // the caller's kernels write the state (same stream)
    put(L.sigma, m->sigma.data(), m->n_padded * 4);
    put(L.volume, m->volume.data(), n * 4);
    put(L.interface, m->interface.data(), n * 4);
    put(L.sum_pos, m->sum_pos.data(), n * D * 8);
    put(L.sum_outer, m->sum_outer.data(), n * D * D * 8);
    put(L.contact, m->contact.data(), (n + 1) * (n + 1) * 4);
    if (m->cfg.act) put(L.birth, m->birth.data(), m->n_padded * 4);
    if (edges) {
        put(L.edge_sites, m->edge_sites.data(), m->n_sites * 4);
        put(L.edge_pos, m->edge_pos.data(), m->n_padded * 4);
    }
    HIP_OK(hipStreamSynchronize(e->stream));
    m->device_dirty = false;
}

size_t device_footprint(Engine* e, const Model& m) {
    uint32_t bytes = 0;
    bool gsig = false;
    return layout_of(m, lds_eligible(e, m, &bytes, &gsig)).total;
}

// LDS bytes of k_sublattice for a model: sigma in LDS when it fits with the
// tables and priority words (*sg = false), else the tables alone.
static uint32_t sublattice_lds(Engine* e, const Model& m, bool* sg) {
    const uint32_t T = (uint32_t)m.cfg.types.size();
    const LdsLayout in = lds_layout((uint32_t)m.n_padded, (uint32_t)m.n_cells, T, true);
    const uint32_t with_sigma = sl_lds_bytes(in, (uint32_t)m.n_cells);
    *sg = with_sigma + kLdsStaticReserve > e->max_lds;
    if (!*sg) return with_sigma;
    const LdsLayout out = lds_layout((uint32_t)m.n_padded, (uint32_t)m.n_cells, T, false);
    return sl_lds_bytes(out, (uint32_t)m.n_cells);
}

bool sublattice_supported(Engine* e, const Model& m) {
    if (m.cfg.acceptance != CPM_ACCEPT_METROPOLIS || m.cfg.act) return false;
    uint32_t bytes = 0;
    bool gsig = false;
    if (!lds_eligible(e, m, &bytes, &gsig)) return false;  // packed sigma words
    if (m.cfg.boundary == CPM_BOUNDARY_PERIODIC)
        for (int d = 0; d < m.D; ++d)
            if (m.dims[d] % 2) return false;
    return true;
}

size_t device_free_bytes(Engine* e) {
    HIP_OK(hipSetDevice(e->device));
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    return free_b;
}

void device_release(Model* m) {
    if (!m || !m->dev) return;
    (void)hipFree(m->dev->base);
    delete m->dev;
    m->dev = nullptr;
}

// The copy stream of the calling host thread: threads take the engine's
// copy streams round robin.
static hipStream_t copy_stream_for(Engine* e) {
    static std::atomic<uint32_t> next_slot{0};
    thread_local const uint32_t slot = next_slot.fetch_add(1) % Engine::kCopyStreams;
    return e->copy_stream[slot];
}

void sync_to_host(Model* m) {
    if (!m->dev || !m->device_dirty) return;
    const Layout L = layout_of(*m, m->dev->lds);
    const size_t n = m->n_cells, D = m->D;
    // a model initialised on the device has host arrays of the right size
    // only from its first sync on
    m->volume.resize(n);
    m->interface.resize(n);
    m->sum_pos.resize(n * D);
    m->sum_outer.resize(n * D * D);
    m->contact.resize((n + 1) * (n + 1));
    const size_t contact_bytes = (n + 1) * (n + 1) * 4;
    // sigma through the position sums are adjacent in the allocation: one
    // copy for them (each copy call costs more than a few KiB of traffic),
    // and the contact matrix straight into place when it is large
    const bool contact_apart = contact_bytes > (1u << 20);
    const size_t end = contact_apart ? L.contact : L.contact + contact_bytes;
    std::vector<uint8_t> stage(end - L.sigma);
    hipStream_t st = copy_stream_for(m->dev->eng);
    HIP_OK(hipStreamWaitEvent(st, m->dev->eng->work_done, 0));
    HIP_OK(hipMemcpyAsync(stage.data(), m->dev->base + L.sigma, stage.size(),
                          hipMemcpyDeviceToHost, st));
    if (contact_apart)
        HIP_OK(hipMemcpyAsync(m->contact.data(), m->dev->base + L.contact, contact_bytes,
                              hipMemcpyDeviceToHost, st));
    if (m->cfg.act)
        HIP_OK(hipMemcpyAsync(m->birth.data(), m->dev->base + L.birth, m->n_padded * 4,
                              hipMemcpyDeviceToHost, st));
    if (m->cfg.proposal == CPM_PROPOSAL_EDGE_LIST) {
        HIP_OK(hipMemcpyAsync(m->edge_sites.data(), m->dev->base + L.edge_sites,
                              m->n_sites * 4, hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(m->edge_pos.data(), m->dev->base + L.edge_pos, m->n_padded * 4,
                              hipMemcpyDeviceToHost, st));
    }
    HIP_OK(hipStreamSynchronize(st));
    auto put = [&](void* dst, size_t off, size_t bytes) {
        if (bytes) std::memcpy(dst, stage.data() + (off - L.sigma), bytes);
    };
    put(m->sigma.data(), L.sigma, m->n_padded * 4);
    put(m->volume.data(), L.volume, n * 4);
    put(m->interface.data(), L.interface, n * 4);
    put(m->sum_pos.data(), L.sum_pos, n * D * 8);
    put(m->sum_outer.data(), L.sum_outer, n * D * D * 8);
    if (!contact_apart) put(m->contact.data(), L.contact, contact_bytes);
    m->device_dirty = false;
}

template <class T>
static void ensure_buf(T*& ptr, size_t& cap, size_t n) {
    if (n <= cap) return;
    if (ptr) HIP_OK(hipFree(ptr));
    ptr = nullptr;
    size_t c = std::max(n, (size_t)64);
    HIP_OK(hipMalloc(&ptr, c * sizeof(T)));
    cap = c;
}

// The lattice, type indices, volumes and interfaces only (adjacent at the
// start of the allocation): what the energy needs.  The rest of the host
// copy stays stale (device_dirty unchanged) until a full sync.
static void sync_lattice_tables(Model* m) {
    if (!m->dev || !m->device_dirty) return;
    const Layout L = layout_of(*m, m->dev->lds);
    const size_t n = m->n_cells;
    m->volume.resize(n);
    m->interface.resize(n);
    std::vector<uint8_t> stage(L.interface + n * 4 - L.sigma);
    hipStream_t st = copy_stream_for(m->dev->eng);
    HIP_OK(hipStreamWaitEvent(st, m->dev->eng->work_done, 0));
    HIP_OK(hipMemcpyAsync(stage.data(), m->dev->base + L.sigma, stage.size(),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    std::memcpy(m->sigma.data(), stage.data(), m->n_padded * 4);
    std::memcpy(m->volume.data(), stage.data() + (L.volume - L.sigma), n * 4);
    std::memcpy(m->interface.data(), stage.data() + (L.interface - L.sigma), n * 4);
}

void sync_scalars(Model*) {}  // scalars are written back after every call

Engine* device_engine(const Model* m) { return m->dev ? m->dev->eng : nullptr; }

// take_sample of the device-resident state of each model: the readout
// kernel writes the labelling and per-cell shape into the model's readout
// area; each model's volumes, interfaces, contacts and readout area (and
// lattice when asked for) are then copied back on the copy streams, and
// the samples assembled on the host (3D eigenvalues included).
void device_samples(Engine* e, Model* const* models, size_t n, bool include_lattice,
                    const uint64_t* at_mcs, const double* at_energy, Sample* out) {
    if (!n) return;
    {
        std::lock_guard<std::mutex> lock(e->mu);
        HIP_OK(hipSetDevice(e->device));
        std::vector<Rep> reps(n);
        std::vector<uint32_t> order(n);
        for (size_t i = 0; i < n; ++i) {
            if (!models[i]->dev || models[i]->dev->eng != e)
                throw Error(CPM_ERR_ARG, "device readout of a model not resident on this engine");
            reps[i] = models[i]->dev->rep;
            order[i] = (uint32_t)i;
        }
        ensure_buf(e->reps_dev, e->reps_cap, n);
        ensure_buf(e->idx_dev, e->idx_cap, n);
        HIP_OK(hipMemcpyAsync(e->reps_dev, reps.data(), n * sizeof(Rep), hipMemcpyHostToDevice,
                              e->stream));
        HIP_OK(hipMemcpyAsync(e->idx_dev, order.data(), n * 4, hipMemcpyHostToDevice, e->stream));
        hipLaunchKernelGGL(k_readout, dim3((uint32_t)n), dim3(kRoThreads), 0, e->stream,
                           e->reps_dev, e->idx_dev);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(e->work_done, e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
    }
    parallel_for(n, [&](size_t i) {
        Model* m = models[i];
        const Layout L = layout_of(*m, m->dev->lds);
        const uint32_t D = (uint32_t)m->D, nc = (uint32_t)m->n_cells;
        const ReadoutLayout Ro = readout_layout(D, nc, (uint32_t)m->n_sites);
        // the readout area up to a first share of the medium components
        constexpr uint32_t kFirstMedium = 1024;
        const uint32_t first = Ro.medium + 4 * std::min<uint32_t>(kFirstMedium, (uint32_t)m->n_sites);
        std::vector<uint8_t> ro(Ro.total);
        std::vector<uint8_t> vi(L.interface + nc * 4 - L.volume);
        Sample& s = out[i];
        s.contact.resize((size_t)(nc + 1) * (nc + 1));
        std::vector<uint32_t> lat;
        if (include_lattice) lat.resize(m->n_padded);
        hipStream_t st = copy_stream_for(e);
        HIP_OK(hipStreamWaitEvent(st, e->work_done, 0));
        HIP_OK(hipMemcpyAsync(ro.data(), m->dev->base + L.readout, first, hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(vi.data(), m->dev->base + L.volume, vi.size(), hipMemcpyDeviceToHost,
                              st));
        HIP_OK(hipMemcpyAsync(s.contact.data(), m->dev->base + L.contact, s.contact.size() * 4,
                              hipMemcpyDeviceToHost, st));
        if (include_lattice)
            HIP_OK(hipMemcpyAsync(lat.data(), m->dev->base + L.sigma, lat.size() * 4,
                                  hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        uint32_t hdr[2];
        std::memcpy(hdr, ro.data() + Ro.header, 8);
        if (hdr[1] > kFirstMedium) {
            HIP_OK(hipMemcpyAsync(ro.data() + first, m->dev->base + L.readout + first,
                                  4 * (size_t)(hdr[1] - kFirstMedium), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
        }
        s.mcs = at_mcs[i];
        s.energy = at_energy[i];
        s.ids.resize(nc);
        s.volume.resize(nc);
        s.interface.resize(nc);
        std::memcpy(s.volume.data(), vi.data(), nc * 4);
        std::memcpy(s.interface.data(), vi.data() + (L.interface - L.volume), nc * 4);
        s.centroid.resize((size_t)nc * D);
        s.eigen.resize((size_t)nc * D);
        s.rg2.resize(nc);
        s.kappa.resize(nc);
        const uint32_t per = readout_doubles_per_cell(D);
        const double* cells = (const double*)(ro.data() + Ro.cells);
        for (uint32_t c = 0; c < nc; ++c) {
            s.ids[c] = c + 1;
            const double* x = cells + (size_t)c * per;
            for (uint32_t d = 0; d < D; ++d) s.centroid[(size_t)c * D + d] = x[d];
            if (D == 2) {
                s.eigen[(size_t)c * 2] = x[6];
                s.eigen[(size_t)c * 2 + 1] = x[7];
                s.rg2[c] = x[8];
                s.kappa[c] = x[9];
            } else {
                shape_from_tensor((int)D, x + D, &s.eigen[(size_t)c * D], &s.rg2[c], &s.kappa[c]);
            }
        }
        const uint32_t* frag = (const uint32_t*)(ro.data() + Ro.frag);
        s.fragmented.assign(frag, frag + hdr[0]);
        const uint32_t* med = (const uint32_t*)(ro.data() + Ro.medium);
        s.medium_volumes.resize(hdr[1]);
        s.medium_touch.resize(hdr[1]);
        for (uint32_t k = 0; k < hdr[1]; ++k) {
            s.medium_volumes[k] = med[k] & 0x7fffffffu;
            s.medium_touch[k] = (uint8_t)(med[k] >> 31);
        }
        if (include_lattice) {
            s.lattice = m->unpadded(lat.data());
            s.has_lattice = true;
        }
    });
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

static void run_sublattice(Engine* e, Model* const* models, size_t n, const uint64_t* todo,
                           cpm_counters* counters);


static void launch_lds_engine(int key, uint32_t lds, uint32_t count, hipStream_t st, Rep* reps,
                              const uint32_t* idx, const uint64_t* sp3d) {
    const dim3 g(count), b(key % 4000 >= 2000 ? 128 : 64);
    switch (key) {
#define CPM_CASE(k, d, ne, gs, pw, tb)                                                        \
    case k:                                                                                   \
        hipLaunchKernelGGL((k_lds_engine<d, ne, gs, pw, tb>), g, b, lds, st, reps, idx, sp3d); \
        break;
        CPM_WAVE_KERNELS(CPM_CASE)
#undef CPM_CASE
        default: throw Error(CPM_ERR_DEVICE, "no exact-replay kernel for this stencil");
    }
}

// One exact-replay call with a given sigma placement fill (caller holds the
// engine lock).
// noalloc: allocate no device memory and create no events — every model
// must be resident already, the descriptors and launch order go to
// `scratch`, and launches are not timed one by one.
constexpr double kHbmSpeed = 0.73;     // chain speed with sigma in HBM over in LDS (measured)
constexpr double kStragglerMargin = 1.05;  // an HBM chain past the LDS chains by this much
constexpr double kStragglerShare = 0.20;   // at most this share of a batch's chains
static void run_exact(Engine* e, Model* const* models, size_t n, const uint64_t* todo,
                      Placement placement, cpm_counters* counters, double budget_s = 0.0,
                      bool noalloc = false) {
    bool any3d = false;
    for (size_t i = 0; i < n; ++i) {
        if (noalloc && (!models[i]->dev || models[i]->dev->eng != e))
            throw Error(CPM_ERR_ARG,
                        "cpm_run_mcs_noalloc: model " + std::to_string(i) +
                            " is not resident on this engine's device (advance it once with "
                            "cpm_run_mcs)");
        upload(e, models[i]);
        any3d |= models[i]->D == 3;
    }
    if (any3d && noalloc && !e->sp3d)
        throw Error(CPM_ERR_ARG, "cpm_run_mcs_noalloc: the 3D simple-point table is not built yet "
                                 "(advance a 3D model once with cpm_run_mcs)");
    if (any3d) ensure_sp3d(e);
    Rep* reps_dev;
    uint32_t* idx_dev;
    if (noalloc) {
        if (n > kNoallocMaxReplicas)
            throw Error(CPM_ERR_ARG, "cpm_run_mcs_noalloc: " + std::to_string(n) +
                                         " models exceed the maximum of " +
                                         std::to_string(kNoallocMaxReplicas) + " per call");
        reps_dev = (Rep*)e->noalloc_scratch;
        idx_dev = (uint32_t*)(e->noalloc_scratch + kNoallocRepBytes);
    } else {
        ensure_buf(e->reps_dev, e->reps_cap, n);
        ensure_buf(e->idx_dev, e->idx_cap, n);
        reps_dev = e->reps_dev;
        idx_dev = e->idx_dev;
    }
    std::vector<Rep> reps(n);
    // Sigma placement for this batch: a lattice too big for LDS keeps sigma
    // in HBM; the others keep it in LDS in the placement's order (smallest
    // lattices or longest chains first) while the whole batch's LDS still
    // fits the placement's share of the chip (so every replica is resident
    // from the start and the largest lattices' serial chains set the time);
    // the rest keep sigma in HBM (read through L2) and use LDS for their
    // per-cell tables only.
    // The replicas with the longest chains (and a lone replica) run on the
    // pipelined kernel, a second wave decoding and evaluating ahead of the
    // walking one, while the batch leaves SIMDs without a wave: where every
    // SIMD is busy the second wave would only take issue slots from others.
    std::vector<char> piped(n, 0);
    {
        double max_work = 0.0;
        size_t waves = 0;
        std::vector<uint32_t> cand;
        for (size_t i = 0; i < n; ++i) {
            const Model* m = models[i];
            if (!todo[i] || !m->dev->wave) continue;
            ++waves;
            max_work = std::max(max_work, (double)todo[i] * (double)m->n_sites);
            if (m->dev->lds_bytes_hbm + pipe_extra_bytes(m->D) + kLdsStaticReserve <= e->max_lds)
                cand.push_back((uint32_t)i);
        }
        auto work = [&](uint32_t i) { return (double)todo[i] * (double)models[i]->n_sites; };
        std::stable_sort(cand.begin(), cand.end(),
                         [&](uint32_t a, uint32_t b) { return work(a) > work(b); });
        const size_t simds = 4 * (size_t)e->n_cu;
        // a time-limited call runs every chain for the same time: any of
        // them gains from a SIMD of its own
        const double share = budget_s > 0 ? 0.0 : kPipeShare;
        for (uint32_t i : cand) {
            if (!e->pipelining || work(i) < share * max_work || waves + 1 > simds) break;
            piped[i] = 1;
            ++waves;
        }
    }
    auto lds_need = [&](size_t i, bool hbm) {
        const Device* d = models[i]->dev;
        return (hbm ? d->lds_bytes_hbm : d->lds_bytes) +
               (piped[i] ? pipe_extra_bytes(models[i]->D) : 0u);
    };
    std::vector<char> in_hbm(n, 0);
    auto place = [&]() {
        std::fill(in_hbm.begin(), in_hbm.end(), 0);
        std::vector<uint32_t> cand;
        double demand = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const Model* m = models[i];
            if (!todo[i] || !m->dev->wave) continue;
            in_hbm[i] = 1;
            demand += (double)lds_class(e, lds_need(i, true));
            if (!m->dev->gsig && lds_need(i, false) + kLdsStaticReserve <= e->max_lds)
                cand.push_back((uint32_t)i);
        }
        if (placement.longest_first)
            std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
                return (double)todo[a] * (double)models[a]->n_sites >
                       (double)todo[b] * (double)models[b]->n_sites;
            });
        else
            std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
                return models[a]->dev->lds_bytes < models[b]->dev->lds_bytes;
            });
        const double capacity = placement.fill * (double)e->n_cu * (double)e->max_lds;
        for (uint32_t i : cand) {
            const double extra =
                (double)lds_class(e, lds_need(i, false)) - (double)lds_class(e, lds_need(i, true));
            if (demand + extra > capacity) {
                // smallest first: nothing later fits either; longest first:
                // a later, smaller lattice may still fit
                if (placement.longest_first) continue;
                break;
            }
            demand += extra;
            in_hbm[i] = 0;
        }
    };
    place();
    // Kernels of more instantiations than there are class streams would
    // share streams and run one after another: then no replica is pipelined.
    {
        std::set<int> keys;
        bool any_piped = false;
        for (size_t i = 0; i < n; ++i)
            if (todo[i] && models[i]->dev->wave) {
                keys.insert((piped[i] ? 2000 : 0) + (in_hbm[i] ? 1000 : 0) + models[i]->D * 100 +
                            models[i]->n_energy);
                any_piped |= piped[i] != 0;
            }
        if (any_piped && keys.size() > (size_t)Engine::kClassStreams) {
            std::fill(piped.begin(), piped.end(), 0);
            place();
        }
    }
    // The longest chains left with sigma in HBM by the LDS budget would run
    // well past every chain whose lattice is in LDS (an HBM lattice costs a
    // single-wave chain ~1/kHbmSpeed per attempt).  They run on the
    // pipelined kernel instead, whose second wave decodes and evaluates
    // ahead and hides the HBM latency (in a full batch ~180 ns per attempt
    // against ~280 single-wave): their extra waves take SIMD slots from the
    // batch's short chains, which start when the first chains end and
    // still finish long before the longest ones.  At most kStragglerShare
    // of the batch's chains is piped this way, and only while the kernels
    // keep a class stream each.
    size_t n_stragglers = 0;
    if (budget_s == 0.0 && e->pipelining) {
        auto work = [&](size_t i) { return (double)todo[i] * (double)models[i]->n_sites; };
        double lds_max = 0.0;
        size_t waves = 0;
        std::set<int> keys;
        for (size_t i = 0; i < n; ++i)
            if (todo[i] && models[i]->dev->wave) {
                ++waves;
                if (!in_hbm[i] || piped[i]) lds_max = std::max(lds_max, work(i));
                keys.insert((piped[i] ? 2000 : 0) + (in_hbm[i] ? 1000 : 0) + models[i]->D * 100 +
                            models[i]->n_energy);
            }
        std::vector<size_t> cand;
        for (size_t i = 0; i < n; ++i)
            if (todo[i] && models[i]->dev->wave && in_hbm[i] && !piped[i] &&
                work(i) > kStragglerMargin * kHbmSpeed * lds_max &&
                lds_need(i, true) + pipe_extra_bytes(models[i]->D) + kLdsStaticReserve <=
                    e->max_lds) {
                cand.push_back(i);
                keys.insert(3000 + models[i]->D * 100 + models[i]->n_energy);
            }
        if (!cand.empty() && (double)cand.size() <= kStragglerShare * (double)waves &&
            keys.size() <= (size_t)Engine::kClassStreams)
            for (size_t i : cand) {
                piped[i] = 1;
                ++n_stragglers;
            }
    }
    // partition: global-path replicas, and wave-path replicas by kernel
    // instantiation (the key of launch_lds_engine)
    std::vector<uint32_t> order, global_idx;
    std::map<int, std::vector<uint32_t>> by_kernel;
    for (size_t i = 0; i < n; ++i) {
        Model* m = models[i];
        Rep R = m->dev->rep;
        for (int k = 0; k < 4; ++k) R.rng[k] = m->rng.s[k];
        R.energy = m->energy;
        R.mcs = m->mcs;
        R.n_edge = m->n_edge;
        R.todo = todo[i];
        R.attempt_in_mcs = m->attempt_in_mcs;
        R.done = todo[i] == 0;
        R.error = 0;
        for (int k = 0; k < 6; ++k) R.cnt[k] = 0;
        R.t_run = 0;
        // 100 MHz ticks (s_memrealtime); the lane path runs whole MCS only,
        // one per time-limited call
        R.budget = budget_s > 0 && m->dev->wave ? std::max<uint64_t>(1, (uint64_t)(budget_s * 1e8)) : 0;
        if (budget_s > 0 && !m->dev->wave) R.todo = std::min<uint64_t>(R.todo, 1);
        if (m->attempt_in_mcs && !m->dev->wave)
            throw Error(CPM_ERR_DEVICE, "a lane-path model stopped mid-MCS");
        reps[i] = R;
        if (!todo[i]) continue;
        if (m->dev->wave)
            by_kernel[(budget_s > 0 ? 4000 : 0) + (piped[i] ? 2000 : 0) + (in_hbm[i] ? 1000 : 0) +
                      m->D * 100 + m->n_energy]
                .push_back((uint32_t)i);
        else
            global_idx.push_back((uint32_t)i);
    }
    // One launch per group, the groups of all kernels running concurrently
    // on the class streams.  A launch's LDS allocation is its largest
    // member's, so each kernel's replicas are sorted by footprint class and
    // cut into contiguous groups (as many as the streams allow, shared
    // between kernels in proportion to their replica counts) that minimise
    // the LDS allocated in total.
    struct Launch {
        int D;
        uint32_t lds, first, count;
    };
    std::vector<Launch> launches;
    {
        size_t n_lds = 0;
        for (auto& kv : by_kernel) n_lds += kv.second.size();
        const size_t n_kernels = by_kernel.size();
        const size_t spare =
            Engine::kClassStreams > n_kernels ? Engine::kClassStreams - n_kernels : 0;
        for (auto& [key, v] : by_kernel) {
            std::vector<std::pair<uint32_t, uint32_t>> cls(v.size());  // (class bytes, replica)
            for (size_t j = 0; j < v.size(); ++j)
                cls[j] = {lds_class(e, lds_need(v[j], in_hbm[v[j]] != 0)), v[j]};
            std::stable_sort(cls.begin(), cls.end(),
                             [](auto& a, auto& b) { return a.first < b.first; });
            // distinct classes with counts
            std::vector<uint32_t> size_k;
            std::vector<size_t> count_k, start_k;
            for (size_t j = 0; j < cls.size(); ++j) {
                if (size_k.empty() || cls[j].first != size_k.back()) {
                    size_k.push_back(cls[j].first);
                    count_k.push_back(0);
                    start_k.push_back(j);
                }
                ++count_k.back();
            }
            const size_t K = size_k.size();
            const size_t G = std::min(K, 1 + (n_lds ? spare * v.size() / n_lds : 0));
            // best[g][k]: least allocation covering classes [0, k) in g groups
            const double inf = 1e300;
            std::vector<std::vector<double>> best(G + 1, std::vector<double>(K + 1, inf));
            std::vector<std::vector<size_t>> cut(G + 1, std::vector<size_t>(K + 1, 0));
            best[0][0] = 0.0;
            for (size_t g = 1; g <= G; ++g)
                for (size_t k = 1; k <= K; ++k) {
                    double cnt = 0.0;
                    for (size_t a = k; a-- > 0;) {  // last group = classes [a, k)
                        cnt += (double)count_k[a];
                        const double c = best[g - 1][a] + cnt * (double)size_k[k - 1];
                        if (c < best[g][k]) {
                            best[g][k] = c;
                            cut[g][k] = a;
                        }
                    }
                }
            size_t g_best = 1;
            for (size_t g = 1; g <= G; ++g)
                if (best[g][K] < best[g_best][K]) g_best = g;
            std::vector<std::pair<size_t, size_t>> groups;  // class ranges
            for (size_t g = g_best, k = K; g > 0; --g) {
                groups.push_back({cut[g][k], k});
                k = cut[g][k];
            }
            // largest footprints first, and within a launch the longest
            // chains first: the dispatcher spreads consecutive workgroups
            // over the CUs and SIMDs, so the replicas that set the launch's
            // length do not share a SIMD while others are left idle
            for (auto [a, b] : groups) {
                const size_t first = start_k[a], last = start_k[b - 1] + count_k[b - 1];
                launches.push_back({key, size_k[b - 1], (uint32_t)order.size(),
                                    (uint32_t)(last - first)});
                const size_t at = order.size();
                for (size_t j = first; j < last; ++j) order.push_back(cls[j].second);
                std::stable_sort(order.begin() + at, order.end(), [&](uint32_t x, uint32_t y) {
                    return (double)todo[x] * (double)models[x]->n_sites >
                           (double)todo[y] * (double)models[y]->n_sites;
                });
            }
        }
    }
    // pipelined launches are enqueued first: their workgroups take two
    // waves each and are dispatched before the others fill the SIMDs
    std::stable_partition(launches.begin(), launches.end(),
                          [](const Launch& l) { return l.D % 4000 >= 2000; });
    const uint32_t global_first = (uint32_t)order.size();
    order.insert(order.end(), global_idx.begin(), global_idx.end());
    HIP_OK(hipMemcpyAsync(reps_dev, reps.data(), n * sizeof(Rep), hipMemcpyHostToDevice,
                          e->stream));
    if (!order.empty())
        HIP_OK(hipMemcpyAsync(idx_dev, order.data(), order.size() * 4, hipMemcpyHostToDevice,
                              e->stream));
    HIP_OK(hipEventRecord(e->ev0, e->stream));
    HIP_OK(hipEventRecord(e->cls_done[0], e->stream));  // descriptors uploaded
    if (!global_idx.empty()) {
        const uint32_t ng = (uint32_t)global_idx.size();
        hipLaunchKernelGGL(k_exact_replay, dim3((ng + 63) / 64), dim3(64), 0, e->stream,
                           reps_dev, idx_dev + global_first, ng, e->sp3d);
        HIP_OK(hipGetLastError());
    }
    double replay_ms = 0.0;
    uint64_t replay_launches = 0;
    if (!launches.empty()) {
        // one launch per group: each wave runs its replica to its MCS
        // target, refilling its own stream buffer as it goes
        while (!noalloc && e->launch_ev.size() < 2 * launches.size()) {
            hipEvent_t ev;
            HIP_OK(hipEventCreate(&ev));
            e->launch_ev.push_back(ev);
        }
        const int ns = (int)std::min<size_t>(launches.size(), Engine::kClassStreams);
        for (int k = 0; k < ns; ++k) HIP_OK(hipStreamWaitEvent(e->cls_stream[k], e->cls_done[0], 0));
        for (size_t li = 0; li < launches.size(); ++li) {
            const Launch& l = launches[li];
            hipStream_t st = e->cls_stream[li % Engine::kClassStreams];
            if (!noalloc) HIP_OK(hipEventRecord(e->launch_ev[2 * li], st));
            launch_lds_engine(l.D, l.lds, l.count, st, reps_dev, idx_dev + l.first, e->sp3d);
            HIP_OK(hipGetLastError());
            e->launches_by_key[l.D] += 1;
            if (!noalloc) HIP_OK(hipEventRecord(e->launch_ev[2 * li + 1], st));
        }
        for (int k = 0; k < ns; ++k) {
            HIP_OK(hipEventRecord(e->cls_done[k], e->cls_stream[k]));
            HIP_OK(hipStreamWaitEvent(e->stream, e->cls_done[k], 0));
        }
    }
    HIP_OK(hipEventRecord(e->ev1, e->stream));
    HIP_OK(hipEventRecord(e->work_done, e->stream));
    HIP_OK(hipMemcpyAsync(reps.data(), reps_dev, n * sizeof(Rep), hipMemcpyDeviceToHost,
                          e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    for (size_t li = 0; li < (noalloc ? 0 : launches.size()); ++li) {
        float lms = 0.f;
        HIP_OK(hipEventElapsedTime(&lms, e->launch_ev[2 * li], e->launch_ev[2 * li + 1]));
        replay_ms += lms;
        ++replay_launches;
    }
    if (counters) {
        counters->kernel_ms += ms;
        counters->replay_launch_ms += replay_ms;
        counters->replay_launches += replay_launches;
    }
    for (size_t i = 0; i < n; ++i)
        if (reps[i].error)
            throw Error(CPM_ERR_DEVICE,
                        reps[i].error == 2 ? "the binomial sampler of the EdgeList proposal ran away"
                                           : "an attempt needed more than 128 stream draws");
    for (size_t i = 0; i < n; ++i) {
        Model* m = models[i];
        if (todo[i]) {
            for (int k = 0; k < 4; ++k) m->rng.s[k] = reps[i].rng[k];
            m->energy = reps[i].energy;
            m->mcs = reps[i].mcs;
            m->n_edge = reps[i].n_edge;
            m->attempt_in_mcs = reps[i].attempt_in_mcs;
            m->device_dirty = true;
            // the chain's own speed in this call (wave path), for schedulers
            if (reps[i].t_run > 0 && reps[i].cnt[0] > 0)
                m->attempt_rate = (double)reps[i].cnt[0] / ((double)reps[i].t_run * 1e-8);
        }
        add_counters(counters, reps[i]);
    }
    if (counters) counters->piped_stragglers += n_stragglers;
}

static void run_sublattice(Engine* e, Model* const* models, size_t n, const uint64_t* todo,
                           cpm_counters* counters) {
    {
        std::lock_guard<std::mutex> lock(e->mu);
        HIP_OK(hipSetDevice(e->device));
        bool any3d = false;
        for (size_t i = 0; i < n; ++i) {
            Model* m = models[i];
            if (m->cfg.acceptance != CPM_ACCEPT_METROPOLIS || m->cfg.act)
                throw Error(CPM_ERR_ARG,
                            "sublattice mode runs the Metropolis rule without the Act term; "
                            "use exact replay for Metropolis-Hastings or Act models");
            upload(e, m);
            if (!m->dev->lds)
                throw Error(CPM_ERR_ARG,
                            "sublattice mode needs packed lattice words (at most 4095 cells "
                            "and 15 cell types)");
            if (m->cfg.boundary == CPM_BOUNDARY_PERIODIC)
                for (int d = 0; d < m->D; ++d)
                    if (m->dims[d] % 2)
                        throw Error(CPM_ERR_ARG,
                                    "sublattice mode needs even side lengths on periodic axes");
            any3d |= m->D == 3;
        }
        if (any3d) ensure_sp3d(e);
        ensure_buf(e->reps_dev, e->reps_cap, n);
        ensure_buf(e->idx_dev, e->idx_cap, n);
        std::vector<Rep> reps(n);
        // (kernel key, LDS class) -> replicas; key = D * 100 + 10 for the big
        // workgroups, + 1 for sigma in HBM
        std::map<std::pair<int, uint32_t>, std::vector<uint32_t>> classes;
        for (size_t i = 0; i < n; ++i) {
            Model* m = models[i];
            Rep R = m->dev->rep;
            R.mcs = m->mcs;
            R.todo = todo[i];
            R.done = todo[i] == 0;
            for (int k = 0; k < 6; ++k) R.cnt[k] = 0;
            if (todo[i]) {
                // one draw of the replica's own stream keys its
                // counter-based randomness for every later call
                if (!m->has_sublattice_key) {
                    m->sublattice_key = xo_next(m->rng);
                    m->has_sublattice_key = true;
                }
                R.fill_state[0] = m->sublattice_key;
                bool sg = false;
                const uint32_t bytes = sublattice_lds(e, *m, &sg);
                if (bytes + kLdsStaticReserve > e->max_lds)
                    throw Error(CPM_ERR_ARG,
                                "sublattice mode: the per-cell tables do not fit in LDS");
                const bool big = sg || m->n_sites >= kSlBigSites;
                const int key = m->D * 100 + (big ? 10 : 0) + (sg ? 1 : 0);
                classes[{key, lds_class(e, bytes)}].push_back((uint32_t)i);
            }
            reps[i] = R;
        }
        std::vector<uint32_t> order;
        HIP_OK(hipMemcpyAsync(e->reps_dev, reps.data(), n * sizeof(Rep), hipMemcpyHostToDevice,
                              e->stream));
        for (auto& [key, v] : classes) order.insert(order.end(), v.begin(), v.end());
        if (!order.empty())
            HIP_OK(hipMemcpyAsync(e->idx_dev, order.data(), order.size() * 4,
                                  hipMemcpyHostToDevice, e->stream));
        HIP_OK(hipEventRecord(e->ev0, e->stream));
        uint32_t first = 0;
        for (auto& [key, v] : classes) {
            const dim3 g((uint32_t)v.size());
            const uint32_t lds = key.second;
            Rep* rp = e->reps_dev;
            const uint32_t* ix = e->idx_dev + first;
            switch (key.first) {
#define CPM_SL(k, d, sg, nt)                                                                 \
    case k: hipLaunchKernelGGL((k_sublattice<d, sg, nt>), g, dim3(nt), lds, e->stream, rp, ix, \
                               e->sp3d); break;
                CPM_SL_KERNELS(CPM_SL)
#undef CPM_SL
                default: throw Error(CPM_ERR_DEVICE, "no sublattice kernel for this key");
            }
            HIP_OK(hipGetLastError());
            first += (uint32_t)v.size();
        }
        // contact counts of every replica that ran (at most 65535 per launch)
        for (size_t at = 0; at < order.size(); at += 65535) {
            const dim3 g(64, (uint32_t)std::min<size_t>(order.size() - at, 65535));
            hipLaunchKernelGGL(k_contact_clear, g, dim3(256), 0, e->stream, e->reps_dev,
                               e->idx_dev + at);
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_contact_count, g, dim3(256), 0, e->stream, e->reps_dev,
                               e->idx_dev + at);
            HIP_OK(hipGetLastError());
        }
        // the running energy is recomputed from scratch (a running FP64 sum
        // of concurrently accepted moves would depend on their order), on
        // the device in the host's summation order
        if (!order.empty()) {
            hipLaunchKernelGGL(k_energy, dim3((uint32_t)order.size()), dim3(64), 0, e->stream,
                               e->reps_dev, e->idx_dev);
            HIP_OK(hipGetLastError());
        }
        HIP_OK(hipEventRecord(e->ev1, e->stream));
        HIP_OK(hipEventRecord(e->work_done, e->stream));
        HIP_OK(hipMemcpyAsync(reps.data(), e->reps_dev, n * sizeof(Rep), hipMemcpyDeviceToHost,
                              e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        if (counters) counters->kernel_ms += ms;
        for (size_t i = 0; i < n; ++i) {
            if (!todo[i]) continue;
            models[i]->mcs = reps[i].mcs;
            models[i]->energy = reps[i].energy;
            models[i]->device_dirty = true;
            add_counters(counters, reps[i]);
        }
    }
    parallel_for(n, [&](size_t i) {
        if (!todo[i]) return;
        Model* m = models[i];
        // EdgeList models: the mode's proposal is the uniform one, which the
        // edge list's is statistically (config.rs:42-70: the same chain
        // without its null attempts); the list is rebuilt from the new
        // lattice as EdgeList::build would (edge_list.rs:121-131)
        if (m->cfg.proposal == CPM_PROPOSAL_EDGE_LIST) {
            sync_lattice_tables(m);
            m->build_edge_list();
            const Layout L = layout_of(*m, m->dev->lds);
            hipStream_t st = copy_stream_for(e);
            HIP_OK(hipMemcpyAsync(m->dev->base + L.edge_sites, m->edge_sites.data(),
                                  m->n_sites * 4, hipMemcpyHostToDevice, st));
            HIP_OK(hipMemcpyAsync(m->dev->base + L.edge_pos, m->edge_pos.data(),
                                  m->n_padded * 4, hipMemcpyHostToDevice, st));
            HIP_OK(hipStreamSynchronize(st));
        }
    });
}

// Coarse signature of an exact-replay batch for the placement autotuner:
// dimensions present, log2 of the replica count and the batch's LDS demand
// (sigma in LDS everywhere it fits) in quarters of the chip's LDS.
static uint64_t batch_signature(Engine* e, Model* const* models, size_t n, const uint64_t* todo) {
    double demand = 0.0;
    uint64_t dims = 0, count = 0;
    for (size_t i = 0; i < n; ++i) {
        if (!todo[i] || !models[i]->dev || !models[i]->dev->wave) continue;
        demand += (double)models[i]->dev->lds_bytes;
        dims |= 1ull << models[i]->D;
        ++count;
    }
    const uint64_t q = (uint64_t)std::min(64.0, 4.0 * demand / ((double)e->n_cu * e->max_lds));
    uint64_t lg = 0;
    while ((1ull << lg) < count) ++lg;
    return dims | (lg << 8) | (q << 16);
}

// A fixed sigma placement from the environment, CPM_SIGMA_PLACEMENT =
// "smallest:<share>" or "longest:<share>" (share > 0), which skips the
// autotuner; fill 0 when unset.  Profilers that serialise kernels (counter
// collection) change the tuner's timings; fixing the placement keeps a
// profiled run's launches those of an unprofiled one.
static Placement fixed_placement() {
    static const Placement v = [] {
        const char* s = std::getenv("CPM_SIGMA_PLACEMENT");
        if (!s || !*s) return Placement{0.0, false};
        const char* colon = std::strchr(s, ':');
        const std::string order = colon ? std::string(s, colon) : std::string();
        char* end = nullptr;
        const double f = colon ? std::strtod(colon + 1, &end) : 0.0;
        if ((order != "smallest" && order != "longest") || end == colon + 1 || *end || !(f > 0.0))
            throw Error(CPM_ERR_ARG,
                        "CPM_SIGMA_PLACEMENT must be smallest:<share> or longest:<share> with "
                        "share > 0");
        return Placement{f, order == "longest"};
    }();
    return v;
}

// Exact replay with an autotuned sigma placement.  Results do not depend on
// the placement or on how the MCS are split into launches, so the first
// large enough call of a new batch signature is split: two slices of its
// MCS under each placement candidate, timed, then the rest under the fastest,
// which later calls with that signature reuse.
static void add_call_counters(cpm_counters* to, const cpm_counters& cc) {
    if (!to) return;
    to->attempts += cc.attempts;
    to->non_null += cc.non_null;
    to->topology += cc.topology;
    to->energy += cc.energy;
    to->accepted += cc.accepted;
    to->exp_near_ties += cc.exp_near_ties;
    to->kernel_ms += cc.kernel_ms;
    to->replay_launches += cc.replay_launches;
    to->replay_launch_ms += cc.replay_launch_ms;
    to->piped_stragglers += cc.piped_stragglers;
}

void run_mcs_each(Engine* e, Model* const* models, size_t n, const uint64_t* todo, int mode,
                  cpm_counters* counters, double budget_s) {
    if (mode != CPM_MODE_EXACT_REPLAY && mode != CPM_MODE_SUBLATTICE)
        throw Error(CPM_ERR_ARG, "unknown mode");
    for (size_t i = 0; i < n; ++i)
        if (todo[i] && models[i]->attempt_in_mcs && mode != CPM_MODE_EXACT_REPLAY)
            throw Error(CPM_ERR_ARG, "a model stopped mid-MCS continues in exact replay only");
    if (mode == CPM_MODE_SUBLATTICE) {
        run_sublattice(e, models, n, todo, counters);
        return;
    }
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    for (size_t i = 0; i < n; ++i) upload(e, models[i]);
    const Placement fixed = e->forced.fill > 0.0 ? e->forced : fixed_placement();
    if (budget_s > 0) {  // a time-limited call: no tuning slices
        e->last_placement = fixed.fill > 0.0 ? fixed : kDefaultPlacement;
        run_exact(e, models, n, todo, e->last_placement, counters, budget_s);
        return;
    }
    if (fixed.fill > 0.0) {
        e->last_placement = fixed;
        run_exact(e, models, n, todo, fixed, counters);
        return;
    }
    const uint64_t sig = batch_signature(e, models, n, todo);
    auto it = e->tuned_placement.find(sig);
    uint64_t max_todo = 0;
    for (size_t i = 0; i < n; ++i) max_todo = std::max(max_todo, todo[i]);
    // two slices per candidate, the second round in reverse order so that
    // the lattices' drift during the call (early MCS accept more) cancels
    const uint64_t slice = max_todo / (4 * kNumPlacements);  // 25 % of the call in all
    // short calls (serving rounds, sampling intervals of a few MCS) are not
    // worth the extra launches: they run with the default placement
    if (it != e->tuned_placement.end() || slice < kTuneMinSliceMcs || n < 2) {
        e->last_placement =
            it != e->tuned_placement.end() ? kPlacements[it->second] : kDefaultPlacement;
        run_exact(e, models, n, todo, e->last_placement, counters);
        return;
    }
    std::vector<uint64_t> left(todo, todo + n), part(n);
    double att[kNumPlacements] = {}, ms[kNumPlacements] = {};
    for (int r = 0; r < 2 * kNumPlacements; ++r) {
        const int c = r < kNumPlacements ? r : 2 * kNumPlacements - 1 - r;
        for (size_t i = 0; i < n; ++i) {
            part[i] = std::min(left[i], slice);
            left[i] -= part[i];
        }
        cpm_counters cc{};
        run_exact(e, models, n, part.data(), kPlacements[c], &cc);
        att[c] += (double)cc.attempts;
        ms[c] += cc.kernel_ms;
        add_call_counters(counters, cc);
    }
    int best = 0;
    for (int c = 1; c < kNumPlacements; ++c)
        if (att[c] * ms[best] > att[best] * ms[c]) best = c;
    // the runner-up and the best again on slices twice as long (best,
    // runner-up, runner-up, best), which decide between them
    int second = best == 0 ? 1 : 0;
    for (int c = 0; c < kNumPlacements; ++c)
        if (c != best && att[c] * ms[second] > att[second] * ms[c]) second = c;
    double att2[2] = {}, ms2[2] = {};
    for (int r = 0; r < 4; ++r) {
        const int w = (r == 0 || r == 3) ? 0 : 1;
        bool any = false;
        for (size_t i = 0; i < n; ++i) {
            part[i] = std::min(left[i], 2 * slice);
            left[i] -= part[i];
            any |= part[i] != 0;
        }
        if (!any) break;
        cpm_counters cc{};
        run_exact(e, models, n, part.data(), kPlacements[w ? second : best], &cc);
        att2[w] += (double)cc.attempts;
        ms2[w] += cc.kernel_ms;
        add_call_counters(counters, cc);
    }
    if (att2[0] > 0 && att2[1] > 0 && att2[1] * ms2[0] > att2[0] * ms2[1]) best = second;
    e->tuned_placement[sig] = best;
    e->last_placement = kPlacements[best];
    bool more = false;
    for (size_t i = 0; i < n; ++i) more |= left[i] != 0;
    if (more) run_exact(e, models, n, left.data(), kPlacements[best], counters);
}

void engine_last_placement(Engine* e, int* order, double* share) {
    std::lock_guard<std::mutex> lock(e->mu);
    *order = e->last_placement.fill > 0.0 ? (e->last_placement.longest_first ? 2 : 1) : 0;
    *share = e->last_placement.fill;
}

void engine_set_placement(Engine* e, int order, double share) {
    std::lock_guard<std::mutex> lock(e->mu);
    if (order == 0) {
        e->forced = Placement{0.0, false};
        return;
    }
    if ((order != 1 && order != 2) || !(share > 0.0))
        throw Error(CPM_ERR_ARG, "sigma placement: order must be 0, 1 or 2 and share > 0");
    e->forced = Placement{share, order == 2};
}

// exp(-dh) with the device libm, the correctly rounded exp(-dh), and the
// Metropolis decision for u, each exactly as the exact-replay kernels
// compute them (metropolis.hpp).
__global__ void k_metropolis_check(const double* dh, const double* u, uint64_t n, double* p_libm,
                                   double* p_cr, uint8_t* acc) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = dh[i];
    p_libm[i] = exp(-x);
    p_cr[i] = exp_cr(-x);
    if (u) {
        bool tie = false;
        acc[i] = metropolis_accept(u[i], x, &tie) ? 1 : 0;
    }
}

void engine_metropolis_check(Engine* e, const double* dh, const double* u, size_t n,
                             double* p_libm, double* p_cr, uint8_t* acc) {
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    if (!n) return;
    const size_t bytes = n * sizeof(double);
    char* buf = nullptr;
    const size_t total = (u ? 4 : 3) * bytes + n;
    HIP_OK(hipMalloc(&buf, total));
    double* d_dh = (double*)buf;
    double* d_pl = (double*)(buf + bytes);
    double* d_pc = (double*)(buf + 2 * bytes);
    double* d_u = u ? (double*)(buf + 3 * bytes) : nullptr;
    uint8_t* d_acc = u ? (uint8_t*)(buf + 4 * bytes) : nullptr;
    HIP_OK(hipMemcpyAsync(d_dh, dh, bytes, hipMemcpyHostToDevice, e->stream));
    if (u) HIP_OK(hipMemcpyAsync(d_u, u, bytes, hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_metropolis_check, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                       e->stream, d_dh, d_u, (uint64_t)n, d_pl, d_pc, d_acc);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(p_libm, d_pl, bytes, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipMemcpyAsync(p_cr, d_pc, bytes, hipMemcpyDeviceToHost, e->stream));
    if (u) HIP_OK(hipMemcpyAsync(acc, d_acc, n, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    HIP_OK(hipFree(buf));
}

__global__ void k_act_root_check(const double* x, const uint32_t* cnt, uint64_t n, double* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = act_root(x[i], cnt[i]);
}

__global__ void k_mh_check(const uint32_t* na, const uint32_t* nb, const double* dh,
                           const double* u, uint64_t n, uint8_t* acc) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool tie = false;
    acc[i] = mh_accept(u[i], na[i], nb[i], dh[i], &tie) ? 1 : 0;
}

__global__ void k_binomial_check(uint64_t* st, const uint64_t* n, const double* p, uint64_t cnt,
                                 uint64_t* k, uint32_t* bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    Xoshiro r;
    for (int j = 0; j < 4; ++j) r.s[j] = st[4 * i + j];
    if (!binomial_sample(r, n[i], p[i], &k[i])) atomicOr(bad, 1u);
    for (int j = 0; j < 4; ++j) st[4 * i + j] = r.s[j];
}

void engine_binomial_check(Engine* e, uint64_t* rng_state, const uint64_t* n, const double* p,
                           size_t count, uint64_t* k) {
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    if (!count) return;
    char* buf = nullptr;
    HIP_OK(hipMalloc(&buf, count * 56 + 256));
    uint64_t* d_st = (uint64_t*)buf;
    uint64_t* d_n = (uint64_t*)(buf + count * 32);
    double* d_p = (double*)(buf + count * 40);
    uint64_t* d_k = (uint64_t*)(buf + count * 48);
    uint32_t* d_bad = (uint32_t*)(buf + count * 56);
    const uint32_t zero = 0;
    HIP_OK(hipMemcpyAsync(d_st, rng_state, count * 32, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_n, n, count * 8, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_p, p, count * 8, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_bad, &zero, 4, hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_binomial_check, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0,
                       e->stream, d_st, d_n, d_p, (uint64_t)count, d_k, d_bad);
    HIP_OK(hipGetLastError());
    uint32_t bad = 0;
    HIP_OK(hipMemcpyAsync(rng_state, d_st, count * 32, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipMemcpyAsync(k, d_k, count * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    HIP_OK(hipFree(buf));
    if (bad) throw Error(CPM_ERR_DEVICE, "the binomial sampler ran away");
}

void engine_act_root_check(Engine* e, const double* product, const uint32_t* count, size_t n,
                           double* out) {
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    if (!n) return;
    char* buf = nullptr;
    HIP_OK(hipMalloc(&buf, n * 20));
    double* d_x = (double*)buf;
    double* d_out = (double*)(buf + n * 8);
    uint32_t* d_c = (uint32_t*)(buf + n * 16);
    HIP_OK(hipMemcpyAsync(d_x, product, n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_c, count, n * 4, hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_act_root_check, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                       e->stream, d_x, d_c, (uint64_t)n, d_out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, d_out, n * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    HIP_OK(hipFree(buf));
}

void engine_mh_check(Engine* e, const uint32_t* n_a, const uint32_t* n_b, const double* dh,
                     const double* u, size_t n, uint8_t* accept) {
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    if (!n) return;
    char* buf = nullptr;
    HIP_OK(hipMalloc(&buf, n * 25));
    double* d_dh = (double*)buf;
    double* d_u = (double*)(buf + n * 8);
    uint32_t* d_na = (uint32_t*)(buf + n * 16);
    uint32_t* d_nb = (uint32_t*)(buf + n * 20);
    uint8_t* d_acc = (uint8_t*)(buf + n * 24);
    HIP_OK(hipMemcpyAsync(d_dh, dh, n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_u, u, n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_na, n_a, n * 4, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(d_nb, n_b, n * 4, hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_mh_check, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, e->stream,
                       d_na, d_nb, d_dh, d_u, (uint64_t)n, d_acc);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(accept, d_acc, n, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    HIP_OK(hipFree(buf));
}

void engine_set_pipelining(Engine* e, bool on) {
    std::lock_guard<std::mutex> lock(e->mu);
    e->pipelining = on;
}

std::vector<std::pair<int, uint64_t>> engine_launch_counts(Engine* e) {
    std::lock_guard<std::mutex> lock(e->mu);
    return {e->launches_by_key.begin(), e->launches_by_key.end()};
}

void run_mcs_noalloc(Engine* e, Model* const* models, size_t n, uint64_t n_mcs,
                     cpm_counters* counters) {
    for (size_t i = 0; i < n; ++i)
        if (models[i]->attempt_in_mcs)
            throw Error(CPM_ERR_ARG, "a model stopped mid-MCS continues in exact replay only");
    std::lock_guard<std::mutex> lock(e->mu);
    HIP_OK(hipSetDevice(e->device));
    std::vector<uint64_t> todo(n, n_mcs);
    const Placement fixed = e->forced.fill > 0.0 ? e->forced : fixed_placement();
    Placement pl = kDefaultPlacement;
    if (fixed.fill > 0.0) {
        pl = fixed;
    } else {
        auto it = e->tuned_placement.find(batch_signature(e, models, n, todo.data()));
        if (it != e->tuned_placement.end()) pl = kPlacements[it->second];
    }
    run_exact(e, models, n, todo.data(), pl, counters, 0.0, true);
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

namespace cpm {

// initialization::initialize of many models: scatter_and_grow on the device
// for the default-initialised ones (at most kInitMaxCells cells, uniform
// proposal, no Act term), placed and grown in one launch (one wave per
// model) straight into each model's own device allocation, and finished
// there as initialization.rs:296-305 does (halo, interfaces, position sums,
// contacts, energy): the models start device-resident, their host arrays
// filled at the first sync.  Explicit placements and the other models
// initialise on the host threads.  On failure the error of the lowest
// failing index is thrown, with the reference's message.
void device_initialize(Engine* e, Model* const* ms, size_t n) {
    std::vector<std::exception_ptr> errs(n);
    std::vector<ScatterPlan> plans(n);
    std::vector<size_t> on_dev, on_host;
    std::vector<uint64_t> cap_of(n, 0);
    for (size_t i = 0; i < n; ++i) {
        Model& m = *ms[i];
        bool dev = !m.cfg.init_explicit && m.n_padded < (1ull << 31) &&
                   m.cfg.proposal == CPM_PROPOSAL_UNIFORM && !m.cfg.act;
        if (dev) {
            try {
                plans[i] = m.scatter_plan();
            } catch (...) {
                errs[i] = std::current_exception();
                continue;
            }
            dev = plans[i].type_of.size() <= kInitMaxCells;
            // a cell's frontier holds at most n_copy sites per site it claims
            uint64_t cap = 0;
            for (uint32_t t : plans[i].type_of)
                cap += (uint64_t)m.n_copy * std::max<uint64_t>(m.cfg.types[t].target_volume, 1);
            cap_of[i] = cap;
            dev = dev && cap < (1ull << 28);
        }
        (dev ? on_dev : on_host).push_back(i);
    }
    parallel_for(on_host.size(), [&](size_t k) {
        try {
            ms[on_host[k]]->initialize();
        } catch (...) {
            errs[on_host[k]] = std::current_exception();
        }
    });
    if (!on_dev.empty()) {
        const size_t nj = on_dev.size();
        std::lock_guard<std::mutex> lock(e->mu);
        HIP_OK(hipSetDevice(e->device));
        // each model's device allocation and parameters (the state follows)
        for (size_t k = 0; k < nj; ++k) {
            Model& m = *ms[on_dev[k]];
            const ScatterPlan& p = plans[on_dev[k]];
            device_release(&m);
            m.initialized = false;
            m.n_cells = p.type_of.size();
            m.type_index = p.type_of;
            upload(e, &m, false);
        }
        // one scratch allocation: jobs, then per model seeds, frontier
        // offsets and targets, then the frontiers
        std::vector<size_t> aux_off(nj), fr_off(nj);
        size_t off = align256(nj * sizeof(InitJob));
        const size_t aux_begin = off;
        for (size_t k = 0; k < nj; ++k) {
            aux_off[k] = off;
            off += align256(plans[on_dev[k]].type_of.size() * 4 * 5);
        }
        const size_t aux_end = off;
        for (size_t k = 0; k < nj; ++k) {
            fr_off[k] = off;
            off += align256(cap_of[on_dev[k]] * 4);
        }
        char* buf = nullptr;
        HIP_OK(hipMalloc(&buf, off));
        std::vector<InitJob> jobs(nj);
        std::vector<Rep> reps(nj);
        std::vector<char> host_aux(aux_end - aux_begin);  // fr_start and target, staged
        uint32_t max_cells = 1;
        for (size_t k = 0; k < nj; ++k) {
            const Model& m = *ms[on_dev[k]];
            const ScatterPlan& p = plans[on_dev[k]];
            const uint32_t cells = (uint32_t)p.type_of.size();
            max_cells = std::max(max_cells, cells);
            reps[k] = m.dev->rep;
            InitJob& J = jobs[k];
            std::memset(&J, 0, sizeof J);
            char* a = buf + aux_off[k];
            J.sigma = reps[k].sigma;
            J.volume = reps[k].volume;
            J.seeds = (uint32_t*)a;
            J.fr_start = (const uint32_t*)(a + cells * 12);
            J.target = (const uint32_t*)(a + cells * 16);
            J.frontier = (uint32_t*)(buf + fr_off[k]);
            for (int q = 0; q < 4; ++q) J.rng[q] = m.rng.s[q];
            J.sep_sq = p.sep_sq;
            for (int d = 0; d < 3; ++d) {
                J.dims[d] = d < m.D ? (uint32_t)m.dims[d] : 1u;
                J.strides[d] = d < m.D ? (uint32_t)m.strides[d] : 1u;
            }
            int32_t o[26 * 3];
            const int nc = stencil(m.D, m.cfg.copy_nb, o);
            for (int q = 0; q < nc; ++q)
                for (int d = 0; d < m.D; ++d) J.copy_vec[q][d] = o[q * m.D + d];
            J.n_cells = cells;
            J.n_copy = (uint32_t)nc;
            J.D = (uint32_t)m.D;
            J.periodic = m.cfg.boundary == CPM_BOUNDARY_PERIODIC ? 1u : 0u;
            J.n_padded = (uint32_t)m.n_padded;
            char* h = host_aux.data() + (aux_off[k] - aux_begin);
            uint32_t* fs = (uint32_t*)(h + cells * 12);
            uint32_t* tg = (uint32_t*)(h + cells * 16);
            uint64_t at = 0;
            for (uint32_t c = 0; c < cells; ++c) {
                const uint32_t t = m.cfg.types[p.type_of[c]].target_volume;
                fs[c] = (uint32_t)at;
                tg[c] = t;
                at += (uint64_t)nc * std::max<uint32_t>(t, 1u);
            }
        }
        HIP_OK(hipMemcpyAsync(buf, jobs.data(), nj * sizeof(InitJob), hipMemcpyHostToDevice,
                              e->stream));
        HIP_OK(hipMemcpyAsync(buf + aux_begin, host_aux.data(), host_aux.size(),
                              hipMemcpyHostToDevice, e->stream));
        hipLaunchKernelGGL(k_scatter_grow, dim3((uint32_t)nj), dim3(64), max_cells * 12, e->stream,
                           (InitJob*)buf);
        HIP_OK(hipGetLastError());
        // the derived state, over the models' descriptors (at most 65535 per launch)
        ensure_buf(e->reps_dev, e->reps_cap, nj);
        ensure_buf(e->idx_dev, e->idx_cap, nj);
        std::vector<uint32_t> order(nj);
        for (size_t k = 0; k < nj; ++k) order[k] = (uint32_t)k;
        HIP_OK(hipMemcpyAsync(e->reps_dev, reps.data(), nj * sizeof(Rep), hipMemcpyHostToDevice,
                              e->stream));
        HIP_OK(hipMemcpyAsync(e->idx_dev, order.data(), nj * 4, hipMemcpyHostToDevice, e->stream));
        for (size_t at = 0; at < nj; at += 65535) {
            const dim3 g(64, (uint32_t)std::min<size_t>(nj - at, 65535));
            hipLaunchKernelGGL(k_halo, g, dim3(256), 0, e->stream, e->reps_dev + at);
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_state_clear, g, dim3(256), 0, e->stream, e->reps_dev + at);
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_recompute_cells, g, dim3(256), 0, e->stream, e->reps_dev + at);
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_contact_count, g, dim3(256), 0, e->stream, e->reps_dev,
                               e->idx_dev + at);
            HIP_OK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_energy, dim3((uint32_t)nj), dim3(64), 0, e->stream, e->reps_dev,
                           e->idx_dev);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(e->work_done, e->stream));
        HIP_OK(hipMemcpyAsync(jobs.data(), buf, nj * sizeof(InitJob), hipMemcpyDeviceToHost,
                              e->stream));
        HIP_OK(hipMemcpyAsync(reps.data(), e->reps_dev, nj * sizeof(Rep), hipMemcpyDeviceToHost,
                              e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
        HIP_OK(hipFree(buf));
        for (size_t k = 0; k < nj; ++k) {
            Model& m = *ms[on_dev[k]];
            if (jobs[k].status) {
                device_release(&m);
                try {
                    m.crowded_error(plans[on_dev[k]]);
                } catch (...) {
                    errs[on_dev[k]] = std::current_exception();
                }
                continue;
            }
            for (int q = 0; q < 4; ++q) m.rng.s[q] = jobs[k].rng[q];
            m.energy = reps[k].energy;
            m.used_default = true;
            m.initialized = true;
            m.device_dirty = true;  // the host arrays are filled at the first sync
        }
    }
    for (auto& err : errs)
        if (err) std::rethrow_exception(err);
}

}  // namespace cpm

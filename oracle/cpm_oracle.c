/*
 * CPU oracle for the Metropolis copy-attempt engine — TEST INFRASTRUCTURE
 * ONLY (tests, smoke check, CPU-baseline leg of the benchmark).  See
 * cpm_oracle.h for what each function restates and what pins it.
 *
 * Build: make -C oracle   (gcc -O3 -ffp-contract=off, no fast-math: the
 * reference's FP64 energy arithmetic is never contracted or reassociated).
 */
#include "cpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* RNG: rand_xoshiro 0.6.0 Xoshiro256PlusPlus + rand 0.8 sampling       */
/* (rng.rs:7-14; monte_carlo.rs:39,57,158; initialization.rs:235,290,313) */
/* ------------------------------------------------------------------ */

static inline uint64_t rotl64(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
}

uint64_t oc_rng_next(uint64_t s[4]) {
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* SeedableRng::seed_from_u64: SplitMix64 expansion into a 32-byte seed,
   little-endian words, all-zero seed remapped to seed_from_u64(0). */
void oc_rng_seed(uint64_t seed, uint64_t out[4]) {
    uint64_t state = seed;
    for (int i = 0; i < 4; ++i) {
        state += 0x9e3779b97f4a7c15ULL;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        out[i] = z ^ (z >> 31);
    }
    if ((out[0] | out[1] | out[2] | out[3]) == 0) oc_rng_seed(0, out);
}

/* UniformInt<usize>::sample_single_inclusive(0, hi): widening multiply
   with the conservative rejection zone (range << lz(range)) - 1. */
uint64_t oc_rng_gen_range_incl(uint64_t s[4], uint64_t hi) {
    const uint64_t range = hi + 1;
    if (range == 0) return oc_rng_next(s);
    const uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
        const uint64_t v = oc_rng_next(s);
        const unsigned __int128 m = (unsigned __int128)v * range;
        const uint64_t lo = (uint64_t)m;
        if (lo <= zone) return (uint64_t)(m >> 64);
    }
}

/* gen_range(0..n) == sample_single(0, n) == sample_single_inclusive(0, n-1) */
uint64_t oc_rng_gen_range(uint64_t s[4], uint64_t n) {
    return oc_rng_gen_range_incl(s, n - 1);
}

/* Standard f64: 53 high bits scaled by 2^-53. */
double oc_rng_gen_f64(uint64_t s[4]) {
    const uint64_t v = oc_rng_next(s);
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------ */
/* BLAKE3 (the published specification: 7-round compression over a     */
/* binary tree of 1 KiB chunks), used by rng.rs:20-40 and config.rs.    */
/* ------------------------------------------------------------------ */

static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u,
                                  0xA54FF53Au, 0x510E527Fu, 0x9B05688Cu,
                                  0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13,
                                    1, 11, 12, 5, 9, 14, 15, 8};
enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

static inline uint32_t rotr32(uint32_t x, int n) {
    return (x >> n) | (x << (32 - n));
}

static inline void b3_g(uint32_t* st, int a, int b, int c, int d, uint32_t mx,
                        uint32_t my) {
    st[a] = st[a] + st[b] + mx;
    st[d] = rotr32(st[d] ^ st[a], 16);
    st[c] = st[c] + st[d];
    st[b] = rotr32(st[b] ^ st[c], 12);
    st[a] = st[a] + st[b] + my;
    st[d] = rotr32(st[d] ^ st[a], 8);
    st[c] = st[c] + st[d];
    st[b] = rotr32(st[b] ^ st[c], 7);
}

static void b3_compress(const uint32_t cv[8], const uint32_t block[16],
                        uint64_t counter, uint32_t block_len, uint32_t flags,
                        uint32_t out[16]) {
    uint32_t st[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                       B3_IV[0], B3_IV[1], B3_IV[2], B3_IV[3],
                       (uint32_t)counter, (uint32_t)(counter >> 32),
                       block_len, flags};
    uint32_t m[16];
    memcpy(m, block, sizeof m);
    for (int r = 0; r < 7; ++r) {
        b3_g(st, 0, 4, 8, 12, m[0], m[1]);
        b3_g(st, 1, 5, 9, 13, m[2], m[3]);
        b3_g(st, 2, 6, 10, 14, m[4], m[5]);
        b3_g(st, 3, 7, 11, 15, m[6], m[7]);
        b3_g(st, 0, 5, 10, 15, m[8], m[9]);
        b3_g(st, 1, 6, 11, 12, m[10], m[11]);
        b3_g(st, 2, 7, 8, 13, m[12], m[13]);
        b3_g(st, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            uint32_t p[16];
            for (int i = 0; i < 16; ++i) p[i] = m[B3_PERM[i]];
            memcpy(m, p, sizeof m);
        }
    }
    for (int i = 0; i < 8; ++i) {
        out[i] = st[i] ^ st[i + 8];
        out[i + 8] = st[i + 8] ^ cv[i];
    }
}

static void b3_words(const uint8_t* bytes, size_t len, uint32_t w[16]) {
    uint8_t buf[64];
    memset(buf, 0, sizeof buf);
    memcpy(buf, bytes, len);
    for (int i = 0; i < 16; ++i)
        w[i] = (uint32_t)buf[4 * i] | ((uint32_t)buf[4 * i + 1] << 8) |
               ((uint32_t)buf[4 * i + 2] << 16) | ((uint32_t)buf[4 * i + 3] << 24);
}

/* Chaining value of one chunk (chunk_counter), or, when is_root, the root
   output words of the single-chunk tree. */
static void b3_chunk(const uint8_t* data, size_t len, uint64_t chunk_counter,
                     int is_root, uint32_t out_cv[8]) {
    uint32_t cv[8];
    memcpy(cv, B3_IV, sizeof cv);
    size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (size_t b = 0; b < nblocks; ++b) {
        size_t off = b * 64;
        size_t blen = len - off < 64 ? len - off : 64;
        uint32_t w[16], out[16];
        b3_words(data + off, blen, w);
        uint32_t flags = 0;
        if (b == 0) flags |= B3_CHUNK_START;
        if (b == nblocks - 1) {
            flags |= B3_CHUNK_END;
            if (is_root) flags |= B3_ROOT;
        }
        b3_compress(cv, w, chunk_counter, (uint32_t)blen, flags, out);
        memcpy(cv, out, sizeof cv);
    }
    memcpy(out_cv, cv, sizeof cv);
}

static void b3_parent(const uint32_t l[8], const uint32_t r[8], int is_root,
                      uint32_t out_cv[8]) {
    uint32_t block[16], out[16];
    memcpy(block, l, 32);
    memcpy(block + 8, r, 32);
    b3_compress(B3_IV, block, 0, 64, B3_PARENT | (is_root ? B3_ROOT : 0), out);
    memcpy(out_cv, out, 32);
}

/* Recursive tree: the left subtree holds the largest power-of-two number of
   chunks strictly less than the total (the spec's tree shape). */
static void b3_subtree(const uint8_t* data, size_t len, uint64_t chunk0,
                       int is_root, uint32_t out_cv[8]) {
    if (len <= 1024) {
        b3_chunk(data, len, chunk0, is_root, out_cv);
        return;
    }
    size_t nchunks = (len + 1023) / 1024;
    size_t left_chunks = 1;
    while (left_chunks * 2 < nchunks) left_chunks *= 2;
    size_t left_len = left_chunks * 1024;
    uint32_t l[8], r[8];
    b3_subtree(data, left_len, chunk0, 0, l);
    b3_subtree(data + left_len, len - left_len, chunk0 + left_chunks, 0, r);
    b3_parent(l, r, is_root, out_cv);
}

void oc_blake3(const uint8_t* data, size_t len, uint8_t out[32]) {
    uint32_t cv[8];
    b3_subtree(data, len, 0, 1, cv);
    for (int i = 0; i < 8; ++i) {
        out[4 * i] = (uint8_t)cv[i];
        out[4 * i + 1] = (uint8_t)(cv[i] >> 8);
        out[4 * i + 2] = (uint8_t)(cv[i] >> 16);
        out[4 * i + 3] = (uint8_t)(cv[i] >> 24);
    }
}

/* rng.rs:20-25, 37-40: BLAKE3(LE(master) || LE(index)), first 8 bytes LE */
uint64_t oc_derive_seed_u64(uint64_t master_seed, uint64_t index) {
    uint8_t in[16], h[32];
    for (int i = 0; i < 8; ++i) {
        in[i] = (uint8_t)(master_seed >> (8 * i));
        in[8 + i] = (uint8_t)(index >> (8 * i));
    }
    oc_blake3(in, 16, h);
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | h[i];
    return v;
}

/* ------------------------------------------------------------------ */
/* Stencils (neighborhood.rs:45-115)                                   */
/* ------------------------------------------------------------------ */

int oc_full_offsets(int dim, int32_t* out) {
    int total = dim == 2 ? 9 : 27, n = 0;
    for (int code = 0; code < total; ++code) {
        int c = code, nz = 0, o[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            o[d] = c % 3 - 1;
            c /= 3;
            if (o[d]) ++nz;
        }
        if (!nz) continue;
        for (int d = 0; d < dim; ++d) out[n * dim + d] = o[d];
        ++n;
    }
    return n;
}

/* Fills offs (n x dim) for a stencil kind; returns count. */
static int stencil_offsets(int dim, int kind, int32_t* offs) {
    if (kind == OC_NB_VON_NEUMANN) {
        int n = 0;
        for (int axis = 0; axis < dim; ++axis) {
            for (int sgn = 1; sgn >= -1; sgn -= 2) {
                for (int d = 0; d < dim; ++d) offs[n * dim + d] = d == axis ? sgn : 0;
                ++n;
            }
        }
        return n;
    }
    int32_t full[26 * 3];
    int nf = oc_full_offsets(dim, full);
    if (kind == OC_NB_EIGHTEEN) {
        int n = 0;
        for (int i = 0; i < nf; ++i) {
            int nz = 0;
            for (int d = 0; d < dim; ++d) nz += full[i * dim + d] != 0;
            if (nz <= 2) {
                memcpy(offs + n * dim, full + i * dim, sizeof(int32_t) * dim);
                ++n;
            }
        }
        return n;
    }
    memcpy(offs, full, sizeof(int32_t) * nf * dim);
    return nf;
}

/* ------------------------------------------------------------------ */
/* Simple points (simple_point.rs:35-366)                              */
/* ------------------------------------------------------------------ */

static int face_adj(const int32_t* a, const int32_t* b, int dim) {
    int m = 0;
    for (int d = 0; d < dim; ++d) m += abs(a[d] - b[d]);
    return m == 1;
}

static int full_adj(const int32_t* a, const int32_t* b, int dim) {
    int same = 1;
    for (int d = 0; d < dim; ++d) {
        if (abs(a[d] - b[d]) > 1) return 0;
        if (a[d] != b[d]) same = 0;
    }
    return !same;
}

/* Slow path: explicit stack flood fill over index lists. */
static int flood_count(const int32_t* offs, int n, int dim, const int* mask,
                       int use_face, const int32_t* touch_origin_of,
                       int* n_touching) {
    int visited[26] = {0}, stack[26], count = 0, touching = 0;
    for (int start = 0; start < n; ++start) {
        if (!mask[start] || visited[start]) continue;
        int sp = 0, touches = 0;
        visited[start] = 1;
        stack[sp++] = start;
        while (sp) {
            int i = stack[--sp];
            if (touch_origin_of && face_adj(offs + i * dim, touch_origin_of, dim)) touches = 1;
            for (int j = 0; j < n; ++j) {
                if (!mask[j] || visited[j]) continue;
                int adj = use_face ? face_adj(offs + i * dim, offs + j * dim, dim)
                                   : full_adj(offs + i * dim, offs + j * dim, dim);
                if (adj) {
                    visited[j] = 1;
                    stack[sp++] = j;
                }
            }
        }
        ++count;
        touching += touches;
    }
    if (n_touching) *n_touching = touching;
    return count;
}

int oc_sp_slow(int dim, uint32_t pattern) {
    int32_t offs[26 * 3];
    int n = oc_full_offsets(dim, offs);
    int obj[26], comp[26];
    for (int i = 0; i < n; ++i) {
        obj[i] = (pattern >> i) & 1;
        comp[i] = !obj[i];
    }
    if (flood_count(offs, n, dim, comp, 0, NULL, NULL) != 1) return 0;
    const int32_t origin[3] = {0, 0, 0};
    int touching = 0;
    flood_count(offs, n, dim, obj, 1, origin, &touching);
    return touching == 1;
}

typedef struct {
    int n;
    uint32_t face[26], full[26], origin_face, all;
} sp_checker;

static void sp_checker_init(sp_checker* c, int dim) {
    int32_t offs[26 * 3];
    const int32_t origin[3] = {0, 0, 0};
    c->n = oc_full_offsets(dim, offs);
    c->origin_face = 0;
    for (int i = 0; i < c->n; ++i) {
        c->face[i] = c->full[i] = 0;
        for (int j = 0; j < c->n; ++j) {
            if (i == j) continue;
            if (face_adj(offs + i * dim, offs + j * dim, dim)) c->face[i] |= 1u << j;
            if (full_adj(offs + i * dim, offs + j * dim, dim)) c->full[i] |= 1u << j;
        }
        if (face_adj(offs + i * dim, origin, dim)) c->origin_face |= 1u << i;
    }
    c->all = (uint32_t)((1ull << c->n) - 1);
}

static uint32_t sp_fill(uint32_t seed, uint32_t universe, const uint32_t* adj) {
    uint32_t comp = seed, frontier = seed;
    while (frontier) {
        uint32_t next = 0;
        while (frontier) {
            int i = __builtin_ctz(frontier);
            frontier &= frontier - 1;
            next |= adj[i];
        }
        next &= universe & ~comp;
        comp |= next;
        frontier = next;
    }
    return comp;
}

static int sp_checker_is_simple(const sp_checker* c, uint32_t obj) {
    uint32_t comp = ~obj & c->all;
    if (!comp) return 0;
    uint32_t first = sp_fill(comp & -comp, comp, c->full);
    if (first != comp) return 0;
    uint32_t rem = obj;
    int touching = 0;
    while (rem) {
        uint32_t part = sp_fill(rem & -rem, obj, c->face);
        rem &= ~part;
        if (part & c->origin_face) ++touching;
    }
    return touching == 1;
}

int oc_sp_fast(int dim, uint32_t pattern) {
    sp_checker c;
    sp_checker_init(&c, dim);
    return sp_checker_is_simple(&c, pattern);
}

void oc_sp_table2d(uint8_t out[256]) {
    sp_checker c;
    sp_checker_init(&c, 2);
    for (uint32_t p = 0; p < 256; ++p) out[p] = (uint8_t)sp_checker_is_simple(&c, p);
}

void oc_sp_table3d(uint64_t* words) {
    sp_checker c;
    sp_checker_init(&c, 3);
#pragma omp parallel for schedule(static)
    for (long w = 0; w < (1L << 20); ++w) {
        uint64_t bits = 0;
        for (int b = 0; b < 64; ++b)
            if (sp_checker_is_simple(&c, (uint32_t)(w * 64 + b))) bits |= 1ull << b;
        words[w] = bits;
    }
}

static const uint64_t* table3d_cached(void) {
    static uint64_t* t = NULL;
    if (!t) {
        t = (uint64_t*)malloc(sizeof(uint64_t) << 20);
        oc_sp_table3d(t);
    }
    return t;
}

/* ------------------------------------------------------------------ */
/* Model                                                               */
/* ------------------------------------------------------------------ */

struct ocpm {
    int D, boundary, halo_mode;
    size_t dims[3], pdims[3], strides[3], n_sites, n_padded;
    uint32_t* cells;
    size_t n_cells;
    uint32_t *type_index, *volume, *interface, *contact;
    int64_t *sum_pos, *sum_outer;
    double energy;
    int n_types;
    double *lambda_v, *lambda_i, *J;
    uint32_t *target_v, *target_i;
    ptrdiff_t copy_off[26], energy_off[26], topo_off[26], conn_off[26], full_off[26];
    double energy_w[26];
    int n_copy, n_energy, n_topo, n_conn;
    uint8_t table2d[256];
    uint64_t rng[4];
    uint64_t mcs;
    uint64_t counters[5];
    /* optional record of the Metropolis draws: (dH, u) per uphill test */
    double* mlog;
    uint64_t mlog_cap, mlog_n;
    /* acceptance rule (OC_METROPOLIS / OC_METROPOLIS_HASTINGS), the Act
       term (config.rs:84-103) and the connectivity check of run_mcs
       (off only for the exact-Boltzmann checks: monte_carlo.rs:30-42) */
    int acceptance, act, check_conn;
    double* lambda_act;
    uint32_t* max_act;
    uint32_t* birth; /* energy.rs:337-344: padded-lattice indexed, all 0 */
    /* ProposalMode::EdgeList (edge_list.rs:42-186): the edge sites in
       their list order and each padded site's position in it */
    int edge_list;
    size_t* edge_sites;
    uint32_t* edge_pos;
    size_t n_edge;
};

static ptrdiff_t off_delta(const ocpm* m, const int32_t* o) {
    ptrdiff_t d = 0;
    for (int k = 0; k < m->D; ++k) d += (ptrdiff_t)o[k] * (ptrdiff_t)m->strides[k];
    return d;
}

static void padded_coord(const ocpm* m, size_t flat, size_t* pc) {
    for (int d = 0; d < m->D; ++d) {
        pc[d] = flat / m->strides[d];
        flat %= m->strides[d];
    }
}

static int is_interior(const ocpm* m, size_t flat) {
    size_t pc[3];
    padded_coord(m, flat, pc);
    for (int d = 0; d < m->D; ++d)
        if (pc[d] == 0 || pc[d] == m->pdims[d] - 1) return 0;
    return 1;
}

static size_t flat_of(const ocpm* m, const size_t* c) {
    size_t f = 0;
    for (int d = 0; d < m->D; ++d) f += (c[d] + 1) * m->strides[d];
    return f;
}

/* lattice.rs:227-231 (wrapped_interior_flat) */
static size_t wrapped_flat(const ocpm* m, size_t flat) {
    size_t pc[3], ic[3];
    padded_coord(m, flat, pc);
    for (int d = 0; d < m->D; ++d) {
        if (pc[d] == 0) ic[d] = m->dims[d] - 1;
        else if (pc[d] == m->pdims[d] - 1) ic[d] = 0;
        else ic[d] = pc[d] - 1;
    }
    return flat_of(m, ic);
}

/* lattice.rs:240-254: whole-halo rescan */
static void refresh_halo_full(ocpm* m) {
    for (size_t f = 0; f < m->n_padded; ++f) {
        if (is_interior(m, f)) continue;
        m->cells[f] = m->cells[wrapped_flat(m, f)];
    }
}

/* Only the halo mirrors of interior coordinate c (equivalent to the full
   rescan when the halo was consistent before the write). */
static void refresh_halo_mirrors(ocpm* m, const size_t* c) {
    int nopt[3];
    size_t opt[3][3];
    for (int d = 0; d < m->D; ++d) {
        nopt[d] = 1;
        opt[d][0] = c[d] + 1;
        if (c[d] == 0) opt[d][nopt[d]++] = m->dims[d] + 1;
        if (c[d] == m->dims[d] - 1) opt[d][nopt[d]++] = 0;
    }
    size_t src = flat_of(m, c);
    int total = 1;
    for (int d = 0; d < m->D; ++d) total *= nopt[d];
    for (int k = 1; k < total; ++k) {
        int r = k;
        size_t f = 0;
        for (int d = m->D - 1; d >= 0; --d) {
            f += opt[d][r % nopt[d]] * m->strides[d];
            r /= nopt[d];
        }
        m->cells[f] = m->cells[src];
    }
}

static void interior_coord(const ocpm* m, size_t flat, size_t* c) {
    padded_coord(m, flat, c);
    for (int d = 0; d < m->D; ++d) c[d] -= 1;
}

ocpm* oc_model_new(int dim, const uint64_t* dims, int boundary, int copy_nb,
                   int energy_nb, int n_types, const double* lambda_v,
                   const double* lambda_i, const uint32_t* target_v,
                   const uint32_t* target_i, const double* adhesion,
                   uint64_t seed) {
    ocpm* m = (ocpm*)calloc(1, sizeof(ocpm));
    m->D = dim;
    m->boundary = boundary;
    m->halo_mode = OC_HALO_FULL_RESCAN;
    m->n_sites = 1;
    m->n_padded = 1;
    for (int d = 0; d < dim; ++d) {
        m->dims[d] = dims[d];
        m->pdims[d] = dims[d] + 2;
        m->n_sites *= dims[d];
        m->n_padded *= m->pdims[d];
    }
    m->strides[dim - 1] = 1;
    for (int d = dim - 2; d >= 0; --d) m->strides[d] = m->strides[d + 1] * m->pdims[d + 1];
    m->cells = (uint32_t*)calloc(m->n_padded, sizeof(uint32_t));
    m->n_types = n_types;
    m->lambda_v = (double*)malloc(sizeof(double) * n_types);
    m->lambda_i = (double*)malloc(sizeof(double) * n_types);
    m->target_v = (uint32_t*)malloc(sizeof(uint32_t) * n_types);
    m->target_i = (uint32_t*)malloc(sizeof(uint32_t) * n_types);
    memcpy(m->lambda_v, lambda_v, sizeof(double) * n_types);
    memcpy(m->lambda_i, lambda_i, sizeof(double) * n_types);
    memcpy(m->target_v, target_v, sizeof(uint32_t) * n_types);
    memcpy(m->target_i, target_i, sizeof(uint32_t) * n_types);
    m->J = (double*)malloc(sizeof(double) * (n_types + 1) * (n_types + 1));
    memcpy(m->J, adhesion, sizeof(double) * (n_types + 1) * (n_types + 1));
    int32_t o[26 * 3];
    m->n_copy = stencil_offsets(dim, copy_nb, o);
    for (int i = 0; i < m->n_copy; ++i) m->copy_off[i] = off_delta(m, o + i * dim);
    m->n_energy = stencil_offsets(dim, energy_nb, o);
    for (int i = 0; i < m->n_energy; ++i) {
        m->energy_off[i] = off_delta(m, o + i * dim);
        m->energy_w[i] = 1.0;
    }
    m->n_conn = stencil_offsets(dim, OC_NB_VON_NEUMANN, o);
    for (int i = 0; i < m->n_conn; ++i) m->conn_off[i] = off_delta(m, o + i * dim);
    m->n_topo = oc_full_offsets(dim, o);
    for (int i = 0; i < m->n_topo; ++i) {
        m->topo_off[i] = off_delta(m, o + i * dim);
        m->full_off[i] = m->topo_off[i];
    }
    oc_sp_table2d(m->table2d);
    oc_rng_seed(seed, m->rng);
    m->check_conn = 1;
    return m;
}

void oc_model_free(ocpm* m) {
    if (!m) return;
    free(m->cells);
    free(m->type_index);
    free(m->volume);
    free(m->interface);
    free(m->contact);
    free(m->sum_pos);
    free(m->sum_outer);
    free(m->lambda_v);
    free(m->lambda_i);
    free(m->target_v);
    free(m->target_i);
    free(m->J);
    free(m->lambda_act);
    free(m->max_act);
    free(m->birth);
    free(m->edge_sites);
    free(m->edge_pos);
    free(m);
}

void oc_set_modes(ocpm* m, int acceptance, int act, const double* lambda_act,
                  const uint32_t* max_act) {
    m->acceptance = acceptance;
    m->act = act;
    free(m->lambda_act);
    free(m->max_act);
    free(m->birth);
    m->lambda_act = NULL;
    m->max_act = NULL;
    m->birth = NULL;
    if (act) {
        m->lambda_act = (double*)malloc(sizeof(double) * m->n_types);
        m->max_act = (uint32_t*)malloc(sizeof(uint32_t) * m->n_types);
        memcpy(m->lambda_act, lambda_act, sizeof(double) * m->n_types);
        memcpy(m->max_act, max_act, sizeof(uint32_t) * m->n_types);
        m->birth = (uint32_t*)calloc(m->n_padded, sizeof(uint32_t));
    }
}

void oc_set_connectivity_check(ocpm* m, int on) { m->check_conn = on; }

void oc_set_proposal(ocpm* m, int edge_list) {
    m->edge_list = edge_list;
    free(m->edge_sites);
    free(m->edge_pos);
    m->edge_sites = NULL;
    m->edge_pos = NULL;
    m->n_edge = 0;
    if (edge_list) {
        m->edge_sites = (size_t*)malloc(sizeof(size_t) * m->n_sites);
        m->edge_pos = (uint32_t*)malloc(sizeof(uint32_t) * m->n_padded);
        for (size_t i = 0; i < m->n_padded; ++i) m->edge_pos[i] = UINT32_MAX;
    }
}

/* ---- edge list (edge_list.rs:77-186) ---- */
static void edge_insert(ocpm* m, size_t site) {
    if (m->edge_pos[site] != UINT32_MAX) return;
    m->edge_pos[site] = (uint32_t)m->n_edge;
    m->edge_sites[m->n_edge++] = site;
}
static void edge_remove(ocpm* m, size_t site) {
    const uint32_t idx = m->edge_pos[site];
    if (idx == UINT32_MAX) return;
    const size_t last = m->n_edge - 1;
    if (idx != last) {
        const size_t moved = m->edge_sites[last];
        m->edge_sites[idx] = moved;
        m->edge_pos[moved] = idx;
    }
    m->n_edge--;
    m->edge_pos[site] = UINT32_MAX;
}
static int edge_site(const ocpm* m, size_t site) {
    const uint32_t owner = m->cells[site];
    for (int k = 0; k < m->n_copy; ++k)
        if (m->cells[(size_t)((ptrdiff_t)site + m->copy_off[k])] != owner) return 1;
    return 0;
}
/* from-scratch build, interior sites in row-major order */
static void edge_build(ocpm* m) {
    for (size_t i = 0; i < m->n_padded; ++i) m->edge_pos[i] = UINT32_MAX;
    m->n_edge = 0;
    for (size_t f = 0; f < m->n_padded; ++f)
        if (is_interior(m, f) && edge_site(m, f)) edge_insert(m, f);
}
/* after an accepted copy onto t: t, then each copy neighbour (a periodic
   halo site canonicalised, a fixed one skipped) */
static void edge_update_around(ocpm* m, size_t t) {
    for (int k = -1; k < m->n_copy; ++k) {
        size_t site = k < 0 ? t : (size_t)((ptrdiff_t)t + m->copy_off[k]);
        if (!is_interior(m, site)) {
            if (m->boundary != OC_PERIODIC) continue;
            site = wrapped_flat(m, site);
        }
        if (edge_site(m, site)) edge_insert(m, site);
        else edge_remove(m, site);
    }
}

/* ---- rand 0.8 UniformFloat<f64> (uniform.rs): low + scale * [0, 1) with
   the [0, 1) value from 52 random bits as (1.x) - 1 ---- */
static double uniform_f64_scale(double low, double high) {
    const double max_rand = 1.0 - 0x1p-52;
    double scale = high - low;
    while (scale * max_rand + low >= high) scale = nextafter(scale, 0.0);
    return scale;
}
static double uniform_f64_sample(uint64_t* rng, double low, double scale) {
    const uint64_t bits = (oc_rng_next(rng) >> 12) | 0x3FF0000000000000ull;
    double v12;
    memcpy(&v12, &bits, 8);
    return (v12 - 1.0) * scale + low;
}

/* compiler-rt __powidf2: x^n by binary powering (Rust's f64::powi) */
static double powi_f64(double a, int32_t b) {
    const int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* Rust's `x as i64`: truncation toward zero, saturating, NaN -> 0 (a
   sample of v = 0 gives log(v) = -inf in the tails) */
static inline int64_t f64_to_i64(double x) {
    if (x != x) return 0;
    if (x <= -0x1p63) return INT64_MIN;
    if (x >= 0x1p63) return INT64_MAX;
    return (int64_t)x;
}

static double btpe_stirling(double a) {
    const double a2 = a * a;
    return (13860.0 - (462.0 - (132.0 - (99.0 - 140.0 / a2) / a2) / a2) / a2) / a / 166320.0;
}

/* rand_distr 0.4.3 Binomial::sample (binomial.rs): BINV inversion when
   n * min(p, 1 - p) < 10, else BTPE (Kachitvichyanukul & Schmeiser 1988,
   with GSL's signs in the final test); results for p > 1/2 mirrored.
   Returns 0 without a draw for p == 0 and n for p == 1. */
uint64_t oc_binomial(uint64_t rng[4], uint64_t n_u, double p_in) {
    if (p_in == 0.0) return 0;
    if (p_in == 1.0) return n_u;
    const double p = p_in <= 0.5 ? p_in : 1.0 - p_in;
    const double q = 1.0 - p;
    uint64_t result;
    if ((double)n_u * p < 10.0 && n_u <= (uint64_t)INT32_MAX) {
        const double s = p / q;
        const double a = (double)(n_u + 1) * s;
        double r = powi_f64(q, (int32_t)n_u);
        double u = oc_rng_gen_f64(rng);
        uint64_t x = 0;
        while (u > r) {
            u -= r;
            x += 1;
            r *= a / (double)x - s;
        }
        result = x;
    } else {
        const double n = (double)n_u;
        const double np = n * p;
        const double npq = np * q;
        const double f_m = np + p;
        const int64_t m = f64_to_i64(f_m);
        const double p1 = floor(2.195 * sqrt(npq) - 4.6 * q) + 0.5;
        const double x_m = (double)m + 0.5;
        const double x_l = x_m - p1;
        const double x_r = x_m + p1;
        const double c = 0.134 + 20.5 / (15.3 + (double)m);
        const double p2 = p1 * (1.0 + 2.0 * c);
        const double al = (f_m - x_l) / (f_m - x_l * p);
        const double lambda_l = al * (1.0 + 0.5 * al);
        const double ar = (x_r - f_m) / (x_r * q);
        const double lambda_r = ar * (1.0 + 0.5 * ar);
        const double p3 = p2 + c / lambda_l;
        const double p4 = p3 + c / lambda_r;
        const double scale_u = uniform_f64_scale(0.0, p4);
        const double scale_v = uniform_f64_scale(0.0, 1.0);
        int64_t y;
        for (;;) {
            const double u = uniform_f64_sample(rng, 0.0, scale_u);
            double v = uniform_f64_sample(rng, 0.0, scale_v);
            if (!(u > p1)) {
                y = f64_to_i64(x_m - p1 * v + u);
                break;
            }
            if (!(u > p2)) {
                const double x = x_l + (u - p1) / c;
                v = v * c + 1.0 - fabs(x - x_m) / p1;
                if (v > 1.0) continue;
                y = f64_to_i64(x);
            } else if (!(u > p3)) {
                y = f64_to_i64(x_l + log(v) / lambda_l);
                if (y < 0) continue;
                v *= (u - p2) * lambda_l;
            } else {
                y = f64_to_i64(x_r - log(v) / lambda_r);
                if (y > 0 && (uint64_t)y > n_u) continue;
                v *= (u - p3) * lambda_r;
            }
            const int64_t k = y > m ? y - m : m - y;
            if (!(k > 20 && (double)k < 0.5 * npq - 1.0)) {
                const double s = p / q;
                const double a = s * (n + 1.0);
                double f = 1.0;
                if (m < y) {
                    for (int64_t i = m + 1; i <= y; ++i) f *= a / (double)i - s;
                } else if (m > y) {
                    for (int64_t i = y + 1; i <= m; ++i) f /= a / (double)i - s;
                }
                if (v > f) continue;
                break;
            }
            const double kf = (double)k;
            const double rho = (kf / npq) * ((kf * (kf / 3.0 + 0.625) + 1.0 / 6.0) / npq + 0.5);
            const double t = -0.5 * kf * kf / npq;
            const double alpha = log(v);
            if (alpha < t - rho) break;
            if (alpha > t + rho) continue;
            const double x1 = (double)(y + 1);
            const double f1 = (double)(m + 1);
            const double z = (double)(f64_to_i64(n) + 1 - m);
            const double w = (double)(f64_to_i64(n) - y + 1);
            if (alpha > x_m * log(f1 / x1) + (n - (double)m + 0.5) * log(z / w) +
                            (double)(y - m) * log(w * p / (x1 * q)) + btpe_stirling(f1) +
                            btpe_stirling(z) - btpe_stirling(x1) - btpe_stirling(w))
                continue;
            break;
        }
        result = (uint64_t)y;
    }
    return p != p_in ? n_u - result : result;
}

void oc_set_halo_mode(ocpm* m, int mode) { m->halo_mode = mode; }

static inline size_t ckey(const ocpm* m, uint32_t a, uint32_t b) {
    uint32_t lo = a <= b ? a : b, hi = a <= b ? b : a;
    return (size_t)lo * (m->n_cells + 1) + hi;
}

static inline int adh_idx(const ocpm* m, uint32_t id) {
    return id == 0 ? 0 : (int)m->type_index[id - 1] + 1;
}

/* energy.rs:87-90 / 150-153: lambda * (x - target)^2 with powi(2) == x*x */
static inline double vterm(const ocpm* m, int t, double v) {
    double d = v - (double)m->target_v[t];
    return m->lambda_v[t] * (d * d);
}
static inline double iterm(const ocpm* m, int t, double v) {
    double d = v - (double)m->target_i[t];
    return m->lambda_i[t] * (d * d);
}

/* energy.rs:167-186 */
static void interface_deltas(const ocpm* m, size_t t, uint32_t L, uint32_t G,
                             double* dl, double* dg) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < m->n_energy; ++k) {
        double w = m->energy_w[k];
        uint32_t n = m->cells[t + m->energy_off[k]];
        a += n == L ? w : -w;
        b += n == G ? -w : w;
    }
    *dl = a;
    *dg = b;
}

/* energy.rs:521-523 (volume 92-103, interface 190-202, adhesion 242-264) */
static double total_delta(const ocpm* m, size_t t, uint32_t L, uint32_t G) {
    double vol = 0.0;
    if (L) {
        int ty = (int)m->type_index[L - 1];
        double v = (double)m->volume[L - 1];
        vol += vterm(m, ty, v - 1.0) - vterm(m, ty, v);
    }
    if (G) {
        int ty = (int)m->type_index[G - 1];
        double v = (double)m->volume[G - 1];
        vol += vterm(m, ty, v + 1.0) - vterm(m, ty, v);
    }
    double dl, dg;
    interface_deltas(m, t, L, G, &dl, &dg);
    double itf = 0.0;
    if (L) {
        int ty = (int)m->type_index[L - 1];
        double i = (double)m->interface[L - 1];
        itf += iterm(m, ty, i + dl) - iterm(m, ty, i);
    }
    if (G) {
        int ty = (int)m->type_index[G - 1];
        double i = (double)m->interface[G - 1];
        itf += iterm(m, ty, i + dg) - iterm(m, ty, i);
    }
    const int T1 = m->n_types + 1;
    const int li = adh_idx(m, L), gi = adh_idx(m, G);
    double adh = 0.0;
    for (int k = 0; k < m->n_energy; ++k) {
        double w = m->energy_w[k];
        uint32_t n = m->cells[t + m->energy_off[k]];
        int ni = adh_idx(m, n);
        double before = n != L ? m->J[li * T1 + ni] : 0.0;
        double after = n != G ? m->J[gi * T1 + ni] : 0.0;
        adh += w * (after - before);
    }
    return vol + itf + adh;
}

/* dynamics.rs:106-110 */
static inline uint32_t apply_delta(uint32_t cur, double delta) {
    return (uint32_t)((int64_t)cur + (int64_t)round(delta));
}

/* dynamics.rs:187-202 */
static int64_t nearest_image(int64_t coord, double centroid, int64_t extent,
                             int boundary) {
    if (boundary == OC_FIXED) return coord;
    int64_t best = coord;
    double best_d = fabs((double)coord - centroid);
    const int64_t cands[2] = {coord - extent, coord + extent};
    for (int i = 0; i < 2; ++i) {
        double dd = fabs((double)cands[i] - centroid);
        if (dd < best_d) {
            best = cands[i];
            best_d = dd;
        }
    }
    return best;
}

/* dynamics.rs:23-58, 73-98, 125-183, 206-224 */
static void commit(ocpm* m, size_t t, uint32_t L, uint32_t G, double dl, double dg) {
    const int D = m->D;
    for (int k = 0; k < m->n_energy; ++k) {
        uint32_t n = m->cells[t + m->energy_off[k]];
        if (n != L) {
            size_t key = ckey(m, L, n);
            if (m->contact[key]) m->contact[key] -= 1;
        }
        if (n != G) m->contact[ckey(m, G, n)] += 1;
    }
    m->cells[t] = G;
    if (L) {
        m->volume[L - 1] -= 1;
        m->interface[L - 1] = apply_delta(m->interface[L - 1], dl);
    }
    if (G) {
        m->volume[G - 1] += 1;
        m->interface[G - 1] = apply_delta(m->interface[G - 1], dg);
    }
    size_t c[3];
    interior_coord(m, t, c);
    if (L) {
        int64_t* s = m->sum_pos + (size_t)(L - 1) * D;
        int64_t* o = m->sum_outer + (size_t)(L - 1) * D * D;
        for (int d = 0; d < D; ++d) s[d] -= (int64_t)c[d];
        for (int d = 0; d < D; ++d)
            for (int e = 0; e < D; ++e) o[d * D + e] -= (int64_t)c[d] * (int64_t)c[e];
    }
    if (G) {
        int64_t* s = m->sum_pos + (size_t)(G - 1) * D;
        int64_t* o = m->sum_outer + (size_t)(G - 1) * D * D;
        uint32_t v = m->volume[G - 1];
        uint32_t vbg = v > 0 ? v - 1 : 0;
        if (vbg < 1) vbg = 1;
        double cen[3];
        for (int d = 0; d < D; ++d) cen[d] = (double)s[d] / (double)vbg;
        int64_t img[3];
        for (int d = 0; d < D; ++d) {
            img[d] = nearest_image((int64_t)c[d], cen[d], (int64_t)m->dims[d], m->boundary);
            s[d] += img[d];
        }
        for (int d = 0; d < D; ++d)
            for (int e = 0; e < D; ++e) o[d * D + e] += img[d] * img[e];
    }
    if (m->boundary == OC_PERIODIC) {
        int on_boundary = 0;
        for (int d = 0; d < D; ++d)
            if (c[d] == 0 || c[d] == m->dims[d] - 1) on_boundary = 1;
        if (on_boundary) {
            if (m->halo_mode == OC_HALO_FULL_RESCAN) refresh_halo_full(m);
            else refresh_halo_mirrors(m, c);
        }
    }
}

static int topo_ok(ocpm* m, size_t t, uint32_t L) {
    if (L == 0) return 1;
    uint32_t pattern = 0;
    for (int i = 0; i < m->n_topo; ++i)
        if (m->cells[t + m->topo_off[i]] == L) pattern |= 1u << i;
    m->counters[2]++;
    if (m->D == 2) return m->table2d[pattern];
    const uint64_t* t3 = table3d_cached();
    return (int)((t3[pattern >> 6] >> (pattern & 63)) & 1);
}

/* energy.rs:364-376: a flat index canonicalised to the interior site it
   stands for (periodic only; a fixed halo stays where it is) */
static inline size_t act_canonical(const ocpm* m, size_t flat) {
    return m->boundary == OC_PERIODIC ? wrapped_flat(m, flat) : flat;
}

/* energy.rs:403-444: geometric mean of activity over {site} and its copy
   neighbours, restricted to sites owned by rel, with the gaining type's
   max_act for every retained site */
static double act_geometric_mean(const ocpm* m, size_t site, uint32_t rel, uint32_t max_act,
                                 uint64_t mcs) {
    const size_t centre = act_canonical(m, site);
    double product = 1.0;
    uint32_t count = 0;
    for (int k = -1; k < m->n_copy; ++k) {
        const size_t f = act_canonical(
            m, k < 0 ? centre : (size_t)((ptrdiff_t)centre + m->copy_off[k]));
        if (m->cells[f] != rel) continue;
        const uint64_t birth = m->birth[f];
        const uint64_t elapsed = mcs > birth ? mcs - birth : 0;
        const uint64_t activity = (uint64_t)max_act > elapsed ? (uint64_t)max_act - elapsed : 0;
        product *= (double)activity;
        count += 1;
    }
    return pow(product, 1.0 / (double)count);
}

/* energy.rs:446-457: delta_H_Act = -(lambda_Act / Max_Act) (GM_src - GM_tgt),
   0 when medium gains */
static double act_delta(const ocpm* m, size_t t, size_t s, uint32_t L, uint32_t G,
                        uint64_t mcs) {
    if (G == 0) return 0.0;
    const uint32_t ty = m->type_index[G - 1];
    const double lambda = m->lambda_act[ty];
    const uint32_t max_act = m->max_act[ty];
    const double gm_src = act_geometric_mean(m, s, G, max_act, mcs);
    const double gm_tgt = act_geometric_mean(m, t, L, max_act, mcs);
    return -(lambda / (double)max_act) * (gm_src - gm_tgt);
}

/* monte_carlo.rs:200-222: copy neighbours of t owned by L (n_A) and by G (n_B) */
static void mh_counts(const ocpm* m, size_t t, uint32_t L, uint32_t G, uint32_t* na,
                      uint32_t* nb) {
    uint32_t a = 0, b = 0;
    for (int k = 0; k < m->n_copy; ++k) {
        const uint32_t c = m->cells[(size_t)((ptrdiff_t)t + m->copy_off[k])];
        a += c == L;
        b += c == G;
    }
    *na = a;
    *nb = b;
}

/* monte_carlo.rs:52-145, 147-158: one attempt at target t */
static inline void attempt_at(ocpm* m, uint64_t* restrict rng, uint64_t* restrict cnt, size_t t) {
    size_t k = (size_t)oc_rng_gen_range(rng, (uint64_t)m->n_copy);
    size_t s = (size_t)((ptrdiff_t)t + m->copy_off[k]);
    cnt[0]++;
    uint32_t L = m->cells[t], G = m->cells[s];
    if (L == G) return;
    cnt[1]++;
    if (m->check_conn && !topo_ok(m, t, L)) return;
    if (L != 0 && m->volume[L - 1] <= 1) return;
    cnt[3]++;
    /* monte_carlo.rs:80-97: Act enters the acceptance test only */
    const double dh = total_delta(m, t, L, G);
    const double dh_acc = dh + (m->act ? act_delta(m, t, s, L, G, m->mcs) : 0.0);
    int accept;
    if (m->acceptance == OC_METROPOLIS_HASTINGS) {
        /* monte_carlo.rs:168-196: no downhill shortcut; n_A == 0 rejects
           without a draw */
        uint32_t na, nb;
        mh_counts(m, t, L, G, &na, &nb);
        if (na == 0) return;
        const double ratio = ((double)na / (double)nb) * exp(-dh_acc);
        const double u = oc_rng_gen_f64(rng);
        accept = u < fmin(ratio, 1.0);
    } else {
        accept = dh_acc <= 0.0;
        if (!accept) {
            const double u = oc_rng_gen_f64(rng);
            accept = u < exp(-dh_acc);
            if (m->mlog) {
                if (m->mlog_n < m->mlog_cap) {
                    m->mlog[2 * m->mlog_n] = dh_acc;
                    m->mlog[2 * m->mlog_n + 1] = u;
                }
                m->mlog_n++;
            }
        }
    }
    if (!accept) return;
    cnt[4]++;
    /* energy.rs:459-464 then dynamics.rs; only the conservative part
       reaches the energy (monte_carlo.rs:265-270) */
    if (m->act) m->birth[t] = (uint32_t)m->mcs;
    double dl, dg;
    interface_deltas(m, t, L, G, &dl, &dg);
    commit(m, t, L, G, dl, dg);
    if (m->edge_list) edge_update_around(m, t); /* monte_carlo.rs:257-264 */
    m->energy += dh;
}

/* monte_carlo.rs:27-42: a uniformly drawn interior target */
static inline void attempt(ocpm* m, uint64_t* restrict rng, uint64_t* restrict cnt) {
    size_t c[3];
    for (int d = 0; d < m->D; ++d) c[d] = (size_t)oc_rng_gen_range(rng, m->dims[d]);
    attempt_at(m, rng, cnt, flat_of(m, c));
}

void oc_run_mcs(ocpm* m, uint64_t n) {
    /* the stream state and the A/N/E/C counters live in locals for the
       loop (the topology counter stays in the model: topo_ok updates it) */
    uint64_t rng[4], cnt[5];
    memcpy(rng, m->rng, sizeof rng);
    memcpy(cnt, m->counters, sizeof cnt);
    for (uint64_t i = 0; i < n; ++i) {
        if (!m->edge_list) {
            for (size_t a = 0; a < m->n_sites; ++a) attempt(m, rng, cnt);
        } else {
            /* monte_carlo.rs:336-351: k ~ Binomial(N_sites, |E| / N_sites)
               with |E| read once, then k targets drawn from E */
            const uint64_t k = m->n_edge == 0
                                   ? 0
                                   : oc_binomial(rng, m->n_sites,
                                                 (double)m->n_edge / (double)m->n_sites);
            for (uint64_t a = 0; a < k; ++a) {
                if (m->n_edge == 0) break;
                const size_t t = m->edge_sites[oc_rng_gen_range(rng, m->n_edge)];
                attempt_at(m, rng, cnt, t);
            }
        }
        m->mcs += 1;
    }
    memcpy(m->rng, rng, sizeof rng);
    cnt[2] = m->counters[2];
    memcpy(m->counters, cnt, sizeof cnt);
}

/* Visit interior coordinates row-major, last axis fastest. */
#define FOR_INTERIOR(m, c, ...)                                            \
    do {                                                                   \
        size_t c[3] = {0, 0, 0};                                           \
        for (size_t _i = 0; _i < (m)->n_sites; ++_i) {                     \
            __VA_ARGS__;                                                   \
            for (int _d = (m)->D - 1; _d >= 0; --_d) {                     \
                if (++c[_d] < (m)->dims[_d]) break;                        \
                c[_d] = 0;                                                 \
            }                                                              \
        }                                                                  \
    } while (0)

/* energy.rs:106-114, 205-213, 269-308, 561-565 */
double oc_global_energy(const ocpm* m) {
    double vol = 0.0, itf = 0.0, adh = 0.0;
    for (size_t i = 0; i < m->n_cells; ++i) vol += vterm(m, (int)m->type_index[i], (double)m->volume[i]);
    for (size_t i = 0; i < m->n_cells; ++i) itf += iterm(m, (int)m->type_index[i], (double)m->interface[i]);
    const int T1 = m->n_types + 1;
    FOR_INTERIOR(m, c, {
        size_t fi = flat_of(m, c);
        uint32_t ci = m->cells[fi];
        int ii = adh_idx(m, ci);
        for (int k = 0; k < m->n_energy; ++k) {
            size_t fj = (size_t)((ptrdiff_t)fi + m->energy_off[k]);
            int has_key = 1;
            size_t key = fj;
            if (!is_interior(m, fj)) {
                if (m->boundary == OC_PERIODIC) key = wrapped_flat(m, fj);
                else has_key = 0;
            }
            if (has_key && key <= fi) continue;
            uint32_t cj = m->cells[fj];
            if (cj == ci) continue;
            adh += m->energy_w[k] * m->J[ii * T1 + adh_idx(m, cj)];
        }
    });
    return vol + itf + adh;
}

void oc_recompute(const ocpm* m, uint32_t* interface, int64_t* sum_pos,
                  int64_t* sum_outer, uint32_t* contact) {
    const int D = m->D;
    const size_t n = m->n_cells;
    memset(interface, 0, sizeof(uint32_t) * n);
    memset(sum_pos, 0, sizeof(int64_t) * n * D);
    memset(sum_outer, 0, sizeof(int64_t) * n * D * D);
    memset(contact, 0, sizeof(uint32_t) * (n + 1) * (n + 1));
    FOR_INTERIOR(m, c, {
        size_t fi = flat_of(m, c);
        uint32_t ci = m->cells[fi];
        if (ci) {
            size_t i = ci - 1;
            for (int d = 0; d < D; ++d) {
                sum_pos[i * D + d] += (int64_t)c[d];
                for (int e = 0; e < D; ++e)
                    sum_outer[i * D * D + d * D + e] += (int64_t)c[d] * (int64_t)c[e];
            }
            for (int k = 0; k < m->n_energy; ++k)
                if (m->cells[fi + m->energy_off[k]] != ci) interface[i] += 1;
        }
        for (int k = 0; k < m->n_energy; ++k) {
            size_t fj = (size_t)((ptrdiff_t)fi + m->energy_off[k]);
            int has_key = 1;
            size_t key = fj;
            if (!is_interior(m, fj)) {
                if (m->boundary == OC_PERIODIC) key = wrapped_flat(m, fj);
                else has_key = 0;
            }
            if (has_key && key <= fi) continue;
            uint32_t cj = m->cells[fj];
            if (cj == ci) continue;
            contact[ckey(m, ci, cj)] += 1;
        }
    });
}

static void alloc_cells(ocpm* m, size_t n) {
    m->n_cells = n;
    m->type_index = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    m->volume = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    m->interface = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    m->sum_pos = (int64_t*)calloc((n ? n : 1) * m->D, sizeof(int64_t));
    m->sum_outer = (int64_t*)calloc((n ? n : 1) * m->D * m->D, sizeof(int64_t));
    m->contact = (uint32_t*)calloc((n + 1) * (n + 1), sizeof(uint32_t));
}

static void finish_init(ocpm* m) {
    oc_recompute(m, m->interface, m->sum_pos, m->sum_outer, m->contact);
    m->energy = oc_global_energy(m);
    /* initialization.rs:427-435 */
    if (m->edge_list) edge_build(m);
}

/* initialization.rs:145-178: BFS connectivity under von Neumann */
static int cell_connected(const ocpm* m, uint32_t id) {
    size_t total = 0, first = (size_t)-1;
    FOR_INTERIOR(m, c, {
        size_t f = flat_of(m, c);
        if (m->cells[f] == id) {
            if (first == (size_t)-1) first = f;
            ++total;
        }
    });
    if (!total) return 0;
    uint8_t* seen = (uint8_t*)calloc(m->n_padded, 1);
    size_t* stack = (size_t*)malloc(sizeof(size_t) * (total + 1));
    size_t sp = 0, nvis = 1;
    stack[sp++] = first;
    seen[first] = 1;
    while (sp) {
        size_t f = stack[--sp];
        for (int k = 0; k < m->n_conn; ++k) {
            size_t nb = (size_t)((ptrdiff_t)f + m->conn_off[k]);
            if (!is_interior(m, nb) && m->boundary == OC_PERIODIC) nb = wrapped_flat(m, nb);
            if (m->cells[nb] == id && !seen[nb]) {
                seen[nb] = 1;
                stack[sp++] = nb;
                ++nvis;
            }
        }
    }
    free(seen);
    free(stack);
    return nvis == total;
}

int oc_init_explicit(ocpm* m, const uint32_t* lattice, uint64_t n_cells,
                     const uint64_t* cell_type_of) {
    for (uint64_t i = 0; i < n_cells; ++i)
        if (cell_type_of[i] >= (uint64_t)m->n_types) return 2;
    for (size_t i = 0; i < m->n_sites; ++i)
        if (lattice[i] > n_cells) return 3;
    memset(m->cells, 0, sizeof(uint32_t) * m->n_padded);
    alloc_cells(m, (size_t)n_cells);
    size_t idx = 0;
    FOR_INTERIOR(m, c, {
        uint32_t v = lattice[idx++];
        m->cells[flat_of(m, c)] = v;
        if (v) m->volume[v - 1] += 1;
    });
    if (m->boundary == OC_PERIODIC) refresh_halo_full(m);
    for (size_t i = 0; i < n_cells; ++i)
        if (!m->volume[i]) return 4;
    for (size_t i = 0; i < n_cells; ++i) m->type_index[i] = (uint32_t)cell_type_of[i];
    for (uint32_t id = 1; id <= n_cells; ++id)
        if (!cell_connected(m, id)) return 5;
    finish_init(m);
    return 0;
}

/* Places an arbitrary lattice (no connectivity or volume checks) and
   recomputes every derived quantity: the brute-force checker's view of a
   state produced elsewhere (checker.rs:36-136). */
void oc_load_state(ocpm* m, const uint32_t* lattice, uint64_t n_cells,
                   const uint64_t* cell_type_of) {
    memset(m->cells, 0, sizeof(uint32_t) * m->n_padded);
    free(m->type_index);
    free(m->volume);
    free(m->interface);
    free(m->sum_pos);
    free(m->sum_outer);
    free(m->contact);
    alloc_cells(m, (size_t)n_cells);
    size_t idx = 0;
    FOR_INTERIOR(m, c, {
        uint32_t v = lattice[idx++];
        m->cells[flat_of(m, c)] = v;
        if (v) m->volume[v - 1] += 1;
    });
    if (m->boundary == OC_PERIODIC) refresh_halo_full(m);
    for (size_t i = 0; i < n_cells; ++i) m->type_index[i] = (uint32_t)cell_type_of[i];
    finish_init(m);
}

/* initialization.rs:308-327 (canonical_growable_neighbour) */
static int growable(const ocpm* m, size_t flat, ptrdiff_t off, size_t* out) {
    size_t raw = (size_t)((ptrdiff_t)flat + off);
    if (is_interior(m, raw)) {
        *out = raw;
        return 1;
    }
    if (m->boundary == OC_PERIODIC) {
        *out = wrapped_flat(m, raw);
        return 1;
    }
    return 0;
}

/* initialization.rs:202-330 */
int oc_init_scatter(ocpm* m, const uint64_t* cell_counts) {
    size_t n = 0;
    for (int t = 0; t < m->n_types; ++t) n += (size_t)cell_counts[t];
    if (!n) return 6;
    uint32_t* type_of = (uint32_t*)malloc(sizeof(uint32_t) * n);
    {
        size_t k = 0;
        for (int t = 0; t < m->n_types; ++t)
            for (uint64_t j = 0; j < cell_counts[t]; ++j) type_of[k++] = (uint32_t)t;
    }
    double avg = 0.0;
    for (int t = 0; t < m->n_types; ++t) avg += (double)m->target_v[t];
    avg /= (double)m->n_types;
    const double min_sep = sqrt(avg / 3.141592653589793);
    size_t* seeds = (size_t*)malloc(sizeof(size_t) * 3 * n);
    for (size_t i = 0; i < n; ++i) {
        int placed = 0;
        for (int a = 0; a < 20000 && !placed; ++a) {
            size_t cand[3] = {0, 0, 0};
            for (int d = 0; d < m->D; ++d) cand[d] = (size_t)oc_rng_gen_range(m->rng, m->dims[d]);
            int far = 1;
            for (size_t j = 0; j < i && far; ++j) {
                double sq = 0.0;
                for (int d = 0; d < m->D; ++d) {
                    double raw = fabs((double)seeds[3 * j + d] - (double)cand[d]);
                    double delta = raw;
                    if (m->boundary == OC_PERIODIC) {
                        double alt = (double)m->dims[d] - raw;
                        delta = fmin(raw, alt);
                    }
                    sq += delta * delta;
                }
                if (!(sqrt(sq) >= min_sep)) far = 0;
            }
            if (far) {
                for (int d = 0; d < m->D; ++d) seeds[3 * i + d] = cand[d];
                placed = 1;
            }
        }
        if (!placed) {
            free(type_of);
            free(seeds);
            return 7;
        }
    }
    memset(m->cells, 0, sizeof(uint32_t) * m->n_padded);
    alloc_cells(m, n);
    for (size_t i = 0; i < n; ++i) {
        m->cells[flat_of(m, seeds + 3 * i)] = (uint32_t)(i + 1);
        m->volume[i] += 1;
        m->type_index[i] = type_of[i];
    }
    /* frontiers: growable dynamic arrays */
    size_t** fr = (size_t**)malloc(sizeof(size_t*) * n);
    size_t* flen = (size_t*)calloc(n, sizeof(size_t));
    size_t* fcap = (size_t*)malloc(sizeof(size_t) * n);
    for (size_t i = 0; i < n; ++i) {
        fcap[i] = 64;
        fr[i] = (size_t*)malloc(sizeof(size_t) * fcap[i]);
        size_t sf = flat_of(m, seeds + 3 * i);
        for (int k = 0; k < m->n_copy; ++k) {
            size_t nb;
            if (growable(m, sf, m->copy_off[k], &nb) && m->cells[nb] == 0) fr[i][flen[i]++] = nb;
        }
    }
    size_t* order = (size_t*)malloc(sizeof(size_t) * n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    for (;;) {
        int grew = 0;
        for (size_t oi = 0; oi < n; ++oi) {
            size_t i = order[oi];
            if (m->volume[i] >= m->target_v[type_of[i]]) continue;
            size_t w = 0;
            for (size_t r = 0; r < flen[i]; ++r)
                if (m->cells[fr[i][r]] == 0) fr[i][w++] = fr[i][r];
            flen[i] = w;
            if (!w) continue;
            size_t pick = (size_t)oc_rng_gen_range(m->rng, flen[i]);
            size_t claimed = fr[i][pick];
            fr[i][pick] = fr[i][flen[i] - 1];
            flen[i] -= 1;
            m->cells[claimed] = (uint32_t)(i + 1);
            m->volume[i] += 1;
            for (int k = 0; k < m->n_copy; ++k) {
                size_t nb;
                if (growable(m, claimed, m->copy_off[k], &nb) && m->cells[nb] == 0) {
                    if (flen[i] == fcap[i]) {
                        fcap[i] *= 2;
                        fr[i] = (size_t*)realloc(fr[i], sizeof(size_t) * fcap[i]);
                    }
                    fr[i][flen[i]++] = nb;
                }
            }
            grew = 1;
        }
        if (!grew) break;
        for (size_t i = n - 1; i >= 1; --i) {
            size_t j = (size_t)oc_rng_gen_range_incl(m->rng, i);
            size_t tmp = order[i];
            order[i] = order[j];
            order[j] = tmp;
        }
    }
    if (m->boundary == OC_PERIODIC) refresh_halo_full(m);
    for (size_t i = 0; i < n; ++i) free(fr[i]);
    free(fr);
    free(flen);
    free(fcap);
    free(order);
    free(seeds);
    free(type_of);
    finish_init(m);
    return 0;
}

/* labelling.rs:70-154 */
uint64_t oc_label(const ocpm* m, uint32_t* fragmented, uint64_t max_frag,
                  uint32_t* medium_volume, uint8_t* medium_touch,
                  uint64_t max_components, uint64_t* n_medium) {
    uint8_t* visited = (uint8_t*)calloc(m->n_padded, 1);
    size_t* queue = (size_t*)malloc(sizeof(size_t) * (m->n_sites + 1));
    uint32_t* comps = (uint32_t*)calloc(m->n_cells + 1, sizeof(uint32_t));
    uint64_t nm = 0;
    FOR_INTERIOR(m, c, {
        size_t seed = flat_of(m, c);
        if (!visited[seed]) {
            uint32_t owner = m->cells[seed];
            const ptrdiff_t* offs = owner == 0 ? m->full_off : m->conn_off;
            int noff = owner == 0 ? m->n_topo : m->n_conn;
            size_t head = 0, tail = 0;
            visited[seed] = 1;
            queue[tail++] = seed;
            uint32_t volume = 0;
            int touches = 0;
            while (head < tail) {
                size_t site = queue[head++];
                volume += 1;
                for (int k = 0; k < noff; ++k) {
                    size_t nb = (size_t)((ptrdiff_t)site + offs[k]);
                    size_t can = nb;
                    if (!is_interior(m, nb)) {
                        if (owner == 0 && m->boundary == OC_FIXED) touches = 1;
                        if (m->boundary != OC_PERIODIC) continue;
                        can = wrapped_flat(m, nb);
                    }
                    if (visited[can]) continue;
                    if (m->cells[can] != owner) continue;
                    visited[can] = 1;
                    queue[tail++] = can;
                }
            }
            if (owner == 0) {
                if (nm < max_components) {
                    medium_volume[nm] = volume;
                    medium_touch[nm] = (uint8_t)touches;
                }
                ++nm;
            } else {
                comps[owner] += 1;
            }
        }
    });
    uint64_t nf = 0;
    for (size_t id = 1; id <= m->n_cells; ++id)
        if (comps[id] > 1) {
            if (nf < max_frag) fragmented[nf] = (uint32_t)id;
            ++nf;
        }
    *n_medium = nm;
    free(visited);
    free(queue);
    free(comps);
    return nf;
}

/* ---- accessors ---- */
uint64_t oc_n_sites(const ocpm* m) { return m->n_sites; }
uint64_t oc_n_cells(const ocpm* m) { return m->n_cells; }
uint64_t oc_mcs(const ocpm* m) { return m->mcs; }
double oc_energy(const ocpm* m) { return m->energy; }
void oc_rng_state(const ocpm* m, uint64_t out[4]) { memcpy(out, m->rng, 32); }
void oc_lattice(const ocpm* m, uint32_t* out) {
    size_t i = 0;
    FOR_INTERIOR(m, c, { out[i++] = m->cells[flat_of(m, c)]; });
}
void oc_padded_lattice(const ocpm* m, uint32_t* out) {
    memcpy(out, m->cells, sizeof(uint32_t) * m->n_padded);
}
void oc_volume(const ocpm* m, uint32_t* out) { memcpy(out, m->volume, sizeof(uint32_t) * m->n_cells); }
void oc_interface(const ocpm* m, uint32_t* out) { memcpy(out, m->interface, sizeof(uint32_t) * m->n_cells); }
void oc_type_index(const ocpm* m, uint32_t* out) { memcpy(out, m->type_index, sizeof(uint32_t) * m->n_cells); }
void oc_sum_position(const ocpm* m, int64_t* out) {
    memcpy(out, m->sum_pos, sizeof(int64_t) * m->n_cells * m->D);
}
void oc_sum_outer(const ocpm* m, int64_t* out) {
    memcpy(out, m->sum_outer, sizeof(int64_t) * m->n_cells * m->D * m->D);
}
void oc_contact(const ocpm* m, uint32_t* out) {
    memcpy(out, m->contact, sizeof(uint32_t) * (m->n_cells + 1) * (m->n_cells + 1));
}
void oc_counters(const ocpm* m, uint64_t out[5]) { memcpy(out, m->counters, sizeof(uint64_t) * 5); }

/* Records the (dH, u) pair of every uphill Metropolis test from now on into
   buf (2 doubles each, up to cap pairs); returns the number of tests seen
   since the last call (buf NULL: stop recording). */
uint64_t oc_metropolis_log(ocpm* m, double* buf, uint64_t cap) {
    const uint64_t seen = m->mlog_n;
    m->mlog = buf;
    m->mlog_cap = cap;
    m->mlog_n = 0;
    return seen;
}

/* Act state and probes (energy.rs:329-471) */
void oc_birth_mcs(const ocpm* m, uint32_t* out) {
    size_t idx = 0;
    FOR_INTERIOR(m, c, { out[idx++] = m->birth ? m->birth[flat_of(m, c)] : 0; });
}
void oc_set_birth_mcs(ocpm* m, const uint32_t* in) {
    if (!m->birth) return;
    size_t idx = 0;
    FOR_INTERIOR(m, c, { m->birth[flat_of(m, c)] = in[idx++]; });
}
/* delta_H_Act of copying the copy neighbour k of the interior site at
   coordinate tc onto it, at the given MCS (the model's Act parameters) */
double oc_act_delta(const ocpm* m, const uint64_t* tc, int k, uint64_t mcs) {
    size_t c[3];
    for (int d = 0; d < m->D; ++d) c[d] = (size_t)tc[d];
    const size_t t = flat_of(m, c);
    const size_t s = (size_t)((ptrdiff_t)t + m->copy_off[k]);
    return act_delta(m, t, s, m->cells[t], m->cells[s], mcs);
}
/* the edge list in its order, as interior row-major indices; returns |E| */
uint64_t oc_edge_sites(const ocpm* m, uint64_t* out, uint64_t cap) {
    if (!m->edge_list) return 0;
    for (size_t i = 0; i < m->n_edge && i < cap; ++i) {
        size_t c[3];
        interior_coord(m, m->edge_sites[i], c);
        uint64_t r = 0;
        for (int d = 0; d < m->D; ++d) r = r * m->dims[d] + c[d];
        out[i] = r;
    }
    return m->n_edge;
}
/* replaces the edge list by the given order (interior row-major indices) */
void oc_set_edge_sites(ocpm* m, const uint64_t* in, uint64_t n) {
    if (!m->edge_list) return;
    for (size_t i = 0; i < m->n_padded; ++i) m->edge_pos[i] = UINT32_MAX;
    m->n_edge = 0;
    for (uint64_t i = 0; i < n; ++i) {
        size_t c[3];
        uint64_t r = in[i];
        for (int d = m->D - 1; d >= 0; --d) {
            c[d] = (size_t)(r % m->dims[d]);
            r /= m->dims[d];
        }
        edge_insert(m, flat_of(m, c));
    }
}

/* pow(x[i], 1 / n[i]) with this platform's libm (the reference's
   f64::powf in energy.rs:443) */
void oc_root(const double* x, const uint32_t* n, uint64_t len, double* out) {
    for (uint64_t i = 0; i < len; ++i) out[i] = pow(x[i], 1.0 / (double)n[i]);
}

/* exp(-x[i]) with this platform's libm (the reference's f64::exp). */
void oc_exp_neg(const double* x, uint64_t n, double* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = exp(-x[i]);
}

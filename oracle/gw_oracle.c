/*
 * gw_oracle.c -- CPU restatement of GOWalker's walk step logic (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this file,
 * as the checker.  The product never links it.
 *
 * Normative walk semantics restated here (the GPU kernels must agree bit for bit):
 *
 *   Walker w of type t starts at source s with cur = prev = s, hop = 0
 *   (engine/walk.hpp:32-49 sets source = previous = current, hop = 0).
 *   While hop < length:
 *     1. termination (build addition; off when t_term == 0):
 *          r = philox(key=seed, ctr=(w, hop, 0, 1)).x ; r < t_term  => end, status 2
 *     2. dv = deg(cur); dv == 0 => end without moving, status 3
 *        (gpu/gpuwalk.hpp:112-115 ends the walk; the reference's bogus move to 0 is dropped)
 *     3. choose x in N(cur) (below); a step that reaches ORC_MAX_TRIALS trials ends the
 *        walk with status 4 (never reached for p, q in [2^-16, 2^16]).
 *     4. prev = cur, cur = x, hop++, visits[x]++, hash = (hash ^ x) * FNV_PRIME
 *   hop == length => status 1.  hash starts as (FNV_OFFSET ^ s) * FNV_PRIME.
 *
 *   Draws: trial k of a step uses philox(seed, (w, hop, k/2, 0)); even k takes
 *   (pos word, accept word) = (.x, .y), odd k takes (.z, .w).
 *   pos = (pos_word * dv) >> 32 (multiply-shift, always < dv; replaces the unguarded
 *   float scaling of gpu/gpuwalk.hpp:126,246).
 *
 *   Node2vec (gpu/gpuwalk.hpp:118-150 control flow, engine/engine.hpp:224-243 search):
 *     weights 1/p (x == prev), 1 (x in N(prev)), 1/q (otherwise); envelope
 *     M = max(1/p, 1, 1/q); a class with weight w accepts iff acc < T(w) with
 *     T(w) = floor(w/M * 2^32) (T = 2^32 when w == M).  acc < min T accepts before the
 *     candidate is classified (the reference's rv <= min_val fast path).  For x != prev,
 *     prev's list is searched only when min(T(1), T(1/q)) <= acc < max(T(1), T(1/q)):
 *     outside that band both classes give the same answer (the walk is unchanged; the
 *     probe counter counts the searches made).
 *
 *   SOPR / SOSR (gpu/gpuwalk.hpp:241-266, engine/engine.hpp:323-376):
 *     P(x) proportional to (1-a)/dv + a/du * [x in N(prev), x != prev], a = A/2^16.
 *     Rejection mode: envelope (1-a)/dv + a/du; a non-neighbour (or prev) accepts iff
 *     acc * (B*du + A*dv) < (B*du) << 32, a neighbour always accepts.  The first test
 *     is applied before classification (the rv <= min_val fast path).
 *     Mixture mode (build addition, same distribution): count c = number of positions
 *     of N(cur) whose value is in N(prev) and differs from prev; with draws
 *     d = philox(seed, (w, hop, 0, 0)): if c == 0 or d.x * (B*du + A*c) < (B*du) << 32
 *     take N(cur)[(d.y*dv)>>32], else take the k-th matching position, k = (d.z*c)>>32.
 *     Mode: mixture iff du*(4*L(dv)+2) < E*(1+L(du)) where L(n) = bit length of n and
 *     E = ceil((B*du + A*dv) / (B*du)) bounds the rejection trial count.
 *     (Sampler 1.)  Sampler 0, the default, is a proposal rejection that needs
 *     duplicate-free adjacency lists: attempt k draws d = philox(seed, (w, hop, k, 0));
 *     if d.x < B << 16 (probability 1-a) take N(cur)[(d.y*dv)>>32]; otherwise
 *     y = N(prev)[(d.y*du)>>32] is accepted iff y != prev and y is in N(cur) (textbook
 *     search), else the attempt is retried.  Every accepted x has mass
 *     (1-a)/dv + a/du [x in N(prev), x != prev], i.e. the same law, at ~1/(1-a) attempts.
 */
#include "gw_oracle.h"
#include <string.h>
#include <math.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FNV_OFFSET 1469598103934665603ULL
#define FNV_PRIME 1099511628211ULL
#define ORC_MAX_TRIALS (1u << 24)

typedef unsigned __int128 u128;

/* Philox-4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11.  Round: (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2),
 * hi(M0*c0)^c3^k1, lo(M0*c0)); the key is bumped by the Weyl constants between rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void draw(uint64_t seed, uint32_t w, uint32_t hop, uint32_t k, uint32_t tag,
                        uint32_t out[4])
{
    uint32_t ctr[4] = { w, hop, k, tag };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    orc_philox4x32_10(ctr, key, out);
}

/* RandNum, util/util.hpp:229-241 (splitmix64 seeding), :262-271 (xoroshiro128+).
 * The reference holds z in a signed long long (`LL z`), so its right shifts are
 * arithmetic; that differs from textbook splitmix64 and is reproduced here. */
static inline int64_t asr(int64_t z, int k) { return z >> k; }   /* gcc: arithmetic */

void orc_randnum_init(orc_randnum* r, uint64_t seed)
{
    for (int i = 0; i < 2; i++) {
        int64_t z = (int64_t)(seed += 0x9E3779B97F4A7C15ULL);
        z = (int64_t)((uint64_t)(z ^ asr(z, 30)) * 0xBF58476D1CE4E5B9ULL);
        z = (int64_t)((uint64_t)(z ^ asr(z, 27)) * 0x94D049BB133111EBULL);
        if (i == 0) r->s0 = (uint64_t)(z ^ asr(z, 31)); else r->s1 = (uint64_t)(z ^ asr(z, 31));
    }
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_randnum_lrand(orc_randnum* r)
{
    const uint64_t s0 = r->s0;
    uint64_t s1 = r->s1;
    const uint64_t result = s0 + s1;
    s1 ^= s0;
    r->s0 = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
    r->s1 = rotl64(s1, 36);
    return result;
}

/* util/util.hpp:274-282 */
static double randnum_drand(orc_randnum* r)
{
    uint64_t bits = (0x3FFULL << 52) | (orc_randnum_lrand(r) >> 12);
    double d;
    memcpy(&d, &bits, sizeof d);
    return d - 1.0;
}

/* util/util.hpp:291-301.  The reference's union {uint64_t i; float d;} reads the float
 * from the low 32 bits of the 64-bit pattern, which are the low 32 bits of r. */
void orc_randnum_pair(orc_randnum* r, float max_val, uint64_t bound, int literal,
                      float* rand_val, uint64_t* rand_pos)
{
    uint64_t x = orc_randnum_lrand(r);
    *rand_pos = (x >> 32) % bound;
    if (literal) {
        uint64_t bits = (0x3FFULL << 52) | (x & ((1ULL << 52) - 1));
        uint32_t lo = (uint32_t)bits;
        float f;
        memcpy(&f, &lo, sizeof f);
        *rand_val = (float)(((double)f - 1.0) * (double)max_val);
    } else {
        double u = (double)(x & ((1ULL << 52) - 1)) * (1.0 / 4503599627370496.0);
        *rand_val = (float)(u * (double)max_val);
    }
}

static uint64_t thr(double w, double m)
{
    if (w >= m) return 1ULL << 32;
    return (uint64_t)floor(w / m * 4294967296.0);
}

void orc_n2v_thresholds(double p, double q, orc_typecfg* t)
{
    double ip = 1.0 / p, iq = 1.0 / q;
    double m = ip > 1.0 ? ip : 1.0;
    if (iq > m) m = iq;
    t->t_p = thr(ip, m);
    t->t_one = thr(1.0, m);
    t->t_q = thr(iq, m);
    uint64_t mn = t->t_p < t->t_one ? t->t_p : t->t_one;
    t->t_min = mn < t->t_q ? mn : t->t_q;
}

uint32_t orc_alpha_fx(double alpha) { return (uint32_t)llround(alpha * 65536.0); }

uint64_t orc_term_threshold(double term)
{
    if (term <= 0.0) return 0;
    return (uint64_t)floor(term * 4294967296.0);
}

static inline uint32_t bitlen(uint64_t n) { return n ? 64u - (uint32_t)__builtin_clzll(n) : 0u; }

int orc_sopr_use_mixture(uint64_t du, uint64_t dv, uint32_t a, uint32_t b)
{
    uint64_t bdu = (uint64_t)b * du;
    uint64_t num = bdu + (uint64_t)a * dv;
    uint64_t e = (num + bdu - 1) / bdu;
    uint64_t rej = e * (1 + bitlen(du));
    uint64_t mix = du * (4 * (uint64_t)bitlen(dv) + 2);
    return mix < rej;
}

/* textbook membership search, engine/engine.hpp:224-243 */
static int member(const uint32_t* a, uint64_t n, uint32_t x, uint64_t* probes)
{
    int64_t lo = 0, hi = (int64_t)n - 1;
    while (lo <= hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        uint32_t m = a[mid];
        (*probes)++;
        if (m == x) return 1;
        if (m < x) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* The same membership test with a different probe order (not in the reference; its
 * answer equals member()'s for every sorted list, so walks do not change): the next probe
 * is interpolated from the value bounds [vlo, vhi) of the remaining range in IEEE single
 * precision (the kernel computes the same positions), which lands near x in a few probes
 * when ids are spread evenly (randomly permuted graphs); when a probe fails to halve the
 * range the next one bisects, and ranges of at most 16 entries are bisected. */
static int member_interp(const uint32_t* a, uint64_t n, uint32_t x, uint32_t id_end, uint64_t* probes)
{
    int64_t lo = 0, hi = (int64_t)n - 1;
    uint32_t vlo = 0, vhi = id_end;
    int bisect = 0;
    if (x >= vhi) return 0;
    while (lo <= hi) {
        const uint32_t len = (uint32_t)(hi - lo + 1);
        int64_t mid;
        if (bisect || len <= 16 || vhi - vlo < 2) {
            mid = lo + ((hi - lo) >> 1);
        } else {
            volatile float num = (float)(x - vlo) * (float)len;   /* no excess precision */
            const uint32_t off = (uint32_t)(num / (float)(vhi - vlo));
            mid = lo + (off < len ? off : len - 1);
        }
        uint32_t m = a[mid];
        (*probes)++;
        if (m == x) return 1;
        if (m < x) { lo = mid + 1; vlo = m + 1; } else { hi = mid - 1; vhi = m; }
        bisect = (uint64_t)(hi - lo + 1) * 2 > len;
    }
    return 0;
}

/* Window orders (not in the reference; same answer as member() for every sorted list):
 * each probe reads the 16 entries of the list that lie in one 16-entry piece of the whole
 * graph's edge array, aligned to 16 there (head = index of a[0] in that array), around
 * the guessed position -- one 64-byte read of a 128-byte line, so the last few levels of
 * a search cost one probe instead of one each.  The guess bisects the remaining range
 * (search 2) or is member_interp's interpolated position (search 3, bisecting after a
 * probe that failed to halve the range, with no small-range rule).  A probe that does
 * not find x either shows it absent (x falls between two entries read) or cuts the range
 * to one side of the entries read, whose values become the new bound. */
static int member_window(const uint32_t* a, uint64_t n, uint32_t x, int interp, uint32_t id_end, uint64_t head,
                         uint64_t* probes)
{
    uint64_t lo = 0, hi = n;   /* remaining range [lo, hi) */
    uint32_t vlo = 0, vhi = id_end;
    int bisect = 0;
    if (interp && x >= vhi) return 0;
    while (lo < hi) {
        const uint32_t len = (uint32_t)(hi - lo);
        uint64_t g;
        if (!interp || bisect || vhi - vlo < 2) {
            g = lo + ((len - 1) >> 1);
        } else {
            volatile float num = (float)(x - vlo) * (float)len;   /* no excess precision */
            const uint32_t off = (uint32_t)(num / (float)(vhi - vlo));
            g = lo + (off < len ? off : len - 1);
        }
        const int64_t w0 = (int64_t)g - (int64_t)((head + g) & 15);
        const uint64_t s = w0 > (int64_t)lo ? (uint64_t)w0 : lo;
        const uint64_t e = (uint64_t)(w0 + 16) < hi ? (uint64_t)(w0 + 16) : hi;
        (*probes)++;
        for (uint64_t i = s; i < e; ++i)
            if (a[i] == x) return 1;
        if (x < a[s]) { hi = s; vhi = a[s]; }
        else if (x > a[e - 1]) { lo = e; vlo = a[e - 1] + 1; }
        else return 0;
        bisect = (uint64_t)(hi - lo) * 2 > len;
    }
    return 0;
}

int orc_member(const uint32_t* a, uint64_t n, uint32_t x, uint32_t search, uint32_t id_end, uint64_t head,
               uint64_t* probes)
{
    if (search & 2) return member_window(a, n, x, (int)(search & 1), id_end, head, probes);
    return search ? member_interp(a, n, x, id_end, probes) : member(a, n, x, probes);
}

static uint64_t lower_bound(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t x, uint64_t* probes)
{
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        (*probes)++;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static uint64_t upper_bound(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t x, uint64_t* probes)
{
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        (*probes)++;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* Walk the distinct values y of N(u) (y != u) in ascending order and find their runs
 * in N(v).  select == UINT64_MAX: return the total count c.  Otherwise return the
 * index (into N(v)) of the select-th matching position. */
static uint64_t intersect(const uint32_t* nu, uint64_t du, const uint32_t* nv, uint64_t dv,
                          uint32_t u, uint64_t select, uint64_t* probes)
{
    uint64_t c = 0, base = 0;
    uint32_t last = 0;
    int have_last = 0;
    for (uint64_t i = 0; i < du && base < dv; i++) {
        uint32_t y = nu[i];
        (*probes)++;
        if ((have_last && y == last) || y == u) { last = y; have_last = 1; continue; }
        last = y; have_last = 1;
        uint64_t lb = lower_bound(nv, base, dv, y, probes);
        if (lb < dv) {
            (*probes)++;
            if (nv[lb] == y) {
                uint64_t ub = upper_bound(nv, lb, dv, y, probes);
                uint64_t cnt = ub - lb;
                if (select != UINT64_MAX && select < c + cnt) return lb + (select - c);
                c += cnt;
                base = ub;
                continue;
            }
        }
        base = lb;
    }
    return c;
}

static void walk_impl(const orc_graph* g, const orc_walkcfg* cfg, uint32_t wid, uint32_t type,
                      uint32_t source, uint32_t* path, uint32_t* visits, orc_walk_result* res)
{
    const orc_typecfg* tc = &cfg->type[type & 1];
    const uint32_t* csr = g->csr;
    uint32_t cur = source, prev = source, hop = 0, status = 1;
    uint64_t hash = (FNV_OFFSET ^ source) * FNV_PRIME;
    uint64_t steps = 0, trials = 0, probes = 0, mix = 0;
    uint64_t u_head = g->beg[prev], du = g->beg[prev + 1] - u_head;
    if (path) {
        for (uint32_t i = 0; i <= tc->length; i++) path[i] = 0xFFFFFFFFu;
        path[0] = source;
    }
    while (hop < tc->length) {
        uint32_t d[4];
        if (tc->t_term) {
            draw(cfg->seed, wid, hop, 0, 1, d);
            if ((uint64_t)d[0] < tc->t_term) { status = 2; break; }
        }
        uint64_t v_head = g->beg[cur], dv = g->beg[cur + 1] - v_head;
        if (dv == 0) { status = 3; break; }
        uint32_t x = 0;
        int ok = 0;
        if (tc->algo == ORC_N2V) {
            /* prev's list decides only between the thresholds of the in-N(prev) and other classes */
            const uint64_t band_lo = tc->t_one < tc->t_q ? tc->t_one : tc->t_q;
            const uint64_t band_hi = tc->t_one < tc->t_q ? tc->t_q : tc->t_one;
            for (uint32_t k = 0; k < ORC_MAX_TRIALS; k++) {
                if ((k & 1) == 0) draw(cfg->seed, wid, hop, k >> 1, 0, d);
                uint32_t rp = (k & 1) ? d[2] : d[0];
                uint64_t ra = (k & 1) ? d[3] : d[1];
                trials++;
                x = csr[v_head + (((uint64_t)rp * dv) >> 32)];
                if (ra < tc->t_min) { ok = 1; break; }
                uint64_t t;
                if (x == prev) t = tc->t_p;
                else if (ra < band_lo) { ok = 1; break; }   /* both x != prev classes accept */
                else if (ra >= band_hi) continue;           /* both reject */
                else if (orc_member(csr + u_head, du, x, tc->search, (uint32_t)g->nverts, u_head, &probes)) t = tc->t_one;
                else t = tc->t_q;
                if (ra < t) { ok = 1; break; }
            }
        } else if (tc->sopr_sampler == 0) {
            const uint64_t tb = (uint64_t)tc->alpha_b << 16;
            for (uint32_t k = 0; k < ORC_MAX_TRIALS; k++) {
                draw(cfg->seed, wid, hop, k, 0, d);
                trials++;
                if ((uint64_t)d[0] < tb) {
                    x = csr[v_head + (((uint64_t)d[1] * dv) >> 32)];
                    ok = 1;
                    break;
                }
                uint32_t y = csr[u_head + (((uint64_t)d[1] * du) >> 32)];
                if (y != prev && orc_member(csr + v_head, dv, y, tc->search, (uint32_t)g->nverts, v_head, &probes)) {
                    x = y;
                    ok = 1;
                    break;
                }
            }
        } else {
            uint64_t A = tc->alpha_a, B = tc->alpha_b;
            u128 num32 = (u128)(B * du) << 32;
            if (orc_sopr_use_mixture(du, dv, tc->alpha_a, tc->alpha_b)) {
                mix++;
                trials++;
                draw(cfg->seed, wid, hop, 0, 0, d);
                uint64_t c = intersect(csr + u_head, du, csr + v_head, dv, prev, UINT64_MAX, &probes);
                if (c == 0 || (u128)d[0] * (B * du + A * c) < num32) {
                    x = csr[v_head + (((uint64_t)d[1] * dv) >> 32)];
                } else {
                    uint64_t k = ((uint64_t)d[2] * c) >> 32;
                    uint64_t pos = intersect(csr + u_head, du, csr + v_head, dv, prev, k, &probes);
                    x = csr[v_head + pos];
                }
                ok = 1;
            } else {
                uint64_t den = B * du + A * dv;
                for (uint32_t k = 0; k < ORC_MAX_TRIALS; k++) {
                    if ((k & 1) == 0) draw(cfg->seed, wid, hop, k >> 1, 0, d);
                    uint32_t rp = (k & 1) ? d[2] : d[0];
                    uint64_t ra = (k & 1) ? d[3] : d[1];
                    trials++;
                    x = csr[v_head + (((uint64_t)rp * dv) >> 32)];
                    if ((u128)ra * den < num32) { ok = 1; break; }
                    if (x != prev && orc_member(csr + u_head, du, x, tc->search, (uint32_t)g->nverts, u_head, &probes)) { ok = 1; break; }
                }
            }
        }
        if (!ok) { status = 4; break; }
        prev = cur;
        cur = x;
        u_head = v_head;
        du = dv;
        hop++;
        steps++;
        hash = (hash ^ x) * FNV_PRIME;
        if (path) path[hop] = x;
        if (visits) __atomic_fetch_add(&visits[x], 1u, __ATOMIC_RELAXED);
    }
    res->cur = cur; res->prev = prev; res->hop = hop; res->status = status;
    res->hash = hash;
    res->steps = steps; res->trials = trials; res->probes = probes; res->mix_steps = mix;
}

void orc_walk(const orc_graph* g, const orc_walkcfg* cfg, uint32_t wid, uint32_t type,
              uint32_t source, uint32_t* path, orc_walk_result* res)
{
    walk_impl(g, cfg, wid, type, source, path, NULL, res);
}

void orc_run(const orc_graph* g, const orc_walkcfg* cfg, uint64_t nwalkers,
             const uint32_t* sources, uint32_t wpv, const uint32_t* types, uint32_t ntypes_mod,
             uint32_t* paths, uint64_t* hashes, uint32_t* final_state,
             uint32_t* visits, orc_totals* totals, int nthreads)
{
    uint64_t s_steps = 0, s_trials = 0, s_probes = 0, s_mix = 0;
    if (wpv == 0) wpv = 1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) reduction(+ : s_steps, s_trials, s_probes, s_mix)
#endif
    for (int64_t i = 0; i < (int64_t)nwalkers; i++) {
        uint32_t w = (uint32_t)i;
        uint32_t src = sources ? sources[w] : (uint32_t)(w / wpv);
        uint32_t ty = types ? types[w] : (ntypes_mod ? (w % ntypes_mod) : 0);
        uint32_t len = cfg->type[ty & 1].length;
        orc_walk_result r;
        uint32_t* path = paths ? paths + (uint64_t)w * (len + 1) : NULL;
        walk_impl(g, cfg, w, ty, src, path, visits, &r);
        if (hashes) hashes[w] = r.hash;
        if (final_state) {
            final_state[4 * (uint64_t)w + 0] = r.cur;
            final_state[4 * (uint64_t)w + 1] = r.prev;
            final_state[4 * (uint64_t)w + 2] = r.hop;
            final_state[4 * (uint64_t)w + 3] = r.status;
        }
        s_steps += r.steps; s_trials += r.trials; s_probes += r.probes; s_mix += r.mix_steps;
    }
    if (totals) {
        totals->steps = s_steps; totals->trials = s_trials; totals->probes = s_probes;
        totals->mix_steps = s_mix; totals->walkers = nwalkers;
    }
}

/* Transcription of sidewalkupdate (engine/engine.hpp:158-287) and SOPR_sidewalkupdate
 * (engine/engine.hpp:289-404) over one fully resident block: the block-exit `else break`
 * (:272-273, :389-390) is removed so that the walk runs to `length`. */
static uint64_t ref_walk(const orc_graph* g, const orc_refcfg* c, uint32_t source,
                         orc_randnum* rn, uint32_t* visits, uint64_t* trials)
{
    const uint32_t* csr = g->csr;
    uint32_t cur = source, prev = source, hop = 0;
    uint64_t moves = 0, tr = 0;
    float inv_p = 1.0f / c->p, inv_q = 1.0f / c->q;
    float n2v_max = inv_p > (inv_q > 1.0f ? inv_q : 1.0f) ? inv_p : (inv_q > 1.0f ? inv_q : 1.0f);
    float n2v_min = inv_p < (inv_q < 1.0f ? inv_q : 1.0f) ? inv_p : (inv_q < 1.0f ? inv_q : 1.0f);
    float alpha = c->alpha;
    while (hop < c->length) {
        if (c->term > 0.0 && randnum_drand(rn) < c->term) break;
        uint64_t adj_head = g->beg[cur], adj_tail = g->beg[cur + 1];
        uint64_t prev_adj_head = g->beg[prev], prev_adj_tail = g->beg[prev + 1];
        uint64_t deg = adj_tail - adj_head;
        uint32_t next = 0;
        int real = 1;
        if (deg == 0) {
            hop = c->length - 1;
            real = 0;
        } else {
            float max_val, min_val, inv_deg = 0.0f, inv_prev_deg = 0.0f;
            if (c->algo == ORC_N2V) {
                max_val = n2v_max;
                min_val = n2v_min;
            } else {
                uint64_t prev_deg = prev_adj_tail - prev_adj_head;
                inv_deg = (float)((1.0 - alpha) / (double)deg);
                inv_prev_deg = alpha / (float)prev_deg;
                float s = inv_deg + inv_prev_deg;
                max_val = inv_deg > s ? inv_deg : s;
                min_val = inv_deg < s ? inv_deg : s;
            }
            uint64_t pos = 0;
            for (;;) {
                float rv;
                orc_randnum_pair(rn, max_val, deg, c->literal, &rv, &pos);
                tr++;
                uint32_t cand = csr[adj_head + pos];
                if (rv <= min_val) break;
                if (cand == prev) {
                    if (c->algo == ORC_N2V) { if (rv < inv_p) break; }
                    else { if ((double)rv < (1.0 - alpha) / (double)deg) break; }
                } else {
                    int found = 0;
                    int64_t left = (int64_t)prev_adj_head, right = (int64_t)prev_adj_tail - 1;
                    while (left <= right) {
                        int64_t mid = left + ((right - left) >> 1);
                        uint32_t mv = csr[mid];
                        if (mv == cand) { found = 1; break; }
                        if (mv < cand) left = mid + 1; else right = mid - 1;
                    }
                    if (c->algo == ORC_N2V) {
                        if (found) { if (rv < 1.0f) break; }
                        else { if (rv < inv_q) break; }
                    } else {
                        if (found) { if (rv < inv_deg + inv_prev_deg) break; }
                        else { if (rv < inv_deg) break; }
                    }
                }
            }
            next = csr[adj_head + pos];
        }
        prev = cur;
        cur = next;
        hop++;
        if (real) {
            moves++;
            if (visits) __atomic_fetch_add(&visits[next], 1u, __ATOMIC_RELAXED);
        }
    }
    *trials += tr;
    return moves;
}

/* Per-walker RandNum seed: the reference seeds per thread (time(NULL) + tid,
 * engine/engine.hpp:98-101); here each walker gets its own stream, decorrelated across
 * run seeds by a 64-bit finaliser (MurmurHash3 fmix64). */
static uint64_t fmix64(uint64_t k)
{
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

static uint64_t walker_seed(uint64_t seed, uint64_t w) { return fmix64(seed) ^ fmix64(w + 0x632BE59BD9B4E019ULL); }

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

uint64_t orc_ref_run(const orc_graph* g, const orc_refcfg* cfg, uint64_t w0, uint64_t nw,
                     const uint32_t* sources, uint32_t wpv, uint32_t* visits,
                     uint64_t* trials, double max_seconds, uint64_t* done, int nthreads)
{
    uint64_t moves = 0, tr = 0, fin = 0;
    int stop = 0;
    double t0 = now_s();
    if (wpv == 0) wpv = 1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : moves, tr, fin)
#endif
    for (int64_t i = 0; i < (int64_t)nw; i++) {
        if (__atomic_load_n(&stop, __ATOMIC_RELAXED)) continue;
        uint64_t w = w0 + (uint64_t)i;
        uint32_t src = sources ? sources[w] : (uint32_t)(w / wpv);
        orc_randnum rn;
        orc_randnum_init(&rn, walker_seed(cfg->seed, w));
        moves += ref_walk(g, cfg, src, &rn, visits, &tr);
        fin++;
        if (max_seconds > 0.0 && (i & 255) == 0 && now_s() - t0 > max_seconds)
            __atomic_store_n(&stop, 1, __ATOMIC_RELAXED);
    }
    if (trials) *trials = tr;
    if (done) *done = fin;
    return moves;
}

/* split_blocks, preprocess/split.hpp:15-87, over a fully loaded .beg array.  The
 * reference streams .beg in VERT_SIZE chunks; for graphs under 2^30 ids that is the
 * same single pass. */
size_t orc_split_blocks(const uint64_t* beg, uint64_t entries, uint64_t block_size,
                        uint32_t* vblocks, uint64_t* eblocks, size_t cap)
{
    uint64_t max_nedges = block_size / 4, max_nverts = block_size / 8;
    uint32_t cur_pos = 0;
    uint64_t rd_edges = 0;
    size_t n = 0;
    if (cap < 2 || entries == 0) return 0;
    vblocks[n] = 0; eblocks[n] = 0; n++;
    for (uint64_t v = 0; v < entries; v++) {
        if (beg[v] - rd_edges > max_nedges) {
            if (n >= cap) return 0;
            cur_pos = (uint32_t)(v - 1);
            rd_edges = beg[v - 1];
            vblocks[n] = cur_pos; eblocks[n] = rd_edges; n++;
        } else if ((uint32_t)v - cur_pos >= max_nverts) {
            if (n >= cap) return 0;
            cur_pos = (uint32_t)v;
            rd_edges = beg[v];
            vblocks[n] = cur_pos; eblocks[n] = rd_edges; n++;
        }
    }
    if (n >= cap) return 0;
    vblocks[n] = (uint32_t)(entries - 1);
    eblocks[n] = beg[entries - 1];
    n++;
    return n;
}

/* GOwalker_scheduler_t::max_choose_blocks, engine/schedule.hpp:70-138.  std::sort's
 * unspecified tie order (:103,:127) is replaced by "lowest candidate id wins". */
int orc_max_choose(const uint64_t* W, uint32_t nblocks, uint32_t* buckets, uint32_t* nbuckets,
                   uint32_t ncblock, uint64_t zero_threshold)
{
    uint32_t nc = *nbuckets;
    int found = 0;
    uint64_t best = 0;
    uint32_t best_blk = 0, best_pos = 0;
    for (uint32_t c = 0; c < nblocks; c++) {
        int cached = 0;
        for (uint32_t j = 0; j < nc; j++) if (buckets[j] == c) { cached = 1; break; }
        if (cached) continue;
        uint64_t score = 0;
        uint32_t pos = 0;
        if (nc >= ncblock) {
            uint64_t mn = 0xffffffffULL, all = 0;
            for (uint32_t j = 0; j < nc; j++) {
                uint64_t s = W[(uint64_t)c * nblocks + buckets[j]] + W[(uint64_t)buckets[j] * nblocks + c];
                all += s;
                if (s < mn) { mn = s; pos = j; }
            }
            score = all - mn;
        } else {
            for (uint32_t j = 0; j < nc; j++)
                score += W[(uint64_t)c * nblocks + buckets[j]] + W[(uint64_t)buckets[j] * nblocks + c];
            score += W[(uint64_t)c * nblocks + c];
        }
        if (!found || score > best) { found = 1; best = score; best_blk = c; best_pos = pos; }
    }
    if (!found || best == 0 || best <= zero_threshold) return 0;
    if (nc >= ncblock) buckets[best_pos] = best_blk;
    else buckets[(*nbuckets)++] = best_blk;
    return 1;
}

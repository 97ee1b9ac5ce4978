/*
 * Multithreaded C oracle of the four benchmark workloads.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU-side checker and the CPU baseline of the
 * benchmark.  Nothing on the product path links or calls it.
 *
 * Each function restates the plan semantics of the reference for one plan
 * shape, over plain column arrays, with pthreads:
 *   selection  -> keep rows whose predicate holds (cudaTranslator.py:358-366)
 *   map        -> C double arithmetic, no fused multiply-add
 *                 (cudaTranslator.py:410-422; compiled with -ffp-contract=off)
 *   equi-join  -> unique build table for multimatch=False (hashJoins.py:95-153),
 *                 multimap count / scan / insert for multimatch=True
 *                 (hashJoins.py:214-240, 465-534)
 *   group-by   -> hash aggregation (cudaTranslator.py:425-525); FP64 sums use
 *                 Neumaier-compensated accumulation per group and thread, so
 *                 the oracle's own error stays far below 1e-9 relative.
 * Hash tables here are ordinary open addressing with a multiplicative hash;
 * nothing is shared with the GPU implementation.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double s, c;
} ksum;

static inline void kadd(ksum* k, double v) {
    double t = k->s + v;
    if ((k->s >= 0 ? k->s : -k->s) >= (v >= 0 ? v : -v))
        k->c += (k->s - t) + v;
    else
        k->c += (v - t) + k->s;
    k->s = t;
}

static inline void kmerge(ksum* a, const ksum* b) {
    kadd(a, b->s);
    kadd(a, b->c);
}

static inline double kval(const ksum* k) { return k->s + k->c; }

static inline uint64_t hmix(uint64_t x) {
    x ^= x >> 31;
    x *= 0x7fb5d329728ea185ull;
    x ^= x >> 27;
    x *= 0x81dadef4bc2dd44dull;
    x ^= x >> 33;
    return x;
}

typedef struct {
    int nthreads;
    int64_t n;
    void* arg;
    void (*fn)(void* arg, int t, int64_t lo, int64_t hi);
} pool_job;

typedef struct {
    pool_job* job;
    int t;
} pool_slot;

static void* pool_run(void* p) {
    pool_slot* s = (pool_slot*)p;
    pool_job* j = s->job;
    int64_t per = (j->n + j->nthreads - 1) / j->nthreads;
    int64_t lo = (int64_t)s->t * per, hi = lo + per;
    if (lo > j->n) lo = j->n;
    if (hi > j->n) hi = j->n;
    j->fn(j->arg, s->t, lo, hi);
    return NULL;
}

static void parallel_for(int nthreads, int64_t n, void* arg, void (*fn)(void*, int, int64_t, int64_t)) {
    if (nthreads < 1) nthreads = 1;
    pool_job job = {nthreads, n, arg, fn};
    pthread_t th[256];
    pool_slot slots[256];
    if (nthreads > 256) nthreads = job.nthreads = 256;
    for (int t = 0; t < nthreads; ++t) {
        slots[t].job = &job;
        slots[t].t = t;
        pthread_create(&th[t], NULL, pool_run, &slots[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------ c0 */
/* scan cfg0 -> date <= cutoff -> disc_price, charge -> group (g1, g2):
 * SUM qty, price, disc_price, charge, disc; COUNT.  Keys (g1,g2) index a
 * 65536-entry direct table (two int8 codes). */
typedef struct {
    const int8_t *g1, *g2;
    const int32_t* qty;
    const double *price, *disc, *tax;
    const uint32_t* date;
    uint32_t cutoff;
    ksum (*acc)[65536][5];
    int64_t (*cnt)[65536];
} c0_arg;

static void c0_part(void* p, int t, int64_t lo, int64_t hi) {
    c0_arg* a = (c0_arg*)p;
    ksum(*acc)[5] = a->acc[t];
    int64_t* cnt = a->cnt[t];
    for (int64_t i = lo; i < hi; ++i) {
        if (!(a->date[i] <= a->cutoff)) continue;
        double price = a->price[i], disc = a->disc[i];
        double charge = (price * (1.0f - disc)) * (1.0f + a->tax[i]);
        double disc_price = price * (1.0f - disc);
        int key = ((uint8_t)a->g1[i]) | ((int)(uint8_t)a->g2[i] << 8);
        kadd(&acc[key][0], (double)a->qty[i]);
        kadd(&acc[key][1], price);
        kadd(&acc[key][2], disc_price);
        kadd(&acc[key][3], charge);
        kadd(&acc[key][4], disc);
        cnt[key] += 1;
    }
}

/* out_sums: 65536 x 5 doubles, out_cnt: 65536 counts (indexed by g1 | g2<<8) */
int dqc_oracle_c0(const int8_t* g1, const int8_t* g2, const int32_t* qty, const double* price, const double* disc,
                  const double* tax, const uint32_t* date, int64_t n, uint32_t cutoff, int nthreads,
                  double* out_sums, int64_t* out_cnt) {
    c0_arg a = {g1, g2, qty, price, disc, tax, date, cutoff, NULL, NULL};
    a.acc = calloc((size_t)nthreads, sizeof(*a.acc));
    a.cnt = calloc((size_t)nthreads, sizeof(*a.cnt));
    if (!a.acc || !a.cnt) return -1;
    parallel_for(nthreads, n, &a, c0_part);
    for (int k = 0; k < 65536; ++k) {
        ksum tot[5];
        memset(tot, 0, sizeof(tot));
        int64_t c = 0;
        for (int t = 0; t < nthreads; ++t) {
            for (int j = 0; j < 5; ++j) kmerge(&tot[j], &a.acc[t][k][j]);
            c += a.cnt[t][k];
        }
        for (int j = 0; j < 5; ++j) out_sums[k * 5 + j] = kval(&tot[j]);
        out_cnt[k] = c;
    }
    free(a.acc);
    free(a.cnt);
    return 0;
}

/* ------------------------------------------------------------------ c1 */
/* dim: d_cat == cat -> unique table on d_key; fact: f_date in [lo, hi) ->
 * probe f_fk -> group by d_key, SUM f_price * (1 - f_disc). */
typedef struct {
    int64_t cap;
    int32_t* keys;
    uint8_t* used;
} keyset;

static int keyset_init(keyset* s, int64_t n) {
    s->cap = 64;
    while (s->cap < 2 * n) s->cap <<= 1;
    s->keys = malloc((size_t)s->cap * sizeof(int32_t));
    s->used = calloc((size_t)s->cap, 1);
    return (s->keys && s->used) ? 0 : -1;
}

static void keyset_add(keyset* s, int32_t k) {
    uint64_t h = hmix((uint64_t)(uint32_t)k) & (uint64_t)(s->cap - 1);
    while (s->used[h] && s->keys[h] != k) h = (h + 1) & (uint64_t)(s->cap - 1);
    s->used[h] = 1;
    s->keys[h] = k;
}

static int keyset_has(const keyset* s, int32_t k) {
    uint64_t h = hmix((uint64_t)(uint32_t)k) & (uint64_t)(s->cap - 1);
    while (s->used[h]) {
        if (s->keys[h] == k) return 1;
        h = (h + 1) & (uint64_t)(s->cap - 1);
    }
    return 0;
}

typedef struct {
    int64_t cap, n;
    int32_t* keys;
    uint8_t* used;
    ksum* sums;
} groupmap;

static int groupmap_init(groupmap* g, int64_t cap) {
    g->cap = 1024;
    while (g->cap < cap) g->cap <<= 1;
    g->n = 0;
    g->keys = malloc((size_t)g->cap * sizeof(int32_t));
    g->used = calloc((size_t)g->cap, 1);
    g->sums = calloc((size_t)g->cap, sizeof(ksum));
    return (g->keys && g->used && g->sums) ? 0 : -1;
}

static void groupmap_free(groupmap* g) {
    free(g->keys);
    free(g->used);
    free(g->sums);
}

static ksum* groupmap_slot(groupmap* g, int32_t k);

static void groupmap_grow(groupmap* g) {
    groupmap o = *g;
    groupmap_init(g, o.cap * 2);
    for (int64_t i = 0; i < o.cap; ++i)
        if (o.used[i]) *groupmap_slot(g, o.keys[i]) = o.sums[i];
    groupmap_free(&o);
}

static ksum* groupmap_slot(groupmap* g, int32_t k) {
    if (2 * (g->n + 1) > g->cap) groupmap_grow(g);
    uint64_t h = hmix((uint64_t)(uint32_t)k) & (uint64_t)(g->cap - 1);
    while (g->used[h]) {
        if (g->keys[h] == k) return &g->sums[h];
        h = (h + 1) & (uint64_t)(g->cap - 1);
    }
    g->used[h] = 1;
    g->keys[h] = k;
    g->n++;
    return &g->sums[h];
}

typedef struct {
    const int32_t* fk;
    const double *price, *disc;
    const uint32_t* date;
    uint32_t lo, hi;
    const keyset* dims;
    groupmap* maps;
} c1_arg;

static void c1_part(void* p, int t, int64_t lo, int64_t hi) {
    c1_arg* a = (c1_arg*)p;
    groupmap* g = &a->maps[t];
    for (int64_t i = lo; i < hi; ++i) {
        uint32_t d = a->date[i];
        if (!(d >= a->lo && d < a->hi)) continue;
        int32_t k = a->fk[i];
        if (!keyset_has(a->dims, k)) continue;
        double rev = a->price[i] * (1.0f - a->disc[i]);
        kadd(groupmap_slot(g, k), rev);
    }
}

/* returns the number of groups written to out_keys / out_sums (capacity cap) */
int64_t dqc_oracle_c1(const int32_t* d_key, const int8_t* d_cat, int64_t nd, int8_t cat, const int32_t* f_fk,
                      const double* f_price, const double* f_disc, const uint32_t* f_date, int64_t nf,
                      uint32_t date_lo, uint32_t date_hi, int nthreads, int32_t* out_keys, double* out_sums,
                      int64_t cap) {
    keyset dims;
    int64_t nsel = 0;
    for (int64_t i = 0; i < nd; ++i) nsel += d_cat[i] == cat;
    if (keyset_init(&dims, nsel)) return -1;
    for (int64_t i = 0; i < nd; ++i)
        if (d_cat[i] == cat) keyset_add(&dims, d_key[i]);
    groupmap* maps = calloc((size_t)nthreads, sizeof(groupmap));
    for (int t = 0; t < nthreads; ++t) groupmap_init(&maps[t], 4096);
    c1_arg a = {f_fk, f_price, f_disc, f_date, date_lo, date_hi, &dims, maps};
    parallel_for(nthreads, nf, &a, c1_part);
    groupmap total;
    groupmap_init(&total, 4096);
    for (int t = 0; t < nthreads; ++t) {
        for (int64_t i = 0; i < maps[t].cap; ++i)
            if (maps[t].used[i]) kmerge(groupmap_slot(&total, maps[t].keys[i]), &maps[t].sums[i]);
        groupmap_free(&maps[t]);
    }
    free(maps);
    int64_t n = 0;
    for (int64_t i = 0; i < total.cap; ++i) {
        if (!total.used[i]) continue;
        if (n < cap) {
            out_keys[n] = total.keys[i];
            out_sums[n] = kval(&total.sums[i]);
        }
        ++n;
    }
    groupmap_free(&total);
    free(dims.keys);
    free(dims.used);
    return n;
}

/* ------------------------------------------------------------------ c2 */
/* build (b_key, b_pay) multimap; probe p_f < fmax -> every match:
 * COUNT, SUM b_pay * p_pay (int64, wrapping). */
typedef struct {
    int64_t nb;
    uint32_t mask;
    uint32_t *off, *cnt;
    int32_t* ekey;
    int64_t* epay;
} multimap;

typedef struct {
    const multimap* mm;
    const int32_t *pk, *pf;
    const int64_t* pp;
    int32_t fmax;
    int64_t* counts;
    uint64_t* sums;
} c2_arg;

static void c2_part(void* p, int t, int64_t lo, int64_t hi) {
    c2_arg* a = (c2_arg*)p;
    const multimap* m = a->mm;
    int64_t c = 0;
    uint64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) {
        if (!(a->pf[i] < a->fmax)) continue;
        int32_t k = a->pk[i];
        uint32_t b = (uint32_t)(hmix((uint64_t)(uint32_t)k) & m->mask);
        for (uint32_t j = m->off[b], e = m->off[b] + m->cnt[b]; j < e; ++j) {
            if (m->ekey[j] != k) continue;
            c += 1;
            s += (uint64_t)m->epay[j] * (uint64_t)a->pp[i];
        }
    }
    a->counts[t] = c;
    a->sums[t] = s;
}

int dqc_oracle_c2(const int32_t* b_key, const int64_t* b_pay, int64_t nb, const int32_t* p_key, const int32_t* p_f,
                  const int64_t* p_pay, int64_t np, int32_t fmax, int nthreads, int64_t* out_count,
                  int64_t* out_sum) {
    multimap m;
    uint32_t nbk = 64;
    while (nbk < 2 * nb) nbk <<= 1;
    m.mask = nbk - 1;
    m.off = calloc(nbk, sizeof(uint32_t));
    m.cnt = calloc(nbk, sizeof(uint32_t));
    m.ekey = malloc((size_t)(nb ? nb : 1) * sizeof(int32_t));
    m.epay = malloc((size_t)(nb ? nb : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < nb; ++i) m.cnt[hmix((uint64_t)(uint32_t)b_key[i]) & m.mask]++;
    uint32_t run = 0;
    for (uint32_t b = 0; b < nbk; ++b) {
        m.off[b] = run;
        run += m.cnt[b];
        m.cnt[b] = 0;
    }
    for (int64_t i = 0; i < nb; ++i) {
        uint32_t b = (uint32_t)(hmix((uint64_t)(uint32_t)b_key[i]) & m.mask);
        uint32_t pos = m.off[b] + m.cnt[b]++;
        m.ekey[pos] = b_key[i];
        m.epay[pos] = b_pay[i];
    }
    int64_t counts[256];
    uint64_t sums[256];
    c2_arg a = {&m, p_key, p_f, p_pay, fmax, counts, sums};
    parallel_for(nthreads, np, &a, c2_part);
    int64_t c = 0;
    uint64_t s = 0;
    for (int t = 0; t < nthreads && t < 256; ++t) {
        c += counts[t];
        s += sums[t];
    }
    *out_count = c;
    *out_sum = (int64_t)s;
    free(m.off);
    free(m.cnt);
    free(m.ekey);
    free(m.epay);
    return 0;
}

/* ------------------------------------------------------------------ c3 */
/* item: LIKE '%word%' on a STRING column (offsets + heap): for a pattern
 * whose only wildcards are a leading and a trailing '%', SQL LIKE is a
 * substring test (util.cuh:144-199 reduces to that case).  sales probes
 * item (unique, filtered), cust (unique), nation (unique) and every joined
 * row is aggregated twice: by (n_key, c_year) and by i_key, each with
 * SUM(s_qty) (exact in int64, a double aggregate in the plan), SUM(s_price)
 * (compensated) and COUNT. */
typedef struct {
    int64_t cap, n;
    int64_t* keys;
    int64_t* vals;
    uint8_t* used;
} i64map;

static int i64map_init(i64map* m, int64_t expect) {
    int64_t cap = 64;
    while (cap < 2 * expect + 1) cap *= 2;
    m->cap = cap;
    m->n = 0;
    m->keys = calloc((size_t)cap, sizeof(int64_t));
    m->vals = calloc((size_t)cap, sizeof(int64_t));
    m->used = calloc((size_t)cap, 1);
    return (m->keys && m->vals && m->used) ? 0 : -1;
}

static void i64map_free(i64map* m) {
    free(m->keys);
    free(m->vals);
    free(m->used);
}

/* slot of key k: existing, or claimed (used=1, vals=-1) when insert != 0, else -1 */
static int64_t i64map_slot(i64map* m, int64_t k, int insert) {
    uint64_t s = hmix((uint64_t)k) & (uint64_t)(m->cap - 1);
    for (;;) {
        if (!m->used[s]) {
            if (!insert) return -1;
            m->used[s] = 1;
            m->keys[s] = k;
            m->vals[s] = -1;
            m->n++;
            return (int64_t)s;
        }
        if (m->keys[s] == k) return (int64_t)s;
        s = (s + 1) & (uint64_t)(m->cap - 1);
    }
}

typedef struct {
    int64_t qty, cnt;
    ksum rev;
} c3_acc;

typedef struct {
    const int32_t *s_item, *s_cust, *s_nation, *s_qty;
    const double* s_price;
    const i64map *items, *custs, *nations;
    int64_t nitems_pass, ngroups_ny;
    c3_acc** by_item;   /* per thread: nitems_pass */
    c3_acc** by_ny;     /* per thread: ngroups_ny = nations x years (dense) */
    int32_t year_lo;
    int32_t nyears;
} c3_arg;

static void c3_part(void* p, int t, int64_t lo, int64_t hi) {
    c3_arg* a = (c3_arg*)p;
    c3_acc* bi = a->by_item[t];
    c3_acc* bn = a->by_ny[t];
    for (int64_t i = lo; i < hi; ++i) {
        int64_t si = i64map_slot((i64map*)a->items, a->s_item[i], 0);
        if (si < 0) continue;
        int64_t sc = i64map_slot((i64map*)a->custs, a->s_cust[i], 0);
        if (sc < 0) continue;
        int64_t sn = i64map_slot((i64map*)a->nations, a->s_nation[i], 0);
        if (sn < 0) continue;
        int64_t q = a->s_qty[i];
        double pr = a->s_price[i];
        c3_acc* x = &bi[a->items->vals[si]];
        x->qty += q;
        x->cnt += 1;
        kadd(&x->rev, pr);
        int64_t g = a->nations->vals[sn] * a->nyears + (a->custs->vals[sc] - a->year_lo);
        c3_acc* y = &bn[g];
        y->qty += q;
        y->cnt += 1;
        kadd(&y->rev, pr);
    }
}

static int contains(const char* s, int64_t n, const char* w, int64_t wn) {
    if (wn == 0) return 1;
    for (int64_t i = 0; i + wn <= n; ++i)
        if (s[i] == w[0] && memcmp(s + i, w, (size_t)wn) == 0) return 1;
    return 0;
}

/* Outputs (capacities from the caller):
 *   item groups:  out_item_key / _qty / _rev / _cnt, at most cap_item rows;
 *   (n_key, c_year) groups: out_ny_nkey / _year / _qty / _rev / _cnt, at most cap_ny rows.
 * Returns the number of item groups (and *n_ny), or -1 on error. */
int64_t dqc_oracle_c3(const int32_t* i_key, const uint64_t* i_offs, const char* i_chars, int64_t ni, const char* word,
                      const int32_t* c_key, const int32_t* c_year, int64_t nc, const int32_t* n_key, int64_t nn,
                      const int32_t* s_item, const int32_t* s_cust, const int32_t* s_nation, const int32_t* s_qty,
                      const double* s_price, int64_t ns, int nthreads, int32_t* out_item_key, double* out_item_qty,
                      double* out_item_rev, int32_t* out_item_cnt, int64_t cap_item, int32_t* out_ny_nkey,
                      int32_t* out_ny_year, double* out_ny_qty, double* out_ny_rev, int32_t* out_ny_cnt,
                      int64_t cap_ny, int64_t* n_ny) {
    if (nthreads < 1) nthreads = 1;
    const int64_t wn = (int64_t)strlen(word);
    i64map items, custs, nations;
    if (i64map_init(&items, ni / 8 + 16) || i64map_init(&custs, nc) || i64map_init(&nations, nn)) return -1;
    int64_t npass = 0;
    for (int64_t i = 0; i < ni; ++i) {
        if (!contains(i_chars + i_offs[i], (int64_t)(i_offs[i + 1] - i_offs[i]), word, wn)) continue;
        if (items.n * 2 + 2 > items.cap) {   /* grow */
            i64map bigger;
            if (i64map_init(&bigger, items.cap)) return -1;
            for (int64_t s = 0; s < items.cap; ++s)
                if (items.used[s]) bigger.vals[i64map_slot(&bigger, items.keys[s], 1)] = items.vals[s];
            i64map_free(&items);
            items = bigger;
        }
        int64_t s = i64map_slot(&items, i_key[i], 1);
        if (items.vals[s] < 0) items.vals[s] = npass++;
    }
    int32_t ylo = 0, yhi = -1;
    for (int64_t i = 0; i < nc; ++i) {
        if (i == 0 || c_year[i] < ylo) ylo = c_year[i];
        if (i == 0 || c_year[i] > yhi) yhi = c_year[i];
        custs.vals[i64map_slot(&custs, c_key[i], 1)] = c_year[i];
    }
    const int32_t nyears = nc > 0 ? yhi - ylo + 1 : 1;
    int64_t nnat = 0;
    for (int64_t i = 0; i < nn; ++i) {
        int64_t s = i64map_slot(&nations, n_key[i], 1);
        if (nations.vals[s] < 0) nations.vals[s] = nnat++;
    }
    const int64_t ngroups = (nnat > 0 ? nnat : 1) * (int64_t)nyears;
    c3_arg a = {s_item, s_cust, s_nation, s_qty, s_price, &items, &custs, &nations, npass, ngroups, NULL, NULL,
                ylo, nyears};
    a.by_item = calloc((size_t)nthreads, sizeof(c3_acc*));
    a.by_ny = calloc((size_t)nthreads, sizeof(c3_acc*));
    for (int t = 0; t < nthreads; ++t) {
        a.by_item[t] = calloc((size_t)(npass + 1), sizeof(c3_acc));
        a.by_ny[t] = calloc((size_t)ngroups, sizeof(c3_acc));
    }
    parallel_for(nthreads, ns, &a, c3_part);
    for (int t = 1; t < nthreads; ++t) {
        for (int64_t g = 0; g < npass; ++g) {
            a.by_item[0][g].qty += a.by_item[t][g].qty;
            a.by_item[0][g].cnt += a.by_item[t][g].cnt;
            kmerge(&a.by_item[0][g].rev, &a.by_item[t][g].rev);
        }
        for (int64_t g = 0; g < ngroups; ++g) {
            a.by_ny[0][g].qty += a.by_ny[t][g].qty;
            a.by_ny[0][g].cnt += a.by_ny[t][g].cnt;
            kmerge(&a.by_ny[0][g].rev, &a.by_ny[t][g].rev);
        }
    }
    int64_t nout = 0;
    for (int64_t s = 0; s < items.cap; ++s) {
        if (!items.used[s]) continue;
        const c3_acc* x = &a.by_item[0][items.vals[s]];
        if (x->cnt == 0) continue;
        if (nout < cap_item) {
            out_item_key[nout] = (int32_t)items.keys[s];
            out_item_qty[nout] = (double)x->qty;
            out_item_rev[nout] = kval(&x->rev);
            out_item_cnt[nout] = (int32_t)x->cnt;
        }
        ++nout;
    }
    int64_t nny = 0;
    for (int64_t s = 0; s < nations.cap; ++s) {
        if (!nations.used[s]) continue;
        for (int32_t y = 0; y < nyears; ++y) {
            const c3_acc* x = &a.by_ny[0][nations.vals[s] * nyears + y];
            if (x->cnt == 0) continue;
            if (nny < cap_ny) {
                out_ny_nkey[nny] = (int32_t)nations.keys[s];
                out_ny_year[nny] = ylo + y;
                out_ny_qty[nny] = (double)x->qty;
                out_ny_rev[nny] = kval(&x->rev);
                out_ny_cnt[nny] = (int32_t)x->cnt;
            }
            ++nny;
        }
    }
    *n_ny = nny;
    for (int t = 0; t < nthreads; ++t) {
        free(a.by_item[t]);
        free(a.by_ny[t]);
    }
    free(a.by_item);
    free(a.by_ny);
    i64map_free(&items);
    i64map_free(&custs);
    i64map_free(&nations);
    return nout;
}

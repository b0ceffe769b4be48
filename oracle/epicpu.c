/*
 * epicpu -- CPU oracle for EpiRust's per-hour agent step in deterministic mode.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never shipped or
 * linked by it. Structure deliberately follows the reference engine: an AoS
 * citizen array, a start-of-hour occupancy snapshot ("current_locations"), a
 * per-agent routine evaluated against the snapshot, then a sequential merge
 * into "upcoming_locations" and a swap (engine/src/allocation_map.rs:67-134).
 *
 * Deterministic mode (project-defined; the reference has no seeded mode):
 *   - agent ids are dense u32 in creation order (replaces Uuid::new_v4,
 *     engine/src/citizen/mod.rs:75);
 *   - every random decision reads Philox4x32-10 (Salmon et al., SC'11) with
 *     key = (seed_lo, seed_hi) and counter = (agent id, hour, site, block);
 *   - uniform integer on [0,n) = (u32 * n) >> 32;
 *   - Bernoulli(p) = rand 0.8's rule: p == 1 -> true, else u64 < floor(p * 2^64);
 *   - a cell claimed by several agents in one hour goes to the lowest id.
 *
 * Built with -DEPICPU_FNV_STORE (librefcpu.so) the same routine runs on the reference's
 * store instead: citizens held by value in two FNV hash maps keyed by cell, a parallel
 * map over the current map's buckets, a single-thread merge in bucket order (hash-order
 * collision winners) and draws from each agent-hour's sequential stream in the
 * reference's call order. That build is the timed CPU baseline and the stand-in for the
 * unmodified engine in distributional checks; it is not bit-comparable with
 * deterministic mode.
 */
#include "epicpu.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ Philox */
/* Philox4x32-10 as published in "Parallel Random Numbers: As Easy as 1, 2, 3"
   (Salmon, Moraes, Dror, Shaw, SC'11): multipliers and Weyl key increments are
   the paper's constants. */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void epicpu_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

#define INIT_HOUR 0xFFFFFFFFu
enum {
    SITE_MOVE = 1,      /* w0 x, w1 y of the hour's one random area point; w2 neighbour choice */
    SITE_DISEASE = 2,   /* w0 exposed random factor; (w2,w3) the hour's disease Bernoulli */
    SITE_INFECT = 3,    /* blocks 0..3: neighbour k uses block k>>1, words 2(k&1), 2(k&1)+1 */
    SITE_VACCINATE = 4, /* (w0,w1) */
    SITE_INIT_A = 16,   /* (w0,w1) working, (w2,w3) public transport */
    SITE_INIT_B = 17,   /* (w0,w1) hospital staff, w2 immunity */
    SITE_INIT_C = 18,   /* (w0,w1) essential worker, (w2,w3) starting-infection key */
    SITE_INIT_HOUSE = 19, /* counter id = house index: w0,w1 x permutation, w2,w3 y permutation */
    SITE_REF_STREAM = 32, /* reference-order mode: the agent-hour's sequential stream, block = draw index */
    SITE_REF_MERGE = 33   /* reference-order mode: counter id = hour; the merge permutation's stream */
};

static void draw(uint64_t seed, uint32_t id, uint32_t hour, uint32_t site, uint32_t block, uint32_t w[4]) {
    uint32_t ctr[4] = {id, hour, site, block};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    epicpu_philox4x32_10(ctr, key, w);
}

static inline uint64_t u64_of(uint32_t hi, uint32_t lo) { return ((uint64_t)hi << 32) | lo; }
static inline uint32_t uniform_u32(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

/* rand 0.8.5 Bernoulli::new: p_int = (p * 2^64) as u64, with p == 1 mapped to "always". */
#define ALWAYS_TRUE 0xFFFFFFFFFFFFFFFFull
uint64_t epicpu_bernoulli_threshold(double p) {
    if (p == 1.0) return ALWAYS_TRUE;
    return (uint64_t)(p * 18446744073709551616.0);
}
static inline int bernoulli(uint64_t thr, uint64_t v) { return thr == ALWAYS_TRUE || v < thr; }
static int valid_p(double p) { return p >= 0.0 && p <= 1.0; }

/* ------------------------------------------------------------------ geography */
typedef struct { int32_t x0, y0, x1, y1; } rect;
typedef struct { int32_t x, y; } point;

static inline int rect_contains(rect a, int32_t x, int32_t y) {
    return a.x0 <= x && a.x1 >= x && a.y0 <= y && a.y1 >= y; /* area.rs:83-88 */
}
static inline int rect_eq(rect a, rect b) { /* Area equality ignores location_id, area.rs:38-42 */
    return a.x0 == b.x0 && a.y0 == b.y0 && a.x1 == b.x1 && a.y1 == b.y1;
}
static const int32_t MOORE_DX[8] = {-1, 0, 1, -1, 1, -1, 0, 1}; /* point.rs:57-61 */
static const int32_t MOORE_DY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

static void define_geography(uint32_t g, rect* housing, rect* transport, rect* work, rect* hospital) {
    /* geography/mod.rs:33-48 */
    int32_t home_w = (int32_t)ceil((double)g * 0.4);
    int32_t t_end = home_w + (int32_t)ceil((double)g * 0.2);
    int32_t w_end = t_end + (int32_t)ceil((double)g * 0.2);
    int32_t h_end = w_end + (int32_t)ceil((double)g * 0.1);
    *housing = (rect){0, 0, home_w - 1, (int32_t)g};
    *transport = (rect){home_w, 0, t_end - 1, (int32_t)g};
    *work = (rect){t_end, 0, w_end - 1, (int32_t)g};
    *hospital = (rect){w_end, 0, h_end - 1, (int32_t)g};
}

void epicpu_define_geography(uint32_t grid_size, int32_t rects[16]) {
    rect r[4];
    define_geography(grid_size, &r[0], &r[1], &r[2], &r[3]);
    for (int i = 0; i < 4; i++) {
        rects[4 * i] = r[i].x0; rects[4 * i + 1] = r[i].y0;
        rects[4 * i + 2] = r[i].x1; rects[4 * i + 3] = r[i].y1;
    }
}

/* area.rs:95-117: buildings of size x size laid out row-major from the start point. */
typedef struct { int32_t sx, sy; int32_t nx, ny; int32_t size; } building_grid;
static building_grid make_buildings(rect a, int32_t size) {
    building_grid b;
    b.sx = a.x0; b.sy = a.y0; b.size = size;
    b.nx = (a.x1 - a.x0 + 1) / size;
    b.ny = (a.y1 - a.y0 + 1) / size;
    if (b.nx < 0) b.nx = 0;
    if (b.ny < 0) b.ny = 0;
    return b;
}
static inline uint32_t building_count(building_grid b) { return (uint32_t)b.nx * (uint32_t)b.ny; }
static inline rect building_at(building_grid b, uint32_t k) {
    int32_t x0 = b.sx + (int32_t)(k % (uint32_t)b.nx) * b.size;
    int32_t y0 = b.sy + (int32_t)(k / (uint32_t)b.nx) * b.size;
    return (rect){x0, y0, x0 + b.size - 1, y0 + b.size - 1};
}

uint32_t epicpu_area_factory(int32_t sx, int32_t sy, int32_t ex, int32_t ey, uint32_t size, int32_t* out, uint32_t cap) {
    building_grid b = make_buildings((rect){sx, sy, ex, ey}, (int32_t)size);
    uint32_t n = building_count(b);
    for (uint32_t k = 0; k < n && k < cap; k++) {
        rect r = building_at(b, k);
        out[4 * k] = r.x0; out[4 * k + 1] = r.y0; out[4 * k + 2] = r.x1; out[4 * k + 3] = r.y1;
    }
    return n;
}

/* grid.rs:240-261. Returns 0, or -1 on the reference's division by zero. */
static int resize_hospital(rect* h, int32_t n, double staff_pct, double beds_pct) {
    double beds_f = ceil((double)n * beds_pct + (double)n * staff_pct);
    uint32_t beds = beds_f <= 0.0 ? 0u : (beds_f >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)beds_f);
    uint32_t cells = (uint32_t)((h->x1 - h->x0) * (h->y1 - h->y0)); /* area.rs:90-92 */
    if (beds <= cells) {
        uint32_t w = (uint32_t)(h->x1 - h->x0);
        if (w == 0) return -1;
        h->y1 = (int32_t)(beds / w);
    }
    return 0;
}
void epicpu_resize_hospital(int32_t r[4], int32_t n, double staff_pct, double beds_pct) {
    rect h = {r[0], r[1], r[2], r[3]};
    resize_hospital(&h, n, staff_pct, beds_pct);
    r[0] = h.x0; r[1] = h.y0; r[2] = h.x1; r[3] = h.y1;
}
void epicpu_increase_hospital_size(int32_t r[4], uint32_t g) { /* grid.rs:233-238 */
    r[2] = (int32_t)g;
    r[3] = (int32_t)g;
}

uint32_t epicpu_neighbors_in_area(int32_t px, int32_t py, const int32_t a[4], int32_t* out) {
    rect ar = {a[0], a[1], a[2], a[3]};
    uint32_t n = 0;
    for (int k = 0; k < 8; k++) {
        int32_t x = px + MOORE_DX[k], y = py + MOORE_DY[k];
        if (rect_contains(ar, x, y)) { out[2 * n] = x; out[2 * n + 1] = y; n++; }
    }
    return n;
}

uint32_t epicpu_area_iter(const int32_t a[4], int32_t* out, uint32_t cap) { /* area.rs:119-146 */
    uint32_t n = 0;
    for (int32_t y = a[1]; y <= a[3]; y++)
        for (int32_t x = a[0]; x <= a[2]; x++) {
            if (n < cap) { out[2 * n] = x; out[2 * n + 1] = y; }
            n++;
        }
    return n;
}

/* ------------------------------------------------------------------ disease */
enum { ST_S = 0, ST_E = 1, ST_I = 2, ST_R = 3, ST_D = 4 };
enum { SEV_PRE = 0, SEV_ASYM = 1, SEV_MILD = 2, SEV_SEVERE = 3 };
enum { WS_NORMAL = 0, WS_ESSENTIAL = 1, WS_STAFF = 2, WS_NA = 3 };

/* common/src/disease/mod.rs:88-95 */
static double transmission_rate(const epicpu_config* c, uint32_t d) {
    if (c->regular_transmission_start_day < d && d <= c->high_transmission_start_day) return c->regular_transmission_rate;
    if (c->high_transmission_start_day < d && d <= c->last_day) return c->high_transmission_rate;
    return 0.0;
}
double epicpu_transmission_rate(const epicpu_config* c, uint32_t d) { return transmission_rate(c, d); }
/* 1 = regular window, 2 = high window, 0 = rate 0.0 */
static int transmission_window(const epicpu_config* c, uint32_t d) {
    if (c->regular_transmission_start_day < d && d <= c->high_transmission_start_day) return 1;
    if (c->high_transmission_start_day < d && d <= c->last_day) return 2;
    return 0;
}
int epicpu_is_to_be_hospitalized(const epicpu_config* c, uint32_t d) { /* disease/mod.rs:97-99 */
    return transmission_rate(c, d) >= c->high_transmission_rate;
}

/* ------------------------------------------------------------------ citizen */
typedef struct {
    uint32_t id;
    int32_t immunity;
    rect home, work, current_area;
    uint8_t vaccinated, uses_pt, hospitalized, isolated, work_quarantined;
    uint8_t work_status;
    uint32_t work_start_at;
    uint8_t state, severity;
    uint32_t at_hour, infection_day;
} citizen;

typedef struct {
    uint32_t hour;
    uint32_t kind;
} event;

#ifdef EPICPU_FNV_STORE
/* ------------------------------------------------------------------ refcpu store
   Compiled only into librefcpu.so, the reference-structure CPU baseline. The store is the
   reference's CitizenLocationMap (allocation_map.rs:44-65): two FnvHashMap<Point, Citizen>,
   citizens held by value in the table, swapped and cleared every hour. The table is
   open-addressed like hashbrown's: one control byte per bucket, a power-of-two bucket
   count sized for at most 7/8 of the buckets full, home bucket = hash & mask, iteration in
   bucket order. The hash is FNV-1a 64 (fnv 1.0.6) over the little-endian bytes of x, then
   y (Point's derived Hash, geography/point.rs:23-28). Probing is linear where hashbrown
   probes 16-bucket groups triangularly, so the iteration order differs from the Rust
   engine's; it is a hash order all the same. */
typedef struct { point key; citizen val; } fnv_entry;
typedef struct { uint8_t* ctrl; fnv_entry* e; uint64_t mask; } fnv_map;
typedef struct { point old_cell, new_cell; citizen c; } fnv_update;
#endif

struct epicpu_sim {
    epicpu_config cfg;
    uint32_t n, g, stride, height;
    rect housing, transport, work, hospital;
    building_grid houses, offices;
    uint32_t n_tpoints;
    citizen* ag;
    point* pos;
    uint32_t* current;  /* cell -> agent index + 1 (0 = vacant): the start-of-hour snapshot */
    uint32_t* upcoming;
    citizen* upd;
    point* tgt;
    uint64_t thr_regular, thr_high, thr_death, thr_symptoms, thr_severe;
    uint32_t counts[7];
    uint32_t hour;
    int stopped;
    int locked;
    uint32_t zero_infection_hour;
    uint32_t hosp_new_infections;
    int hosp_applied;
    event* events;
    uint32_t n_events, cap_events;
    uint32_t threads;
    int ref_order;  /* epicpu_set_reference_order */
    int random_merge; /* epicpu_set_random_merge */
    uint32_t* perm;
#ifdef EPICPU_FNV_STORE
    fnv_map cur, nxt;    /* current_locations, upcoming_locations */
    fnv_update* updates; /* the parallel map's output, in iteration order */
#endif
};

#ifdef EPICPU_FNV_STORE
static uint64_t fnv_point(point p) {
    const uint32_t v[2] = {(uint32_t)p.x, (uint32_t)p.y};
    uint64_t h = 0xcbf29ce484222325ull;
    for (int k = 0; k < 2; k++)
        for (int b = 0; b < 4; b++) {
            h ^= (v[k] >> (8 * b)) & 0xFFu;
            h *= 0x100000001b3ull;
        }
    return h;
}
static int fnv_alloc(fnv_map* m, uint32_t n) {
    uint64_t want = ((uint64_t)n * 8 + 6) / 7, cap = 16;
    while (cap < want) cap <<= 1;
    m->mask = cap - 1;
    m->ctrl = (uint8_t*)calloc(cap, 1);
    m->e = (fnv_entry*)malloc(cap * sizeof(fnv_entry));
    return m->ctrl && m->e ? 0 : -1;
}
static void fnv_free(fnv_map* m) { free(m->ctrl); free(m->e); m->ctrl = NULL; m->e = NULL; }
static inline const citizen* fnv_get(const fnv_map* m, point p) {
    for (uint64_t i = fnv_point(p) & m->mask;; i = (i + 1) & m->mask) {
        if (!m->ctrl[i]) return NULL;
        if (m->e[i].key.x == p.x && m->e[i].key.y == p.y) return &m->e[i].val;
    }
}
/* HashMap::entry(p).or_insert(c): the value now stored at p */
static const citizen* fnv_or_insert(fnv_map* m, point p, const citizen* c) {
    for (uint64_t i = fnv_point(p) & m->mask;; i = (i + 1) & m->mask) {
        if (!m->ctrl[i]) {
            m->ctrl[i] = 1;
            m->e[i].key = p;
            m->e[i].val = *c;
            return &m->e[i].val;
        }
        if (m->e[i].key.x == p.x && m->e[i].key.y == p.y) return &m->e[i].val;
    }
}
/* HashMap::insert(p, c): replaces the value at p */
static void fnv_insert(fnv_map* m, point p, const citizen* c) {
    for (uint64_t i = fnv_point(p) & m->mask;; i = (i + 1) & m->mask) {
        if (!m->ctrl[i] || (m->e[i].key.x == p.x && m->e[i].key.y == p.y)) {
            m->ctrl[i] = 1;
            m->e[i].key = p;
            m->e[i].val = *c;
            return;
        }
    }
}
/* HashMap::clear: every control byte, O(capacity) */
static void fnv_clear(fnv_map* m) { memset(m->ctrl, 0, m->mask + 1); }
#define FOR_EACH_CITIZEN(s, c)                                 \
    for (uint64_t b_ = 0; b_ <= (s)->cur.mask; b_++)           \
        if ((s)->cur.ctrl[b_])                                 \
            for (citizen* c = &(s)->cur.e[b_].val; c; c = NULL)
#else
#define FOR_EACH_CITIZEN(s, c) \
    for (uint32_t i_ = 0; i_ < (s)->n; i_++) for (citizen* c = &(s)->ag[i_]; c; c = NULL)
#endif

int epicpu_set_reference_order(epicpu_sim* s, int on) {
#ifdef EPICPU_FNV_STORE
    return on ? 0 : -2; /* refcpu always draws in the reference's order */
#endif
    if (on && !s->perm) {
        s->perm = (uint32_t*)malloc((size_t)(s->n ? s->n : 1) * sizeof(uint32_t));
        if (!s->perm) return -1;
    }
    s->ref_order = on ? 1 : 0;
    return 0;
}

int epicpu_set_random_merge(epicpu_sim* s, int on) {
#ifdef EPICPU_FNV_STORE
    return on ? -2 : 0; /* refcpu merges in bucket order */
#endif
    if (on && !s->perm) {
        s->perm = (uint32_t*)malloc((size_t)(s->n ? s->n : 1) * sizeof(uint32_t));
        if (!s->perm) return -1;
    }
    s->random_merge = on ? 1 : 0;
    return 0;
}

void epicpu_set_threads(epicpu_sim* s, uint32_t threads) {
    s->threads = threads < 1 ? 1 : (threads > 64 ? 64 : threads);
}

static inline uint32_t cell_index(const epicpu_sim* s, int32_t x, int32_t y) {
    return (uint32_t)y * s->stride + (uint32_t)x;
}
/* The snapshot's agent at a cell, or NULL (allocation_map.rs:152-154) */
static inline const citizen* agent_at(const epicpu_sim* s, point p) {
#ifdef EPICPU_FNV_STORE
    return fnv_get(&s->cur, p);
#else
    const uint32_t occ = s->current[cell_index(s, p.x, p.y)];
    return occ ? &s->ag[occ - 1] : NULL;
#endif
}
static inline int is_vacant(const epicpu_sim* s, point p) { /* allocation_map.rs:161-163 */
#ifdef EPICPU_FNV_STORE
    return fnv_get(&s->cur, p) == NULL;
#else
    return s->current[cell_index(s, p.x, p.y)] == 0;
#endif
}
static inline int in_grid(const epicpu_sim* s, int32_t x, int32_t y) { /* allocation_map.rs:156-159 */
    return x >= 0 && y >= 0 && x < (int32_t)s->g && y < (int32_t)s->g;
}
static inline int is_symptomatic(const citizen* c) {
    return c->state == ST_I && (c->severity == SEV_MILD || c->severity == SEV_SEVERE);
}
static inline int can_move(const citizen* c) { /* citizen/mod.rs:452-454 */
    return !(is_symptomatic(c) || c->hospitalized || c->state == ST_D || c->isolated);
}
static inline int is_working(const citizen* c) { return c->work_status != WS_NA; }
static inline uint32_t infection_day(const citizen* c) { return c->state == ST_I ? c->infection_day : 0; }
static inline uint32_t immunity_day(const citizen* c) { /* citizen/mod.rs:184: (day as i32 + immunity) as u32 */
    return (uint32_t)((int32_t)infection_day(c) + c->immunity);
}

/* Per-(agent, hour) random draws with fixed counters; in reference-order mode a single
   sequential stream consumed in the reference's call order instead. */
typedef struct {
    const epicpu_sim* s;
    uint32_t id, hour;
    int have_move, have_disease, infect_block;
    uint32_t mv[4], dz[4], inf[4];
    int ref;
    uint32_t site, sbuf[4], sblock, spos;
} draws;

static uint32_t next_u32(draws* d) {
    if (d->spos == 4) { draw(d->s->cfg.seed, d->id, d->hour, d->site, d->sblock++, d->sbuf); d->spos = 0; }
    return d->sbuf[d->spos++];
}
/* rand's gen::<u64>() from a 32-bit source: low word first (rand_core next_u64_via_u32). */
static uint64_t next_u64(draws* d) {
    uint32_t lo = next_u32(d);
    return u64_of(next_u32(d), lo);
}
/* rand 0.8.5 UniformInt<u32>::sample_single over [0, n), n >= 1: widening multiply with
   rejection zone (n << leading_zeros(n)) - 1. */
static uint32_t gen_range(draws* d, uint32_t n) {
    const uint32_t zone = (n << __builtin_clz(n)) - 1u;
    for (;;) {
        uint64_t m = (uint64_t)next_u32(d) * n;
        if ((uint32_t)m <= zone) return (uint32_t)(m >> 32);
    }
}

static const uint32_t* move_words(draws* d) {
    if (!d->have_move) { draw(d->s->cfg.seed, d->id, d->hour, SITE_MOVE, 0, d->mv); d->have_move = 1; }
    return d->mv;
}
static const uint32_t* disease_words(draws* d) {
    if (!d->have_disease) { draw(d->s->cfg.seed, d->id, d->hour, SITE_DISEASE, 0, d->dz); d->have_disease = 1; }
    return d->dz;
}
static uint64_t infect_u64(draws* d, int k) {
    int b = k >> 1;
    if (d->infect_block != b) { draw(d->s->cfg.seed, d->id, d->hour, SITE_INFECT, (uint32_t)b, d->inf); d->infect_block = b; }
    int w = 2 * (k & 1);
    return u64_of(d->inf[w], d->inf[w + 1]);
}

/* area.rs:76-81: x then y, both inclusive. */
static point random_point(draws* d, rect a) {
    point p;
    if (d->ref) {
        p.x = a.x0 + (int32_t)gen_range(d, (uint32_t)(a.x1 - a.x0 + 1));
        p.y = a.y0 + (int32_t)gen_range(d, (uint32_t)(a.y1 - a.y0 + 1));
        return p;
    }
    const uint32_t* w = move_words(d);
    p.x = a.x0 + (int32_t)uniform_u32(w[0], (uint32_t)(a.x1 - a.x0 + 1));
    p.y = a.y0 + (int32_t)uniform_u32(w[1], (uint32_t)(a.y1 - a.y0 + 1));
    return p;
}

static point move_agent(const epicpu_sim* s, point old_cell, point new_cell) { /* allocation_map.rs:136-142 */
    return is_vacant(s, new_cell) ? new_cell : old_cell;
}

/* citizen/mod.rs:415-432 */
static point move_agent_from(const epicpu_sim* s, const citizen* c, point cell, draws* d) {
    if (!can_move(c)) return cell;
    point loc = cell;
    if (!rect_contains(c->current_area, cell.x, cell.y)) loc = random_point(d, c->current_area);
    point cand[8];
    uint32_t n = 0;
    for (int k = 0; k < 8; k++) {
        point p = {loc.x + MOORE_DX[k], loc.y + MOORE_DY[k]};
        if (rect_contains(c->current_area, p.x, p.y) && in_grid(s, p.x, p.y) && is_vacant(s, p)) cand[n++] = p;
    }
    point nc = cell;
    if (d->ref) {
        /* IteratorRandom::choose on a Filter (size hint lower bound 0): reservoir sampling,
           element k replaces the choice when gen_range(0..k) == 0 */
        for (uint32_t k = 0; k < n; k++)
            if (gen_range(d, k + 1) == 0) nc = cand[k];
    } else if (n > 0) {
        nc = cand[uniform_u32(move_words(d)[2], n)];
    }
    return move_agent(s, cell, nc);
}

/* citizen/mod.rs:367-395 */
static point goto_area(const epicpu_sim* s, const citizen* c, rect target, point cell, draws* d) {
    int override_movement = 0;
    if (c->work_status == WS_NORMAL || c->work_status == WS_ESSENTIAL) {
        if (rect_contains(c->work, cell.x, cell.y) && rect_eq(target, c->home) &&
            c->state == ST_I && (c->severity == SEV_MILD || c->severity == SEV_SEVERE))
            override_movement = 1;
    }
    if (!can_move(c) && !override_movement) return cell;
    if (is_working(c)) {
        point p = random_point(d, target);
        if (!is_vacant(s, p)) p = cell;
        return move_agent(s, cell, p);
    }
    return move_agent_from(s, c, cell, d);
}

/* default_disease_handler.rs:64-86, with the short-circuit `find` replaced by an
   OR over fixed per-neighbour draws (same event probability). */
static int on_susceptible(const epicpu_sim* s, const citizen* c, point cell, draws* d) {
    if (c->work_quarantined || c->vaccinated) return 0;
    for (int k = 0; k < 8; k++) {
        int32_t x = cell.x + MOORE_DX[k], y = cell.y + MOORE_DY[k];
        if (!rect_contains(c->current_area, x, y) || !in_grid(s, x, y)) continue;
        const citizen* nb = agent_at(s, (point){x, y}); /* snapshot state of the neighbour */
        if (!nb) continue;
        if (nb->state != ST_I || nb->hospitalized) continue;
        /* rate(occupant) = get_current_transmission_rate(day + immunity), citizen/mod.rs:182-185 */
        int win = transmission_window(&s->cfg, immunity_day(nb));
        if (d->ref) { /* short-circuit find: one gen_bool per infectious neighbour, rate 0 included */
            uint64_t thr = win == 0 ? 0 : (win == 1 ? s->thr_regular : s->thr_high);
            if (bernoulli(thr, next_u64(d))) return 1;
            continue;
        }
        if (win == 0) continue; /* gen_bool(0.0) never fires */
        uint64_t thr = win == 1 ? s->thr_regular : s->thr_high;
        if (bernoulli(thr, infect_u64(d, k))) return 1;
    }
    return 0;
}

/* disease_state_machine.rs:53-70 with default_disease_handler.rs:41-62 */
static void update_infection_dynamics(const epicpu_sim* s, citizen* c, point cell, uint32_t h, draws* d) {
    const epicpu_config* cfg = &s->cfg;
    if (c->state == ST_S) {
        if (on_susceptible(s, c, cell, d)) { c->state = ST_E; c->at_hour = h; }
    } else if (c->state == ST_E) {
        const uint32_t* w = d->ref ? NULL : disease_words(d);
        /* RANGE_FOR_EXPOSED, constants.rs:43 */
        int32_t rf = (int32_t)(d->ref ? gen_range(d, 3) : uniform_u32(w[0], 3)) - 1;
        if (h - c->at_hour >= (uint32_t)((int32_t)cfg->exposed_duration + rf)) {
            int symptoms = bernoulli(s->thr_symptoms, d->ref ? next_u64(d) : u64_of(w[2], w[3]));
            c->state = ST_I;
            c->infection_day = 0;
            if (symptoms) { c->severity = SEV_PRE; c->at_hour = h; }
            else { c->severity = SEV_ASYM; c->at_hour = 0; }
        }
    } else if (c->state == ST_I) {
        if (c->severity == SEV_PRE && h - c->at_hour >= cfg->pre_symptomatic_duration) {
            const uint64_t v = d->ref ? next_u64(d) : u64_of(disease_words(d)[2], disease_words(d)[3]);
            c->severity = bernoulli(s->thr_severe, v) ? SEV_SEVERE : SEV_MILD;
            c->at_hour = 0;
        }
    }
}

/* citizen/mod.rs:351-365, allocation_map.rs:144-150 */
static point hospitalize(const epicpu_sim* s, citizen* c, point cell, draws* d) {
    point nc = cell;
    if (!c->hospitalized && c->state == ST_I && c->severity == SEV_SEVERE &&
        epicpu_is_to_be_hospitalized(&s->cfg, immunity_day(c))) {
        rect h = s->hospital;
        for (int32_t y = h.y0; y <= h.y1; y++)
            for (int32_t x = h.x0; x <= h.x1; x++) {
                point p = {x, y};
                if (is_vacant(s, p)) {
                    c->hospitalized = 1;
                    return move_agent(s, cell, p);
                }
            }
        nc = move_agent(s, cell, random_point(d, c->home));
        c->hospitalized = 0;
    }
    return nc;
}

/* citizen/mod.rs:397-413 with default_disease_handler.rs:88-102 */
static point deceased(const epicpu_sim* s, citizen* c, point cell, draws* d) {
    point nc = cell;
    if (c->state == ST_I) {
        if (c->severity == SEV_ASYM && c->infection_day == 9) c->state = ST_R;       /* ASYMPTOMATIC_LAST_DAY */
        else if (c->severity == SEV_MILD && c->infection_day == 12) c->state = ST_R; /* MILD_INFECTED_LAST_DAY */
        else if (c->severity == SEV_SEVERE && c->infection_day == s->cfg.last_day) {
            const uint64_t v = d->ref ? next_u64(d) : u64_of(disease_words(d)[2], disease_words(d)[3]);
            c->state = bernoulli(s->thr_death, v) ? ST_D : ST_R;
        }
        if (c->state != ST_I) { c->severity = 0; c->infection_day = 0; c->at_hour = 0; }
    }
    if (c->state == ST_R) nc = move_agent(s, cell, random_point(d, c->home));
    if ((c->state == ST_R || c->state == ST_D) && c->hospitalized) c->hospitalized = 0;
    return nc;
}

/* citizen/mod.rs:257-349 */
static point perform_movements(const epicpu_sim* s, citizen* c, point cell, uint32_t hd, uint32_t h, draws* d) {
    point nc = cell;
    if (c->work_status == WS_NORMAL || c->work_status == WS_ESSENTIAL) {
        if (hd == 7 || hd == 17) {
            if (c->uses_pt) { nc = goto_area(s, c, s->transport, cell, d); c->current_area = s->transport; }
            else nc = move_agent_from(s, c, cell, d);
        } else if (hd == 8) {
            nc = goto_area(s, c, c->work, cell, d);
            c->current_area = c->work;
        } else if (hd == 16) {
            nc = goto_area(s, c, c->home, cell, d);
            c->current_area = c->home;
        } else {
            nc = move_agent_from(s, c, cell, d);
        }
        update_infection_dynamics(s, c, nc, h, d);
    } else if (c->work_status == WS_STAFF) {
        uint32_t ws = c->work_start_at;
        uint32_t since = h > ws ? h - ws : 0; /* saturating_sub */
        if (since == 24u * 14u) { c->work_quarantined = 1; return nc; }
        if (since == 24u * 14u * 2u) {
            nc = goto_area(s, c, c->home, cell, d);
            c->current_area = c->home;
            c->work_start_at = h + 24u * 14u;
            return nc;
        }
        if (hd == 8) {
            if (!rect_eq(c->current_area, s->hospital) && ws <= h) {
                nc = goto_area(s, c, s->hospital, cell, d);
                c->current_area = s->hospital;
                c->work_start_at = h;
            }
            c->work_quarantined = 0;
        } else if (hd == 16) {
            c->work_quarantined = 1;
        } else {
            if (!c->work_quarantined && can_move(c)) nc = move_agent_from(s, c, cell, d);
        }
        update_infection_dynamics(s, c, nc, h, d);
    } else { /* NA */
        if (hd == 8) {
            nc = goto_area(s, c, s->housing, cell, d);
            c->current_area = s->housing;
        } else if (hd == 12) {
            nc = goto_area(s, c, c->home, cell, d);
            c->current_area = c->home;
        } else {
            nc = move_agent_from(s, c, cell, d);
        }
        update_infection_dynamics(s, c, nc, h, d);
    }
    return nc;
}

/* citizen/mod.rs:227-255 */
static point routine(const epicpu_sim* s, citizen* c, point cell, uint32_t h) {
    draws d;
    memset(&d, 0, sizeof d);
    d.s = s; d.id = c->id; d.hour = h; d.infect_block = -1;
    d.ref = s->ref_order; d.site = SITE_REF_STREAM; d.spos = 4;
    point nc = cell;
    uint32_t hd = h % 24;
    if (hd == 0) {
        if (c->state == ST_I) c->infection_day += 1; /* state.rs:69-73 */
        nc = hospitalize(s, c, cell, &d);
    } else if (hd >= 1 && hd <= 6) {
        if (c->work_status != WS_STAFF) c->current_area = c->home;
    } else if (hd == 23) {
        nc = deceased(s, c, cell, &d);
    } else {
        nc = perform_movements(s, c, cell, hd, h, &d);
    }
    return nc;
}

/* ------------------------------------------------------------------ interventions */
int epicpu_lockdown_should_apply(int locked, uint32_t at_number, uint32_t hour, uint32_t infected) {
    return !locked && hour % 24 == 0 && infected > at_number; /* lockdown.rs:55-61 */
}
int epicpu_lockdown_should_unlock(int locked, uint32_t zero_hr, uint32_t hour) {
    /* lockdown.rs:69-73: (14 * 1.5).round() days */
    return locked && hour == zero_hr + 21u * 24u;
}
uint32_t epicpu_hospital_counts_updated(uint32_t n, uint32_t hour, uint32_t infected) {
    if (hour % 24 == 0) return infected > n ? infected - n : 0; /* hospital.rs:68-74 */
    return n;
}
int epicpu_hospital_should_apply(int applied, uint32_t has, uint32_t thr, uint32_t n, uint32_t hour) {
    if (applied) return 0;
    return hour % 24 == 0 && has && n >= thr; /* hospital.rs:51-58 */
}

static void push_event(epicpu_sim* s, uint32_t hour, uint32_t kind) {
    if (s->n_events == s->cap_events) {
        s->cap_events = s->cap_events ? 2 * s->cap_events : 64;
        s->events = (event*)realloc(s->events, s->cap_events * sizeof(event));
    }
    s->events[s->n_events++] = (event){hour, kind};
}

/* allocation_map.rs:306-387 */
static void process_interventions(epicpu_sim* s) {
    uint32_t h = s->counts[0];
    const epicpu_config* cfg = &s->cfg;
    /* VaccinateIntervention keeps a HashMap<hour, percent>: the last entry for an hour wins. */
    int vi = -1;
    for (uint32_t k = 0; k < cfg->n_vaccinate; k++)
        if (cfg->vaccinate_at_hour[k] == h) vi = (int)k;
    if (vi >= 0) {
        uint64_t thr = epicpu_bernoulli_threshold(cfg->vaccinate_percent[vi]);
        FOR_EACH_CITIZEN(s, c) {
            if (c->state != ST_S) continue;
            uint32_t w[4];
            draw(cfg->seed, c->id, h, SITE_VACCINATE, 0, w);
            if (bernoulli(thr, u64_of(w[0], w[1]))) c->vaccinated = 1;
        }
        push_event(s, h, 0);
    }
    if (cfg->has_lockdown && epicpu_lockdown_should_apply(s->locked, cfg->lockdown_at_number_of_infections, h, s->counts[3])) {
        s->locked = 1;
        FOR_EACH_CITIZEN(s, c)
            if (c->work_status != WS_ESSENTIAL) c->isolated = 1;
        push_event(s, h, 1);
    }
    if (epicpu_lockdown_should_unlock(s->locked, s->zero_infection_hour, h)) {
        FOR_EACH_CITIZEN(s, c)
            if (c->isolated) c->isolated = 0;
        s->locked = 0;
        s->zero_infection_hour = 0;
        push_event(s, h, 2);
    }
    s->hosp_new_infections = epicpu_hospital_counts_updated(s->hosp_new_infections, h, s->counts[3]);
    if (epicpu_hospital_should_apply(s->hosp_applied, cfg->has_build_new_hospital, cfg->spread_rate_threshold,
                                     s->hosp_new_infections, h)) {
        s->hospital.x1 = (int32_t)s->g; /* grid.rs:233-238 */
        s->hospital.y1 = (int32_t)s->g;
        s->hosp_applied = 1;
        push_event(s, h, 3);
    }
}

/* ------------------------------------------------------------------ one hour */
static void update_counts(uint32_t counts[7], const citizen* c) { /* counts.rs:126-140 */
    switch (c->state) {
    case ST_S: counts[1]++; break;
    case ST_E: counts[2]++; break;
    case ST_I: if (c->hospitalized) counts[4]++; else counts[3]++; break;
    case ST_R: counts[5]++; break;
    default: counts[6]++; break;
    }
}

#ifdef EPICPU_FNV_STORE
/* refcpu's hour (allocation_map.rs:82-134): rayon's par_iter over the current map splits
   its buckets into contiguous ranges and `collect` keeps iteration order, so thread k
   counts the agents of its bucket range, then writes their updates at its offset. */
typedef struct {
    epicpu_sim* s;
    uint32_t h;
    uint64_t begin, end;
    uint32_t count, offset;
} fnv_job;

static void* fnv_count_range(void* arg) {
    fnv_job* j = (fnv_job*)arg;
    uint32_t n = 0;
    for (uint64_t b = j->begin; b < j->end; b++) n += j->s->cur.ctrl[b];
    j->count = n;
    return NULL;
}

static void* fnv_map_range(void* arg) {
    fnv_job* j = (fnv_job*)arg;
    fnv_update* u = j->s->updates + j->offset;
    for (uint64_t b = j->begin; b < j->end; b++) {
        if (!j->s->cur.ctrl[b]) continue;
        const fnv_entry* e = &j->s->cur.e[b];
        u->old_cell = e->key;
        u->c = e->val;
        u->new_cell = routine(j->s, &u->c, e->key, j->h);
        u++;
    }
    return NULL;
}

static void fnv_run(fnv_job* jobs, uint32_t nt, void* (*fn)(void*)) {
    pthread_t th[64];
    for (uint32_t k = 1; k < nt; k++) pthread_create(&th[k], NULL, fn, &jobs[k]);
    fn(&jobs[0]);
    for (uint32_t k = 1; k < nt; k++) pthread_join(th[k], NULL);
}

static int simulate(epicpu_sim* s, uint32_t h) {
    for (int k = 1; k < 7; k++) s->counts[k] = 0;
    const uint64_t cap = s->cur.mask + 1;
    uint32_t nt = s->threads > 1 && s->n >= 4096 ? s->threads : 1;
    fnv_job jobs[64];
    const uint64_t chunk = (cap + nt - 1) / nt;
    for (uint32_t k = 0; k < nt; k++) {
        const uint64_t b = k * chunk < cap ? k * chunk : cap, e = b + chunk < cap ? b + chunk : cap;
        jobs[k] = (fnv_job){s, h, b, e, 0, 0};
    }
    fnv_run(jobs, nt, fnv_count_range);
    uint32_t off = 0;
    for (uint32_t k = 0; k < nt; k++) { jobs[k].offset = off; off += jobs[k].count; }
    if (off != s->n) return -2;
    fnv_run(jobs, nt, fnv_map_range);
    /* the sequential merge, in iteration order (allocation_map.rs:93-125) */
    for (uint32_t k = 0; k < s->n; k++) {
        const fnv_update* u = &s->updates[k];
        const citizen* at_new = fnv_or_insert(&s->nxt, u->new_cell, &u->c);
        if (at_new->id != u->c.id) fnv_insert(&s->nxt, u->old_cell, &u->c);
        update_counts(s->counts, &u->c);
    }
    /* swap (allocation_map.rs:131-134): the old current map is cleared and becomes upcoming */
    fnv_clear(&s->cur);
    fnv_map t = s->cur; s->cur = s->nxt; s->nxt = t;
    uint32_t total = 0;
    for (int k = 1; k < 7; k++) total += s->counts[k];
    return total == s->n ? 0 : -2;
}

/* CitizenLocationMap::new (allocation_map.rs:52-65): agents inserted in creation order.
   The dense arrays that population synthesis filled are released. */
static int fnv_build(epicpu_sim* s) {
    if (fnv_alloc(&s->cur, s->n) || fnv_alloc(&s->nxt, s->n)) return -1;
    s->updates = (fnv_update*)malloc((size_t)(s->n ? s->n : 1) * sizeof(fnv_update));
    if (!s->updates) return -1;
    for (uint32_t i = 0; i < s->n; i++) fnv_insert(&s->cur, s->pos[i], &s->ag[i]);
    free(s->ag); free(s->upd); free(s->pos); free(s->tgt); free(s->current); free(s->upcoming);
    s->ag = s->upd = NULL;
    s->pos = s->tgt = NULL;
    s->current = s->upcoming = NULL;
    s->ref_order = 1; /* draws from each agent-hour's sequential stream, in the reference's call order */
    return 0;
}
#else
static int fnv_build(epicpu_sim* s) { (void)s; return 0; }

/* The parallel map of allocation_map.rs:82-92: every agent's routine reads only the
   start-of-hour snapshot, so agents can be split over threads like rayon's par_iter. */
typedef struct {
    epicpu_sim* s;
    uint32_t h, begin, end;
} map_job;

static void* map_range(void* arg) {
    map_job* j = (map_job*)arg;
    for (uint32_t i = j->begin; i < j->end; i++) {
        j->s->upd[i] = j->s->ag[i];
        j->s->tgt[i] = routine(j->s, &j->s->upd[i], j->s->pos[i], j->h);
    }
    return NULL;
}

/* allocation_map.rs:67-134 */
static int simulate(epicpu_sim* s, uint32_t h) {
    for (int k = 1; k < 7; k++) s->counts[k] = 0;
    uint32_t nt = s->threads > 1 && s->n >= 4096 ? s->threads : 1;
    map_job jobs[64];
    pthread_t th[64];
    uint32_t chunk = (s->n + nt - 1) / nt;
    for (uint32_t k = 0; k < nt; k++) {
        uint32_t b = k * chunk, e = b + chunk < s->n ? b + chunk : s->n;
        jobs[k] = (map_job){s, h, b < s->n ? b : s->n, e};
    }
    for (uint32_t k = 1; k < nt; k++) pthread_create(&th[k], NULL, map_range, &jobs[k]);
    map_range(&jobs[0]);
    for (uint32_t k = 1; k < nt; k++) pthread_join(th[k], NULL);
    /* Sequential merge. Iterating in id order makes "first inserted wins" equal to
       "lowest agent id wins". Targets are vacant in the snapshot or the agent's own
       cell, so a loser's old cell is always free in upcoming. */
    if (s->ref_order || s->random_merge) { /* a uniformly random merge order (stands in for hash order) */
        draws d;
        memset(&d, 0, sizeof d);
        d.s = s; d.id = 0; d.hour = h; d.site = SITE_REF_MERGE; d.spos = 4;
        for (uint32_t i = 0; i < s->n; i++) s->perm[i] = i;
        for (uint32_t i = s->n; i > 1; i--) {
            uint32_t j = gen_range(&d, i), t = s->perm[i - 1];
            s->perm[i - 1] = s->perm[j];
            s->perm[j] = t;
        }
    }
    for (uint32_t k = 0; k < s->n; k++) {
        const uint32_t i = s->ref_order || s->random_merge ? s->perm[k] : k;
        point oc = s->pos[i], nc = s->tgt[i];
        uint32_t ci = cell_index(s, nc.x, nc.y);
        if (s->upcoming[ci] == 0) {
            s->upcoming[ci] = i + 1;
            s->pos[i] = nc;
        } else {
            uint32_t oi = cell_index(s, oc.x, oc.y);
            if (s->upcoming[oi] != 0) return -1; /* impossible by construction */
            s->upcoming[oi] = i + 1;
        }
        update_counts(s->counts, &s->upd[i]);
    }
    /* swap (allocation_map.rs:131-134) */
    memset(s->current, 0, (size_t)s->stride * s->height * sizeof(uint32_t));
    uint32_t* t = s->current; s->current = s->upcoming; s->upcoming = t;
    citizen* ta = s->ag; s->ag = s->upd; s->upd = ta;
    uint32_t total = 0;
    for (int k = 1; k < 7; k++) total += s->counts[k];
    return total == s->n ? 0 : -2;
}

#endif

int epicpu_step(epicpu_sim* s, uint32_t n_hours, uint32_t* rows, uint32_t* n_written, int* stopped) {
    uint32_t written = 0;
    while (written < n_hours && !s->stopped) {
        uint32_t h = s->hour + 1;
        if (h >= s->cfg.hours) { s->stopped = 1; break; } /* loop is 1..hours (epidemiology_simulation.rs:223) */
        if (s->n == 0) return -3; /* "No citizens!" */
        s->counts[0] = h;
        int rc = simulate(s, h);
        if (rc) return rc;
        s->hour = h;
        if (rows) memcpy(rows + 7 * written, s->counts, 7 * sizeof(uint32_t));
        written++;
        process_interventions(s);
        /* epidemiology_simulation.rs:564-575, standalone branch */
        if (s->counts[2] == 0 && s->counts[3] == 0 && s->counts[4] == 0) s->stopped = 1;
        if (h + 1 >= s->cfg.hours) s->stopped = 1;
    }
    if (n_written) *n_written = written;
    if (stopped) *stopped = s->stopped;
    return 0;
}

/* ------------------------------------------------------------------ init */
static void set_err(char* err, size_t len, const char* msg) {
    if (err && len) { snprintf(err, len, "%s", msg); }
}

typedef struct { uint64_t key; uint32_t id; } sel;
static int sel_cmp(const void* a, const void* b) {
    const sel* x = (const sel*)a; const sel* y = (const sel*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

/* rand's choose_multiple(range, m) restated as a partial Fisher-Yates over the range
   driven by fixed words; returns min(m, len) elements. */
static uint32_t choose_multiple_det(int32_t lo, int32_t hi, uint32_t m, const uint32_t* words, int32_t* out) {
    uint32_t len = (uint32_t)(hi - lo + 1);
    if (m > len) m = len;
    int32_t a[2];
    for (uint32_t k = 0; k < len && k < 2; k++) a[k] = lo + (int32_t)k;
    for (uint32_t t = 0; t < m; t++) {
        uint32_t r = t + uniform_u32(words[t], len - t);
        int32_t tmp = a[t]; a[t] = a[r]; a[r] = tmp;
    }
    for (uint32_t t = 0; t < m; t++) out[t] = a[t];
    return m;
}

epicpu_sim* epicpu_create(const epicpu_config* cfg, char* err, size_t errlen) {
    char msg[256];
    const double ps[] = {cfg->public_transport_percentage, cfg->working_percentage, cfg->regular_transmission_rate,
                         cfg->high_transmission_rate, cfg->death_rate, cfg->percentage_asymptomatic_population,
                         cfg->percentage_severe_infected_population, cfg->hospital_beds_percentage,
                         cfg->lockdown_essential_workers_population};
    for (unsigned k = 0; k < sizeof ps / sizeof ps[0]; k++)
        if (!valid_p(ps[k])) {
            snprintf(msg, sizeof msg, "p=%g is outside range [0.0, 1.0]", ps[k]);
            set_err(err, errlen, msg);
            return NULL;
        }
    for (uint32_t k = 0; k < cfg->n_vaccinate; k++)
        if (!valid_p(cfg->vaccinate_percent[k])) {
            snprintf(msg, sizeof msg, "p=%g is outside range [0.0, 1.0]", cfg->vaccinate_percent[k]);
            set_err(err, errlen, msg);
            return NULL;
        }
    if (cfg->n_vaccinate > EPICPU_MAX_VACC) { set_err(err, errlen, "too many Vaccinate entries"); return NULL; }
    if (cfg->grid_size == 0 || cfg->grid_size > 32000) { set_err(err, errlen, "grid_size out of supported range"); return NULL; }

    epicpu_sim* s = (epicpu_sim*)calloc(1, sizeof(epicpu_sim));
    s->cfg = *cfg;
    s->n = cfg->csv_n > 0 ? cfg->csv_n : cfg->number_of_agents;
    s->cfg.number_of_agents = s->n;
    s->g = cfg->grid_size;
    define_geography(s->g, &s->housing, &s->transport, &s->work, &s->hospital);
    s->houses = make_buildings(s->housing, 2);   /* HOME_SIZE */
    s->offices = make_buildings(s->work, 10);    /* OFFICE_SIZE */
    uint32_t H = building_count(s->houses), O = building_count(s->offices);
    if (s->n > 0 && (H == 0 || O == 0)) {
        set_err(err, errlen, "attempt to calculate the remainder with a divisor of zero");
        free(s);
        return NULL;
    }
    int32_t maxx = (int32_t)s->g;
    if (s->hospital.x1 > maxx) maxx = s->hospital.x1;
    s->stride = (uint32_t)maxx + 1;
    s->height = s->g + 1;

    /* transport points: grid.rs:96-100, area.rs:64-74 (only their count is used by the step) */
    double n_pt = (double)s->n * (cfg->public_transport_percentage + 0.1) * (cfg->working_percentage + 0.1);
    uint64_t n_req = (uint64_t)ceil(n_pt);
    uint64_t nx = (uint64_t)ceilf(sqrtf((float)n_req));
    uint64_t lx = (uint64_t)(s->transport.x1 - s->transport.x0 + 1), ly = (uint64_t)(s->transport.y1 - s->transport.y0 + 1);
    uint64_t cx = nx < lx ? nx : lx, cy = nx < ly ? nx : ly;
    uint64_t np = cx * cy;
    s->n_tpoints = (uint32_t)(n_req < np ? n_req : np);

    uint32_t total_si = cfg->si_exposed + cfg->si_infected_mild_asymptomatic + cfg->si_infected_mild_symptomatic +
                        cfg->si_infected_severe;
    if (total_si > s->n) {
        snprintf(msg, sizeof msg, "There are %u people set to infect, but only %u agents available", total_si, s->n);
        set_err(err, errlen, msg);
        free(s);
        return NULL;
    }
    if (cfg->csv_n > 0 && s->n > 4u * H) { /* grid.rs:212-219 */
        snprintf(msg, sizeof msg, "Cannot accommodate citizens into homes! There are %u citizens, but %u home points",
                 s->n, 4u * H);
        set_err(err, errlen, msg);
        free(s);
        return NULL;
    }
    if (H > 0 && s->n > 4u * H) {
        uint32_t k = (s->n + H - 1) / H;
        snprintf(msg, sizeof msg, "There are %u agents assigned to a house, but house capacity is 4", k);
        set_err(err, errlen, msg);
        free(s);
        return NULL;
    }

    s->thr_regular = epicpu_bernoulli_threshold(cfg->regular_transmission_rate);
    s->thr_high = epicpu_bernoulli_threshold(cfg->high_transmission_rate);
    s->thr_death = epicpu_bernoulli_threshold(cfg->death_rate);
    s->thr_symptoms = epicpu_bernoulli_threshold(1.0 - cfg->percentage_asymptomatic_population);
    s->thr_severe = epicpu_bernoulli_threshold(cfg->percentage_severe_infected_population);

    size_t cells = (size_t)s->stride * s->height;
    s->ag = (citizen*)calloc(s->n ? s->n : 1, sizeof(citizen));
    s->upd = (citizen*)calloc(s->n ? s->n : 1, sizeof(citizen));
    s->pos = (point*)calloc(s->n ? s->n : 1, sizeof(point));
    s->tgt = (point*)calloc(s->n ? s->n : 1, sizeof(point));
    s->current = (uint32_t*)calloc(cells, sizeof(uint32_t));
    s->upcoming = (uint32_t*)calloc(cells, sizeof(uint32_t));

    /* citizen_factory.rs:31-88 */
    uint64_t thr_work = epicpu_bernoulli_threshold(cfg->working_percentage);
    uint64_t thr_pt = epicpu_bernoulli_threshold(cfg->public_transport_percentage);
    uint64_t thr_staff = epicpu_bernoulli_threshold(0.002); /* HOSPITAL_STAFF_PERCENTAGE */
    uint32_t pt_users = 0;
    for (uint32_t i = 0; i < s->n; i++) {
        uint32_t a[4], b[4];
        draw(cfg->seed, i, INIT_HOUR, SITE_INIT_A, 0, a);
        draw(cfg->seed, i, INIT_HOUR, SITE_INIT_B, 0, b);
        citizen* c = &s->ag[i];
        c->id = i;
        rect home = building_at(s->houses, i % H);
        rect office = building_at(s->offices, i % O);
        int working, pt;
        if (cfg->csv_n > 0) { /* Citizen::from_record (citizen/mod.rs:155-180): flags from the record */
            working = cfg->csv_working[i] != 0;
            pt = cfg->csv_pub_transport[i] != 0;
        } else {
            working = bernoulli(thr_work, u64_of(a[0], a[1]));
            pt = bernoulli(thr_pt, u64_of(a[2], a[3])) && working && pt_users < s->n_tpoints;
            if (pt) pt_users++;
        }
        c->home = home;
        c->work = working ? office : home;
        c->uses_pt = (uint8_t)pt;
        if (working) {
            if (bernoulli(thr_staff, u64_of(b[0], b[1]))) { c->work_status = WS_STAFF; c->work_start_at = 8; }
            else c->work_status = WS_NORMAL;
        } else {
            c->work_status = WS_NA;
        }
        c->immunity = (int32_t)uniform_u32(b[2], 5) - 2; /* IMMUNITY_RANGE */
        c->state = ST_S;
        c->current_area = home;
    }
    /* set_starting_infections (citizen_factory.rs:112-134): the `total` agents with the
       smallest Philox keys, in key order, take exposed / asymptomatic / mild / severe. */
    if (total_si > 0) {
        sel* keys = (sel*)malloc((size_t)s->n * sizeof(sel));
        for (uint32_t i = 0; i < s->n; i++) {
            uint32_t c3[4];
            draw(cfg->seed, i, INIT_HOUR, SITE_INIT_C, 0, c3);
            keys[i].key = u64_of(c3[2], c3[3]);
            keys[i].id = i;
        }
        qsort(keys, s->n, sizeof(sel), sel_cmp);
        uint32_t k = 0;
        for (uint32_t j = 0; j < cfg->si_exposed; j++, k++) { citizen* c = &s->ag[keys[k].id]; c->state = ST_E; c->at_hour = 0; }
        for (uint32_t j = 0; j < cfg->si_infected_mild_asymptomatic; j++, k++) {
            citizen* c = &s->ag[keys[k].id]; c->state = ST_I; c->severity = SEV_ASYM; c->infection_day = 1;
        }
        for (uint32_t j = 0; j < cfg->si_infected_mild_symptomatic; j++, k++) {
            citizen* c = &s->ag[keys[k].id]; c->state = ST_I; c->severity = SEV_MILD; c->infection_day = 1;
        }
        for (uint32_t j = 0; j < cfg->si_infected_severe; j++, k++) {
            citizen* c = &s->ag[keys[k].id]; c->state = ST_I; c->severity = SEV_SEVERE; c->infection_day = 1;
        }
        free(keys);
    }
    /* set_start_locations_and_occupancies (grid.rs:125-155): the agents of house h are
       h, h+H, h+2H, ...; they take random_points(k) of the house in id order. */
    for (uint32_t hidx = 0; hidx < H && hidx < s->n; hidx++) {
        uint32_t k = (s->n - hidx + H - 1) / H;
        rect hr = building_at(s->houses, hidx);
        uint32_t w[4];
        draw(cfg->seed, hidx, INIT_HOUR, SITE_INIT_HOUSE, 0, w);
        uint32_t nxp = (uint32_t)ceilf(sqrtf((float)k));
        int32_t xs[2], ys[2];
        uint32_t mx = choose_multiple_det(hr.x0, hr.x1, nxp, w, xs);
        uint32_t my = choose_multiple_det(hr.y0, hr.y1, nxp, w + 2, ys);
        uint32_t j = 0;
        for (uint32_t a = 0; a < mx && j < k; a++)
            for (uint32_t b = 0; b < my && j < k; b++, j++) {
                uint32_t id = hidx + j * H;
                s->pos[id] = (point){xs[a], ys[b]};
            }
    }
    for (uint32_t i = 0; i < s->n; i++) s->current[cell_index(s, s->pos[i].x, s->pos[i].y)] = i + 1;
    /* hospital: grid.rs:240-261 */
    if (resize_hospital(&s->hospital, (int32_t)s->n, 0.002, cfg->hospital_beds_percentage)) {
        set_err(err, errlen, "attempt to divide by zero");
        epicpu_destroy(s);
        return NULL;
    }
    /* init_interventions (epidemiology_simulation.rs:178-192, citizen/mod.rs:434-440) */
    double ess = cfg->has_lockdown ? cfg->lockdown_essential_workers_population : 0.0;
    uint64_t thr_ess = epicpu_bernoulli_threshold(ess);
    for (uint32_t i = 0; i < s->n; i++) {
        citizen* c = &s->ag[i];
        if (c->work_status != WS_NORMAL) continue;
        uint32_t w[4];
        draw(cfg->seed, i, INIT_HOUR, SITE_INIT_C, 0, w);
        if (bernoulli(thr_ess, u64_of(w[0], w[1]))) c->work_status = WS_ESSENTIAL;
    }
    /* counts_at_start (utils/util.rs:45-51) */
    s->counts[0] = 0;
    s->counts[1] = s->n - total_si;
    s->counts[2] = cfg->si_exposed;
    s->counts[3] = total_si - cfg->si_exposed;
    s->hour = 0;
    s->stopped = cfg->hours <= 1;
    if (fnv_build(s)) {
        set_err(err, errlen, "out of memory");
        epicpu_destroy(s);
        return NULL;
    }
    return s;
}

static rect rect_of(const int32_t r[4]) { return (rect){r[0], r[1], r[2], r[3]}; }

epicpu_sim* epicpu_create_scenario(const epicpu_config* cfg, const epicpu_agent_init* agents, uint32_t n,
                                   uint32_t last_hour, const int32_t* hospital, char* err, size_t errlen) {
    char msg[256];
    if (cfg->grid_size == 0 || cfg->grid_size > 32000) { set_err(err, errlen, "grid_size out of supported range"); return NULL; }
    epicpu_sim* s = (epicpu_sim*)calloc(1, sizeof(epicpu_sim));
    s->cfg = *cfg;
    s->cfg.csv_n = 0;
    s->n = n;
    s->cfg.number_of_agents = n;
    s->g = cfg->grid_size;
    define_geography(s->g, &s->housing, &s->transport, &s->work, &s->hospital);
    s->houses = make_buildings(s->housing, 2);
    s->offices = make_buildings(s->work, 10);
    if (hospital) s->hospital = rect_of(hospital);
    else if (resize_hospital(&s->hospital, (int32_t)n, 0.002, cfg->hospital_beds_percentage)) {
        set_err(err, errlen, "attempt to divide by zero");
        free(s);
        return NULL;
    }
    s->stride = s->g + 1;
    s->height = s->g + 1;
    s->thr_regular = epicpu_bernoulli_threshold(cfg->regular_transmission_rate);
    s->thr_high = epicpu_bernoulli_threshold(cfg->high_transmission_rate);
    s->thr_death = epicpu_bernoulli_threshold(cfg->death_rate);
    s->thr_symptoms = epicpu_bernoulli_threshold(1.0 - cfg->percentage_asymptomatic_population);
    s->thr_severe = epicpu_bernoulli_threshold(cfg->percentage_severe_infected_population);
    size_t cells = (size_t)s->stride * s->height;
    s->ag = (citizen*)calloc(n ? n : 1, sizeof(citizen));
    s->upd = (citizen*)calloc(n ? n : 1, sizeof(citizen));
    s->pos = (point*)calloc(n ? n : 1, sizeof(point));
    s->tgt = (point*)calloc(n ? n : 1, sizeof(point));
    s->current = (uint32_t*)calloc(cells, sizeof(uint32_t));
    s->upcoming = (uint32_t*)calloc(cells, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) {
        const epicpu_agent_init* a = &agents[i];
        if (a->x < 0 || a->y < 0 || (uint32_t)a->x >= s->stride || (uint32_t)a->y >= s->height ||
            s->current[cell_index(s, a->x, a->y)] != 0) {
            snprintf(msg, sizeof msg, "agent %u: cell (%d, %d) is outside the grid or taken", i, a->x, a->y);
            set_err(err, errlen, msg);
            epicpu_destroy(s);
            return NULL;
        }
        citizen* c = &s->ag[i];
        c->id = i;
        c->home = rect_of(a->home);
        c->work = rect_of(a->work);
        c->current_area = rect_of(a->current_area);
        c->state = (uint8_t)a->state;
        c->severity = (uint8_t)a->severity;
        c->work_status = (uint8_t)a->work_status;
        c->hospitalized = (a->flags & 1) != 0;
        c->vaccinated = (a->flags & 2) != 0;
        c->isolated = (a->flags & 4) != 0;
        c->work_quarantined = (a->flags & 8) != 0;
        c->uses_pt = (a->flags & 16) != 0;
        c->infection_day = a->infection_day;
        c->at_hour = a->at_hour;
        c->work_start_at = a->work_start_at;
        c->immunity = a->immunity;
        s->pos[i] = (point){a->x, a->y};
        s->current[cell_index(s, a->x, a->y)] = i + 1;
    }
    s->counts[0] = last_hour;
    for (uint32_t i = 0; i < n; i++) update_counts(s->counts, &s->ag[i]);
    s->hour = last_hour;
    s->stopped = last_hour + 1 >= cfg->hours;
    if (fnv_build(s)) {
        set_err(err, errlen, "out of memory");
        epicpu_destroy(s);
        return NULL;
    }
    return s;
}

void epicpu_destroy(epicpu_sim* s) {
    if (!s) return;
    free(s->ag); free(s->upd); free(s->pos); free(s->tgt);
    free(s->current); free(s->upcoming); free(s->events); free(s->perm);
#ifdef EPICPU_FNV_STORE
    fnv_free(&s->cur); fnv_free(&s->nxt); free(s->updates);
#endif
    free(s);
}

void epicpu_counts(const epicpu_sim* s, uint32_t out[7]) { memcpy(out, s->counts, 7 * sizeof(uint32_t)); }

uint32_t epicpu_events(const epicpu_sim* s, uint32_t* hours, uint32_t* kinds, uint32_t cap) {
    for (uint32_t i = 0; i < s->n_events && i < cap; i++) { hours[i] = s->events[i].hour; kinds[i] = s->events[i].kind; }
    return s->n_events;
}

void epicpu_geometry(const epicpu_sim* s, int32_t r[20], uint32_t c3[3]) {
    rect init_h = s->hospital;
    const rect rs[5] = {s->housing, s->transport, s->work, init_h, s->hospital};
    for (int i = 0; i < 5; i++) { r[4 * i] = rs[i].x0; r[4 * i + 1] = rs[i].y0; r[4 * i + 2] = rs[i].x1; r[4 * i + 3] = rs[i].y1; }
    c3[0] = building_count(s->houses);
    c3[1] = building_count(s->offices);
    c3[2] = s->n_tpoints;
}

void epicpu_read_agents(const epicpu_sim* s, epicpu_agents* o) {
#ifdef EPICPU_FNV_STORE
    for (uint64_t b = 0; b <= s->cur.mask; b++) {
        if (!s->cur.ctrl[b]) continue;
        const citizen* c = &s->cur.e[b].val;
        const uint32_t i = c->id;
        o->x[i] = s->cur.e[b].key.x;
        o->y[i] = s->cur.e[b].key.y;
#else
    for (uint32_t i = 0; i < s->n; i++) {
        const citizen* c = &s->ag[i];
        o->x[i] = s->pos[i].x;
        o->y[i] = s->pos[i].y;
#endif
        o->state[i] = c->state;
        o->severity[i] = c->state == ST_I ? c->severity : 0;
        o->infection_day[i] = c->state == ST_I ? c->infection_day : 0;
        o->at_hour[i] = (c->state == ST_E || (c->state == ST_I && c->severity == SEV_PRE)) ? c->at_hour : 0;
        o->flags[i] = (uint8_t)((c->hospitalized ? 1 : 0) | (c->vaccinated ? 2 : 0) | (c->isolated ? 4 : 0) |
                                (c->work_quarantined ? 8 : 0) | (c->uses_pt ? 16 : 0));
        o->work_status[i] = c->work_status;
        o->work_start_at[i] = c->work_status == WS_STAFF ? c->work_start_at : 0;
        o->immunity[i] = (int8_t)c->immunity;
        o->current_area[4 * i] = c->current_area.x0;
        o->current_area[4 * i + 1] = c->current_area.y0;
        o->current_area[4 * i + 2] = c->current_area.x1;
        o->current_area[4 * i + 3] = c->current_area.y1;
    }
}

int epicpu_first_vacant(const int32_t a[4], const int32_t* occ, uint32_t n_occ, int32_t out[2]) {
    for (int32_t y = a[1]; y <= a[3]; y++)
        for (int32_t x = a[0]; x <= a[2]; x++) {
            int taken = 0;
            for (uint32_t k = 0; k < n_occ; k++)
                if (occ[2 * k] == x && occ[2 * k + 1] == y) { taken = 1; break; }
            if (!taken) { out[0] = x; out[1] = y; return 1; }
        }
    return 0;
}

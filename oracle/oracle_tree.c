/*
 * oracle_tree.c -- CPU restatement of the reference's tree-culled
 * verification stage.  TEST INFRASTRUCTURE / CPU BASELINE ONLY: see
 * oracle.h.
 *
 * It follows rknn-inf's dfs_rknn_traverse / get_rknn_candidates
 * (rknn-inf/inf.h:194-249): users sit in an R-tree; a child node is pruned
 * when at least k candidate half-planes contain its whole rectangle on the
 * candidate's side (check_rectangle_line_position and
 * can_prune_node_by_valid_side, inf.h:104-147: all four corners on the
 * invalid side); the users of the leaves that remain are counted one by one
 * with early exit at k.  Two differences, both for exactness against the
 * binary64 oracle:
 *   - the rectangle test is the distance form m(c) = (c-q).(f-q) - |f-q|^2/2
 *     at the four corners with a margin covering binary64 rounding of m and
 *     of the oracle's own D(u,.) (a pruned node never holds a user the
 *     oracle would keep; the reference's slope-intercept test has no margin);
 *   - leaf users are counted with the oracle predicate itself.
 * The tree is bulk-loaded (sort-tile-recursive, 16 entries per node) where
 * the reference inserts one user at a time (RStarTree2D.cpp:106-201); the
 * reference reports the build separately from its query time
 * (inf.cpp:163-179), and so does this.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#define FAN 16

typedef struct {
    double minx, miny, maxx, maxy;
    uint32_t first, count; /* children: nodes of the level below, or user slots */
} onode;

struct oracle_tree {
    uint64_t nu;
    float *x, *y;   /* users in leaf-slot order */
    int32_t* id;
    onode* lv[32];  /* lv[0] = leaves ... lv[nl-1] = top */
    uint64_t nn[32];
    int nl;
};

/* ---- sort-tile-recursive ordering ---- */
typedef struct {
    double cx, cy;
    uint32_t i;
} okey;

static int cmp_x(const void* a, const void* b)
{
    const okey *p = (const okey*)a, *q = (const okey*)b;
    return (p->cx > q->cx) - (p->cx < q->cx);
}
static int cmp_y(const void* a, const void* b)
{
    const okey *p = (const okey*)a, *q = (const okey*)b;
    return (p->cy > q->cy) - (p->cy < q->cy);
}

static void str_sort(okey* k, uint64_t n)
{
    qsort(k, n, sizeof(okey), cmp_x);
    const uint64_t groups = (n + FAN - 1) / FAN;
    uint64_t slabs = (uint64_t)ceil(sqrt((double)groups));
    if (slabs == 0) slabs = 1;
    const uint64_t slab = slabs * FAN;
    for (uint64_t s = 0; s < n; s += slab) qsort(k + s, (n - s < slab ? n - s : slab), sizeof(okey), cmp_y);
}

oracle_tree* oracle_tree_build(const float* ux, const float* uy, const int32_t* uid, uint64_t nu)
{
    oracle_tree* t = (oracle_tree*)calloc(1, sizeof(oracle_tree));
    if (!t) return NULL;
    t->nu = nu;
    if (nu == 0) return t;
    okey* k = (okey*)malloc(nu * sizeof(okey));
    t->x = (float*)malloc(nu * sizeof(float));
    t->y = (float*)malloc(nu * sizeof(float));
    t->id = (int32_t*)malloc(nu * sizeof(int32_t));
    for (uint64_t i = 0; i < nu; ++i) {
        k[i].cx = ux[i];
        k[i].cy = uy[i];
        k[i].i = (uint32_t)i;
    }
    str_sort(k, nu);
    for (uint64_t i = 0; i < nu; ++i) {
        t->x[i] = ux[k[i].i];
        t->y[i] = uy[k[i].i];
        t->id[i] = uid[k[i].i];
    }
    free(k);
    /* leaves over user slots */
    uint64_t n = (nu + FAN - 1) / FAN;
    onode* cur = (onode*)malloc(n * sizeof(onode));
    for (uint64_t j = 0; j < n; ++j) {
        onode* d = &cur[j];
        d->first = (uint32_t)(j * FAN);
        d->count = (uint32_t)(nu - j * FAN < FAN ? nu - j * FAN : FAN);
        d->minx = d->maxx = t->x[d->first];
        d->miny = d->maxy = t->y[d->first];
        for (uint32_t s = d->first; s < d->first + d->count; ++s) {
            d->minx = fmin(d->minx, t->x[s]);
            d->maxx = fmax(d->maxx, t->x[s]);
            d->miny = fmin(d->miny, t->y[s]);
            d->maxy = fmax(d->maxy, t->y[s]);
        }
    }
    t->nl = 0;
    while (1) {
        if (n <= FAN || t->nl == 31) {
            t->lv[t->nl] = cur;
            t->nn[t->nl++] = n;
            break;
        }
        /* order this level by STR on node centres, then group consecutive nodes */
        okey* kk = (okey*)malloc(n * sizeof(okey));
        for (uint64_t j = 0; j < n; ++j) {
            kk[j].cx = 0.5 * (cur[j].minx + cur[j].maxx);
            kk[j].cy = 0.5 * (cur[j].miny + cur[j].maxy);
            kk[j].i = (uint32_t)j;
        }
        str_sort(kk, n);
        onode* sorted = (onode*)malloc(n * sizeof(onode));
        for (uint64_t j = 0; j < n; ++j) sorted[j] = cur[kk[j].i];
        free(kk);
        free(cur);
        t->lv[t->nl] = sorted;
        t->nn[t->nl++] = n;
        const uint64_t up = (n + FAN - 1) / FAN;
        cur = (onode*)malloc(up * sizeof(onode));
        for (uint64_t j = 0; j < up; ++j) {
            onode* d = &cur[j];
            d->first = (uint32_t)(j * FAN);
            d->count = (uint32_t)(n - j * FAN < FAN ? n - j * FAN : FAN);
            d->minx = sorted[d->first].minx;
            d->maxx = sorted[d->first].maxx;
            d->miny = sorted[d->first].miny;
            d->maxy = sorted[d->first].maxy;
            for (uint32_t s = d->first; s < d->first + d->count; ++s) {
                d->minx = fmin(d->minx, sorted[s].minx);
                d->maxx = fmax(d->maxx, sorted[s].maxx);
                d->miny = fmin(d->miny, sorted[s].miny);
                d->maxy = fmax(d->maxy, sorted[s].maxy);
            }
        }
        n = up;
    }
    return t;
}

void oracle_tree_free(oracle_tree* t)
{
    if (!t) return;
    for (int l = 0; l < t->nl; ++l) free(t->lv[l]);
    free(t->x);
    free(t->y);
    free(t->id);
    free(t);
}

/* ---- query ---- */
typedef struct {
    double qx, qy;
    uint32_t k, q;
    const float *fx, *fy;
    uint64_t nf;
    const uint32_t* cand; /* NULL: every facility but q */
    uint64_t nc;
    int32_t* out;
    uint64_t n_out;
} oquery;

static inline uint32_t cand_at(const oquery* Q, uint64_t i) { return Q->cand ? Q->cand[i] : (uint32_t)i; }
static inline uint64_t cand_n(const oquery* Q) { return Q->cand ? Q->nc : Q->nf; }

/* Does f certainly contain the whole rectangle on its side (every point of
 * it strictly closer to f than to q, by the oracle's binary64 decision)? */
static int covers(const oquery* Q, uint32_t f, const onode* d)
{
    const double gx = (double)Q->fx[f] - Q->qx, gy = (double)Q->fy[f] - Q->qy;
    const double h = 0.5 * (gx * gx + gy * gy);
    if (h == 0.0) return 0;
    const double cx[2] = {d->minx - Q->qx, d->maxx - Q->qx}, cy[2] = {d->miny - Q->qy, d->maxy - Q->qy};
    double dmax = 0.0, lmax = 0.0, mmin = INFINITY;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const double m = (cx[a] * gx + cy[b] * gy) - h;
            mmin = fmin(mmin, m);
            dmax = fmax(dmax, cx[a] * cx[a] + cy[b] * cy[b]);
            lmax = fmax(lmax, fabs(cx[a]) + fabs(cy[b]));
        }
    /* oracle's decision error 2.01v(3 Dq + 4H), evaluation error of m, both doubled */
    const double v = 0x1p-53;
    const double margin = 8.0 * v * (3.0 * dmax + 4.0 * h + lmax * (fabs(gx) + fabs(gy))) + 0x1p-1000;
    return mmin > margin;
}

static uint32_t count_leaf_user(const oquery* Q, float x, float y)
{
    const double dxq = (double)x - Q->qx, dyq = (double)y - Q->qy;
    const double dq = dxq * dxq + dyq * dyq;
    uint32_t cnt = 0;
    const uint64_t n = cand_n(Q);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t f = cand_at(Q, i);
        if (f == Q->q || f >= Q->nf) continue;
        const double dx = (double)x - (double)Q->fx[f], dy = (double)y - (double)Q->fy[f];
        if (dx * dx + dy * dy < dq && ++cnt >= Q->k) return cnt;
    }
    return cnt;
}

static void dfs(const oracle_tree* t, oquery* Q, int level, uint64_t node)
{
    const onode* d = &t->lv[level][node];
    uint32_t full = 0;
    const uint64_t n = cand_n(Q);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t f = cand_at(Q, i);
        if (f == Q->q || f >= Q->nf) continue;
#if ORACLE_MUTANT == 6
        if (covers(Q, f, d) && ++full + 1 >= Q->k) return; /* planted: prunes one candidate early */
#else
        if (covers(Q, f, d) && ++full >= Q->k) return; /* pruned */
#endif
    }
    if (level == 0) {
        for (uint32_t s = d->first; s < d->first + d->count; ++s)
            if (count_leaf_user(Q, t->x[s], t->y[s]) < Q->k) {
                if (Q->out) Q->out[Q->n_out] = t->id[s];
                ++Q->n_out;
            }
        return;
    }
    for (uint32_t c = d->first; c < d->first + d->count; ++c) dfs(t, Q, level - 1, c);
}

static int cmp_id(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

int64_t oracle_tree_rknn(const oracle_tree* t, const float* fx, const float* fy, uint64_t nf, uint32_t q,
                         uint32_t k, const uint32_t* cand, uint64_t nc, int32_t* out)
{
    if (!t || q >= nf || k == 0) return -1;
    oquery Q = {fx[q], fy[q], k, q, fx, fy, nf, cand, nc, out, 0};
    if (t->nu == 0) return 0;
    const int top = t->nl - 1;
    for (uint64_t j = 0; j < t->nn[top]; ++j) dfs(t, &Q, top, j);
    if (out) qsort(out, Q.n_out, sizeof(int32_t), cmp_id);
    return (int64_t)Q.n_out;
}

/* ---- many queries, one per thread at a time ---- */
typedef struct {
    const oracle_tree* t;
    const float *fx, *fy;
    uint64_t nf, nq;
    const uint32_t *q, *k;
    const uint64_t* off;
    const uint32_t* cands;
    int64_t* counts;
    uint64_t* next;
} many_job;

static void* many_worker(void* p)
{
    many_job* j = (many_job*)p;
    for (;;) {
        const uint64_t i = __atomic_fetch_add(j->next, 1, __ATOMIC_RELAXED);
        if (i >= j->nq) break;
        const uint32_t* c = j->cands ? j->cands + j->off[i] : NULL;
        const uint64_t nc = j->cands ? j->off[i + 1] - j->off[i] : 0;
        j->counts[i] = oracle_tree_rknn(j->t, j->fx, j->fy, j->nf, j->q[i], j->k[i], c, nc, NULL);
    }
    return NULL;
}

int oracle_tree_rknn_many(const oracle_tree* t, const float* fx, const float* fy, uint64_t nf, uint64_t nq,
                          const uint32_t* q, const uint32_t* k, const uint64_t* off, const uint32_t* cands,
                          int64_t* counts)
{
    if (!t || !q || !k || !counts || (cands && !off)) return -1;
    uint64_t next = 0;
    many_job j = {t, fx, fy, nf, nq, q, k, off, cands, counts, &next};
    int nt = oracle_get_threads();
    if ((uint64_t)nt > nq) nt = nq ? (int)nq : 1;
    if (nt <= 1) {
        many_worker(&j);
        return 0;
    }
    pthread_t th[256];
    for (int i = 0; i < nt; ++i) pthread_create(&th[i], NULL, many_worker, &j);
    for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
    return 0;
}

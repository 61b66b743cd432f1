/*
 * oracle.c -- FP64 brute-force RkNN oracle.  TEST INFRASTRUCTURE ONLY:
 * see oracle.h for what it restates and what pins it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).  The same source is
 * also built with -DORACLE_MUTANT=n into liboracle_mutN.so; each mutant
 * plants one plausible mistake so the pinning tests can show that they
 * catch it.  Mutants (never used outside tests/test_oracle.py):
 *   1  ties lost by q       (D(u,f) <= D(u,q) counts as closer)
 *   2  transposed operand   (dy taken from the x coordinates)
 *   3  off-by-one keep rule (count <= k keeps the user)
 *   4  dropped term         (dy*dy missing from D)
 *   5  wrong exclusion      (skips facility q+1 instead of q)
 *   6  early prune          (oracle_tree.c: a node is pruned at k-1 covering
 *                            candidates)
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

static int g_threads = 1;

void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : (n > 256 ? 256 : n); }
int oracle_get_threads(void) { return g_threads; }

/* Squared distance exactly as the oracle defines it (rknn-naive/rknn.cpp:45-47
 * computes the same differences; the oracle keeps the square in binary64). */
static inline double dist2(float ax, float ay, float bx, float by)
{
    double dx = (double)ax - (double)bx;
#if ORACLE_MUTANT == 2
    double dy = (double)ax - (double)by; /* transposed: x of the user against y of the point */
    (void)ay;
#else
    double dy = (double)ay - (double)by;
#endif
#if ORACLE_MUTANT == 4
    (void)dy;
    return dx * dx;
#else
    return dx * dx + dy * dy;
#endif
}

static inline int closer(double df, double dq)
{
#if ORACLE_MUTANT == 1
    return df <= dq;
#else
    return df < dq; /* strict: q wins ties, inf.h:77-93 / inf_kernel.cu:20-36 */
#endif
}

static inline uint32_t excluded_index(uint32_t q, uint64_t nf)
{
#if ORACLE_MUTANT == 5
    return (uint32_t)((q + 1) % nf);
#else
    (void)nf;
    return q;
#endif
}

/* Count of facilities strictly closer to (x,y) than f_q, stopping at cap
 * (the early exit of inf_kernel.cu:56-63 and rknn_kernel.cu:150-162). */
static uint32_t count_closer(float x, float y, const float* fx, const float* fy, uint64_t nf,
                             uint32_t q, const uint32_t* cand, uint64_t nc, uint32_t cap)
{
    const double dq = dist2(x, y, fx[q], fy[q]);
    const uint32_t skip = excluded_index(q, nf);
    uint32_t cnt = 0;
    if (cand == NULL) {
        for (uint64_t f = 0; f < nf; ++f) {
            if (f == skip) continue;
            if (closer(dist2(x, y, fx[f], fy[f]), dq)) {
                if (++cnt >= cap) return cnt;
            }
        }
    } else {
        for (uint64_t i = 0; i < nc; ++i) {
            uint32_t f = cand[i];
            if (f == skip || f >= nf) continue;
            if (closer(dist2(x, y, fx[f], fy[f]), dq)) {
                if (++cnt >= cap) return cnt;
            }
        }
    }
    return cnt;
}

typedef struct {
    const float *ux, *uy, *fx, *fy;
    uint64_t nf, nc, lo, hi;
    uint32_t q, cap;
    const uint32_t* cand;
    uint32_t* counts;
} count_job;

static void* count_worker(void* p)
{
    count_job* j = (count_job*)p;
    for (uint64_t u = j->lo; u < j->hi; ++u)
        j->counts[u] = count_closer(j->ux[u], j->uy[u], j->fx, j->fy, j->nf, j->q, j->cand, j->nc, j->cap);
    return NULL;
}

static void run_counts(const float* ux, const float* uy, uint64_t nu, const float* fx, const float* fy,
                       uint64_t nf, uint32_t q, const uint32_t* cand, uint64_t nc, uint32_t cap,
                       uint32_t* counts)
{
    int nt = g_threads;
    if ((uint64_t)nt > nu) nt = nu ? (int)nu : 1;
    count_job jobs[256];
    pthread_t th[256];
    uint64_t step = (nu + nt - 1) / (uint64_t)nt;
    for (int t = 0; t < nt; ++t) {
        count_job c = {ux, uy, fx, fy, nf, nc, t * step, (t + 1) * step, q, cap, cand, counts};
        if (c.lo > nu) c.lo = nu;
        if (c.hi > nu) c.hi = nu;
        jobs[t] = c;
    }
    if (nt == 1) {
        count_worker(&jobs[0]);
        return;
    }
    for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, count_worker, &jobs[t]);
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
}

static int cmp_i32(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

int64_t oracle_rknn(const float* ux, const float* uy, const int32_t* uid, uint64_t nu,
                    const float* fx, const float* fy, uint64_t nf,
                    uint32_t q, uint32_t k, const uint32_t* cand, uint64_t nc,
                    int32_t* out)
{
    if (k == 0 || nf == 0 || q >= nf) return -1;
    if (nu == 0) return 0;
    uint32_t* counts = (uint32_t*)malloc(nu * sizeof(uint32_t));
    if (!counts) return -1;
    run_counts(ux, uy, nu, fx, fy, nf, q, cand, nc, k, counts);
    int64_t n = 0;
    for (uint64_t u = 0; u < nu; ++u) {
#if ORACLE_MUTANT == 3
        if (counts[u] <= k) out[n++] = uid[u];
#else
        if (counts[u] < k) out[n++] = uid[u]; /* inf.h:213-216: violations < k */
#endif
    }
    free(counts);
    qsort(out, (size_t)n, sizeof(int32_t), cmp_i32);
    return n;
}

int oracle_counts(const float* ux, const float* uy, uint64_t nu,
                  const float* fx, const float* fy, uint64_t nf,
                  uint32_t q, const uint32_t* cand, uint64_t nc, uint32_t cap,
                  uint32_t* counts)
{
    if (nf == 0 || q >= nf) return -1;
    if (cap == 0) cap = UINT32_MAX;
    run_counts(ux, uy, nu, fx, fy, nf, q, cand, nc, cap, counts);
    return 0;
}

/* ---- literal rknn-naive (rknn-naive/rknn.cpp:37-63,169-198) ---- */

typedef struct {
    float d;
    int32_t x, y, id;
} naive_pair;

/* std::pair<float, Point>::operator< : distance first, then Point::operator<
 * (rknn.cpp:53-57): x, then y, then id. */
static int cmp_naive(const void* a, const void* b)
{
    const naive_pair* p = (const naive_pair*)a;
    const naive_pair* r = (const naive_pair*)b;
    if (p->d < r->d) return -1;
    if (r->d < p->d) return 1;
    if (p->x != r->x) return p->x < r->x ? -1 : 1;
    if (p->y != r->y) return p->y < r->y ? -1 : 1;
    if (p->id != r->id) return p->id < r->id ? -1 : 1;
    return 0;
}

int64_t oracle_rknn_naive_int(const int32_t* ux, const int32_t* uy, const int32_t* uid, uint64_t nu,
                              const int32_t* fx, const int32_t* fy, const int32_t* fid, uint64_t nf,
                              uint32_t q, uint32_t k, int32_t* out)
{
    if (k == 0 || nf == 0 || q >= nf) return -1;
    naive_pair* d = (naive_pair*)malloc(nf * sizeof(naive_pair));
    if (!d) return -1;
    const int32_t qid = fid[q];
    const uint64_t limit = k < nf ? k : nf;
    int64_t n = 0;
    for (uint64_t u = 0; u < nu; ++u) {
        for (uint64_t f = 0; f < nf; ++f) {
            /* int - int, pow(.,2) in double, sqrt in double, returned as float */
            double dx = (double)(ux[u] - fx[f]);
            double dy = (double)(uy[u] - fy[f]);
            d[f].d = (float)sqrt(pow(dx, 2.0) + pow(dy, 2.0));
            d[f].x = fx[f];
            d[f].y = fy[f];
            d[f].id = fid[f];
        }
        qsort(d, (size_t)nf, sizeof(naive_pair), cmp_naive);
        for (uint64_t i = 0; i < limit; ++i) {
            if (d[i].id == qid) {
                out[n++] = uid[u];
                break;
            }
        }
    }
    free(d);
    return n;
}

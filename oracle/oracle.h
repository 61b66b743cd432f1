/*
 * oracle.h -- CPU restatement of the reference's RkNN semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * CLI, the serving loop) links, loads or calls this code.  It is used by
 * tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg (timed, never shipped).
 *
 * The defining predicate ("rknn-naive computed in FP64, ties won by q"):
 *
 *   R(q,k) = { u : #{ f in C : D(u,f) < D(u,q) } < k },  C = F \ {f_q}
 *   D(u,p) = dx*dx + dy*dy,  dx = (double)u.x - (double)p.x,  dy likewise,
 *
 * evaluated in IEEE binary64 in exactly that order with no FMA contraction
 * (built with -ffp-contract=off).  It follows rknn-naive/rknn.cpp:169-198
 * (per-user distances to every facility, q kept iff among the k nearest)
 * with the strict-closer tie rule of rknn-inf/inf.h:77-93 (a point on the
 * bisector is on q's valid side), rknn-inf-gpu/inf_kernel.cu:20-66 and
 * rknn-naive-gpu/rknn_kernel.cu:150-162 (count dist < dist_q, stop at k).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   oracle_rknn / oracle_counts : closed-form invariants (sum over q of
 *     |R(q,k)| = N_u*min(k,N_f) on tie-free data, monotone in k, k>=N_f =>
 *     all users, k=1 => closed Voronoi cell by argmin), sort-based rank of q
 *     (a different algorithm), hand-built tie cases, planted mutants, and
 *     result counts of the reference's own rknn-naive and rknn-inf binaries
 *     compiled from source into oracle/_ref.
 *   oracle_rknn_naive_int : literal rknn-naive (float distance, sort by
 *     (dist,x,y,id)); pinned against the reference rknn-naive binary's
 *     counts on integer data with ties.
 *   oracle_tree_rknn : the reference's tree-culled verification
 *     (rknn-inf/inf.h:194-249); pinned by set equality with oracle_rknn on
 *     random, tie-laden and clustered data, and by the reference rknn-inf
 *     binary's counts where that binary is exact.
 */
#ifndef RKNN_ORACLE_H
#define RKNN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sorted user ids of R(q,k).  cand == NULL means C = F \ {f_q}; otherwise C
 * is the given facility-index list (entries equal to q are skipped).
 * out must hold nu ids.  Returns |R| or -1 on invalid arguments
 * (q >= nf, k == 0). */
int64_t oracle_rknn(const float* ux, const float* uy, const int32_t* uid, uint64_t nu,
                    const float* fx, const float* fy, uint64_t nf,
                    uint32_t q, uint32_t k, const uint32_t* cand, uint64_t nc,
                    int32_t* out);

/* Per-user count of facilities in C strictly closer than f_q, stopping at
 * cap (cap = UINT32_MAX counts all).  counts has nu entries, user order. */
int oracle_counts(const float* ux, const float* uy, uint64_t nu,
                  const float* fx, const float* fy, uint64_t nf,
                  uint32_t q, const uint32_t* cand, uint64_t nc, uint32_t cap,
                  uint32_t* counts);

/* Literal rknn-naive (rknn-naive/rknn.cpp:37-63,169-198): integer
 * coordinates, float distance sqrt(pow(dx,2)+pow(dy,2)), std::sort of
 * (float dist, Point) with Point ordered by (x,y,id), user kept iff the
 * query facility's id is among the first min(k,N_f).  Ids are written in
 * user order (the reference's result order).  Returns |R| or -1. */
int64_t oracle_rknn_naive_int(const int32_t* ux, const int32_t* uy, const int32_t* uid, uint64_t nu,
                              const int32_t* fx, const int32_t* fy, const int32_t* fid, uint64_t nf,
                              uint32_t q, uint32_t k, int32_t* out);

/* Number of worker threads used by the functions above (1 = serial). */
void oracle_set_threads(int n);
int oracle_get_threads(void);

/* ---- cpu-inf: the reference's tree-culled verification (oracle_tree.c) ----
 * Users in a bulk-loaded R-tree (16 entries per node); a node is pruned when
 * >= k candidates certainly contain its rectangle on their side; leaf users
 * are counted with the oracle predicate.  Same sets as oracle_rknn on the
 * same candidate list.  Pinned against oracle_rknn (tests/test_oracle.py). */
typedef struct oracle_tree oracle_tree;
oracle_tree* oracle_tree_build(const float* ux, const float* uy, const int32_t* uid, uint64_t nu);
void oracle_tree_free(oracle_tree* t);
/* Sorted ids of R(q,k) into out (nu entries, or NULL to count only);
 * cand == NULL means C = F \ {f_q}.  Returns |R| or -1. */
int64_t oracle_tree_rknn(const oracle_tree* t, const float* fx, const float* fy, uint64_t nf, uint32_t q,
                         uint32_t k, const uint32_t* cand, uint64_t nc, int32_t* out);
/* counts[i] = |R(q[i], k[i])| with candidates cands[off[i] .. off[i+1]) (or
 * every facility when cands == NULL), queries spread over the worker
 * threads, one query per thread at a time. */
int oracle_tree_rknn_many(const oracle_tree* t, const float* fx, const float* fy, uint64_t nf, uint64_t nq,
                          const uint32_t* q, const uint32_t* k, const uint64_t* off, const uint32_t* cands,
                          int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif

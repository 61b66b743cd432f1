/*
 * rknn_io.h -- input files and seeded synthetic datasets.
 *
 * Text format of rknn-inf/inf.cpp:136-158 and data/preprocess.py:67-73:
 *   fac_cnt \n  fac_cnt lines "id x y" \n  usr_cnt \n  usr_cnt lines "id x y"
 * Coordinates are parsed with strtof (integer and decimal both work; the
 * reference rknn-naive parses int only, rknn-naive/rknn.cpp:39).  The writer
 * prints each fp32 as its exact binary64 value (%.17g), so the reference's
 * double parser (inf.cpp:141-143) reads the identical values.
 *
 * Generators replace data/random_gen.py:3-11 (unseeded randint) with
 * SplitMix64 streams:
 *   config 1, 2  uniform fp32 (float)(1000*u53) in [0,1000)^2
 *   config 3     clustered + road-like, snapped to a 2^-7 grid, with 1% of
 *                users placed exactly on a bisector of a query facility and
 *                a near facility, or duplicated
 *   config 4, 5  half uniform, half config-3 style (ties included)
 * Facility ids are 0..N_f-1, user ids N_f..N_f+N_u-1 (disjoint, as in
 * data/preprocess.py:55-56).  Default seed: 0x524B4E4E ^ config.
 */
#ifndef RKNN_IO_H
#define RKNN_IO_H

#include <stddef.h>
#include <stdint.h>

#include "rknn.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Loads a reference-format file.  The six arrays are allocated by the
 * library; release them with rk_free. */
rk_status rk_load_text(const char* path, float** fx, float** fy, int32_t** fid, uint64_t* nf,
                       float** ux, float** uy, int32_t** uid, uint64_t* nu);
rk_status rk_write_text(const char* path, const float* fx, const float* fy, const int32_t* fid, uint64_t nf,
                        const float* ux, const float* uy, const int32_t* uid, uint64_t nu);
void rk_free(void* p);

/* Distinct query facility indices (partial Fisher-Yates), nq <= nf. */
rk_status rk_generate_queries(uint64_t nf, uint64_t nq, uint64_t seed, uint32_t* q_out);

/* Fills caller-allocated arrays.  nq_ties/seed_ties name the query set
 * whose bisectors carry the tie users (configs 3-5; ignored otherwise);
 * pass the nq and seed later given to rk_generate_queries. */
rk_status rk_generate(int config, uint64_t nu, uint64_t nf, uint64_t seed, uint64_t nq_ties,
                      uint64_t seed_ties, float* ux, float* uy, int32_t* uid, float* fx, float* fy,
                      int32_t* fid);

uint64_t rk_default_seed(int config);

#ifdef __cplusplus
}
#endif
#endif

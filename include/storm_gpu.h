/*
 * storm_gpu.h -- drop-in GPU build of the STORM_contiguous_t half of storm.h.
 *
 * Same names, argument meaning and error behaviour as the reference C ABI
 * (storm.h:224-231, storm.cpp:1159-1285); link libstorm_gpu.so instead of
 * storm.cpp and the pairwise calls answer from an MI355X through storm_hip.h.
 *
 * Differences a caller can observe:
 *   - a device failure (no GPU, out of device memory) makes the pairwise calls
 *     print the reason to stderr and return UINT64_MAX; the reason is also
 *     available from storm_hip_last_error(). The reference has no device and
 *     so no such failure; it returns 0 only for a NULL handle, as here.
 */
#ifndef STORM_GPU_H_
#define STORM_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct STORM_contiguous_s STORM_contiguous_t;

/* storm.cpp:1159-1170: NULL on allocation failure. */
STORM_contiguous_t* STORM_contig_new(size_t vector_length);
/* storm.cpp:1172-1174: NULL-safe. */
void STORM_contig_free(STORM_contiguous_t* bitmap);
/* storm.cpp:1176-1220: -1 NULL handle, -2 NULL values, 0 when n_values == 0
 * (no vector is added), -4 on allocation failure, else n_values. Values should
 * be sorted (the reference asserts it in debug builds only; order does not
 * change the result); repeated values are collapsed; values >= vector_length
 * are out of contract (the reference asserts; this build returns -3 and adds
 * nothing). */
int STORM_contig_add(STORM_contiguous_t* bitmap, const uint32_t* values, const uint32_t n_values);
/* storm.cpp:1222-1230: 1, or -1 for a NULL handle. */
int STORM_contig_clear(STORM_contiguous_t* bitmap);
/* storm.cpp:1232-1285: sum over i<j of |Xi n Xj|; 0 for a NULL handle. The
 * block size of the _blocked forms is a CPU cache knob with no effect on the
 * result; the device tiles by its own geometry. */
uint64_t STORM_contig_pairw_intersect_cardinality(STORM_contiguous_t* bitmap);
uint64_t STORM_contig_pairw_intersect_cardinality_blocked(STORM_contiguous_t* bitmap, uint32_t bsize);
uint64_t STORM_contig_pairw_intersect_cardinality_list(STORM_contiguous_t* bitmap);
uint64_t STORM_contig_pairw_intersect_cardinality_blocked_list(STORM_contiguous_t* bitmap, uint32_t bsize);

/* GPU-build extras (no reference counterpart). */
/* Per-pair counts of the last pairwise call into out[N(N-1)/2], packed upper
 * triangle, idx(i,j) = i*N - i(i+1)/2 + (j-i-1). Returns 0 or a negative
 * storm_hip status. */
int STORM_contig_gpu_pairw_counts(STORM_contiguous_t* bitmap, uint32_t* out);
/* Number of vectors held. */
size_t STORM_contig_gpu_size(const STORM_contiguous_t* bitmap);

/* ---- STORM_t: the 65,536-value block container (storm.h:150-221) ---------
 * Same names and results as storm.cpp:872-1046 for sorted input: one vector
 * per STORM_add, the pairwise calls sum |Ci n Cj| over every pair, the square
 * form over every (vector of the first, vector of the second). The host keeps
 * each vector as a sorted position list and the pairwise calls run the
 * engine's per-tile selection on the device (the corpus is uploaded once per
 * change). Serialized sizes follow the reference's block layout (blocks of
 * 65,536 values; fewer than 4,096 values in a block are u16 scalars).
 *
 * Differences a caller can observe:
 *   - repeated values count once (set semantics, as the reference's bitmap
 *     blocks do; its scalar blocks assume distinct values);
 *   - unsorted input is sorted (the reference requires sorted input);
 *   - the value 4,294,967,295 is out of range (the device universe is u32):
 *     STORM_add returns -3 and adds nothing;
 *   - a device failure makes the pairwise calls print the reason to stderr
 *     and return UINT64_MAX (the reference returns UINT64_MAX only for a NULL
 *     handle), and the square form return 0; the reason is also in
 *     storm_hip_last_error(). */
typedef struct STORM_s STORM_t;

/* storm.cpp:872-879: NULL on allocation failure. */
STORM_t* STORM_new(void);
/* storm.cpp:881-892: NULL-safe. */
void STORM_free(STORM_t* bitmap);
/* storm.cpp:894-916: -1 NULL handle; otherwise one vector is appended (empty
 * when values is NULL or n_values is 0) and 1 is returned. */
int STORM_add(STORM_t* bitmap, const uint32_t* values, const uint32_t n_values);
/* storm.cpp:918-925: 1, or -1 for a NULL handle. */
int STORM_clear(STORM_t* bitmap);
/* storm.cpp:927-1011: UINT64_MAX for a NULL handle. bsize is a CPU cache knob
 * with no effect on the result. */
uint64_t STORM_pairw_intersect_cardinality(STORM_t* bitmap);
uint64_t STORM_pairw_intersect_cardinality_blocked(STORM_t* bitmap, uint32_t bsize);
/* storm.cpp:1027-1046: 0 when either handle is NULL. */
uint64_t STORM_intersect_cardinality_square(const STORM_t* bitmap1, const STORM_t* bitmap2);
/* storm.cpp:1013-1023 with storm.cpp:396-418: 0 for a NULL handle. */
uint64_t STORM_serialized_size(const STORM_t* bitmap);

#ifdef __cplusplus
}
#endif
#endif

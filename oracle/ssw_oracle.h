/* ssw_oracle.h -- CPU oracle (test infrastructure only; see ssw_oracle.c). */
#ifndef SSW_ORACLE_H
#define SSW_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int64_t oracle_slim_chunk(int n, int64_t count, const int32_t *A, const int32_t *B,
                          int32_t *length, uint8_t *sync, int64_t *expanded, int32_t bound);
int64_t oracle_wide_chunk(int n, int64_t count, const int32_t *A, const int32_t *B,
                          int32_t *length, uint8_t *sync, int64_t *expanded, int32_t bound);
int64_t oracle_slim_parallel(int n, int64_t count, const int32_t *A, const int32_t *B,
                             int32_t *length, uint8_t *sync, int64_t *expanded,
                             int32_t bound, int threads);
int64_t oracle_wide_parallel(int n, int64_t count, const int32_t *A, const int32_t *B,
                             int32_t *length, uint8_t *sync, int64_t *expanded,
                             int32_t bound, int threads);
int oracle_witness(int n, const int32_t *a, const int32_t *b, int32_t *length,
                   uint8_t *word, int32_t cap);
int oracle_verify_word(int n, const int32_t *a, const int32_t *b, const uint8_t *word,
                       int32_t len);
int oracle_brute_force(int n, const int32_t *a, const int32_t *b, int32_t maxlen,
                       int32_t *length, uint8_t *word);
int oracle_pair_sync(int n, const int32_t *a, const int32_t *b);
#ifdef __cplusplus
}
#endif
#endif

// Lookup of tuned launch choices (qmx_tune.hip).
#pragma once
#include <cstdint>

namespace qmx {
// Bucket of a basis length used in tuning keys.
int64_t tune_g_bucket(int64_t G);
// Bucket of a batch count (the next power of two).
int64_t tune_b_bucket(int64_t b);
// The tuned choice for (op, key), if one is set: writes n_choice values.
bool tune_lookup(int op, const int64_t* key, int n_key, int64_t* choice, int n_choice);
}  // namespace qmx

// Tuned launch choices: a process-wide table from (operation, shape key) to
// a launch choice, filled by the host-side tuner (or from its cache file)
// and read by the launch code of qmx_dot, qmx_gemm and the plane pass before
// their own cost models. Keys with the basis length use its bucket
// (qmx::tune_g_bucket) and batches their power-of-two bucket, so that one
// entry serves the nearby sizes of a served stream's buckets.
#include <map>
#include <mutex>
#include <vector>

#include "qmx_common.h"
#include "qmx_tune.h"

namespace qmx {
namespace {
std::mutex g_tune_mu;
std::map<std::vector<int64_t>, std::vector<int64_t>> g_tune;
}  // namespace

int64_t tune_g_bucket(int64_t G) {
  // Quarter-octave buckets: the bucket of G is the smallest 2^(q/4) >= G.
  if (G <= 64) return 64;
  int64_t b = 64;
  int q = 0;
  while (true) {
    const double v = 64.0 * __builtin_pow(2.0, 0.25 * (++q));
    b = (int64_t)(v + 0.5);
    if (b >= G) return b;
  }
}

int64_t tune_b_bucket(int64_t b) {
  int64_t v = 1;
  while (v < b) v <<= 1;
  return v;
}

bool tune_lookup(int op, const int64_t* key, int n_key, int64_t* choice, int n_choice) {
  std::vector<int64_t> k(1 + n_key);
  k[0] = op;
  for (int i = 0; i < n_key; ++i) k[1 + i] = key[i];
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (g_tune.empty()) return false;
  const auto it = g_tune.find(k);
  if (it == g_tune.end() || (int)it->second.size() < n_choice) return false;
  for (int i = 0; i < n_choice; ++i) choice[i] = it->second[i];
  return true;
}
}  // namespace qmx

extern "C" {

int64_t qmx_tune_bucket(int64_t G) { return G < 0 ? -1 : qmx::tune_g_bucket(G); }

int64_t qmx_tune_batch_bucket(int64_t b) { return b < 0 ? -1 : qmx::tune_b_bucket(b); }

int qmx_tune_set(int op, const int64_t* key, int n_key, const int64_t* choice, int n_choice) {
  QMX_CHECK_ARG(op >= QMX_TUNE_DOT && op <= QMX_TUNE_PLANE, "qmx_tune_set: unknown operation");
  QMX_CHECK_ARG(key && choice && n_key > 0 && n_key <= 16 && n_choice > 0 && n_choice <= 16,
                "qmx_tune_set: bad key or choice");
  std::vector<int64_t> k(1 + n_key);
  k[0] = op;
  for (int i = 0; i < n_key; ++i) k[1 + i] = key[i];
  std::lock_guard<std::mutex> lock(qmx::g_tune_mu);
  qmx::g_tune[k] = std::vector<int64_t>(choice, choice + n_choice);
  return QMX_OK;
}

int qmx_tune_clear(void) {
  std::lock_guard<std::mutex> lock(qmx::g_tune_mu);
  qmx::g_tune.clear();
  return QMX_OK;
}

int64_t qmx_tune_count(void) {
  std::lock_guard<std::mutex> lock(qmx::g_tune_mu);
  return (int64_t)qmx::g_tune.size();
}

}  // extern "C"

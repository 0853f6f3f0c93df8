// tune.hip -- the autotuner's cache: per operator shape (kind, mask size,
// pixel types, accumulator, boundary mode, contraction, device weights,
// size class) the kernel variant and its tile parameter, with the measured
// time.  The cache
// is process-wide, guarded by a mutex, and can be saved to / loaded from a
// text file (one entry per line).
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

#include "hh_local.hpp"

namespace hh {

static std::mutex g_tune_mu;
using KeyTuple = std::tuple<int, int, int, int, int, int, int, int, int, int>;
static std::map<KeyTuple, TuneVal> g_tune;

static KeyTuple key_of(const TuneKey& k) {
  return std::make_tuple(k.kind, k.kx, k.ky, k.in_dt, k.out_dt, k.acc_f, k.boundary, k.fma, k.wdev,
                         k.sz);
}

bool tune_lookup(const TuneKey& k, TuneVal& v) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(key_of(k));
  if (it == g_tune.end()) return false;
  v = it->second;
  return true;
}

void tune_store(const TuneKey& k, const TuneVal& v) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key_of(k)] = v;
}

int tune_save(const char* path) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(HIPACC_ERR_INVALID_ARG, "tune_save: cannot open %s", path);
  std::fprintf(f, "# kind kx ky in_dtype out_dtype acc_float boundary fp_contract weights_dev "
                 "size_class variant param ms\n");
  for (const auto& e : g_tune) {
    const auto& k = e.first;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d %.6f\n", std::get<0>(k), std::get<1>(k),
                 std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k), std::get<6>(k),
                 std::get<7>(k), std::get<8>(k), std::get<9>(k), e.second.variant,
                 e.second.param, e.second.ms);
  }
  std::fclose(f);
  return HIPACC_OK;
}

int tune_load(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return set_error(HIPACC_ERR_INVALID_ARG, "tune_load: cannot open %s", path);
  char line[512];
  int n = 0;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '#') continue;
    TuneKey k;
    TuneVal v;
    if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d %f", &k.kind, &k.kx, &k.ky,
                    &k.in_dt, &k.out_dt, &k.acc_f, &k.boundary, &k.fma, &k.wdev, &k.sz,
                    &v.variant, &v.param, &v.ms) == 13) {
      if (v.variant < kVarTile || v.variant > kVarRing) continue;
      tune_store(k, v);
      ++n;
    }
  }
  std::fclose(f);
  return HIPACC_OK;
}

}  // namespace hh

// autotune.cpp -- the process-wide cache of launch shapes chosen by the
// library's autotuners (PhaseField::autotune for the stencil and general
// kernels, Lbfgs::autotune for the vector passes).
//
// A shape is a short list of integers stored under a key that names the
// kernel family, the problem shape and the precision.  Lookups happen when a
// potential is initialised or a pass is launched; an entry, once stored,
// applies to every later problem of the same shape in the process.  With
// MINIM_AUTOTUNE_CACHE naming a file, entries are read from it on first use
// and every new entry is appended to it, so later processes start tuned.
#include "autotune.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace minim_amd {

namespace {
std::mutex g_mu;
std::map<std::string, std::vector<int>>* g_cache = nullptr;

const char* cache_file() {
  const char* f = std::getenv("MINIM_AUTOTUNE_CACHE");
  return (f && *f) ? f : nullptr;
}

// Caller holds g_mu.
std::map<std::string, std::vector<int>>& cache() {
  if (!g_cache) {
    g_cache = new std::map<std::string, std::vector<int>>();
    if (const char* f = cache_file()) {
      std::ifstream in(f);
      std::string line;
      while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key.empty() || key[0] == '#') continue;
        std::vector<int> v;
        int x;
        while (ls >> x) v.push_back(x);
        if (!v.empty()) (*g_cache)[key] = v;
      }
    }
  }
  return *g_cache;
}
}  // namespace

bool tune_lookup(const std::string& key, std::vector<int>& out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto& c = cache();
  auto it = c.find(key);
  if (it == c.end()) return false;
  out = it->second;
  return true;
}

void tune_store(const std::string& key, const std::vector<int>& v) {
  std::lock_guard<std::mutex> lk(g_mu);
  cache()[key] = v;
  if (const char* f = cache_file()) {
    std::ofstream o(f, std::ios::app);
    o << key;
    for (int x : v) o << ' ' << x;
    o << '\n';
  }
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(g_mu);
  cache().clear();
}

std::string tune_key(const char* family, const std::vector<long>& shape, int prec) {
  std::ostringstream k;
  k << family;
  for (long s : shape) k << ':' << s;
  k << ':' << (prec == 0 ? "f64" : "f32");
  return k.str();
}

}  // namespace minim_amd

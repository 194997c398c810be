// autotune.hpp -- process-wide cache of launch shapes chosen by the
// library's autotuners (see autotune.cpp).
#pragma once

#include <string>
#include <vector>

namespace minim_amd {

// The shape stored under key, if any.
bool tune_lookup(const std::string& key, std::vector<int>& out);
// Store (and, with MINIM_AUTOTUNE_CACHE set, append to that file).
void tune_store(const std::string& key, const std::vector<int>& v);
// Forget every entry of this process (the file is left as it is).
void tune_clear();
// "family:s0:s1:...:f64|f32"
std::string tune_key(const char* family, const std::vector<long>& shape, int prec);

}  // namespace minim_amd

// tuner.cpp - launch-shape autotuner (see tuner.h).
#include "tuner.h"

#include <cstdio>
#include <cstdlib>

namespace bzg {

namespace {
// TUNE_RADIX_SHAPE: threads per 4096-row tile (16, 8 or 4 rows per thread)
const int kRadixShapes[] = {256, 512, 1024};
// TUNE_MTF_RUN: segments per MTF wave (the list state entering a run's first
// segment is computed, the later ones continue from the previous segment)
const int kMtfRuns[] = {1, 2, 4, 8};
}  // namespace

Tuner& Tuner::get() {
  static Tuner t;
  return t;
}

Tuner::Tuner() {
  for (int k = 0; k < TUNE_KNOBS; k++) forced_[k] = -1;
  const char* path = std::getenv("BZG_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  int k, c, v;
  while (std::fscanf(f, "%d %d %d", &k, &c, &v) == 3)
    if (k >= 0 && k < TUNE_KNOBS && c >= 0 && c < kTuneClasses && v >= 0 && v < variants(k))
      slot_[k][c].chosen = v;
  std::fclose(f);
}

int Tuner::variants(int knob) const {
  if (knob == TUNE_RADIX_SHAPE) return (int)(sizeof(kRadixShapes) / sizeof(kRadixShapes[0]));
  if (knob == TUNE_SORT_LAUNCH) return 3;
  if (knob == TUNE_MTF_RUN) return (int)(sizeof(kMtfRuns) / sizeof(kMtfRuns[0]));
  return 0;
}

int Tuner::value(int knob, int variant) const {
  if (variant < 0 || variant >= variants(knob)) return 0;
  if (knob == TUNE_RADIX_SHAPE) return kRadixShapes[variant];
  if (knob == TUNE_MTF_RUN) return kMtfRuns[variant];
  return variant;  // TUNE_SORT_LAUNCH: 0 per round, 1/2 device-driven (x1/x4 grids)
}

int Tuner::pick(int knob, int cls) {
  std::lock_guard<std::mutex> lk(mu_);
  if (forced_[knob] >= 0) return forced_[knob];
  Slot& s = slot_[knob][cls];
  if (s.chosen >= 0) return s.chosen;
  int best = 0;
  for (int v = 1; v < variants(knob); v++)
    if (s.n[v] < s.n[best]) best = v;
  return best;
}

void Tuner::report(int knob, int cls, int variant, double ns, double units) {
  std::lock_guard<std::mutex> lk(mu_);
  if (forced_[knob] >= 0 || variant < 0 || variant >= variants(knob) || !(ns > 0) || !(units > 0)) return;
  Slot& s = slot_[knob][cls];
  if (s.chosen >= 0) return;
  s.ns[variant] += ns;
  s.units[variant] += units;
  s.n[variant]++;
  for (int v = 0; v < variants(knob); v++)
    if (s.n[v] < kTuneTrials) return;
  int best = 0;
  for (int v = 1; v < variants(knob); v++)
    if (s.ns[v] / s.units[v] < s.ns[best] / s.units[best]) best = v;
  s.chosen = best;
  save_locked();
}

bool Tuner::exploring(int knob, int cls) {
  std::lock_guard<std::mutex> lk(mu_);
  return forced_[knob] < 0 && slot_[knob][cls].chosen < 0;
}

void Tuner::force(int knob, int variant) {
  std::lock_guard<std::mutex> lk(mu_);
  forced_[knob] = variant;
}

int Tuner::state(int knob, int cls, double* ns_per_unit, int* samples) {
  std::lock_guard<std::mutex> lk(mu_);
  const Slot& s = slot_[knob][cls];
  for (int v = 0; v < variants(knob); v++) {
    if (ns_per_unit) ns_per_unit[v] = s.units[v] > 0 ? s.ns[v] / s.units[v] : 0.0;
    if (samples) samples[v] = s.n[v];
  }
  return forced_[knob] >= 0 ? forced_[knob] : s.chosen;
}

void Tuner::reset() {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& row : slot_)
    for (auto& s : row) s = Slot{};
}

void Tuner::save_locked() {
  const char* path = std::getenv("BZG_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "w");
  if (!f) return;
  for (int k = 0; k < TUNE_KNOBS; k++)
    for (int c = 0; c < kTuneClasses; c++)
      if (slot_[k][c].chosen >= 0) std::fprintf(f, "%d %d %d\n", k, c, slot_[k][c].chosen);
  std::fclose(f);
}

}  // namespace bzg

// tuner.h - launch-shape autotuner (host side).
//
// Some kernels come in several launch shapes that compute identical results
// at different speeds on different inputs.  The tuner picks a shape per
// knob and size class: while exploring it hands out every variant in turn,
// timed by the caller with HIP events (nanoseconds and units of work per
// sample), and after kTuneTrials samples of each it keeps the one with the
// least total time per total work (so that small samples, dominated by
// fixed costs, weigh little) for the rest of the process.  Choices can be forced (tests), inspected (benchmarks) and kept
// across processes in a text file named by the BZG_TUNE_CACHE environment
// variable ("knob class variant" per line).
#pragma once

#include <cstdint>
#include <mutex>

namespace bzg {

enum TuneKnob : int {
  TUNE_RADIX_SHAPE = 0,  // threads per large-group radix-scatter workgroup
  TUNE_SORT_LAUNCH = 1,  // suffix-sort rounds: per-round readback with exact
                         // grids (0) or device-driven, later rounds on
                         // persistent grids of one (1) or four (2) residencies
  TUNE_MTF_RUN = 2,      // consecutive 4096-symbol MTF segments per wave
                         // (1, 2, 4, 8); classes: under 32 MiB of block
                         // capacity (0), larger (1)
  TUNE_KNOBS = 3
};
constexpr int kTuneMaxVariants = 4;
constexpr int kTuneClasses = 4;
constexpr int kTuneTrials = 2;

class Tuner {
 public:
  static Tuner& get();
  int variants(int knob) const;
  int value(int knob, int variant) const;  // the launch parameter of a variant
  int pick(int knob, int cls);             // variant to run now
  void report(int knob, int cls, int variant, double ns, double units);
  // whether a class of a knob has no choice yet
  bool exploring(int knob, int cls);
  void force(int knob, int variant);       // -1: automatic
  // chosen variant (-1 while exploring); per-variant mean ns/unit and samples
  int state(int knob, int cls, double* ns_per_unit, int* samples);
  void reset();

 private:
  Tuner();
  void save_locked();
  struct Slot {
    double ns[kTuneMaxVariants] = {};
    double units[kTuneMaxVariants] = {};
    int n[kTuneMaxVariants] = {};
    int chosen = -1;
  };
  std::mutex mu_;
  Slot slot_[TUNE_KNOBS][kTuneClasses];
  int forced_[TUNE_KNOBS];
};

// Size classes of the radix knob: round 0 sorts whole blocks, the later
// rounds the groups left over (their sizes are not known to the host when
// the rounds are enqueued from device-side counts).
enum : int { TUNE_CLASS_ROUND0 = 0, TUNE_CLASS_LATER = 1 };

}  // namespace bzg

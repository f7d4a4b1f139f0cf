// tune.hpp -- online autotuner for launch parameters that do not change results.
//
// A choice (for example the workgroup count of a persistent sweep phase, or the column splits
// of the kNN tile) is keyed by the kernel and a shape bucket. The first calls for a key try each
// candidate once and record its device time; afterwards the fastest is used. Only parameters
// whose every candidate yields bit-identical results are tuned this way, so exploration never
// changes an answer. The table is process-wide and guarded by a mutex.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace scgpu {

struct TuneChoice {
    int index;        // candidate to use for this call
    bool exploring;   // true while candidates are still being measured
};

// Picks the candidate for key (n_candidates > 0): while some candidate has fewer than reps
// measurements, the least-measured one; then the one with the smallest best time (taking the
// minimum over reps measurements discards one-time costs such as a first launch).
TuneChoice tune_pick(const std::string &key, int n_candidates, int reps = 2);
// Records the measured time of candidate index for key.
void tune_record(const std::string &key, int index, double ms);
// One line per key: the candidates' best times and the chosen one.
std::string tune_report();
void tune_reset();
bool tune_enabled();

}  // namespace scgpu

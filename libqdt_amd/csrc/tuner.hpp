// tuner.hpp -- the batched form's autotuner: the LDS cell size (128, 256 or
// 512 points, which sets the cell kernel's LDS per wavefront and so how many
// of its wavefronts share a CU) and the merges per wavefront of the
// one-thread-per-node merge kernel (16, 32 or 64), chosen per workload by
// timing every pair on the workload's first batch.
//
// Choices are kept per shape bucket (set count and total points, each to a
// power of two) in memory and in a text file ($QDT_TUNE_FILE, else
// $XDG_CACHE_HOME/qdt_amd/tune.txt, else ~/.cache/qdt_amd/tune.txt; one line
// "sets_log2 points_log2 cell_points merge_lanes device_ms"), so later
// processes skip the timing.  QDT_CELL_POINTS and QDT_MERGE_LANES override
// the choice.
#pragma once

#include "engine.hpp"

#include <stddef.h>

namespace qdt
{

struct TuneChoice
{
    int cell_points = 256;
    int merge_lanes = 16;
    double device_ms = 0; // the chosen pair's time on the tuning batch (0: not timed here)
    int source = 0;       // 0 default, 1 timed now, 2 from the cache, 3 from the environment
};

// The pair for this batch's shape bucket, timing the candidates on the
// engine (eager runs of this very batch) when the bucket has none yet.
Status tune_batch(Engine& eng, const double* d_pts, const size_t* set_off, int k, TuneChoice& out);

} // namespace qdt

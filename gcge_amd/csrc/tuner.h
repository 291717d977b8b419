// Launch-shape tuner.
//
// Some launch parameters have no single best value across problem shapes:
// the SpMM's column tile, workgroup size and row-block order (plain or one
// contiguous range per XCD), the row-block order of the FP32 BlockPCG's
// gathering kernels, and the row chunking of the MFMA Gram (how many
// workgroups share the rows of one output tile).  For
// each such operation and shape class the tuner keeps a choice in a
// process-wide table.  With tuning on, the first call of a class not yet
// in the table times every candidate on the call's own operands (HIP events
// on the calling stream), keeps the fastest and finishes the call with it;
// with tuning off, classes missing from the table use the built-in
// heuristic.  The table can be saved and loaded as text, and a table tuned
// on MI355X ships next to the library and is loaded when the first device
// context is created.
//
// SpMM and BlockPCG candidates give bit-identical results (each row is
// summed in the same order, partial sums are stored by row block).  Gram chunkings change the grouping of the partial sums, so
// a Gram result depends on the table: results are deterministic for a
// given table, and the shipped table fixes them by default.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>

namespace gcge {

enum TuneOp : int { kTuneSpmm = 1, kTuneGram = 2, kTuneSpmmPw = 3 };

struct TuneKey {
    int op = 0, a = 0, b = 0, c = 0;
};

int ilog2_floor(int64_t v);

// The table's choice for a key, or -1.
int tune_lookup(const TuneKey &k);
bool tune_enabled();
// Times launch(i) for each candidate i in [0, ncand) on stream s, records
// the fastest under k and returns it.  Every candidate must leave the same
// outputs valid (each overwrites them completely).
int tune_run(const TuneKey &k, int ncand, int dflt, const std::function<void(int)> &launch,
             hipStream_t s);
// Loads the table shipped next to the library, once per process.
void tune_load_default();

}  // namespace gcge

// Launch-shape choices: the measured tables written by imod-rs_amd/autotune.py
// (tile_table.inc), with the analytic fallbacks used where a table has no entry.
#pragma once

namespace tiltgpu {

enum { kTuneHeuristic = 0, kTuneTable = 1, kTuneForced = 2 };

// Backprojection tile depth for a kernel variant and shape: the depth that
// minimises padded depth x measured per-slice cost, with the costs of the
// table entry of the same variant nearest in log width.  0 when the table has
// no entry for the variant.
int tuned_tile_z(int two_row, int local, int half, int in_nx, int out_nz);

// Forward-projection bins per lane (4, 8 or 16) for a width: the table entry
// nearest in log width, else the width-only rule (at least 64 lanes per row).
int tuned_fp_bins(int in_nx, int* source);

}  // namespace tiltgpu

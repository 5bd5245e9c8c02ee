// Launch-shape choices: the measured tables written by imod-rs_amd/autotune.py
// (tile_table.inc), with the analytic fallbacks used where a table has no entry.
#pragma once

namespace tiltgpu {

enum { kTuneHeuristic = 0, kTuneTable = 1, kTuneForced = 2 };

// Backprojection tile depth for a kernel variant and shape (1 view per stage,
// 1 x per lane): the depth that minimises padded depth x measured per-slice
// cost, with the costs measured at the table's width and thickness nearest to
// (in_nx, out_nz).  0 when the table has no entry for the variant.
int tuned_tile_z(int two_row, int local, int half, int in_nx, int out_nz);

// The same over every measured launch shape: returns the tile depth (0 without
// a table entry) and sets views per LDS stage and x positions per lane.
int tuned_bp_shape(int two_row, int local, int half, int in_nx, int out_nz, int* vps, int* xpl);

// Forward-projection bins per lane (4, 8 or 16) for a width: the table entry
// nearest in log width, else the width-only rule (at least 64 lanes per row).
// rows (optional): volume rows per workgroup of that entry (1 or 2; 1 without
// a table entry).
int tuned_fp_bins(int in_nx, int* source, int* rows = nullptr, int* slide = nullptr);

}  // namespace tiltgpu

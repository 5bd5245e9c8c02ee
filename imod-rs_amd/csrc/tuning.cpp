// Table lookup of launch shapes (see tuning.h).
#include "tuning.h"

#include <cmath>

namespace tiltgpu {
namespace {

// Measured cost of one backprojection launch shape -- tile depth, views per
// LDS stage, x positions per lane -- for a kernel variant, storage, width and
// thickness: k = launch time / padded depth, relative to shape (128, 1, 1) of
// the same measurement = 1000 (0 = not built or did not fit).
struct BpCost {
    int two_row, local, half, in_nx, thickness, tile_z, vps, xpl, k;
};
struct FpTune {
    int in_nx, bins, rows, slide;
};

#include "tile_table.inc"

double log_dist(int a, int b) { return std::fabs(std::log2(static_cast<double>(a) / b)); }

// The measurement nearest to (in_nx, out_nz) of the variant: width first (log
// distance), then thickness.  Returns its (in_nx, thickness) or false.
bool nearest(int two_row, int local, int half, int in_nx, int out_nz, int* w, int* th) {
    double bd = 1e30;
    bool found = false;
    for (const BpCost& e : kBpCostTable) {
        if (e.two_row < 0 || e.two_row != two_row || e.local != local || e.half != half) continue;
        const double d = log_dist(e.in_nx, in_nx) * 16.0 + log_dist(e.thickness, out_nz);
        if (d < bd) {
            bd = d;
            *w = e.in_nx;
            *th = e.thickness;
            found = true;
        }
    }
    return found;
}

// time ~ padded depth x the measured per-slice cost of the nearest
// measurement; shapes restricted to (1, 1) when !shapes
int choose(int two_row, int local, int half, int in_nx, int out_nz, bool shapes, int* vps,
           int* xpl) {
    *vps = 1;
    *xpl = 1;
    int w = 0, th = 0;
    if (!nearest(two_row, local, half, in_nx, out_nz, &w, &th)) return 0;
    int choice = 0;
    double bt = 1e300;
    for (const BpCost& e : kBpCostTable) {
        if (e.two_row != two_row || e.local != local || e.half != half || e.in_nx != w ||
            e.thickness != th || e.k <= 0)
            continue;
        if (!shapes && (e.vps != 1 || e.xpl != 1)) continue;
        const double t =
            static_cast<double>((out_nz + e.tile_z - 1) / e.tile_z * e.tile_z) * e.k;
        if (t < bt) {
            bt = t;
            choice = e.tile_z;
            *vps = e.vps;
            *xpl = e.xpl;
        }
    }
    return choice;
}

}  // namespace

int tuned_tile_z(int two_row, int local, int half, int in_nx, int out_nz) {
    int v, x;
    return choose(two_row, local, half, in_nx, out_nz, false, &v, &x);
}

int tuned_bp_shape(int two_row, int local, int half, int in_nx, int out_nz, int* vps, int* xpl) {
    return choose(two_row, local, half, in_nx, out_nz, true, vps, xpl);
}

int tuned_fp_bins(int in_nx, int* source, int* rows, int* slide) {
    int best = 0, best_rows = 1, best_slide = 1;
    double bd = 1e30;
    for (const FpTune& e : kFpTable) {
        if (e.in_nx <= 0) continue;
        const double d = log_dist(e.in_nx, in_nx);
        if (d < bd) {
            bd = d;
            best = e.bins;
            best_rows = e.rows == 2 ? 2 : 1;
            best_slide = e.slide != 0;
        }
    }
    if (rows) *rows = 1;
    if (slide) *slide = 1;
    if (best == 4 || best == 8 || best == 16) {
        if (source) *source = kTuneTable;
        if (rows) *rows = best_rows;
        if (slide) *slide = best_slide;
        return best;
    }
    if (source) *source = kTuneHeuristic;
    if (in_nx <= 256) return 4;
    if (in_nx <= 512) return 8;
    return 16;
}

}  // namespace tiltgpu

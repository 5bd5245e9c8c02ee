// Table lookup of launch shapes (see tuning.h).
#include "tuning.h"

#include <cmath>

namespace tiltgpu {
namespace {

struct BpTune {
    int two_row, local, half, in_nx, k32, k64, k128, k256;  // per-padded-slice cost of each depth
};
struct FpTune {
    int in_nx, bins;
};

#include "tile_table.inc"

double log_dist(int a, int b) { return std::fabs(std::log2(static_cast<double>(a) / b)); }

}  // namespace

int tuned_tile_z(int two_row, int local, int half, int in_nx, int out_nz) {
    const BpTune* best = nullptr;
    double bd = 1e30;
    for (const BpTune& e : kBpTable) {
        if (e.two_row < 0 || e.two_row != two_row || e.local != local || e.half != half) continue;
        const double d = log_dist(e.in_nx, in_nx);
        if (d < bd) {
            bd = d;
            best = &e;
        }
    }
    if (!best) return 0;
    // time ~ padded depth x per-slice cost of the depth
    const int tz[4] = {32, 64, 128, 256};
    const int k[4] = {best->k32, best->k64, best->k128, best->k256};
    int choice = 0;
    double bt = 1e300;
    for (int i = 0; i < 4; ++i) {
        if (k[i] <= 0) continue;
        const double t = static_cast<double>((out_nz + tz[i] - 1) / tz[i] * tz[i]) * k[i];
        if (t < bt) {
            bt = t;
            choice = tz[i];
        }
    }
    return choice;
}

int tuned_fp_bins(int in_nx, int* source) {
    int best = 0;
    double bd = 1e30;
    for (const FpTune& e : kFpTable) {
        if (e.in_nx <= 0) continue;
        const double d = log_dist(e.in_nx, in_nx);
        if (d < bd) {
            bd = d;
            best = e.bins;
        }
    }
    if (best == 4 || best == 8 || best == 16) {
        if (source) *source = kTuneTable;
        return best;
    }
    if (source) *source = kTuneHeuristic;
    if (in_nx <= 256) return 4;
    if (in_nx <= 512) return 8;
    return 16;
}

}  // namespace tiltgpu

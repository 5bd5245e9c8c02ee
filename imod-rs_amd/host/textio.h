// Text parsing helpers shared by the host utilities and the `tilt` front end.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tiltgpu_host {
// Per-view local-alignment grids, concatenated in view order.
struct LocalGrids {
    std::vector<int32_t> nx, ny;
    std::vector<float> xs, ys, dx, dy;
};
bool load_local_grids(const std::string& path, int nz, LocalGrids* g, std::string* err);
std::string trim(const std::string& s);
bool parse_f32(const std::string& t, float* out);
bool parse_usize(const std::string& t, long long* out);
}  // namespace tiltgpu_host

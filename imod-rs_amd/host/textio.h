// Text parsing helpers shared by the host utilities and the `tilt` front end.
#pragma once
#include <string>

namespace tiltgpu_host {
std::string trim(const std::string& s);
bool parse_f32(const std::string& t, float* out);
bool parse_usize(const std::string& t, long long* out);
}  // namespace tiltgpu_host

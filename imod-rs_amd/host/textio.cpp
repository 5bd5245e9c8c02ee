// Text inputs of `tilt`: .tlt angles and per-view files.
//   read_tilt_file      crates/imod-transforms/src/tilt.rs:6-23
//   load_per_view_file  apps/tilt/src/main.rs:310-326
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "../../include/tiltgpu.h"
#include "textio.h"

namespace tiltgpu_host {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// str::parse::<f32>: an optional sign, decimal digits with an optional point and
// exponent, or inf / infinity / nan; nothing else (no hex, no spaces).
bool parse_f32(const std::string& t, float* out) {
    if (t.empty()) return false;
    size_t i = 0;
    if (t[i] == '+' || t[i] == '-') ++i;
    std::string rest = t.substr(i);
    std::string low;
    for (char c : rest) low += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    if (low == "inf" || low == "infinity" || low == "nan") {
        float v = low == "nan" ? NAN : INFINITY;
        *out = (t[0] == '-') ? -v : v;
        return true;
    }
    bool digits = false, point = false;
    size_t j = 0;
    for (; j < rest.size(); ++j) {
        const char c = rest[j];
        if (std::isdigit(static_cast<unsigned char>(c))) {
            digits = true;
        } else if (c == '.' && !point) {
            point = true;
        } else {
            break;
        }
    }
    if (!digits) return false;
    if (j < rest.size()) {
        if (rest[j] != 'e' && rest[j] != 'E') return false;
        ++j;
        if (j < rest.size() && (rest[j] == '+' || rest[j] == '-')) ++j;
        bool ed = false;
        for (; j < rest.size(); ++j) {
            if (!std::isdigit(static_cast<unsigned char>(rest[j]))) return false;
            ed = true;
        }
        if (!ed) return false;
    }
    errno = 0;
    *out = std::strtof(t.c_str(), nullptr);
    return true;
}

bool parse_usize(const std::string& t, long long* out) {
    if (t.empty()) return false;
    size_t i = (t[0] == '+') ? 1 : 0;
    if (i == t.size()) return false;
    for (size_t j = i; j < t.size(); ++j)
        if (!std::isdigit(static_cast<unsigned char>(t[j]))) return false;
    *out = std::strtoll(t.c_str() + i, nullptr, 10);
    return true;
}

}  // namespace tiltgpu_host

using namespace tiltgpu_host;

extern "C" int tiltgpu_read_tilt_file(const char* path, float* out, int32_t cap) {
    std::ifstream f(path);
    if (!f) return TILTGPU_ERR_IO;
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        const std::string t = trim(line);
        if (t.empty()) continue;
        float v;
        if (!parse_f32(t, &v)) continue;  // unparsable lines are skipped (tilt.rs:17-19)
        if (n < cap && out) out[n] = v;
        ++n;
    }
    return n;
}

extern "C" int tiltgpu_read_per_view_file(const char* path, float* out, int32_t cap) {
    std::ifstream f(path);
    if (!f) return TILTGPU_ERR_IO;
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        float v;
        if (!parse_f32(t, &v)) return TILTGPU_ERR_IO;  // main.rs:320-322 exits on a parse error
        if (n < cap && out) out[n] = v;
        ++n;
    }
    return n;
}

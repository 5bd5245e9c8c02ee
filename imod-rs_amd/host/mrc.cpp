// MRC 2014 reader / writer for the `tilt` drop-in (see mrc.h).
#include "mrc.h"

#include <cmath>
#include <cstring>

#include "../../include/tiltgpu.h"

namespace tiltgpu_host {

namespace {

// Multi-byte fields of the header as (offset, width): every 4-byte word and
// 2-byte short, in on-disk order (byte arrays are not swapped).
struct Field {
    int off, width;
};
const Field kFields[] = {
    {0, 4},   {4, 4},   {8, 4},   {12, 4},  {16, 4},  {20, 4},  {24, 4},  {28, 4},
    {32, 4},  {36, 4},  {40, 4},  {44, 4},  {48, 4},  {52, 4},  {56, 4},  {60, 4},
    {64, 4},  {68, 4},  {72, 4},  {76, 4},  {80, 4},  {84, 4},  {88, 4},  {92, 4},
    {96, 2},  {108, 4}, {128, 2}, {130, 2}, {132, 2}, {134, 2}, {136, 4}, {140, 4},
    {144, 4}, {148, 4}, {152, 4}, {156, 4}, {160, 2}, {162, 2}, {164, 2}, {166, 2},
    {168, 2}, {170, 2}, {172, 4}, {176, 4}, {180, 4}, {184, 4}, {188, 4}, {192, 4},
    {196, 4}, {200, 4}, {204, 4}, {216, 4}, {220, 4},
};

void swap_bytes(uint8_t* p, int w) {
    for (int i = 0; i < w / 2; ++i) std::swap(p[i], p[w - 1 - i]);
}

void swap_header(MrcHeader* h) {
    auto* b = reinterpret_cast<uint8_t*>(h);
    for (const Field& f : kFields) swap_bytes(b + f.off, f.width);
}

// MrcReader::dimensions_reasonable (read.rs:457-464)
bool reasonable(const MrcHeader& h) {
    return h.nx > 0 && h.ny > 0 && h.nz > 0 && h.nx < 100000 && h.ny < 100000 &&
           h.nz < 1000000;
}

int bytes_per_pixel(int mode) {
    switch (mode) {
        case 0: return 1;
        case 1: return 2;
        case 2: return 4;
        case 3: return 4;
        case 4: return 8;
        case 6: return 2;
        case 12: return 2;
        case 16: return 3;
        case 101: return 1;
        default: return -1;
    }
}

// IEEE half -> float (exact; read.rs:468-498 does the same conversion)
float half_to_float(uint16_t bits) {
    const uint32_t sign = (bits >> 15) & 1u, ex = (bits >> 10) & 0x1fu, man = bits & 0x3ffu;
    uint32_t out;
    if (ex == 0) {
        if (man == 0) {
            out = sign << 31;
        } else {
            uint32_t m = man;
            int e = 0;
            while ((m & 0x400u) == 0) {
                m <<= 1;
                ++e;
            }
            out = (sign << 31) | (static_cast<uint32_t>(127 - 15 + 1 - e) << 23) | ((m & 0x3ffu) << 13);
        }
    } else if (ex == 31) {
        out = (sign << 31) | (0xffu << 23) | (man << 13);
    } else {
        out = (sign << 31) | ((ex - 15 + 127) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &out, 4);
    return f;
}

template <typename T>
T get(const uint8_t* p, bool swapped) {
    uint8_t b[sizeof(T)];
    std::memcpy(b, p, sizeof(T));
    if (swapped) swap_bytes(b, sizeof(T));
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

}  // namespace

MrcHeader MrcHeader::make(int32_t nx, int32_t ny, int32_t nz, int32_t mode) {
    MrcHeader h;
    std::memset(&h, 0, sizeof h);
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.mode = mode;
    h.mx = nx;
    h.my = ny;
    h.mz = nz;
    h.xlen = static_cast<float>(nx);
    h.ylen = static_cast<float>(ny);
    h.zlen = static_cast<float>(nz);
    h.alpha = h.beta = h.gamma = 90.0f;
    h.mapc = 1;
    h.mapr = 2;
    h.maps = 3;
    h.nversion = 20140;
    h.imod_stamp = kImodStamp;
    std::memcpy(h.cmap, "MAP ", 4);
    h.stamp[0] = 0x44;
    h.stamp[1] = 0x44;
    h.rms = -1.0f;
    return h;
}

bool MrcHeader::add_label(const std::string& text) {
    if (nlabl < 0 || nlabl >= 10) return false;
    const size_t len = std::min<size_t>(text.size(), 80);
    std::memcpy(labels[nlabl], text.data(), len);
    for (size_t i = len; i < 80; ++i) labels[nlabl][i] = ' ';  // IMOD pads with spaces
    ++nlabl;
    return true;
}

std::string MrcHeader::label(int i) const {
    if (i < 0 || i >= nlabl || i >= 10) return std::string();
    int end = 80;
    while (end > 0 && (labels[i][end - 1] == ' ' || labels[i][end - 1] == 0)) --end;
    return std::string(labels[i], labels[i] + end);
}

size_t MrcHeader::slice_bytes() const {
    const int m = bytes_per_pixel(mode) < 0 ? 0 : mode;  // unknown mode: sized as bytes
    if (m == 101) return static_cast<size_t>((nx + 1) / 2) * ny;
    return static_cast<size_t>(nx) * ny * bytes_per_pixel(m);
}

MrcReader::~MrcReader() {
    if (f_) std::fclose(f_);
}

bool MrcReader::open(const std::string& path, std::string* err) {
    f_ = std::fopen(path.c_str(), "rb");
    if (!f_) {
        *err = "cannot open " + path;
        return false;
    }
    if (std::fread(&h_, 1, 1024, f_) != 1024) {
        *err = "invalid header: file shorter than 1024 bytes";
        return false;
    }
    swapped_ = (h_.stamp[0] == 0x11 && h_.stamp[1] == 0x11) || !reasonable(h_);
    if (swapped_) {
        swap_header(&h_);
        if (!reasonable(h_)) {
            *err = "invalid header: unreasonable dimensions even after byte-swap";
            return false;
        }
    }
    if (h_.nversion == 0 && std::memcmp(h_.cmap, "MAP ", 4) != 0) {
        // old-style header: origins stored as zorg, xorg, yorg (read.rs:53-80)
        const float ox = h_.yorg, oy = h_.zorg, oz = h_.xorg;
        h_.xorg = ox;
        h_.yorg = oy;
        h_.zorg = oz;
    }
    return true;
}

bool MrcReader::read_slice_f32_into(int z, float* dst, size_t cap, size_t* n, std::string* err) {
    if (z < 0 || z >= h_.nz) {
        *err = "slice index out of range";
        return false;
    }
    if (h_.mode == 2 && !swapped_) {
        const size_t npix = static_cast<size_t>(h_.nx) * h_.ny;
        if (npix > cap) {
            *err = "slice larger than the buffer";
            return false;
        }
        const uint64_t off = h_.data_offset() + static_cast<uint64_t>(z) * h_.slice_bytes();
        if (fseeko(f_, static_cast<off_t>(off), SEEK_SET) != 0 ||
            std::fread(dst, sizeof(float), npix, f_) != npix) {
            *err = "short read of slice " + std::to_string(z);
            return false;
        }
        *n = npix;
        return true;
    }
    std::vector<float> s;
    if (!read_slice_f32(z, &s, err)) return false;
    if (s.size() > cap) {
        *err = "slice larger than the buffer";
        return false;
    }
    std::memcpy(dst, s.data(), s.size() * sizeof(float));
    *n = s.size();
    return true;
}

bool MrcReader::read_slice_f32(int z, std::vector<float>* out, std::string* err) {
    if (z < 0 || z >= h_.nz) {
        *err = "slice index out of range";
        return false;
    }
    const int mode = h_.mode;
    if (bytes_per_pixel(mode) < 0) {
        *err = "unsupported MRC mode " + std::to_string(mode);
        return false;
    }
    const size_t sb = h_.slice_bytes();
    std::vector<uint8_t> raw(sb);
    const uint64_t off = h_.data_offset() + static_cast<uint64_t>(z) * sb;
    if (fseeko(f_, static_cast<off_t>(off), SEEK_SET) != 0 || std::fread(raw.data(), 1, sb, f_) != sb) {
        *err = "short read of slice " + std::to_string(z);
        return false;
    }
    const size_t npix = static_cast<size_t>(h_.nx) * h_.ny;
    const uint8_t* p = raw.data();
    out->clear();
    switch (mode) {
        case 0:
        case 16:
            out->assign(raw.begin(), raw.end());
            break;
        case 1:
            out->resize(npix);
            for (size_t i = 0; i < npix; ++i) (*out)[i] = get<int16_t>(p + 2 * i, swapped_);
            break;
        case 6:
            out->resize(npix);
            for (size_t i = 0; i < npix; ++i) (*out)[i] = get<uint16_t>(p + 2 * i, swapped_);
            break;
        case 2:
            out->resize(npix);
            for (size_t i = 0; i < npix; ++i) (*out)[i] = get<float>(p + 4 * i, swapped_);
            break;
        case 3:
            out->resize(npix);
            for (size_t i = 0; i < npix; ++i) {
                const float re = get<int16_t>(p + 4 * i, swapped_);
                const float im = get<int16_t>(p + 4 * i + 2, swapped_);
                (*out)[i] = std::sqrt(re * re + im * im);
            }
            break;
        case 4:
            out->resize(npix);
            for (size_t i = 0; i < npix; ++i) {
                const float re = get<float>(p + 8 * i, swapped_);
                const float im = get<float>(p + 8 * i + 4, swapped_);
                (*out)[i] = std::sqrt(re * re + im * im);
            }
            break;
        case 12:
            out->resize(npix);
            for (size_t i = 0; i < npix; ++i) (*out)[i] = half_to_float(get<uint16_t>(p + 2 * i, swapped_));
            break;
        case 101: {
            out->reserve(npix);
            const size_t rb = static_cast<size_t>((h_.nx + 1) / 2);
            for (int row = 0; row < h_.ny; ++row)
                for (int col = 0; col < h_.nx; ++col) {
                    const uint8_t b = p[row * rb + col / 2];
                    out->push_back(static_cast<float>(col % 2 == 0 ? (b & 0x0f) : (b >> 4)));
                }
            break;
        }
    }
    return true;
}

MrcWriter::~MrcWriter() {
    if (f_) std::fclose(f_);
}

bool MrcWriter::create(const std::string& path, const MrcHeader& h, std::string* err) {
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) {
        *err = "cannot create " + path;
        return false;
    }
    h_ = h;
    if (std::fwrite(&h_, 1, 1024, f_) != 1024) {
        *err = "write failed";
        return false;
    }
    if (h_.next > 0) {
        std::vector<uint8_t> z(static_cast<size_t>(h_.next), 0);
        if (std::fwrite(z.data(), 1, z.size(), f_) != z.size()) {
            *err = "write failed";
            return false;
        }
    }
    return true;
}

bool MrcWriter::write_slice_f32(const float* data, size_t n, std::string* err) {
    if (h_.mode != 2) {
        *err = "only mode 2 (float) output is written";
        return false;
    }
    if (n != static_cast<size_t>(h_.nx) * h_.ny) {
        *err = "slice size does not match the header";
        return false;
    }
    if (std::fwrite(data, sizeof(float), n, f_) != n) {  // little-endian host
        *err = "write failed";
        return false;
    }
    return true;
}

bool MrcWriter::finish(float amin, float amax, float amean, std::string* err) {
    h_.amin = amin;
    h_.amax = amax;
    h_.amean = amean;
    if (std::fseek(f_, 0, SEEK_SET) != 0 || std::fwrite(&h_, 1, 1024, f_) != 1024 ||
        std::fflush(f_) != 0) {
        *err = "header rewrite failed";
        return false;
    }
    std::fclose(f_);
    f_ = nullptr;
    return true;
}

}  // namespace tiltgpu_host

// ------------------------------------------------------------------ C ABI
using tiltgpu_host::MrcHeader;
using tiltgpu_host::MrcReader;
using tiltgpu_host::MrcWriter;

extern "C" int tiltgpu_mrc_info_read(const char* path, tiltgpu_mrc_info* out) {
    if (!path || !out) return TILTGPU_ERR_ARG;
    MrcReader r;
    std::string err;
    if (!r.open(path, &err)) return TILTGPU_ERR_IO;
    const MrcHeader& h = r.header();
    std::memset(out, 0, sizeof *out);
    out->nx = h.nx;
    out->ny = h.ny;
    out->nz = h.nz;
    out->mode = h.mode;
    out->mx = h.mx;
    out->my = h.my;
    out->mz = h.mz;
    out->next = h.next;
    out->nlabl = h.nlabl;
    out->swapped = r.swapped() ? 1 : 0;
    out->is_imod = h.imod_stamp == MrcHeader::kImodStamp ? 1 : 0;
    out->xlen = h.xlen;
    out->ylen = h.ylen;
    out->zlen = h.zlen;
    out->amin = h.amin;
    out->amax = h.amax;
    out->amean = h.amean;
    const std::string l0 = h.label(0);
    std::memcpy(out->label0, l0.data(), std::min<size_t>(l0.size(), 80));
    return TILTGPU_OK;
}

extern "C" int tiltgpu_mrc_read_sections(const char* path, int32_t z0, int32_t z1, float* out,
                                         int64_t cap) {
    if (!path || !out) return TILTGPU_ERR_ARG;
    MrcReader r;
    std::string err;
    if (!r.open(path, &err)) return TILTGPU_ERR_IO;
    int64_t pos = 0;
    std::vector<float> s;
    for (int z = z0; z < z1; ++z) {
        if (!r.read_slice_f32(z, &s, &err)) return TILTGPU_ERR_IO;
        if (pos + static_cast<int64_t>(s.size()) > cap) return TILTGPU_ERR_ARG;
        std::memcpy(out + pos, s.data(), s.size() * sizeof(float));
        pos += static_cast<int64_t>(s.size());
    }
    return TILTGPU_OK;
}

extern "C" int tiltgpu_mrc_write_float(const char* path, const float* data, int32_t nx, int32_t ny,
                                       int32_t nz, float xlen, float ylen, float zlen,
                                       const char* label) {
    if (!path || !data || nx < 1 || ny < 1 || nz < 1) return TILTGPU_ERR_ARG;
    MrcHeader h = MrcHeader::make(nx, ny, nz, 2);
    if (xlen > 0) h.xlen = xlen;
    if (ylen > 0) h.ylen = ylen;
    if (zlen > 0) h.zlen = zlen;
    if (label) h.add_label(label);
    MrcWriter w;
    std::string err;
    if (!w.create(path, h, &err)) return TILTGPU_ERR_IO;
    const size_t per = static_cast<size_t>(nx) * ny;
    float gmin = 3.40282347e38f, gmax = -3.40282347e38f;
    double gsum = 0;
    for (int z = 0; z < nz; ++z) {
        const float* s = data + z * per;
        float mn = 3.40282347e38f, mx = -3.40282347e38f;
        double sum = 0;
        for (size_t i = 0; i < per; ++i) {
            mn = s[i] < mn ? s[i] : mn;
            mx = s[i] > mx ? s[i] : mx;
            sum += s[i];
        }
        gmin = mn < gmin ? mn : gmin;
        gmax = mx > gmax ? mx : gmax;
        gsum += static_cast<double>(static_cast<float>(sum / per)) * static_cast<double>(per);
        if (!w.write_slice_f32(s, per, &err)) return TILTGPU_ERR_IO;
    }
    const float gmean = static_cast<float>(gsum / (static_cast<double>(per) * nz));
    return w.finish(gmin, gmax, gmean, &err) ? TILTGPU_OK : TILTGPU_ERR_IO;
}

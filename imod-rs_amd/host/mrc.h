// MRC 2014 image files: the subset of imod-rs crates/imod-mrc that `tilt` uses
// (MrcHeader::new / add_label, MrcReader::open / read_slice_f32,
// MrcWriter::create / write_slice_f32 / finish).  On-disk layout: the 1024-byte
// header of crates/imod-mrc/src/header.rs:8-110 (word order fixed by the MRC
// format), little-endian on write, either endianness on read.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace tiltgpu_host {

#pragma pack(push, 1)
struct MrcHeader {
    int32_t nx, ny, nz, mode;
    int32_t nxstart, nystart, nzstart;
    int32_t mx, my, mz;
    float xlen, ylen, zlen;
    float alpha, beta, gamma;
    int32_t mapc, mapr, maps;
    float amin, amax, amean;
    int32_t ispg, next;
    int16_t creatid;
    uint8_t blank[6];
    uint8_t ext_type[4];
    int32_t nversion;
    uint8_t blank2[16];
    int16_t nint, nreal, sub, zfac;
    float min2, max2, min3, max3;
    int32_t imod_stamp, imod_flags;
    int16_t idtype, lens, nd1, nd2, vd1, vd2;
    float tilt_angles[6];
    float xorg, yorg, zorg;
    uint8_t cmap[4];
    uint8_t stamp[4];
    float rms;
    int32_t nlabl;
    char labels[10][80];

    static constexpr int32_t kImodStamp = static_cast<int32_t>(0x25004488u);
    // MrcHeader::new (header.rs:181-241)
    static MrcHeader make(int32_t nx, int32_t ny, int32_t nz, int32_t mode);
    // add_label (header.rs:244-258): false when all 10 slots are used
    bool add_label(const std::string& text);
    std::string label(int i) const;
    size_t slice_bytes() const;
    uint64_t data_offset() const { return 1024u + static_cast<uint64_t>(next > 0 ? next : 0); }
};
#pragma pack(pop)
static_assert(sizeof(MrcHeader) == 1024, "MRC header must be 1024 bytes");

class MrcReader {
  public:
    // MrcReader::open (read.rs:28-93): little-endian first, big-endian when the
    // stamp says so or the dimensions are unreasonable; old-style origin fix.
    bool open(const std::string& path, std::string* err);
    const MrcHeader& header() const { return h_; }
    bool swapped() const { return swapped_; }
    // read_slice_f32 (read.rs:127-231): any supported mode -> f32
    bool read_slice_f32(int z, std::vector<float>* out, std::string* err);
    // The same conversion into dst[0 .. cap): mode 2 in host byte order is read
    // straight into dst.  Returns the value count through *n.
    bool read_slice_f32_into(int z, float* dst, size_t cap, size_t* n, std::string* err);
    ~MrcReader();

  private:
    FILE* f_ = nullptr;
    MrcHeader h_{};
    bool swapped_ = false;
};

class MrcWriter {
  public:
    // MrcWriter::create (write.rs:20-41)
    bool create(const std::string& path, const MrcHeader& h, std::string* err);
    // write_slice_f32 for mode 2 (write.rs:58-104)
    bool write_slice_f32(const float* data, size_t n, std::string* err);
    // finish (write.rs:141-153): header statistics, header rewritten
    bool finish(float amin, float amax, float amean, std::string* err);
    ~MrcWriter();

  private:
    FILE* f_ = nullptr;
    MrcHeader h_{};
};

}  // namespace tiltgpu_host

// `ctfphaseflip` -- the drop-in front end of imod-rs apps/ctfphaseflip: its
// option surface (main.rs:20-89) and file handling (main.rs:212-705), with the
// strip transforms, the CTF flip and the blend run on the GPU through
// libtiltgpu (tiltgpu_ctf_*, include/tiltgpu.h).  The stack streams to the
// device in chunks of sections (file reads overlapping uploads), is corrected
// there, and streams back the same way while sections are written with the
// reference's header statistics.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tiltgpu.h"
#include "mrc.h"
#include "pipeline.h"
#include "textio.h"

using namespace tiltgpu_host;

namespace {

struct Args {
    std::string input, output, tilt_file, defocus_file;
    float defocus = -3000.0f, voltage = 300.0f, cs = 2.7f, pixel = 0.0f;
    bool has_pixel = false;
    int strip_width = 256;
    float tilt_axis = 0.0f;
    int max_strip_width = 0;
    bool has_max = false;
    float amp = 0.07f, cuton = 0.0f, plate = 0.0f;
    bool gpu = false;
    int upload_sections = 0;  // sections per transfer chunk; 0 = about 128 MB
};

[[noreturn]] void usage(const char* msg) {
    if (msg) std::fprintf(stderr, "error: %s\n\n", msg);
    std::fprintf(stderr,
                 "CTF phase-flip correction of a tilt series (MI355X)\n\n"
                 "Usage: ctfphaseflip -i <INPUT> -o <OUTPUT> -t <TILT_FILE> [OPTIONS]\n\n"
                 "  -i, --input <f>               input tilt series (MRC)\n"
                 "  -o, --output <f>              output corrected stack (MRC)\n"
                 "  -t, --tilt-file <f>           tilt angle file (.tlt)\n"
                 "  -d, --defocus <nm>            defocus, negative = underfocus [-3000]\n"
                 "      --defocus-file <f>        per-view 'view defocus1 [defocus2 astig]'\n"
                 "  -v, --voltage <kV>            [300]\n"
                 "  -c, --cs <mm>                 spherical aberration [2.7]\n"
                 "  -p, --pixel-size <A>          pixel size (default: from the header)\n"
                 "  -w, --strip-width <n>         strip width in pixels [256]\n"
                 "      --tilt-axis-angle <deg>   tilt axis rotation from Y [0]\n"
                 "      --max-strip-width <n>     maximum strip width (default: automatic)\n"
                 "      --amplitude-contrast <w>  [0.07]\n"
                 "      --cuton-freq <1/A>        no correction below this frequency [0]\n"
                 "      --plate-phase <deg>       phase plate shift [0]\n"
                 "      --gpu                     accepted (the correction always runs on the GPU)\n"
                 "      --upload-sections <n>     sections per file<->device transfer [about 128 MB]\n");
    std::exit(2);
}

bool to_usize(const std::string& s, int* v) {
    long long x;
    if (!parse_usize(s, &x) || x > 0x7fffffff) return false;
    *v = static_cast<int>(x);
    return true;
}

Args parse_args(int argc, char** argv) {
    Args a;
    struct Opt {
        const char* lng;
        char sht;
        int kind;  // 0 flag, 1 string, 2 usize, 3 float
        void* dst;
        bool* seen;
    };
    const Opt opts[] = {
        {"input", 'i', 1, &a.input, nullptr},
        {"output", 'o', 1, &a.output, nullptr},
        {"tilt-file", 't', 1, &a.tilt_file, nullptr},
        {"defocus", 'd', 3, &a.defocus, nullptr},
        {"defocus-file", 0, 1, &a.defocus_file, nullptr},
        {"voltage", 'v', 3, &a.voltage, nullptr},
        {"cs", 'c', 3, &a.cs, nullptr},
        {"pixel-size", 'p', 3, &a.pixel, &a.has_pixel},
        {"strip-width", 'w', 2, &a.strip_width, nullptr},
        {"tilt-axis-angle", 0, 3, &a.tilt_axis, nullptr},
        {"max-strip-width", 0, 2, &a.max_strip_width, &a.has_max},
        {"amplitude-contrast", 0, 3, &a.amp, nullptr},
        {"cuton-freq", 0, 3, &a.cuton, nullptr},
        {"gpu", 0, 0, &a.gpu, nullptr},
        {"plate-phase", 0, 3, &a.plate, nullptr},
        {"upload-sections", 0, 2, &a.upload_sections, nullptr},
    };
    for (int i = 1; i < argc; ++i) {
        std::string arg = argv[i];
        if (arg == "-h" || arg == "--help") usage(nullptr);
        const Opt* o = nullptr;
        std::string val;
        bool has_val = false;
        if (arg.rfind("--", 0) == 0) {
            std::string name = arg.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
                has_val = true;
            }
            for (const Opt& c : opts)
                if (name == c.lng) o = &c;
        } else if (arg.size() >= 2 && arg[0] == '-') {
            for (const Opt& c : opts)
                if (c.sht && arg[1] == c.sht) o = &c;
            if (o && arg.size() > 2) {
                val = arg.substr(2);
                has_val = true;
            }
        }
        if (!o) usage(("unexpected argument '" + arg + "'").c_str());
        if (o->kind == 0) {
            if (has_val) usage(("flag takes no value: " + arg).c_str());
            *static_cast<bool*>(o->dst) = true;
            continue;
        }
        if (!has_val) {
            if (i + 1 >= argc) usage(("a value is required for " + arg).c_str());
            val = argv[++i];
        }
        if (o->kind == 1) {
            *static_cast<std::string*>(o->dst) = val;
        } else if (o->kind == 2) {
            if (!to_usize(val, static_cast<int*>(o->dst)))
                usage(("invalid value '" + val + "' for " + arg).c_str());
        } else if (!parse_f32(val, static_cast<float*>(o->dst))) {
            usage(("invalid value '" + val + "' for " + arg).c_str());
        }
        if (o->seen) *o->seen = true;
    }
    if (a.input.empty() || a.output.empty() || a.tilt_file.empty())
        usage("the following required arguments were not provided: -i, -o, -t");
    return a;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

int main(int argc, char** argv) {
    const Args a = parse_args(argc, argv);

    // tilt angles (main.rs:215-218)
    const int nt = tiltgpu_read_tilt_file(a.tilt_file.c_str(), nullptr, 0);
    if (nt < 0) {
        std::fprintf(stderr, "Error reading tilt file: %s\n", a.tilt_file.c_str());
        return 1;
    }
    std::vector<float> tilts(std::max(nt, 1));
    tiltgpu_read_tilt_file(a.tilt_file.c_str(), tilts.data(), nt);

    // defocus file (main.rs:221-226)
    std::vector<int32_t> dviews;
    std::vector<float> dent;
    const bool have_file = !a.defocus_file.empty();
    if (have_file) {
        const int n = tiltgpu_read_defocus_file(a.defocus_file.c_str(), nullptr, nullptr, 0);
        if (n < 0) {
            if (n <= -100)
                std::fprintf(stderr, "Error: Defocus file line %d is malformed\n", -100 - n);
            else
                std::fprintf(stderr, "Error: Cannot read defocus file %s\n", a.defocus_file.c_str());
            return 1;
        }
        dviews.resize(std::max(n, 1));
        dent.resize(3 * std::max(n, 1));
        tiltgpu_read_defocus_file(a.defocus_file.c_str(), dviews.data(), dent.data(), n);
        dviews.resize(n);
        dent.resize(3 * n);
    }

    MrcReader reader;
    std::string err;
    if (!reader.open(a.input, &err)) {
        std::fprintf(stderr, "Error: %s\n", err.c_str());
        return 1;
    }
    const MrcHeader& h = reader.header();
    const int nx = h.nx, ny = h.ny, nz = h.nz;
    // pixel_size_x (crates/imod-mrc/src/header.rs:150-152)
    const float pixel = a.has_pixel ? a.pixel : (h.mx > 0 ? h.xlen / static_cast<float>(h.mx) : 1.0f);

    // interpolate_defocus over the stack's views (main.rs:239-248)
    std::vector<float> vdef;
    int n_vdef = 0;
    if (have_file) {
        if (static_cast<int>(dviews.size()) < nz)
            std::fprintf(stderr,
                         "ctfphaseflip: defocus file has %zu entries for %d views, interpolating "
                         "missing values\n",
                         dviews.size(), nz);
        vdef.resize(3 * static_cast<size_t>(std::max(nz, 1)));
        n_vdef = tiltgpu_interpolate_defocus(dviews.data(), dent.data(),
                                             static_cast<int32_t>(dviews.size()), nz, vdef.data());
        if (n_vdef < 0) n_vdef = 0;
    }
    if (nt < nz)
        std::fprintf(stderr, "Warning: tilt file has %d angles but stack has %d sections\n", nt, nz);

    tiltgpu_ctf_params p;
    tiltgpu_ctf_default_params(&p);
    p.voltage_kv = a.voltage;
    p.cs_mm = a.cs;
    p.pixel_a = pixel;
    p.amplitude_contrast = a.amp;
    p.cuton_freq = a.cuton;
    p.plate_phase_deg = a.plate;
    p.tilt_axis_deg = a.tilt_axis;
    p.defocus_nm = a.defocus;
    p.strip_width = a.strip_width;
    p.max_strip_width = a.has_max ? std::max(a.max_strip_width, 0) : 0;
    if (a.has_max && a.max_strip_width == 0) {
        std::fprintf(stderr, "Error: --max-strip-width must be at least 1\n");
        return 1;
    }
    const float* vptr = have_file ? vdef.data() : nullptr;
    const int width = tiltgpu_ctf_strip_width(&p, nx, vptr, n_vdef);
    if (width < 1) {
        std::fprintf(stderr, "Error: %s\n", tiltgpu_last_error());
        return 1;
    }
    std::fprintf(stderr, "ctfphaseflip: GPU acceleration enabled\n");
    std::fprintf(stderr,
                 "ctfphaseflip: voltage=%.0fkV, Cs=%.1fmm, pixel=%.2fA, strips=%dpx, "
                 "tilt_axis=%.1fdeg\n",
                 a.voltage, a.cs, pixel, width, a.tilt_axis);
    if (std::fabs(a.plate) > 1e-6f)
        std::fprintf(stderr, "ctfphaseflip: phase plate correction=%.1fdeg\n", a.plate);
    if (have_file)
        std::fprintf(stderr, "ctfphaseflip: using per-view defocus file with astigmatism support\n");
    else
        std::fprintf(stderr, "ctfphaseflip: defocus=%.0fnm\n", a.defocus);

    // the stack goes to the device in chunks of sections (about 128 MB) read
    // into two page-locked buffers on a worker thread while the previous chunk
    // uploads; the corrected sections come back the same way while the previous
    // chunk is written with the header statistics (main.rs:690-704)
    const double t0 = now_s();
    const size_t plane = static_cast<size_t>(nx) * ny;
    const int chunk = std::max(
        1, std::min(nz, a.upload_sections > 0 ? a.upload_sections
                                              : static_cast<int>((128u << 20) / (plane * 4) + 1)));
    const int nchunks = (nz + chunk - 1) / chunk;
    float* bufs[2] = {nullptr, nullptr};
    for (auto& b : bufs) {
        b = static_cast<float*>(tiltgpu_host_alloc(static_cast<int64_t>(plane * chunk * 4)));
        if (!b) {
            std::fprintf(stderr, "Error: %s\n", tiltgpu_last_error());
            return 1;
        }
    }
    auto release = [&] {
        for (auto& b : bufs) tiltgpu_host_free(b);
    };
    tiltgpu_ctf* c = nullptr;
    if (tiltgpu_ctf_create(nx, ny, nz, &p, tilts.data(), nt, vptr, n_vdef, &c) != 0) {
        std::fprintf(stderr, "ctfphaseflip: correction failed: %s\n", tiltgpu_last_error());
        release();
        return 1;
    }
    std::string perr;
    bool ok = pipeline2(
        nchunks, bufs,
        [&](int k, float* buf, std::string* e) {
            const int z0 = k * chunk, z1 = std::min(nz, z0 + chunk);
            for (int z = z0; z < z1; ++z) {
                size_t n = 0;
                if (!reader.read_slice_f32_into(z, buf + (z - z0) * plane, plane, &n, e) ||
                    n != plane) {
                    if (e->empty()) *e = "unsupported MRC mode";
                    return false;
                }
            }
            return true;
        },
        [&](int k, float* buf, std::string* e) {
            const int z0 = k * chunk, z1 = std::min(nz, z0 + chunk);
            if (tiltgpu_ctf_upload(c, z0, z1 - z0, buf) != 0) {
                *e = std::string("upload: ") + tiltgpu_last_error();
                return false;
            }
            return true;
        },
        &perr);
    const double t1 = now_s();
    double ms = 0;
    if (ok && tiltgpu_ctf_run(c) != 0) {
        perr = std::string("correction failed: ") + tiltgpu_last_error();
        ok = false;
    }
    if (ok) tiltgpu_ctf_last_run(c, &ms, nullptr);
    const double t2 = now_s();

    // output header (main.rs:318-327)
    MrcHeader oh = MrcHeader::make(nx, ny, nz, 2);
    oh.xlen = h.xlen;
    oh.ylen = h.ylen;
    oh.zlen = h.zlen;
    oh.mx = h.mx;
    oh.my = h.my;
    oh.mz = h.mz;
    oh.add_label("ctfphaseflip: CTF phase correction (2D strips)");
    MrcWriter w;
    if (ok && !w.create(a.output, oh, &perr)) ok = false;
    float gmin = 3.40282347e38f, gmax = -3.40282347e38f;
    double gsum = 0.0;
    if (ok)
        ok = pipeline2(
            nchunks, bufs,
            [&](int k, float* buf, std::string* e) {
                const int z0 = k * chunk, z1 = std::min(nz, z0 + chunk);
                if (tiltgpu_ctf_download(c, z0, z1 - z0, buf) != 0) {
                    *e = std::string("download: ") + tiltgpu_last_error();
                    return false;
                }
                return true;
            },
            [&](int k, float* buf, std::string* e) {
                const int z0 = k * chunk, z1 = std::min(nz, z0 + chunk);
                for (int z = z0; z < z1; ++z) {
                    const float* sec = buf + (z - z0) * plane;
                    float mn = 3.40282347e38f, mx = -3.40282347e38f;
                    double sum = 0.0;
                    for (size_t i = 0; i < plane; ++i) {
                        if (sec[i] < mn) mn = sec[i];
                        if (sec[i] > mx) mx = sec[i];
                        sum += static_cast<double>(sec[i]);
                    }
                    const float mean = static_cast<float>(sum / static_cast<double>(plane));
                    if (mn < gmin) gmin = mn;
                    if (mx > gmax) gmax = mx;
                    gsum += static_cast<double>(mean) * static_cast<double>(plane);
                    if (!w.write_slice_f32(sec, plane, e)) return false;
                }
                return true;
            },
            &perr);
    tiltgpu_ctf_destroy(c);
    release();
    if (ok && !w.finish(gmin, gmax, static_cast<float>(gsum / (static_cast<double>(plane) * nz)),
                        &perr))
        ok = false;
    if (!ok) {
        std::fprintf(stderr, "Error: %s\n", perr.c_str());
        return 1;
    }
    const double t3 = now_s();
    std::fprintf(stderr, "ctfphaseflip: corrected %d sections (2D strips)\n", nz);
    std::fprintf(stderr,
                 "ctfphaseflip: wall time read+upload %.3f s, correction %.3f s (device %.3f s), "
                 "read-back+write %.3f s\n",
                 t1 - t0, t2 - t1, ms / 1000.0, t3 - t2);
    return 0;
}

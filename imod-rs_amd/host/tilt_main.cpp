// `tilt` -- the drop-in front end: the option surface and file handling of
// imod-rs apps/tilt/src/main.rs:18-120,540-1095, with the reconstruction run on
// the GPU through libtiltgpu (include/tiltgpu.h).  The stack is read in chunks
// of sections while the previous chunk uploads, and output rows come back in
// slabs while the previous slab is written: two page-locked buffers each way,
// a worker thread on one side of each pair.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tiltgpu.h"
#include "mrc.h"
#include "pipeline.h"
#include "textio.h"

using namespace tiltgpu_host;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Args {
    std::string input, output, tilt_file;
    int thickness = 0;
    int width = -1;
    int threads = -1;
    bool gpu = false;
    std::string radial = "0.35,0.05";
    bool no_weight = false;
    int sirt = 0;
    float sirt_relax = 1.0f;
    bool no_cosine_stretch = false;
    std::string local_file, xtilt_file, zfactor_file;
    bool log = false;
    float densweight = 1.0f, outscale = 1.0f, outadd = 0.0f;
    int exact_filter_size = 0;
    std::string fill_mode = "mean";
    std::string cosine_interp = "linear";
    float max_cosine_tilt = 80.0f;
    bool hamming = false;
    int fake_sirt = 0;
    std::string storage = "f32";
    std::string geometry = "f32";
    std::string xf_file;
    float xf_fill = 0.0f;
    int slab_rows = 64;
    int upload_views = 0;
};

[[noreturn]] void usage(const char* msg) {
    if (msg) std::fprintf(stderr, "error: %s\n\n", msg);
    std::fprintf(stderr,
                 "Weighted back-projection reconstruction (MI355X)\n\n"
                 "Usage: tilt -i <INPUT> -o <OUTPUT> -t <TILT_FILE> [OPTIONS]\n\n"
                 "  -i, --input <f>            aligned tilt series (MRC)\n"
                 "  -o, --output <f>           output reconstruction (MRC)\n"
                 "  -t, --tilt-file <f>        tilt angle file (.tlt)\n"
                 "  -z, --thickness <n>        thickness in Z (default: input width)\n"
                 "  -w, --width <n>            output width (default: input width)\n"
                 "  -j, --threads <n>          host threads (accepted; the GPU does the work)\n"
                 "      --gpu                  accepted (the reconstruction always runs on the GPU)\n"
                 "      --radial <c,f>         R-weighting cutoff,falloff [0.35,0.05]\n"
                 "      --no-weight            disable R-weighting\n"
                 "      --sirt <n>             SIRT iterations [0]\n"
                 "      --sirt-relax <r>       SIRT relaxation [1.0]\n"
                 "      --no-cosine-stretch    disable cosine stretching\n"
                 "      --local-file <f>       local alignment (view x y dx dy per line)\n"
                 "      --xtilt-file <f>       X-axis tilt per view (degrees)\n"
                 "      --zfactor-file <f>     Z scaling factor per view\n"
                 "      --log                  log10 transform of projections\n"
                 "      --densweight <v>       [1.0]   --outscale <v> [1.0]   --outadd <v> [0.0]\n"
                 "      --exact-filter-size <n>\n"
                 "      --fill-mode <m>        mean | zero | edge [mean]\n"
                 "      --cosine-interp <m>    linear | cubic [linear]\n"
                 "      --max-cosine-tilt <d>  [80.0]\n"
                 "      --hamming              Hamming window instead of Gaussian roll-off\n"
                 "      --fake-sirt <n>        low-pass passes after reconstruction [0]\n"
                 "      --storage <s>          f32 | f16 | bf16 projection storage in HBM [f32]\n"
                 "      --xf-file <f>          apply the .xf section transforms first (newstack -x,\n"
                 "                             on the GPU)\n"
                 "      --xf-fill <v>          fill value outside transformed sections [0]\n"
                 "      --geometry <g>         f32 (the reference's tables) | f64 (computed in\n"
                 "                             FP64 on the GPU, rounded once) [f32]\n"
                 "      --slab-rows <n>        output rows per device->file transfer [64]\n"
                 "      --upload-views <n>     sections per file->device transfer [about 128 MB]\n");
    std::exit(2);
}

bool to_int(const std::string& s, int* v) {
    long long x;
    if (s.size() > 1 && s[0] == '-' && parse_usize(s.substr(1), &x)) {
        *v = static_cast<int>(-x);
        return true;
    }
    if (!parse_usize(s, &x)) return false;
    *v = static_cast<int>(x);
    return true;
}

Args parse_args(int argc, char** argv) {
    Args a;
    struct Opt {
        const char* lng;
        char sht;
        int kind;  // 0 flag, 1 string, 2 int, 3 float
        void* dst;
    };
    const Opt opts[] = {
        {"input", 'i', 1, &a.input},
        {"output", 'o', 1, &a.output},
        {"tilt-file", 't', 1, &a.tilt_file},
        {"thickness", 'z', 2, &a.thickness},
        {"width", 'w', 2, &a.width},
        {"threads", 'j', 2, &a.threads},
        {"gpu", 0, 0, &a.gpu},
        {"radial", 0, 1, &a.radial},
        {"no-weight", 0, 0, &a.no_weight},
        {"sirt", 0, 2, &a.sirt},
        {"sirt-relax", 0, 3, &a.sirt_relax},
        {"no-cosine-stretch", 0, 0, &a.no_cosine_stretch},
        {"local-file", 0, 1, &a.local_file},
        {"xtilt-file", 0, 1, &a.xtilt_file},
        {"zfactor-file", 0, 1, &a.zfactor_file},
        {"log", 0, 0, &a.log},
        {"densweight", 0, 3, &a.densweight},
        {"outscale", 0, 3, &a.outscale},
        {"outadd", 0, 3, &a.outadd},
        {"exact-filter-size", 0, 2, &a.exact_filter_size},
        {"fill-mode", 0, 1, &a.fill_mode},
        {"cosine-interp", 0, 1, &a.cosine_interp},
        {"max-cosine-tilt", 0, 3, &a.max_cosine_tilt},
        {"hamming", 0, 0, &a.hamming},
        {"fake-sirt", 0, 2, &a.fake_sirt},
        {"storage", 0, 1, &a.storage},
        {"geometry", 0, 1, &a.geometry},
        {"xf-file", 0, 1, &a.xf_file},
        {"xf-fill", 0, 3, &a.xf_fill},
        {"slab-rows", 0, 2, &a.slab_rows},
        {"upload-views", 0, 2, &a.upload_views},
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
            if (!to_int(val, static_cast<int*>(o->dst)))
                usage(("invalid integer '" + val + "' for " + arg).c_str());
        } else {
            if (!parse_f32(val, static_cast<float*>(o->dst)))
                usage(("invalid number '" + val + "' for " + arg).c_str());
        }
    }
    if (a.input.empty() || a.output.empty() || a.tilt_file.empty())
        usage("the following required arguments were not provided: -i, -o, -t");
    return a;
}

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

std::vector<float> read_values(const std::string& path, bool per_view, const char* what) {
    const int n = per_view ? tiltgpu_read_per_view_file(path.c_str(), nullptr, 0)
                           : tiltgpu_read_tilt_file(path.c_str(), nullptr, 0);
    if (n < 0) {
        std::fprintf(stderr, "Error reading %s: %s\n", what, path.c_str());
        std::exit(1);
    }
    std::vector<float> v(n);
    if (per_view)
        tiltgpu_read_per_view_file(path.c_str(), v.data(), n);
    else
        tiltgpu_read_tilt_file(path.c_str(), v.data(), n);
    return v;
}

void check(int rc, const char* what) {
    if (rc < 0) {
        std::fprintf(stderr, "tilt: %s failed: %s\n", what, tiltgpu_last_error());
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char** argv) {
    const Args a = parse_args(argc, argv);
    const std::vector<float> angles = read_values(a.tilt_file, false, "tilt file");

    MrcReader reader;
    std::string err;
    if (!reader.open(a.input, &err)) {
        std::fprintf(stderr, "Error: %s\n", err.c_str());
        return 1;
    }
    const MrcHeader& h = reader.header();
    const int in_nx = h.nx, in_ny = h.ny, nz = h.nz;
    const int out_nx = a.width >= 0 ? a.width : in_nx;
    const int out_ny = in_ny;
    const int out_nz = a.thickness > 0 ? a.thickness : in_nx;
    std::fprintf(stderr, "tilt: %d projections, %d x %d -> %d x %d x %d (GPU)\n", nz, in_nx,
                 in_ny, out_nx, out_ny, out_nz);

    // --radial "c,f" (main.rs:580-585): unparsable parts keep their defaults
    float cutoff = 0.35f, falloff = 0.05f;
    {
        const std::string& r = a.radial;
        const size_t comma = r.find(',');
        float v;
        if (parse_f32(trim(r.substr(0, comma)), &v)) cutoff = v;
        if (comma != std::string::npos) {
            const std::string rest = r.substr(comma + 1);
            if (parse_f32(trim(rest.substr(0, rest.find(','))), &v)) falloff = v;
        }
    }
    tiltgpu_options o;
    tiltgpu_default_options(&o);
    o.radial_cutoff = cutoff;
    o.radial_falloff = falloff;
    o.no_weight = a.no_weight;
    o.exact_filter_size = a.exact_filter_size;
    o.hamming = a.hamming;
    o.no_cosine_stretch = a.no_cosine_stretch;
    o.cosine_cubic = lower(a.cosine_interp) == "cubic";
    o.max_cosine_tilt = a.max_cosine_tilt;
    o.log = a.log;
    const std::string fm = lower(a.fill_mode);  // unknown modes mean "mean" (main.rs:590-594)
    o.fill_mode = fm == "zero" ? TILTGPU_FILL_ZERO : fm == "edge" ? TILTGPU_FILL_EDGE : TILTGPU_FILL_MEAN;
    o.densweight = a.densweight;
    o.outscale = a.outscale;
    o.outadd = a.outadd;
    o.sirt = a.sirt;
    o.sirt_relax = a.sirt_relax;
    o.fake_sirt = a.fake_sirt;
    const std::string st = lower(a.storage);
    o.storage = st == "f16" ? TILTGPU_STORE_F16 : st == "bf16" ? TILTGPU_STORE_BF16 : TILTGPU_STORE_F32;
    o.geometry = lower(a.geometry) == "f64" ? TILTGPU_GEOMETRY_F64 : TILTGPU_GEOMETRY_F32;

    tiltgpu_dims d{in_nx, in_ny, nz, out_nx, out_nz};
    tiltgpu_session* S = nullptr;
    check(tiltgpu_create(&d, &o, &S), "session creation");
    check(tiltgpu_set_angles(S, angles.data(), static_cast<int32_t>(angles.size())), "angles");
    if (!a.local_file.empty()) {
        std::fprintf(stderr, "tilt: loading local alignment from %s\n", a.local_file.c_str());
        LocalGrids g;
        if (!load_local_grids(a.local_file, nz, &g, &err)) {
            std::fprintf(stderr, "Error reading local alignment file: %s\n", err.c_str());
            return 1;
        }
        check(tiltgpu_set_local_alignment(S, g.nx.data(), g.ny.data(), g.xs.data(), g.ys.data(),
                                          g.dx.data(), g.dy.data()),
              "local alignment");
    }
    if (!a.xtilt_file.empty()) {
        const auto xt = read_values(a.xtilt_file, true, "xtilt file");
        check(tiltgpu_set_xtilt(S, xt.data(), static_cast<int32_t>(xt.size())), "xtilt");
    }
    if (!a.zfactor_file.empty()) {
        const auto zf = read_values(a.zfactor_file, true, "zfactor file");
        check(tiltgpu_set_zfactor(S, zf.data(), static_cast<int32_t>(zf.size())), "zfactor");
    }
    if (!a.xf_file.empty()) {
        const int n = tiltgpu_read_xf_file(a.xf_file.c_str(), nullptr, 0);
        if (n <= 0) {
            std::fprintf(stderr, "Error reading transform file %s\n", a.xf_file.c_str());
            return 1;
        }
        std::vector<float> xf(6 * static_cast<size_t>(n));
        tiltgpu_read_xf_file(a.xf_file.c_str(), xf.data(), n);
        check(tiltgpu_set_section_transforms(S, xf.data(), n, a.xf_fill), "section transforms");
    }

    // projections (main.rs:601-605): chunks of sections read into page-locked
    // buffers on a worker thread while the previous chunk uploads
    const double t_start = now_s();
    const size_t plane = static_cast<size_t>(in_nx) * in_ny;
    const int chunk_views = std::max(
        1, std::min(nz, a.upload_views > 0 ? a.upload_views
                                           : static_cast<int>((128u << 20) / (plane * 4) + 1)));
    const int in_chunks = (nz + chunk_views - 1) / chunk_views;
    float* inbuf[2] = {nullptr, nullptr};
    for (auto& b : inbuf) {
        b = static_cast<float*>(tiltgpu_host_alloc(static_cast<int64_t>(chunk_views * plane * 4)));
        if (!b) {
            std::fprintf(stderr, "Error: %s\n", tiltgpu_last_error());
            return 1;
        }
    }
    {
        std::string perr;
        const bool ok = pipeline2(
            in_chunks, inbuf,
            [&](int k, float* buf, std::string* e) {
                const int v0 = k * chunk_views, v1 = std::min(nz, v0 + chunk_views);
                for (int z = v0; z < v1; ++z) {
                    size_t n = 0;
                    if (!reader.read_slice_f32_into(z, buf + (z - v0) * plane, plane, &n, e) ||
                        n != plane) {
                        if (e->empty()) *e = "unsupported MRC mode";
                        return false;
                    }
                }
                return true;
            },
            [&](int k, float* buf, std::string* e) {
                const int v0 = k * chunk_views, v1 = std::min(nz, v0 + chunk_views);
                if (tiltgpu_upload_views(S, v0, v1 - v0, buf) != 0) {
                    *e = std::string("upload: ") + tiltgpu_last_error();
                    return false;
                }
                return true;
            },
            &perr);
        for (auto& b : inbuf) tiltgpu_host_free(b);
        if (!ok) {
            std::fprintf(stderr, "Error: %s\n", perr.c_str());
            return 1;
        }
    }
    const double t_read = now_s();
    check(tiltgpu_run(S), "reconstruction");
    const double t_run = now_s();

    // output header and sections (main.rs:710-719, 1077-1093)
    MrcHeader oh = MrcHeader::make(out_nx, out_nz, out_ny, 2);
    oh.xlen = h.xlen * static_cast<float>(out_nx) / static_cast<float>(in_nx);
    oh.ylen = h.xlen * static_cast<float>(out_nz) / static_cast<float>(in_nx);
    oh.zlen = h.ylen;
    oh.mx = out_nx;
    oh.my = out_nz;
    oh.mz = out_ny;
    oh.add_label("tilt: " + std::to_string(nz) + " projections, thickness " + std::to_string(out_nz));
    MrcWriter writer;
    if (!writer.create(a.output, oh, &err)) {
        std::fprintf(stderr, "Error: %s\n", err.c_str());
        return 1;
    }
    // output rows come back in slabs into page-locked buffers on a worker
    // thread while the previous slab is written (main.rs:1077-1093)
    const size_t per = static_cast<size_t>(out_nx) * out_nz;
    const int slab = std::max(1, std::min(out_ny, a.slab_rows > 0 ? a.slab_rows : 64));
    const int out_chunks = (out_ny + slab - 1) / slab;
    float* outbuf[2] = {nullptr, nullptr};
    for (auto& b : outbuf) {
        b = static_cast<float*>(tiltgpu_host_alloc(static_cast<int64_t>((per + 3) * slab * 4)));
        if (!b) {
            std::fprintf(stderr, "Error: %s\n", tiltgpu_last_error());
            return 1;
        }
    }
    float gmin = 3.40282347e38f, gmax = -3.40282347e38f;
    double gsum = 0.0;
    {
        std::string perr;
        const bool ok = pipeline2(
            out_chunks, outbuf,
            [&](int k, float* buf, std::string* e) {
                const int y0 = k * slab, y1 = std::min(out_ny, y0 + slab);
                // the slab's per-row statistics follow its rows in the buffer
                if (tiltgpu_read_rows(S, y0, y1, buf, buf + per * slab) != 0) {
                    *e = std::string("read rows: ") + tiltgpu_last_error();
                    return false;
                }
                return true;
            },
            [&](int k, float* buf, std::string* e) {
                const int y0 = k * slab, y1 = std::min(out_ny, y0 + slab);
                const float* stats = buf + per * slab;
                for (int r = 0; r < y1 - y0; ++r) {
                    const float smin = stats[3 * r], smax = stats[3 * r + 1], smean = stats[3 * r + 2];
                    if (smin < gmin) gmin = smin;
                    if (smax > gmax) gmax = smax;
                    gsum += static_cast<double>(smean) * static_cast<double>(per);
                    if (!writer.write_slice_f32(buf + r * per, per, e)) return false;
                }
                std::fprintf(stderr, "  %d/%d rows\n", y1, out_ny);
                return true;
            },
            &perr);
        for (auto& b : outbuf) tiltgpu_host_free(b);
        if (!ok) {
            std::fprintf(stderr, "Error: %s\n", perr.c_str());
            return 1;
        }
    }
    const float gmean = static_cast<float>(gsum / (static_cast<double>(per) * out_ny));
    if (!writer.finish(gmin, gmax, gmean, &err)) {
        std::fprintf(stderr, "Error: %s\n", err.c_str());
        return 1;
    }
    const double t_end = now_s();
    tiltgpu_timing t;
    if (tiltgpu_get_timing(S, &t) == 0)
        std::fprintf(stderr, "tilt: device time filter %.2f ms, backprojection %.2f ms, "
                             "forward projection %.2f ms\n", t.filter_ms, t.bp_ms, t.fp_ms);
    std::fprintf(stderr, "tilt: wall time read+upload %.3f s, reconstruction %.3f s, "
                         "read-back+write %.3f s, total %.3f s\n",
                 t_read - t_start, t_run - t_read, t_end - t_run, t_end - t_start);
    tiltgpu_destroy(S);
    std::fprintf(stderr, "tilt: reconstruction complete -> %s\n", a.output.c_str());
    return 0;
}

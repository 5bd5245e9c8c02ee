// Continuous-batching reconstruction server.
//
// Requests -- whole reconstructions (WBP or SIRT) of series of any shape, and
// row slabs of series kept resident in HBM -- are queued by any host thread.
// One worker thread owns the device.  It admits queued requests in FIFO order
// while their device memory fits the budget, then runs rounds.  A round issues
// one batched launch per stage across every request in flight:
//   forward projection (SIRT iterations) -> filter (new series: R-weighting +
//   stretch; SIRT: residual O - R) -> backprojection (persistent, per-XCD
//   tile queues over the tiles of all requests).
// Requests that arrive while a round is on the device join the next round.  A
// WBP request needs one round, a SIRT-k request k+1, a slab request one.
//
// Tile shapes and every per-tile computation depend only on the request, so a
// request's result is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "session_internal.h"

using namespace tiltgpu_sess;

namespace {

using Clock = std::chrono::steady_clock;

enum { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3 };
enum { kRecon = 0, kSlab = 1, kSeries = 2 };

struct Req {
    int64_t id = 0;
    int kind = kRecon;
    tiltgpu_dims d{};
    tiltgpu_options o{};
    std::vector<float> angles, xtilt, zfac, proj;
    bool has_xtilt = false, has_zfac = false, has_local = false;
    std::vector<int32_t> lnx, lny;
    std::vector<float> lxs, lys, ldx, ldy;
    float* out = nullptr;  // caller's host buffer
    int series = -1, y0 = 0, y1 = 0;
    int64_t need = 0;
    tiltgpu_session* s = nullptr;
    float* slab = nullptr;  // device slab output
    int phase = 0;          // 0: first round; k: SIRT iteration k
    int state = kQueued;
    int rc = 0;
    std::string err;
    Clock::time_point t_submit, t_done;
};

int64_t recon_bytes(const tiltgpu_dims& d, const tiltgpu_options& o) {
    const int64_t np = static_cast<int64_t>(d.nz) * d.in_ny * d.in_nx * 4;
    const int64_t nv = static_cast<int64_t>(d.in_ny) * d.out_nz * d.out_nx * 4;
    int64_t b = 2 * np + nv + (1 << 20);
    if (o.log) b += np;
    if (o.storage != TILTGPU_STORE_F32) b += np / 2;
    if (o.sirt > 0) b += np;
    if (o.fake_sirt > 0) b += nv;
    return b;
}

struct BatchBufs {
    void* dev = nullptr;
    void* host = nullptr;
    size_t cap = 0;
};

int ensure_batch(BatchBufs& b, size_t bytes) {
    if (bytes <= b.cap) return TILTGPU_OK;
    if (b.dev) (void)hipFree(b.dev);
    if (b.host) (void)hipHostFree(b.host);
    b.dev = b.host = nullptr;
    b.cap = 0;
    const size_t cap = std::max<size_t>(bytes, 1 << 16);
    if (hipMalloc(&b.dev, cap) != hipSuccess || hipHostMalloc(&b.host, cap) != hipSuccess)
        return fail(TILTGPU_ERR_HIP, "server: descriptor buffer allocation failed");
    b.cap = cap;
    return TILTGPU_OK;
}

}  // namespace

struct tiltgpu_server {
    int device = 0;
    int n_cu = 256;
    int64_t budget = 0, used = 0;
    hipStream_t stream = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::deque<Req*> queue;
    std::vector<Req*> active;
    std::unordered_map<int64_t, Req*> reqs;
    std::vector<Req*> series;  // resident series (kind kSeries, state kDone)
    int64_t next_id = 1;
    BatchBufs bp_bufs, filt_bufs, fp_bufs;
    // statistics
    std::vector<double> lat_ms;
    int64_t completed = 0, failed = 0, rounds = 0, bp_launches = 0, bp_batched = 0;
    int max_active = 0;
};

namespace {

void finish(tiltgpu_server* S, Req* r, int rc, const char* msg) {
    r->rc = rc;
    if (rc < 0) r->err = msg ? msg : "";
    r->t_done = Clock::now();
}

// Allocate and fill the device state of an admitted request (worker thread).
int setup(tiltgpu_server* S, Req* r) {
    if (r->kind == kSlab) {
        Req* sr = S->series[r->series];
        const int64_t per = static_cast<int64_t>(sr->d.out_nz) * sr->d.out_nx;
        HIP_TRY(hipMalloc(&r->slab, static_cast<size_t>(r->y1 - r->y0) * per * 4));
        return TILTGPU_OK;
    }
    int rc = session_create(&r->d, &r->o, S->stream, &r->s);
    if (rc) return rc;
    tiltgpu_session* s = r->s;
    if ((rc = tiltgpu_set_angles(s, r->angles.data(), static_cast<int32_t>(r->angles.size()))))
        return rc;
    if (r->has_xtilt &&
        (rc = tiltgpu_set_xtilt(s, r->xtilt.data(), static_cast<int32_t>(r->xtilt.size()))))
        return rc;
    if (r->has_zfac &&
        (rc = tiltgpu_set_zfactor(s, r->zfac.data(), static_cast<int32_t>(r->zfac.size()))))
        return rc;
    if (r->has_local &&
        (rc = tiltgpu_set_local_alignment(s, r->lnx.data(), r->lny.data(), r->lxs.data(),
                                          r->lys.data(), r->ldx.data(), r->ldy.data())))
        return rc;
    if ((rc = prepare_pipeline_geometry(s))) return rc;
    if (r->kind == kSeries) dev_free(s, s->vol);  // slab requests bring their own outputs
    HIP_TRY(hipMemcpyAsync(s->raw, r->proj.data(), r->proj.size() * 4, hipMemcpyHostToDevice,
                           S->stream));
    s->uploaded = true;
    return TILTGPU_OK;
}

int grid_for_bp(tiltgpu_server* S, size_t lds) {
    // persistent launch: enough workgroups to fill every CU a few times over;
    // LDS per workgroup bounds the residency (160 KiB per CU)
    const int per_cu = std::max<int>(1, std::min<int>(8, static_cast<int>((160 * 1024) / std::max<size_t>(lds, 1))));
    return S->n_cu * per_cu;
}

struct BpItem {
    Req* r;
    BpParams p;
    int storage, two_row, local, tile_z;
    size_t lds;
};

// One persistent launch per kernel variant; all groups' descriptors go to the
// device in a single copy (the pinned staging is rewritten only after the
// round's stream synchronisation).
int launch_bp_groups(tiltgpu_server* S, std::vector<std::vector<BpItem>>& groups) {
    if (groups.empty()) return TILTGPU_OK;
    struct Layout {
        size_t descs, tiles, counters;
        int n, total;
        size_t lds;
    };
    std::vector<Layout> lay(groups.size());
    size_t bytes = 0;
    for (size_t g = 0; g < groups.size(); ++g) {
        const int n = static_cast<int>(groups[g].size());
        Layout& L = lay[g];
        L.n = n;
        L.descs = bytes;
        bytes += n * sizeof(BpParams);
        L.tiles = bytes;
        bytes += (n + 1) * sizeof(int);
        bytes = (bytes + 63) / 64 * 64;
        L.counters = bytes;
        bytes += 8 * sizeof(int);
        bytes = (bytes + 255) / 256 * 256;
    }
    int rc = ensure_batch(S->bp_bufs, bytes);
    if (rc) return rc;
    auto* h = static_cast<char*>(S->bp_bufs.host);
    for (size_t g = 0; g < groups.size(); ++g) {
        Layout& L = lay[g];
        auto* hp = reinterpret_cast<BpParams*>(h + L.descs);
        auto* ht = reinterpret_cast<int*>(h + L.tiles);
        auto* hc = reinterpret_cast<int*>(h + L.counters);
        int total = 0;
        size_t lds = 0;
        for (int i = 0; i < L.n; ++i) {
            hp[i] = groups[g][i].p;
            ht[i] = total;
            total += hp[i].ntx * hp[i].ntz * (hp[i].y1 - hp[i].y0);
            lds = std::max(lds, groups[g][i].lds);
        }
        // every series of the launch gets the launch's LDS capacity (a larger
        // capacity only moves the second staging buffer; results are unchanged)
        const int entry = groups[g][0].two_row ? 16 : 8;
        for (int i = 0; i < L.n; ++i) hp[i].lds_cap = static_cast<int>(lds / (2 * entry));
        ht[L.n] = total;
        for (int k = 0; k < 8; ++k) hc[k] = 0;
        L.total = total;
        L.lds = lds;
    }
    HIP_TRY(hipMemcpyAsync(S->bp_bufs.dev, h, bytes, hipMemcpyHostToDevice, S->stream));
    auto* d = static_cast<char*>(S->bp_bufs.dev);
    for (size_t g = 0; g < groups.size(); ++g) {
        const Layout& L = lay[g];
        BpBatch b{};
        b.descs = reinterpret_cast<const BpParams*>(d + L.descs);
        b.tile_start = reinterpret_cast<const int*>(d + L.tiles);
        b.nreq = L.n;
        b.total = L.total;
        b.chunk = (L.total + 7) / 8;
        b.counters = reinterpret_cast<int*>(d + L.counters);
        const BpItem& it = groups[g][0];
        HIP_TRY(launch_backproject_batch(b, it.storage, it.two_row, it.local, it.tile_z,
                                         grid_for_bp(S, L.lds), L.lds, S->stream));
        S->bp_launches++;
        S->bp_batched += L.n;
    }
    return TILTGPU_OK;
}

int run_round(tiltgpu_server* S, std::vector<Req*>& act) {
    std::vector<Req*> prep, iter, slabs;
    for (Req* r : act) {
        if (r->state != kActive) continue;
        if (r->kind == kSlab)
            slabs.push_back(r);
        else if (r->phase == 0)
            prep.push_back(r);
        else
            iter.push_back(r);
    }
    int rc;
    // 1. log + means of new series
    for (Req* r : prep)
        if ((rc = run_means(r->s))) return rc;
    // 2. forward projection of SIRT iterations (one-row: one batched launch)
    {
        std::vector<FpParams> g;
        for (Req* r : iter) {
            if (r->s->two_row) {
                if ((rc = run_forward(r->s))) return rc;
            } else {
                g.push_back(make_fp(r->s));
            }
        }
        if (!g.empty()) {
            const int n = static_cast<int>(g.size());
            const size_t off_b = n * sizeof(FpParams);
            const size_t bytes = off_b + (n + 1) * sizeof(int);
            if ((rc = ensure_batch(S->fp_bufs, bytes))) return rc;
            auto* h = static_cast<char*>(S->fp_bufs.host);
            auto* hp = reinterpret_cast<FpParams*>(h);
            auto* hb = reinterpret_cast<int*>(h + off_b);
            int total = 0;
            size_t lds = 0;
            for (int i = 0; i < n; ++i) {
                hp[i] = g[i];
                hb[i] = total;
                total += fp_gather_blocks(g[i]);
                lds = std::max(lds, fp_lds_bytes(g[i]));
            }
            hb[n] = total;
            HIP_TRY(hipMemcpyAsync(S->fp_bufs.dev, h, bytes, hipMemcpyHostToDevice, S->stream));
            FpBatch b{reinterpret_cast<const FpParams*>(S->fp_bufs.dev),
                      reinterpret_cast<const int*>(static_cast<char*>(S->fp_bufs.dev) + off_b), n};
            HIP_TRY(launch_forward_gather_batch(b, total, lds, S->stream));
        }
    }
    // 3. filter: new series (R-weighting + stretch) and SIRT residuals O - R
    {
        std::vector<FilterParams> f;
        std::vector<Req*> fr;
        for (Req* r : prep) {
            const float* O = r->s->o.log ? r->s->logd : r->s->raw;
            f.push_back(make_filter(r->s, O, nullptr, true));
            fr.push_back(r);
        }
        for (Req* r : iter) {
            const float* O = r->s->o.log ? r->s->logd : r->s->raw;
            FilterParams p = make_filter(r->s, O, r->s->sino, false);
            const int nblk = filter_blocks(p);
            if (r->s->partial_n < nblk) {
                dev_free(r->s, r->s->partial);
                if ((rc = dev_alloc(r->s, &r->s->partial, nblk))) return rc;
                r->s->partial_n = nblk;
            }
            p.sq_partial = r->s->partial;
            f.push_back(p);
            fr.push_back(r);
        }
        if (!f.empty()) {
            const int n = static_cast<int>(f.size());
            const size_t off_b = n * sizeof(FilterParams);
            const size_t bytes = off_b + (n + 1) * sizeof(int);
            if ((rc = ensure_batch(S->filt_bufs, bytes))) return rc;
            auto* h = static_cast<char*>(S->filt_bufs.host);
            auto* hp = reinterpret_cast<FilterParams*>(h);
            auto* hb = reinterpret_cast<int*>(h + off_b);
            int total = 0;
            size_t lds = 0;
            for (int i = 0; i < n; ++i) {
                hp[i] = f[i];
                hb[i] = total;
                total += filter_blocks(f[i]);
                lds = std::max(lds, filter_lds_bytes(f[i].n, f[i].pairs_per_wg));
            }
            hb[n] = total;
            HIP_TRY(hipMemcpyAsync(S->filt_bufs.dev, h, bytes, hipMemcpyHostToDevice, S->stream));
            FilterBatch b{reinterpret_cast<const FilterParams*>(S->filt_bufs.dev),
                          reinterpret_cast<const int*>(static_cast<char*>(S->filt_bufs.dev) + off_b),
                          n};
            HIP_TRY(launch_filter_batch(b, total, lds, S->stream));
            for (int i = 0; i < n; ++i) {
                tiltgpu_session* s = fr[i]->s;
                if (s->o.storage != TILTGPU_STORE_F32)
                    HIP_TRY(launch_convert_storage(s->filt, s->store, s->scale, s->absmax,
                                                   static_cast<int64_t>(s->d.in_nx) * s->d.in_ny,
                                                   s->d.nz, s->o.storage, S->stream));
                if (f[i].sq_partial)
                    HIP_TRY(launch_sum_partials(s->partial, filter_blocks(f[i]),
                                                s->resid + (fr[i]->phase - 1), 1, S->stream));
            }
        }
    }
    // 4. backprojection: one persistent launch per kernel variant
    {
        std::vector<BpItem> items;
        auto add = [&](Req* r, tiltgpu_session* s, int y0, int y1, int out_mode, float* out) {
            const void* proj = s->o.storage == TILTGPU_STORE_F32 ? static_cast<const void*>(s->filt)
                                                                 : s->store;
            BpItem it;
            it.r = r;
            it.p = make_bp(s, proj, s->scale, y0, y1, out_mode, out);
            it.storage = s->o.storage;
            it.two_row = s->two_row ? 1 : 0;
            it.local = s->has_local ? 1 : 0;
            it.tile_z = s->tile_z;
            it.lds = bp_lds_bytes(s->lds_cap, it.two_row);
            items.push_back(it);
        };
        for (Req* r : prep) {
            if (r->kind == kSeries) continue;  // a resident series is only filtered
            const bool fused = r->s->o.sirt == 0 && r->s->o.fake_sirt == 0;
            add(r, r->s, 0, r->s->d.in_ny, fused ? 1 : 0, r->s->vol);
        }
        for (Req* r : iter) add(r, r->s, 0, r->s->d.in_ny, 2, r->s->vol);
        for (Req* r : slabs) {
            Req* sr = S->series[r->series];
            add(r, sr->s, r->y0, r->y1, 1, r->slab);
        }
        std::stable_sort(items.begin(), items.end(), [](const BpItem& a, const BpItem& b) {
            return std::make_tuple(a.storage, a.two_row, a.local, a.tile_z) <
                   std::make_tuple(b.storage, b.two_row, b.local, b.tile_z);
        });
        std::vector<std::vector<BpItem>> groups;
        size_t i = 0;
        while (i < items.size()) {
            size_t j = i;
            std::vector<BpItem> grp;
            while (j < items.size() && items[j].storage == items[i].storage &&
                   items[j].two_row == items[i].two_row && items[j].local == items[i].local &&
                   items[j].tile_z == items[i].tile_z)
                grp.push_back(items[j++]);
            groups.push_back(std::move(grp));
            i = j;
        }
        if ((rc = launch_bp_groups(S, groups))) return rc;
    }
    // 5. advance; finished requests: fake-SIRT, output transform, copy out
    for (Req* r : prep) r->phase = 1;
    for (Req* r : iter) r->phase++;
    for (Req* r : act) {
        if (r->state != kActive) continue;
        if (r->kind == kSlab) {
            Req* sr = S->series[r->series];
            const int64_t per = static_cast<int64_t>(sr->d.out_nz) * sr->d.out_nx;
            HIP_TRY(hipMemcpyAsync(r->out, r->slab, (r->y1 - r->y0) * per * 4,
                                   hipMemcpyDeviceToHost, S->stream));
            continue;
        }
        tiltgpu_session* s = r->s;
        if (r->phase <= s->o.sirt) continue;  // more SIRT iterations to go
        s->sirt_done = s->o.sirt;
        if (r->kind == kSeries) continue;  // stays resident for slab requests
        const int ny = s->d.in_ny, onz = s->d.out_nz, onx = s->d.out_nx;
        for (int k = 0; k < s->o.fake_sirt; ++k) {
            HIP_TRY(launch_lowpass_x(s->vol, s->vol_tmp, ny, onz, onx, S->stream));
            HIP_TRY(launch_lowpass_z(s->vol_tmp, s->vol, ny, onz, onx, S->stream));
            HIP_TRY(launch_lowpass_y(s->vol, s->vol_tmp, ny, onz, onx, S->stream));
            std::swap(s->vol, s->vol_tmp);
        }
        if (s->o.sirt > 0 || s->o.fake_sirt > 0)
            HIP_TRY(launch_transform(s->vol, static_cast<int64_t>(ny) * onz * onx,
                                     s->o.densweight, s->o.outscale, s->o.outadd, S->stream));
        HIP_TRY(hipMemcpyAsync(r->out, s->vol, static_cast<size_t>(ny) * onz * onx * 4,
                               hipMemcpyDeviceToHost, S->stream));
    }
    HIP_TRY(hipStreamSynchronize(S->stream));
    S->rounds++;
    return TILTGPU_OK;
}

bool round_finished(const Req* r) {
    if (r->kind == kSlab) return true;
    return r->phase > r->s->o.sirt;
}

void release(tiltgpu_server* S, Req* r) {
    if (r->kind == kSeries) return;
    if (r->s) {
        tiltgpu_destroy(r->s);
        r->s = nullptr;
    }
    if (r->slab) {
        (void)hipFree(r->slab);
        r->slab = nullptr;
    }
    r->proj.clear();
    r->proj.shrink_to_fit();
}

void worker_main(tiltgpu_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<Req*> admitted;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->active.empty(); });
            if (S->stop && S->active.empty() && S->queue.empty()) break;
            // FIFO admission while the device-memory budget allows
            while (!S->queue.empty()) {
                Req* r = S->queue.front();
                if (S->used + r->need > S->budget) {
                    if (S->active.empty() && admitted.empty()) {
                        // nothing in flight will free memory: only resident series
                        // hold it, so this request can never be admitted
                        S->queue.pop_front();
                        finish(S, r, TILTGPU_ERR_BUDGET,
                               "request does not fit the budget left by resident series");
                        r->state = kFailed;
                        S->failed++;
                        S->cv_done.notify_all();
                        continue;
                    }
                    break;
                }
                S->queue.pop_front();
                S->used += r->need;
                admitted.push_back(r);
            }
        }
        for (Req* r : admitted) {
            int rc = setup(S, r);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) {
                finish(S, r, rc, tiltgpu_last_error());
                r->state = kFailed;
                release(S, r);
                S->used -= r->need;
                S->failed++;
                S->cv_done.notify_all();
            } else {
                r->state = kActive;
                S->active.push_back(r);
            }
        }
        std::vector<Req*> act;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            act = S->active;
            S->max_active = std::max<int>(S->max_active, static_cast<int>(act.size()));
        }
        if (act.empty()) continue;
        int rc = run_round(S, act);
        std::string msg = rc ? tiltgpu_last_error() : "";
        if (rc == 0) {
            for (Req* r : act)
                if (r->s && (rc = check_kernel_error(r->s))) {
                    msg = tiltgpu_last_error();
                    break;
                }
        }
        std::lock_guard<std::mutex> lk(S->mu);
        std::vector<Req*> keep;
        for (Req* r : act) {
            if (rc == 0 && !round_finished(r)) {
                keep.push_back(r);
                continue;
            }
            // a failed round fails every request that was in it (the device
            // state of each is unknown); other queued requests are unaffected
            finish(S, r, rc, msg.c_str());
            r->state = rc ? kFailed : kDone;
            release(S, r);
            if (r->kind != kSeries) S->used -= r->need;
            if (rc)
                S->failed++;
            else if (r->kind != kSeries) {
                S->completed++;
                S->lat_ms.push_back(
                    std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
            }
        }
        S->active = keep;
        S->cv_done.notify_all();
    }
}

Req* new_req(tiltgpu_server* S) {
    Req* r = new Req();
    r->id = S->next_id++;
    r->t_submit = Clock::now();
    S->reqs[r->id] = r;
    return r;
}

int enqueue(tiltgpu_server* S, Req* r, int64_t* id) {
    if (r->need > S->budget) {
        const int64_t need = r->need;
        S->reqs.erase(r->id);
        delete r;
        return fail(TILTGPU_ERR_BUDGET, "request needs %lld bytes, budget is %lld",
                    static_cast<long long>(need), static_cast<long long>(S->budget));
    }
    if (id) *id = r->id;
    S->queue.push_back(r);
    S->cv_work.notify_one();
    return TILTGPU_OK;
}

int fill_recon(Req* r, const tiltgpu_dims* d, const tiltgpu_options* o, const float* angles,
               int32_t n_angles, const float* projections) {
    int rc = validate_dims(d);
    if (rc) return rc;
    if (!angles || n_angles < 1 || !projections) return fail(TILTGPU_ERR_ARG, "bad arguments");
    r->d = *d;
    if (o)
        r->o = *o;
    else
        tiltgpu_default_options(&r->o);
    if (r->o.storage < 0 || r->o.storage > 2) return fail(TILTGPU_ERR_ARG, "unknown storage");
    r->angles.assign(angles, angles + n_angles);
    const size_t np = static_cast<size_t>(d->nz) * d->in_ny * d->in_nx;
    r->proj.assign(projections, projections + np);
    r->need = recon_bytes(r->d, r->o);
    return TILTGPU_OK;
}

}  // namespace

extern "C" {

int tiltgpu_server_create(int64_t budget_bytes, tiltgpu_server** out) {
    if (!out || budget_bytes <= 0) return fail(TILTGPU_ERR_ARG, "bad arguments");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(TILTGPU_ERR_NODEV, "no GPU visible to HIP");
    auto* S = new tiltgpu_server();
    S->budget = budget_bytes;
    S->series.reserve(256);
    (void)hipGetDevice(&S->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, S->device) == hipSuccess) S->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return fail(TILTGPU_ERR_HIP, "server stream creation failed");
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return TILTGPU_OK;
}

int tiltgpu_server_submit(tiltgpu_server* S, const tiltgpu_dims* d, const tiltgpu_options* o,
                          const float* angles, int32_t n_angles, const float* projections,
                          float* out, int64_t* id) {
    if (!S || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::unique_ptr<Req> tmp(new Req());
    int rc = fill_recon(tmp.get(), d, o, angles, n_angles, projections);
    if (rc) return rc;
    if (!tmp->o.no_cosine_stretch && n_angles < d->nz)
        return fail(TILTGPU_ERR_ARG, "cosine stretch needs one tilt angle per section");
    std::lock_guard<std::mutex> lk(S->mu);
    Req* r = new_req(S);
    const int64_t rid = r->id;
    const auto t = r->t_submit;
    *r = std::move(*tmp);
    r->id = rid;
    r->t_submit = t;
    r->kind = kRecon;
    r->out = out;
    return enqueue(S, r, id);
}

int tiltgpu_server_add_series(tiltgpu_server* S, const tiltgpu_dims* d, const tiltgpu_options* o,
                              const float* angles, int32_t n_angles, const float* xtilt,
                              int32_t n_xtilt, const int32_t* lnx, const int32_t* lny,
                              const float* lxs, const float* lys, const float* ldx,
                              const float* ldy, const float* projections, int32_t* series_id) {
    if (!S || !series_id) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::unique_ptr<Req> tmp(new Req());
    int rc = fill_recon(tmp.get(), d, o, angles, n_angles, projections);
    if (rc) return rc;
    if (tmp->o.sirt != 0 || tmp->o.fake_sirt != 0)
        return fail(TILTGPU_ERR_ARG, "a resident series serves WBP slabs (sirt = fake_sirt = 0)");
    if (xtilt && n_xtilt > 0) {
        tmp->xtilt.assign(xtilt, xtilt + n_xtilt);
        tmp->has_xtilt = true;
    }
    if (lnx && lny) {
        const int nz = d->nz;
        tmp->lnx.assign(lnx, lnx + nz);
        tmp->lny.assign(lny, lny + nz);
        int64_t nxs = 0, nys = 0, ng = 0;
        for (int v = 0; v < nz; ++v) {
            nxs += lnx[v];
            nys += lny[v];
            ng += static_cast<int64_t>(lnx[v]) * lny[v];
        }
        tmp->lxs.assign(lxs, lxs + nxs);
        tmp->lys.assign(lys, lys + nys);
        tmp->ldx.assign(ldx, ldx + ng);
        tmp->ldy.assign(ldy, ldy + ng);
        tmp->has_local = true;
    }
    // a resident series keeps projections and their filtered copy, not a volume
    tmp->need -= static_cast<int64_t>(d->in_ny) * d->out_nz * d->out_nx * 4;
    int64_t id = 0;
    Req* r = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r = new_req(S);
        const int64_t rid = r->id;
        const auto t = r->t_submit;
        *r = std::move(*tmp);
        r->id = rid;
        r->t_submit = t;
        r->kind = kSeries;
        if ((rc = enqueue(S, r, &id))) return rc;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return r->state == kDone || r->state == kFailed; });
    if (r->state == kFailed) return fail(r->rc, "%s", r->err.c_str());
    if (S->series.size() >= S->series.capacity())
        return fail(TILTGPU_ERR_BUDGET, "too many resident series");
    S->series.push_back(r);  // capacity is reserved: the worker's reads never see a move
    *series_id = static_cast<int32_t>(S->series.size() - 1);
    return TILTGPU_OK;
}

int tiltgpu_server_submit_slab(tiltgpu_server* S, int32_t series_id, int32_t y0, int32_t y1,
                               float* out, int64_t* id) {
    if (!S || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    if (series_id < 0 || series_id >= static_cast<int32_t>(S->series.size()))
        return fail(TILTGPU_ERR_ARG, "unknown series %d", series_id);
    Req* sr = S->series[series_id];
    if (y0 < 0 || y1 > sr->d.in_ny || y0 >= y1) return fail(TILTGPU_ERR_ARG, "bad row range");
    Req* r = new_req(S);
    r->kind = kSlab;
    r->series = series_id;
    r->y0 = y0;
    r->y1 = y1;
    r->out = out;
    r->need = static_cast<int64_t>(y1 - y0) * sr->d.out_nz * sr->d.out_nx * 4;
    return enqueue(S, r, id);
}

int tiltgpu_server_poll(tiltgpu_server* S, int64_t id) {
    if (!S) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(TILTGPU_ERR_ARG, "unknown request %lld", static_cast<long long>(id));
    Req* r = it->second;
    if (r->state == kDone) return 1;
    if (r->state == kFailed) return fail(r->rc, "%s", r->err.c_str());
    return 0;
}

int tiltgpu_server_wait(tiltgpu_server* S, int64_t id, double* latency_ms) {
    if (!S) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(TILTGPU_ERR_ARG, "unknown request %lld", static_cast<long long>(id));
    Req* r = it->second;
    S->cv_done.wait(lk, [&] { return r->state == kDone || r->state == kFailed; });
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int rc = r->rc;
    const std::string err = r->err;
    if (r->kind != kSeries) {
        S->reqs.erase(it);
        delete r;
    }
    if (rc < 0) return fail(rc, "%s", err.c_str());
    return TILTGPU_OK;
}

int tiltgpu_server_get_stats(tiltgpu_server* S, tiltgpu_server_stats* st) {
    if (!S || !st) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof *st);
    st->completed = S->completed;
    st->failed = S->failed;
    st->rounds = S->rounds;
    st->bp_launches = S->bp_launches;
    st->bp_batched_requests = S->bp_batched;
    st->max_active = S->max_active;
    st->active = static_cast<int32_t>(S->active.size());
    st->queued = static_cast<int32_t>(S->queue.size());
    st->bytes_in_use = S->used;
    st->budget_bytes = S->budget;
    if (!S->lat_ms.empty()) {
        std::vector<double> v = S->lat_ms;
        std::sort(v.begin(), v.end());
        auto q = [&](double f) {
            const double pos = f * (v.size() - 1);
            const size_t i = static_cast<size_t>(pos);
            const double t = pos - i;
            return i + 1 < v.size() ? v[i] * (1 - t) + v[i + 1] * t : v[i];
        };
        st->p50_ms = q(0.50);
        st->p99_ms = q(0.99);
        double s = 0;
        for (double x : v) s += x;
        st->mean_ms = s / v.size();
    }
    return TILTGPU_OK;
}

int tiltgpu_server_reset_stats(tiltgpu_server* S) {
    if (!S) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    S->lat_ms.clear();
    S->completed = S->failed = S->rounds = S->bp_launches = S->bp_batched = 0;
    S->max_active = 0;
    return TILTGPU_OK;
}

void tiltgpu_server_destroy(tiltgpu_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    for (Req* r : S->series) {
        if (r->s) tiltgpu_destroy(r->s);
        r->s = nullptr;
    }
    for (auto& kv : S->reqs) delete kv.second;
    for (Req* r : S->queue) (void)r;  // queued requests are owned by reqs
    for (BatchBufs* b : {&S->bp_bufs, &S->filt_bufs, &S->fp_bufs}) {
        if (b->dev) (void)hipFree(b->dev);
        if (b->host) (void)hipHostFree(b->host);
    }
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
}

}  // extern "C"

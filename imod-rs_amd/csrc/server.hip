// Continuous-batching reconstruction server.
//
// Requests -- whole reconstructions (WBP or SIRT) of series of any shape, and
// row slabs of series kept resident in HBM -- are queued by any host thread.
// A scheduler thread admits queued requests in FIFO order while their device
// memory fits the budget and issues rounds.  A round is one batched launch per
// stage across every request in flight:
//   forward projection (SIRT iterations) -> filter (new series: R-weighting +
//   stretch; SIRT: residual O - R) -> backprojection (persistent, per-XCD
//   tile queues over the tiles of all requests).
// Requests that arrive while a round is on the device join the next round: a
// new round is issued only when fewer than max_inflight rounds are in flight
// (one by default), so rounds coalesce arrivals.  A WBP request needs one
// round, a SIRT-k request k+1, a slab request one.
//
// Admission work -- buffers from a device-buffer cache that never calls hipFree
// on the hot path, geometry tables, the projection upload on a stream of its
// own -- overlaps the round on the device.
// Finisher threads wait for each round, copy finished results to the callers'
// buffers (one copy stream each) and return their buffers to the cache.
//
// Tile shapes and every per-tile computation depend only on the request, so a
// request's result is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
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
constexpr int kSlots = 3;  // round slots: at most this many rounds in flight
constexpr size_t kAdmitMax = 8;              // admissions per round issued
constexpr int64_t kAdmitBytes = 2ll << 30;   // projection bytes uploaded per round issued

// An item of at most this many voxel-views runs at the smallest tile depth of
// its kernel variant in the round, in that depth's launch, instead of a launch
// of its own depth.  A shallower tile costs a deep-tuned item up to about 40%
// more per slice, while a launch of its own costs about a tile's time at the
// tail plus the launch gap (tens of microseconds): at the backprojection's
// 4 Gvoxel-view/ms, 5e8 voxel-views is where the two meet
// (TILTGPU_SERVER_MERGE_WORK overrides it).
constexpr double kMergeWork = 5e8;
// A round's SIRT iterations carry at most this much work (about 20 ms of the
// device at 2000 Gvoxel-view/s), so that a round, which every newly arrived
// request waits for, stays short when large SIRT requests are in flight.
constexpr double kRoundWork = 4e10;

struct Req {
    int64_t id = 0;
    int kind = kRecon;
    tiltgpu_dims d{};
    tiltgpu_options o{};
    std::vector<float> angles, xtilt, zfac;
    const float* proj = nullptr;  // caller's projections, read at admission
    size_t proj_n = 0;
    bool has_xtilt = false, has_zfac = false, has_local = false;
    std::vector<int32_t> lnx, lny;
    std::vector<float> lxs, lys, ldx, ldy;
    float* out = nullptr;  // caller's host buffer
    int series = -1, y0 = 0, y1 = 0;
    int64_t need = 0;
    tiltgpu_session* s = nullptr;
    float* slab = nullptr;  // device slab output
    size_t slab_bytes = 0;
    int phase = 0;          // 0: first round; k: SIRT iteration k
    int64_t last_round = -1;  // the last round that ran one of its SIRT iterations
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

// Grows only when a round needs more descriptor space than any before it (the
// slot's previous round has completed by then).
int ensure_batch(BatchBufs& b, size_t bytes) {
    if (bytes <= b.cap) return TILTGPU_OK;
    if (b.dev) (void)hipFree(b.dev);
    if (b.host) (void)hipHostFree(b.host);
    b.dev = b.host = nullptr;
    b.cap = 0;
    const size_t cap = std::max<size_t>(bytes, 1 << 20);
    if (hipMalloc(&b.dev, cap) != hipSuccess || hipHostMalloc(&b.host, cap) != hipSuccess)
        return fail(TILTGPU_ERR_HIP, "server: descriptor buffer allocation failed");
    b.cap = cap;
    return TILTGPU_OK;
}

// Device-buffer cache.  Sizes are rounded up to quarter-octave classes so that
// buffers of similar requests are reused.  A returned buffer becomes reusable
// once the newest round issued at its return has completed, so no reuse can
// overtake work still reading or writing it.  hipFree (which synchronises the
// device) runs only to trim the cache when it would exceed its cap.
class DevCache : public BufferPool {
public:
    int64_t cap = 0;

    void* get(size_t bytes) override {
        std::lock_guard<std::mutex> lk(mu_);
        reclaim();
        const size_t cls = size_class(bytes);
        auto it = free_.lower_bound(cls);
        if (it != free_.end() && it->first <= cls + cls / 2) {
            void* p = it->second;
            live_[p] = it->first;
            live_bytes_ += static_cast<int64_t>(it->first);
            cached_bytes_ -= static_cast<int64_t>(it->first);
            free_.erase(it);
            return p;
        }
        while (!free_.empty() && live_bytes_ + cached_bytes_ + static_cast<int64_t>(cls) > cap) {
            auto last = std::prev(free_.end());
            (void)hipFree(last->second);
            cached_bytes_ -= static_cast<int64_t>(last->first);
            free_.erase(last);
        }
        void* p = nullptr;
        if (hipMalloc(&p, cls) != hipSuccess) {
            for (auto& kv : free_) (void)hipFree(kv.second);
            for (auto& kv : free_) cached_bytes_ -= static_cast<int64_t>(kv.first);
            free_.clear();
            if (hipMalloc(&p, cls) != hipSuccess) return nullptr;
        }
        live_[p] = cls;
        live_bytes_ += static_cast<int64_t>(cls);
        return p;
    }

    void put(void* p, size_t) override {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        auto it = live_.find(p);
        if (it == live_.end()) return;
        pending_.push_back({p, it->second, enq_seq_});
        live_bytes_ -= static_cast<int64_t>(it->second);
        cached_bytes_ += static_cast<int64_t>(it->second);
        live_.erase(it);
    }

    void set_enqueued(uint64_t seq) {
        std::lock_guard<std::mutex> lk(mu_);
        enq_seq_ = seq;
    }
    void set_done(uint64_t seq) {
        std::lock_guard<std::mutex> lk(mu_);
        done_seq_ = std::max(done_seq_, seq);
    }
    void release_all() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& kv : free_) (void)hipFree(kv.second);
        for (auto& e : pending_) (void)hipFree(e.p);
        for (auto& kv : live_) (void)hipFree(kv.first);
        free_.clear();
        pending_.clear();
        live_.clear();
        live_bytes_ = cached_bytes_ = 0;
    }

private:
    struct Pending {
        void* p;
        size_t cls;
        uint64_t seq;
    };
    static size_t size_class(size_t bytes) {
        if (bytes <= (64u << 10)) return 64u << 10;
        int k = 63 - __builtin_clzll(static_cast<unsigned long long>(bytes));
        const size_t step = static_cast<size_t>(1) << (k - 2);
        return (bytes + step - 1) / step * step;
    }
    void reclaim() {
        size_t w = 0;
        for (size_t i = 0; i < pending_.size(); ++i) {
            if (pending_[i].seq <= done_seq_)
                free_.emplace(pending_[i].cls, pending_[i].p);
            else
                pending_[w++] = pending_[i];
        }
        pending_.resize(w);
    }
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    std::vector<Pending> pending_;
    std::unordered_map<void*, size_t> live_;
    int64_t live_bytes_ = 0, cached_bytes_ = 0;
    uint64_t enq_seq_ = 0, done_seq_ = 0;
};

struct Slot {
    BatchBufs bp, filt, fp;
    hipEvent_t done = nullptr;
    uint64_t seq = 0;  // round last issued with this slot (0: never)
};

// A request in its last round, handed to a finisher thread; the round's end
// event is shared by the requests of that round.
struct FinItem {
    std::shared_ptr<hipEvent_t> ev;
    uint64_t seq = 0;
    int rc = 0;
    std::string msg;
    Req* r = nullptr;
};

constexpr int kFinishers = 4;  // result copies of pageable buffers proceed in parallel

}  // namespace

struct tiltgpu_server {
    int device = 0;
    // Rounds allowed on the device at once (1..kSlots).  A new round is issued
    // only when fewer are in flight; requests arriving meanwhile are admitted
    // and join it, so rounds carry every request that arrived while the device
    // was busy.  TILTGPU_SERVER_ROUNDS overrides.
    int max_inflight = 1;
    // Backprojection launches group requests by storage, geometry variant and
    // tile depth; requests of at most merge_work voxel-views join the launch of
    // their variant's smallest depth (the depths are powers of two, so a request
    // prepared for a deeper tile divides evenly; results do not depend on the
    // depth).  TILTGPU_SERVER_MERGE_TILES=0 keeps one launch per depth.
    bool merge_tiles = true;
    double merge_work = kMergeWork;
    // SIRT iterations per round: at most this many voxel-views of forward and
    // back projection (at least one iteration); the iterations of other requests
    // wait for later rounds, the longest-waiting first.  0: every request in
    // flight iterates every round.  TILTGPU_SERVER_ROUND_WORK overrides.
    double round_work = kRoundWork;
    int n_cu = 256;
    int64_t budget = 0, used = 0;
    hipStream_t stream = nullptr;     // rounds
    hipStream_t up_stream = nullptr;  // projection uploads
    hipStream_t down_stream[kFinishers] = {};  // result copies, one per finisher
    std::thread worker, finisher[kFinishers];
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_fin;
    bool stop = false, fin_stop = false;
    std::deque<Req*> queue;
    std::vector<Req*> active;
    std::unordered_map<int64_t, Req*> reqs;
    std::vector<Req*> series;  // resident series (kind kSeries, state kDone)
    int64_t next_id = 1;
    int finishing = 0;         // requests handed to the finisher, not yet released
    Slot slots[kSlots];
    uint64_t round_seq = 0;
    uint64_t completed_seq = 0;  // newest round known to have completed (scheduler)
    std::deque<FinItem> fin;
    std::vector<hipEvent_t> events;  // spare events (under mu)
    DevCache cache;
    // statistics
    std::vector<double> lat_ms;
    int64_t completed = 0, failed = 0, rounds = 0, bp_launches = 0, bp_batched = 0;
    int max_active = 0;
};

namespace {

hipEvent_t take_event(tiltgpu_server* S) {
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->events.empty()) {
        hipEvent_t e = S->events.back();
        S->events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    return e;
}

void give_event(tiltgpu_server* S, hipEvent_t e) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->events.push_back(e);
}

void finish(tiltgpu_server*, Req* r, int rc, const char* msg) {
    r->rc = rc;
    if (rc < 0) r->err = msg ? msg : "";
    r->t_done = Clock::now();
}

// Device state of an admitted request (scheduler thread): buffers from the
// cache, geometry, and the projection upload on the upload stream, which the
// round stream waits for before its next kernels.
int setup(tiltgpu_server* S, Req* r) {
    if (r->kind == kSlab) {
        Req* sr = S->series[r->series];
        const int64_t per = static_cast<int64_t>(sr->d.out_nz) * sr->d.out_nx;
        r->slab_bytes = static_cast<size_t>(r->y1 - r->y0) * per * 4;
        r->slab = static_cast<float*>(S->cache.get(r->slab_bytes));
        if (!r->slab) return fail(TILTGPU_ERR_BUDGET, "no device memory for the slab");
        return TILTGPU_OK;
    }
    int rc = session_create(&r->d, &r->o, S->stream, &r->s, &S->cache);
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
    s->fixed_shape = true;  // the batched backprojection runs launch shape (1, 1)
    if ((rc = prepare_pipeline_geometry(s))) return rc;
    if (r->kind == kSeries) dev_free(s, s->vol);  // slab requests bring their own outputs
    HIP_TRY(hipMemcpyAsync(s->raw, r->proj, r->proj_n * 4, hipMemcpyHostToDevice, S->up_stream));
    hipEvent_t up = take_event(S);
    HIP_TRY(hipEventRecord(up, S->up_stream));
    HIP_TRY(hipStreamWaitEvent(S->stream, up, 0));
    give_event(S, up);  // reusable: the wait has been enqueued
    s->uploaded = true;
    r->proj = nullptr;
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
    double work;  // voxel-views of the item's rows
};

// One persistent launch per kernel variant; all groups' descriptors go to the
// device in a single copy from the slot's pinned staging (rewritten only once
// the slot's previous round has completed).
int launch_bp_groups(tiltgpu_server* S, std::vector<std::vector<BpItem>>& groups,
                     BatchBufs& bufs) {
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
    int rc = ensure_batch(bufs, bytes);
    if (rc) return rc;
    auto* h = static_cast<char*>(bufs.host);
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
    HIP_TRY(hipMemcpyAsync(bufs.dev, h, bytes, hipMemcpyHostToDevice, S->stream));
    auto* d = static_cast<char*>(bufs.dev);
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

// Issue one round (no waiting): stages 1-4 for every request in flight, then the
// finishing work of requests whose last round this is.
int run_round(tiltgpu_server* S, std::vector<Req*>& act, Slot& sl) {
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
    if (S->round_work > 0 && iter.size() > 1) {
        std::stable_sort(iter.begin(), iter.end(),
                         [](const Req* a, const Req* b) { return a->last_round < b->last_round; });
        std::vector<Req*> run;
        double work = 0;
        for (Req* r : iter) {
            const tiltgpu_session* s = r->s;
            const double w = 2.0 * s->d.in_ny * s->d.out_nz * s->d.out_nx * s->nviews;
            if (!run.empty() && work + w > S->round_work) continue;  // waits for a later round
            run.push_back(r);
            work += w;
        }
        iter.swap(run);
    }
    for (Req* r : iter) r->last_round = S->rounds;
    int rc;
    // 1. log + means of new series
    for (Req* r : prep)
        if ((rc = run_means(r->s))) return rc;
    // 2. forward projection of SIRT iterations (one-row: one batched launch per
    // walk form, one or two volume rows per workgroup, as each series' tuned
    // launch would take, and per LDS class, so that series with 16 bins per
    // lane do not lower the occupancy of the others; every form gives the same
    // bits)
    {
        constexpr int kGroups = 4;  // (rows per workgroup - 1) * 2 + (16 bins per lane)
        std::vector<FpParams> grp[kGroups];
        for (Req* r : iter) {
            if (r->s->two_row) {
                if ((rc = run_forward(r->s))) return rc;
            } else {
                const FpParams q = make_fp(r->s);
                grp[(q.rows_per_wg == 2 ? 2 : 0) + (q.bins_per_lane > 8 ? 1 : 0)].push_back(q);
            }
        }
        size_t off_d[kGroups] = {}, off_b[kGroups] = {}, bytes = 0;
        for (int k = 0; k < kGroups; ++k) {
            if (grp[k].empty()) continue;
            off_d[k] = bytes;
            bytes += grp[k].size() * sizeof(FpParams);
            off_b[k] = bytes;
            bytes += (grp[k].size() + 1) * sizeof(int);
            bytes = (bytes + 255) / 256 * 256;
        }
        if (bytes > 0) {
            if ((rc = ensure_batch(sl.fp, bytes))) return rc;
            auto* h = static_cast<char*>(sl.fp.host);
            int total[kGroups] = {};
            size_t lds[kGroups] = {};
            for (int k = 0; k < kGroups; ++k) {
                const int n = static_cast<int>(grp[k].size());
                if (n == 0) continue;
                auto* hp = reinterpret_cast<FpParams*>(h + off_d[k]);
                auto* hb = reinterpret_cast<int*>(h + off_b[k]);
                // each series' workgroup range starts at a multiple of 8 and is
                // padded to one: its workgroups then spread evenly over the 8 XCDs
                for (int i = 0; i < n; ++i) {
                    hp[i] = grp[k][i];
                    hb[i] = total[k];
                    total[k] += (fp_gather_blocks(grp[k][i]) + 7) / 8 * 8;
                    lds[k] = std::max(lds[k], fp_lds_bytes(grp[k][i]));
                }
                hb[n] = total[k];
            }
            HIP_TRY(hipMemcpyAsync(sl.fp.dev, h, bytes, hipMemcpyHostToDevice, S->stream));
            auto* d = static_cast<char*>(sl.fp.dev);
            for (int k = 0; k < kGroups; ++k) {
                if (grp[k].empty()) continue;
                FpBatch b{reinterpret_cast<const FpParams*>(d + off_d[k]),
                          reinterpret_cast<const int*>(d + off_b[k]),
                          static_cast<int>(grp[k].size())};
                HIP_TRY(launch_forward_gather_batch(b, total[k], lds[k], k / 2 + 1, S->stream));
            }
        }
    }
    // 3. filter: new series (R-weighting + stretch) and SIRT residuals O - R
    {
        std::vector<FilterParams> f;
        std::vector<Req*> fr;
        for (Req* r : prep) {
            const float* O = r->s->o.log ? r->s->logd : stack_src(r->s);
            f.push_back(make_filter(r->s, O, nullptr, true));
            fr.push_back(r);
        }
        for (Req* r : iter) {
            const float* O = r->s->o.log ? r->s->logd : stack_src(r->s);
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
            if ((rc = ensure_batch(sl.filt, bytes))) return rc;
            auto* h = static_cast<char*>(sl.filt.host);
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
            HIP_TRY(hipMemcpyAsync(sl.filt.dev, h, bytes, hipMemcpyHostToDevice, S->stream));
            FilterBatch b{reinterpret_cast<const FilterParams*>(sl.filt.dev),
                          reinterpret_cast<const int*>(static_cast<char*>(sl.filt.dev) + off_b),
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
            it.work = static_cast<double>(y1 - y0) * s->d.out_nz * s->d.out_nx * s->d.nz;
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
        // small items join the launch of their variant's smallest depth; every
        // other item runs at its own depth
        if (S->merge_tiles) {
            std::map<std::tuple<int, int, int>, int> dmin;
            for (const BpItem& it : items) {
                auto k = std::make_tuple(it.storage, it.two_row, it.local);
                auto f = dmin.find(k);
                if (f == dmin.end() || it.tile_z < f->second) dmin[k] = it.tile_z;
            }
            for (BpItem& it : items) {
                if (it.work > S->merge_work) continue;
                // the padded thickness is a multiple of every smaller depth, and
                // the LDS bound of a deeper tile covers a shallower one
                const int tz = dmin[std::make_tuple(it.storage, it.two_row, it.local)];
                it.p.ntz = it.p.nz_pad / tz;
                it.tile_z = tz;
            }
        }
        auto key = [](const BpItem& a) {
            return std::make_tuple(a.storage, a.two_row, a.local, a.tile_z);
        };
        std::stable_sort(items.begin(), items.end(),
                         [&](const BpItem& a, const BpItem& b) { return key(a) < key(b); });
        std::vector<std::vector<BpItem>> groups;
        size_t i = 0;
        while (i < items.size()) {
            size_t j = i;
            std::vector<BpItem> grp;
            while (j < items.size() && key(items[j]) == key(items[i])) grp.push_back(items[j++]);
            groups.push_back(std::move(grp));
            i = j;
        }
        if ((rc = launch_bp_groups(S, groups, sl.bp))) return rc;
    }
    // 5. advance; requests in their last round: fake-SIRT and output transform
    // (the finisher copies results out once the round has completed)
    for (Req* r : prep) r->phase = 1;
    for (Req* r : iter) r->phase++;
    for (Req* r : act) {
        if (r->state != kActive || r->kind == kSlab) continue;
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
    }
    return TILTGPU_OK;
}

bool round_finished(const Req* r) {
    if (r->kind == kSlab) return true;
    return r->phase > r->s->o.sirt;
}

// Return a finished request's device state to the cache (finisher thread, after
// its last round and its result copy have completed).
void release(tiltgpu_server* S, Req* r, bool keep_series) {
    if (r->kind == kSeries && keep_series) return;
    if (r->s) {
        tiltgpu_destroy(r->s);
        r->s = nullptr;
    }
    if (r->slab) {
        S->cache.put(r->slab, r->slab_bytes);
        r->slab = nullptr;
    }
    r->proj = nullptr;
}

// Admission is possible when the head of the queue fits, or when nothing holds
// memory that could still be freed (then it fails instead of waiting forever).
bool admissible(const tiltgpu_server* S) {
    if (S->queue.empty()) return false;
    const Req* r = S->queue.front();
    return S->used + r->need <= S->budget || (S->active.empty() && S->finishing == 0);
}

void worker_main(tiltgpu_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<Req*> admitted;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            const auto drained = [&] { return S->stop && S->queue.empty() && S->active.empty(); };
            S->cv_work.wait(lk, [&] { return drained() || !S->active.empty() || admissible(S); });
            if (drained()) break;
            // FIFO admission while the device-memory budget allows; a few at a
            // time, so that the first rounds start while later uploads proceed
            int64_t admit_bytes = 0;
            while (!S->queue.empty() && admitted.size() < kAdmitMax && admit_bytes < kAdmitBytes) {
                Req* r = S->queue.front();
                if (S->used + r->need > S->budget) {
                    if (S->active.empty() && admitted.empty() && S->finishing == 0) {
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
                admit_bytes += static_cast<int64_t>(r->proj_n) * 4;
                admitted.push_back(r);
            }
        }
        // rounds that have completed free their returned buffers for reuse
        for (Slot& sl : S->slots)
            if (sl.seq && hipEventQuery(sl.done) == hipSuccess) S->cache.set_done(sl.seq);
        for (Req* r : admitted) {
            int rc = setup(S, r);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) {
                finish(S, r, rc, tiltgpu_last_error());
                r->state = kFailed;
                release(S, r, false);
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
        // Coalescing: while max_inflight rounds are on the device, wait for the
        // oldest, then admit what arrived meanwhile before issuing the next round.
        {
            int inflight = 0;
            Slot* oldest = nullptr;
            for (Slot& q : S->slots) {
                if (!q.seq || q.seq <= S->completed_seq) continue;
                if (hipEventQuery(q.done) == hipSuccess) {
                    S->completed_seq = std::max(S->completed_seq, q.seq);
                    S->cache.set_done(q.seq);
                    continue;
                }
                ++inflight;
                if (!oldest || q.seq < oldest->seq) oldest = &q;
            }
            if (inflight >= S->max_inflight && oldest) {
                (void)hipEventSynchronize(oldest->done);
                S->completed_seq = std::max(S->completed_seq, oldest->seq);
                S->cache.set_done(oldest->seq);
                continue;
            }
        }
        Slot& sl = S->slots[S->round_seq % kSlots];
        if (sl.seq) {
            (void)hipEventSynchronize(sl.done);
            S->cache.set_done(sl.seq);
        }
        const uint64_t seq = ++S->round_seq;
        S->cache.set_enqueued(seq);
        int rc = run_round(S, act, sl);
        std::string msg = rc ? tiltgpu_last_error() : "";
        if (hipEventRecord(sl.done, S->stream) == hipSuccess) sl.seq = seq;
        std::vector<Req*> keep, done;
        for (Req* r : act) {
            // a failed round fails every request that was in it (the device
            // state of each is unknown); other queued requests are unaffected
            if (rc == 0 && !round_finished(r))
                keep.push_back(r);
            else
                done.push_back(r);
        }
        std::shared_ptr<hipEvent_t> ev;
        if (!done.empty()) {
            ev = std::shared_ptr<hipEvent_t>(new hipEvent_t(take_event(S)), [S](hipEvent_t* e) {
                give_event(S, *e);
                delete e;
            });
            (void)hipEventRecord(*ev, S->stream);
        }
        std::lock_guard<std::mutex> lk(S->mu);
        S->active = keep;
        S->rounds++;
        for (Req* r : done) {
            FinItem it;
            it.ev = ev;
            it.seq = seq;
            it.rc = rc;
            it.msg = msg;
            it.r = r;
            S->fin.push_back(std::move(it));
        }
        S->finishing += static_cast<int>(done.size());
        if (!done.empty()) S->cv_fin.notify_all();
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->fin_stop = true;
    S->cv_fin.notify_all();
}

void finisher_main(tiltgpu_server* S, int k) {
    (void)hipSetDevice(S->device);
    for (;;) {
        FinItem it;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_fin.wait(lk, [&] { return S->fin_stop || !S->fin.empty(); });
            if (S->fin.empty()) break;
            it = std::move(S->fin.front());
            S->fin.pop_front();
        }
        Req* r = it.r;
        int rc = it.rc;
        std::string msg = it.msg;
        hipError_t e = hipEventSynchronize(*it.ev);
        S->cache.set_done(it.seq);
        if (rc == 0 && e != hipSuccess) {
            rc = TILTGPU_ERR_HIP;
            msg = hipGetErrorString(e);
        }
        if (rc == 0 && r->s && (rc = check_kernel_error(r->s))) msg = tiltgpu_last_error();
        if (rc == 0 && r->kind != kSeries) {
            const void* src = r->kind == kSlab ? static_cast<const void*>(r->slab) : r->s->vol;
            const size_t bytes = r->kind == kSlab ? r->slab_bytes
                                                  : static_cast<size_t>(r->s->d.in_ny) *
                                                        r->s->d.out_nz * r->s->d.out_nx * 4;
            hipError_t c =
                hipMemcpyAsync(r->out, src, bytes, hipMemcpyDeviceToHost, S->down_stream[k]);
            if (c == hipSuccess) c = hipStreamSynchronize(S->down_stream[k]);
            if (c != hipSuccess) {
                rc = TILTGPU_ERR_HIP;
                msg = hipGetErrorString(c);
            }
        }
        release(S, r, rc == 0);
        it.ev.reset();
        std::lock_guard<std::mutex> lk(S->mu);
        finish(S, r, rc, msg.c_str());
        r->state = rc ? kFailed : kDone;
        if (r->kind != kSeries || rc) S->used -= r->need;
        if (rc)
            S->failed++;
        else if (r->kind != kSeries) {
            S->completed++;
            S->lat_ms.push_back(
                std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
        }
        S->finishing--;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
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
    // borrowed: the caller keeps the projections until the request completes
    r->proj = projections;
    r->proj_n = static_cast<size_t>(d->nz) * d->in_ny * d->in_nx;
    r->need = recon_bytes(r->d, r->o);
    return TILTGPU_OK;
}

// Copies a request's x-axis tilts, z factors and local grids (each optional).
int fill_geometry(Req* r, int nz, const tiltgpu_view_geometry* g) {
    if (!g) return TILTGPU_OK;
    if (g->xtilt && g->n_xtilt > 0) {
        r->xtilt.assign(g->xtilt, g->xtilt + g->n_xtilt);
        r->has_xtilt = true;
    }
    if (g->zfactor && g->n_zfactor > 0) {
        r->zfac.assign(g->zfactor, g->zfactor + g->n_zfactor);
        r->has_zfac = true;
    }
    if (g->lnx && g->lny) {
        r->lnx.assign(g->lnx, g->lnx + nz);
        r->lny.assign(g->lny, g->lny + nz);
        int64_t nxs = 0, nys = 0, ng = 0;
        for (int v = 0; v < nz; ++v) {
            if (g->lnx[v] < 0 || g->lny[v] < 0) return fail(TILTGPU_ERR_ARG, "negative grid size");
            nxs += g->lnx[v];
            nys += g->lny[v];
            ng += static_cast<int64_t>(g->lnx[v]) * g->lny[v];
        }
        if ((nxs && !g->lxs) || (nys && !g->lys) || (ng && (!g->ldx || !g->ldy)))
            return fail(TILTGPU_ERR_ARG, "local grid arrays missing");
        r->lxs.assign(g->lxs, g->lxs + nxs);
        r->lys.assign(g->lys, g->lys + nys);
        r->ldx.assign(g->ldx, g->ldx + ng);
        r->ldy.assign(g->ldy, g->ldy + ng);
        r->has_local = true;
    }
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
    S->cache.cap = budget_bytes + budget_bytes / 4;
    if (const char* e = std::getenv("TILTGPU_SERVER_ROUNDS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= kSlots) S->max_inflight = v;
    }
    if (const char* e = std::getenv("TILTGPU_SERVER_MERGE_TILES")) S->merge_tiles = std::atoi(e) != 0;
    if (const char* e = std::getenv("TILTGPU_SERVER_MERGE_WORK")) S->merge_work = std::atof(e);
    if (const char* e = std::getenv("TILTGPU_SERVER_ROUND_WORK")) S->round_work = std::atof(e);
    bool ok = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&S->up_stream, hipStreamNonBlocking) == hipSuccess;
    for (hipStream_t& st : S->down_stream)
        ok = ok && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    for (Slot& sl : S->slots)
        ok = ok && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        for (Slot& sl : S->slots)
            if (sl.done) (void)hipEventDestroy(sl.done);
        for (hipStream_t st : {S->stream, S->up_stream})
            if (st) (void)hipStreamDestroy(st);
        for (hipStream_t st : S->down_stream)
            if (st) (void)hipStreamDestroy(st);
        delete S;
        return fail(TILTGPU_ERR_HIP, "server stream/event creation failed");
    }
    S->worker = std::thread(worker_main, S);
    for (int k = 0; k < kFinishers; ++k) S->finisher[k] = std::thread(finisher_main, S, k);
    *out = S;
    return TILTGPU_OK;
}

int tiltgpu_server_submit(tiltgpu_server* S, const tiltgpu_dims* d, const tiltgpu_options* o,
                          const float* angles, int32_t n_angles, const float* projections,
                          float* out, int64_t* id) {
    return tiltgpu_server_submit_geometry(S, d, o, angles, n_angles, nullptr, projections, out,
                                          id);
}

int tiltgpu_server_submit_geometry(tiltgpu_server* S, const tiltgpu_dims* d,
                                   const tiltgpu_options* o, const float* angles,
                                   int32_t n_angles, const tiltgpu_view_geometry* geom,
                                   const float* projections, float* out, int64_t* id) {
    if (!S || !out) return fail(TILTGPU_ERR_ARG, "bad arguments");
    std::unique_ptr<Req> tmp(new Req());
    int rc = fill_recon(tmp.get(), d, o, angles, n_angles, projections);
    if (rc) return rc;
    if ((rc = fill_geometry(tmp.get(), d->nz, geom))) return rc;
    // the two-row reprojection of SIRT keeps 64-bit fixed-point sums
    if (tmp->o.sirt > 0 && (tmp->has_xtilt || tmp->has_local))
        tmp->need += static_cast<int64_t>(d->nz) * d->in_ny * d->in_nx * 8;
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
    const tiltgpu_view_geometry g{xtilt, n_xtilt, nullptr, 0, lnx, lny, lxs, lys, ldx, ldy};
    if ((rc = fill_geometry(tmp.get(), d->nz, &g))) return rc;
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
    // the scheduler drains the queue, then stops the finisher once it is empty
    if (S->worker.joinable()) S->worker.join();
    for (std::thread& t : S->finisher)
        if (t.joinable()) t.join();
    (void)hipStreamSynchronize(S->stream);
    for (Req* r : S->series) {
        if (r->s) tiltgpu_destroy(r->s);
        r->s = nullptr;
    }
    for (auto& kv : S->reqs) delete kv.second;
    for (Slot& sl : S->slots) {
        for (BatchBufs* b : {&sl.bp, &sl.filt, &sl.fp}) {
            if (b->dev) (void)hipFree(b->dev);
            if (b->host) (void)hipHostFree(b->host);
        }
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    for (hipEvent_t e : S->events) (void)hipEventDestroy(e);
    S->cache.release_all();
    for (hipStream_t st : {S->stream, S->up_stream})
        if (st) (void)hipStreamDestroy(st);
    for (hipStream_t st : S->down_stream)
        if (st) (void)hipStreamDestroy(st);
    delete S;
}

}  // extern "C"

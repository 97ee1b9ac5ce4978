// Serving loop of libdqcrt.so (dqc_serve_* in include/dogqc_amd.h):
// continuous batching of plan runs on one GPU.
//
// One scheduler thread owns every HIP call of a server.  Each loop turn it
//   1. polls the completion event of every busy slot; a finished run's
//      status (copied to pinned host memory on the slot's stream) decides
//      between done and re-admission with enlarged tables,
//   2. admits arrived queries in arrival order while a slot is free and the
//      query's workspace fits the arena (a first-fit allocator over one
//      device allocation made at creation, so admission never calls
//      hipMalloc); a query waiting for memory may be passed by later, smaller
//      ones for at most bypass_ms,
//   3. enqueues an admitted plan, its status readback and an event on the
//      slot's stream.
// With size feedback on, a query that needed larger tables teaches the
// server the sizes per input row for its plan module, and later queries of
// the module start with tables scaled to their own inputs, so they neither
// retry nor get admitted with too small a footprint.
// The client thread only submits requests and reads records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "dogqc_amd.h"
#include "dqc_runtime_internal.h"

static_assert(sizeof(dqc_serve_config) == 56, "dqc_serve_config layout");
static_assert(sizeof(dqc_serve_request) == 96, "dqc_serve_request layout");
static_assert(sizeof(dqc_serve_record) == 64 + sizeof(dqc_plan_layout) + sizeof(dqc_plan_status),
              "dqc_serve_record layout");

namespace {

using dqcrt::fail;
using dqcrt::Module;

constexpr uint64_t ARENA_ALIGN = 256;

// First-fit allocator over [0, size) with coalescing of neighbours on free.
class Arena {
  public:
    explicit Arena(uint64_t size) { free_[0] = size; }

    bool alloc(uint64_t bytes, uint64_t* off) {
        bytes = (bytes + ARENA_ALIGN - 1) & ~(ARENA_ALIGN - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const uint64_t rest = it->second - bytes;
                const uint64_t next = it->first + bytes;
                free_.erase(it);
                if (rest) free_[next] = rest;
                used_[*off] = bytes;
                return true;
            }
        }
        return false;
    }

    void release(uint64_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        uint64_t start = off, len = u->second;
        used_.erase(u);
        auto nx = free_.lower_bound(start);
        if (nx != free_.end() && start + len == nx->first) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == start) {
                start = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = len;
    }

  private:
    std::map<uint64_t, uint64_t> free_;   // offset -> length
    std::map<uint64_t, uint64_t> used_;   // offset -> length
};

struct Request {
    int64_t id = 0;
    Module* module = nullptr;
    std::vector<int64_t> rows;
    std::vector<const void*> cols;
    dqc_plan_inputs inputs{};
    dqc_plan_tuning tuning{};
    dqc_serve_record rec{};
    uint64_t arenaOff = 0;
    bool hasWs = false;
    int64_t blockedSince = -1;   // first time it waited for memory at the head of the queue
    bool keep = false;
    bool heavy = false;          // counted against max_heavy while running

    // a query that may run in a batch (see dqc_serve_request)
    Module* batchModule = nullptr;
    std::vector<const void*> batchCols;
    std::vector<int32_t> batchColPrefix;
    std::vector<int32_t> batchTableMap;
    int64_t batchOf = -1;        // the batch it ran in

    // a batch: a request of the server's own for several queries
    bool internal = false;
    std::vector<int64_t> members;       // query ids, by growing prefixes
    std::vector<int32_t> prefixCols;    // batched-module columns filled with the members' prefixes
    std::vector<int64_t> prefix;        // [prefix column][member] row counts (host copy)
    uint64_t prefixOff = 0;             // prefix block, after the workspace in the allocation
    int keepRefs = 0;                   // members that keep the result and have not released it

    void bindInputs() {
        inputs.table_rows = rows.data();
        inputs.cols = cols.data();
    }
};

// Table sizes per driving input row learned for one plan module.
struct Feedback {
    double ht[DQC_MAX_HTS] = {};
    double mm[DQC_MAX_HTS] = {};
    double split[DQC_MAX_HTS] = {};
    double temp[DQC_MAX_TEMPS] = {};
    double result = 0;
};

// the largest input table: the size the plan's tables grow with
int64_t drivingRows(const Request& r) {
    int64_t m = 1;
    for (int64_t v : r.rows) m = std::max(m, v);
    return m;
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int64_t req = -1;
};

struct Server {
    dqc_serve_config cfg{};
    int device = 0;
    void* arenaBase = nullptr;
    std::unique_ptr<Arena> arena;
    std::vector<Slot> slots;
    dqc_plan_status* pinned = nullptr;   // one status per slot
    // per slot: a batch's prefix row counts staged in pinned host memory, so
    // their copy to the device is asynchronous (a copy from pageable memory
    // is not); a slot's stage is reused only after its query completed
    static constexpr size_t PREFIX_STAGE_WORDS = 512;
    int64_t* prefixPinned = nullptr;

    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<Request>> reqs;
    std::deque<int64_t> pending;          // not yet admitted: arrived ones first, then by arrival time
    int64_t unfinished = 0;
    std::vector<int64_t> toRelease;       // kept workspaces the client released
    std::map<Module*, Feedback> feedback;

    std::thread worker;
    std::atomic<bool> started{false};
    std::atomic<bool> stop{false};
    std::chrono::steady_clock::time_point t0;
    int fatal = 0;

    int64_t now() const {
        return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    }

    void finish(Request& r, int state, int error) {
        r.rec.state = state;
        r.rec.error = error;
        r.rec.done_ns = now();
        if (!r.internal) {
            --unfinished;
            return;
        }
        // a batch ends its members with it
        for (size_t k = 0; k < r.members.size(); ++k) {
            Request& m = *reqs[r.members[k]];
            m.rec.batch_size = (int32_t)r.members.size();
            m.rec.batch_index = (int32_t)k;
            m.rec.retries = r.rec.retries;
            m.rec.layout = r.rec.layout;
            m.rec.status = r.rec.status;
            m.rec.ws_bytes = r.rec.ws_bytes;
            m.rec.ws = (state == DQC_SERVE_DONE && m.keep) ? r.rec.ws : nullptr;
            finish(m, state, error);
        }
    }

    static constexpr uint64_t align256(uint64_t x) { return (x + 255u) & ~(uint64_t)255u; }

    // A batch of the waiting queries `group` (same module and batched
    // module), or -1 when their prefixes are not nested in every table or
    // the batch would not fit the arena.
    int64_t makeBatch(std::vector<int64_t> group) {
        std::stable_sort(group.begin(), group.end(), [this](int64_t a, int64_t b) {
            return drivingRows(*reqs[a]) < drivingRows(*reqs[b]);
        });
        const Request& m0 = *reqs[group[0]];
        const size_t N = group.size(), nt = m0.batchTableMap.size(), nc = m0.batchCols.size();
        std::unique_ptr<Request> b(new Request());
        b->internal = true;
        b->module = m0.batchModule;
        b->members = group;
        b->rows.assign(nt, 0);
        for (size_t j = 0; j < nt; ++j) {
            const int32_t src = m0.batchTableMap[j];
            if (src < 0) {
                b->rows[j] = (int64_t)N;
                continue;
            }
            int64_t prev = 0;
            for (int64_t id : group) {
                const Request& m = *reqs[id];
                if (src >= (int32_t)m.rows.size()) return -1;
                const int64_t v = m.rows[src];
                if (v < prev) return -1;   // prefixes not nested: no batch
                prev = v;
            }
            b->rows[j] = prev;
        }
        b->cols = m0.batchCols;
        for (size_t c = 0; c < nc; ++c) {
            const int32_t j = m0.batchColPrefix[c];
            if (j < 0) continue;
            if (j >= (int32_t)nt || m0.batchTableMap[j] < 0) return -1;
            b->prefixCols.push_back((int32_t)c);
            for (int64_t id : group) b->prefix.push_back(reqs[id]->rows[m0.batchTableMap[j]]);
        }
        if (b->prefix.size() > PREFIX_STAGE_WORDS) return -1;   // more than a slot's stage holds: no batch
        b->inputs.ntables = (int32_t)nt;
        b->inputs.ncols = (int32_t)nc;
        b->bindInputs();
        int64_t arrival = m0.rec.arrival_ns;
        for (int64_t id : group) {
            arrival = std::min(arrival, reqs[id]->rec.arrival_ns);
            b->keep = b->keep || reqs[id]->keep;
        }
        b->rec.arrival_ns = arrival;
        b->rec.state = DQC_SERVE_PENDING;
        b->rec.slot = -1;
        dqc_plan_layout L;
        if (b->module->layout(&b->inputs, &b->tuning, &L)) return -1;
        if (align256(L.ws_bytes) + b->prefix.size() * sizeof(int64_t) > cfg.arena_bytes) return -1;
        const int64_t id = (int64_t)reqs.size();
        b->id = id;
        reqs.push_back(std::move(b));
        return id;
    }

    void freeWs(Request& r) {
        if (r.hasWs) arena->release(r.arenaOff);
        r.hasWs = false;
        r.rec.ws = nullptr;
    }

    // Sizes from what earlier queries of the module needed, scaled to this
    // query's inputs; only ever enlarges.
    void applyFeedback(Request& r) {
        auto it = feedback.find(r.module);
        if (it == feedback.end()) return;
        const Feedback& f = it->second;
        dqc_plan_layout L;
        if (r.module->layout(&r.inputs, &r.tuning, &L)) return;
        const double rows = (double)drivingRows(r);
        auto grow = [rows](int64_t& field, int64_t current, double ratio, double margin) {
            if (ratio <= 0) return;
            const int64_t want = (int64_t)(ratio * rows * margin) + 1;
            if (want > std::max<int64_t>(field, current)) field = want;
        };
        for (int i = 0; i < L.nhts && i < DQC_MAX_HTS; ++i) {
            grow(r.tuning.ht_capacity[i], L.ht_capacity[i], f.ht[i], 1.0);
            grow(r.tuning.mm_entries[i], L.mm_entries[i], f.mm[i], 1.1);
            if (L.split_capacity[i] > 0) grow(r.tuning.split_capacity[i], L.split_capacity[i], f.split[i], 1.1);
        }
        for (int i = 0; i < L.ntemps && i < DQC_MAX_TEMPS; ++i)
            grow(r.tuning.temp_capacity[i], L.temp_capacity[i], f.temp[i], 1.1);
        if (L.nresult > 0) grow(r.tuning.result_capacity, L.result_capacity, f.result, 1.1);
    }

    // After a run that had to enlarge tables: remember the enlarged sizes.
    void learn(const Request& r) {
        Feedback& f = feedback[r.module];
        const double rows = (double)drivingRows(r);
        const dqc_plan_tuning& t = r.tuning;
        for (int i = 0; i < DQC_MAX_HTS; ++i) {
            if (t.ht_capacity[i] > 0) f.ht[i] = std::max(f.ht[i], t.ht_capacity[i] / rows);
            if (t.mm_entries[i] > 0) f.mm[i] = std::max(f.mm[i], t.mm_entries[i] / rows);
            if (t.split_capacity[i] > 0) f.split[i] = std::max(f.split[i], t.split_capacity[i] / rows);
        }
        for (int i = 0; i < DQC_MAX_TEMPS; ++i)
            if (t.temp_capacity[i] > 0) f.temp[i] = std::max(f.temp[i], t.temp_capacity[i] / rows);
        if (t.result_capacity > 0) f.result = std::max(f.result, t.result_capacity / rows);
    }

    int runningHeavy() const {
        int n = 0;
        for (const Slot& sl : slots)
            if (sl.req >= 0 && reqs[sl.req]->heavy) ++n;
        return n;
    }

    int freeSlots() const {
        int n = 0;
        for (const Slot& sl : slots)
            if (sl.req < 0) ++n;
        return n;
    }

    enum Admit { ADMITTED, HELD_HEAVY, HELD_LIGHT, NO_MEMORY };

    // Admission of one request on slot `si`.  HELD_HEAVY: the heavy limit is
    // reached; HELD_LIGHT: a light query leaves the last free slot to a
    // waiting heavy one (`heavyWaiting`); NO_MEMORY: the arena is too full.
    // Failing queries count as admitted (they are finished).
    Admit tryAdmit(int64_t id, int si, int64_t t, bool heavyWaiting) {
        Request& r = *reqs[id];
        // every (re)admission: what any query of the module needed so far
        if (cfg.size_feedback) applyFeedback(r);
        dqc_plan_layout L;
        int rc = r.module->layout(&r.inputs, &r.tuning, &L);
        if (rc) {
            fail(rc, "plan layout failed (%d)", rc);
            finish(r, DQC_SERVE_FAILED, rc);
            return ADMITTED;
        }
        if (L.ws_bytes > cfg.arena_bytes) {
            finish(r, DQC_SERVE_FAILED, DQC_ENOMEM);
            return ADMITTED;
        }
        const bool heavy = cfg.heavy_bytes > 0 && cfg.max_heavy > 0 && L.ws_bytes >= cfg.heavy_bytes;
        if (heavy && runningHeavy() >= cfg.max_heavy) return HELD_HEAVY;
        if (!heavy && heavyWaiting && freeSlots() <= 1) return HELD_LIGHT;
        uint64_t off = 0;
        const uint64_t prefixBytes = r.prefix.size() * sizeof(int64_t);
        const uint64_t allocBytes = prefixBytes ? align256(L.ws_bytes) + prefixBytes : L.ws_bytes;
        if (!arena->alloc(allocBytes, &off)) return NO_MEMORY;
        r.arenaOff = off;
        r.hasWs = true;
        r.heavy = heavy;
        r.rec.ws = (char*)arenaBase + off;
        r.rec.ws_bytes = L.ws_bytes;
        r.rec.layout = L;
        r.rec.admit_ns = t;
        r.rec.slot = si;
        r.rec.state = DQC_SERVE_RUNNING;
        Slot& s = slots[si];
        hipError_t e = hipSuccess;
        if (prefixBytes) {
            // the members' prefix row counts after the workspace; the copy
            // precedes the plan on the slot's stream
            r.prefixOff = align256(L.ws_bytes);
            int64_t* dev = (int64_t*)((char*)r.rec.ws + r.prefixOff);
            const size_t N = r.members.size();
            for (size_t i = 0; i < r.prefixCols.size(); ++i) r.cols[r.prefixCols[i]] = dev + i * N;
            r.bindInputs();
            int64_t* stage = prefixPinned + (size_t)si * PREFIX_STAGE_WORDS;
            std::memcpy(stage, r.prefix.data(), prefixBytes);
            e = hipMemcpyAsync(dev, stage, prefixBytes, hipMemcpyHostToDevice, s.stream);
        }
        rc = e == hipSuccess ? r.module->enqueue(&r.inputs, &r.tuning, r.rec.ws, L.ws_bytes, (void*)s.stream, nullptr)
                             : DQC_EHIP;
        if (rc == DQC_OK) e = hipMemcpyAsync(&pinned[si], r.rec.ws, sizeof(dqc_plan_status), hipMemcpyDeviceToHost,
                                            s.stream);
        if (rc == DQC_OK && e == hipSuccess) e = hipEventRecord(s.done, s.stream);
        if (rc != DQC_OK || e != hipSuccess) {
            // nothing of this query can be pending on a stream it failed to use
            (void)hipStreamSynchronize(s.stream);
            freeWs(r);
            finish(r, DQC_SERVE_FAILED, rc != DQC_OK ? rc : DQC_EHIP);
            return ADMITTED;
        }
        s.req = id;
        return ADMITTED;
    }

    // One scheduler turn; returns whether anything happened.
    bool turn() {
        bool progress = false;
        std::unique_lock<std::mutex> lk(mu);
        for (int64_t id : toRelease) freeWs(*reqs[id]);
        toRelease.clear();
        // 1. completions
        for (size_t si = 0; si < slots.size(); ++si) {
            Slot& s = slots[si];
            if (s.req < 0) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            Request& r = *reqs[s.req];
            s.req = -1;
            progress = true;
            if (q != hipSuccess) {
                freeWs(r);
                finish(r, DQC_SERVE_FAILED, DQC_EHIP);
                fatal = DQC_EHIP;
                continue;
            }
            r.rec.status = pinned[si];
            if (dqcrt::adjustForRetry(r.rec.layout, r.rec.status, &r.tuning)) {
                freeWs(r);
                if (r.rec.retries >= cfg.max_retries) {
                    finish(r, DQC_SERVE_FAILED, DQC_EOVERFLOW);
                } else {
                    ++r.rec.retries;
                    r.rec.state = DQC_SERVE_PENDING;
                    // the enlarged sizes already tell later queries of the
                    // module what they need; the query itself runs again next
                    if (cfg.size_feedback) learn(r);
                    pending.push_front(r.id);
                }
                continue;
            }
            if (r.internal) {
                r.keepRefs = 0;
                for (int64_t id : r.members) r.keepRefs += reqs[id]->keep ? 1 : 0;
            }
            if (!r.keep) freeWs(r);
            if (cfg.size_feedback && r.rec.retries > 0) learn(r);
            finish(r, DQC_SERVE_DONE, 0);
        }
        // 2. admissions in arrival order (stamped after the completions of
        // this turn, whose slots and memory they may take over)
        const int64_t ta = now();
        bool heavyWaiting = false;
        for (auto it = pending.begin(); it != pending.end();) {
            int si = freeSlot();
            if (si < 0) break;
            Request& r = *reqs[*it];
            if (r.rec.arrival_ns > ta) break;   // pending is kept in arrival order
            if (cfg.max_batch > 1 && r.batchModule && !r.internal) {
                // waiting queries of the same plan join this one in a batch
                std::vector<int64_t> group{*it};
                const int maxb = std::min(cfg.max_batch, DQC_SERVE_MAX_BATCH);
                const double head = (double)drivingRows(r);
                auto similar = [this, head](const Request& o) {
                    if (cfg.batch_ratio <= 0) return true;
                    const double x = (double)drivingRows(o), q = (double)cfg.batch_ratio;
                    return x <= head * q && head <= x * q;
                };
                for (auto jt = std::next(it); jt != pending.end() && (int)group.size() < maxb; ++jt) {
                    const Request& o = *reqs[*jt];
                    if (o.rec.arrival_ns > ta) break;
                    if (!o.internal && o.module == r.module && o.batchModule == r.batchModule && similar(o))
                        group.push_back(*jt);
                }
                // the largest member leaves a batch whose other members hold
                // too little work to pay for the batch's extra cost
                while (cfg.batch_min_share > 0 && group.size() >= 2) {
                    size_t big = 0;
                    double sum = 0;
                    for (size_t k = 0; k < group.size(); ++k) {
                        const double x = (double)drivingRows(*reqs[group[k]]);
                        sum += x;
                        if (x > (double)drivingRows(*reqs[group[big]])) big = k;
                    }
                    const double largest = (double)drivingRows(*reqs[group[big]]);
                    if (sum - largest >= cfg.batch_min_share * largest) break;
                    group.erase(group.begin() + (std::ptrdiff_t)big);
                }
                const int64_t bid = group.size() >= 2 ? makeBatch(group) : -1;
                if (bid >= 0) {
                    const Admit a = tryAdmit(bid, si, ta, heavyWaiting);
                    if (a == ADMITTED) {
                        progress = true;
                        for (int64_t id : group) {
                            Request& m = *reqs[id];
                            m.batchOf = bid;
                            m.rec.state = reqs[bid]->rec.state;
                            m.rec.admit_ns = ta;
                            m.rec.slot = si;
                        }
                        for (auto kt = pending.begin(); kt != pending.end();)
                            kt = std::find(group.begin(), group.end(), *kt) != group.end() ? pending.erase(kt)
                                                                                          : std::next(kt);
                        it = pending.begin();
                        continue;
                    }
                    reqs.pop_back();   // not admitted now: the queries wait as they are
                }
            }
            const Admit a = tryAdmit(*it, si, ta, heavyWaiting);
            if (a == ADMITTED) {
                progress = true;
                it = pending.erase(it);
                continue;
            }
            if (a == HELD_HEAVY) {   // later light queries may still run now
                heavyWaiting = true;
                ++it;
                continue;
            }
            if (a == HELD_LIGHT) break;
            // does not fit: wait, or let later queries pass for a while
            if (r.blockedSince < 0) r.blockedSince = ta;
            if (cfg.bypass_ms <= 0 || ta - r.blockedSince > (int64_t)cfg.bypass_ms * 1000000) break;
            ++it;
        }
        if (progress) cv.notify_all();
        return progress;
    }

    // insert into `pending`, which is kept in arrival order
    void enqueuePending(int64_t id) {
        const int64_t a = reqs[id]->rec.arrival_ns;
        auto pos = pending.end();
        while (pos != pending.begin() && reqs[*std::prev(pos)]->rec.arrival_ns > a) --pos;
        pending.insert(pos, id);
    }

    int freeSlot() const {
        for (size_t i = 0; i < slots.size(); ++i)
            if (slots[i].req < 0) return (int)i;
        return -1;
    }

    bool busy() const {
        for (const Slot& s : slots)
            if (s.req >= 0) return true;
        return false;
    }

    void run() {
        (void)hipSetDevice(device);
        while (true) {
            const bool progress = turn();
            if (stop.load()) {
                std::lock_guard<std::mutex> lk(mu);
                if (!busy()) break;
            }
            if (!progress) std::this_thread::yield();
        }
    }
};

}  // namespace

extern "C" {

int dqc_serve_create(const dqc_serve_config* cfg, void** server) {
    if (!cfg || !server) return fail(DQC_EINVAL, "dqc_serve_create: null argument");
    if (cfg->slots < 1 || cfg->slots > 64) return fail(DQC_EINVAL, "slots must be in 1..64, got %d", cfg->slots);
    if (cfg->arena_bytes == 0) return fail(DQC_EINVAL, "arena_bytes must be positive");
    std::unique_ptr<Server> s(new Server());
    s->cfg = *cfg;
    HIP_CHECK(hipGetDevice(&s->device), "hipGetDevice");
    void* base = nullptr;
    hipError_t e = hipMalloc(&base, cfg->arena_bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(DQC_ENOMEM, "arena of %llu bytes: %s", (unsigned long long)cfg->arena_bytes,
                    hipGetErrorString(e));
    }
    s->arenaBase = base;
    s->arena.reset(new Arena(cfg->arena_bytes));
    e = hipHostMalloc((void**)&s->pinned, sizeof(dqc_plan_status) * cfg->slots, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(base);
        return dqcrt::hipFail(e, "hipHostMalloc");
    }
    e = hipHostMalloc((void**)&s->prefixPinned, sizeof(int64_t) * Server::PREFIX_STAGE_WORDS * cfg->slots,
                      hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipHostFree(s->pinned);
        (void)hipFree(base);
        return dqcrt::hipFail(e, "hipHostMalloc");
    }
    s->slots.resize(cfg->slots);
    for (Slot& sl : s->slots) {
        HIP_CHECK(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
        HIP_CHECK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
    }
    *server = s.release();
    return DQC_OK;
}

int dqc_serve_submit(void* server, const dqc_serve_request* req, int64_t* id) {
    Server* s = (Server*)server;
    if (!s || !req || !req->module) return fail(DQC_EINVAL, "dqc_serve_submit: null argument");
    if (req->inputs.ntables < 0 || req->inputs.ncols < 0) return fail(DQC_EINVAL, "bad input binding");
    std::unique_ptr<Request> r(new Request());
    r->module = (Module*)req->module;
    r->rows.assign(req->inputs.table_rows, req->inputs.table_rows + req->inputs.ntables);
    r->cols.assign(req->inputs.cols, req->inputs.cols + req->inputs.ncols);
    r->inputs.ntables = req->inputs.ntables;
    r->inputs.ncols = req->inputs.ncols;
    r->bindInputs();
    if (req->tuning) r->tuning = *req->tuning;
    r->keep = req->keep_result != 0;
    if (req->batch_module) {
        if (req->batch_ncols <= 0 || req->batch_ntables <= 0 || !req->batch_cols || !req->batch_col_prefix ||
            !req->batch_table_map)
            return fail(DQC_EINVAL, "dqc_serve_submit: incomplete batch binding");
        r->batchModule = (Module*)req->batch_module;
        r->batchCols.assign(req->batch_cols, req->batch_cols + req->batch_ncols);
        r->batchColPrefix.assign(req->batch_col_prefix, req->batch_col_prefix + req->batch_ncols);
        r->batchTableMap.assign(req->batch_table_map, req->batch_table_map + req->batch_ntables);
    }
    r->rec.arrival_ns = req->arrival_ns;
    r->rec.state = DQC_SERVE_PENDING;
    r->rec.slot = -1;
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t n = (int64_t)s->reqs.size();
    r->id = n;
    s->reqs.push_back(std::move(r));
    s->enqueuePending(n);
    ++s->unfinished;
    if (id) *id = n;
    return DQC_OK;
}

int dqc_serve_start(void* server) {
    Server* s = (Server*)server;
    if (!s) return fail(DQC_EINVAL, "null server");
    if (s->started.exchange(true)) return fail(DQC_EINVAL, "server already started");
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->run(); });
    return DQC_OK;
}

int dqc_serve_wait(void* server, int64_t timeout_ms) {
    Server* s = (Server*)server;
    if (!s) return fail(DQC_EINVAL, "null server");
    if (!s->started.load()) return fail(DQC_EINVAL, "server not started");
    std::unique_lock<std::mutex> lk(s->mu);
    auto pred = [s] { return s->unfinished == 0 || s->fatal != 0; };
    if (timeout_ms < 0) {
        s->cv.wait(lk, pred);
    } else if (!s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) {
        return fail(DQC_EINVAL, "dqc_serve_wait: %lld queries unfinished after %lld ms", (long long)s->unfinished,
                    (long long)timeout_ms);
    }
    if (s->fatal) return fail(s->fatal, "a query failed on the device");
    return DQC_OK;
}

int dqc_serve_record_of(void* server, int64_t id, dqc_serve_record* out) {
    Server* s = (Server*)server;
    if (!s || !out) return fail(DQC_EINVAL, "dqc_serve_record_of: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->reqs.size()) return fail(DQC_EINVAL, "no query %lld", (long long)id);
    *out = s->reqs[id]->rec;
    return DQC_OK;
}

int dqc_serve_release(void* server, int64_t id) {
    Server* s = (Server*)server;
    if (!s) return fail(DQC_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->reqs.size()) return fail(DQC_EINVAL, "no query %lld", (long long)id);
    Request& r = *s->reqs[id];
    if (r.rec.state != DQC_SERVE_DONE) return fail(DQC_EINVAL, "query %lld is not done", (long long)id);
    if (r.batchOf >= 0) {
        // the batch's workspace goes when its last keeping member lets go
        Request& b = *s->reqs[r.batchOf];
        if (r.rec.ws) {
            r.rec.ws = nullptr;
            if (--b.keepRefs <= 0 && b.hasWs) {
                if (s->started.load() && !s->stop.load())
                    s->toRelease.push_back(b.id);
                else
                    s->freeWs(b);
            }
        }
        return DQC_OK;
    }
    if (!r.hasWs) return DQC_OK;
    if (s->started.load() && !s->stop.load())
        s->toRelease.push_back(id);   // the scheduler thread owns the arena
    else
        s->freeWs(r);
    r.rec.ws = nullptr;
    return DQC_OK;
}

int dqc_serve_destroy(void* server) {
    Server* s = (Server*)server;
    if (!s) return DQC_OK;
    if (s->started.load()) {
        s->stop.store(true);
        s->worker.join();
    }
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->pinned) (void)hipHostFree(s->pinned);
    if (s->prefixPinned) (void)hipHostFree(s->prefixPinned);
    if (s->arenaBase) (void)hipFree(s->arenaBase);
    delete s;
    return DQC_OK;
}

}  // extern "C"

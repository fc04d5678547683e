// Request server: continuous batching of products and convolutions on one
// GPU under a device-memory budget.
//
// Requests enter one FIFO queue. Worker slots, each with its own stream, take
// work as soon as they are free:
//   - small products (both operands short) are taken in groups of up to
//     kGroupMax queued requests and run as one batched launch sequence over
//     slot-local memory;
//   - every other request is admitted when its device working set (operands,
//     spectra, carries, result) fits in what the budget has left, in queue
//     order, and runs alone on the slot's stream.
// Several slots run at once, so requests that arrive while others are on the
// GPU overlap with them. Root tables are shared and stay resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "engine.h"

namespace nttamd {

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint64_t kSmallLimbs = 65536;      // a + b limbs of a groupable product
constexpr size_t kGroupMax = 512;            // requests per group
constexpr uint64_t kGroupSpectra = 256ull << 20;  // spectra bytes per group

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit sub-allocator over one device allocation, with coalescing.
class DeviceArena {
   public:
    bool init(uint64_t bytes) {
        bytes = align_up(bytes);
        if (hipMalloc(&base_, bytes) != hipSuccess) {
            base_ = nullptr;
            return false;
        }
        capacity_ = bytes;
        free_[0] = bytes;
        return true;
    }
    ~DeviceArena() {
        if (base_) hipFree(base_);
    }
    // Returns the offset, or UINT64_MAX when no free range is large enough.
    uint64_t alloc(uint64_t bytes) {
        bytes = align_up(std::max<uint64_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t off = it->first, size = it->second;
            free_.erase(it);
            if (size > bytes) free_[off + bytes] = size - bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return UINT64_MAX;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = align_up(std::max<uint64_t>(bytes, 1));
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    char* at(uint64_t off) const { return static_cast<char*>(base_) + off; }
    uint64_t capacity() const { return capacity_; }
    uint64_t peak() const { return peak_; }

   private:
    void* base_ = nullptr;
    uint64_t capacity_ = 0, used_ = 0, peak_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    ntt_request r{};
    Geometry g;
    ntt_conv_plan plan{};
    bool square = false;
    bool device = false;
    bool small = false;
    uint64_t out_len = 0;
    uint64_t bytes = 0;  // arena bytes of the working set
    uint64_t offset = UINT64_MAX;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 0;  // 0 queued or running, 1 done, -1 failed
    bool waited = false;  // has waited for device memory (counted once)
};

struct Slot {
    Batch batch;  // its stream is the slot's stream
    DevBuf dev_in, dev_out;
    HostBuf host_in, host_out;  // a group's operands and results, gathered
    HostStage stage;            // a large request's pageable host buffers
    ~Slot() {
        dev_in.release();
        dev_out.release();
        host_in.release();
        host_out.release();
    }
};

}  // namespace

struct Server {
    std::mutex lock;
    std::condition_variable work, done;
    std::deque<Request*> queue;
    std::unordered_map<int64_t, Request*> requests;
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<std::thread> threads;
    DeviceArena arena;
    int64_t next_id = 0;
    bool stopping = false;
    int device = 0;
    ntt_server_stats stats{};

    ~Server() {
        {
            std::lock_guard<std::mutex> g(lock);
            stopping = true;
        }
        work.notify_all();
        for (auto& t : threads) t.join();
        for (auto& kv : requests) delete kv.second;
    }

    // Picks work for a free slot (lock held), in arrival order: the small
    // products queued ahead of the oldest large request go as one group;
    // when that large request is the oldest, it is admitted as soon as its
    // working set fits the budget. While it waits for memory, small products
    // queued behind it may run (they use the slot's own arena, not the
    // budget), but none is taken ahead of it once it fits, so a steady
    // stream of small products cannot starve it.
    bool take(std::vector<Request*>* group) {
        group->clear();
        uint64_t spectra = 0;
        auto add_small = [&](std::deque<Request*>::iterator& it) -> bool {
            Request* q = *it;
            if (!group->empty() && q->device != (*group)[0]->device) {
                ++it;
                return true;
            }
            const uint64_t sb = op_spectra_bytes(q->g) + (q->square ? 0 : op_spectra_bytes_b(q->g));
            if (!group->empty() && (group->size() >= kGroupMax || spectra + sb > kGroupSpectra)) return false;
            spectra += sb;
            group->push_back(q);
            it = queue.erase(it);
            return true;
        };
        auto it = queue.begin();
        while (it != queue.end() && (*it)->small)
            if (!add_small(it)) return true;
        if (!group->empty() || it == queue.end()) return !group->empty();
        // The oldest request is large.
        Request* big = *it;
        const uint64_t off = arena.alloc(big->bytes);
        if (off != UINT64_MAX) {
            big->offset = off;
            group->push_back(big);
            queue.erase(it);
            return true;
        }
        if (!big->waited) {
            big->waited = true;
            ++stats.waits_for_memory;
        }
        // It waits for memory: small products behind it may run meanwhile.
        for (++it; it != queue.end();) {
            if (!(*it)->small) {
                ++it;
                continue;
            }
            if (!add_small(it)) break;
        }
        return !group->empty();
    }

    void finish(const std::vector<Request*>& group, bool ok) {
        const double t = now_s();
        std::lock_guard<std::mutex> g(lock);
        for (Request* q : group) {
            q->t_done = t;
            q->status = ok ? 1 : -1;
            if (ok)
                ++stats.completed;
            else
                ++stats.failed;
            if (q->offset != UINT64_MAX) {
                arena.release(q->offset, q->bytes);
                q->offset = UINT64_MAX;
            }
        }
        if (group.size() > 1 || (group.size() == 1 && group[0]->small)) {
            ++stats.groups;
            stats.grouped += group.size();
        }
        done.notify_all();
        work.notify_all();  // memory was freed
    }

    bool run_large(Slot& slot, Request* q);
    bool run_group(Slot& slot, const std::vector<Request*>& group);

    void worker(int index) {
        hipSetDevice(device);
        Slot& slot = *slots[index];
        std::vector<Request*> group;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(lock);
                while (!take(&group)) {
                    if (stopping && queue.empty()) return;
                    work.wait(g);
                }
            }
            const double t = now_s();
            for (Request* q : group) q->t_start = t;
            bool ok;
            if (group.size() == 1 && !group[0]->small)
                ok = run_large(slot, group[0]);
            else
                ok = run_group(slot, group);
            finish(group, ok);
            group.clear();
        }
    }
};

// One request alone on the slot's stream, with its working set carved from
// the arena range reserved at admission.
bool Server::run_large(Slot& slot, Request* q) {
    const hipStream_t s = slot.batch.ex.stream;
    char* p = arena.at(q->offset);
    const ntt_request& r = q->r;
    const uint64_t la = r.len_a, lb = q->square ? r.len_a : r.len_b;
    const uint64_t* da = r.a;
    const uint64_t* db = q->square ? r.a : r.b;
    uint64_t* dout = r.dst;
    uint64_t off = 0;
    auto carve = [&](uint64_t bytes) {
        char* x = p + off;
        off += align_up(bytes);
        return x;
    };
    if (!q->device) {
        uint64_t* ia = reinterpret_cast<uint64_t*>(carve(la * 8));
        if (!copy_to_device(slot.stage, s, ia, r.a, la * 8)) return false;
        da = ia;
        if (q->square) {
            db = ia;
        } else {
            uint64_t* ib = reinterpret_cast<uint64_t*>(carve(lb * 8));
            if (!copy_to_device(slot.stage, s, ib, r.b, lb * 8)) return false;
            db = ib;
        }
        dout = reinterpret_cast<uint64_t*>(carve(q->out_len * 8));
    }
    const bool conv = r.kind == NTT_REQ_CONV;
    // ConvPlan::convolve rejects coefficients above their bounds
    // (src/convolve.rs:148-162); checked here, off the submitting thread.
    if (conv && !q->device && !conv_admits(&q->plan, r.a, la, r.b, lb)) {
        set_error("server: convolution coefficient above its bound");
        return false;
    }
    OpBuffers buf{nullptr, nullptr, nullptr};
    if (!(conv && q->plan.schoolbook)) {
        buf.spec_a = reinterpret_cast<double*>(carve(op_spectra_bytes(q->g)));
        buf.spec_b = q->square ? buf.spec_a : reinterpret_cast<double*>(carve(op_spectra_bytes_b(q->g)));
        buf.carry = reinterpret_cast<uint32_t*>(carve(op_carry_bytes(q->out_len, q->g.school)));
    }
    const bool ok = conv ? convolution_on(s, q->g, &q->plan, da, db, q->square, buf, dout)
                         : product_on(s, q->g, da, la, db, lb, q->square, buf, dout);
    // dst is written only once every kernel of the request has completed.
    if (hipStreamSynchronize(s) != hipSuccess || !ok) return false;
    if (!q->device) return copy_to_host(slot.stage, s, r.dst, dout, q->out_len * 8);
    return true;
}

// A group of small products as one batch: operands gathered into pinned
// staging, one copy in, the batched passes, one copy out.
bool Server::run_group(Slot& slot, const std::vector<Request*>& group) {
    const size_t n = group.size();
    std::vector<size_t> la(n), lb(n);
    std::vector<uint8_t> sq(n);
    uint64_t in_limbs = 0, out_limbs = 0;
    bool device = true;
    for (size_t i = 0; i < n; ++i) {
        const Request* q = group[i];
        la[i] = q->r.len_a;
        lb[i] = q->square ? q->r.len_a : q->r.len_b;
        sq[i] = q->square ? 1 : 0;
        in_limbs += la[i] + (q->square ? 0 : lb[i]);
        out_limbs += q->out_len;
        device = device && q->device;
    }
    Batch& B = slot.batch;
    if (!batch_plan(B, n, la.data(), lb.data(), sq.data())) return false;
    const hipStream_t s = B.ex.stream;
    std::vector<const uint64_t*> pa(n), pb(n);
    std::vector<uint64_t*> po(n);
    if (device) {
        for (size_t i = 0; i < n; ++i) {
            pa[i] = group[i]->r.a;
            pb[i] = group[i]->square ? group[i]->r.a : group[i]->r.b;
            po[i] = group[i]->r.dst;
        }
    } else {
        if (!slot.host_in.ensure(in_limbs * 8) || !slot.host_out.ensure(out_limbs * 8) ||
            !slot.dev_in.ensure(in_limbs * 8) || !slot.dev_out.ensure(out_limbs * 8))
            return false;
        uint64_t* hin = slot.host_in.as<uint64_t>();
        uint64_t* din = slot.dev_in.as<uint64_t>();
        uint64_t* dout = slot.dev_out.as<uint64_t>();
        uint64_t ofs = 0, oofs = 0;
        for (size_t i = 0; i < n; ++i) {
            const Request* q = group[i];
            memcpy(hin + ofs, q->r.a, la[i] * 8);
            pa[i] = din + ofs;
            ofs += la[i];
            if (q->square) {
                pb[i] = pa[i];
            } else {
                memcpy(hin + ofs, q->r.b, lb[i] * 8);
                pb[i] = din + ofs;
                ofs += lb[i];
            }
            po[i] = dout + oofs;
            oofs += q->out_len;
        }
        if (hipMemcpyAsync(din, hin, in_limbs * 8, hipMemcpyHostToDevice, s) != hipSuccess) return false;
    }
    if (!batch_bind(B, pa.data(), pb.data(), po.data())) return false;
    if (!batch_run(B, s, false)) {
        hipStreamSynchronize(s);
        return false;
    }
    if (!device) {
        uint64_t* hout = slot.host_out.as<uint64_t>();
        if (hipMemcpyAsync(hout, slot.dev_out.ptr, out_limbs * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
            return false;
        if (hipStreamSynchronize(s) != hipSuccess) return false;
        uint64_t oofs = 0;
        for (size_t i = 0; i < n; ++i) {
            memcpy(group[i]->r.dst, hout + oofs, group[i]->out_len * 8);
            oofs += group[i]->out_len;
        }
        return true;
    }
    return hipStreamSynchronize(s) == hipSuccess;
}

// Validates a request and sizes its working set.
static bool prepare_request(const ntt_request* r, Request* q) {
    q->r = *r;
    q->device = (r->flags & NTT_REQ_DEVICE) != 0;
    if (!r->a || !r->dst || r->len_a == 0) return false;
    switch (r->kind) {
        case NTT_REQ_SQR:
            q->square = true;
            q->r.b = r->a;
            q->r.len_b = r->len_a;
            break;
        case NTT_REQ_MUL:
            if (!r->b || r->len_b == 0) return false;
            q->square = r->a == r->b && r->len_a == r->len_b;
            break;
        case NTT_REQ_CONV:
            if (!r->b || r->len_b == 0) return false;
            break;
        default:
            return false;
    }
    const uint64_t la = q->r.len_a, lb = q->r.len_b;
    if (r->kind == NTT_REQ_CONV) {
        if (!ntt_conv_plan_new(la, lb, r->max_a, r->max_b, (r->flags & NTT_REQ_SIGNED) != 0, &q->plan)) return false;
        if (!conv_geometry(&q->plan, &q->g)) return false;
        q->out_len = q->plan.output_len;
        q->square = r->a == r->b && la == lb && !q->plan.schoolbook;
    } else {
        q->small = la + lb <= kSmallLimbs;
        if (!select_product(la, lb, &q->g, q->small)) return false;
        q->out_len = la + lb;
    }
    if (r->dst_len < q->out_len) return false;
    uint64_t bytes = 0;
    if (!q->device) bytes += align_up(la * 8) + (q->square ? 0 : align_up(lb * 8)) + align_up(q->out_len * 8);
    if (!(r->kind == NTT_REQ_CONV && q->plan.schoolbook))
        bytes += align_up(op_spectra_bytes(q->g)) + (q->square ? 0 : align_up(op_spectra_bytes_b(q->g))) +
                 align_up(op_carry_bytes(q->out_len, q->g.school));
    q->bytes = bytes;
    return true;
}

}  // namespace nttamd

using namespace nttamd;

struct ntt_server {
    Server s;
};

extern "C" {

ntt_server* ntt_server_new(uint64_t hbm_budget_bytes, uint32_t slots) {
    clear_error();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("server: no device");
        return nullptr;
    }
    if (hbm_budget_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            set_error("server: hipMemGetInfo failed");
            return nullptr;
        }
        hbm_budget_bytes = free_b / 2;
    }
    if (slots == 0) slots = 4;
    if (slots > 16) slots = 16;
    ntt_server* h = new (std::nothrow) ntt_server();
    if (!h) return nullptr;
    Server& S = h->s;
    S.device = dev;
    if (!S.arena.init(hbm_budget_bytes)) {
        set_error("server: device budget allocation failed");
        delete h;
        return nullptr;
    }
    S.stats.budget_bytes = S.arena.capacity();
    for (uint32_t i = 0; i < slots; ++i) {
        S.slots.emplace_back(new Slot());
        if (!S.slots.back()->batch.ex.init()) {
            set_error("server: stream creation failed");
            delete h;
            return nullptr;
        }
    }
    for (uint32_t i = 0; i < slots; ++i) S.threads.emplace_back(&Server::worker, &S, (int)i);
    return h;
}

void ntt_server_free(ntt_server* h) { delete h; }

int64_t ntt_server_submit(ntt_server* h, const ntt_request* r) {
    clear_error();
    if (!h || !r) return -1;
    Request* q = new (std::nothrow) Request();
    if (!q) return -1;
    if (!prepare_request(r, q)) {
        delete q;
        set_error("server: request rejected (shape, bounds or buffer lengths)");
        return -1;
    }
    Server& S = h->s;
    if (!q->small && q->bytes > S.arena.capacity()) {
        delete q;
        set_error("server: request exceeds the device budget");
        return -1;
    }
    {
        std::lock_guard<std::mutex> g(S.lock);
        q->id = S.next_id++;
        q->t_submit = now_s();
        S.requests[q->id] = q;
        S.queue.push_back(q);
        ++S.stats.submitted;
    }
    S.work.notify_one();
    return q->id;
}

int ntt_server_poll(ntt_server* h, int64_t ticket) {
    if (!h) return -1;
    std::lock_guard<std::mutex> g(h->s.lock);
    auto it = h->s.requests.find(ticket);
    return it == h->s.requests.end() ? -1 : it->second->status;
}

int ntt_server_wait(ntt_server* h, int64_t ticket) {
    if (!h) return -1;
    Server& S = h->s;
    std::unique_lock<std::mutex> g(S.lock);
    auto it = S.requests.find(ticket);
    if (it == S.requests.end()) return -1;
    Request* q = it->second;
    S.done.wait(g, [&] { return q->status != 0; });
    return q->status;
}

bool ntt_server_drain(ntt_server* h) {
    if (!h) return false;
    Server& S = h->s;
    std::unique_lock<std::mutex> g(S.lock);
    S.done.wait(g, [&] { return S.stats.completed + S.stats.failed == S.stats.submitted; });
    return S.stats.failed == 0;
}

bool ntt_server_times(ntt_server* h, int64_t ticket, double* submit_s, double* start_s, double* done_s) {
    if (!h) return false;
    std::lock_guard<std::mutex> g(h->s.lock);
    auto it = h->s.requests.find(ticket);
    if (it == h->s.requests.end()) return false;
    const Request* q = it->second;
    if (submit_s) *submit_s = q->t_submit;
    if (start_s) *start_s = q->t_start;
    if (done_s) *done_s = q->t_done;
    return true;
}

void ntt_server_release(ntt_server* h, int64_t ticket) {
    if (!h) return;
    std::lock_guard<std::mutex> g(h->s.lock);
    auto it = h->s.requests.find(ticket);
    if (it == h->s.requests.end() || it->second->status == 0) return;
    delete it->second;
    h->s.requests.erase(it);
}

void ntt_server_get_stats(ntt_server* h, ntt_server_stats* out) {
    if (!h || !out) return;
    std::lock_guard<std::mutex> g(h->s.lock);
    *out = h->s.stats;
    out->peak_bytes = h->s.arena.peak();
}

double ntt_clock_s(void) { return now_s(); }

void* ntt_host_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void ntt_host_free(void* p) {
    if (p) hipHostFree(p);
}

}  // extern "C"

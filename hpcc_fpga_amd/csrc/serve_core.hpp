// Request loop of the serving engine: queue, admission by HBM budget,
// concurrent execution slots, resident factors and tickets.
//
// Host-only C++ (no HIP): the device work is behind the Executor interface, so
// the same admission logic runs under the HIP executor of the library and
// under a CPU stand-in in the tests.
//
// Continuous batching: `slots` worker threads each run one solve at a time on
// their own device stream.  A request is admitted, strictly in arrival order,
// as soon as a worker is free and the memory of its factor fits in the arena;
// requests that arrive while others run are admitted as slots and memory free
// up, and their device work overlaps.  A request whose factor can never fit
// (it exceeds the arena, or it waits with nothing in flight that could free
// memory) fails with HPL_ERR_BUDGET instead of blocking the queue.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/hpl_hip.h"

namespace hplserve {

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Best-fit sub-allocator over one contiguous region, offsets aligned to
// kAlign; freed neighbours are coalesced.
class Arena {
   public:
    static constexpr size_t kAlign = 256;
    static size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

    void init(size_t size) {
        size_ = size / kAlign * kAlign;
        free_.clear();
        by_size_.clear();
        used_ = peak_ = 0;
        if (size_) {
            free_[0] = size_;
            by_size_.emplace(size_, 0);
        }
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = round_up(std::max<size_t>(bytes, 1));
        auto it = by_size_.lower_bound(bytes);
        if (it == by_size_.end()) return false;
        const size_t sz = it->first, o = it->second;
        by_size_.erase(it);
        free_.erase(o);
        if (sz > bytes) {
            free_[o + bytes] = sz - bytes;
            by_size_.emplace(sz - bytes, o + bytes);
        }
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        *off = o;
        return true;
    }
    void release(size_t off, size_t bytes) {
        bytes = round_up(std::max<size_t>(bytes, 1));
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            erase_size(next->second, next->first);
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                erase_size(prev->second, prev->first);
                off = prev->first;
                bytes += prev->second;
                free_.erase(prev);
            }
        }
        free_[off] = bytes;
        by_size_.emplace(bytes, off);
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const { return by_size_.empty() ? 0 : by_size_.rbegin()->first; }
    size_t free_blocks() const { return free_.size(); }

   private:
    void erase_size(size_t sz, size_t off) {
        auto r = by_size_.equal_range(sz);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == off) {
                by_size_.erase(it);
                return;
            }
    }
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;          // offset -> size
    std::multimap<size_t, size_t> by_size_;  // size -> offset
};

// Shape of a system on the device, fixed by its order alone (so a request is
// computed the same way whatever else is running).
struct Plan {
    uint64_t npad = 0, lda = 0;
    uint32_t block = 0;
    size_t factor_bytes = 0;  // original + factor + inverses
};

struct Factor {
    uint64_t id = 0;
    uint64_t n = 0;
    Plan plan;
    hpl_prec prec = HPL_PREC_FP32;
    size_t off = 0;
    int users = 0;          // queued or running requests that hold it
    bool kept = false;      // resident for later requests
    bool ready = false;     // factorised successfully
    bool released = false;  // the client gave it up
};

struct Job {
    uint64_t ticket = 0;
    hpl_request req{};
    Factor* factor = nullptr;
    bool fresh = false;  // factorises a new system
    double t_submit = 0, t_admit = 0, t_done = 0;
    hpl_response resp{};
    std::string err;
    bool done = false;
};

class Executor {
   public:
    virtual ~Executor() = default;
    virtual Plan plan(uint64_t n) const = 0;
    // Run job on worker `slot`; the factor's memory starts at arena offset
    // factor->off.  Fills job.resp (except the times) and job.err.
    virtual hpl_status run(int slot, Job& job) = 0;
};

class ServeCore {
   public:
    ServeCore(Executor* ex, size_t arena_bytes, int slots, int max_nrhs, uint64_t max_n,
              hpl_prec prec)
        : ex_(ex), max_nrhs_(max_nrhs), max_n_(max_n), prec_(prec) {
        arena_.init(arena_bytes);
        for (int s = 0; s < slots; ++s) workers_.emplace_back([this, s] { worker_loop(s); });
    }
    ~ServeCore() { shutdown(); }

    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (stop_) return;
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto& t : workers_) t.join();
        workers_.clear();
    }

    hpl_status submit(const hpl_request& r, uint64_t* ticket, std::string* err) {
        if (!ticket) return HPL_ERR_INVALID_ARG;
        if (r.n == 0 || r.n > max_n_ || r.nrhs < 1 || r.nrhs > max_nrhs_ || !r.b || !r.x ||
            r.ldb < r.n || r.ldx < r.n || (r.a && r.lda < r.n)) {
            *err = "bad request (n=" + std::to_string(r.n) + ", nrhs=" + std::to_string(r.nrhs) +
                   ", max_n=" + std::to_string(max_n_) + ")";
            return HPL_ERR_INVALID_ARG;
        }
        auto job = std::make_unique<Job>();
        job->req = r;
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) {
            *err = "server is shutting down";
            return HPL_ERR_STATE;
        }
        if (r.factor) {
            auto it = factors_.find(r.factor);
            if (it == factors_.end() || it->second->released || !it->second->ready) {
                *err = "factor " + std::to_string(r.factor) + " is not resident";
                return HPL_ERR_STATE;
            }
            if (it->second->n != r.n) {
                *err = "factor order " + std::to_string(it->second->n) + " != request order " +
                       std::to_string(r.n);
                return HPL_ERR_INVALID_ARG;
            }
            job->factor = it->second.get();
            job->factor->users++;
        } else {
            const Plan p = ex_->plan(r.n);
            if (Arena::round_up(p.factor_bytes) > arena_.size()) {
                *err = "system of order " + std::to_string(r.n) + " needs " +
                       std::to_string(p.factor_bytes) + " bytes, more than the " +
                       std::to_string(arena_.size()) + "-byte factor budget";
                return HPL_ERR_BUDGET;
            }
            job->fresh = true;
        }
        job->ticket = ++next_ticket_;
        job->t_submit = now_s();
        *ticket = job->ticket;
        queue_.push_back(job.get());
        jobs_.emplace(job->ticket, std::move(job));
        submitted_++;
        cv_work_.notify_one();
        return HPL_SUCCESS;
    }

    // Block until the ticket's request is done; the ticket is then retired.
    hpl_status wait(uint64_t ticket, hpl_response* out, std::string* err) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = jobs_.find(ticket);
        if (it == jobs_.end()) {
            *err = "unknown ticket " + std::to_string(ticket);
            return HPL_ERR_INVALID_ARG;
        }
        Job* j = it->second.get();
        cv_done_.wait(lk, [&] { return j->done; });
        if (out) *out = j->resp;
        if (j->resp.status != HPL_SUCCESS) *err = j->err;
        const hpl_status st = j->resp.status;
        // the wait released the lock: a submit meanwhile may have rehashed
        // jobs_ and invalidated `it`, so erase by key
        jobs_.erase(ticket);
        return st;
    }

    hpl_status release(uint64_t id, std::string* err) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = factors_.find(id);
        if (it == factors_.end() || it->second->released) {
            *err = "factor " + std::to_string(id) + " is not resident";
            return HPL_ERR_INVALID_ARG;
        }
        it->second->released = true;
        if (it->second->users == 0) drop_factor(it);
        cv_work_.notify_all();
        return HPL_SUCCESS;
    }

    void stats(hpl_serve_stats* s) {
        std::lock_guard<std::mutex> lk(mu_);
        s->submitted = submitted_;
        s->completed = completed_;
        s->failed = failed_;
        s->queued = (int)queue_.size();
        s->in_flight = in_flight_;
        s->max_in_flight = max_in_flight_;
        uint64_t resident = 0;
        for (auto& f : factors_)
            if (f.second->kept && f.second->ready && !f.second->released) resident++;
        s->resident_factors = resident;
        s->factor_bytes_in_use = arena_.used();
        s->peak_factor_bytes = arena_.peak();
        s->factor_budget_bytes = arena_.size();
    }

    // Host pointer of an arena offset is the executor's business; the arena
    // only hands out offsets.
    const Arena& arena() const { return arena_; }

   private:
    using FactorMap = std::unordered_map<uint64_t, std::unique_ptr<Factor>>;

    void drop_factor(FactorMap::iterator it) {
        arena_.release(it->second->off, it->second->plan.factor_bytes);
        factors_.erase(it);
    }

    // Admit the head of the queue if it fits; called with mu_ held.
    bool try_admit(Job* j) {
        if (!j->fresh) return true;
        const Plan p = ex_->plan(j->req.n);
        size_t off = 0;
        if (!arena_.alloc(p.factor_bytes, &off)) return false;
        auto f = std::make_unique<Factor>();
        f->id = ++next_factor_;
        f->n = j->req.n;
        f->plan = p;
        f->prec = prec_;
        f->off = off;
        f->users = 1;
        f->kept = j->req.keep != 0;
        j->factor = f.get();
        factors_.emplace(f->id, std::move(f));
        return true;
    }

    void finish(Job* j, hpl_status st) {
        j->resp.status = st;
        j->done = true;
        completed_++;
        if (st != HPL_SUCCESS) failed_++;
        cv_done_.notify_all();
    }

    void worker_loop(int slot) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            Job* j = nullptr;
            while (!j) {
                if (!queue_.empty()) {
                    Job* head = queue_.front();
                    if (try_admit(head)) {
                        queue_.pop_front();
                        j = head;
                        break;
                    }
                    if (in_flight_ == 0) {
                        // nothing running can free memory: the head cannot be served
                        queue_.pop_front();
                        head->err = "factor memory exhausted by resident factors (budget " +
                                    std::to_string(arena_.size()) + " bytes)";
                        head->t_admit = head->t_done = now_s();
                        fill_times(head);
                        finish(head, HPL_ERR_BUDGET);
                        continue;
                    }
                }
                if (stop_ && queue_.empty()) return;
                cv_work_.wait(lk);
            }
            j->t_admit = now_s();
            in_flight_++;
            max_in_flight_ = std::max(max_in_flight_, in_flight_);
            lk.unlock();
            hpl_status st = ex_->run(slot, *j);
            lk.lock();
            in_flight_--;
            j->t_done = now_s();
            fill_times(j);
            Factor* f = j->factor;
            if (j->fresh && st == HPL_SUCCESS) f->ready = true;
            j->resp.factor = (f->kept && f->ready) ? f->id : 0;
            f->users--;
            if (f->users == 0 && (!f->kept || !f->ready || f->released))
                drop_factor(factors_.find(f->id));
            finish(j, st);
            cv_work_.notify_all();
        }
    }

    static void fill_times(Job* j) {
        j->resp.queue_s = j->t_admit - j->t_submit;
        j->resp.service_s = j->t_done - j->t_admit;
        j->resp.latency_s = j->t_done - j->t_submit;
        j->resp.reused = j->fresh ? 0 : 1;
    }

    Executor* ex_;
    const int max_nrhs_;
    const uint64_t max_n_;
    const hpl_prec prec_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Job*> queue_;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs_;
    FactorMap factors_;
    Arena arena_;
    std::vector<std::thread> workers_;
    bool stop_ = false;
    uint64_t next_ticket_ = 0, next_factor_ = 0;
    uint64_t submitted_ = 0, completed_ = 0, failed_ = 0;
    int in_flight_ = 0, max_in_flight_ = 0;
};

}  // namespace hplserve

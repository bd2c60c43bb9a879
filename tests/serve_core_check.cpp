// CPU check of the serving engine's host logic (hpcc_fpga_amd/csrc/serve_core.hpp):
// the arena allocator and the request loop, driven by a stand-in executor
// that "solves" x = 2 b after a short sleep.  Built and run by
// tests/test_serve_core.py; prints "OK" and exits 0 when every check holds.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../hpcc_fpga_amd/csrc/serve_core.hpp"

using namespace hplserve;

#define CHECK(cond)                                                         \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            std::exit(1);                                                   \
        }                                                                   \
    } while (0)

// ---------------------------------------------------------------- arena
static void check_arena() {
    Arena a;
    const size_t size = 1 << 20;
    a.init(size);
    std::mt19937 g(3);
    std::map<size_t, size_t> live;  // off -> rounded size
    for (int step = 0; step < 20000; ++step) {
        if (live.empty() || (g() % 3 != 0 && live.size() < 200)) {
            const size_t want = 1 + g() % 9000;
            size_t off;
            if (a.alloc(want, &off)) {
                const size_t sz = Arena::round_up(want);
                CHECK(off % Arena::kAlign == 0 && off + sz <= a.size());
                // no overlap with live blocks
                auto nx = live.lower_bound(off);
                if (nx != live.end()) CHECK(off + sz <= nx->first);
                if (nx != live.begin()) {
                    auto pv = std::prev(nx);
                    CHECK(pv->first + pv->second <= off);
                }
                live[off] = sz;
            } else {
                CHECK(a.largest_free() < Arena::round_up(want));
            }
        } else {
            auto it = live.begin();
            std::advance(it, g() % live.size());
            a.release(it->first, it->second);
            live.erase(it);
        }
        size_t used = 0;
        for (auto& kv : live) used += kv.second;
        CHECK(a.used() == used);
        CHECK(a.peak() >= used);
    }
    for (auto& kv : live) a.release(kv.first, kv.second);
    CHECK(a.used() == 0);
    CHECK(a.free_blocks() == 1 && a.largest_free() == a.size());  // fully coalesced
}

// ---------------------------------------------------------------- loop
struct FakeExecutor : Executor {
    std::atomic<int> running{0}, max_running{0};
    std::mutex mu;
    std::vector<std::pair<uint64_t, double>> admits;  // (ticket, t_admit) in run order
    Plan plan(uint64_t n) const override {
        Plan p;
        p.block = 64;
        p.npad = (n + 63) / 64 * 64;
        p.lda = p.npad;
        p.factor_bytes = p.npad * 1000;
        return p;
    }
    hpl_status run(int, Job& j) override {
        const int r = ++running;
        int m = max_running.load();
        while (r > m && !max_running.compare_exchange_weak(m, r)) {
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            admits.emplace_back(j.ticket, j.t_admit);
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200 + j.req.n % 500));
        for (int c = 0; c < j.req.nrhs; ++c)
            for (uint64_t i = 0; i < j.req.n; ++i)
                j.req.x[c * j.req.ldx + i] = 2.0 * j.req.b[c * j.req.ldb + i];
        j.resp.iterations = 1;
        j.resp.backward_error = 0.0;
        --running;
        return HPL_SUCCESS;
    }
};

static hpl_request make_req(uint64_t n, int nrhs, std::vector<double>& b, std::vector<double>& x) {
    b.resize(n * nrhs);
    x.assign(n * nrhs, 0.0);
    for (size_t i = 0; i < b.size(); ++i) b[i] = (double)(i % 97) - 48.0;
    hpl_request r{};
    r.n = n;
    r.nrhs = nrhs;
    r.seed = n;
    r.b = b.data();
    r.ldb = n;
    r.x = x.data();
    r.ldx = n;
    return r;
}

static void check_loop() {
    FakeExecutor ex;
    const int slots = 4;
    // factors of up to 1024 * 1000 bytes; the arena holds about three of the largest
    ServeCore core(&ex, 3200000, slots, 16, 1024, HPL_PREC_FP8);
    std::mt19937 g(11);
    const int N = 300;
    std::vector<std::vector<double>> bs(N), xs(N);
    std::vector<uint64_t> tickets(N);
    std::vector<int> nrhs(N);
    std::vector<uint64_t> ns(N);
    std::string err;
    for (int i = 0; i < N; ++i) {
        ns[i] = 64 + g() % 961;
        nrhs[i] = 1 + g() % 16;
        hpl_request r = make_req(ns[i], nrhs[i], bs[i], xs[i]);
        CHECK(core.submit(r, &tickets[i], &err) == HPL_SUCCESS);
    }
    for (int i = 0; i < N; ++i) {
        hpl_response resp{};
        CHECK(core.wait(tickets[i], &resp, &err) == HPL_SUCCESS);
        CHECK(resp.status == HPL_SUCCESS && resp.reused == 0 && resp.factor == 0);
        CHECK(resp.latency_s >= resp.service_s && resp.queue_s >= 0);
        for (size_t k = 0; k < bs[i].size(); ++k) CHECK(xs[i][k] == 2.0 * bs[i][k]);
    }
    CHECK(ex.max_running.load() <= slots);
    CHECK(ex.max_running.load() >= 2);  // requests overlapped
    // admission in arrival order: t_admit non-decreasing in ticket order
    std::sort(ex.admits.begin(), ex.admits.end());
    for (size_t i = 1; i < ex.admits.size(); ++i)
        CHECK(ex.admits[i].second >= ex.admits[i - 1].second);
    hpl_serve_stats st{};
    core.stats(&st);
    CHECK(st.completed == (uint64_t)N && st.failed == 0 && st.in_flight == 0 && st.queued == 0);
    CHECK(st.factor_bytes_in_use == 0);                    // every transient factor freed
    CHECK(st.peak_factor_bytes <= st.factor_budget_bytes);  // never above the budget
    CHECK(st.max_in_flight <= slots);

    // a system larger than the whole factor budget is refused at submission
    std::vector<double> b, x;
    {
        ServeCore small(&ex, 100000, 2, 16, 1024, HPL_PREC_FP8);
        uint64_t t;
        hpl_request r = make_req(512, 1, b, x);
        CHECK(small.submit(r, &t, &err) == HPL_ERR_BUDGET);
    }

    // resident factors: keep, reuse, release
    {
        ServeCore c2(&ex, 3200000, 2, 16, 1024, HPL_PREC_FP8);
        uint64_t t;
        hpl_response resp{};
        hpl_request r = make_req(1000, 2, b, x);
        r.keep = 1;
        CHECK(c2.submit(r, &t, &err) == HPL_SUCCESS);
        CHECK(c2.wait(t, &resp, &err) == HPL_SUCCESS);
        const uint64_t fid = resp.factor;
        CHECK(fid != 0);
        c2.stats(&st);
        CHECK(st.resident_factors == 1 && st.factor_bytes_in_use >= 1024 * 1000);
        std::vector<double> b2, x2;
        hpl_request r2 = make_req(1000, 3, b2, x2);
        r2.factor = fid;
        CHECK(c2.submit(r2, &t, &err) == HPL_SUCCESS);
        CHECK(c2.wait(t, &resp, &err) == HPL_SUCCESS);
        CHECK(resp.reused == 1 && resp.factor == fid);
        // order mismatch is refused
        std::vector<double> b3, x3;
        hpl_request r3 = make_req(999, 1, b3, x3);
        r3.factor = fid;
        CHECK(c2.submit(r3, &t, &err) == HPL_ERR_INVALID_ARG);
        // fill the arena with kept factors: the next fresh request cannot be
        // admitted while nothing runs, and fails instead of blocking
        std::vector<uint64_t> kept{fid};
        for (int k = 0; k < 2; ++k) {
            hpl_request rk = make_req(1000, 1, b3, x3);
            rk.keep = 1;
            CHECK(c2.submit(rk, &t, &err) == HPL_SUCCESS);
            CHECK(c2.wait(t, &resp, &err) == HPL_SUCCESS);
            kept.push_back(resp.factor);
        }
        hpl_request rf = make_req(1000, 1, b3, x3);
        CHECK(c2.submit(rf, &t, &err) == HPL_SUCCESS);
        CHECK(c2.wait(t, &resp, &err) == HPL_ERR_BUDGET);
        // after a release it fits again
        CHECK(c2.release(kept[0], &err) == HPL_SUCCESS);
        CHECK(c2.release(kept[0], &err) == HPL_ERR_INVALID_ARG);
        hpl_request rr = make_req(1000, 1, b3, x3);
        rr.factor = kept[0];
        CHECK(c2.submit(rr, &t, &err) == HPL_ERR_STATE);  // released factor
        CHECK(c2.submit(rf, &t, &err) == HPL_SUCCESS);
        CHECK(c2.wait(t, &resp, &err) == HPL_SUCCESS);
        c2.stats(&st);
        CHECK(st.resident_factors == 2);
        // bad requests
        hpl_request bad = make_req(1000, 17, b3, x3);
        CHECK(c2.submit(bad, &t, &err) == HPL_ERR_INVALID_ARG);
        bad = make_req(2048, 1, b3, x3);
        CHECK(c2.submit(bad, &t, &err) == HPL_ERR_INVALID_ARG);
        CHECK(c2.wait(987654, &resp, &err) == HPL_ERR_INVALID_ARG);
    }
}

// A waiter blocks while another thread submits many more requests (the job
// table grows and rehashes during the wait); the waiter must still retire its
// own ticket and every other ticket stays valid.
static void check_concurrent_wait() {
    FakeExecutor ex;
    ServeCore core(&ex, 64000000, 2, 16, 1024, HPL_PREC_FP8);
    std::string err;
    const int N = 400;
    std::vector<std::vector<double>> bs(N), xs(N);
    std::vector<uint64_t> tickets(N);
    hpl_request r0 = make_req(1024, 16, bs[0], xs[0]);
    CHECK(core.submit(r0, &tickets[0], &err) == HPL_SUCCESS);
    std::atomic<int> waited{0};
    std::thread waiter([&] {
        hpl_response resp{};
        std::string e;
        CHECK(core.wait(tickets[0], &resp, &e) == HPL_SUCCESS);
        waited = 1;
    });
    for (int i = 1; i < N; ++i) {
        hpl_request r = make_req(64 + (uint64_t)i % 200, 1, bs[i], xs[i]);
        CHECK(core.submit(r, &tickets[i], &err) == HPL_SUCCESS);
    }
    waiter.join();
    CHECK(waited.load() == 1);
    for (int i = 1; i < N; ++i) {
        hpl_response resp{};
        CHECK(core.wait(tickets[i], &resp, &err) == HPL_SUCCESS);
        for (size_t k = 0; k < bs[i].size(); ++k) CHECK(xs[i][k] == 2.0 * bs[i][k]);
    }
    hpl_response resp{};
    CHECK(core.wait(tickets[0], &resp, &err) == HPL_ERR_INVALID_ARG);  // retired once
}

int main() {
    check_arena();
    check_loop();
    check_concurrent_wait();
    std::printf("OK\n");
    return 0;
}

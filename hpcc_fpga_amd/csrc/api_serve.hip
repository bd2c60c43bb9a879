// Serving engine: the C ABI hpl_serve_* and its HIP executor.
//
// The server holds one device allocation of hbm_budget bytes.  Its front part
// is split into one workspace per slot (panel copies, refinement buffers and
// the right-hand sides of the running request, sized for max_n and
// max_nrhs); the rest is the arena of factors (original, factor and per-block
// inverses of each admitted system).  Every slot owns a solver context in
// external mode whose buffers point into this allocation, so serving never
// calls hipMalloc or hipFree and never synchronises the device as a whole.
// A slot runs its solves on its own stream without look-ahead: no event waits
// cross streams, so slots sharing a hardware queue cannot wait on each other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hpl_hip.h"
#include "ctx.hpp"
#include "kernels.hpp"
#include "serve_core.hpp"

namespace {

using hplserve::Arena;
using hplserve::Job;
using hplserve::Plan;

thread_local std::string tl_err;  // hpl_serve_last_error of the calling thread

hpl_status set_err(hpl_status st, const std::string& msg) {
    tl_err = msg;
    return st;
}

size_t al(size_t b) { return Arena::round_up(b); }

uint32_t block_for(uint64_t n) { return n <= 256 ? 64 : n <= 2048 ? 128 : 256; }

struct Slot {
    hpl_ctx* c = nullptr;
    double* b64 = nullptr;  // right-hand sides, npad x nrhs
    double* x64 = nullptr;  // solutions, npad x nrhs
};

class HipExecutor : public hplserve::Executor {
   public:
    HipExecutor(const hpl_serve_config& cfg) : cfg_(cfg) {}

    Plan plan(uint64_t n) const override {
        Plan p;
        p.block = block_for(n);
        p.npad = (n + p.block - 1) / p.block * p.block;
        p.lda = hplx::device_lda(p.npad);
        const hplx::WsPlan w = hplx::plan_factor_ws(p.npad, p.block, cfg_.prec, 1);
        p.factor_bytes = 2 * al(p.npad * p.lda * sizeof(float)) + al(w.inv) + al(w.binv);
        return p;
    }

    // Workspace bytes of one slot: each buffer sized for the largest system
    // (over the block sizes in use), carved in the order of bind_slot.
    size_t slot_bytes() {
        SlotCaps& k = caps_;
        k = SlotCaps{};
        for (uint32_t B : {64u, 128u, 256u}) {
            if (B > 64 && block_for(cfg_.max_n) < B) continue;
            const uint64_t np = (cfg_.max_n + B - 1) / B * B;
            const hplx::WsPlan w = hplx::plan_factor_ws(np, B, cfg_.prec, 1);
            const hplx::IrPlan ir = hplx::plan_ir(np, cfg_.max_nrhs);
            k.lt = std::max(k.lt, w.lt);
            k.dt = std::max(k.dt, w.dt);
            k.inner = std::max(k.inner, w.inner);
            k.lp = std::max(k.lp, w.lp);
            k.sc = std::max(k.sc, w.sc);
            k.work = std::max(k.work, ir.work);
            k.r = std::max(k.r, ir.r);
            k.xs = std::max(k.xs, ir.xs);
            k.out = std::max(k.out, ir.out);
            k.vec = std::max(k.vec, (size_t)np * cfg_.max_nrhs * sizeof(double));
        }
        return al(k.lt) + al(k.dt) + al(k.inner) + 2 * al(k.lp) + 2 * al(k.sc) + al(k.work) +
               al(k.r) + al(k.xs) + al(k.out) + 2 * al(k.vec) + al(kStatusBytes);
    }

    // Point slot s's context at its workspace.
    void bind_slot(int s, char* base) {
        Slot& S = slots_[s];
        hpl_ctx* c = S.c;
        char* p = base;
        auto take = [&](size_t bytes) {
            char* q = bytes ? p : nullptr;
            p += al(bytes);
            return q;
        };
        const SlotCaps& k = caps_;
        c->dLt = reinterpret_cast<float*>(take(k.lt));
        c->capLt = k.lt;
        c->dDT = reinterpret_cast<float*>(take(k.dt));
        c->capDT = k.dt;
        c->dInner = reinterpret_cast<float*>(take(k.inner));
        c->capInner = k.inner;
        c->dLpT = take(k.lp);
        c->capLpT = k.lp;
        c->dLpL = take(k.lp);
        c->capLpL = k.lp;
        c->dScT = reinterpret_cast<float*>(take(k.sc));
        c->capScT = k.sc;
        c->dScL = reinterpret_cast<float*>(take(k.sc));
        c->capScL = k.sc;
        c->dWork = reinterpret_cast<double*>(take(k.work));
        c->capWork = k.work;
        c->dIrR = reinterpret_cast<double*>(take(k.r));
        c->capIrR = k.r;
        c->dIrXs = reinterpret_cast<float*>(take(k.xs));
        c->capIrXs = k.xs;
        c->dIrOut = reinterpret_cast<double*>(take(k.out));
        c->capIrOut = k.out;
        S.b64 = reinterpret_cast<double*>(take(k.vec));
        S.x64 = reinterpret_cast<double*>(take(k.vec));
        c->dStatus = reinterpret_cast<int*>(take(kStatusBytes));
        c->capStatus = kStatusBytes;
    }

    hpl_status make_slots(int count) {
        for (int s = 0; s < count; ++s) {
            Slot S;
            S.c = new hpl_ctx();
            S.c->device = cfg_.device;
            S.c->external = true;
            S.c->lookahead = false;
            S.c->depth_fp32 = S.c->depth_lp = 1;
            slots_.push_back(S);
            if (hplx::init_streams(S.c) != HPL_SUCCESS)
                return set_err(HPL_ERR_HIP, "stream creation failed for a serving slot");
        }
        return HPL_SUCCESS;
    }

    void set_factor_base(char* p) { factor_base_ = p; }

    ~HipExecutor() override {
        for (Slot& S : slots_) hpl_destroy(S.c);  // external: frees no buffers
    }

    hpl_status run(int slot, Job& job) override {
        Slot& S = slots_[slot];
        hpl_ctx* c = S.c;
        const hpl_request& rq = job.req;
        hplserve::Factor* f = job.factor;
        const Plan& p = f->plan;
        const uint64_t n = rq.n, np = p.npad, lda = p.lda;
        job.resp.block = p.block;
        job.resp.npad = np;
        if (hipSetDevice(c->device) != hipSuccess) return fail(job, c, HPL_ERR_HIP, "hipSetDevice");
        char* mem = factor_base_ + f->off;
        const size_t abytes = al(np * lda * sizeof(float));
        c->n = np;
        c->lda = lda;
        c->dA0 = reinterpret_cast<float*>(mem);
        c->capA0 = abytes;
        c->dA = reinterpret_cast<float*>(mem + abytes);
        c->capA = abytes;
        const hplx::WsPlan w = hplx::plan_factor_ws(np, p.block, f->prec, 1);
        c->dInv = reinterpret_cast<float*>(mem + 2 * abytes);
        c->capInv = w.inv;
        c->dBInv = reinterpret_cast<float*>(mem + 2 * abytes + al(w.inv));
        c->capBInv = w.binv;
        c->keep_orig = c->have_orig = true;
        hipStream_t s0 = c->s0;
        if (job.fresh) {
            hipError_t e;
            if (rq.a)
                e = hipMemcpy2DAsync(c->dA0, lda * sizeof(float), rq.a, rq.lda * sizeof(float),
                                     n * sizeof(float), n, hipMemcpyHostToDevice, s0);
            else
                e = launch_generate(c->dA0, (int64_t)lda, (int64_t)n, 0, rq.seed, 1.0, nullptr,
                                    nullptr, nullptr, s0);
            if (e == hipSuccess) e = launch_pad_identity(c->dA0, (int64_t)lda, (int64_t)n,
                                                         (int64_t)np, s0);
            if (e == hipSuccess)
                e = hipMemcpyAsync(c->dA, c->dA0, np * lda * sizeof(float),
                                   hipMemcpyDeviceToDevice, s0);
            if (e != hipSuccess) return fail(job, c, HPL_ERR_HIP, hipGetErrorString(e));
            const hpl_status st = hpl_gefa_dev(c, p.block, f->prec, nullptr);
            if (st) return fail(job, c, st, c->err.c_str());
        } else {
            // a resident factor: its block inverses were computed by the
            // request that factorised it
            c->B = p.block;
            c->prec = f->prec;
            c->factored = true;
            c->binv_ready = true;
        }
        const int nr = rq.nrhs;
        hipError_t e = hipSuccess;
        if (np > n) e = launch_fill_f64(S.b64, (int64_t)(np * nr), 0.0, s0);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(S.b64, np * sizeof(double), rq.b, rq.ldb * sizeof(double),
                                 n * sizeof(double), (size_t)nr, hipMemcpyHostToDevice, s0);
        if (e != hipSuccess) return fail(job, c, HPL_ERR_HIP, hipGetErrorString(e));
        hpl_ir_result res[16];
        hpl_status st =
            hplx::ir_solve(c, nr, S.b64, S.x64, cfg_.max_iter, cfg_.tol_per_n * (double)n, res);
        if (st != HPL_SUCCESS && st != HPL_ERR_NOT_CONVERGED) return fail(job, c, st, c->err.c_str());
        e = hipMemcpy2DAsync(rq.x, rq.ldx * sizeof(double), S.x64, np * sizeof(double),
                             n * sizeof(double), (size_t)nr, hipMemcpyDeviceToHost, s0);
        if (e == hipSuccess) e = hipStreamSynchronize(s0);
        if (e != hipSuccess) return fail(job, c, HPL_ERR_HIP, hipGetErrorString(e));
        job.resp.iterations = 0;
        job.resp.backward_error = 0.0;
        for (int j = 0; j < nr; ++j) {
            job.resp.iterations = std::max(job.resp.iterations, res[j].iterations);
            job.resp.backward_error = std::max(job.resp.backward_error, res[j].backward_error);
        }
        if (st) job.err = c->err;
        return st;
    }

   private:
    static hpl_status fail(Job& job, hpl_ctx* c, hpl_status st, const char* msg) {
        job.err = msg ? msg : "";
        (void)hipGetLastError();
        (void)hipStreamSynchronize(c->s0);
        return st;
    }

    hpl_serve_config cfg_;
    std::vector<Slot> slots_;
    char* factor_base_ = nullptr;
    struct SlotCaps {
        size_t lt = 0, dt = 0, inner = 0, lp = 0, sc = 0, work = 0, r = 0, xs = 0, out = 0,
               vec = 0;
    } caps_;
    static constexpr size_t kStatusBytes = 64;
};

}  // namespace

struct hpl_server {
    hpl_serve_config cfg{};
    char* mem = nullptr;
    size_t ws_bytes = 0;
    std::unique_ptr<HipExecutor> ex;
    std::unique_ptr<hplserve::ServeCore> core;
};

extern "C" {

void hpl_serve_config_default(hpl_serve_config* cfg) {
    if (!cfg) return;
    cfg->device = 0;
    cfg->slots = 4;
    cfg->hbm_budget = 200000000000ULL;
    cfg->prec = HPL_PREC_FP8;
    cfg->max_iter = 30;
    cfg->tol_per_n = 1e-14;
    cfg->max_n = 8192;
    cfg->max_nrhs = 16;
}

hpl_status hpl_serve_create(const hpl_serve_config* cfg, hpl_server** out) {
    if (!cfg || !out) return set_err(HPL_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 8 || cfg->max_nrhs < 1 || cfg->max_nrhs > 16 ||
        cfg->max_n < 1 || cfg->max_n > 65536 || cfg->max_iter < 0 || !(cfg->tol_per_n > 0) ||
        (cfg->prec != HPL_PREC_FP32 && cfg->prec != HPL_PREC_BF16 && cfg->prec != HPL_PREC_FP16 &&
         cfg->prec != HPL_PREC_FP8))
        return set_err(HPL_ERR_INVALID_ARG,
                       "bad serving configuration (slots 1..8, nrhs 1..16, n 1..65536)");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || cfg->device < 0 || cfg->device >= count)
        return set_err(HPL_ERR_NO_DEVICE, "no such HIP device");
    if (hipSetDevice(cfg->device) != hipSuccess) return set_err(HPL_ERR_NO_DEVICE, "hipSetDevice");
    auto srv = std::make_unique<hpl_server>();
    srv->cfg = *cfg;
    srv->ex = std::make_unique<HipExecutor>(*cfg);
    srv->ws_bytes = srv->ex->slot_bytes();
    const size_t ws_total = srv->ws_bytes * (size_t)cfg->slots;
    if (ws_total + srv->ex->plan(std::min<uint64_t>(cfg->max_n, 64)).factor_bytes >
        cfg->hbm_budget)
        return set_err(HPL_ERR_BUDGET, "hbm_budget " + std::to_string(cfg->hbm_budget) +
                                           " below the slot workspaces (" +
                                           std::to_string(ws_total) + " bytes)");
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&srv->mem), cfg->hbm_budget);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        srv->mem = nullptr;
        return set_err(HPL_ERR_OUT_OF_MEMORY,
                       std::string("hipMalloc of the serving budget failed: ") +
                           hipGetErrorString(e));
    }
    hpl_status st = srv->ex->make_slots(cfg->slots);
    if (st) {
        srv->ex.reset();
        (void)hipFree(srv->mem);
        return st;
    }
    for (int s = 0; s < cfg->slots; ++s) srv->ex->bind_slot(s, srv->mem + s * srv->ws_bytes);
    srv->ex->set_factor_base(srv->mem + ws_total);
    srv->core = std::make_unique<hplserve::ServeCore>(srv->ex.get(), cfg->hbm_budget - ws_total,
                                                      cfg->slots, cfg->max_nrhs, cfg->max_n,
                                                      cfg->prec);
    *out = srv.release();
    return HPL_SUCCESS;
}

void hpl_serve_destroy(hpl_server* srv) {
    if (!srv) return;
    (void)hipSetDevice(srv->cfg.device);
    if (srv->core) srv->core->shutdown();
    srv->core.reset();
    srv->ex.reset();
    if (srv->mem) (void)hipFree(srv->mem);
    delete srv;
}

const char* hpl_serve_last_error(const hpl_server*) { return tl_err.c_str(); }

hpl_status hpl_serve_submit(hpl_server* srv, const hpl_request* req, uint64_t* ticket) {
    if (!srv || !req || !ticket) return set_err(HPL_ERR_INVALID_ARG, "null argument");
    std::string err;
    const hpl_status st = srv->core->submit(*req, ticket, &err);
    return st ? set_err(st, err) : st;
}

hpl_status hpl_serve_wait(hpl_server* srv, uint64_t ticket, hpl_response* resp) {
    if (!srv) return set_err(HPL_ERR_INVALID_ARG, "null argument");
    std::string err;
    const hpl_status st = srv->core->wait(ticket, resp, &err);
    return st ? set_err(st, err) : st;
}

hpl_status hpl_serve_release(hpl_server* srv, uint64_t factor) {
    if (!srv) return set_err(HPL_ERR_INVALID_ARG, "null argument");
    std::string err;
    const hpl_status st = srv->core->release(factor, &err);
    return st ? set_err(st, err) : st;
}

hpl_status hpl_serve_stats_get(hpl_server* srv, hpl_serve_stats* s) {
    if (!srv || !s) return set_err(HPL_ERR_INVALID_ARG, "null argument");
    std::memset(s, 0, sizeof *s);
    srv->core->stats(s);
    s->workspace_bytes = srv->ws_bytes * (uint64_t)srv->cfg.slots;
    return HPL_SUCCESS;
}

}  // extern "C"

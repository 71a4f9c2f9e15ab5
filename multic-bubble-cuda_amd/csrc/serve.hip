// Serving loop of libmcb (mcb_serve_run, include/mcb.h): bubble cases that
// arrive over time share one device as one continuously re-formed batch.
//
// Each scheduling round: requests that have arrived join the waiting queue;
// the head of the queue is admitted while its footprint fits the device-memory
// budget next to the resident cases (arrival order, no overtaking); the
// resident cases then advance together by min(tick, fewest steps left) steps
// (one batch launch per pass, graph-replayed); cases that reached their step
// count are read out (FP64 sum of phi and kinetic energy, optional phi) and
// leave, returning their memory to a pool that later cases of the same size
// reuse. Cases never exchange data, so a request's numbers are the same
// whatever else is resident.
//
// With cfg->sequential the same trace is served one request at a time by
// single-case contexts (the plain fused path for FP32), one kept per lattice
// shape: the baseline the batched loop is compared with.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <tuple>
#include <thread>
#include <vector>

#include "batch.hpp"

struct mcb_ctx;

namespace mcb {
mcb_status ctx_init_bubbles_device(mcb_ctx* c, const BubbleSet& s);
mcb_status ctx_reset_case(mcb_ctx* c, const mcb_params& p, const BubbleSet& s);
mcb_status host_validate(const mcb_params* p, std::string* why);
}  // namespace mcb

using namespace mcb;

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

void wait_until(Clock::time_point t0, double t) {
    const double now = seconds_since(t0);
    if (t > now) std::this_thread::sleep_for(std::chrono::duration<double>(t - now));
}

bool bubbles_of(const mcb_case& c, BubbleSet* s) {
    if (c.n_bubbles < 1 || c.n_bubbles > kMaxBubbleSet) return false;
    for (int i = 0; i < c.n_bubbles; ++i) s->b[i] = c.bubbles[i];
    s->n = c.n_bubbles;
    s->width = c.width;
    return true;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // Nearest rank.
    const size_t k = (size_t)std::ceil(q * (double)v.size());
    return v[std::min(v.size() - 1, k == 0 ? 0 : k - 1)];
}

void finish_stats(const mcb_request* reqs, int n, const mcb_request_result* res, mcb_serve_stats* st) {
    std::vector<double> ttr;
    double first = 1e300, last = 0.0, lu = 0.0;
    int done = 0, rej = 0;
    for (int i = 0; i < n; ++i) {
        first = std::min(first, reqs[i].arrival_s);
        if (res[i].status == MCB_OK) {
            ++done;
            ttr.push_back(res[i].finish_s - res[i].arrival_s);
            last = std::max(last, res[i].finish_s);
            lu += (double)res[i].lattice_updates;
        } else {
            ++rej;
        }
    }
    st->completed = done;
    st->rejected = rej;
    st->lattice_updates = lu;
    st->makespan_s = done ? last - first : 0.0;
    st->mlups = st->makespan_s > 0 ? lu / st->makespan_s / 1e6 : 0.0;
    st->p50_s = percentile(ttr, 0.50);
    st->p99_s = percentile(ttr, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double t : ttr) {
        sum += t;
        mx = std::max(mx, t);
    }
    st->mean_s = ttr.empty() ? 0.0 : sum / (double)ttr.size();
    st->max_s = mx;
}

mcb_params request_params(const mcb_request& r, int storage, int device) {
    mcb_params p = r.c.params;
    p.storage = storage;
    p.device = device;
    p.bc_mode = MCB_BC_PERIODIC;
    return p;
}

// One request at a time on single-case contexts. A context is kept per
// lattice shape and re-targeted at each request of that shape (parameters and
// initial state reset on the device), so the baseline pays allocation once per
// shape, not once per request; peak_bytes counts every context it holds.
mcb_status serve_sequential(const mcb_serve_config& cfg, const mcb_request* reqs, int n, mcb_request_result* res,
                            mcb_serve_stats* st) {
    std::map<std::tuple<int, int, int>, mcb_ctx*> ctxs;
    double held = 0.0;
    const Clock::time_point t0 = Clock::now();
    for (int i = 0; i < n; ++i) {
        mcb_request_result& r = res[i];
        wait_until(t0, reqs[i].arrival_s);
        r.admit_s = seconds_since(t0);
        mcb_params p = request_params(reqs[i], cfg.storage, cfg.device);
        p.path = MCB_PATH_FUSED;
        BubbleSet s{};
        mcb_status rs = bubbles_of(reqs[i].c, &s) && reqs[i].c.nsteps >= 1 ? MCB_OK : MCB_ERR_INVALID_ARG;
        if (rs == MCB_OK) rs = host_validate(&p, nullptr);
        mcb_ctx* ctx = nullptr;
        if (rs == MCB_OK) {
            const auto key = std::make_tuple(p.nx, p.ny, p.nz);
            auto it = ctxs.find(key);
            if (it == ctxs.end()) {
                rs = mcb_create(&p, nullptr, &ctx);
                if (rs == MCB_OK) {
                    ctxs.emplace(key, ctx);
                    held += (double)mcb_footprint_bytes(&p);
                    rs = ctx_init_bubbles_device(ctx, s);
                }
            } else {
                ctx = it->second;
                rs = ctx_reset_case(ctx, p, s);
            }
        }
        if (rs == MCB_OK) rs = mcb_step(ctx, reqs[i].c.nsteps);
        if (rs == MCB_OK) rs = mcb_compute_macros(ctx);
        if (rs == MCB_OK) rs = mcb_reduce(ctx, &r.sum_phi, &r.kinetic);
        if (rs == MCB_OK && reqs[i].phi_out)
            rs = mcb_read_field(ctx, MCB_FIELD_PHI, reqs[i].phi_out, (size_t)p.nx * p.ny * p.nz);
        r.status = rs;
        r.finish_s = seconds_since(t0);
        r.lattice_updates = rs == MCB_OK ? (long long)p.nx * p.ny * p.nz * reqs[i].c.nsteps : 0;
        st->rounds += 1;
        st->peak_cases = 1;
        st->peak_bytes = std::max(st->peak_bytes, held);
    }
    for (auto& kv : ctxs) mcb_destroy(kv.second);
    finish_stats(reqs, n, res, st);
    return MCB_OK;
}

struct Resident {
    int req;
    int id;
    long long left;
    size_t bytes;
};

mcb_status serve_batched(const mcb_serve_config& cfg, const mcb_request* reqs, int n, mcb_request_result* res,
                         mcb_serve_stats* st, hipStream_t stream) {
    size_t budget = cfg.hbm_budget;
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return MCB_ERR_HIP;
        budget = (size_t)(0.9 * (double)free_b);
    }
    const int tick = cfg.tick_steps > 0 ? cfg.tick_steps : 100;
    Batch batch(cfg.device, stream, true);
    // Freed case buffers are kept for later cases of the same sizes, within
    // the budget that resident cases leave unused.
    batch.set_pool_limit(budget);
    std::vector<Resident> active;
    std::deque<int> waiting;
    size_t in_use = 0;
    int next = 0, settled = 0;
    const Clock::time_point t0 = Clock::now();
    while (settled < n) {
        const double now = seconds_since(t0);
        while (next < n && reqs[next].arrival_s <= now) waiting.push_back(next++);
        // Admission, in arrival order.
        const Clock::time_point ta = Clock::now();
        while (!waiting.empty()) {
            const int i = waiting.front();
            const mcb_params p = request_params(reqs[i], cfg.storage, cfg.device);
            const size_t need = case_bytes(p);
            BubbleSet s{};
            if (host_validate(&p, nullptr) != MCB_OK || !bubbles_of(reqs[i].c, &s) || reqs[i].c.nsteps < 1) {
                res[i].status = MCB_ERR_INVALID_ARG;
                waiting.pop_front();
                ++settled;
                continue;
            }
            if (need > budget) {
                res[i].status = MCB_ERR_CAPACITY;
                waiting.pop_front();
                ++settled;
                continue;
            }
            if ((cfg.max_cases > 0 && (int)active.size() >= cfg.max_cases) || in_use + need > budget) break;
            if (in_use + need + batch.pool_bytes() > budget) {
                mcb_status ts = batch.trim_pool(budget - in_use - need);
                if (ts != MCB_OK) return ts;
            }
            int id = 0;
            mcb_status as = batch.add_bubbles(p, s.b, s.n, s.width, &id);
            if (as != MCB_OK) {
                if (active.empty()) {
                    res[i].status = as;   // fails alone: give up on this request
                    waiting.pop_front();
                    ++settled;
                    continue;
                }
                break;   // retry once resident cases have left
            }
            res[i].admit_s = seconds_since(t0);
            active.push_back(Resident{i, id, reqs[i].c.nsteps, need});
            in_use += need;
            waiting.pop_front();
            st->peak_cases = std::max(st->peak_cases, (int)active.size());
            st->peak_bytes = std::max(st->peak_bytes, (double)in_use);
        }
        st->admit_s += seconds_since(ta);
        if (active.empty()) {
            const Clock::time_point ti = Clock::now();
            if (next < n) wait_until(t0, reqs[next].arrival_s);
            st->idle_s += seconds_since(ti);
            continue;
        }
        long long k = tick;
        for (const Resident& a : active) k = std::min(k, a.left);
        const Clock::time_point te = Clock::now();
        mcb_status ss = batch.step((int)k);
        const Clock::time_point tw = Clock::now();
        st->enqueue_s += std::chrono::duration<double>(tw - te).count();
        if (ss == MCB_OK) ss = batch.synchronize();
        st->wait_s += seconds_since(tw);
        if (ss != MCB_OK) return ss;
        st->rounds += 1;
        const Clock::time_point tr = Clock::now();
        // Read out every case that is done, then release them all: a removal
        // re-forms the batch (work lists, views, step graphs) before the next
        // macro pass, so removing between read-outs would re-form it once per
        // finished case instead of once per round.
        for (Resident& a : active) {
            a.left -= k;
            if (a.left > 0) continue;
            mcb_request_result& r = res[a.req];
            const mcb_request& q = reqs[a.req];
            mcb_status rs = batch.macros(a.id);
            if (rs == MCB_OK) rs = batch.reduce(a.id, &r.sum_phi, &r.kinetic);
            if (rs == MCB_OK && q.phi_out) {
                const CaseState* c = batch.get(a.id);
                rs = batch.read(a.id, MCB_FIELD_PHI, q.phi_out, (size_t)c->geo.n);
            }
            const CaseState* c = batch.get(a.id);
            r.lattice_updates = rs == MCB_OK ? c->geo.n * (long long)q.c.nsteps : 0;
            r.status = rs;
            r.finish_s = seconds_since(t0);
        }
        for (size_t j = 0; j < active.size();) {
            const Resident& a = active[j];
            if (a.left > 0) {
                ++j;
                continue;
            }
            mcb_status rm = batch.remove(a.id);
            if (rm != MCB_OK) return rm;
            in_use -= a.bytes;
            ++settled;
            active.erase(active.begin() + (long)j);
        }
        st->retire_s += seconds_since(tr);
    }
    finish_stats(reqs, n, res, st);
    return MCB_OK;
}

}  // namespace

extern "C" mcb_status mcb_serve_run(const mcb_serve_config* cfg, const mcb_request* reqs, int n,
                                    mcb_request_result* results, mcb_serve_stats* stats) {
    if (!cfg || n < 0 || (n > 0 && (!reqs || !results)) || !stats) return MCB_ERR_INVALID_ARG;
    if (cfg->storage < MCB_STORE_FP32 || cfg->storage > MCB_STORE_BF16) return MCB_ERR_INVALID_ARG;
    for (int i = 1; i < n; ++i)
        if (reqs[i].arrival_s < reqs[i - 1].arrival_s) return MCB_ERR_INVALID_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || cfg->device < 0 || cfg->device >= count) return MCB_ERR_NO_DEVICE;
    if (hipSetDevice(cfg->device) != hipSuccess) return MCB_ERR_NO_DEVICE;
    *stats = mcb_serve_stats{};
    for (int i = 0; i < n; ++i) {
        results[i] = mcb_request_result{};
        results[i].status = MCB_ERR_INVALID_ARG;
        results[i].arrival_s = reqs[i].arrival_s;
    }
    if (cfg->sequential) return serve_sequential(*cfg, reqs, n, results, stats);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MCB_ERR_HIP;
    mcb_status st = serve_batched(*cfg, reqs, n, results, stats, s);
    (void)hipStreamDestroy(s);
    return st;
}

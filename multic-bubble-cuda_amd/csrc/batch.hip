// Batched execution (batch.hpp) and the tile autotuner.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <tuple>

#include "batch.hpp"
#include "state.hpp"
#include "tile.hpp"

namespace mcb {

namespace {

long long plane_of(const mcb_params& p) {
    return plane_stride_cells((long long)p.nx * p.ny * p.nz);
}

unsigned pop_bytes(int storage) { return storage == MCB_STORE_FP32 ? 4u : 2u; }

constexpr int kMacroFields = 10;

// Batches of at least this many cells launch their steps one by one instead
// of replaying captured step graphs (see rebuild): a step of 8M cells takes
// ~0.2 ms, several times what enqueueing its launches takes.
constexpr long long kCellsEager = 8LL << 20;

// Work items of a case (x fastest, then y, then z chunks), renumbered by
// xcd_runs: each XCD takes runs of 8 consecutive tiles, so x-neighbouring
// tiles meet in one L2 while the device works on one contiguous stretch of
// the list (against xcd_order, where each XCD walks its own range: the 64
// FP32 'batch' presets 22.0k -> 22.4k MLUPS, FP16 at 256^3 41.5k -> 42.7k,
// 128^3 and 144^3 within 0.5%; without renumbering FP16 ran 2-7% slower).
std::vector<WorkItem> tiles_of(const CaseState& c, int slot) {
    std::vector<WorkItem> v;
    const Tiling& t = c.tiling;
    for (int z0 = 0; z0 < c.p.nz; z0 += t.zc)
        for (int y0 = 0; y0 < c.p.ny; y0 += t.ty)
            for (int x0 = 0; x0 < c.p.nx; x0 += t.tx) v.push_back(WorkItem{slot, x0, y0, z0});
    return v;
}

std::vector<WorkItem> xcd_permute(const std::vector<WorkItem>& in) {
    std::vector<WorkItem> out(in.size());
    const unsigned n = (unsigned)in.size();
    for (unsigned lid = 0; lid < n; ++lid) out[lid] = in[xcd_runs(lid, n, 8)];
    return out;
}

}  // namespace

size_t case_bytes(const mcb_params& p) {
    const size_t P = (size_t)plane_of(p);
    const bool moments = p.tau_g == 1.0f;
    size_t b = (moments ? 2 : 4) * 19 * P * pop_bytes(p.storage);          // f (and g) A/B
    b += (moments || p.storage != MCB_STORE_FP32 ? 2 : 1) * P * sizeof(float);   // phi
    if (moments) b += p.storage == MCB_STORE_FP32 ? 6 * P * sizeof(float) : 3 * P * 2;   // u, n (FP32) or B (FP16)
    return b;
}

CaseView CaseState::view(int q) const {
    CaseView v{};
    v.geo = geo;
    v.phys = phys;
    v.zc = tiling.zc;
    v.tx = tiling.tx;
    v.ty = tiling.ty;
    v.fsrc = f[q];
    v.fdst = f[q ^ 1];
    v.gsrc = g[q];
    v.gdst = g[q ^ 1];
    v.phi = phi[q];
    v.phi_prev = phi[q ^ 1];
    v.mslot = q;
    v.mass = mass;
    v.massq = massq;
    v.partial = partial;
    // Moment form: only the parity of the initial state (cur = 0 before the
    // first step) sees GM_INITIAL.
    v.gm = gm == GM_INITIAL && q != cur ? GM_MOMENTS : gm;
    for (int k = 0; k < 3; ++k) {
        v.gm_u[k] = gmu[k];
        v.gm_n[k] = gmn[k];
    }
    return v;
}

CaseState::~CaseState() {
    for (auto& a : allocs) (void)hipFree(a.first);
    if (d_work) (void)hipFree(d_work);
    if (partial) (void)hipFree(partial);
}

// ---------------------------------------------------------------- tilings

namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, Tiling> g_tuned;

bool supported_for(const Tiling& t, int storage) {
    return storage == MCB_STORE_FP32 ? tiling_supported(t) : tiling_supported16(t);
}
}  // namespace

Tiling choose_tiling(const mcb_params& p) {
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(std::make_tuple(p.nx, p.ny, p.nz, p.storage));
        if (it != g_tuned.end()) return it->second;
    }
    Tiling t = p.nx >= 64 ? Tiling{64, 8, 16} : (p.nx >= 32 ? Tiling{32, 8, 16} : Tiling{16, 16, 16});
    // 16-bit forms: half the bytes per cell, so longer z runs pay off
    // (measured at 256^3: 64x8x32 fastest, 64x4x32 4% and 32x8x32 17%
    // slower); a 32-wide tile where it pads x less, since padded lanes do no
    // work. Measured per x side on batches of equal cubes (48 ... 192): a tile
    // spanning the whole row where one is instantiated (48, 80, 96, 112, 144,
    // 176; 1.1-1.9x), 64x4 for the other multiples of 32 up to 192; both with
    // z-chunks of 16, which give a lattice of this size alone enough
    // workgroups to fill the GPU (serving at 20 requests/s: p99 0.13 s
    // against 0.28 s with chunks of 32, and the same throughput at saturation).
    if (p.storage != MCB_STORE_FP32 && p.nx >= 48) {
        const int pad64 = (p.nx + 63) / 64 * 64, pad32 = (p.nx + 31) / 32 * 32;
        t = pad32 < pad64 ? Tiling{32, 8, 32} : Tiling{64, 8, 32};
        const Tiling row{p.nx, 4, 16};
        if (tiling_supported16(row)) t = row;
        else if (p.nx <= 192 && p.nx % 32 == 0) t = Tiling{64, 4, 16};
    }
    t.zc = std::min(t.zc, p.nz);
    return t;
}

Tiling batch_tiling(const mcb_params& p) {
    // FP32 cases of a batch share one tile shape, so a batch step is one phase
    // and one collide launch for every case: on the 64 'batch' presets (sides
    // 32-128) 64x8 tiles with 8-plane chunks ran 21.8k MLUPS against 20.5k
    // with each side's autotuned single-case tile (four launch groups), whose
    // small groups leave CUs idle while they drain.
    if (p.storage == MCB_STORE_FP32) return Tiling{64, 8, std::min(8, p.nz)};
    return choose_tiling(p);
}

Tiling choose_tiling_single(const mcb_params& p) {
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(std::make_tuple(p.nx, p.ny, p.nz, p.storage));
        if (it != g_tuned.end()) return it->second;
    }
    // A lattice alone on the device: shorter z-chunks until the pass has two
    // workgroups per CU (256 CUs), down to 4 planes. The reference preset
    // (64x64x256) gets 8 tiles per z layer at 64x8, i.e. 128 workgroups with
    // 16-plane chunks, which leaves half the CUs idle.
    Tiling t = choose_tiling(p);
    auto tiles = [&](const Tiling& q) {
        return (long long)((p.nx + q.tx - 1) / q.tx) * ((p.ny + q.ty - 1) / q.ty) * ((p.nz + q.zc - 1) / q.zc);
    };
    while (t.zc > 4 && tiles(t) < 512) t.zc /= 2;
    // Where the chunks are already down to 4 planes, 4-row tiles give twice
    // the workgroups for the same halo restaging per plane (the reference
    // preset: 64x4x4 0.0591-0.0595 ms per step against 0.0610-0.0613 with
    // 64x8x4; 64^3: 0.0317 against 0.0363 ms).
    if (t.zc <= 4 && t.tx == 64 && t.ty == 8 && p.storage == MCB_STORE_FP32) t.ty = 4;
    return t;
}

// ------------------------------------------------------------------ Batch

Batch::Batch(int device, hipStream_t stream, bool use_graphs)
    : device_(device), stream_(stream), use_graphs_(use_graphs) {}

Batch::~Batch() {
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    for (auto& g : graph_)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto& kv : groups_)
        if (kv.second.d_work) (void)hipFree(kv.second.d_work);
    for (auto& v : d_views_)
        if (v) (void)hipFree(v);
    if (d_partial_) (void)hipFree(d_partial_);
    cases_.clear();
    for (auto& q : macro_bufs_) (void)hipFree(q);
    for (auto& kv : pool_) (void)hipFree(kv.second);
}

mcb_status Batch::fail(mcb_status s, const std::string& m) {
    err_ = m;
    return s;
}

#define B_HIP(call)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return fail(hip_status(e_), std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

CaseState* Batch::get(int id) {
    auto it = cases_.find(id);
    return it == cases_.end() ? nullptr : it->second.get();
}

size_t Batch::bytes() const {
    size_t b = 0;
    for (auto& kv : cases_) b += kv.second->bytes;
    return b + macro_bytes_;
}

mcb_status Batch::dev_alloc(size_t bytes, void** out) {
    auto it = pool_.find(bytes);
    if (it != pool_.end()) {
        *out = it->second;
        pool_.erase(it);
        pool_bytes_ -= bytes;
        return MCB_OK;
    }
    hipError_t e = hipMalloc(out, bytes);
    if (e == hipErrorOutOfMemory && !pool_.empty()) {
        (void)hipGetLastError();
        mcb_status st = trim_pool(0);
        if (st != MCB_OK) return st;
        e = hipMalloc(out, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(hip_status(e), "hipMalloc failed for a case");
    }
    return MCB_OK;
}

void Batch::dev_release(void* ptr, size_t bytes) {
    if (pool_bytes_ + bytes <= pool_limit_) {
        pool_.emplace(bytes, ptr);
        pool_bytes_ += bytes;
    } else {
        (void)hipFree(ptr);
    }
}

mcb_status Batch::trim_pool(size_t keep) {
    B_HIP(hipSetDevice(device_));
    while (pool_bytes_ > keep && !pool_.empty()) {
        auto it = std::prev(pool_.end());   // largest first
        pool_bytes_ -= it->first;
        void* q = it->second;
        pool_.erase(it);
        B_HIP(hipFree(q));
    }
    return MCB_OK;
}

mcb_status Batch::ensure_partial() {
    if (!d_partial_) B_HIP(hipMalloc(reinterpret_cast<void**>(&d_partial_), (2 * kReduceBlocks + 2) * sizeof(double)));
    return MCB_OK;
}

mcb_status Batch::add(const mcb_params& p, const float* phi_host, int* id) {
    if (!phi_host) return fail(MCB_ERR_INVALID_ARG, "null argument");
    return add_impl(p, phi_host, nullptr, id);
}

mcb_status Batch::add_bubbles(const mcb_params& p, const mcb_bubble* bubbles, int n, float width, int* id) {
    if (!bubbles || n < 1 || n > kMaxBubbleSet || !(width > 0.0f) || !std::isfinite(width))
        return fail(MCB_ERR_INVALID_ARG, "invalid bubble set");
    BubbleSet s{};
    for (int i = 0; i < n; ++i) {
        const mcb_bubble& b = bubbles[i];
        if (!std::isfinite(b.cx) || !std::isfinite(b.cy) || !std::isfinite(b.cz) || !std::isfinite(b.r) || b.r < 0.0f)
            return fail(MCB_ERR_INVALID_ARG, "invalid bubble");
        s.b[i] = b;
    }
    s.n = n;
    s.width = width;
    return add_impl(p, nullptr, &s, id);
}

mcb_status Batch::add_impl(const mcb_params& p, const float* phi_host, const BubbleSet* bubbles, int* id) {
    if (!id) return fail(MCB_ERR_INVALID_ARG, "null argument");
    std::string why;
    if (host_validate(&p, &why) != MCB_OK) return fail(MCB_ERR_INVALID_ARG, why);
    if (p.bc_mode != MCB_BC_PERIODIC) return fail(MCB_ERR_UNSUPPORTED, "batched cases are periodic");
    if (p.f_form != MCB_F_POPULATIONS) return fail(MCB_ERR_UNSUPPORTED, "batched cases keep f as populations");
    if (storage_ >= 0 && p.storage != storage_) return fail(MCB_ERR_UNSUPPORTED, "one storage form per batch");
    const long long P = plane_of(p);
    if (19LL * P * pop_bytes(p.storage) > 0xFFFFFFFFLL) return fail(MCB_ERR_UNSUPPORTED, "lattice too large");
    Tiling t = batch_tiling(p);
    if (!supported_for(t, p.storage)) return fail(MCB_ERR_UNSUPPORTED, "no tiling for this lattice");
    B_HIP(hipSetDevice(device_));

    // On any failure below, cs goes out of scope and frees what was allocated.
    auto cs = std::make_unique<CaseState>();
    CaseState* c = cs.get();
    c->p = p;
    c->storage = p.storage;
    c->mode = p.tau_g == 1.0f ? COLLIDE_STEP : COLLIDE_STEP_BGK;
    c->tiling = t;
    c->geo.nx = p.nx;
    c->geo.ny = p.ny;
    c->geo.nz = p.nz;
    c->geo.nxy = p.nx * p.ny;
    c->geo.n = (long long)p.nx * p.ny * p.nz;
    c->geo.plane = P;
    // FP32 populations in 64-cell blocks where a z plane holds whole blocks
    // (16-bit ones stay in planes: with blocks the FP16 batch ran 3% slower).
    c->geo.blocked = MCB_POP_BLOCKS && p.storage == MCB_STORE_FP32 && c->geo.nxy % 64 == 0;
    c->phys = host_phys(p);
    auto alloc = [&](size_t bytes, void** out) -> mcb_status {
        mcb_status s = dev_alloc(bytes, out);
        if (s != MCB_OK) return s;
        c->allocs.emplace_back(*out, bytes);
        c->bytes += bytes;
        return MCB_OK;
    };
    mcb_status st;
    const size_t pops = 19 * (size_t)P * pop_bytes(p.storage);
    // At tau_g = 1 g is kept in moment form instead of two sets of 19
    // planes, for every storage form.
    const bool moments = c->mode == COLLIDE_STEP;
    for (int b = 0; b < 2; ++b) {
        if ((st = alloc(pops, &c->f[b])) != MCB_OK) return st;
        if (!moments && (st = alloc(pops, &c->g[b])) != MCB_OK) return st;
    }
    void* q = nullptr;
    if ((st = alloc((size_t)P * sizeof(float), &q)) != MCB_OK) return st;
    c->phi[0] = c->phi[1] = static_cast<float*>(q);
    if (moments || p.storage != MCB_STORE_FP32) {
        // phi[1] holds the state's phi (moment form) or the g anchor (16 bits).
        if ((st = alloc((size_t)P * sizeof(float), &q)) != MCB_OK) return st;
        c->phi[1] = static_cast<float*>(q);
    }
    if (moments) {
        // FP32: u and n (collisionPhase's inputs, bit-exact; a batch's
        // 8-plane chunks are too short for the phase pass to recompute n
        // profitably); 16 bits: the vector B = 3 phi u + phi_n n in FP16
        // (phase_tile_ab).
        const bool fp32 = p.storage == MCB_STORE_FP32;
        for (int k = 0; k < (fp32 ? 6 : 3); ++k) {
            if ((st = alloc((size_t)P * (fp32 ? sizeof(float) : 2), &q)) != MCB_OK) return st;
            (k < 3 ? c->gmu[k] : c->gmn[k - 3]) = static_cast<float*>(q);
        }
        if (!fp32) {
            // Mass-drift correction of the 16-bit moment form (MassQ).
            if ((st = alloc(sizeof(MassQ), &q)) != MCB_OK) return st;
            c->massq = static_cast<MassQ*>(q);
            const MassQ zero{0, 0, 0, {0.0, 0.0}};
            B_HIP(hipMemcpyAsync(c->massq, &zero, sizeof(zero), hipMemcpyHostToDevice, stream_));
        }
        c->gm = GM_INITIAL;
    } else if (p.storage != MCB_STORE_FP32) {
        // 16-bit phase populations: per-step FP64 mass-drift correction.
        if ((st = alloc(sizeof(MassCell), &q)) != MCB_OK) return st;
        c->mass = static_cast<MassCell*>(q);
        double target = 0.0;
        if (phi_host)
            for (long long i = 0; i < c->geo.n; ++i) target += (double)phi_host[i];
        if ((st = set_mass_target(c, target)) != MCB_OK) return st;
    }
    // Initial state: the current populations are in buffer 0, the initial
    // phi (state or g anchor) in phi[1].
    if (phi_host) {
        B_HIP(hipMemcpyAsync(c->phi[1], phi_host, (size_t)c->geo.n * sizeof(float), hipMemcpyHostToDevice, stream_));
    } else {
        B_HIP(launch_bubble_phi(c->phi[1], c->geo, *bubbles, stream_));
        if (c->mass) {
            if ((st = ensure_partial()) != MCB_OK) return st;
            B_HIP(launch_sum(c->phi[1], c->geo.n, d_partial_, &c->mass->target, stream_));
        }
    }
    if (p.storage == MCB_STORE_FP32) {
        B_HIP(launch_init_pull(c->phi[1], static_cast<float*>(c->f[0]), static_cast<float*>(c->g[0]), c->geo, c->phys,
                               stream_));
    } else {
        B_HIP(launch_init16(c->phi[1], c->f[0], c->g[0], c->geo, c->phys, p.storage, stream_));
    }
    c->cur = 0;
    c->parity_offset = parity_;
    c->phi_last = c->phi[1];
    c->steps = 0;
    if ((st = build_own_table(c)) != MCB_OK) return st;
    B_HIP(hipStreamSynchronize(stream_));
    storage_ = p.storage;
    *id = next_id_++;
    cases_[*id] = std::move(cs);
    dirty_ = true;
    return MCB_OK;
}

// The case's own work table (case index 0, used with the case's entry of the
// view arrays as the base) and its partial-sum buffer, for its tiling.
mcb_status Batch::build_own_table(CaseState* c) {
    auto own = xcd_permute(tiles_of(*c, 0));
    c->n_work = (unsigned)own.size();
    if (own.size() > c->work_cap) {
        if (c->d_work) B_HIP(hipFree(c->d_work));
        c->d_work = nullptr;
        c->work_cap = 0;
        B_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_work), own.size() * sizeof(WorkItem)));
        c->work_cap = own.size();
    }
    B_HIP(hipMemcpyAsync(c->d_work, own.data(), own.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream_));
    if ((c->mass || c->massq) && c->n_work > c->partial_cap) {
        if (c->partial) B_HIP(hipFree(c->partial));
        c->partial = nullptr;
        c->partial_cap = 0;
        B_HIP(hipMalloc(reinterpret_cast<void**>(&c->partial), c->n_work * sizeof(double)));
        c->partial_cap = c->n_work;
    }
    B_HIP(hipStreamSynchronize(stream_));   // own is a host temporary
    return MCB_OK;
}

mcb_status Batch::remove(int id) {
    auto it = cases_.find(id);
    if (it == cases_.end()) return fail(MCB_ERR_INVALID_ARG, "no such case");
    B_HIP(hipSetDevice(device_));
    B_HIP(hipStreamSynchronize(stream_));
    CaseState* c = it->second.get();
    for (auto& a : c->allocs) dev_release(a.first, a.second);
    c->allocs.clear();
    cases_.erase(it);   // frees the case's tables
    if (cases_.empty()) storage_ = -1;
    if (macro_id_ == id) macro_id_ = 0;
    dirty_ = true;
    return MCB_OK;
}

mcb_status Batch::set_mass_target(CaseState* c, double target) {
    const MassCell m{target, {0.0, 0.0}};
    B_HIP(hipMemcpyAsync(c->mass, &m, sizeof(m), hipMemcpyHostToDevice, stream_));
    B_HIP(hipStreamSynchronize(stream_));
    return MCB_OK;
}

mcb_status Batch::ensure_macro_scratch() {
    size_t need = 0;
    for (auto& kv : cases_) need = std::max(need, (size_t)kv.second->geo.plane);
    if (need <= macro_plane_) return MCB_OK;
    for (void* q : macro_bufs_) B_HIP(hipFree(q));
    macro_bufs_.clear();
    macro_bytes_ = 0;
    macro_plane_ = 0;
    macro_id_ = 0;
    for (int k = 0; k < kMacroFields; ++k) {
        void* q = nullptr;
        B_HIP(hipMalloc(&q, need * sizeof(float)));
        macro_bufs_.push_back(q);
        macro_bytes_ += need * sizeof(float);
    }
    macro_plane_ = need;
    return MCB_OK;
}

mcb_status Batch::rebuild() {
    B_HIP(hipSetDevice(device_));
    B_HIP(hipStreamSynchronize(stream_));
    mcb_status st = ensure_macro_scratch();
    if (st != MCB_OK) return st;
    order_.clear();
    for (auto& kv : cases_) order_.push_back(kv.first);
    const size_t n = order_.size();
    if (n > views_cap_) {
        for (auto& v : d_views_) {
            if (v) B_HIP(hipFree(v));
            v = nullptr;
        }
        views_cap_ = 0;
        const size_t cap = std::max<size_t>(n, 2 * n);
        for (auto& v : d_views_) B_HIP(hipMalloc(reinterpret_cast<void**>(&v), cap * sizeof(CaseView)));
        views_cap_ = cap;
    }
    std::map<std::tuple<int, int, int>, std::vector<WorkItem>> by_group;
    any_initial_ = false;
    any_mass_ = false;
    for (int q = 0; q < 2; ++q) hv_[q].clear();
    for (size_t i = 0; i < n; ++i) {
        CaseState* c = cases_[order_[i]].get();
        for (int q = 0; q < 2; ++q) {
            CaseView v = c->view(q ^ c->parity_offset);
            float** m = reinterpret_cast<float**>(&v.rho);
            for (int k = 0; k < kMacroFields; ++k) m[k] = static_cast<float*>(macro_bufs_[k]);
            hv_[q].push_back(v);
        }
        if (c->gm == GM_INITIAL) any_initial_ = true;
        if (c->massq || c->mass) any_mass_ = true;
        auto tl = tiles_of(*c, (int)i);
        auto& vec = by_group[std::make_tuple(c->tiling.tx, c->tiling.ty, c->mode)];
        vec.insert(vec.end(), tl.begin(), tl.end());
        c->slot = (int)i;
    }
    for (int q = 0; q < 2; ++q)
        if (n) B_HIP(hipMemcpyAsync(d_views_[q], hv_[q].data(), n * sizeof(CaseView), hipMemcpyHostToDevice, stream_));
    for (auto& kv : groups_) kv.second.n = 0;
    std::vector<std::vector<WorkItem>> staged;
    for (auto& kv : by_group) {
        Group& gr = groups_[kv.first];
        gr.t = Tiling{std::get<0>(kv.first), std::get<1>(kv.first), 0};
        gr.mode = std::get<2>(kv.first);
        staged.push_back(xcd_permute(kv.second));
        const auto& items = staged.back();
        if (items.size() > gr.cap) {
            if (gr.d_work) B_HIP(hipFree(gr.d_work));
            gr.d_work = nullptr;
            gr.cap = 0;
            const size_t cap = 2 * items.size();
            B_HIP(hipMalloc(reinterpret_cast<void**>(&gr.d_work), cap * sizeof(WorkItem)));
            gr.cap = cap;
        }
        gr.n = (unsigned)items.size();
        B_HIP(hipMemcpyAsync(gr.d_work, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream_));
    }
    B_HIP(hipStreamSynchronize(stream_));   // hv and staged are host temporaries
    for (auto& g : graph_) {
        if (g) B_HIP(hipGraphExecDestroy(g));
        g = nullptr;
    }
    // Step graphs pay where a step is short next to its launches (a few small
    // cases); a batch of kCellsEager cells or more keeps the device busy with
    // launches enqueued one by one, and a serving loop that re-forms it every
    // round would only pay the capture.
    long long cells = 0;
    for (const auto& kv : cases_) cells += kv.second->geo.n;
    if (use_graphs_ && n && cells < kCellsEager) {
        for (int q = 0; q < 2; ++q) {
            hipGraph_t gr = nullptr;
            if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                (void)hipGetLastError();
                use_graphs_ = false;
                break;
            }
            mcb_status s1 = launch_step(q), s2 = launch_step(q ^ 1);
            hipError_t e = hipStreamEndCapture(stream_, &gr);
            if (s1 != MCB_OK || s2 != MCB_OK || e != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                (void)hipGetLastError();
                use_graphs_ = false;
                break;
            }
            e = hipGraphInstantiate(&graph_[q], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                graph_[q] = nullptr;
                use_graphs_ = false;
                break;
            }
        }
    }
    dirty_ = false;
    return MCB_OK;
}

mcb_status Batch::launch_step(int q) {
    for (auto& kv : groups_) {
        const Group& gr = kv.second;
        if (gr.n) B_HIP(launch_phase_batch(d_views_[q], gr.d_work, gr.n, gr.t, storage_, stream_));
    }
    if (any_mass_) B_HIP(launch_mass_finish(d_views_[q], (unsigned)order_.size(), stream_));
    for (auto& kv : groups_) {
        const Group& gr = kv.second;
        if (gr.n) B_HIP(launch_collide_batch(d_views_[q], gr.d_work, gr.n, gr.t, gr.mode, storage_, stream_));
    }
    return MCB_OK;
}

mcb_status Batch::step(int nsteps) {
    if (nsteps < 0) return fail(MCB_ERR_INVALID_ARG, "negative step count");
    if (cases_.empty()) return MCB_OK;
    B_HIP(hipSetDevice(device_));
    if (dirty_) {
        mcb_status st = rebuild();
        if (st != MCB_OK) return st;
    }
    if (nsteps == 0) return MCB_OK;
    int left = nsteps;
    if (any_initial_) {
        // The first step of moment-form cases reads their initial state; the
        // view of that parity then switches to the moment form (stream-ordered
        // copy from the persistent host array, so graphs see it).
        mcb_status st = launch_step(parity_);
        if (st != MCB_OK) return st;
        const int q = parity_;
        parity_ ^= 1;
        left -= 1;
        for (size_t i = 0; i < order_.size(); ++i) {
            CaseState* c = cases_[order_[i]].get();
            if (c->gm != GM_INITIAL) continue;
            c->gm = GM_MOMENTS;
            hv_[q][i].gm = GM_MOMENTS;
            B_HIP(hipMemcpyAsync(d_views_[q] + i, &hv_[q][i], sizeof(CaseView), hipMemcpyHostToDevice, stream_));
        }
        any_initial_ = false;
    }
    while (left > 0) {
        if (use_graphs_ && graph_[parity_] && left >= 2) {
            B_HIP(hipGraphLaunch(graph_[parity_], stream_));
            left -= 2;
        } else {
            mcb_status st = launch_step(parity_);
            if (st != MCB_OK) return st;
            parity_ ^= 1;
            left -= 1;
        }
    }
    batch_steps_ += nsteps;
    macro_id_ = 0;  // the macro fields no longer match any case's state
    for (auto& kv : cases_) {
        CaseState* c = kv.second.get();
        c->cur = parity_ ^ c->parity_offset;
        c->steps += nsteps;
        c->phi_last = c->phi[c->cur ^ 1];
    }
    return MCB_OK;
}

mcb_status Batch::macros(int id) {
    CaseState* c = get(id);
    if (!c) return fail(MCB_ERR_INVALID_ARG, "no such case");
    B_HIP(hipSetDevice(device_));
    if (dirty_) {
        mcb_status st = rebuild();
        if (st != MCB_OK) return st;
    }
    // The case's own table indexes case 0 of a view array based at its slot.
    const CaseView* views = d_views_[parity_] + c->slot;
    B_HIP(launch_phase_batch(views, c->d_work, c->n_work, c->tiling, storage_, stream_));
    B_HIP(launch_collide_batch(views, c->d_work, c->n_work, c->tiling, COLLIDE_MACROS, storage_, stream_));
    c->phi_last = c->phi[c->cur];
    macro_id_ = id;
    return MCB_OK;
}

mcb_status Batch::read(int id, int field, float* host, size_t n) {
    CaseState* c = get(id);
    if (!c || !host) return fail(MCB_ERR_INVALID_ARG, "no such case or null buffer");
    B_HIP(hipSetDevice(device_));
    const size_t cells = (size_t)c->geo.n;
    const size_t P = (size_t)c->geo.plane;
    if (field == MCB_FIELD_F || field == MCB_FIELD_G) {
        if (n != 19 * cells) return fail(MCB_ERR_INVALID_ARG, "population buffer must hold 19*nx*ny*nz floats");
        const bool isf = field == MCB_FIELD_F;
        float* tmp = nullptr;
        B_HIP(hipMalloc(reinterpret_cast<void**>(&tmp), 19 * P * sizeof(float)));
        hipError_t e;
        if (!isf && c->gm != GM_POPULATIONS)
            e = c->storage == MCB_STORE_FP32 ? launch_gm_to_ref(c->view(c->cur), tmp, stream_)
                                             : launch_ab_to_ref(c->view(c->cur), tmp, stream_);
        else if (c->storage == MCB_STORE_FP32)
            e = launch_link_shift(static_cast<const float*>(isf ? c->f[c->cur] : c->g[c->cur]), c->geo.blocked, tmp, 0,
                                  c->geo, +1, stream_);
        else
            e = launch_decode16(isf ? c->f[c->cur] : c->g[c->cur], isf ? nullptr : c->phi[c->cur ^ 1], tmp, c->geo,
                                c->phys, c->storage, stream_);
        for (int l = 0; l < 19 && e == hipSuccess; ++l)
            e = hipMemcpyAsync(host + l * cells, tmp + l * P, cells * sizeof(float), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        (void)hipFree(tmp);
        B_HIP(e);
        return MCB_OK;
    }
    if (n != cells) return fail(MCB_ERR_INVALID_ARG, "field buffer must hold nx*ny*nz floats");
    const float* src = nullptr;
    if (field == MCB_FIELD_PHI) {
        src = c->phi_last;
    } else if (field >= MCB_FIELD_RHO && field <= MCB_FIELD_NORMZ) {
        if (macro_id_ != id) return fail(MCB_ERR_INVALID_ARG, "compute the macros of this case first");
        src = static_cast<const float*>(macro_bufs_[field - MCB_FIELD_RHO]);
    } else {
        return fail(MCB_ERR_INVALID_ARG, "unknown field");
    }
    B_HIP(hipMemcpyAsync(host, src, cells * sizeof(float), hipMemcpyDeviceToHost, stream_));
    B_HIP(hipStreamSynchronize(stream_));
    return MCB_OK;
}

mcb_status Batch::to_populations(CaseState* c) {
    B_HIP(hipStreamSynchronize(stream_));
    const size_t pops = 19 * (size_t)c->geo.plane * pop_bytes(c->storage);
    for (int b = 0; b < 2; ++b) {
        mcb_status st = dev_alloc(pops, &c->g[b]);
        if (st != MCB_OK) return st;
        c->allocs.emplace_back(c->g[b], pops);
        c->bytes += pops;
    }
    if (c->storage != MCB_STORE_FP32) {
        void* q = nullptr;
        mcb_status st = dev_alloc(sizeof(MassCell), &q);
        if (st != MCB_OK) return st;
        c->allocs.emplace_back(q, sizeof(MassCell));
        c->mass = static_cast<MassCell*>(q);
        if ((st = build_own_table(c)) != MCB_OK) return st;   // partial sums
    }
    c->gm = GM_POPULATIONS;
    dirty_ = true;
    return MCB_OK;
}

// --------------------------------------------------------------- state

// What the next step of a case reads (view(cur)): f of the current buffer;
// g as populations, or its moment form (u in FP32; B in 16 bits); the
// state's phi (the moment form's phi, the 16-bit g anchor; with FP32 g
// populations the phi the last step reported, which the populations do not
// give back); the 16-bit mass target and the correction pending for the next
// phase pass. n is not part of an image: it is gradCalc of the state's phi,
// rebuilt on loading. FP32 images have the parts of a single context's (same
// form), so either resumes in the other.
std::vector<StatePart> Batch::state_parts(CaseState* c, double* mass) {
    const size_t N = (size_t)c->geo.n, P = (size_t)c->geo.plane;
    const unsigned eb = pop_bytes(c->storage);
    // Images keep populations in planes: blocked ones pass through the idle
    // A/B buffers (stage_pops), which the next step overwrites.
    const int side = c->geo.blocked ? c->cur ^ 1 : c->cur;
    std::vector<StatePart> parts;
    parts.push_back({"f", c->f[side], 19, N, P, eb, false});
    if (c->gm == GM_POPULATIONS) parts.push_back({"g", c->g[side], 19, N, P, eb, false});
    parts.push_back({"phi", c->phi[c->cur ^ 1], 1, N, N, 4, false});
    if (c->gm != GM_POPULATIONS) {
        static const char* const un[3] = {"u0", "u1", "u2"};
        for (int k = 0; k < 3; ++k)
            if (c->gmu[k]) parts.push_back({un[k], c->gmu[k], 1, N, N, c->storage == MCB_STORE_FP32 ? 4u : 2u, false});
    }
    if (c->mass) parts.push_back({"mass", mass, 1, 2, 2, 8, true});
    if (c->massq) parts.push_back({"massq", mass + 2, 1, 3, 3, 8, true});
    return parts;
}

mcb_status Batch::state_size(int id, size_t* bytes) {
    CaseState* c = get(id);
    if (!c || !bytes) return fail(MCB_ERR_INVALID_ARG, "no such case or null size pointer");
    double m5[5];
    *bytes = state_image_bytes(state_parts(c, m5));
    return MCB_OK;
}

mcb_status Batch::save_state(int id, void* buf, size_t bytes) {
    CaseState* c = get(id);
    if (!c) return fail(MCB_ERR_INVALID_ARG, "no such case");
    B_HIP(hipSetDevice(device_));
    B_HIP(hipStreamSynchronize(stream_));
    double m5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (c->mass) {
        MassCell m;
        B_HIP(hipMemcpy(&m, c->mass, sizeof(m), hipMemcpyDeviceToHost));
        m5[0] = m.target;
        m5[1] = m.delta[c->cur];
    }
    if (c->massq) {
        MassQ m;
        B_HIP(hipMemcpy(&m, c->massq, sizeof(m), hipMemcpyDeviceToHost));
        std::memcpy(&m5[2], &m.target, sizeof(double));   // the integer's bits
        m5[3] = (double)m.set;
        m5[4] = m.delta[c->cur];
    }
    StateHeader h{};
    h.params = c->p;
    h.path = MCB_PATH_FUSED;
    h.gm = c->gm;
    h.steps = c->steps;
    if (c->geo.blocked) B_HIP(stage_pops(c, true));
    return state_image_save(stream_, h, state_parts(c, m5), buf, bytes, &err_);
}

// Blocked FP32 populations around a state image: the current f (and g
// populations) into the idle buffers in planes, or back.
hipError_t Batch::stage_pops(CaseState* c, bool to_image) {
    const int cur = c->cur, idle = c->cur ^ 1;
    void* const* bufs[2] = {c->f, c->gm == GM_POPULATIONS ? c->g : nullptr};
    for (void* const* b : bufs) {
        if (!b) continue;
        const hipError_t e = to_image ? launch_pop_relayout(b[cur], 1, b[idle], 0, c->geo, 4, stream_)
                                      : launch_pop_relayout(b[idle], 0, b[cur], 1, c->geo, 4, stream_);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

mcb_status Batch::load_state(int id, const void* buf, size_t bytes) {
    CaseState* c = get(id);
    if (!c) return fail(MCB_ERR_INVALID_ARG, "no such case");
    StateHeader h;
    mcb_status st = state_image_header(buf, bytes, &h, &err_);
    if (st != MCB_OK) return st;
    if (!same_state_params(h.params, c->p) || h.path != MCB_PATH_FUSED)
        return fail(MCB_ERR_INVALID_ARG, "state image of another lattice or parameter set");
    if (h.gm < GM_POPULATIONS || h.gm > GM_INITIAL || (h.gm != GM_POPULATIONS && c->mode != COLLIDE_STEP))
        return fail(MCB_ERR_INVALID_ARG, "state image with an invalid g form");
    B_HIP(hipSetDevice(device_));
    B_HIP(hipStreamSynchronize(stream_));
    if (h.gm == GM_POPULATIONS && c->gm != GM_POPULATIONS && (st = to_populations(c)) != MCB_OK) return st;
    // A case that holds populations and gets a moment-form image keeps its
    // g buffers (unused) and reads the moment fields again.
    c->gm = h.gm;
    double m5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if ((st = state_image_load(stream_, buf, bytes, state_parts(c, m5), &err_)) != MCB_OK) return st;
    if (c->geo.blocked) {
        B_HIP(stage_pops(c, false));
        B_HIP(hipStreamSynchronize(stream_));
    }
    if (c->mass) {
        MassCell m{m5[0], {0.0, 0.0}};
        m.delta[c->cur] = m5[1];
        B_HIP(hipMemcpy(c->mass, &m, sizeof(m), hipMemcpyHostToDevice));
    }
    if (c->massq) {
        MassQ m{0, m5[3] != 0.0 ? 1 : 0, 0, {0.0, 0.0}};
        std::memcpy(&m.target, &m5[2], sizeof(double));
        m.delta[c->cur] = m5[4];
        B_HIP(hipMemcpy(c->massq, &m, sizeof(m), hipMemcpyHostToDevice));
    }
    if (c->gm != GM_POPULATIONS && c->gmn[0]) {
        B_HIP(launch_normals(c->view(c->cur), stream_));
        B_HIP(hipStreamSynchronize(stream_));
    }
    c->steps = h.steps;
    c->phi_last = c->phi[c->cur ^ 1];
    if (macro_id_ == id) macro_id_ = 0;
    dirty_ = true;
    return MCB_OK;
}

mcb_status Batch::add_state(const void* buf, size_t bytes, int* id) {
    StateHeader h;
    mcb_status st = state_image_header(buf, bytes, &h, &err_);
    if (st != MCB_OK) return st;
    if (h.path != MCB_PATH_FUSED) return fail(MCB_ERR_INVALID_ARG, "state image of a staged-path context");
    mcb_params p = h.params;
    p.device = device_;
    std::string why;
    if (host_validate(&p, &why) != MCB_OK) return fail(MCB_ERR_INVALID_ARG, "state image: " + why);
    std::vector<float> zero;
    try {
        zero.assign((size_t)p.nx * p.ny * p.nz, 0.0f);
    } catch (const std::bad_alloc&) {
        return fail(MCB_ERR_OUT_OF_MEMORY, "host memory for the case's placeholder phi");
    }
    int nid = 0;
    if ((st = add(p, zero.data(), &nid)) != MCB_OK) return st;
    if ((st = load_state(nid, buf, bytes)) != MCB_OK) {
        const std::string keep = err_;
        (void)remove(nid);
        err_ = keep;
        return st;
    }
    if (id) *id = nid;
    return MCB_OK;
}

mcb_status Batch::write(int id, int field, const float* host, size_t n) {
    CaseState* c = get(id);
    if (!c || !host) return fail(MCB_ERR_INVALID_ARG, "no such case or null buffer");
    B_HIP(hipSetDevice(device_));
    const size_t cells = (size_t)c->geo.n;
    const size_t P = (size_t)c->geo.plane;
    if (field == MCB_FIELD_F || field == MCB_FIELD_G) {
        if (n != 19 * cells) return fail(MCB_ERR_INVALID_ARG, "population buffer must hold 19*nx*ny*nz floats");
        const bool isf = field == MCB_FIELD_F;
        if (!isf && c->gm != GM_POPULATIONS) {
            // Arbitrary g populations have no moment form: the case keeps g as
            // populations from here on (16 bits: anchored to the state's phi,
            // with the mass-drift correction).
            mcb_status st = to_populations(c);
            if (st != MCB_OK) return st;
        }
        float* tmp = nullptr;
        B_HIP(hipMalloc(reinterpret_cast<void**>(&tmp), 19 * P * sizeof(float)));
        hipError_t e = hipSuccess;
        for (int l = 0; l < 19 && e == hipSuccess; ++l)
            e = hipMemcpyAsync(tmp + l * P, host + l * cells, cells * sizeof(float), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) {
            if (c->storage == MCB_STORE_FP32)
                e = launch_link_shift(tmp, 0, static_cast<float*>(isf ? c->f[c->cur] : c->g[c->cur]), c->geo.blocked,
                                      c->geo, -1, stream_);
            else
                e = launch_encode16(tmp, isf ? nullptr : c->phi[c->cur ^ 1], isf ? c->f[c->cur] : c->g[c->cur], c->geo,
                                    c->phys, c->storage, stream_);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        (void)hipFree(tmp);
        B_HIP(e);
        if (c->mass && !isf) {
            // The written phase populations define the mass to keep.
            double target = 0.0;
            for (size_t i = 0; i < 19 * cells; ++i) target += (double)host[i];
            return set_mass_target(c, target);
        }
        return MCB_OK;
    }
    if (field != MCB_FIELD_PHI) return fail(MCB_ERR_INVALID_ARG, "only PHI, F and G can be written");
    if (c->storage != MCB_STORE_FP32 && c->gm == GM_POPULATIONS)
        return fail(MCB_ERR_UNSUPPORTED, "with 16-bit populations phi anchors the g populations; write G instead");
    if (n != cells) return fail(MCB_ERR_INVALID_ARG, "field buffer must hold nx*ny*nz floats");
    // In moment form phi of the state is phi[cur ^ 1]; like a write with
    // populations, the written phi is what the next phase pass replaces.
    if (c->gm != GM_POPULATIONS) c->phi_last = c->phi[c->cur];
    B_HIP(hipMemcpyAsync(c->phi_last, host, cells * sizeof(float), hipMemcpyHostToDevice, stream_));
    B_HIP(hipStreamSynchronize(stream_));
    return MCB_OK;
}

mcb_status Batch::reduce(int id, double* sum_phi, double* kinetic) {
    CaseState* c = get(id);
    if (!c) return fail(MCB_ERR_INVALID_ARG, "no such case");
    if (macro_id_ != id) return fail(MCB_ERR_INVALID_ARG, "compute the macros of this case first");
    B_HIP(hipSetDevice(device_));
    mcb_status st = ensure_partial();
    if (st != MCB_OK) return st;
    float* const* m = reinterpret_cast<float* const*>(macro_bufs_.data());
    B_HIP(launch_reduce(c->phi_last, m[0], m[1], m[2], m[3], c->geo.n, d_partial_, d_partial_ + 2 * kReduceBlocks,
                        stream_));
    double h[2] = {0.0, 0.0};
    B_HIP(hipMemcpyAsync(h, d_partial_ + 2 * kReduceBlocks, sizeof(h), hipMemcpyDeviceToHost, stream_));
    B_HIP(hipStreamSynchronize(stream_));
    if (sum_phi) *sum_phi = h[0];
    if (kinetic) *kinetic = h[1];
    if (!std::isfinite(h[0]) || !std::isfinite(h[1])) return fail(MCB_ERR_NONFINITE, "non-finite diagnostic");
    return MCB_OK;
}

mcb_status Batch::set_tiling(int id, const Tiling& t) {
    CaseState* c = get(id);
    if (!c) return fail(MCB_ERR_INVALID_ARG, "no such case");
    if (!supported_for(t, c->storage) || t.zc > c->p.nz) return fail(MCB_ERR_INVALID_ARG, "unsupported tiling");
    B_HIP(hipSetDevice(device_));
    B_HIP(hipStreamSynchronize(stream_));
    c->tiling = t;
    mcb_status st = build_own_table(c);
    if (st != MCB_OK) return st;
    dirty_ = true;
    return MCB_OK;
}

mcb_status Batch::synchronize() {
    B_HIP(hipSetDevice(device_));
    B_HIP(hipStreamSynchronize(stream_));
    return MCB_OK;
}

// --------------------------------------------------------------- autotune

mcb_status autotune(const mcb_params& p0, int device, Tiling* best, double* ms_per_step) {
    mcb_params p = p0;
    p.bc_mode = MCB_BC_PERIODIC;
    if (hipSetDevice(device) != hipSuccess) return MCB_ERR_NO_DEVICE;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MCB_ERR_HIP;
    mcb_status result = MCB_OK;
    double best_ms = 1e30;
    Tiling best_t = choose_tiling(p);
    {
        Batch b(device, s, false);
        // A centred bubble of radius nx/4 as the scratch state.
        mcb_bubble bub{p.nx * 0.5f, p.ny * 0.5f, p.nz * 0.5f, p.nx * 0.25f};
        int id = 0;
        mcb_status st = b.add_bubbles(p, &bub, 1, 3.0f, &id);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (st == MCB_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = MCB_ERR_HIP;
        const long long cells = (long long)p.nx * p.ny * p.nz;
        const int steps = (int)std::max<long long>(4, std::min<long long>(40, 400000000LL / std::max<long long>(cells, 1)));
        for (int i = 0; st == MCB_OK && i < tiling_count(); ++i) {
            for (int zc : {8, 16, 32}) {
                Tiling t = tiling_at(i);
                t.zc = std::min(zc, p.nz);
                if (!supported_for(t, p.storage) || t.tx > 2 * p.nx || t.ty > 2 * p.ny) continue;
                if ((st = b.set_tiling(id, t)) != MCB_OK) break;
                if ((st = b.step(2)) != MCB_OK) break;
                float ms = 0.0f;
                if (hipEventRecord(e0, s) != hipSuccess) { st = MCB_ERR_HIP; break; }
                if ((st = b.step(steps)) != MCB_OK) break;
                if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    st = MCB_ERR_HIP;
                    break;
                }
                if (ms / steps < best_ms) {
                    best_ms = ms / steps;
                    best_t = t;
                }
            }
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        result = st;
    }
    (void)hipStreamDestroy(s);
    if (result != MCB_OK) return result;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[std::make_tuple(p0.nx, p0.ny, p0.nz, p0.storage)] = best_t;
    }
    if (best) *best = best_t;
    if (ms_per_step) *ms_per_step = best_ms;
    return MCB_OK;
}

}  // namespace mcb

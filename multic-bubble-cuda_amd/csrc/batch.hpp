// Batched execution of many independent cases on one device.
//
// Every case keeps its own lattice, parameters and buffers resident in HBM.
// A batch step runs the phase pass of all cases, then the collide pass of all
// cases, each as one launch per tile-shape group: workgroups read a work
// item (case, tile origin) from a table built in XCD order, then a CaseView
// from a device array. Two batch steps (both A/B parities) are captured in a
// hipGraph and replayed; the graph is rebuilt when cases join or leave.
// Cases never exchange data, so each case's numbers do not depend on what
// else is in the batch.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "fused.hpp"
#include "state.hpp"

namespace mcb {

struct CaseState {
    mcb_params p{};
    Geo geo{};
    Phys phys{};
    int storage = MCB_STORE_FP32;
    int mode = COLLIDE_STEP;       // COLLIDE_STEP or COLLIDE_STEP_BGK
    Tiling tiling{64, 8, 16};
    int cur = 0;                    // population buffer of the current state
    int parity_offset = 0;          // cur = batch parity ^ parity_offset
    long long steps = 0;            // steps taken since init
    void* f[2] = {nullptr, nullptr};
    void* g[2] = {nullptr, nullptr};
    float* phi[2] = {nullptr, nullptr};
    float* phi_last = nullptr;      // written by the most recent phase pass
    int gm = GM_POPULATIONS;        // FP32 at tau_g = 1: g in moment form (GMode)
    float* gmu[3] = {nullptr, nullptr, nullptr};
    float* gmn[3] = {nullptr, nullptr, nullptr};   // FP32 moment form: n, stored by the collide pass
    MassCell* mass = nullptr;       // 16-bit storage: mass-drift correction state
    MassQ* massq = nullptr;         // 16-bit storage in moment form: its correction state
    double* partial = nullptr;      // one FP64 partial per phase work item
    size_t partial_cap = 0;
    std::vector<std::pair<void*, size_t>> allocs;   // device blocks and their sizes
    size_t bytes = 0;
    int slot = 0;                   // index in the batch's view arrays
    WorkItem* d_work = nullptr;     // this case's tiles with case index 0 (its macro pass)
    size_t work_cap = 0;
    unsigned n_work = 0;

    // Buffers seen by the kernels at A/B parity q (q = cur before a step).
    CaseView view(int q) const;

    CaseState() = default;
    CaseState(const CaseState&) = delete;
    CaseState& operator=(const CaseState&) = delete;
    // Frees the device memory still listed in allocs (the owning batch makes
    // its device current and drains its stream first).
    ~CaseState();
};

// Host-side constants of a parameter set (host.cpp).
Phys host_phys(const mcb_params& p);
mcb_status host_validate(const mcb_params* p, std::string* why);

// Device bytes of a case with these parameters (populations, phase field(s)
// and, at tau_g = 1, the moment form of g).
size_t case_bytes(const mcb_params& p);

class Batch {
public:
    Batch(int device, hipStream_t stream, bool use_graphs);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // Allocate a case and set its initial state from phi (host, nx*ny*nz
    // floats), or from bubbles on the device (serving; bubble_phi_cell, the
    // bits of the host initialisation).
    mcb_status add(const mcb_params& p, const float* phi_host, int* id);
    mcb_status add_bubbles(const mcb_params& p, const mcb_bubble* bubbles, int n, float width, int* id);
    mcb_status remove(int id);
    mcb_status step(int nsteps);
    mcb_status macros(int id);
    mcb_status read(int id, int field, float* host, size_t n);
    mcb_status write(int id, int field, const float* host, size_t n);
    mcb_status reduce(int id, double* sum_phi, double* kinetic);
    mcb_status set_tiling(int id, const Tiling& t);
    // State image of a case at a step boundary (state.hpp): size, save, load
    // into a case of the same parameters, or a new case from an image.
    mcb_status state_size(int id, size_t* bytes);
    mcb_status save_state(int id, void* buf, size_t bytes);
    mcb_status load_state(int id, const void* buf, size_t bytes);
    mcb_status add_state(const void* buf, size_t bytes, int* id);
    mcb_status synchronize();
    // Keep up to limit bytes of freed case buffers for reuse by later cases
    // of the same sizes (0: free them at once). trim_pool frees cached blocks
    // until at most keep bytes remain cached.
    void set_pool_limit(size_t limit) { pool_limit_ = limit; }
    mcb_status trim_pool(size_t keep);
    size_t pool_bytes() const { return pool_bytes_; }
    CaseState* get(int id);
    int count() const { return (int)cases_.size(); }
    size_t bytes() const;
    const std::string& error() const { return err_; }
    long long batch_steps() const { return batch_steps_; }
    hipStream_t stream() const { return stream_; }

private:
    mcb_status fail(mcb_status s, const std::string& m);
    mcb_status add_impl(const mcb_params& p, const float* phi_host, const BubbleSet* bubbles, int* id);
    mcb_status dev_alloc(size_t bytes, void** out);
    void dev_release(void* ptr, size_t bytes);
    mcb_status build_own_table(CaseState* c);
    mcb_status rebuild();
    mcb_status launch_step(int q);
    // 16-bit storage: reset the mass target and the pending corrections.
    mcb_status set_mass_target(CaseState* c, double target);
    // Keep g as populations from here on (arbitrary g has no moment form):
    // allocates the g buffers and, for 16 bits, the mass-drift state.
    mcb_status to_populations(CaseState* c);
    // The arrays of a case's state image; mass receives / supplies (5 slots)
    // the 16-bit mass target and pending correction of stored g (0, 1) and
    // of the moment form (2: fixed-point target, 3: target recorded, 4:
    // pending correction).
    std::vector<StatePart> state_parts(CaseState* c, double* mass);
    hipError_t stage_pops(CaseState* c, bool to_image);
    // One set of macro fields per batch, sized for its largest case; it holds
    // the observables of the case of the most recent macros() call.
    mcb_status ensure_macro_scratch();
    mcb_status ensure_partial();

    int device_;
    hipStream_t stream_;
    bool use_graphs_;
    std::string err_;
    std::map<int, std::unique_ptr<CaseState>> cases_;
    std::vector<int> order_;          // case ids in view-array order
    int next_id_ = 1;
    int parity_ = 0;
    long long batch_steps_ = 0;
    bool dirty_ = true;
    int storage_ = -1;                // all cases of a batch share one storage form
    // device arrays
    CaseView* d_views_[2] = {nullptr, nullptr};
    size_t views_cap_ = 0;
    struct Group {
        Tiling t{};
        int mode = COLLIDE_STEP;
        WorkItem* d_work = nullptr;
        size_t cap = 0;
        unsigned n = 0;
    };
    // Launch groups by (tile x, tile y, collide mode), kept across rebuilds.
    std::map<std::tuple<int, int, int>, Group> groups_;
    std::vector<CaseView> hv_[2];     // host copies of the view arrays
    bool any_initial_ = false;        // a moment-form case has not stepped yet
    bool any_mass_ = false;          // a case keeps a mass-drift correction (MassCell or MassQ)
    hipGraphExec_t graph_[2] = {nullptr, nullptr};
    double* d_partial_ = nullptr;     // reduction scratch (2 * kReduceBlocks + 2 doubles)
    std::vector<void*> macro_bufs_;
    size_t macro_plane_ = 0;
    size_t macro_bytes_ = 0;
    int macro_id_ = 0;
    std::multimap<size_t, void*> pool_;
    size_t pool_bytes_ = 0;
    size_t pool_limit_ = 0;
};

// Tile shape for a lattice: the autotuned choice for (nx, ny, nz, storage)
// when one exists, else a size rule.
Tiling choose_tiling(const mcb_params& p);
// The tiling of a single-case context: as choose_tiling, with z-chunks short
// enough that a small lattice alone still fills the device.
Tiling choose_tiling_single(const mcb_params& p);
// The tiling of a batch case: one shape for all FP32 cases (one launch per
// pass), the 16-bit rule of choose_tiling otherwise.
Tiling batch_tiling(const mcb_params& p);
// Measure every supported tile shape and z-chunk on a scratch case and
// remember the fastest for (nx, ny, nz, storage).
mcb_status autotune(const mcb_params& p, int device, Tiling* best, double* ms_per_step);

}  // namespace mcb

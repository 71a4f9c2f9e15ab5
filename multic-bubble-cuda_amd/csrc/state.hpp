// Byte image of a lattice state at a step boundary (mcb_save_state,
// mcb_load_state, mcb_batch_*_state): a header, then named parts, each the
// rows of one device (or host) array copied densely. A context or batch case
// lists the arrays its next step reads; loading the image into a context or
// case with the same parameters restores them, so the run continues bit for
// bit. The reference has no counterpart: it writes phi snapshots only
// (src/main.cu:184-189), which cannot restart a run.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace mcb {

constexpr uint32_t kStateVersion = 3;

struct StateHeader {
    char magic[8];          // "MCBSTATE"
    uint32_t version;       // kStateVersion
    uint32_t nparts;
    uint64_t bytes;         // whole image
    mcb_params params;      // of the context or case that saved it
    int32_t path;           // resolved path: MCB_PATH_FUSED or MCB_PATH_STAGED
    int32_t gm;             // representation of g (GMode)
    int64_t steps;          // steps since the initial state
};

struct StatePartHeader {
    char name[8];
    uint32_t rows;
    uint32_t elem_bytes;
    uint64_t row_elems;
};

// One array of the state: rows x row_elems elements of elem_bytes, rows
// row_stride elements apart (populations: 19 planes of the padded plane
// stride). host: ptr is host memory.
struct StatePart {
    const char* name;
    void* ptr;
    unsigned rows;
    size_t row_elems;
    size_t row_stride;
    unsigned elem_bytes;
    bool host;
};

size_t state_image_bytes(const std::vector<StatePart>& parts);

// Writes the image (header fields from info, parts in order) into buf.
mcb_status state_image_save(hipStream_t s, const StateHeader& info, const std::vector<StatePart>& parts, void* buf,
                            size_t bytes, std::string* err);

// Checks the header of an image (magic, version, size) and returns it.
mcb_status state_image_header(const void* buf, size_t bytes, StateHeader* h, std::string* err);

// Copies the image's parts into parts, which must match the image's part
// list (names, shapes, element sizes) in order.
mcb_status state_image_load(hipStream_t s, const void* buf, size_t bytes, const std::vector<StatePart>& parts,
                            std::string* err);

// The parameters that determine a state's meaning: everything but the
// device, the requested path and the normal mode (n is not in an image).
bool same_state_params(const mcb_params& a, const mcb_params& b);

}  // namespace mcb

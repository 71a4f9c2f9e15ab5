// State images (state.hpp): format, save and load.
#include "state.hpp"

#include <cstring>

namespace mcb {

namespace {

constexpr char kMagic[8] = {'M', 'C', 'B', 'S', 'T', 'A', 'T', 'E'};

size_t part_data_bytes(const StatePart& p) { return (size_t)p.rows * p.row_elems * p.elem_bytes; }

mcb_status bad(std::string* err, const std::string& m) {
    if (err) *err = m;
    return MCB_ERR_INVALID_ARG;
}

mcb_status hip_bad(std::string* err, hipError_t e, const char* what) {
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    return hip_status(e);
}

}  // namespace

size_t state_image_bytes(const std::vector<StatePart>& parts) {
    size_t n = sizeof(StateHeader);
    for (const StatePart& p : parts) n += sizeof(StatePartHeader) + part_data_bytes(p);
    return n;
}

mcb_status state_image_save(hipStream_t s, const StateHeader& info, const std::vector<StatePart>& parts, void* buf,
                            size_t bytes, std::string* err) {
    const size_t need = state_image_bytes(parts);
    if (!buf) return bad(err, "null state buffer");
    if (bytes < need) return bad(err, "state buffer too small (see mcb_state_size)");
    char* out = static_cast<char*>(buf);
    StateHeader h = info;
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kStateVersion;
    h.nparts = (uint32_t)parts.size();
    h.bytes = need;
    std::memcpy(out, &h, sizeof(h));
    size_t off = sizeof(h);
    for (const StatePart& p : parts) {
        StatePartHeader ph{};
        std::strncpy(ph.name, p.name, sizeof(ph.name));
        ph.rows = p.rows;
        ph.elem_bytes = p.elem_bytes;
        ph.row_elems = p.row_elems;
        std::memcpy(out + off, &ph, sizeof(ph));
        off += sizeof(ph);
        const size_t w = p.row_elems * p.elem_bytes;
        if (p.host) {
            for (unsigned r = 0; r < p.rows; ++r)
                std::memcpy(out + off + r * w, static_cast<const char*>(p.ptr) + r * p.row_stride * p.elem_bytes, w);
        } else {
            hipError_t e = hipMemcpy2DAsync(out + off, w, p.ptr, p.row_stride * p.elem_bytes, w, p.rows,
                                            hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) return hip_bad(err, e, "state copy to host");
        }
        off += part_data_bytes(p);
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_bad(err, e, "state copy to host");
    return MCB_OK;
}

mcb_status state_image_header(const void* buf, size_t bytes, StateHeader* h, std::string* err) {
    if (!buf) return bad(err, "null state buffer");
    if (bytes < sizeof(StateHeader)) return bad(err, "state image truncated");
    std::memcpy(h, buf, sizeof(*h));
    if (std::memcmp(h->magic, kMagic, sizeof(kMagic)) != 0) return bad(err, "not a state image");
    if (h->version != kStateVersion) return bad(err, "state image of another version");
    if (h->bytes != bytes) return bad(err, "state image size does not match its header");
    return MCB_OK;
}

mcb_status state_image_load(hipStream_t s, const void* buf, size_t bytes, const std::vector<StatePart>& parts,
                            std::string* err) {
    StateHeader h;
    mcb_status st = state_image_header(buf, bytes, &h, err);
    if (st != MCB_OK) return st;
    if (h.nparts != parts.size() || state_image_bytes(parts) != bytes)
        return bad(err, "state image does not match this lattice's state layout");
    const char* in = static_cast<const char*>(buf);
    size_t off = sizeof(StateHeader);
    // Check every part before copying any.
    for (const StatePart& p : parts) {
        StatePartHeader ph;
        std::memcpy(&ph, in + off, sizeof(ph));
        if (std::strncmp(ph.name, p.name, sizeof(ph.name)) != 0 || ph.rows != p.rows || ph.elem_bytes != p.elem_bytes ||
            ph.row_elems != p.row_elems)
            return bad(err, std::string("state image part mismatch at '") + p.name + "'");
        off += sizeof(ph) + part_data_bytes(p);
    }
    off = sizeof(StateHeader);
    for (const StatePart& p : parts) {
        off += sizeof(StatePartHeader);
        const size_t w = p.row_elems * p.elem_bytes;
        if (p.host) {
            for (unsigned r = 0; r < p.rows; ++r)
                std::memcpy(static_cast<char*>(p.ptr) + r * p.row_stride * p.elem_bytes, in + off + r * w, w);
        } else {
            hipError_t e = hipMemcpy2DAsync(p.ptr, p.row_stride * p.elem_bytes, in + off, w, w, p.rows,
                                            hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return hip_bad(err, e, "state copy to device");
        }
        off += part_data_bytes(p);
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_bad(err, e, "state copy to device");
    return MCB_OK;
}

bool same_state_params(const mcb_params& a, const mcb_params& b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.bc_mode == b.bc_mode && a.weight_set == b.weight_set &&
           a.storage == b.storage && a.tau_f == b.tau_f && a.tau_g == b.tau_g && a.sigma == b.sigma &&
           a.sharp_c == b.sharp_c && a.moving_wall == b.moving_wall && a.f_form == b.f_form;
}

}  // namespace mcb

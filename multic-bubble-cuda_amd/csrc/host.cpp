// Host-side part of libmcb: parameter checks, physical constants, the
// initial bubble field and the reference's output formats. No device calls,
// so all of it runs (and is tested) without a GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mcb.h"
#include "../../include/mcb_ops.h"
#include "bubble.hpp"
#include "lattice.hpp"

namespace mcb {

mcb_status host_validate(const mcb_params* p, std::string* why) {
    auto bad = [&](const char* m) {
        if (why) *why = m;
        return MCB_ERR_INVALID_ARG;
    };
    if (!p) return bad("null parameters");
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return bad("every side must be >= 3");
    // Kernels address a population plane with 32-bit byte offsets.
    if ((long long)p->nx * p->ny * p->nz >= (1LL << 30)) return bad("lattice exceeds 2^30 cells");
    if (p->nz > 65535 || p->ny > 65535) return bad("ny and nz must be <= 65535");
    if (p->bc_mode != MCB_BC_PERIODIC && p->bc_mode != MCB_BC_REFERENCE) return bad("unknown bc_mode");
    if (p->weight_set != MCB_WEIGHTS_LITERAL && p->weight_set != MCB_WEIGHTS_UNIT_SUM) return bad("unknown weight_set");
    if (p->storage < MCB_STORE_FP32 || p->storage > MCB_STORE_BF16) return bad("unknown storage");
    if (p->path < MCB_PATH_AUTO || p->path > MCB_PATH_STAGED) return bad("unknown path");
    if (p->normals < MCB_NORMALS_AUTO || p->normals > MCB_NORMALS_RECOMPUTED) return bad("unknown normals");
    if (p->f_form != MCB_F_POPULATIONS && p->f_form != MCB_F_MOMENTS) return bad("unknown f_form");
    if (!(p->tau_f > 0.5f) || !std::isfinite(p->tau_f)) return bad("tau_f must be a finite value > 0.5");
    if (!(p->tau_g > 0.5f) || !std::isfinite(p->tau_g)) return bad("tau_g must be a finite value > 0.5");
    if (!std::isfinite(p->sigma) || p->sigma < 0.0f) return bad("sigma must be finite and >= 0");
    if (!std::isfinite(p->sharp_c)) return bad("sharp_c must be finite");
    if (p->device < 0) return bad("device must be >= 0");
    return MCB_OK;
}

// src/var.cu:21-25 and 40-45, evaluated in float.
Phys host_phys(const mcb_params& p) {
    Phys ph;
    if (p.weight_set == MCB_WEIGHTS_UNIT_SUM) {
        // Nearest floats to 1/3, 1/18, 1/36 whose exact sum w0 + 6 w1 + 12 w2 is 1.
        ph.w[0] = std::ldexp(11184812.0f, -25);
        ph.w[1] = std::ldexp(14913080.0f, -28);
        ph.w[2] = std::ldexp(14913080.0f, -29);
    } else {
        ph.w[0] = 1.0f / 3.0f;
        ph.w[1] = 1.0f / 18.0f;
        ph.w[2] = 1.0f / 36.0f;
    }
    ph.omega = 1.0f / p.tau_f;
    ph.sharp_c = p.sharp_c < 0.0f ? 0.15f * 3.0f : p.sharp_c;
    ph.sigma = p.sigma;
    ph.omega_g = 1.0f / p.tau_g;
    return ph;
}

}  // namespace mcb

using namespace mcb;

extern "C" {

mcb_status mcb_params_reference(mcb_params* p) {
    if (!p) return MCB_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof(*p));
    p->nx = 64;                        // src/var.cu:3-4
    p->ny = 64;
    p->nz = 256;
    p->bc_mode = MCB_BC_REFERENCE;
    p->weight_set = MCB_WEIGHTS_LITERAL;
    p->storage = MCB_STORE_FP32;
    p->path = MCB_PATH_AUTO;
    p->device = 0;
    p->tau_f = 0.505f;                 // src/var.cu:21
    p->tau_g = 1.0f;                   // full relaxation, src/lbm.cu:324-326
    p->sigma = 0.1f;                   // src/var.cu:25
    p->sharp_c = -1.0f;                // 0.15f*3.0f, src/var.cu:24
    p->moving_wall = 0;
    return MCB_OK;
}

mcb_status mcb_validate_params(const mcb_params* p) { return host_validate(p, nullptr); }

mcb_status mcb_op_consts_from_params(const mcb_params* p, mcb_op_consts* c) {
    if (!p || !c) return MCB_ERR_INVALID_ARG;
    mcb_status st = host_validate(p, nullptr);
    if (st != MCB_OK) return st;
    const Phys ph = host_phys(*p);
    std::memset(c, 0, sizeof(*c));
    for (int q = 0; q < 3; ++q) c->w[q] = ph.w[q];
    c->omega = ph.omega;
    c->sharp_c = ph.sharp_c;
    c->sigma = ph.sigma;
    c->omega_g = ph.omega_g;
    c->tau_g_is_one = p->tau_g == 1.0f;
    c->bc_mode = p->bc_mode;
    // initPhase defaults of src/init.cu:39-46: one bubble, R = 20, at N/2.
    c->n_bubbles = 1;
    c->bubbles[0] = (float)p->nx * 0.5f;
    c->bubbles[1] = (float)p->ny * 0.5f;
    c->bubbles[2] = (float)p->nz * 0.5f;
    c->bubbles[3] = 20.0f;
    c->width = 3.0f;
    return MCB_OK;
}

// initPhase (src/init.cu:27-49) on the host, generalised to several
// bubbles (combined by maximum), a width, and minimum-image distances in
// PERIODIC mode. REFERENCE mode fills the interior only.
mcb_status mcb_host_init_phi(const mcb_params* p, int n, const mcb_bubble* bubbles, float width, float* phi_out,
                             size_t len) {
    if (!p || host_validate(p, nullptr) != MCB_OK) return MCB_ERR_INVALID_ARG;
    if (n < 1 || !bubbles || !phi_out || !(width > 0.0f) || !std::isfinite(width)) return MCB_ERR_INVALID_ARG;
    const int NX = p->nx, NY = p->ny, NZ = p->nz;
    if (len != (size_t)NX * NY * NZ) return MCB_ERR_INVALID_ARG;
    for (int b = 0; b < n; ++b) {
        const mcb_bubble& q = bubbles[b];
        if (!std::isfinite(q.cx) || !std::isfinite(q.cy) || !std::isfinite(q.cz) || !std::isfinite(q.r) || q.r < 0.0f)
            return MCB_ERR_INVALID_ARG;
    }
    const bool periodic = p->bc_mode == MCB_BC_PERIODIC;
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const size_t idx = (size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * k);
                const bool face = i == 0 || i == NX - 1 || j == 0 || j == NY - 1 || k == 0 || k == NZ - 1;
                // REFERENCE mode: faces keep cudaMemset(d_phi, 0) (src/var.cu:86).
                phi_out[idx] = !periodic && face ? 0.0f : bubble_phi_cell(i, j, k, NX, NY, NZ, periodic, bubbles, n, width);
            }
    return MCB_OK;
}

// generateSimulationInfoFile (src/auxFunctions.cu:12-47): same lines, same
// key order, so post/fileTreat.py's getSimInfo parses it unchanged.
mcb_status mcb_write_info_file(const char* path, const char* sim_id, const char* velocity_set, int nx, int ny, int nz,
                               int save_steps, int nsteps, float tau, double mlups) {
    if (!path || !sim_id || !velocity_set) return MCB_ERR_INVALID_ARG;
    std::ofstream file(path);
    if (!file.is_open()) return MCB_ERR_IO;
    std::ostringstream ml;
    if (mlups > 0.0) ml << std::scientific << std::setprecision(6) << mlups;
    else ml << "1.187970e+01";
    const char* pad = "                           ";
    file << "---------------------------- SIMULATION INFORMATION ----------------------------\n"
         << pad << "Simulation ID: " << sim_id << '\n'
         << pad << "Velocity set: " << velocity_set << '\n'
         << pad << "Precision: " << "float" << '\n'
         << pad << "NX: " << nx << '\n'
         << pad << "NY: " << ny << '\n'
         << pad << "NZ: " << nz << '\n'
         << pad << "NZ_TOTAL: " << nz << '\n'
         << pad << "Tau: " << tau << '\n'
         << pad << "Umax: 0.000000e+00\n"
         << pad << "FX: 0.000000e+00\n"
         << pad << "FY: 0.000000e+00\n"
         << pad << "FZ: 0.000000e+00\n"
         << pad << "Save steps: " << save_steps << '\n'
         << pad << "Nsteps: " << nsteps << '\n'
         << pad << "MLUPS: " << ml.str() << '\n'
         << "--------------------------------------------------------------------------------\n";
    file.close();
    return file.fail() ? MCB_ERR_IO : MCB_OK;
}

// copyAndSaveToBinary's file part (src/auxFunctions.cu:57-66).
mcb_status mcb_write_field_file(const char* dir, const char* sim_id, const char* var_name, int step, const float* data,
                                size_t n) {
    if (!dir || !sim_id || !var_name || !data || step < 0) return MCB_ERR_INVALID_ARG;
    std::ostringstream name;
    name << dir << sim_id << "_" << var_name << std::setw(6) << std::setfill('0') << step << ".bin";
    std::ofstream file(name.str(), std::ios::binary);
    if (!file) return MCB_ERR_IO;
    file.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(float)));
    file.close();
    return file.fail() ? MCB_ERR_IO : MCB_OK;
}

int mcb_abi_version(void) { return MCB_ABI_VERSION; }

}  // extern "C"

namespace {

// splitmix64 (Steele, Lea, Flood 2014) and its top-53-bit uniform double.
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double min_image(double d, double n) {
    d = std::fmod(d, n);
    if (d > 0.5 * n) d -= n;
    if (d < -0.5 * n) d += n;
    return d;
}

void periodic_case(mcb_case* c, int n, float tau_f, float tau_g, float sigma, int nsteps) {
    std::memset(c, 0, sizeof(*c));
    mcb_params_reference(&c->params);
    c->params.nx = c->params.ny = c->params.nz = n;
    c->params.bc_mode = MCB_BC_PERIODIC;
    c->params.weight_set = MCB_WEIGHTS_UNIT_SUM;
    c->params.tau_f = tau_f;
    c->params.tau_g = tau_g;
    c->params.sigma = sigma;
    c->width = 3.0f;
    c->nsteps = nsteps;
    c->save_every = 100;
}

}  // namespace

extern "C" {

mcb_status mcb_case_preset(const char* name, uint64_t seed, mcb_case* out) {
    if (!name || !out) return MCB_ERR_INVALID_ARG;
    const std::string nm(name);
    SplitMix64 rng{seed};
    if (nm == "reference") {
        std::memset(out, 0, sizeof(*out));
        mcb_params_reference(&out->params);
        out->n_bubbles = 1;
        // src/init.cu:39-43: R = 20 at (NX*0.5f, NY*0.5f, NZ*0.5f).
        out->bubbles[0] = mcb_bubble{64 * 0.5f, 64 * 0.5f, 256 * 0.5f, 20.0f};
        out->width = 3.0f;
        out->nsteps = 25000;
        out->save_every = 100;
        return MCB_OK;
    }
    if (nm == "cfg0") {
        periodic_case(out, 64, 0.55f, 0.8f, 0.1f, 1000);
        out->n_bubbles = 1;
        out->bubbles[0] = mcb_bubble{32.0f, 32.0f, 32.0f, 10.0f};
        return MCB_OK;
    }
    if (nm == "cfg1") {
        periodic_case(out, 128, 0.55f, 0.8f, 0.1f, 10000);
        out->n_bubbles = 1;
        const float r = (float)(12.0 + 12.0 * rng.uniform());
        const float cx = (float)(128.0 * rng.uniform());
        const float cy = (float)(128.0 * rng.uniform());
        const float cz = (float)(128.0 * rng.uniform());
        out->bubbles[0] = mcb_bubble{cx, cy, cz, r};
        return MCB_OK;
    }
    if (nm == "cfg2") {
        periodic_case(out, 256, 0.52f, 1.0f, 0.1f, 20000);
        out->n_bubbles = 2;
        const double n = 256.0, w = 3.0;
        for (int attempt = 0; attempt < 100000; ++attempt) {
            double r[2], c[2][3];
            for (int b = 0; b < 2; ++b) {
                r[b] = 16.0 + 16.0 * rng.uniform();
                for (int a = 0; a < 3; ++a) c[b][a] = n * rng.uniform();
            }
            double d2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double d = min_image(c[0][a] - c[1][a], n);
                d2 += d * d;
            }
            if (std::sqrt(d2) >= r[0] + r[1] + 4.0 * w) {
                for (int b = 0; b < 2; ++b)
                    out->bubbles[b] = mcb_bubble{(float)c[b][0], (float)c[b][1], (float)c[b][2], (float)r[b]};
                return MCB_OK;
            }
        }
        return MCB_ERR_INVALID_ARG;
    }
    if (nm == "batch" || nm == "serve") {
        const bool serve = nm == "serve";
        int side;
        if (serve) {
            side = 48 + 16 * (int)(10.0 * rng.uniform());          // 48, 64, ..., 192
        } else {
            static const int sides[5] = {32, 48, 64, 96, 128};
            side = sides[(int)(5.0 * rng.uniform())];
        }
        const double rlo = serve ? 0.1 : 0.15;
        const float r = (float)(side * (rlo + (0.3 - rlo) * rng.uniform()));
        const float sigma = (float)(0.01 + 0.09 * rng.uniform());
        const float tau_f = (float)(0.52 + 0.28 * rng.uniform());
        int nsteps = 2000;
        if (serve) nsteps = (int)std::lround(500.0 * std::pow(10.0, rng.uniform()));
        periodic_case(out, side, tau_f, 1.0f, sigma, nsteps);
        if (serve) out->params.storage = MCB_STORE_FP16;
        out->n_bubbles = 1;
        const float cx = (float)(side * rng.uniform()), cy = (float)(side * rng.uniform());
        const float cz = (float)(side * rng.uniform());
        out->bubbles[0] = mcb_bubble{cx, cy, cz, r};
        return MCB_OK;
    }
    return MCB_ERR_INVALID_ARG;
}

mcb_status mcb_serve_trace(uint64_t seed, int n, double rate_per_s, double step_scale, mcb_request* out) {
    if (n < 0 || (n > 0 && !out) || !(rate_per_s > 0.0) || !(step_scale > 0.0)) return MCB_ERR_INVALID_ARG;
    SplitMix64 gaps{seed ^ 0xA5A5A5A5DEADBEEFULL};
    double t = 0.0;
    for (int i = 0; i < n; ++i) {
        std::memset(&out[i], 0, sizeof(out[i]));
        mcb_status st = mcb_case_preset("serve", seed * 1000003ULL + (uint64_t)i, &out[i].c);
        if (st != MCB_OK) return st;
        out[i].c.nsteps = std::max(1, (int)std::lround(out[i].c.nsteps * step_scale));
        // Exponential gap: -ln(1 - U) / rate, U in [0, 1).
        t += -std::log1p(-gaps.uniform()) / rate_per_s;
        out[i].arrival_s = t;
    }
    return MCB_OK;
}

}  // extern "C"

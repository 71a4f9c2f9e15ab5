// mcb — host driver with the reference's command line and output tree.
//
//   mcb <velocity_set> <ID> [extra positional args ignored] [options]
//
// Mirrors src/main.cu: output directory <root>bin/<VS>/<ID>/ (root "../"
// as in src/main.cu:14-21), <ID>_info.txt (src/auxFunctions.cu:12-47), then
// NSTEPS+1 iterations (src/main.cu:74) with <ID>_phi<STEP:06d>.bin written
// whenever STEP % MACRO_SAVE == 0 (src/main.cu:184-189). The file holds d_phi
// after iteration STEP, as in the reference. Only libmcb's C ABI is used.
//
// Options:
//   --preset reference|cfg0|cfg1|cfg2   case (default: reference)
//   --seed S            seed of the seeded presets (default 0)
//   --steps N           NSTEPS override
//   --save-every K      MACRO_SAVE override
//   --root DIR          output root (default "../")
//   --tau-f T --tau-g T --sigma S       physics overrides
//   --path auto|fused|staged            kernel path
//   --extra-fields      also write rho, ux, uy, uz (names without "phi")
//   --device D          HIP device ordinal
//   --load-state FILE   resume from a state image (mcb_save_state) of the
//                       same case: iterations continue after those it holds
//   --save-state FILE   write the state image at the end of the run
//   --state-every K     also write it after every K iterations (replacing
//                       FILE through a temporary file and a rename)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/mcb.h"

namespace {

int die(const char* what, mcb_status st, mcb_ctx* ctx) {
    std::cerr << "mcb: " << what << ": " << mcb_status_string(st);
    if (ctx) std::cerr << " (" << mcb_last_error(ctx) << ")";
    std::cerr << std::endl;
    return 2;
}

bool read_file(const std::string& path, std::vector<char>* out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    in.seekg(0, std::ios::end);
    const std::streamoff n = in.tellg();
    if (n < 0) return false;
    out->resize((size_t)n);
    in.seekg(0);
    return (bool)in.read(out->data(), n);
}

// The context's state image into path, through path.tmp and a rename.
mcb_status write_state(mcb_ctx* ctx, const std::string& path) {
    size_t n = 0;
    mcb_status st = mcb_state_size(ctx, &n);
    if (st != MCB_OK) return st;
    std::vector<char> buf(n);
    if ((st = mcb_save_state(ctx, buf.data(), n)) != MCB_OK) return st;
    const std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
        if (!out || !out.write(buf.data(), (std::streamsize)n)) return MCB_ERR_IO;
    }
    std::error_code ec;
    std::filesystem::rename(tmp, path, ec);
    return ec ? MCB_ERR_IO : MCB_OK;
}

}  // namespace

int main(int argc, char** argv) {
    auto start = std::chrono::steady_clock::now();
    std::vector<std::string> pos;
    std::string preset = "reference", root = "../", path = "auto";
    unsigned long long seed = 0;
    long long steps_override = -1, save_override = -1;
    float tau_f = -1, tau_g = -1, sigma = -1;
    bool extra = false;
    int device = 0;
    std::string load_state, save_state;
    long long state_every = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const char* nm) -> std::string {
            if (i + 1 >= argc) {
                std::cerr << "mcb: " << nm << " needs a value" << std::endl;
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--preset") preset = val("--preset");
        else if (a == "--seed") seed = std::strtoull(val("--seed").c_str(), nullptr, 10);
        else if (a == "--steps") steps_override = std::atoll(val("--steps").c_str());
        else if (a == "--save-every") save_override = std::atoll(val("--save-every").c_str());
        else if (a == "--root") root = val("--root");
        else if (a == "--tau-f") tau_f = std::strtof(val("--tau-f").c_str(), nullptr);
        else if (a == "--tau-g") tau_g = std::strtof(val("--tau-g").c_str(), nullptr);
        else if (a == "--sigma") sigma = std::strtof(val("--sigma").c_str(), nullptr);
        else if (a == "--path") path = val("--path");
        else if (a == "--extra-fields") extra = true;
        else if (a == "--device") device = std::atoi(val("--device").c_str());
        else if (a == "--load-state") load_state = val("--load-state");
        else if (a == "--save-state") save_state = val("--save-state");
        else if (a == "--state-every") state_every = std::atoll(val("--state-every").c_str());
        else if (a.rfind("--", 0) == 0) {
            std::cerr << "mcb: unknown option " << a << std::endl;
            return 1;
        } else pos.push_back(a);
    }
    // src/main.cu:7-10
    if (pos.size() < 2) {
        std::cerr << "Usage: " << argv[0] << " <velocity set> <ID> [options]" << std::endl;
        return 1;
    }
    const std::string VS = pos[0], ID = pos[1];
    if (VS != "D3Q19") {
        std::cerr << "mcb: velocity set " << VS << " is not supported (D3Q19 only)" << std::endl;
        return 1;
    }

    mcb_case cs;
    mcb_status st = mcb_case_preset(preset.c_str(), seed, &cs);
    if (st != MCB_OK) return die("unknown preset", st, nullptr);
    if (steps_override >= 0) cs.nsteps = (int)steps_override;
    if (save_override > 0) cs.save_every = (int)save_override;
    if (tau_f > 0) cs.params.tau_f = tau_f;
    if (tau_g > 0) cs.params.tau_g = tau_g;
    if (sigma >= 0) cs.params.sigma = sigma;
    cs.params.device = device;
    cs.params.path = path == "fused" ? MCB_PATH_FUSED : (path == "staged" ? MCB_PATH_STAGED : MCB_PATH_AUTO);

    const std::string sim_dir = root + "bin/" + VS + "/" + ID + "/";
    std::error_code ec;
    std::filesystem::create_directories(sim_dir, ec);
    if (ec) {
        std::cerr << "mcb: cannot create " << sim_dir << ": " << ec.message() << std::endl;
        return 2;
    }

    if (state_every > 0 && save_state.empty()) {
        std::cerr << "mcb: --state-every needs --save-state" << std::endl;
        return 1;
    }
    mcb_ctx* ctx = nullptr;
    if ((st = mcb_create(&cs.params, nullptr, &ctx)) != MCB_OK) return die("mcb_create", st, nullptr);
    // Iterations already taken: 0, or those a loaded state image holds.
    long long first = 0;
    if (!load_state.empty()) {
        std::vector<char> img;
        if (!read_file(load_state, &img)) {
            std::cerr << "mcb: cannot read " << load_state << std::endl;
            return 2;
        }
        if ((st = mcb_load_state(ctx, img.data(), img.size())) != MCB_OK) return die("load state", st, ctx);
        if ((st = mcb_state_info(img.data(), img.size(), nullptr, &first)) != MCB_OK)
            return die("state info", st, ctx);
        std::cout << "Resumed from " << load_state << " after " << first << " iterations" << std::endl;
    } else if ((st = mcb_init_bubbles(ctx, cs.n_bubbles, cs.bubbles, cs.width)) != MCB_OK) {
        return die("init", st, ctx);
    }

    const int NX = cs.params.nx, NY = cs.params.ny, NZ = cs.params.nz;
    const size_t cells = (size_t)NX * NY * NZ;
    const std::string info = sim_dir + ID + "_info.txt";
    if ((st = mcb_write_info_file(info.c_str(), ID.c_str(), VS.c_str(), NX, NY, NZ, cs.save_every, cs.nsteps,
                                  cs.params.tau_f, -1.0)) != MCB_OK)
        return die("info file", st, ctx);

    std::vector<float> host(cells);
    auto save = [&](int step) -> mcb_status {
        mcb_status s = mcb_read_field(ctx, MCB_FIELD_PHI, host.data(), cells);
        if (s != MCB_OK) return s;
        s = mcb_write_field_file(sim_dir.c_str(), ID.c_str(), "phi", step, host.data(), cells);
        if (s != MCB_OK || !extra) return s;
        if ((s = mcb_compute_macros(ctx)) != MCB_OK) return s;
        const struct { int f; const char* nm; } xs[] = {
            {MCB_FIELD_RHO, "rho"}, {MCB_FIELD_UX, "ux"}, {MCB_FIELD_UY, "uy"}, {MCB_FIELD_UZ, "uz"}};
        for (auto& x : xs) {
            if ((s = mcb_read_field(ctx, x.f, host.data(), cells)) != MCB_OK) return s;
            if ((s = mcb_write_field_file(sim_dir.c_str(), ID.c_str(), x.nm, step, host.data(), cells)) != MCB_OK)
                return s;
        }
        return MCB_OK;
    };

    // STEP = 0..NSTEPS inclusive; after iteration STEP, save if STEP % save == 0.
    // A state image written after iteration STEP holds STEP + 1 iterations.
    double device_ms = 0.0;
    long long done = 0;
    for (long long step = first; step <= cs.nsteps;) {
        long long next_save = ((step + cs.save_every - 1) / cs.save_every) * cs.save_every;
        long long last = std::min<long long>(next_save, cs.nsteps);
        if (state_every > 0) last = std::min(last, ((step + state_every) / state_every) * state_every - 1);
        int chunk = (int)(last - step + 1);
        double ms = 0.0;
        if ((st = mcb_step_timed(ctx, chunk, &ms)) != MCB_OK) return die("step", st, ctx);
        device_ms += ms;
        done += chunk;
        step = last + 1;
        if (last % cs.save_every == 0) {
            if ((st = save((int)last)) != MCB_OK) return die("save", st, ctx);
            std::cout << "Step " << last << " of " << cs.nsteps << ": data saved in " << sim_dir << std::endl;
        }
        if (state_every > 0 && (last + 1) % state_every == 0 && (st = write_state(ctx, save_state)) != MCB_OK)
            return die("save state", st, ctx);
    }
    if (!save_state.empty()) {
        if ((st = write_state(ctx, save_state)) != MCB_OK) return die("save state", st, ctx);
        std::cout << "State after " << std::max(first, cs.nsteps + 1LL) << " iterations saved in " << save_state
                  << std::endl;
    }
    mcb_destroy(ctx);

    auto end = std::chrono::steady_clock::now();
    const double wall = std::chrono::duration<double>(end - start).count();
    const double mlups_dev = device_ms > 0 ? (double)cells * done / (device_ms * 1e3) : 0.0;
    std::cout << "\n// =============================================== //\n"
              << "     Total wall time: " << wall << " s\n"
              << "     Iterations: " << done << (first > 0 ? " (after the loaded state's)" : " (NSTEPS+1)") << "\n"
              << "     Device time: " << device_ms / 1e3 << " s, " << mlups_dev << " MLUPS\n"
              << "// =============================================== //\n"
              << std::endl;
    return 0;
}

// TEST INFRASTRUCTURE — NOT PART OF THE PRODUCT.
// C entry points over the transcription in oracle.hpp, for ctypes (tests/,
// smoke() and bench.py's cpu_baseline leg). Every entry point takes an opaque
// handle holding either an FP32 or an FP64 state.
#include "oracle.hpp"
#include <chrono>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace mcbo;

namespace {

struct Handle {
    int prec;  // 32 or 64
    std::unique_ptr<State<float>> s32;
    std::unique_ptr<State<double>> s64;
};

template <class T> std::vector<T>* field_of(State<T>& s, int id) {
    switch (id) {
        case 0: return &s.f;
        case 1: return &s.g;
        case 2: return &s.g_out;
        case 3: return &s.phi;
        case 4: return &s.normx;
        case 5: return &s.normy;
        case 6: return &s.normz;
        case 7: return &s.indicator;
        case 8: return &s.curvature;
        case 9: return &s.ffx;
        case 10: return &s.ffy;
        case 11: return &s.ffz;
        case 12: return &s.ux;
        case 13: return &s.uy;
        case 14: return &s.uz;
        case 15: return &s.rho;
        case 16: return &s.pxx;
        case 17: return &s.pyy;
        case 18: return &s.pzz;
        case 19: return &s.pxy;
        case 20: return &s.pxz;
        case 21: return &s.pyz;
        default: return nullptr;
    }
}

template <class T> void run_stage(State<T>& s, int stage) {
    switch (stage) {
        case 0: phi_calc(s); break;
        case 1: grad_calc(s); break;
        case 2: curvature_calc(s); break;
        case 3: momenti_calc(s); break;
        case 4: collision_fluid(s); break;
        case 5: collision_phase(s); break;
        case 6: streaming_calc(s); break;
        case 7: fg_boundary(s); break;
        case 8: boundary_conditions(s); break;
        case 9: moving_wall_bc(s); break;
        case 10: init_tensor(s); break;
        case 11: init_dist(s); break;
        default: break;
    }
}

template <class T> void init_case(State<T>& s, int nb, const double* b, double width) {
    std::vector<Bubble> bubbles;
    for (int i = 0; i < nb; ++i) bubbles.push_back(Bubble{b[4 * i], b[4 * i + 1], b[4 * i + 2], b[4 * i + 3]});
    // main.cu:56-68: initTensor, initPhase, initDist.
    init_tensor(s);
    init_phase(s, bubbles, T(width));
    init_dist(s);
}

template <class T> double sum_phi(const State<T>& s) {
    double acc = 0.0;
    for (size_t n = 0; n < s.N; ++n) acc += (double)s.phi[n];
    return acc;
}

template <class T> double kinetic(const State<T>& s) {
    double acc = 0.0;
    for (size_t n = 0; n < s.N; ++n) {
        double u2 = (double)s.ux[n] * s.ux[n] + (double)s.uy[n] * s.uy[n] + (double)s.uz[n] * s.uz[n];
        acc += 0.5 * (double)s.rho[n] * u2;
    }
    return acc;
}

}  // namespace

extern "C" {

void* ora_create(int prec, const Params* p) {
    if (!p || p->nx < 3 || p->ny < 3 || p->nz < 3) return nullptr;
    auto* h = new Handle();
    h->prec = prec;
    if (prec == 64) h->s64.reset(new State<double>(*p));
    else { h->prec = 32; h->s32.reset(new State<float>(*p)); }
    return h;
}

void ora_destroy(void* hp) { delete static_cast<Handle*>(hp); }

int ora_prec(void* hp) { return static_cast<Handle*>(hp)->prec; }

void ora_init(void* hp, int nb, const double* bubbles, double width) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) init_case(*h->s64, nb, bubbles, width);
    else init_case(*h->s32, nb, bubbles, width);
}

void ora_step(void* hp, int n) {
    auto* h = static_cast<Handle*>(hp);
    for (int t = 0; t < n; ++t) {
        if (h->prec == 64) step(*h->s64);
        else step(*h->s32);
    }
}

void ora_macros(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) macros(*h->s64);
    else macros(*h->s32);
}

void ora_stage(void* hp, int stage) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) run_stage(*h->s64, stage);
    else run_stage(*h->s32, stage);
}

// Direct pointer into a field (element type = the handle's precision).
void* ora_field_ptr(void* hp, int id) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) { auto* v = field_of(*h->s64, id); return v ? v->data() : nullptr; }
    auto* v = field_of(*h->s32, id);
    return v ? v->data() : nullptr;
}

size_t ora_field_len(void* hp, int id) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) { auto* v = field_of(*h->s64, id); return v ? v->size() : 0; }
    auto* v = field_of(*h->s32, id);
    return v ? v->size() : 0;
}

double ora_sum_phi(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    return h->prec == 64 ? sum_phi(*h->s64) : sum_phi(*h->s32);
}

double ora_kinetic(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    return h->prec == 64 ? kinetic(*h->s64) : kinetic(*h->s32);
}

// Weights and derived constants actually used (for table checks).
void ora_weights(void* hp, double* w19) {
    auto* h = static_cast<Handle*>(hp);
    for (int l = 0; l < 19; ++l) w19[l] = h->prec == 64 ? h->s64->c.w[l] : (double)h->s32->c.w[l];
}

void ora_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int ora_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Wall time of n full steps (steady_clock), for the CPU baseline.
double ora_time_steps(void* hp, int n) {
    auto t0 = std::chrono::steady_clock::now();
    ora_step(hp, n);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// FP32 tanh of the initial profile (tanh_f32) against the C library's double
// tanh rounded to float, for every float bit pattern in [lo, hi] (non-negative
// floats order like their bit patterns). out[0] = values compared, out[1] =
// results that differ, out[2] = results more than one ulp apart.
void ora_tanh_scan(uint32_t lo, uint32_t hi, long long* out) {
    long long n = 0, diff = 0, far = 0;
#pragma omp parallel for reduction(+ : n, diff, far) schedule(static, 1 << 16)
    for (long long u = (long long)lo; u <= (long long)hi; ++u) {
        const uint32_t bits = (uint32_t)u;
        float x;
        std::memcpy(&x, &bits, 4);
        const float a = mcb::tanh_f32(x);
        const float b = (float)std::tanh((double)x);
        ++n;
        if (std::memcmp(&a, &b, 4) != 0) {
            ++diff;
            int32_t ia, ib;
            std::memcpy(&ia, &a, 4);
            std::memcpy(&ib, &b, 4);
            if (ia - ib > 1 || ib - ia > 1) ++far;
        }
    }
    out[0] = n;
    out[1] = diff;
    out[2] = far;
}

}  // extern "C"

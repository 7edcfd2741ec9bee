// time_step.hip - the time-stepped beta loop (hase_simulation_*).
//
// The reference's CompiledSimulationRunner (include/core/timeSteppedSimulation.hpp:132-615)
// on the prism mesh: beta per prism is the FP64 state, resident on the
// device; every stage refreshes the context's walk copies of beta from it
// (hase::refresh_beta_device), evaluates the ASE flux at the sample points
// into a device array (phi_keep), and one kernel averages that flux over each
// prism's six vertices and composes the derivative
// (include/kernels/derivativeComposition.hpp:27-60).  The integrator
// combinations are the reference's (include/kernels/timeIntegrationUpdateKernels.hpp:21-123).
//
// The elementwise kernels are a few FP64 operations per prism (HBM-bound,
// microseconds at 10^5-10^6 prisms); floating-point contraction is off in
// them so that they round exactly as the formulas are written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <vector>

#include "../../include/hase_ase.h"
#include "common.h"
#include "context.h"

using hase::DevBuf;
using hase::HaseError;

namespace {

constexpr uint32_t kTsBlock = 256;

enum FluxMode : int { kFluxAverage = 0, kFluxZero = 1, kFluxKeep = 2 };

struct ComposeParams {
    const double* beta;          // [P] stage beta, device prism order
    const float* phi_samples;    // [S] flux of the last ASE evaluation (kFluxAverage)
    double* phi_prism;           // [P] written (average / zero) or read (keep)
    const double* pump_rate;     // [P]
    const double* pump_sat;      // [P] or null
    const uint32_t* tri;         // [3T] vertices of device triangle t at 3t + k
    const uint32_t* clad;        // [T] 1 = cladding
    double* dndt_pump;
    double* dndt_ase;
    double* deriv;
    double sigma_sum;            // sigma_e + sigma_a
    double sigma_a;
    double tau;
    uint32_t T, layers, np;
    int flux_mode;
    int pump_on;
};

// Prism flux and derivative (derivativeComposition.hpp:34-59): cladding
// prisms get zeros; otherwise
//   dndt_ase = (beta (sigma_e + sigma_a) - sigma_a) phi,
//   deriv    = pump - dndt_ase - beta / tau.
__global__ void __launch_bounds__(kTsBlock) hase_compose_kernel(ComposeParams p)
{
#pragma clang fp contract(off)
    const uint32_t P = p.T * p.layers;
    const uint32_t j = blockIdx.x * kTsBlock + threadIdx.x;
    if (j >= P) return;
    const uint32_t t = j % p.T, l = j / p.T;
    double phi;
    if (p.flux_mode == kFluxAverage) {
        const uint32_t v0 = p.tri[3u * t], v1 = p.tri[3u * t + 1u], v2 = p.tri[3u * t + 2u];
        const float* lo = p.phi_samples + (size_t)l * p.np;
        const float* hi = lo + p.np;
        const double sum = (double)lo[v0] + (double)lo[v1] + (double)lo[v2] + (double)hi[v0] + (double)hi[v1] +
                           (double)hi[v2];
        phi = sum / 6.0;
        p.phi_prism[j] = phi;
    } else if (p.flux_mode == kFluxZero) {
        phi = 0.0;
        p.phi_prism[j] = 0.0;
    } else {
        phi = p.phi_prism[j];
    }
    if (p.clad[t]) {
        p.dndt_pump[j] = 0.0;
        p.dndt_ase[j] = 0.0;
        p.deriv[j] = 0.0;
        return;
    }
    const double b = p.beta[j];
    double pump = 0.0;
    if (p.pump_on) pump = p.pump_sat ? p.pump_rate[j] - p.pump_sat[j] * b : p.pump_rate[j];
    const double gain = b * p.sigma_sum - p.sigma_a;
    const double ase = gain * phi;
    p.dndt_pump[j] = pump;
    p.dndt_ase[j] = ase;
    p.deriv[j] = pump - ase - b / p.tau;
}

enum CombineKind : int { kAddScaled = 0, kHeun = 1, kRk4 = 2, kExpEuler = 3 };

struct CombineParams {
    const double* base;
    const double* k0;
    const double* k1;
    const double* k2;
    const double* k3;
    double* out;
    double scale;     // kAddScaled: scale; kHeun, kRk4: time step
    double decay;     // kExpEuler: exp(-time_step / tau)
    double tau;
    uint32_t n;
    int kind;
    int clip;         // clamp the result to [0, 1] (ClipBeta)
};

// timeIntegrationUpdateKernels.hpp: AddScaled (:21-40), CombineHeun (:42-62),
// CombineRungeKutta4 (:64-86), ExponentialEulerUpdate with k0 = dndt_pump and
// k1 = dndt_ase (:88-111), ClipBeta (:113-123).
__global__ void __launch_bounds__(kTsBlock) hase_combine_kernel(CombineParams p)
{
#pragma clang fp contract(off)
    const uint32_t j = blockIdx.x * kTsBlock + threadIdx.x;
    if (j >= p.n) return;
    double v;
    switch (p.kind) {
    case kAddScaled: v = p.base[j] + p.scale * p.k0[j]; break;
    case kHeun: v = p.base[j] + 0.5 * p.scale * (p.k0[j] + p.k1[j]); break;
    case kRk4: v = p.base[j] + (p.scale / 6.0) * (p.k0[j] + 2.0 * p.k1[j] + 2.0 * p.k2[j] + p.k3[j]); break;
    default: {
        const double source = p.k0[j] - p.k1[j];
        v = p.tau * source * (1.0 - p.decay) + p.base[j] * p.decay;
    }
    }
    if (p.clip) v = fmin(1.0, fmax(0.0, v));
    p.out[j] = v;
}

}  // namespace

struct hase_simulation {
    hase_context* ctx = nullptr;
    hase_run_control_t run{};
    uint32_t P = 0, S = 0;
    uint32_t completed = 0;
    uint32_t ase_evaluations = 0;
    double sigma_a = 0.0, tau = 1.0;
    bool has_sat = false;
    DevBuf<double> beta, beta_next, stage, deriv, dndt_pump, dndt_ase, phi_prism, k1, k2, k3, k4;
    DevBuf<double> pump_rate, pump_sat, block_sums;
    DevBuf<uint32_t> tri, clad;
    // last ASE evaluation (host)
    std::vector<float> phi_s;
    std::vector<double> mse, dndt_s, se;
    std::vector<uint32_t> rays, dropped;
    double ase_ms = 0.0, integrator_ms = 0.0;
    uint64_t rays_total = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ~hase_simulation()
    {
        if (ctx) {
            (void)hipSetDevice(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);
            ctx->phi_keep = false;
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

namespace {

// sigma_a at the peak of sigma_e, or max_sigma_a without spectra
// (timeSteppedSimulation.hpp:45-54).
double absorption_at_emission_peak(const hase_experiment_t& e)
{
    if (e.number_of_wavelengths == 0 || !e.sigma_a || !e.sigma_e) return e.max_sigma_a;
    uint32_t best = 0;
    for (uint32_t i = 1; i < e.number_of_wavelengths; ++i)
        if (e.sigma_e[i] > e.sigma_e[best]) best = i;
    return e.sigma_a[best];
}

void check_run(const hase_run_control_t* run)
{
    if (!run) throw HaseError(HASE_E_INVALID, "run control must not be NULL");
    if (!(run->time_step > 0.0) || !std::isfinite(run->time_step))
        throw HaseError(HASE_E_INVALID, "time_step must be positive");
    if (run->integrator > HASE_INTEGRATOR_EXPONENTIAL_EULER)
        throw HaseError(HASE_E_INVALID, "unsupported time integrator");
}

// caller prism order -> device prism order
std::vector<double> to_device_order(const hase_context* ctx, const double* user, const char* what, bool finite)
{
    const uint32_t T = ctx->T, P = ctx->P;
    std::vector<double> out(P);
    for (uint32_t j = 0; j < P; ++j) {
        const double v = user[ctx->dev2user[j % T] + (size_t)(j / T) * T];
        if (finite && !std::isfinite(v)) throw HaseError(HASE_E_INVALID, std::string(what) + " must be finite");
        out[j] = v;
    }
    return out;
}

void to_user_order(const hase_context* ctx, const double* dev, double* user)
{
    const uint32_t T = ctx->T, P = ctx->P;
    for (uint32_t j = 0; j < P; ++j) user[ctx->dev2user[j % T] + (size_t)(j / T) * T] = dev[j];
}

void set_pump(hase_simulation* sim, const double* rate, const double* sat)
{
    if (!rate) throw HaseError(HASE_E_INVALID, "pump_rate is required (zeros for no pump)");
    hase_context* ctx = sim->ctx;
    auto r = to_device_order(ctx, rate, "pump_rate", true);
    sim->pump_rate.upload(r.data(), sim->P, ctx->stream);
    sim->has_sat = sat != nullptr;
    if (sat) {
        auto s = to_device_order(ctx, sat, "pump_saturation", true);
        sim->pump_sat.upload(s.data(), sim->P, ctx->stream);
    }
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void set_beta(hase_simulation* sim, const double* beta)
{
    hase_context* ctx = sim->ctx;
    auto b = to_device_order(ctx, beta, "beta_volume", true);
    sim->beta.upload(b.data(), sim->P, ctx->stream);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

uint32_t grid(uint32_t n) { return (n + kTsBlock - 1) / kTsBlock; }

void timed_begin(hase_simulation* sim) { HIP_CHECK(hipEventRecord(sim->ev0, sim->ctx->stream)); }

void timed_end(hase_simulation* sim)
{
    HIP_CHECK(hipEventRecord(sim->ev1, sim->ctx->stream));
    HIP_CHECK(hipEventSynchronize(sim->ev1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, sim->ev0, sim->ev1));
    sim->integrator_ms += ms;
}

void combine(hase_simulation* sim, int kind, const double* base, const double* k0, const double* k1,
             const double* k2, const double* k3, double* out, double scale, bool clip)
{
    CombineParams p{};
    p.base = base;
    p.k0 = k0;
    p.k1 = k1;
    p.k2 = k2;
    p.k3 = k3;
    p.out = out;
    p.scale = scale;
    p.tau = sim->tau;
    p.decay = std::exp(-sim->run.time_step / sim->tau);
    p.n = sim->P;
    p.kind = kind;
    p.clip = clip ? 1 : 0;
    timed_begin(sim);
    hipLaunchKernelGGL(hase_combine_kernel, dim3(grid(sim->P)), dim3(kTsBlock), 0, sim->ctx->stream, p);
    HIP_CHECK(hipGetLastError());
    timed_end(sim);
}

void copy(hase_simulation* sim, double* dst, const double* src)
{
    HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)sim->P * sizeof(double), hipMemcpyDeviceToDevice, sim->ctx->stream));
}

// One stage (the evaluateStage lambda of timeSteppedSimulation.hpp:255-314):
// pump and derivative at `beta`; with refresh, the ASE flux of `beta` first
// (zero when ASE is off in this step).
void evaluate_stage(hase_simulation* sim, const double* beta, bool pump_on, bool ase_on, bool refresh)
{
    hase_context* ctx = sim->ctx;
    int mode = kFluxKeep;
    if (refresh) {
        if (ase_on) {
            auto t0 = std::chrono::steady_clock::now();
            hase::refresh_beta_device(ctx, beta, sim->block_sums.p);
            hase_result_t res{sim->phi_s.data(), sim->mse.data(), sim->rays.data(), sim->dndt_s.data(),
                              sim->se.data(), sim->dropped.data()};
            hase::evaluate_context(ctx, &res);
            sim->ase_evaluations++;
            sim->rays_total += ctx->stats.rays;
            sim->ase_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            mode = kFluxAverage;
        } else {
            std::fill(sim->phi_s.begin(), sim->phi_s.end(), 0.0f);
            std::fill(sim->mse.begin(), sim->mse.end(), 0.0);
            std::fill(sim->rays.begin(), sim->rays.end(), 0u);
            std::fill(sim->se.begin(), sim->se.end(), 0.0);
            std::fill(sim->dropped.begin(), sim->dropped.end(), 0u);
            mode = kFluxZero;
        }
    }
    ComposeParams p{};
    p.beta = beta;
    p.phi_samples = ctx->phi_all.p;
    p.phi_prism = sim->phi_prism.p;
    p.pump_rate = sim->pump_rate.p;
    p.pump_sat = sim->has_sat ? sim->pump_sat.p : nullptr;
    p.tri = sim->tri.p;
    p.clad = sim->clad.p;
    p.dndt_pump = sim->dndt_pump.p;
    p.dndt_ase = sim->dndt_ase.p;
    p.deriv = sim->deriv.p;
    p.sigma_sum = ctx->exp.max_sigma_e + sim->sigma_a;
    p.sigma_a = sim->sigma_a;
    p.tau = sim->tau;
    p.T = ctx->T;
    p.layers = ctx->L - 1;
    p.np = ctx->np;
    p.flux_mode = mode;
    p.pump_on = pump_on ? 1 : 0;
    timed_begin(sim);
    hipLaunchKernelGGL(hase_compose_kernel, dim3(grid(sim->P)), dim3(kTsBlock), 0, ctx->stream, p);
    HIP_CHECK(hipGetLastError());
    timed_end(sim);
}

// One physical step (advanceTimeStep, timeSteppedSimulation.hpp:253-365, and
// the integrators of :367-415); ends with beta_next clipped to [0, 1].
void advance(hase_simulation* sim, bool pump_on, bool ase_on)
{
    const double h = sim->run.time_step;
    double* beta = sim->beta.p;
    double* next = sim->beta_next.p;
    double* stage = sim->stage.p;
    double* d = sim->deriv.p;
    auto stage_eval = [&](const double* b, bool refresh) { evaluate_stage(sim, b, pump_on, ase_on, refresh); };
    switch (sim->run.integrator) {
    case HASE_INTEGRATOR_EXPLICIT_EULER:
        stage_eval(beta, true);
        combine(sim, kAddScaled, beta, d, nullptr, nullptr, nullptr, next, h, true);
        break;
    case HASE_INTEGRATOR_HEUN:
        stage_eval(beta, true);
        copy(sim, sim->k1.p, d);
        combine(sim, kAddScaled, beta, sim->k1.p, nullptr, nullptr, nullptr, stage, h, false);
        stage_eval(stage, true);
        combine(sim, kHeun, beta, sim->k1.p, d, nullptr, nullptr, next, h, true);
        break;
    case HASE_INTEGRATOR_MIDPOINT:
        stage_eval(beta, true);
        combine(sim, kAddScaled, beta, d, nullptr, nullptr, nullptr, stage, 0.5 * h, false);
        stage_eval(stage, true);
        combine(sim, kAddScaled, beta, d, nullptr, nullptr, nullptr, next, h, true);
        break;
    case HASE_INTEGRATOR_RUNGE_KUTTA_4:
    case HASE_INTEGRATOR_FROZEN_PHI_ASE_RUNGE_KUTTA_4: {
        const bool refresh = sim->run.integrator == HASE_INTEGRATOR_RUNGE_KUTTA_4;
        stage_eval(beta, true);
        copy(sim, sim->k1.p, d);
        combine(sim, kAddScaled, beta, sim->k1.p, nullptr, nullptr, nullptr, stage, 0.5 * h, false);
        stage_eval(stage, refresh);
        copy(sim, sim->k2.p, d);
        combine(sim, kAddScaled, beta, sim->k2.p, nullptr, nullptr, nullptr, stage, 0.5 * h, false);
        stage_eval(stage, refresh);
        copy(sim, sim->k3.p, d);
        combine(sim, kAddScaled, beta, sim->k3.p, nullptr, nullptr, nullptr, stage, h, false);
        stage_eval(stage, refresh);
        copy(sim, sim->k4.p, d);
        combine(sim, kRk4, beta, sim->k1.p, sim->k2.p, sim->k3.p, sim->k4.p, next, h, true);
        break;
    }
    case HASE_INTEGRATOR_IMPLICIT_EULER: {
        // fixed-point iterations beta_next = beta + h f(stage), stage <- beta_next
        const uint32_t iters = std::max(1u, sim->run.implicit_iterations);
        copy(sim, stage, beta);
        for (uint32_t i = 0; i < iters; ++i) {
            stage_eval(stage, true);
            const bool last = i + 1 == iters;
            combine(sim, kAddScaled, beta, d, nullptr, nullptr, nullptr, next, h, last);
            if (!last) copy(sim, stage, next);
        }
        break;
    }
    default:   // exponential Euler: exact decay, pump and ASE frozen at beta
        stage_eval(beta, true);
        combine(sim, kExpEuler, beta, sim->dndt_pump.p, sim->dndt_ase.p, nullptr, nullptr, next, h, true);
        break;
    }
    std::swap(sim->beta.p, sim->beta_next.p);
}

void step(hase_simulation* sim, uint32_t n)
{
    if (n == 0) throw HaseError(HASE_E_INVALID, "steps must be positive");
    HIP_CHECK(hipSetDevice(sim->ctx->device));
    const hase_run_control_t& r = sim->run;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t s = r.first_simulation_step + sim->completed;
        const bool pump_on = s < r.pump_steps;
        const bool ase_on = s < r.ase_steps && !(r.pre_pump && s == 0u);
        advance(sim, pump_on, ase_on);
        sim->completed++;
    }
    HIP_CHECK(hipStreamSynchronize(sim->ctx->stream));
}

std::unique_ptr<hase_simulation> create(hase_context* ctx, const hase_run_control_t* run, const double* beta,
                                        const double* rate, const double* sat)
{
    if (!ctx) throw HaseError(HASE_E_INVALID, "context must not be NULL");
    check_run(run);
    if (ctx->comp.min_sample_range != 0 || ctx->comp.max_sample_range + 1u != ctx->S)
        throw HaseError(HASE_E_INVALID, "the time-stepped loop needs a context over all sample points");
    HIP_CHECK(hipSetDevice(ctx->device));
    auto sim = std::make_unique<hase_simulation>();
    sim->ctx = ctx;
    sim->run = *run;
    sim->P = ctx->P;
    sim->S = ctx->S;
    sim->sigma_a = absorption_at_emission_peak(ctx->exp);
    sim->tau = std::max(ctx->tau, std::numeric_limits<double>::min());
    const size_t P = sim->P;
    for (DevBuf<double>* b : {&sim->beta, &sim->beta_next, &sim->stage, &sim->deriv, &sim->dndt_pump,
                              &sim->dndt_ase, &sim->phi_prism, &sim->k1, &sim->k2, &sim->k3, &sim->k4})
        b->alloc(P);
    sim->block_sums.alloc(grid(sim->P));
    sim->tri.upload(ctx->tri_vertex.data(), ctx->tri_vertex.size(), ctx->stream);
    sim->clad.upload(ctx->clad_flag.data(), ctx->T, ctx->stream);
    for (DevBuf<double>* b : {&sim->dndt_pump, &sim->dndt_ase, &sim->phi_prism, &sim->deriv})
        HIP_CHECK(hipMemsetAsync(b->p, 0, P * sizeof(double), ctx->stream));
    sim->phi_s.assign(sim->S, 0.0f);
    sim->mse.assign(sim->S, 0.0);
    sim->dndt_s.assign(sim->S, 0.0);
    sim->se.assign(sim->S, 0.0);
    sim->rays.assign(sim->S, 0u);
    sim->dropped.assign(sim->S, 0u);
    if (beta) {
        set_beta(sim.get(), beta);
    } else {
        set_beta(sim.get(), ctx->beta_host.data());
    }
    set_pump(sim.get(), rate, sat);
    HIP_CHECK(hipEventCreate(&sim->ev0));
    HIP_CHECK(hipEventCreate(&sim->ev1));
    ctx->phi_keep = true;
    ctx->phi_all.alloc(sim->S);
    HIP_CHECK(hipMemsetAsync(ctx->phi_all.p, 0, (size_t)sim->S * sizeof(float), ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return sim;
}

void state(const hase_simulation* sim, hase_time_state_t* st)
{
    if (!st) throw HaseError(HASE_E_INVALID, "state must not be NULL");
    const hase_context* ctx = sim->ctx;
    HIP_CHECK(hipSetDevice(ctx->device));
    st->step = sim->completed;
    st->ase_evaluations = sim->ase_evaluations;
    st->time = (double)sim->completed * sim->run.time_step;
    std::vector<double> tmp(sim->P);
    auto fetch = [&](const double* dev, double* user) {
        if (!user) return;
        HIP_CHECK(hipMemcpy(tmp.data(), dev, (size_t)sim->P * sizeof(double), hipMemcpyDeviceToHost));
        to_user_order(ctx, tmp.data(), user);
    };
    fetch(sim->beta.p, st->beta_volume);
    fetch(sim->phi_prism.p, st->phi_ase);
    fetch(sim->dndt_ase.p, st->dndt_ase);
    fetch(sim->dndt_pump.p, st->dndt_pump);
    if (st->phi_samples) std::copy(sim->phi_s.begin(), sim->phi_s.end(), st->phi_samples);
    if (st->mse) std::copy(sim->mse.begin(), sim->mse.end(), st->mse);
    if (st->total_rays) std::copy(sim->rays.begin(), sim->rays.end(), st->total_rays);
    if (st->standard_error) std::copy(sim->se.begin(), sim->se.end(), st->standard_error);
}

}  // namespace

extern "C" {

int hase_simulation_create(hase_context_t* context, const hase_run_control_t* run, const double* beta_volume,
                           const double* pump_rate, const double* pump_saturation, hase_simulation_t** out_sim)
{
    return hase::guarded([&] {
        if (!out_sim) throw HaseError(HASE_E_INVALID, "out_sim must not be NULL");
        *out_sim = nullptr;
        *out_sim = create(context, run, beta_volume, pump_rate, pump_saturation).release();
    });
}

int hase_simulation_set_pump(hase_simulation_t* sim, const double* pump_rate, const double* pump_saturation)
{
    return hase::guarded([&] {
        if (!sim) throw HaseError(HASE_E_INVALID, "simulation must not be NULL");
        HIP_CHECK(hipSetDevice(sim->ctx->device));
        set_pump(sim, pump_rate, pump_saturation);
    });
}

int hase_simulation_set_beta(hase_simulation_t* sim, const double* beta_volume)
{
    return hase::guarded([&] {
        if (!sim || !beta_volume) throw HaseError(HASE_E_INVALID, "simulation and beta_volume must not be NULL");
        HIP_CHECK(hipSetDevice(sim->ctx->device));
        set_beta(sim, beta_volume);
    });
}

int hase_simulation_step(hase_simulation_t* sim, uint32_t n_steps)
{
    return hase::guarded([&] {
        if (!sim) throw HaseError(HASE_E_INVALID, "simulation must not be NULL");
        step(sim, n_steps);
    });
}

int hase_simulation_state(const hase_simulation_t* sim, hase_time_state_t* st)
{
    return hase::guarded([&] {
        if (!sim) throw HaseError(HASE_E_INVALID, "simulation must not be NULL");
        state(sim, st);
    });
}

int hase_simulation_times(const hase_simulation_t* sim, double* ase_ms, double* integrator_ms, uint64_t* rays)
{
    return hase::guarded([&] {
        if (!sim) throw HaseError(HASE_E_INVALID, "simulation must not be NULL");
        if (ase_ms) *ase_ms = sim->ase_ms;
        if (integrator_ms) *integrator_ms = sim->integrator_ms;
        if (rays) *rays = sim->rays_total;
    });
}

void hase_simulation_destroy(hase_simulation_t* sim) { delete sim; }

}  // extern "C"

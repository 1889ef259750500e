// gpu_backend.hpp — the `gpu` backend a reference FluidSolver holds, as a C++
// class over the C-ABI of include/shhd_gpu.h.
//
// It keeps FluidSolver's contract (src/fluid_solver.h:52-107):
//   - construction from the lowered configuration (FluidSolver::FluidSolver,
//     src/fluid_solver.h:52-95; the lowering is include/shhd_reference_lowering.hpp
//     where the reference's Params and plugins are at hand);
//   - step(state, time) advances the caller's State by one time step
//     (FluidSolver::step, src/fluid_solver.h:97-99), the state in the
//     reference layout (src/state.cc:8-38) as interleaved complex doubles;
//   - set_flux_request / flux() (src/fluid_solver.h:101-107, FluxBuffer slot
//     order of src/buffer_flux.cc:10-48);
//   - errors are std::runtime_error carrying the reference's message, as the
//     reference throws them (src/main.cc:39-42 turns them into exit code 1).
// The device keeps the state between calls: it is uploaded at the first step
// and after the host changed it (host_state_changed(), e.g. after a host-side
// initial condition), and downloaded after every step so measures that read
// the host State see the current one; run() advances many steps with one
// call and one download.
#pragma once

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>

#include "shhd_gpu.h"

namespace shhd_host {

class GpuBackend {
public:
    explicit GpuBackend(const shhd_config& config) {
        check(shhd_gpu_create(&config, &h_), nullptr);
        check(shhd_gpu_get_info(h_, &info_), h_);
    }
    ~GpuBackend() { shhd_gpu_destroy(h_); }
    GpuBackend(const GpuBackend&) = delete;
    GpuBackend& operator=(const GpuBackend&) = delete;

    // complex values of a whole state: (N + 3) fields of (Mx/2 + 1) My
    size_t state_size() const { return info_.state_doubles / 2; }
    const shhd_info& info() const { return info_; }

    // FluidSolver::step(State&, double time)
    void step(std::complex<double>* state, double time) { run(state, 1, time); }

    // nsteps steps with one call (the `run N` segment of Project::run_steps,
    // src/project.h:128-145, when no measure is due in between)
    void run(std::complex<double>* state, int nsteps, double time) {
        double* s = reinterpret_cast<double*>(state);
        if (upload_) {
            check(shhd_gpu_set_state(h_, s, info_.state_doubles), h_);
            upload_ = false;
        }
        const shhd_status st = shhd_gpu_step(h_, nsteps, time);
        // a failed step leaves the device state at the failing step's start:
        // the host copy follows it before the error propagates
        check(shhd_gpu_get_state(h_, s, info_.state_doubles), h_);
        check(st, h_);
    }

    // the host State was written outside the backend (an initial condition)
    void host_state_changed() { upload_ = true; }

    // FluidSolver::set_flux_request (src/fluid_solver.h:101-103)
    void set_flux_request(bool density, bool order_parameter, bool momentum) {
        check(shhd_gpu_set_flux_request(h_, density, order_parameter, momentum), h_);
    }
    // FluidSolver::flux_buffer (src/fluid_solver.h:105-107): 2 + 2N + 3 slots of
    // the spectral size each, unrequested slots zero
    size_t flux_size() const { return static_cast<size_t>(2 + 2 * (info_.num_fields - 3) + 3) * spectral_size(); }
    void flux(std::complex<double>* out) const {
        check(shhd_gpu_get_flux(h_, reinterpret_cast<double*>(out), 2 * flux_size()), h_);
    }

private:
    size_t spectral_size() const { return state_size() / static_cast<size_t>(info_.num_fields); }
    static void check(shhd_status st, const shhd_gpu* h) {
        if (st != SHHD_OK) throw std::runtime_error(shhd_gpu_last_error(h));
    }

    shhd_gpu* h_ = nullptr;
    shhd_info info_{};
    bool upload_ = true;
};

}  // namespace shhd_host

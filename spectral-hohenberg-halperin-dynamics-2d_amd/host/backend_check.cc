// backend_check — drives GpuBackend (gpu_backend.hpp) the way a reference
// FluidSolver run does, for the tests that hold it against the Python binding.
//
//   backend_check layout
//       the sizes and field offsets of the C-ABI structs, one "name value" per
//       line (no device use)
//   backend_check run <config.bin> <state.bin> <nsteps> <flux> <out.bin>
//       a shhd_config image and a reference-layout state in, nsteps calls of
//       step() (one step each), then the state and, when flux is 1, the step
//       flux of all slots out; a runtime error prints "error: <message>" and
//       exits with status 1, as the reference's main does (src/main.cc:39-42)
#include <complex>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_backend.hpp"

namespace {

std::vector<char> read_file(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

#define SHHD_SIZE(T) std::printf("%s %zu\n", #T, sizeof(T));
#define SHHD_OFF(T, F) std::printf("%s.%s %zu\n", #T, #F, offsetof(T, F));

void layout() {
    SHHD_SIZE(shhd_config)
    SHHD_OFF(shhd_config, dt)
    SHHD_OFF(shhd_config, precision)
    SHHD_OFF(shhd_config, mobility)
    SHHD_OFF(shhd_config, phi4_u)
    SHHD_OFF(shhd_config, seed)
    SHHD_OFF(shhd_config, sine_forces)
    SHHD_OFF(shhd_config, gradient_forces)
    SHHD_SIZE(shhd_sine_force)
    SHHD_SIZE(shhd_gradient_force)
    SHHD_SIZE(shhd_info)
    SHHD_OFF(shhd_info, state_doubles)
    SHHD_OFF(shhd_info, device_bytes)
    SHHD_OFF(shhd_info, forward_column_lanes)
    SHHD_SIZE(shhd_probe)
    SHHD_SIZE(shhd_initial_condition)
    SHHD_SIZE(shhd_tuning)
    SHHD_OFF(shhd_tuning, row_waves)
    SHHD_OFF(shhd_tuning, ms)
}

int run(int argc, char** argv) {
    if (argc != 7) throw std::runtime_error("usage: backend_check run <config.bin> <state.bin> <nsteps> <flux> <out.bin>");
    const std::vector<char> cfg_bytes = read_file(argv[2]);
    if (cfg_bytes.size() != sizeof(shhd_config)) throw std::runtime_error("config image size differs from shhd_config");
    shhd_config cfg;
    std::memcpy(&cfg, cfg_bytes.data(), sizeof(cfg));
    shhd_host::GpuBackend backend(cfg);
    const std::vector<char> st_bytes = read_file(argv[3]);
    std::vector<std::complex<double>> state(backend.state_size());
    if (st_bytes.size() != state.size() * sizeof(std::complex<double>)) throw std::runtime_error("state size differs");
    std::memcpy(state.data(), st_bytes.data(), st_bytes.size());
    const int nsteps = std::atoi(argv[4]);
    const bool flux = std::atoi(argv[5]) != 0;
    if (flux) backend.set_flux_request(true, true, true);
    const double dt = cfg.dt;
    for (int n = 0; n < nsteps; ++n) backend.step(state.data(), n * dt);
    std::ofstream out(argv[6], std::ios::binary);
    out.write(reinterpret_cast<const char*>(state.data()), static_cast<std::streamsize>(state.size() * sizeof(state[0])));
    if (flux) {
        std::vector<std::complex<double>> fl(backend.flux_size());
        backend.flux(fl.data());
        out.write(reinterpret_cast<const char*>(fl.data()), static_cast<std::streamsize>(fl.size() * sizeof(fl[0])));
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "layout") {
            layout();
            return 0;
        }
        if (argc >= 2 && std::string(argv[1]) == "run") return run(argc, argv);
        throw std::runtime_error("usage: backend_check layout | run ...");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

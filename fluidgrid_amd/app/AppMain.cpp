// AppMain.cpp -- the reference application's simulation set-up and loop
// (AppMain.cpp:43-62 and 78-133) as a C++ host program over this library:
// the parameter blocks go to ApplicationData, a BlockEmitter fills a
// GPUSPHFluid, HeterogeneousSimulator::doTimeStep advances it, exactly the
// calls the reference makes, without the window, camera and renderer.
//
//   fg_appmain <output file> [steps]
//
// The output file holds the parameter blocks and the initial and final
// particle state (layout below), so a checker can replay the same scene.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fluidgrid/HeterogeneousSimulator.h"

using namespace Simulation;

namespace {

// File layout: "FGAPP1\0\0", int32 steps, int32 n, CommonFluidParams (104 B),
// FluidParams (28 B), KernelsCache (12 B), then float4[n] initial
// positions, initial velocities, final positions, final velocities.
bool write_output(const char* path, int steps, const Application::CommonFluidParams& cp,
                  const Application::FluidParams& fp, const Application::KernelsCache& kc,
                  const std::vector<float4>* arrays) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const char magic[8] = {'F', 'G', 'A', 'P', 'P', '1', 0, 0};
    const int n = fp.max_particles;
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&steps, 4, 1, f) == 1 && std::fwrite(&n, 4, 1, f) == 1 &&
              std::fwrite(&cp, sizeof(cp), 1, f) == 1 && std::fwrite(&fp, sizeof(fp), 1, f) == 1 &&
              std::fwrite(&kc, sizeof(kc), 1, f) == 1;
    for (int k = 0; ok && k < 4; k++) ok = std::fwrite(arrays[k].data(), sizeof(float4), n, f) == (size_t)n;
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <output file> [steps]\n", argv[0]);
        return 2;
    }
    const int steps = argc > 2 ? std::atoi(argv[2]) : 100;

    // AppMain.cpp:84-115
    Application::CommonFluidParams cp{};
    cp.volume_min = make_float3(-20.0f, -20.0f, -20.0f);
    cp.volume_max = make_float3(20.0f, 85.0f, 40.0f);
    cp.global_external_force = make_float3(0, -9.8f, 0);
    cp.simulation_scale = 0.009f;
    cp.deltaTimeIntegration = 0.004f;
    cp.bucket_size = make_int3(64, 64, 64);
    cp.keyBits = 24;
    cp.velocity_limit = 600;
    cp.pointScale = (float)(800 / tanf(90.0 * 0.5f * 3.1415f / 180.0f));
    cp.fluid_to_srb_relation = make_int3(1, 1, 1);

    Application::FluidParams fp{};
    fp.max_particles = 24000;
    fp.mass = ((128 * 1024.0f) / fp.max_particles) * 0.0002f;
    fp.restDensity = 1000.0f;
    fp.stiffness = 0.2f;
    fp.viscosity = 1.0f;
    fp.gasConstant = 1.5;
    fp.particle_radius = 0.3f;

    const float dist = std::pow(fp.mass / fp.restDensity, 1.0f / 3.0f);
    const float spacing = dist * 0.87f / cp.simulation_scale;
    cp.kernel_radius = 2.0f * dist * 0.87f;
    cp.kernel_radius_squared = cp.kernel_radius * cp.kernel_radius;
    cp.pointRadius = spacing;
    const float pv = cp.kernel_radius / cp.simulation_scale;
    cp.cell_size = make_float3(pv, pv, pv);
    Application::ApplicationData::instance()->setCommonFluidParams(cp);

    // The reference hashes cells with HashKeyD (K_DeviceFunctions.cu:17-26).
    if (fg_set_key_mode(FG_KEY_REF_XOR) != FG_OK) return 3;

    Emitter::BlockEmitter emitter(make_float3(-20.0f, -20.0f, -20.0f), make_float3(20.0f, 85.0f, 0.0f), spacing);
    Fluid::GPUSPHFluid fluid(fp, &emitter);
    Simulators::HeterogeneousSimulator* sim = Simulators::HeterogeneousSimulator::instance();
    sim->addSphFluid(&fluid);

    std::vector<float4> state[4];
    for (auto& a : state) a.resize(fp.max_particles);
    fluid.downloadState(state[0].data(), state[1].data());
    for (int s = 0; s < steps; s++) sim->doTimeStep();  // AppMain::OnUpdate, one step per frame
    fluid.downloadState(state[2].data(), state[3].data());
    if (fg_peek_error() != FG_OK) {
        std::fprintf(stderr, "fg_appmain: %s\n", fg_last_error_string());
        return 4;
    }
    const Application::KernelsCache kc = Application::ApplicationData::instance()->getKernelsCache();
    Simulators::HeterogeneousSimulator::destroy();
    if (!write_output(argv[1], steps, cp, fp, kc, state)) {
        std::fprintf(stderr, "fg_appmain: cannot write %s\n", argv[1]);
        return 5;
    }
    std::printf("fg_appmain: %d particles, %d steps -> %s\n", fp.max_particles, steps, argv[1]);
    return 0;
}

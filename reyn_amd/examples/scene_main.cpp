// The reference's main loop (src/main.cu:44-78) without its GUI, on the C++
// host API: a scene from an OBJ file whose shape "fluid" is filled with fluid
// inside the other shapes (`Scene::from_obj`), an IISPH solver with the B3
// kernel, and per step the time-step controller, the grid with the reorder of
// the state and the warm-start pressure, the solver step and the bounds.
//
//   scene_main <file.obj> [N=100000] [steps=100] [out]
//
// With `out`, the final positions go to out.ply and the step trace to out.csv.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "reyn/reyn.hpp"

using namespace reyn;

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.obj> [N=100000] [steps=100] [out]\n", argv[0]);
        return 2;
    }
    const uint32_t N_desired = argc > 2 ? uint32_t(std::atoi(argv[2])) : 100000u;
    const int steps = argc > 3 ? std::atoi(argv[3]) : 100;
    const char* out = argc > 4 ? argv[4] : nullptr;
    try {
        // as the reference: state and buffers first, the scene fills them
        Particles state(1, 1000.f);
        DeviceBuffer<float> tmp1(1), tmp2(1), tmp3(1), tmp4(1), tmp5(1), tmp6(1), prs(1);
        DeviceBuffer<uint32_t> tmp_u32(1);
        Scene scene { Scene::from_obj(argv[1], N_desired, 1., state, tmp1, 3., 1.) };

        const B3 W(2.f * scene.h);
        auto solver { IISPH<B3, Resort::yes>(
            W, scene.N, 0.001f, scene.h, scene.ρ₀, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6, prs, tmp_u32) };
        std::printf("scene: N=%u, Nbdy=%zu, h=%g\n", scene.N, scene.bdy.xs.size(), scene.h);

        double time = 0.0;
        uint32_t iters = 0;
        float dt = 0.0005f;
        StepTrace trace;
        for (int s = 0; s < steps; ++s) {
            dt = simple_dt_controller(iters, dt, 1.0f, scene.h, state, v3(0.f, -9.81f, 0.f), 10, 0.001f);
            const auto grid { scene.get_grid(state, tmp1, 2.f, &prs) };
            iters = solver.step(state, grid, scene.bdy, dt);
            scene.hard_enforce_bounds(state);
            trace.record(dt, iters);
            time += dt;
            if (s % 10 == 0 || s + 1 == steps)
                std::printf("step %4d  t %.4f  dt %.3e  iters %u  mean rho %.3f\n", s, time, dt, iters, tmp1.avg());
        }
        if (out) {
            trace.write_csv(std::string(out) + ".csv");
            save_to_ply(std::string(out) + ".ply", state.xx, state.xy, state.xz);
            std::printf("wrote %s.csv and %s.ply\n", out, out);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// Headless dam break driven through the C++ host API, in the order of the
// reference's main loop (src/main.cu:59-78): the time-step controller (dt
// starts at 0.5 ms; lambda 1, 10 target iterations, 0.1 % change per step),
// grid + reorder of the state and the warm-start pressure, IISPH step, bounds.
//
//   dambreak [side=32] [steps=50] [out]
//
// With `out`, the time-step / iteration trace goes to out.csv and the final
// positions to out.ply.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "reyn/reyn.hpp"

using namespace reyn;

int main(int argc, char** argv)
{
    const int side = argc > 1 ? std::atoi(argv[1]) : 32;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 50;
    const char* out = argc > 3 ? argv[3] : nullptr;
    const float h = 1.0f / side, rho_0 = 1000.f;
    const float3 box = v3(2.f, 1.25f, 2.f);

    // fluid block [0,1]^3 at (i + 1/2) h, walls of 2 layers around the box
    const size_t N = size_t(side) * side * side;
    Particles state(N, rho_0);
    std::vector<float> x(N), y(N), z(N);
    for (int k = 0, i = 0; k < side; ++k)
        for (int j = 0; j < side; ++j)
            for (int l = 0; l < side; ++l, ++i) {
                x[i] = (l + 0.5f) * h;
                y[i] = (j + 0.5f) * h;
                z[i] = (k + 0.5f) * h;
            }
    state.xx.upload(x);
    state.xy.upload(y);
    state.xz.upload(z);
    state.m.fill(rho_0 * h * h * h);
    std::vector<float3> wall;
    const int nx = int(std::lround(box.x / h)) + 4, ny = int(std::lround(box.y / h)) + 4,
              nz = int(std::lround(box.z / h)) + 4;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int l = 0; l < nx; ++l) {
                const float3 p = v3((l - 2 + 0.5f) * h, (j - 2 + 0.5f) * h, (k - 2 + 0.5f) * h);
                const bool inside = p.x > 0 && p.x < box.x && p.y > 0 && p.y < box.y && p.z > 0 && p.z < box.z;
                if (!inside)
                    wall.push_back(p);
            }
    BoundarySamples bdy(wall, h, rho_0);

    DeviceBuffer<float> tmp(N), rho(1), aii(1), si(1), ax(1), ay(1), az(1), prs(1);
    DeviceBuffer<uint32_t> bits(1);
    const B3 W(2.f * h);
    IISPH<B3, Resort::yes> solver(W, uint32_t(N), 0.001f, h, rho_0, rho, aii, si, ax, ay, az, prs, bits);
    UniformGridBuilder builder(bdy.bound_min, bdy.bound_max, 2.f * h);

    float dt = 0.0005f;
    uint32_t iters = 0;
    double t = 0.0;
    StepTrace trace;
    for (int s = 0; s < steps; ++s) {
        dt = simple_dt_controller(iters, dt, 1.f, h, state, v3(0.f, -9.81f, 0.f), 10, 0.001f);
        const auto grid = builder.construct_and_reorder(2.f * h, tmp, state, &prs);
        iters = solver.step(state, grid, bdy, dt);
        hard_enforce_bounds(bdy.bound_min, bdy.bound_max, state);
        trace.record(dt, iters);
        t += dt;
        if (s % 10 == 0 || s + 1 == steps)
            std::printf("step %4d  t %.4f  dt %.3e  iters %u  mean rho %.3f\n", s, t, dt, iters, rho.avg());
    }
    if (out) { // the dt / iteration trace and the final positions
        trace.write_csv(std::string(out) + ".csv");
        save_to_ply(std::string(out) + ".ply", state.xx, state.xy, state.xz);
        std::printf("wrote %s.csv and %s.ply\n", out, out);
    }
    return 0;
}

// Tests of the C++ host API (include/reyn/reyn.hpp) on the GPU, written in
// the style of the reference's doctest cases. Run by tests/test_gpu_cpp_api.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <algorithm>
#include <functional>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "reyn/reyn.hpp"

using namespace reyn;

static int g_failed = 0;
static int g_checks = 0;

#define CHECK(cond)                                                                                                 \
    do {                                                                                                            \
        ++g_checks;                                                                                                 \
        if (!(cond)) {                                                                                              \
            ++g_failed;                                                                                             \
            std::printf("  CHECK failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__);                                 \
        }                                                                                                           \
    } while (0)

static void run(const char* name, const std::function<void()>& f)
{
    const int before = g_failed;
    std::printf("TEST_CASE %s\n", name);
    std::fflush(stdout);
    try {
        f();
    } catch (const std::exception& e) {
        ++g_failed;
        std::printf("  exception: %s\n", e.what());
    }
    std::printf("  %s\n", g_failed == before ? "ok" : "FAILED");
    std::fflush(stdout);
}

struct Lattice {
    float h;
    std::vector<float> x, y, z;
    std::vector<float3> wall;
    float3 box;
};

static Lattice dam_break(int side, float h, float3 box)
{
    Lattice s { h, {}, {}, {}, {}, box };
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> jit(-0.01f * h, 0.01f * h);
    for (int k = 0; k < side; ++k)
        for (int j = 0; j < side; ++j)
            for (int i = 0; i < side; ++i) {
                s.x.push_back((i + 0.5f) * h + jit(rng));
                s.y.push_back((j + 0.5f) * h + jit(rng));
                s.z.push_back((k + 0.5f) * h + jit(rng));
            }
    const int nx = int(std::lround(box.x / h)) + 4, ny = int(std::lround(box.y / h)) + 4,
              nz = int(std::lround(box.z / h)) + 4;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const float3 p = v3((i - 2 + 0.5f) * h, (j - 2 + 0.5f) * h, (k - 2 + 0.5f) * h);
                if (!(p.x > 0 && p.x < box.x && p.y > 0 && p.y < box.y && p.z > 0 && p.z < box.z))
                    s.wall.push_back(p);
            }
    return s;
}

static void upload(Particles& st, const Lattice& s, float rho_0)
{
    st.xx.upload(s.x);
    st.xy.upload(s.y);
    st.xz.upload(s.z);
    st.m.fill(rho_0 * s.h * s.h * s.h);
}

static bool all_finite(const DeviceBuffer<float>& b)
{
    for (float v : b.download())
        if (!std::isfinite(v))
            return false;
    return true;
}

int main()
{
    run("DeviceBuffer upload, resize, reductions", [] {
        std::vector<float> h(1000);
        std::iota(h.begin(), h.end(), 1.f);
        DeviceBuffer<float> b(10);
        b.upload(h);
        CHECK(b.size() == 1000);
        CHECK(b.sum() == 500500.0);
        CHECK(b.min() == 1.0 && b.max() == 1000.0);
        CHECK(b.avg() == 500.5f);
        b.resize(10);
        const auto d = b.download();
        CHECK(d.size() == 10 && d[9] == 10.f);
        DeviceBuffer<uint32_t> u(5, 7u);
        CHECK(u.download()[4] == 7u);
    });

    run("UniformGridBuilder rejects an inverted box", [] {
        bool threw = false;
        try {
            UniformGridBuilder b(v3(1.f), v3(0.f), 0.1f);
        } catch (const std::runtime_error& e) {
            threw = std::string(e.what()).find("Negative number of grid cells") != std::string::npos;
        }
        CHECK(threw);
    });

    run("construct_and_reorder leaves the state in cell order", [] {
        const float rho_0 = 1000.f;
        Lattice s = dam_break(12, 0.05f, v3(1.f, 1.f, 1.f));
        // shuffle the input order
        std::vector<size_t> perm(s.x.size());
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), std::mt19937(3));
        Lattice t = s;
        for (size_t i = 0; i < perm.size(); ++i) {
            t.x[i] = s.x[perm[i]];
            t.y[i] = s.y[perm[i]];
            t.z[i] = s.z[perm[i]];
        }
        Particles st(t.x.size(), rho_0);
        upload(st, t, rho_0);
        DeviceBuffer<float> tmp(1);
        UniformGridBuilder b(v3(0.f), v3(1.f), 0.1f);
        const auto g1 = b.construct_and_reorder(0.1f, tmp, st);
        CHECK(g1.grid.sorted == nullptr && g1.grid.n == st.size());
        // rebuilding the reordered state gives the identity permutation
        const auto g2 = b.construct(0.1f, st.xx, st.xy, st.xz);
        std::vector<uint32_t> sorted(st.size());
        reyn_memcpy(Context::current().get(), sorted.data(), g2.grid.sorted, sorted.size() * 4, 1);
        bool identity = true;
        for (uint32_t i = 0; i < sorted.size(); ++i)
            identity = identity && sorted[i] == i;
        CHECK(identity);
        // the positions are the same multiset
        auto xs = st.xx.download();
        std::sort(xs.begin(), xs.end());
        auto ref = t.x;
        std::sort(ref.begin(), ref.end());
        CHECK(xs == ref);
    });

    run("UniformGridBuilder with deferred build checks", [] {
        const float rho_0 = 1000.f;
        Lattice s = dam_break(10, 0.05f, v3(1.f, 1.f, 1.f));
        Particles a(s.x.size(), rho_0), c(s.x.size(), rho_0);
        upload(a, s, rho_0);
        upload(c, s, rho_0);
        DeviceBuffer<float> ta(1), tc(1);
        UniformGridBuilder ba(v3(0.f), v3(1.f), 0.1f), bc(v3(0.f), v3(1.f), 0.1f);
        bc.defer_check();
        ba.construct_and_reorder(0.1f, ta, a);
        bc.construct_and_reorder(0.1f, tc, c);
        bc.check();
        CHECK(a.xx.download() == c.xx.download() && a.vy.download() == c.vy.download());
        // a point outside the grid: the deferred construct returns, the check throws
        s.x[5] = 50.f;
        Particles bad(s.x.size(), rho_0);
        upload(bad, s, rho_0);
        bc.construct(0.1f, bad.xx, bad.xy, bad.xz);
        bool threw = false;
        try {
            bc.check();
        } catch (const std::runtime_error& e) {
            threw = std::string(e.what()).find("outside the grid") != std::string::npos;
        }
        CHECK(threw);
    });

    run("IISPH dam break through the host API", [] {
        const float rho_0 = 1000.f, h = 1.f / 24;
        Lattice s = dam_break(16, h, v3(1.5f, 1.f, 1.f));
        const size_t N = s.x.size();
        Particles state(N, rho_0);
        upload(state, s, rho_0);
        BoundarySamples bdy(s.wall, h, rho_0);
        DeviceBuffer<float> tmp(N), rho(1), aii(1), si(1), ax(1), ay(1), az(1), prs(1);
        DeviceBuffer<uint32_t> bits(1);
        IISPH<B3, Resort::yes> solver(B3(2.f * h), uint32_t(N), 0.001f, h, rho_0, rho, aii, si, ax, ay, az, prs,
            bits);
        UniformGridBuilder builder(bdy.bound_min, bdy.bound_max, 2.f * h);
        const double m0 = state.m.sum();
        unsigned it_min = 1000;
        for (int step = 0; step < 20; ++step) {
            const float dt = cfl_time_step(0.4f, h, state, v3(0.f, -9.81f, 0.f));
            CHECK(dt > 0.f && std::isfinite(dt));
            const auto grid = builder.construct_and_reorder(2.f * h, tmp, state, &prs);
            it_min = std::min(it_min, solver.step(state, grid, bdy, dt));
            hard_enforce_bounds(bdy.bound_min, bdy.bound_max, state);
        }
        CHECK(it_min >= 3);
        CHECK(solver.info().n_violating == 0);
        CHECK(state.m.sum() == m0);
        CHECK(all_finite(state.xx) && all_finite(state.vy) && all_finite(prs));
        CHECK(std::fabs(rho.avg() / rho_0 - 1.f) < 0.05f);
        CHECK(state.xy.min() >= bdy.bound_min.y && state.xy.max() <= bdy.bound_max.y);
    });

    run("SESPH and PCISPH steps through the host API", [] {
        const float rho_0 = 1000.f, h = 0.05f;
        Lattice s = dam_break(12, h, v3(1.f, 1.f, 1.f));
        const size_t N = s.x.size();
        BoundarySamples bdy(s.wall, h, rho_0);
        UniformGridBuilder builder(bdy.bound_min, bdy.bound_max, 2.f * h, 1.0f);
        DeviceBuffer<float> tmp(N), rho(1), ax(1), ay(1), az(1);
        {
            Particles state(N, rho_0);
            upload(state, s, rho_0);
            SESPH<C2, Resort::yes> sesph(C2(2.f * h), uint32_t(N), 0.001f, h, rho_0, rho, ax, ay, az,
                v3(0.f, -9.81f, 0.f), 1000.f);
            for (int step = 0; step < 5; ++step) {
                const auto grid = builder.construct_and_reorder(2.f * h, tmp, state);
                sesph.step(state, grid, bdy, 1e-3f);
                hard_enforce_bounds(bdy.bound_min, bdy.bound_max, state);
            }
            CHECK(all_finite(state.xy) && all_finite(state.vy));
            CHECK(state.vy.avg() < 0.f);
        }
        {
            Particles state(N, rho_0);
            upload(state, s, rho_0);
            PCISPH<W6, Resort::no> pcisph(W6(2.f * h), uint32_t(N), 0.001f, h, rho_0, rho, ax, ay, az);
            unsigned iters = 0;
            for (int step = 0; step < 5; ++step) {
                const auto grid = builder.construct(2.f * h, state.xx, state.xy, state.xz);
                iters = pcisph.step(state, grid, bdy, 1e-3f);
                hard_enforce_bounds(bdy.bound_min, bdy.bound_max, state);
            }
            CHECK(iters >= 3);
            CHECK(all_finite(state.xy) && all_finite(rho));
        }
    });

    run("SceneBatch: scenes alone and in a batch that they join and leave", [] {
        Lattice a = dam_break(10, 0.05f, v3(1.f, 1.f, 1.f));
        Lattice b = dam_break(8, 0.06f, v3(1.2f, 1.f, 0.9f));
        std::vector<float> abx, aby, abz, bbx, bby, bbz;
        for (const float3& w : a.wall) {
            abx.push_back(w.x);
            aby.push_back(w.y);
            abz.push_back(w.z);
        }
        for (const float3& w : b.wall) {
            bbx.push_back(w.x);
            bby.push_back(w.y);
            bbz.push_back(w.z);
        }
        auto desc = [](const Lattice& s, const std::vector<float>& bx, const std::vector<float>& by,
                        const std::vector<float>& bz) {
            reyn_batch_scene d {};
            d.n = uint32_t(s.x.size());
            d.x = s.x.data();
            d.y = s.y.data();
            d.z = s.z.data();
            d.mass = 1000.f * s.h * s.h * s.h;
            d.n_bdy = uint32_t(bx.size());
            d.bx = bx.data();
            d.by = by.data();
            d.bz = bz.data();
            d.bound_min = v3(*std::min_element(bx.begin(), bx.end()), *std::min_element(by.begin(), by.end()),
                *std::min_element(bz.begin(), bz.end()));
            d.bound_max = v3(*std::max_element(bx.begin(), bx.end()), *std::max_element(by.begin(), by.end()),
                *std::max_element(bz.begin(), bz.end()));
            d.h = s.h;
            return d;
        };
        reyn_batch_params prm {};
        prm.solver = REYN_SOLVER_IISPH;
        prm.kernel = REYN_KERNEL_B3;
        prm.nu = 0.001f;
        prm.rho_0 = 1000.f;
        prm.g = v3(0.f, -9.81f, 0.f);
        prm.cell_ratio = 0.5f;
        prm.dt0 = 5e-4f;
        prm.cfl_lambda = 1.f;
        prm.change_factor = 0.001f;
        prm.iter_target = 10;
        prm.eta_rho_max = 0.001f;
        prm.omega = 0.5f;
        prm.min_iter = 3;
        prm.max_iter = 100;
        prm.stop_rule = REYN_IISPH_STOP_MEAN;
        prm.warm_start = 1;
        const reyn_batch_scene da = desc(a, abx, aby, abz), db = desc(b, bbx, bby, bbz);
        // a runs 6 steps; b joins after 2 and leaves after 3 of its own
        SceneBatch mix(prm, { da });
        mix.step(2);
        const uint32_t ib = mix.add({ db })[0];
        mix.step(3);
        const BatchSceneState gb = mix.download(ib);
        CHECK(mix.info(ib).steps == 3);
        mix.remove({ ib });
        mix.step(1);
        const BatchSceneState ga = mix.download(0);
        CHECK(mix.info(0).steps == 6);
        SceneBatch alone_a(prm, { da }), alone_b(prm, { db });
        alone_a.step(6);
        alone_b.step(3);
        const BatchSceneState ra = alone_a.download(0), rb = alone_b.download(0);
        CHECK(ga.x == ra.x && ga.y == ra.y && ga.z == ra.z && ga.vx == ra.vx && ga.p == ra.p && ga.rho == ra.rho);
        CHECK(gb.x == rb.x && gb.vy == rb.vy && gb.p == rb.p && gb.rho == rb.rho);
        CHECK(mix.info(0).iterations == alone_a.info(0).iterations && mix.info(0).dt == alone_a.info(0).dt);
    });

    run("SceneBatch: a dam break generated on the device steps like its host twin", [] {
        reyn_dam_break d {};
        d.counts[0] = 12;
        d.counts[1] = 10;
        d.counts[2] = 8;
        d.spacing = 0.05;
        d.box[0] = 1.2;
        d.box[1] = 0.8;
        d.box[2] = 0.6;
        d.jitter = 0.01;
        d.seed = 9;
        d.layers = 2;
        uint32_t nl = 0, nb = 0;
        CHECK(reyn_dam_break_sizes(&d, &nl, &nb) == REYN_OK);
        CHECK(nl == 960u && nb > 0u);
        reyn_batch_params prm {};
        prm.solver = REYN_SOLVER_PCISPH;
        prm.kernel = REYN_KERNEL_W6;
        prm.nu = 0.001f;
        prm.rho_0 = 1000.f;
        prm.g = v3(0.f, -9.81f, 0.f);
        prm.cell_ratio = 1.f;
        prm.dt0 = 5e-4f;
        prm.cfl_lambda = 1.f;
        prm.change_factor = 0.001f;
        prm.iter_target = 10;
        prm.eta_rho_max = 0.001f;
        prm.min_iter = 3;
        prm.max_iter = 100;
        SceneBatch b(prm);
        const uint32_t id = b.add_dam_breaks({ d })[0];
        b.step(4);
        const BatchSceneState st = b.download(id);
        CHECK(st.x.size() == nl);
        CHECK(b.info(id).steps == 4 && b.info(id).iterations >= 3);
        bool finite = true;
        for (float v : st.y)
            finite = finite && std::isfinite(v);
        CHECK(finite);
    });

    run("Mesh loads the reference's cube and samples it as a boundary", [] {
        const char* root = std::getenv("REYN_TEST_GOLDEN");
        const std::string dir = root ? root : "tests/golden";
        const Mesh cube = Mesh::load_obj(dir + "/cube.obj");
        CHECK(cube.face_count() == 12u);
        CHECK(cube.get_volume() == 8.0);
        BoundarySamples bdy = sample_mesh(cube, 0.1f, 1000.f);
        CHECK(bdy.xs.size() > 0);
        CHECK(bdy.bound_min.x >= -1.f && bdy.bound_max.x <= 1.f);
        CHECK(bdy.m.min() > 0.0);
        // relaxed samples stay on the cube's surface and spread more evenly
        float r1 = 0.f, r8 = 0.f;
        BoundarySamples one = sample_mesh(cube, 0.1f, 1000.f, 1.f, 1614201312u, Context::current(), 1, 1.f, &r1);
        BoundarySamples relaxed = sample_mesh(cube, 0.1f, 1000.f, 1.f, 1614201312u, Context::current(), 8, 1.f, &r8);
        CHECK(relaxed.xs.size() == bdy.xs.size());
        CHECK(r8 < r1);
        CHECK(relaxed.bound_min.x >= -1.f && relaxed.bound_max.x <= 1.f);
        CHECK(relaxed.bound_min.y >= -1.f && relaxed.bound_max.y <= 1.f);
        CHECK(relaxed.m.min() > 0.0);
        // the LBVH: projection to the cube's faces and the inside test
        const LBVH bvh(cube);
        CHECK(bvh.height() >= 4u && bvh.height() <= 12u);
        DeviceBuffer<float> px(3), py(3), pz(3);
        px.upload(std::vector<float> { 2.f, 0.25f, -0.5f });
        py.upload(std::vector<float> { 0.5f, 0.1f, 0.f });
        pz.upload(std::vector<float> { -0.5f, 0.2f, 3.f });
        DeviceBuffer<uint8_t> in(3);
        bvh.inside(px, py, pz, in);
        const std::vector<uint8_t> inv = in.download();
        CHECK(inv[0] == 0 && inv[1] == 1 && inv[2] == 0);
        bvh.project(px, py, pz);
        const std::vector<float> qx = px.download(), qy = py.download(), qz = pz.download();
        CHECK(qx[0] == 1.f && qy[0] == 0.5f && qz[0] == -0.5f);
        CHECK(qx[1] == 1.f && qy[1] == 0.1f && qz[1] == 0.2f);
        CHECK(qx[2] == -0.5f && qy[2] == 0.f && qz[2] == 1.f);
    });

    run("Scene: a fluid box in a mesh boundary, and a fluid volume from an OBJ shape, stepped by the main loop", [] {
        const char* root = std::getenv("REYN_TEST_GOLDEN");
        const std::string dir = root ? root : "tests/golden";
        // box scene: the cube's surface as the boundary, fluid in [-0.9, 0.1]^3
        Particles state(1, 1000.f);
        DeviceBuffer<float> tmp(1);
        Scene box(dir + "/cube.obj", 8000, v3(-0.9f), v3(0.1f), 1000.f, state, tmp, 2.f, 1.f, 0.01f, 5);
        CHECK(std::fabs(box.h - 0.05f) < 1e-6f);
        CHECK(box.N == state.size() && (box.N == 8000u || box.N == 6859u));
        CHECK(box.bound_min.x >= -1.f && box.bound_max.x <= 1.f);
        const std::vector<float> x = state.xx.download(), m = state.m.download(), vx = state.vx.download();
        CHECK(*std::min_element(x.begin(), x.end()) > -0.91f && *std::max_element(x.begin(), x.end()) < 0.11f);
        CHECK(m[0] == 1000.f * box.h * box.h * box.h && vx[0] == 0.f);
        // the reference's main loop on it (src/main.cu:59-78)
        DeviceBuffer<float> rho(1), aii(1), si(1), ax(1), ay(1), az(1), prs(1);
        DeviceBuffer<uint32_t> bits(1);
        IISPH<B3, Resort::yes> solver(B3(2.f * box.h), box.N, 0.001f, box.h, box.ρ₀, rho, aii, si, ax, ay, az, prs,
            bits);
        float dt = 0.0005f;
        uint32_t iters = 0;
        for (int s = 0; s < 3; ++s) {
            dt = simple_dt_controller(iters, dt, 1.f, box.h, state, v3(0.f, -9.81f, 0.f), 10, 0.001f);
            const auto grid = box.get_grid(state, tmp, 2.f, &prs);
            iters = solver.step(state, grid, box.bdy, dt);
            box.hard_enforce_bounds(state);
        }
        CHECK(iters > 0 && std::isfinite(rho.avg()) && rho.avg() > 500.0);
        // fluid volume: a cube named fluid, [-0.5, 0.5]^3, inside the cube [-1, 1]^3
        const std::string obj = "/tmp/reyn_scene_test.obj";
        {
            std::ifstream in(dir + "/cube.obj");
            std::vector<std::string> v, f;
            for (std::string line; std::getline(in, line);) {
                if (line.rfind("v ", 0) == 0)
                    v.push_back(line);
                if (line.rfind("f ", 0) == 0)
                    f.push_back(line);
            }
            std::ofstream out(obj);
            out << "o box\n";
            for (const auto& l : v)
                out << l << "\n";
            for (const auto& l : f)
                out << l << "\n";
            out << "o fluid\n";
            for (const auto& l : v) {
                std::istringstream is(l.substr(2));
                double a, b, c;
                is >> a >> b >> c;
                out << "v " << 0.5 * a << " " << 0.5 * b << " " << 0.5 * c << "\n";
            }
            for (const auto& l : f) {
                std::istringstream is(l.substr(2));
                out << "f";
                for (int k; is >> k;)
                    out << " " << k + 8;
                out << "\n";
            }
        }
        Particles st2(1, 1000.f);
        Scene vol = Scene::from_obj(obj, 8000, 1000.f, st2, tmp, 2.f, 1.f, 0.f, 5);
        CHECK(std::fabs(vol.h - 0.05f) < 1e-6f);
        // lattice points on the fluid's faces count as outside: 19 per axis remain
        CHECK(vol.N == 6859u && st2.size() == 6859u);
        const std::vector<float> y = st2.xy.download();
        CHECK(*std::min_element(y.begin(), y.end()) > -0.5f && *std::max_element(y.begin(), y.end()) < 0.5f);
        std::remove(obj.c_str());
    });

    run("save_to_ply writes the reference's binary PLY layout", [] {
        std::vector<float> x { 1.f, 2.f, 3.f }, y { -1.f, 0.5f, 7.f }, z { 0.f, 1e-3f, 2.f };
        DeviceBuffer<float> xs(1), ys(1), zs(1);
        xs.upload(x);
        ys.upload(y);
        zs.upload(z);
        const std::filesystem::path path = std::filesystem::temp_directory_path() / "reyn_amd_test.ply";
        save_to_ply(path, xs, ys, zs);
        std::ifstream in(path, std::ios::binary);
        const std::string all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        const size_t end = all.find("end_header\n");
        CHECK(all.rfind("ply\nformat binary_", 0) == 0);
        CHECK(all.find("element vertex 3\n") != std::string::npos);
        CHECK(end != std::string::npos && all.size() == end + 11 + 9 * sizeof(float));
        float back[9];
        std::memcpy(back, all.data() + end + 11, sizeof(back));
        CHECK(back[0] == 1.f && back[1] == -1.f && back[2] == 0.f && back[7] == 7.f && back[8] == 2.f);
        std::filesystem::remove(path);
    });

    std::printf("%d checks, %d failed\n", g_checks, g_failed);
    return g_failed == 0 ? 0 : 1;
}

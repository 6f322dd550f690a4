// Host-code check of libvoronoid_hip's C ABI, built with AddressSanitizer and UndefinedBehavior-
// Sanitizer on the host side (tests/native/build.py) and run by tests/test_host_sanitizers.py.
//
//   host_check            entry points that need no GPU: argument checks and their error paths,
//                         the grid rules, the host shell-order build, the static entry's limits,
//                         and where no GPU is visible, handle creation failing (its cleanup)
//   host_check --device   full handle lifecycles on the current GPU: a tessellation with built-in
//                         and caller-evaluated walls through calculate, sealed calculate, prune,
//                         relax, single-generator moves, random generators and every getter; a
//                         batch; a server with tessellation and Lloyd jobs; the static entry twice
//
// Every call's return code is checked; any sanitizer report or failed check ends the program
// with a non-zero status.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include <hip/hip_runtime.h>

#include "vd.h"

#define CHECK(cond)                                                                               \
    do {                                                                                          \
        if (!(cond)) {                                                                            \
            std::fprintf(stderr, "check failed at %s:%d: %s (last error: %s)\n", __FILE__,         \
                         __LINE__, #cond, vd_last_error());                                       \
            std::exit(1);                                                                         \
        }                                                                                         \
    } while (0)
#define OK(call) CHECK((call) == VD_OK)

namespace {

const double kLo[3] = {0.0, 0.0, 0.0};
const double kHi[3] = {1.0, 1.0, 1.0};

std::vector<double> uniform_points(size_t n, uint32_t seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> p(3 * n);
    for (double &x : p) x = u(rng);
    return p;
}

std::vector<double> clustered_points(size_t n, uint32_t seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<double> g(0.0, 0.05);
    std::vector<double> c(24);
    for (double &x : c) x = u(rng);
    std::vector<double> p(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            double x;
            do x = c[3 * (i % 8) + a] + g(rng); while (x < 0.0 || x > 1.0);
            p[3 * i + a] = x;
        }
    return p;
}

// A caller-evaluated wall: the half-space x + y + z <= 2 with its one cut plane.
int half_contains(void *, const double p[3]) { return p[0] + p[1] + p[2] <= 2.0; }
int half_cut(void *, const double g[3], double *planes, int cap) {
    (void)g;
    if (cap >= 1) {
        const double s = 1.0 / std::sqrt(3.0);
        const double pl[6] = {2.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, s, s, s};
        std::memcpy(planes, pl, sizeof(pl));
    }
    return 1;
}

void host_checks() {
    // argument checks: every one returns before any device call
    vd_tess *t = nullptr;
    CHECK(vd_tess_create(kLo, kHi, 0, 4, 4, &t) == VD_EINVAL && t == nullptr);
    CHECK(std::strstr(vd_last_error(), "grid") != nullptr);
    CHECK(vd_tess_create(kLo, kHi, 20000, 4, 4, &t) == VD_EINVAL);
    CHECK(vd_tess_create(nullptr, kHi, 4, 4, 4, &t) == VD_EINVAL);
    CHECK(vd_tess_create(kLo, kHi, 4, 4, 4, nullptr) == VD_EINVAL);
    CHECK(vd_tess_add_wall_plane(nullptr, -1000, kLo, kHi) == VD_EINVAL);
    CHECK(vd_tess_set_generators(nullptr, nullptr, 0) == VD_EINVAL);
    CHECK(vd_tess_set_generator(nullptr, 0, kLo) == VD_EINVAL);
    CHECK(vd_tess_calculate(nullptr) == VD_EINVAL);
    CHECK(vd_tess_count_cells(nullptr) == 0);
    vd_tess_destroy(nullptr);
    CHECK(vd_batch_create(nullptr) == VD_EINVAL);
    vd_batch_destroy(nullptr);
    vd_server *s = nullptr;
    CHECK(vd_server_create(0, 4, &s) == VD_EINVAL && s == nullptr);
    CHECK(vd_server_create(1 << 30, 0, &s) == VD_EINVAL);
    vd_server_destroy(nullptr);

    // grid rules
    uint32_t g[3];
    OK(vd_grid_for(kLo, kHi, 1000000, 1.0, g));
    CHECK(g[0] == 100 && g[1] == 100 && g[2] == 100);
    CHECK(vd_grid_for(kLo, kHi, 1000, -1.0, g) == VD_EINVAL);
    const std::vector<double> up = uniform_points(200000, 1), cp = clustered_points(200000, 2);
    uint32_t gu[3], gc[3];
    OK(vd_grid_for_points(kLo, kHi, up.data(), up.size() / 3, 0.35, gu));
    OK(vd_grid_for_points(kLo, kHi, cp.data(), cp.size() / 3, 0.35, gc));
    CHECK(gu[0] > 0 && gu[0] == gu[1] && gu[1] == gu[2]);
    CHECK(gc[0] > 0 && gc[0] != gu[0]);
    OK(vd_grid_for_points(kLo, kHi, nullptr, 0, 0.35, g));

    // host shell-order prefix: size, then fill; keys non-decreasing, the next key above the last
    for (uint32_t n : {4u, 9u, 31u}) {
        size_t count = 0;
        double next = 0.0;
        OK(vd_shell_prefix(kLo, kHi, n, n + 1, n + 2, 3, nullptr, nullptr, 0, &count, &next));
        CHECK(count > 0);
        std::vector<int32_t> d(3 * count);
        std::vector<double> k(count);
        size_t c2 = 0;
        OK(vd_shell_prefix(kLo, kHi, n, n + 1, n + 2, 3, d.data(), k.data(), count, &c2, &next));
        CHECK(c2 == count);
        for (size_t i = 1; i < count; ++i) CHECK(k[i - 1] <= k[i]);
        CHECK(next > k[count - 1]);
    }

    // the static entry's limits are checked before the device
    size_t nc = 0;
    double dummy = 0.0;
    uint32_t du = 0;
    CHECK(vd_calculate_static(kLo, kHi, 4, 4, 4, nullptr, 0, &dummy, (size_t)VD_STATIC_MAX_POINTS + 1,
                              &dummy, &dummy, &du, &nc, nullptr) == VD_EINVAL);
    CHECK(vd_calculate_static(kLo, kHi, 2048, 2048, 2048, nullptr, 0, &dummy, 1, &dummy, &dummy, &du, &nc,
                              nullptr) == VD_EINVAL);
    std::printf("host checks passed\n");
}

// Handle creation where no GPU is visible: the call fails and releases what it made.
void no_device_checks() {
    vd_tess *t = nullptr;
    CHECK(vd_tess_create(kLo, kHi, 4, 4, 4, &t) != VD_OK && t == nullptr);
    vd_batch *b = nullptr;
    CHECK(vd_batch_create(&b) != VD_OK && b == nullptr);
    vd_server *s = nullptr;
    CHECK(vd_server_create(size_t(1) << 34, 8, &s) != VD_OK && s == nullptr);
    std::printf("no-device checks passed\n");
}

void tess_checks() {
    const size_t n = 20000;
    const std::vector<double> p = uniform_points(n, 3);
    vd_tess *t = nullptr;
    OK(vd_tess_create(kLo, kHi, 16, 16, 16, &t));
    const double c[3] = {0.5, 0.5, 0.5};
    OK(vd_tess_add_wall_sphere(t, -1000, c, 0.5));
    CHECK(vd_tess_add_wall_sphere(t, -999, c, 0.5) == VD_EINVAL);   // id above WALL_ID_MAX
    const vd_wall_host hw{nullptr, half_contains, half_cut, 1};
    OK(vd_tess_add_wall_planes_per_generator(t, -1001, &hw));
    OK(vd_tess_set_generators(t, p.data(), n));
    CHECK(vd_tess_count_cells(t) == 0);
    double vol0 = 0.0;
    CHECK(vd_tess_get_volumes(t, &vol0, 1) == VD_ESTATE);
    OK(vd_tess_calculate(t));
    const size_t m = vd_tess_count_cells(t);
    CHECK(m > 0 && m == vd_tess_count_generators(t));
    std::vector<double> vol(m), cen(3 * m), gen(3 * m);
    std::vector<uint32_t> nf(m);
    OK(vd_tess_get_volumes(t, vol.data(), m));
    OK(vd_tess_get_centroids(t, cen.data(), m));
    OK(vd_tess_get_face_counts(t, nf.data(), m));
    OK(vd_tess_get_generators(t, gen.data(), m));
    CHECK(vd_tess_get_volumes(t, vol.data(), m - 1) == VD_EINVAL);
    double sum = 0.0;
    for (double v : vol) sum += v;
    // the ball of radius 0.5 (0.5236) less the cap beyond x + y + z = 2 (0.0603), plus the
    // tangent-plane excess of the sphere wall's cells
    CHECK(sum > 0.462 && sum < 0.466);
    std::vector<uint64_t> off(m + 1);
    OK(vd_tess_get_neighbors(t, off.data(), m, nullptr, 0));
    std::vector<int32_t> ids(off[m]);
    OK(vd_tess_get_neighbors(t, off.data(), m, ids.data(), ids.size()));
    for (size_t i = 0; i < m; i += 97) {
        uint64_t sz[3];
        OK(vd_tess_get_cell(t, i, sz, nullptr, nullptr, nullptr, nullptr));
        std::vector<double> vx(3 * sz[0]);
        std::vector<uint32_t> fc(sz[1]), fi(sz[2]);
        std::vector<int32_t> fn(sz[1]);
        OK(vd_tess_get_cell(t, i, sz, vx.data(), fc.data(), fi.data(), fn.data()));
        CHECK(sz[1] == nf[i]);
        for (size_t f = 0; f < sz[1]; ++f) CHECK(fn[f] == ids[off[i] + f]);
    }
    uint64_t sz[3];
    CHECK(vd_tess_get_cell(t, m, sz, nullptr, nullptr, nullptr, nullptr) == VD_ERANGE);
    OK(vd_tess_calculate_sealed(t));
    size_t ns = 0;
    OK(vd_tess_get_seal_log(t, nullptr, 0, &ns));
    std::vector<int32_t> slog(3 * ns + 3);
    OK(vd_tess_get_seal_log(t, slog.data(), slog.size(), &ns));
    OK(vd_tess_prune_boundaries(t));
    size_t pc[2];
    OK(vd_tess_get_prune_log(t, nullptr, 0, nullptr, 0, pc));
    std::vector<int32_t> pi(pc[0] + 1);
    std::vector<double> pp(pc[1] + 1);
    OK(vd_tess_get_prune_log(t, pi.data(), pi.size(), pp.data(), pp.size(), pc));
    for (int k = 0; k < 2; ++k) {
        OK(vd_tess_relax(t));
        OK(vd_tess_calculate(t));
    }
    // single moves: across bins in both directions, within a bin, rejected by a wall, out of range
    for (size_t k = 0; k < 400; ++k) {
        const double q[3] = {p[3 * ((k * 7) % n)], p[3 * ((k * 13) % n) + 1], p[3 * ((k * 31) % n) + 2]};
        OK(vd_tess_set_generator(t, (k * 11) % m, q));
    }
    const double far[3] = {0.02, 0.02, 0.02};
    OK(vd_tess_set_generator(t, 0, far));   // outside the sphere wall: ignored
    CHECK(vd_tess_set_generator(t, m, c) == VD_ERANGE);
    CHECK(vd_tess_set_generator(t, 0, nullptr) == VD_EINVAL);
    OK(vd_tess_calculate(t));
    OK(vd_tess_set_generator(t, 1, c));
    OK(vd_tess_calculate(t));
    OK(vd_tess_random_generators(t, 5000, 12345));   // a few device batches through both walls
    CHECK(vd_tess_count_generators(t) == 5000);
    OK(vd_tess_calculate(t));
    OK(vd_tess_set_generators(t, p.data(), n));
    OK(vd_tess_calculate(t));
    int picks[4];
    OK(vd_tess_autotune_launch(t, picks));
    OK(vd_tess_set_option(t, VD_OPT_KEEP_CELLS, 0));
    OK(vd_tess_set_option(t, VD_OPT_ARENA, 4));   // small arena: the doubling path
    OK(vd_tess_calculate(t));
    double st[10];
    OK(vd_tess_get_stats(t, st));
    OK(vd_tess_clear_walls(t));
    double *d = nullptr;
    CHECK(hipMalloc(&d, n * 24) == hipSuccess);
    CHECK(hipMemcpy(d, p.data(), n * 24, hipMemcpyHostToDevice) == hipSuccess);
    OK(vd_tess_set_generators_device(t, d, n));
    OK(vd_tess_calculate(t));
    CHECK(vd_tess_count_cells(t) == n);
    vd_tess_destroy(t);
    CHECK(hipFree(d) == hipSuccess);
    std::printf("tessellation checks passed\n");
}

void batch_checks() {
    vd_batch *b = nullptr;
    OK(vd_batch_create(&b));
    OK(vd_batch_set_option(b, VD_BOPT_KEEP_CELLS, 1));
    std::vector<std::vector<double>> pts;
    for (int k = 0; k < 12; ++k) {
        pts.push_back(uniform_points(200 + 700 * k, 10 + k));
        vd_wall w{};
        w.id = -1000;
        w.kind = (k % 3 == 0) ? VD_WALL_SPHERE : VD_WALL_PLANE;
        w.a[0] = w.a[1] = w.a[2] = 0.5;
        w.b[2] = 1.0;
        w.r = 0.45;
        uint32_t idx = 0;
        OK(vd_batch_add(b, kLo, kHi, 0, 0, 0, k % 3 == 2 ? nullptr : &w, k % 3 == 2 ? 0 : 1, pts.back().data(),
                        pts.back().size() / 3, &idx));
        CHECK(idx == (uint32_t)k);
    }
    CHECK(vd_batch_count(b) == 12);
    OK(vd_batch_calculate(b));
    for (uint32_t t = 0; t < 12; ++t) {
        uint64_t info[5];
        OK(vd_batch_member_info(b, t, info));
        const size_t m = info[0];
        std::vector<double> vol(m), cen(3 * m), gen(3 * m);
        std::vector<uint32_t> nf(m);
        OK(vd_batch_get_volumes(b, t, vol.data(), m));
        OK(vd_batch_get_centroids(b, t, cen.data(), m));
        OK(vd_batch_get_face_counts(b, t, nf.data(), m));
        OK(vd_batch_get_generators(b, t, gen.data(), m));
        std::vector<uint64_t> off(m + 1);
        OK(vd_batch_get_neighbors(b, t, off.data(), m, nullptr, 0));
        std::vector<int32_t> ids(off[m] + 1);
        OK(vd_batch_get_neighbors(b, t, off.data(), m, ids.data(), ids.size()));
        if (m) {
            uint64_t sz[3];
            OK(vd_batch_get_cell(b, t, m - 1, sz, nullptr, nullptr, nullptr, nullptr));
        }
    }
    uint64_t info[5];
    CHECK(vd_batch_member_info(b, 12, info) == VD_ERANGE);
    OK(vd_batch_relax(b));
    OK(vd_batch_calculate(b));
    const double cand[3] = {0.35, 1.0, 2.0};
    double best = 0.0;
    OK(vd_batch_autotune(b, cand, 3, &best));
    int picks[4];
    OK(vd_batch_autotune_launch(b, picks));
    OK(vd_batch_clear(b));
    CHECK(vd_batch_count(b) == 0);
    vd_batch_destroy(b);
    std::printf("batch checks passed\n");
}

void server_checks() {
    vd_server *s = nullptr;
    OK(vd_server_create(size_t(8) << 30, 16, &s));
    OK(vd_server_set_option(s, VD_SOPT_SMALL_JOB, 20000));
    std::vector<std::vector<double>> pts;
    std::vector<uint64_t> jobs;
    for (int k = 0; k < 6; ++k) {
        pts.push_back(k % 2 ? clustered_points(5000 + 20000 * k, 30 + k) : uniform_points(5000 + 20000 * k, 30 + k));
        uint64_t j = 0;
        OK(vd_server_submit(s, kLo, kHi, 0, 0, 0, nullptr, 0, pts.back().data(), pts.back().size() / 3,
                            k % 3 == 0 ? 4 : 1, &j));
        jobs.push_back(j);
    }
    for (int step = 0; step < 10000; ++step) {
        uint32_t fin = 0;
        OK(vd_server_step(s, &fin));
        uint64_t pend[3];
        OK(vd_server_pending(s, pend));
        if (pend[0] == 0 && pend[1] == 0) break;
    }
    for (uint64_t j : jobs) {
        int32_t state = -1;
        uint64_t info[3];
        double times[3];
        OK(vd_server_job_info(s, j, &state, info, times));
        CHECK(state == 2);   // finished
        const size_t m = info[0];
        std::vector<double> vol(m), cen(3 * m), gen(3 * m);
        OK(vd_server_result(s, j, vol.data(), cen.data(), gen.data(), m));
        uint32_t g[3];
        OK(vd_server_job_grid(s, j, g));
        OK(vd_server_release(s, j));
    }
    uint64_t mem[4];
    OK(vd_server_memory(s, mem));
    double st[4];
    OK(vd_server_stats(s, st));
    vd_server_destroy(s);
    std::printf("server checks passed\n");
}

void static_checks() {
    for (size_t n : {30000u, 7000u}) {
        const std::vector<double> p = uniform_points(n, 40 + (uint32_t)n);
        double *d = nullptr, *vol = nullptr, *cen = nullptr;
        uint32_t *nf = nullptr;
        CHECK(hipMalloc(&d, n * 24) == hipSuccess && hipMalloc(&vol, n * 8) == hipSuccess);
        CHECK(hipMalloc(&cen, n * 24) == hipSuccess && hipMalloc(&nf, n * 4) == hipSuccess);
        CHECK(hipMemcpy(d, p.data(), n * 24, hipMemcpyHostToDevice) == hipSuccess);
        size_t nc = 0;
        OK(vd_calculate_static(kLo, kHi, 24, 24, 24, nullptr, 0, d, n, vol, cen, nf, &nc, nullptr));
        CHECK(hipDeviceSynchronize() == hipSuccess);
        CHECK(nc == n);
        std::vector<double> hv(n);
        CHECK(hipMemcpy(hv.data(), vol, n * 8, hipMemcpyDeviceToHost) == hipSuccess);
        double sum = 0.0;
        for (double v : hv) sum += v;
        CHECK(std::fabs(sum - 1.0) < 1e-9);
        CHECK(hipFree(d) == hipSuccess && hipFree(vol) == hipSuccess);
        CHECK(hipFree(cen) == hipSuccess && hipFree(nf) == hipSuccess);
    }
    std::printf("static entry checks passed\n");
}

}  // namespace

int main(int argc, char **argv) {
    const bool device = argc > 1 && std::strcmp(argv[1], "--device") == 0;
    host_checks();
    if (!device) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) no_device_checks();
        else std::printf("a device is visible: no-device checks skipped\n");
        return 0;
    }
    tess_checks();
    batch_checks();
    server_checks();
    static_checks();
    return 0;
}

// The float64 restatement under AddressSanitizer and UndefinedBehaviorSanitizer (test
// infrastructure only): built by `make -C oracle asan` together with sph_oracle.cpp, it runs every
// entry point on a jittered 2D block of water in a box, a periodic 3D block, mixed smoothing lengths
// and a lone particle, and steps the 2D block. Any out-of-bounds access, use after free or
// undefined arithmetic aborts with the sanitizer's report; a clean run prints "asan check passed".

#include "sph_oracle.cpp"

#include <cstdio>
#include <random>

namespace {

struct Cloud {
    int n = 0;
    std::vector<double> pos, vel, dens, press, mass, len, acc;
    std::vector<uint8_t> fluid;
};

// A block of `nx` x `ny` (x `nz`) particles a spacing `s` apart, jittered by a tenth of a spacing,
// its bottom layer boundary particles; lengths alternate between h and h / 2 when `mixed`.
Cloud block(int dims, int nx, int ny, int nz, double s, double h, bool mixed, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> jitter(-0.1 * s, 0.1 * s), speed(-0.3, 0.3), rho(995.0, 1010.0);
    Cloud c;
    for (int k = 0; k < (dims == 3 ? nz : 1); ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = (i + 0.5) * s + jitter(rng), y = (j + 0.5) * s + jitter(rng);
                const double z = dims == 3 ? (k + 0.5) * s + jitter(rng) : 0.0;
                c.pos.insert(c.pos.end(), {x, y, z});
                c.vel.insert(c.vel.end(), {speed(rng), speed(rng), dims == 3 ? speed(rng) : 0.0});
                c.dens.push_back(rho(rng));
                c.mass.push_back(1000.0 * std::pow(s, dims));
                c.len.push_back(mixed && (i + j) % 2 ? 0.5 * h : h);
                c.fluid.push_back((dims == 3 ? k : j) > 0 ? 1 : 0);
                ++c.n;
            }
    c.press.assign(c.n, 0.0);
    c.acc.assign(3 * (size_t)c.n, 0.0);
    return c;
}

OracleParams params(int dims, double s, double h, double extent, double period) {
    OracleParams p{};
    p.dims = dims;
    p.kernel = 1;
    p.cell = 2.0 * h;
    const int cells = (int)std::ceil(extent / p.cell) + 3;
    p.grid[0] = cells;
    p.grid[1] = period > 0.0 ? (int)std::floor(period / p.cell) : cells;
    p.grid[2] = dims == 3 ? cells : 1;
    p.hmax = h;
    p.hmin = 0.5 * h;
    p.period = period;
    p.rho0 = 1000.0;
    p.c0 = 20.0;
    p.alpha = 0.1;
    p.boundaryFactor = 0.5;
    p.gravity[dims - 1] = -9.81;
    p.diffusion = 2;
    p.delta = 0.1;
    p.extrapolatedWalls = 1;
    p.correction = 1;
    p.limiter = 1;
    p.densityFloor = 900.0;
    p.cfl = 0.25;
    p.maxTimeStep = 1e-3;
    (void)s;
    return p;
}

void everySum(OracleParams p, Cloud& c) {
    const int n = c.n;
    std::vector<int64_t> counts(n), offsets(n + 1, 0), keys(n);
    oracle_cell_keys(&p, n, c.pos.data(), keys.data());
    oracle_neighbors(&p, n, c.pos.data(), c.len.data(), counts.data(), nullptr, nullptr);
    for (int i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + counts[i];
    std::vector<int32_t> rows(offsets[n] > 0 ? offsets[n] : 1);
    oracle_neighbors(&p, n, c.pos.data(), c.len.data(), nullptr, offsets.data(), rows.data());
    oracle_neighbors32(&p, n, c.pos.data(), c.len.data(), counts.data(), nullptr, nullptr);
    std::vector<double> outP(n), outRho(n), g9(9 * (size_t)n), l9(9 * (size_t)n), acc(3 * (size_t)n), rate(n);
    std::vector<double> grad(3 * (size_t)n), shepard(n), xsph(3 * (size_t)n), alphas(n), deltas(n);
    oracle_wall_pressure(&p, n, c.pos.data(), c.vel.data(), c.dens.data(), c.press.data(), c.mass.data(), c.len.data(),
                         c.fluid.data(), outP.data(), outRho.data());
    oracle_velocity_gradients(&p, n, c.pos.data(), c.vel.data(), c.dens.data(), c.mass.data(), c.len.data(),
                              c.fluid.data(), g9.data());
    oracle_correction(&p, n, c.pos.data(), c.dens.data(), c.mass.data(), c.len.data(), c.fluid.data(), l9.data());
    oracle_accelerations(&p, n, c.pos.data(), c.vel.data(), c.dens.data(), c.press.data(), c.mass.data(), c.len.data(),
                         c.fluid.data(), g9.data(), nullptr, acc.data());
    oracle_density_gradients(&p, n, c.pos.data(), c.dens.data(), c.mass.data(), c.len.data(), c.fluid.data(),
                             l9.data(), grad.data());
    oracle_closure_coefficients(&p, n, g9.data(), c.len.data(), c.fluid.data(), p.c0, alphas.data(), deltas.data());
    const OracleClosure closure{alphas.data(), deltas.data(), p.c0, p.rho0, 0.0};
    for (int mode = 0; mode < 4; ++mode) {
        p.diffusion = mode;
        oracle_continuity(&p, n, c.pos.data(), c.vel.data(), c.dens.data(), c.mass.data(), c.len.data(), c.fluid.data(),
                          l9.data(), grad.data(), mode == 3 ? &closure : nullptr, rate.data());
    }
    p.diffusion = 2;
    oracle_accelerations(&p, n, c.pos.data(), c.vel.data(), c.dens.data(), c.press.data(), c.mass.data(), c.len.data(),
                         c.fluid.data(), g9.data(), &closure, acc.data());
    oracle_shepard(&p, n, c.pos.data(), c.dens.data(), c.mass.data(), c.len.data(), shepard.data());
    oracle_xsph(&p, n, c.pos.data(), c.vel.data(), c.dens.data(), c.mass.data(), c.len.data(), c.fluid.data(), 0.5,
                xsph.data());
}

}  // namespace

int main() {
    oracle_threads(2);
    const double s = 0.01, h = 0.013;
    // 2D: sums with mixed lengths, then shifting (one length) and steps in a box
    Cloud flat = block(2, 24, 12, 1, s, h, true, 1);
    everySum(params(2, s, h, 0.3, 0.0), flat);
    Cloud uniform = block(2, 24, 12, 1, s, h, false, 2);
    OracleParams p2 = params(2, s, h, 0.3, 0.0);
    p2.hmin = h;
    std::vector<double> shift(3 * (size_t)uniform.n), change(uniform.n), dv(3 * (size_t)uniform.n);
    oracle_shifting(&p2, uniform.n, uniform.pos.data(), uniform.vel.data(), uniform.dens.data(), uniform.mass.data(),
                    uniform.fluid.data(), 1e-5, 1, 0.5, shift.data(), change.data(), dv.data());
    OracleBoundary box{};
    box.kind = 1;
    box.clearance = 0.25 * s;
    box.upper[0] = box.upper[1] = 0.24;
    box.openTop = 1;
    double time = 0.0, lastDt = 0.0;
    std::vector<double> budget(3 * 10);
    const int taken = oracle_advance(&p2, &box, uniform.n, uniform.pos.data(), uniform.vel.data(), uniform.dens.data(),
                                     uniform.press.data(), uniform.acc.data(), uniform.mass.data(), uniform.len.data(),
                                     uniform.fluid.data(), &time, 1.0, 10, &lastDt, budget.data());
    oracle_enforce(&p2, &box, uniform.n, uniform.pos.data(), uniform.vel.data(), uniform.fluid.data());
    // 3D: periodic along y, mixed lengths
    Cloud deep = block(3, 10, 8, 8, s, h, true, 3);
    everySum(params(3, s, h, 0.1, 0.08), deep);
    // a lone particle: no neighbors at all
    Cloud lone = block(2, 1, 2, 1, s, h, false, 4);
    everySum(params(2, s, h, 0.02, 0.0), lone);
    std::printf("asan check passed: %d steps to t = %.6f s\n", taken, time);
    return taken == 10 ? 0 : 1;
}

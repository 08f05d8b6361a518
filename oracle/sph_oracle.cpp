// FP64 CPU restatement of pySurf's weakly compressible delta-SPH time step.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker or the timed CPU baseline; the
// HIP library never calls it.
//
// The design follows the reference's numba backend (sphCompiled.py:133-234): a counting-sorted
// cell list and a per-particle gather that evaluates each pair from both sides, threaded over
// particles (OpenMP here, numba prange there). The equations and their order are the device
// step's (sphGpu.py:2237-2315), each written in float64:
//
//   rates      cell build (sphGpu.py:1635-1641, sphSolver.py:1379-1408), EOS (sphGpu.py:1058-1070),
//              extrapolated wall pressure (sphGpu.py:804-854), velocity gradients (sphGpu.py:643-697),
//              accelerations with the limited viscosity (sphGpu.py:303-414), fluid maxima (:1072-1082)
//   time step  cflTimeStep (sphSolver.py:1410-1424)
//   integrate  kick (sphGpu.py:1084-1109), correction matrices (sphGpu.py:593-641), continuity with
//              delta-SPH diffusion (sphGpu.py:430-545), density advance with floor (:1111-1140),
//              drift (:1142-1186), boundary enforcement (:1197-1229, :1246-1310, :1332-1419)
//   options    the host-step sums the delta-SPH scheme leaves off: the Shepard density
//              (sphGpu.py:768-802), the XSPH velocity (:725-766), the renormalized density gradient
//              and its diffusion (:548-591, :521-524), the delta-LES-SPH closure's coefficients,
//              viscosity and diffusion (:699-723, :388-393, :526-531), particle shifting (:856-1004)
//
// Parity status of each function is pinned in tests/test_oracle.py (invariants, closed forms,
// brute force, and the numpy pair-list restatement in oracle/array_formulation.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Physics and grid of one run. Mirrors the fields of the HIP library's run description.
typedef struct {
    int32_t dims;            // 2 or 3
    int32_t kernel;          // 0 cubic spline, 1 Wendland C2
    int32_t grid[3];         // cells along x, y, z (z = 1 in 2D)
    double cell;             // cell width, 2 h_max [m]
    double hmax;             // the longest smoothing length [m]
    double hmin;             // the shortest smoothing length, which sets the time step [m]
    double period;           // periodic y length [m], 0 for none
    double rho0, c0, alpha, boundaryFactor;
    double gravity[3];
    int32_t diffusion;       // 0 none, 1 molteni, 2 hydrostatic, 3 antuono (host-step sums)
    double delta;
    int32_t extrapolatedWalls, negativePressure, correction, limiter;
    double densityFloor;     // smallest fluid density [kg/m^3], 0 for none
    double cfl, maxTimeStep;
} OracleParams;

// The boundary the step enforces (sphGpu.py:2121-2171).
typedef struct {
    int32_t kind;            // 0 none, 1 container, 2 bed, 3 height field
    double clearance;        // a quarter spacing [m]
    // container
    double lower[3], upper[3];
    int32_t openTop;
    // bed
    int32_t nBed;
    const double* bedX;
    const double* bedZ;
    int32_t sideWalls;
    double width;
    int32_t offshoreWall;
    double wallX, pistonSpeed, endLimit;
    // height field: table[ny][nx] of bed elevations on a regular grid
    int32_t fieldNx, fieldNy;
    const double* table;
    double originX, originY, stepX, stepY;
    double lowY, highY, lowX, highX;
} OracleBoundary;

// The delta-LES-SPH closure a sum carries (sphGpu.py:1754-1778): each particle's alpha and delta
// (oracle_closure_coefficients), the closure's c0 and rho0, and K nu / c0 with K = 2 (D + 2).
typedef struct {
    const double* alphas;
    const double* deltas;
    double soundSpeed, referenceDensity, waterTerm;
} OracleClosure;

// The options a device-resident step may carry besides the delta-SPH scheme (sphGpu.py:2237-2315):
// the Shepard filter every `filterInterval` steps (sphSolver.py:1979-1981: at step > 0 with
// step % interval == 0) and the delta-LES-SPH closure with its c0, rho0 and K nu / c0. The
// 'antuono' diffusion comes with the run's diffusion mode.
typedef struct {
    int32_t filterInterval;
    int32_t closure;
    double closureSoundSpeed, closureDensity, closureWater;
} OracleOptions;

}  // extern "C"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSoftening = 0.01;      // sphSolver.py:228 (viscositySoftening)
constexpr double kCloseDistance = 0.3;   // sphSolver.py:234 (limiterCloseDistance)
constexpr double kFoldWidth = 0.2;       // sphSolver.py:235 (limiterFoldWidth)
constexpr double kCorrectionFloor = 0.25;  // sphSolver.py:262 (correctionFloor)

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct M3 {
    double m[3][3];
};
inline M3 zeroM() { M3 r; std::memset(&r, 0, sizeof r); return r; }
inline M3 identityM() { M3 r = zeroM(); r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0; return r; }
inline V3 mul(const M3& a, V3 v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
inline void addOuter(M3& a, double s, V3 u, V3 v) {
    const double uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a.m[r][c] += s * uu[r] * vv[c];
}
inline double det(const M3& a) {
    const auto& m = a.m;
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
inline M3 inverse(const M3& a) {
    const auto& m = a.m;
    const double d = det(a);
    M3 r;
    r.m[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / d;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / d;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / d;
    r.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / d;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / d;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / d;
    r.m[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / d;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / d;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / d;
    return r;
}
inline M3 matmul(const M3& a, const M3& b) {
    M3 r = zeroM();
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) r.m[i][j] += a.m[i][k] * b.m[k][j];
    return r;
}

// Kernel normalization at h = 1 (sphGpu.py:1424-1444): sigma(h) = unit / h^dims.
double sigmaUnit(int kernel, int dims) {
    if (kernel == 1) return dims == 2 ? 7.0 / (4.0 * kPi) : 21.0 / (16.0 * kPi);
    return dims == 2 ? 10.0 / (7.0 * kPi) : 1.0 / kPi;
}

// W(r) and dW/dr (sphGpu.py:139-186, sphKernels.py:387-644).
double kernelValue(int kind, double r, double h, double sigma) {
    const double q = r / h;
    if (kind == 1) {
        if (q < 2.0) {
            const double rest = 1.0 - 0.5 * q, sq = rest * rest;
            return sigma * sq * sq * (2.0 * q + 1.0);
        }
        return 0.0;
    }
    if (q < 1.0) return sigma * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
    if (q < 2.0) {
        const double rest = 2.0 - q;
        return sigma * 0.25 * rest * rest * rest;
    }
    return 0.0;
}
double kernelSlope(int kind, double r, double h, double sigma) {
    const double q = r / h;
    if (kind == 1) {
        if (q < 2.0) {
            const double rest = 1.0 - 0.5 * q;
            return -5.0 * sigma * q * rest * rest * rest / h;
        }
        return 0.0;
    }
    if (q < 1.0) return sigma * (-3.0 * q + 2.25 * q * q) / h;
    if (q < 2.0) {
        const double rest = 2.0 - q;
        return sigma * (-0.75 * rest * rest) / h;
    }
    return 0.0;
}
inline double sigmaAt(const OracleParams& p, double h) {
    const double unit = sigmaUnit(p.kernel, p.dims);
    return p.dims == 2 ? unit / (h * h) : unit / (h * h * h);
}

// Python's floor modulus, as np.mod gives it for the wrapped cell index.
inline int64_t wrap(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

// The cell-linked list over one run's particles: a counting sort by cellOrderKey
// (sphSolver.py:1379-1408), stable in particle index, as sphCompiled.py:133-180 builds its list.
struct CellList {
    int64_t nx, ny, nz;
    double cell;
    std::vector<int64_t> start;   // cells + 1
    std::vector<int32_t> order;   // particle ids by cell

    int64_t cellOf(double v) const { return (int64_t)std::floor(v / cell); }

    void build(const OracleParams& p, int n, const double* pos) {
        nx = p.grid[0];
        ny = p.grid[1];
        nz = p.dims == 3 ? p.grid[2] : 1;
        cell = p.cell;
        const int64_t cells = nx * ny * nz;
        start.assign(cells + 1, 0);
        std::vector<int64_t> key(n);
        for (int i = 0; i < n; ++i) {
            const double* x = pos + 3 * i;
            const int64_t cx = wrap(cellOf(x[0]), nx), cy = wrap(cellOf(x[1]), ny);
            const int64_t cz = p.dims == 3 ? wrap(cellOf(x[2]), nz) : 0;
            key[i] = (cz * ny + cy) * nx + cx;
            start[key[i] + 1]++;
        }
        for (int64_t c = 0; c < cells; ++c) start[c + 1] += start[c];
        std::vector<int64_t> cursor(start.begin(), start.end() - 1);
        order.resize(n);
        for (int i = 0; i < n; ++i) order[cursor[key[i]]++] = i;
    }

    // The distinct offsets along an axis of n cells: a grid of fewer than three cells would visit
    // one cell twice.
    static int offsets(int64_t n, int* out) {
        if (n >= 3) { out[0] = -1; out[1] = 0; out[2] = 1; return 3; }
        if (n == 2) { out[0] = 0; out[1] = 1; return 2; }
        out[0] = 0;
        return 1;
    }

    // Calls visit(j) for every particle in the cells around `origin`.
    template <class F>
    void around(const OracleParams& p, V3 origin, F&& visit) const {
        int ox[3], oy[3], oz[3];
        const int kx = offsets(nx, ox), ky = offsets(ny, oy), kz = p.dims == 3 ? offsets(nz, oz) : (oz[0] = 0, 1);
        const int64_t bx = cellOf(origin.x), by = cellOf(origin.y), bz = p.dims == 3 ? cellOf(origin.z) : 0;
        for (int a = 0; a < kz; ++a)
            for (int b = 0; b < ky; ++b)
                for (int c = 0; c < kx; ++c) {
                    const int64_t cx = wrap(bx + ox[c], nx), cy = wrap(by + oy[b], ny);
                    const int64_t cz = p.dims == 3 ? wrap(bz + oz[a], nz) : 0;
                    const int64_t key = (cz * ny + cy) * nx + cx;
                    for (int64_t k = start[key]; k < start[key + 1]; ++k) visit(order[k]);
                }
    }
};

// State views over padded (N, 3) float64 arrays.
inline V3 vec(const double* a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
inline void store(double* a, int i, V3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }

// Searches a particle makes across the periodic seam (sphGpu.py:188-204).
inline int imageCount(const OracleParams& p, V3 point, double reach) {
    if (p.period > 0.0 && (point.y < reach || point.y > p.period - reach)) return 2;
    return 1;
}
inline double imageShift(const OracleParams& p, V3 point, int image) {
    if (image == 0) return 0.0;
    return point.y < 0.5 * p.period ? p.period : -p.period;
}

// One neighbor of particle i: d = r_i - r_j (nearest image), r, h_ij.
struct Pair {
    int j;
    V3 d;
    double r, h;
};

// Calls f(Pair) for every j != i with r < 2 h_ij (sphGpu.py:352-363), r evaluated in float64 as
// sqrt(fma(dz, dz, fma(dy, dy, dx * dx))). `skipCoincident` drops r < 1e-12 as every kernel but
// the wall pressure and Shepard sums does.
template <class F>
void forNeighbors(const OracleParams& p, const CellList& cells, const double* pos, const double* lengths, int i,
                  bool skipCoincident, F&& f) {
    const V3 xi = vec(pos, i);
    const double hi = lengths[i];
    const double reach = hi + p.hmax;
    for (int image = 0; image < imageCount(p, xi, reach); ++image) {
        const V3 offset = {0.0, imageShift(p, xi, image), 0.0};
        const V3 origin = xi + offset;
        cells.around(p, origin, [&](int j) {
            if (j == i) return;
            const V3 d = origin - vec(pos, j);
            const double r = std::sqrt(std::fma(d.z, d.z, std::fma(d.y, d.y, d.x * d.x)));
            const double h = 0.5 * (hi + lengths[j]);
            if (r >= 2.0 * h) return;
            if (skipCoincident && r < 1e-12) return;
            f(Pair{j, d, r, h});
        });
    }
}

struct Fields {
    int n;
    const double* pos;
    const double* vel;
    const double* dens;
    const double* press;
    const double* mass;
    const double* len;
    const uint8_t* fluid;
};

// Tait pressure, clamped at zero unless the fluid carries tension (sphGpu.py:1058-1070).
inline double tait(const OracleParams& p, double rho, bool isFluid) {
    const double stiffness = p.rho0 * p.c0 * p.c0 / 7.0;
    const double ratio = rho / p.rho0, sq = ratio * ratio;
    double pressure = stiffness * (sq * sq * sq * ratio - 1.0);
    if (!p.negativePressure || !isFluid) pressure = std::max(pressure, 0.0);
    return pressure;
}

// Extrapolated wall pressure on boundary particle i (sphGpu.py:804-854, sphSolver.py:1045-1106).
void wallPressureAt(const OracleParams& p, const CellList& cells, const Fields& s, int i, double* outP,
                    double* outRho) {
    double weight = 0.0, total = 0.0;
    const V3 g = {p.gravity[0], p.gravity[1], p.gravity[2]};
    forNeighbors(p, cells, s.pos, s.len, i, false, [&](const Pair& q) {
        if (!s.fluid[q.j]) return;
        const double value = kernelValue(p.kernel, q.r, q.h, sigmaAt(p, q.h));
        weight += value;
        total += value * (s.press[q.j] + s.dens[q.j] * dot(g, q.d));
    });
    const double pressure = weight > 0.0 ? std::max(total / weight, 0.0) : 0.0;
    const double stiffness = p.rho0 * p.c0 * p.c0 / 7.0;
    *outP = pressure;
    *outRho = p.rho0 * std::pow(1.0 + pressure / stiffness, 1.0 / 7.0);
}

// G_i = S_i M_i^-1 over fluid neighbors (sphGpu.py:643-697, sphSolver.py:1151-1190).
M3 velocityGradientAt(const OracleParams& p, const CellList& cells, const Fields& s, int i) {
    if (!s.fluid[i]) return zeroM();
    const V3 vi = vec(s.vel, i);
    M3 rises = zeroM(), offsets = zeroM();
    forNeighbors(p, cells, s.pos, s.len, i, true, [&](const Pair& q) {
        if (!s.fluid[q.j]) return;
        const double w = s.mass[q.j] / s.dens[q.j] * kernelSlope(p.kernel, q.r, q.h, sigmaAt(p, q.h)) / q.r;
        addOuter(rises, w, vec(s.vel, q.j) - vi, q.d);
        addOuter(offsets, -w, q.d, q.d);
    });
    if (p.dims == 2) offsets.m[2][2] += 1.0;
    const double floor = std::pow(kCorrectionFloor, p.dims);
    if (det(offsets) >= floor) return matmul(rises, inverse(offsets));
    return zeroM();
}

// L_i = M_i^-1 over every neighbor (sphGpu.py:593-641, sphSolver.py:656-689).
M3 correctionAt(const OracleParams& p, const CellList& cells, const Fields& s, int i) {
    if (!s.fluid[i]) return identityM();
    M3 total = zeroM();
    forNeighbors(p, cells, s.pos, s.len, i, true, [&](const Pair& q) {
        const double w = -s.mass[q.j] / s.dens[q.j] * kernelSlope(p.kernel, q.r, q.h, sigmaAt(p, q.h)) / q.r;
        addOuter(total, w, q.d, q.d);
    });
    if (p.dims == 2) total.m[2][2] += 1.0;
    if (det(total) >= std::pow(kCorrectionFloor, p.dims)) return inverse(total);
    return identityM();
}

// a_i = -sum_j m_j (p_i / rho_i^2 + p_j / rho_j^2 + Pi_ij) grad W_ij + g (sphGpu.py:303-414;
// sphSolver.py:1122-1149, 1192-1230, 1328-1356).
// With `parts` (4 vectors), the pair terms without gravity split for the energy budget: pressure and
// viscosity from fluid neighbors, then from boundary neighbors (sphGpu.py:396-414).
V3 accelerationAt(const OracleParams& p, const CellList& cells, const Fields& s, const M3* gradients, int i,
                  const OracleClosure* closure = nullptr, V3* parts = nullptr) {
    if (parts)
        for (int k = 0; k < 4; ++k) parts[k] = {0.0, 0.0, 0.0};
    if (!s.fluid[i]) return {0.0, 0.0, 0.0};
    const V3 vi = vec(s.vel, i);
    const double rhoI = s.dens[i];
    const double own = s.press[i] / (rhoI * rhoI);
    const M3 gi = p.limiter ? gradients[i] : zeroM();
    V3 total = {0.0, 0.0, 0.0};
    forNeighbors(p, cells, s.pos, s.len, i, true, [&](const Pair& q) {
        const int j = q.j;
        const double rhoJ = s.dens[j];
        const V3 dv = vi - vec(s.vel, j);
        double approach = dot(dv, q.d);
        double viscosity = 0.0;
        if (approach < 0.0 && p.limiter && s.fluid[j]) {
            const V3 rampI = mul(gi, q.d), rampJ = mul(gradients[j], q.d);
            const double slopeI = dot(q.d, rampI), slopeJ = dot(q.d, rampJ);
            double limiter = 0.0;
            if (slopeI * slopeJ > 0.0) {
                const double ratio = slopeI / slopeJ;
                limiter = std::min(1.0, 4.0 * ratio / ((1.0 + ratio) * (1.0 + ratio)));
                const double eta = q.r / q.h;
                if (eta < kCloseDistance) {
                    const double fade = (eta - kCloseDistance) / kFoldWidth;
                    limiter *= std::exp(-fade * fade);
                }
            }
            approach = dot(dv - (0.5 * limiter) * (rampI + rampJ), q.d);
        }
        if (approach < 0.0) {
            const double mu = q.h * approach / (q.r * q.r + kSoftening * q.h * q.h);
            viscosity = -p.alpha * p.c0 * mu / (0.5 * (rhoI + rhoJ));
            if (!s.fluid[j]) viscosity *= p.boundaryFactor;
        }
        if (closure && s.fluid[j]) {
            // alpha_ij = K nu / (c0 h_ij) + the harmonic mean of the eddy alphas, on every fluid pair,
            // with the unlimited v_ij . r_ij (sphGpu.py:388-393, sphSolver.py:1283-1330)
            const double alphaI = closure->alphas[i], alphaJ = closure->alphas[j];
            const double both = alphaI + alphaJ;
            const double mean = both > 0.0 ? 2.0 * alphaI * alphaJ / both : 0.0;
            viscosity += -(closure->waterTerm / q.h + mean) * closure->soundSpeed * q.h * closure->referenceDensity *
                         dot(dv, q.d) / (q.r * q.r * rhoI * rhoJ);
        }
        const double coefficient = -(own + s.press[j] / (rhoJ * rhoJ) + viscosity) * s.mass[j];
        const double slope = kernelSlope(p.kernel, q.r, q.h, sigmaAt(p, q.h));
        total = total + (coefficient * slope / q.r) * q.d;
        if (parts) {
            const int side = s.fluid[j] ? 0 : 2;
            parts[side] = parts[side] + (-(own + s.press[j] / (rhoJ * rhoJ)) * s.mass[j] * slope / q.r) * q.d;
            parts[side + 1] = parts[side + 1] + (-viscosity * s.mass[j] * slope / q.r) * q.d;
        }
    });
    return total + V3{p.gravity[0], p.gravity[1], p.gravity[2]};
}

// d rho_i / dt with the delta-SPH diffusive term (sphGpu.py:416-545; sphSolver.py:623-654, 718-822).
// `deviation` holds rho - rho0.
// With `parts` (4 values) on a fluid particle, the rate split for the energy budget: the plain sum
// over fluid and over boundary neighbors, the correction's share and the diffusion (sphGpu.py:479-545).
double continuityAt(const OracleParams& p, const CellList& cells, const Fields& s, const double* deviation,
                    const M3* corrections, int i, const V3* gradients = nullptr,
                    const OracleClosure* closure = nullptr, double* parts = nullptr) {
    const V3 vi = vec(s.vel, i);
    const bool boundaryI = !s.fluid[i];
    const bool diffuses = p.diffusion != 0 && !boundaryI;
    const M3 li = (p.correction && corrections) ? corrections[i] : identityM();
    const int vertical = p.dims - 1;
    const double gOverC2 = std::fabs(p.gravity[vertical]) / (p.c0 * p.c0);
    const double factor = 2.0 * p.delta * p.c0;
    double total = 0.0, continuity = 0.0, plainFluid = 0.0, plainWall = 0.0, diffusive = 0.0;
    forNeighbors(p, cells, s.pos, s.len, i, true, [&](const Pair& q) {
        const int j = q.j;
        if (boundaryI && !s.fluid[j]) return;
        const double slope = kernelSlope(p.kernel, q.r, q.h, sigmaAt(p, q.h));
        const V3 dv = vi - vec(s.vel, j);
        const V3 dd = p.correction ? mul(li, q.d) : q.d;
        const double increment = s.mass[j] * dot(dv, dd) * slope / q.r;
        total += increment;
        if (parts) {
            continuity += increment;
            (s.fluid[j] ? plainFluid : plainWall) += s.mass[j] * dot(dv, q.d) * slope / q.r;
        }
        if (diffuses && s.fluid[j]) {
            double difference = deviation[j] - deviation[i];
            if (p.diffusion == 2) {
                const double vert = vertical == 1 ? q.d.y : q.d.z;
                const double ratio = 1.0 + 0.5 * (deviation[i] + deviation[j]) / p.rho0;
                const double sq = ratio * ratio;
                difference -= p.rho0 * gOverC2 * vert / (sq * sq * ratio);
            } else if (p.diffusion == 3 && gradients) {
                // less 1/2 (grad rho_i + grad rho_j) . (r_j - r_i), with d = r_i - r_j (sphGpu.py:521-524)
                difference += 0.5 * dot(gradients[i] + gradients[j], q.d);
            }
            double pairFactor = factor;
            if (closure) {
                // 2 c0 times the harmonic mean of the two deltas (sphGpu.py:526-531)
                const double deltaI = closure->deltas[i], deltaJ = closure->deltas[j];
                const double both = deltaI + deltaJ;
                pairFactor = both > 0.0 ? 2.0 * closure->soundSpeed * 2.0 * deltaI * deltaJ / both : 0.0;
            }
            const double increment = pairFactor * q.h * s.mass[j] / s.dens[j] * difference * (-slope * q.r) /
                                     (q.r * q.r + kSoftening * q.h * q.h);
            total += increment;
            diffusive += increment;
        }
    });
    if (parts) {
        const bool tracked = !boundaryI;
        parts[0] = tracked ? plainFluid : 0.0;
        parts[1] = tracked ? plainWall : 0.0;
        parts[2] = tracked && p.correction ? continuity - plainFluid - plainWall : 0.0;
        parts[3] = tracked ? diffusive : 0.0;
    }
    return total;
}

// The elastic energy per unit mass of the Tait fluid, e(rho) = integral from rho0 to rho of p / s^2
// ds = (c0^2 / 7) x^2 (21 + 35 x + 35 x^2 + 21 x^3 + 7 x^4 + x^5) / (6 (1 + x)), x = rho / rho0 - 1,
// zero below rho0 unless the fluid carries tension (sphEnergy.py:116-160). Checked against a
// quadrature of p / rho^2 in tests/test_oracle_energy.py.
double elasticEnergy(const OracleParams& p, double rho) {
    const double x = rho / p.rho0 - 1.0;
    if (x < 0.0 && !p.negativePressure) return 0.0;
    const double polynomial = 21.0 + x * (35.0 + x * (35.0 + x * (21.0 + x * (7.0 + x))));
    return p.c0 * p.c0 / 7.0 * x * x * polynomial / (6.0 * (1.0 + x));
}

// (e(after) - e(before)) / (after - before), or p / rho^2 at the midpoint for a change of at most
// 1e-10 rho0, where the secant would lose its digits (sphEnergy.py:162-183).
double elasticSlope(const OracleParams& p, double before, double after) {
    const double change = after - before;
    if (std::fabs(change) > 1e-10 * p.rho0) return (elasticEnergy(p, after) - elasticEnergy(p, before)) / change;
    const double middle = 0.5 * (before + after);
    const double ratio = middle / p.rho0, sq = ratio * ratio;
    double pressure = p.rho0 * p.c0 * p.c0 / 7.0 * (sq * sq * sq * ratio - 1.0);
    if (!p.negativePressure) pressure = std::max(pressure, 0.0);
    return pressure / (middle * middle);
}

// The renormalized density gradient of fluid particle i, L_i sum_j V_j (rho_j - rho_i) grad_i W_ij
// over every neighbor, boundary particles included, from the deviations rho - rho0; zero on the
// boundary (sphGpu.py:548-591, sphSolver.py:691-716).
V3 densityGradientAt(const OracleParams& p, const CellList& cells, const Fields& s, const double* deviation,
                     const M3& li, int i) {
    if (!s.fluid[i]) return {0.0, 0.0, 0.0};
    V3 total = {0.0, 0.0, 0.0};
    forNeighbors(p, cells, s.pos, s.len, i, true, [&](const Pair& q) {
        const double slope = kernelSlope(p.kernel, q.r, q.h, sigmaAt(p, q.h));
        total = total + (s.mass[q.j] / s.dens[q.j] * (deviation[q.j] - deviation[i]) * slope / q.r) * q.d;
    });
    return mul(li, total);
}

// The bed segment k with bedX[k] <= x < bedX[k + 1], clipped (sphGpu.py:1231-1244).
int bedSegment(const OracleBoundary& b, double x) {
    int low = 0, high = b.nBed - 1;
    while (high - low > 1) {
        const int middle = (low + high) / 2;
        if (b.bedX[middle] <= x) low = middle; else high = middle;
    }
    return low;
}

void convolutionWeights(double u, double w[4]) {
    const double u2 = u * u, u3 = u2 * u;
    w[0] = -0.5 * u3 + u2 - 0.5 * u;
    w[1] = 1.5 * u3 - 2.5 * u2 + 1.0;
    w[2] = -1.5 * u3 + 2.0 * u2 + 0.5 * u;
    w[3] = 0.5 * u3 - 0.5 * u2;
}
void convolutionSlopeWeights(double u, double w[4]) {
    const double u2 = u * u;
    w[0] = -1.5 * u2 + 2.0 * u - 0.5;
    w[1] = 4.5 * u2 - 5.0 * u;
    w[2] = -4.5 * u2 + 4.0 * u + 0.5;
    w[3] = 1.5 * u2 - u;
}

// Boundary enforcement on fluid particle i (sphGpu.py:1197-1229, 1246-1310, 1332-1419;
// sphBoundaries.py:129-147, 295-344, 597-640).
void enforceAt(const OracleParams& p, const OracleBoundary& b, double* x, double* v) {
    const double clearance = b.clearance;
    if (b.kind == 1) {
        for (int axis = 0; axis < p.dims; ++axis) {
            if (x[axis] < b.lower[axis]) { x[axis] = b.lower[axis]; v[axis] = std::max(v[axis], 0.0); }
            if (axis < p.dims - 1 || b.openTop == 0) {
                if (x[axis] > b.upper[axis]) { x[axis] = b.upper[axis]; v[axis] = std::min(v[axis], 0.0); }
            }
        }
    } else if (b.kind == 2) {
        const int vertical = p.dims - 1;
        const int k = bedSegment(b, x[0]);
        const double run = b.bedX[k + 1] - b.bedX[k], rise = b.bedZ[k + 1] - b.bedZ[k];
        const double share = std::min(std::max((x[0] - b.bedX[k]) / run, 0.0), 1.0);
        const double bed = b.bedZ[k] + share * rise;
        if (x[vertical] < bed + clearance) {
            x[vertical] = bed + clearance;
            const double slope = rise / run, length = std::sqrt(1.0 + slope * slope);
            const double nx = -slope / length, nz = 1.0 / length;
            const double inward = v[0] * nx + v[vertical] * nz;
            if (inward < 0.0) { v[0] -= inward * nx; v[vertical] -= inward * nz; }
        }
        if (b.sideWalls) {
            if (x[1] < clearance) { x[1] = clearance; v[1] = std::max(v[1], 0.0); }
            if (x[1] > b.width - clearance) { x[1] = b.width - clearance; v[1] = std::min(v[1], 0.0); }
        }
        if (b.offshoreWall && x[0] < b.wallX + clearance) {
            x[0] = b.wallX + clearance;
            v[0] = std::max(v[0], b.pistonSpeed);
        }
        if (x[0] > b.endLimit) { x[0] = b.endLimit; v[0] = std::min(v[0], 0.0); }
    } else if (b.kind == 3) {
        const int nx = b.fieldNx, ny = b.fieldNy;
        const double rawX = (x[0] - b.originX) / b.stepX, rawY = (x[1] - b.originY) / b.stepY;
        const double fx = std::min(std::max(rawX, 0.0), (double)(nx - 1));
        const double fy = std::min(std::max(rawY, 0.0), (double)(ny - 1));
        const int ix = (int)std::floor(fx), iy = (int)std::floor(fy);
        double wx[4], wy[4], sx[4], sy[4];
        convolutionWeights(fx - ix, wx);
        convolutionWeights(fy - iy, wy);
        convolutionSlopeWeights(fx - ix, sx);
        convolutionSlopeWeights(fy - iy, sy);
        double bed = 0.0, slopeX = 0.0, slopeY = 0.0;
        for (int a = 0; a < 4; ++a) {
            const int row = std::min(std::max(iy - 1 + a, 0), ny - 1);
            double along = 0.0, alongSlope = 0.0;
            for (int c = 0; c < 4; ++c) {
                const double value = b.table[(int64_t)row * nx + std::min(std::max(ix - 1 + c, 0), nx - 1)];
                along += wx[c] * value;
                alongSlope += sx[c] * value;
            }
            bed += wy[a] * along;
            slopeX += wy[a] * alongSlope;
            slopeY += sy[a] * along;
        }
        slopeX /= b.stepX;
        slopeY /= b.stepY;
        if (rawX < 0.0 || rawX > (double)(nx - 1)) slopeX = 0.0;
        if (rawY < 0.0 || rawY > (double)(ny - 1)) slopeY = 0.0;
        if (x[2] < bed + clearance) {
            x[2] = bed + clearance;
            const double length = std::sqrt(slopeX * slopeX + slopeY * slopeY + 1.0);
            const double n[3] = {-slopeX / length, -slopeY / length, 1.0 / length};
            const double inward = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
            if (inward < 0.0)
                for (int a = 0; a < 3; ++a) v[a] -= inward * n[a];
        }
        if (x[1] < b.lowY) { x[1] = b.lowY; v[1] = std::max(v[1], 0.0); }
        if (x[1] > b.highY) { x[1] = b.highY; v[1] = std::min(v[1], 0.0); }
        if (b.offshoreWall && x[0] < b.lowX) { x[0] = b.lowX; v[0] = std::max(v[0], 0.0); }
        if (x[0] > b.highX) { x[0] = b.highX; v[0] = std::min(v[0], 0.0); }
    }
}

// cflTimeStep (sphSolver.py:1410-1424).
double cflTimeStep(double h, double c0, double vmax, double amax, double cfl, double maxdt) {
    double dt = cfl * h / (c0 + vmax);
    if (amax > 1e-12) dt = std::min(dt, cfl * std::sqrt(h / amax));
    return std::min(dt, maxdt);
}

}  // namespace

extern "C" {

int oracle_threads(int requested) {
#ifdef _OPENMP
    if (requested > 0) omp_set_num_threads(requested);
    return omp_get_max_threads();
#else
    (void)requested;
    return 1;
#endif
}

// cellOrderKey for every position (sphSolver.py:1379-1408).
void oracle_cell_keys(const OracleParams* p, int n, const double* pos, int64_t* keys) {
    const int64_t nx = p->grid[0], ny = p->grid[1], nz = p->dims == 3 ? p->grid[2] : 1;
    for (int i = 0; i < n; ++i) {
        const double* x = pos + 3 * i;
        const int64_t cx = wrap((int64_t)std::floor(x[0] / p->cell), nx);
        const int64_t cy = wrap((int64_t)std::floor(x[1] / p->cell), ny);
        const int64_t cz = p->dims == 3 ? wrap((int64_t)std::floor(x[2] / p->cell), nz) : 0;
        keys[i] = (cz * ny + cy) * nx + cx;
    }
}

// Neighbor counts (pass 1) or rows (pass 2, `offsets` filled by the caller from the counts) of
// the float64 neighbor set {j != i : r_ij < h_i + h_j}, each row ascending in j.
void oracle_neighbors(const OracleParams* p, int n, const double* pos, const double* lengths, int64_t* counts,
                      const int64_t* offsets, int32_t* rows) {
    CellList cells;
    cells.build(*p, n, pos);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        std::vector<int32_t> row;
        forNeighbors(*p, cells, pos, lengths, i, false, [&](const Pair& q) { row.push_back(q.j); });
        std::sort(row.begin(), row.end());
        if (counts) counts[i] = (int64_t)row.size();
        if (rows) std::copy(row.begin(), row.end(), rows + offsets[i]);
    }
}

// The float32 neighbor predicate of the fast path, evaluated as the device evaluates it: each
// float64 position split into a float32 point and a float32 remainder (sphGpu.py:1010-1019),
// d = (p_i - p_j) + offset + (l_i - l_j) in float32 (sphGpu.py:227-242), r = sqrtf(fmaf(dz, dz,
// fmaf(dy, dy, dx * dx))), h_ij = 0.5f (h_i + h_j), and j a neighbor when r < 2 h_ij. Candidate
// cells come from the float64 cell key, as on the device.
void oracle_neighbors32(const OracleParams* p, int n, const double* pos, const double* lengths, int64_t* counts,
                        const int64_t* offsets, int32_t* rows) {
    CellList cells;
    cells.build(*p, n, pos);
    std::vector<float> pt(3 * (size_t)n), lo(3 * (size_t)n), h32(n);
    for (int i = 0; i < 3 * n; ++i) {
        pt[i] = (float)pos[i];
        lo[i] = (float)(pos[i] - (double)pt[i]);
    }
    for (int i = 0; i < n; ++i) h32[i] = (float)lengths[i];
    const float hmax = (float)p->hmax;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        std::vector<int32_t> row;
        const V3 xi = vec(pos, i);
        const float reach = h32[i] + hmax;
        const int images = (p->period > 0.0 && (pt[3 * i + 1] < reach || pt[3 * i + 1] > (float)p->period - reach)) ? 2 : 1;
        for (int image = 0; image < images; ++image) {
            float shift = 0.0f;
            if (image == 1) shift = pt[3 * i + 1] < 0.5f * (float)p->period ? (float)p->period : -(float)p->period;
            const V3 origin = {xi.x, xi.y + (double)shift, xi.z};
            cells.around(*p, origin, [&](int j) {
                if (j == i) return;
                const float dx = (pt[3 * i] - pt[3 * j]) + 0.0f + (lo[3 * i] - lo[3 * j]);
                const float dy = (pt[3 * i + 1] - pt[3 * j + 1]) + shift + (lo[3 * i + 1] - lo[3 * j + 1]);
                const float dz = (pt[3 * i + 2] - pt[3 * j + 2]) + 0.0f + (lo[3 * i + 2] - lo[3 * j + 2]);
                const float r = std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx)));
                const float h = 0.5f * (h32[i] + h32[j]);
                if (r < 2.0f * h) row.push_back(j);
            });
        }
        std::sort(row.begin(), row.end());
        if (counts) counts[i] = (int64_t)row.size();
        if (rows) std::copy(row.begin(), row.end(), rows + offsets[i]);
    }
}

// Host-step sums: each takes the whole state and writes one result per particle.
void oracle_wall_pressure(const OracleParams* p, int n, const double* pos, const double* vel, const double* dens,
                          const double* press, const double* mass, const double* len, const uint8_t* fluid,
                          double* outPress, double* outDens) {
    CellList cells;
    cells.build(*p, n, pos);
    const Fields s{n, pos, vel, dens, press, mass, len, fluid};
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        if (fluid[i]) { outPress[i] = press[i]; outDens[i] = dens[i]; continue; }
        wallPressureAt(*p, cells, s, i, outPress + i, outDens + i);
    }
}

void oracle_velocity_gradients(const OracleParams* p, int n, const double* pos, const double* vel,
                               const double* dens, const double* mass, const double* len, const uint8_t* fluid,
                               double* out9) {
    CellList cells;
    cells.build(*p, n, pos);
    const Fields s{n, pos, vel, dens, nullptr, mass, len, fluid};
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const M3 g = velocityGradientAt(*p, cells, s, i);
        std::memcpy(out9 + 9 * (size_t)i, g.m, sizeof g.m);
    }
}

void oracle_correction(const OracleParams* p, int n, const double* pos, const double* dens, const double* mass,
                       const double* len, const uint8_t* fluid, double* out9) {
    CellList cells;
    cells.build(*p, n, pos);
    const Fields s{n, pos, nullptr, dens, nullptr, mass, len, fluid};
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const M3 l = correctionAt(*p, cells, s, i);
        std::memcpy(out9 + 9 * (size_t)i, l.m, sizeof l.m);
    }
}

// `gradients9` may be null when the limiter is off.
void oracle_accelerations(const OracleParams* p, int n, const double* pos, const double* vel, const double* dens,
                          const double* press, const double* mass, const double* len, const uint8_t* fluid,
                          const double* gradients9, const OracleClosure* closure, double* outAcc) {
    CellList cells;
    cells.build(*p, n, pos);
    const Fields s{n, pos, vel, dens, press, mass, len, fluid};
    std::vector<M3> g;
    if (p->limiter) {
        g.resize(n);
        for (int i = 0; i < n; ++i) std::memcpy(g[i].m, gradients9 + 9 * (size_t)i, sizeof g[i].m);
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) store(outAcc, i, accelerationAt(*p, cells, s, g.data(), i, closure));
}

// Particle shifting, after Lind et al. (sphGpu.py:856-1004, sphSolver.py:824-1010), on one
// smoothing length h = hmax. Pass 1, on the fluid over every neighbor but the coincident: grad C
// with the anti-clumping weight 1 + 0.2 (W / W(dx))^4, grad C0 without it, div r = -sum V_j r W'(r)
// (set to -1 where the support is lopsided, h |grad C0| above 0.3 for the cubic spline and 0.4 for
// the Wendland C2) and grad rho; a boundary particle's div r is 1e6. Pass 2, on a fluid particle
// whose div r is at least 1.5 (2D) or 2.4 (3D): the surface neighbors' normals -grad C0_j / |grad
// C0_j| weighted by W; the shift -size grad C_i (size times |v_i| when `local`), zero with a surface
// neighbor and h |grad C_i| above 0.03, else projected off the normal; the density change
// shift . grad rho_i and the velocity change sum_j V_j (v_j - v_i) (grad W_ij . shift) over fluid
// neighbors but the coincident.
void oracle_shifting(const OracleParams* p, int n, const double* pos, const double* vel, const double* dens,
                     const double* mass, const uint8_t* fluid, double size, int32_t local, double spacingKernel,
                     double* outShift3, double* outDensity, double* outVelocity3) {
    CellList cells;
    cells.build(*p, n, pos);
    const double h = p->hmax, sigma = sigmaAt(*p, h);
    std::vector<double> len(n, h);
    const double lopsided = p->kernel == 1 ? 0.4 : 0.3, threshold = p->dims == 2 ? 1.5 : 2.4;
    std::vector<V3> gradient(n), plain(n), densityGradient(n);
    std::vector<double> divergence(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        gradient[i] = plain[i] = densityGradient[i] = {0.0, 0.0, 0.0};
        if (!fluid[i]) { divergence[i] = 1.0e6; continue; }
        double total = 0.0;
        forNeighbors(*p, cells, pos, len.data(), i, true, [&](const Pair& q) {
            const double volume = mass[q.j] / dens[q.j];
            const double slope = kernelSlope(p->kernel, q.r, h, sigma);
            const double ratio = kernelValue(p->kernel, q.r, h, sigma) / spacingKernel;
            const double square = ratio * ratio;
            const V3 weighted = (volume * slope / q.r) * q.d;
            plain[i] = plain[i] + weighted;
            gradient[i] = gradient[i] + (1.0 + 0.2 * square * square) * weighted;
            densityGradient[i] = densityGradient[i] + ((dens[q.j] - p->rho0) - (dens[i] - p->rho0)) * weighted;
            total -= volume * slope * q.r;
        });
        divergence[i] = h * std::sqrt(dot(plain[i], plain[i])) > lopsided ? -1.0 : total;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const V3 zero = {0.0, 0.0, 0.0};
        store(outShift3, i, zero);
        store(outVelocity3, i, zero);
        outDensity[i] = 0.0;
        if (!fluid[i] || divergence[i] < threshold) continue;
        bool vicinity = false;
        V3 normal = zero;
        forNeighbors(*p, cells, pos, len.data(), i, false, [&](const Pair& q) {
            if (divergence[q.j] >= threshold) return;
            vicinity = true;
            const double length = std::sqrt(dot(plain[q.j], plain[q.j]));
            if (length > 0.0) normal = normal - (kernelValue(p->kernel, q.r, h, sigma) / length) * plain[q.j];
        });
        const V3 vi = vec(vel, i);
        const double scale = local ? size * std::sqrt(dot(vi, vi)) : size;
        V3 shift = -scale * gradient[i];
        if (vicinity) {
            if (h * std::sqrt(dot(gradient[i], gradient[i])) > 0.03) shift = zero;
            const double weight = std::sqrt(dot(normal, normal));
            if (weight > 0.0) {
                normal = (1.0 / weight) * normal;
                shift = shift - dot(shift, normal) * normal;
            }
        }
        store(outShift3, i, shift);
        outDensity[i] = dot(shift, densityGradient[i]);
        if (dot(shift, shift) > 0.0) {
            V3 change = zero;
            forNeighbors(*p, cells, pos, len.data(), i, true, [&](const Pair& q) {
                if (!fluid[q.j]) return;
                const double slope = kernelSlope(p->kernel, q.r, h, sigma);
                change = change + (mass[q.j] / dens[q.j] * slope / q.r * dot(q.d, shift)) * (vec(vel, q.j) - vi);
            });
            store(outVelocity3, i, change);
        }
    }
}

// The closure's coefficients from each particle's G (9 per particle, row-major): alpha_i =
// 2 (D + 2) (C_S l / h)^2 h ||D|| / c0 and delta_i = (C_delta l / h)^2 h ||D|| / c0, C_S = 0.12,
// C_delta = 1.5, l / h = 2, ||D|| = sqrt(2 D:D) of G's symmetric part, each at most 0.2; zero on the
// boundary (sphGpu.py:699-723, sphSolver.py:276-282).
void oracle_closure_coefficients(const OracleParams* p, int n, const double* gradients9, const double* len,
                                 const uint8_t* fluid, double soundSpeed, double* alphas, double* deltas) {
    const double smagorinsky = (0.12 * 2.0) * (0.12 * 2.0), diffusion = (1.5 * 2.0) * (1.5 * 2.0), cap = 0.2;
    for (int i = 0; i < n; ++i) {
        if (!fluid[i]) { alphas[i] = deltas[i] = 0.0; continue; }
        const double* g = gradients9 + 9 * (size_t)i;
        double dd = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const double strain = 0.5 * (g[3 * a + b] + g[3 * b + a]);
                dd += strain * strain;
            }
        const double scale = len[i] * std::sqrt(2.0 * dd) / soundSpeed;
        alphas[i] = std::min(2.0 * (p->dims + 2) * smagorinsky * scale, cap);
        deltas[i] = std::min(diffusion * scale, cap);
    }
}

// `corrections9` may be null when the correction is off.
void oracle_continuity(const OracleParams* p, int n, const double* pos, const double* vel, const double* dens,
                       const double* mass, const double* len, const uint8_t* fluid, const double* corrections9,
                       const double* gradients3, const OracleClosure* closure, double* outRate) {
    CellList cells;
    cells.build(*p, n, pos);
    const Fields s{n, pos, vel, dens, nullptr, mass, len, fluid};
    std::vector<M3> l;
    if (p->correction) {
        l.resize(n);
        for (int i = 0; i < n; ++i) std::memcpy(l[i].m, corrections9 + 9 * (size_t)i, sizeof l[i].m);
    }
    std::vector<double> deviation(n);
    for (int i = 0; i < n; ++i) deviation[i] = dens[i] - p->rho0;
    std::vector<V3> g;
    if (p->diffusion == 3 && gradients3) {
        g.resize(n);
        for (int i = 0; i < n; ++i) g[i] = vec(gradients3, i);
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        outRate[i] = continuityAt(*p, cells, s, deviation.data(), l.data(), i, g.empty() ? nullptr : g.data(), closure);
}

// Each particle's renormalized density gradient, (N, 3) rows, with the L of `corrections9` (the
// caller's, from oracle_correction).
void oracle_density_gradients(const OracleParams* p, int n, const double* pos, const double* dens, const double* mass,
                              const double* len, const uint8_t* fluid, const double* corrections9, double* out3) {
    CellList cells;
    cells.build(*p, n, pos);
    const Fields s{n, pos, nullptr, dens, nullptr, mass, len, fluid};
    std::vector<double> deviation(n);
    for (int i = 0; i < n; ++i) deviation[i] = dens[i] - p->rho0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        M3 li;
        std::memcpy(li.m, corrections9 + 9 * (size_t)i, sizeof li.m);
        store(out3, i, densityGradientAt(*p, cells, s, deviation.data(), li, i));
    }
}

// The Shepard-filtered density of every particle: sum_j m_j W_ij / sum_j (m_j / rho_j) W_ij over
// every neighbor and the particle itself at W(0, h_i), coincident pairs included (sphGpu.py:768-802,
// sphSolver.py:1023-1043).
void oracle_shepard(const OracleParams* p, int n, const double* pos, const double* dens, const double* mass,
                    const double* len, double* out) {
    CellList cells;
    cells.build(*p, n, pos);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const double own = sigmaAt(*p, len[i]);
        double numerator = mass[i] * own, denominator = mass[i] / dens[i] * own;
        forNeighbors(*p, cells, pos, len, i, false, [&](const Pair& q) {
            const double value = kernelValue(p->kernel, q.r, q.h, sigmaAt(*p, q.h));
            numerator += mass[q.j] * value;
            denominator += mass[q.j] / dens[q.j] * value;
        });
        out[i] = numerator / denominator;
    }
}

// The XSPH transport velocity, (N, 3) rows: v_i + epsilon sum_j m_j W_ij / ((rho_i + rho_j) / 2)
// (v_j - v_i) over the fluid neighbors of a fluid particle, coincident pairs included; a boundary
// particle keeps its velocity (sphGpu.py:725-766, sphSolver.py:1358-1377).
void oracle_xsph(const OracleParams* p, int n, const double* pos, const double* vel, const double* dens,
                 const double* mass, const double* len, const uint8_t* fluid, double epsilon, double* out3) {
    CellList cells;
    cells.build(*p, n, pos);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const V3 vi = vec(vel, i);
        if (!fluid[i]) { store(out3, i, vi); continue; }
        V3 correction = {0.0, 0.0, 0.0};
        forNeighbors(*p, cells, pos, len, i, false, [&](const Pair& q) {
            if (!fluid[q.j]) return;
            const double value = kernelValue(p->kernel, q.r, q.h, sigmaAt(*p, q.h));
            const double weight = mass[q.j] * value / (0.5 * (dens[i] + dens[q.j]));
            correction = correction + weight * (vec(vel, q.j) - vi);
        });
        store(out3, i, vi + epsilon * correction);
    }
}

void oracle_enforce(const OracleParams* p, const OracleBoundary* b, int n, double* pos, double* vel,
                    const uint8_t* fluid) {
    for (int i = 0; i < n; ++i)
        if (fluid[i]) enforceAt(*p, *b, pos + 3 * i, vel + 3 * i);
}

double oracle_cfl_time_step(double h, double c0, double vmax, double amax, double cfl, double maxdt) {
    return cflTimeStep(h, c0, vmax, amax, cfl, maxdt);
}

// Device-resident steps (sphSolver.py:1941-1992 with sphGpu.py:2237-2315) until `*time` reaches
// `untilTime` or `maxSteps` steps are taken. State arrays are updated in place; `press` and `acc`
// hold the last step's pressures and accelerations. Returns the steps taken; `*lastDt` the last
// step's size. `budget` (may be null) receives per-step (vmax, amax, dt) triples.
//
// With `options`: the Shepard filter when due, counted by `*stepCount` (the run's steps before
// this call, advanced by the steps taken), and the closure. With `ledger` (14 values, added to):
// the energy budget's device terms over the fluid (sphEnergy.py deviceTerms; sphGpu.py:1084-1419):
//   0-3   m a_part . (v + v') / 2 dt, a_part the pressure and viscosity from fluid and from
//         boundary neighbors; 4 the same for gravity
//   5-8   m e' R_part dt, R_part the plain continuity over fluid and boundary neighbors, the
//         correction's share and the diffusion, e' the elastic secant slope over the step
//   9     the floor's change of m e;  10  the drift's -m g . (x' - x);  11  the enforcement's change
//         of m (v^2 / 2 - g . x);  12  shifting (not a device-resident option: 0);  13  the filter's
//         change of m e
static int advanceWith(const OracleParams* p, const OracleBoundary* b, const OracleOptions* options, int n, double* pos,
                double* vel, double* dens, double* press, double* acc, const double* mass, const double* len,
                const uint8_t* fluid, double* time, double untilTime, int maxSteps, double* lastDt, double* budget,
                int64_t* stepCount, double* ledger) {
    CellList cells;
    std::vector<M3> gradients(n), corrections(n);
    std::vector<double> deviation(n), rate(n), filtered, alphas, deltas, startDensity;
    std::vector<V3> densityGradients, accParts, startVelocity;
    std::vector<double> rateParts;
    const bool closureOn = options && options->closure;
    const bool antuono = p->diffusion == 3;
    const int interval = options ? options->filterInterval : 0;
    OracleClosure closure{};
    if (closureOn) {
        alphas.resize(n);
        deltas.resize(n);
        closure = OracleClosure{alphas.data(), deltas.data(), options->closureSoundSpeed, options->closureDensity,
                                options->closureWater};
    }
    if (antuono) densityGradients.resize(n);
    if (ledger) {
        accParts.resize(4 * (size_t)n);
        rateParts.resize(4 * (size_t)n);
        startVelocity.resize(n);
        startDensity.resize(n);
    }
    const V3 g = {p->gravity[0], p->gravity[1], p->gravity[2]};
    int steps = 0;
    while (*time < untilTime && steps < maxSteps) {
        // Rates (residentRates)
        cells.build(*p, n, pos);
        const Fields s{n, pos, vel, dens, press, mass, len, fluid};
        const int64_t step = stepCount ? *stepCount : steps;
        if (interval > 0 && step > 0 && step % interval == 0) {
            filtered.resize(n);
            oracle_shepard(p, n, pos, dens, mass, len, filtered.data());
            for (int i = 0; i < n; ++i) {
                if (!fluid[i]) continue;
                if (ledger) ledger[13] += mass[i] * (elasticEnergy(*p, filtered[i]) - elasticEnergy(*p, dens[i]));
                dens[i] = filtered[i];
            }
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) press[i] = tait(*p, dens[i], fluid[i] != 0);
        if (p->extrapolatedWalls) {
            // Reads fluid entries only and writes boundary entries only, as on the device
#pragma omp parallel for schedule(dynamic, 64)
            for (int i = 0; i < n; ++i)
                if (!fluid[i]) wallPressureAt(*p, cells, s, i, press + i, dens + i);
        }
        if (p->limiter || closureOn) {
#pragma omp parallel for schedule(dynamic, 64)
            for (int i = 0; i < n; ++i) gradients[i] = velocityGradientAt(*p, cells, s, i);
        }
        if (closureOn) {
            std::vector<double> flat(9 * (size_t)n);
            for (int i = 0; i < n; ++i) std::memcpy(flat.data() + 9 * (size_t)i, gradients[i].m, sizeof gradients[i].m);
            oracle_closure_coefficients(p, n, flat.data(), len, fluid, options->closureSoundSpeed, alphas.data(),
                                        deltas.data());
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i)
            store(acc, i, accelerationAt(*p, cells, s, gradients.data(), i, closureOn ? &closure : nullptr,
                                         ledger ? accParts.data() + 4 * (size_t)i : nullptr));
        double vmax = 0.0, amax = 0.0;
        for (int i = 0; i < n; ++i) {
            if (!fluid[i]) continue;
            const V3 v = vec(vel, i), a = vec(acc, i);
            vmax = std::max(vmax, std::sqrt(dot(v, v)));
            amax = std::max(amax, std::sqrt(dot(a, a)));
        }
        const double dt = cflTimeStep(p->hmin, p->c0, vmax, amax, p->cfl, p->maxTimeStep);
        if (budget) { budget[3 * steps] = vmax; budget[3 * steps + 1] = amax; budget[3 * steps + 2] = dt; }

        // Integration (residentIntegrate)
        for (int i = 0; i < n; ++i) {
            if (!fluid[i]) continue;
            const V3 start = vec(vel, i), next = start + dt * vec(acc, i);
            store(vel, i, next);
            if (ledger) {
                const V3 mean = 0.5 * (start + next);
                for (int k = 0; k < 4; ++k) ledger[k] += mass[i] * dot(accParts[4 * (size_t)i + k], mean) * dt;
                ledger[4] += mass[i] * dot(g, mean) * dt;
            }
        }
        for (int i = 0; i < n; ++i) deviation[i] = dens[i] - p->rho0;
        if (p->correction || antuono) {
#pragma omp parallel for schedule(dynamic, 64)
            for (int i = 0; i < n; ++i) corrections[i] = correctionAt(*p, cells, s, i);
        }
        if (antuono) {
#pragma omp parallel for schedule(dynamic, 64)
            for (int i = 0; i < n; ++i)
                densityGradients[i] = densityGradientAt(*p, cells, s, deviation.data(), corrections[i], i);
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i)
            rate[i] = continuityAt(*p, cells, s, deviation.data(), corrections.data(), i,
                                   antuono ? densityGradients.data() : nullptr, closureOn ? &closure : nullptr,
                                   ledger ? rateParts.data() + 4 * (size_t)i : nullptr);
        for (int i = 0; i < n; ++i) {
            const double before = dens[i];
            double rho = dens[i] + rate[i] * dt;
            const double unfloored = rho;
            if (p->densityFloor > 0.0 && fluid[i]) rho = std::max(rho, p->densityFloor);
            dens[i] = rho;
            if (ledger && fluid[i]) {
                const double slope = elasticSlope(*p, before, unfloored);
                for (int k = 0; k < 4; ++k) ledger[5 + k] += mass[i] * slope * rateParts[4 * (size_t)i + k] * dt;
                ledger[9] += mass[i] * (elasticEnergy(*p, rho) - elasticEnergy(*p, unfloored));
            }
        }
        for (int i = 0; i < n; ++i) {
            if (!fluid[i]) continue;
            const V3 start = vec(pos, i);
            store(pos, i, start + dt * vec(vel, i));
            if (ledger) ledger[10] -= mass[i] * dot(g, vec(pos, i) - start);
            if (p->period > 0.0) pos[3 * i + 1] -= p->period * std::floor(pos[3 * i + 1] / p->period);
            const V3 x0 = vec(pos, i), v0 = vec(vel, i);
            enforceAt(*p, *b, pos + 3 * i, vel + 3 * i);
            if (ledger) {
                const V3 x1 = vec(pos, i), v1 = vec(vel, i);
                ledger[11] += mass[i] * (0.5 * (dot(v1, v1) - dot(v0, v0)) - dot(g, x1 - x0));
            }
        }
        *time += dt;
        *lastDt = dt;
        ++steps;
        if (stepCount) ++*stepCount;
    }
    return steps;
}


int oracle_advance(const OracleParams* p, const OracleBoundary* b, int n, double* pos, double* vel, double* dens,
                   double* press, double* acc, const double* mass, const double* len, const uint8_t* fluid,
                   double* time, double untilTime, int maxSteps, double* lastDt, double* budget) {
    return advanceWith(p, b, nullptr, n, pos, vel, dens, press, acc, mass, len, fluid, time, untilTime, maxSteps,
                       lastDt, budget, nullptr, nullptr);
}

// oracle_advance with the resident options and the energy ledger (either may be null).
int oracle_advance_options(const OracleParams* p, const OracleBoundary* b, const OracleOptions* options, int n,
                           double* pos, double* vel, double* dens, double* press, double* acc, const double* mass,
                           const double* len, const uint8_t* fluid, double* time, double untilTime, int maxSteps,
                           double* lastDt, int64_t* stepCount, double* ledger) {
    return advanceWith(p, b, options, n, pos, vel, dens, press, acc, mass, len, fluid, time, untilTime, maxSteps,
                       lastDt, nullptr, stepCount, ledger);
}

// The elastic energy per unit mass at each density (for the ledger's checks).
void oracle_elastic_energy(const OracleParams* p, int n, const double* dens, double* out) {
    for (int i = 0; i < n; ++i) out[i] = elasticEnergy(*p, dens[i]);
}

}  // extern "C"

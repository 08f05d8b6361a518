// Device-side building blocks of the delta-SPH engine: the run description, the sorted particle
// records the pair sweeps read, the smoothing kernels and the cell-stencil walk.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_hip.h"

namespace sph {

// One run as the kernels read it: the ABI's sph_run plus the constants the kernels derive once.
struct RunDev {
    int64_t first, count;     // slots, and sorted positions: the sort keeps each run's range
    int64_t cellFirst;        // the run's first cell in the context's cell array
    int32_t nx, ny, nz;
    int32_t kernel;
    int32_t diffusion;
    int32_t extrapolatedWalls, negativePressure, correction, limiter;
    int32_t uniform;          // every particle has the longest smoothing length
    int32_t filterInterval;   // the Shepard filter every this many steps (0 never)
    int32_t closure;          // the context's delta-LES-SPH closure is on: G on every fluid particle
    double cell, hmax, hmin, period;
    double rho0, c0, alpha, boundaryFactor;
    double gravity[3];
    double delta, densityFloor, cfl, maxTimeStep;
    double sigmaUnit;         // kernel normalization at h = 1
    double stiffness;         // Tait B = rho0 c0^2 / 7
    sph_boundary boundary;
};

// Per-run clock and step reductions, in device memory.
struct Clock {
    double* time;
    double* until;
    double* dt;
    int64_t* steps;
    int64_t* maxSteps;
    unsigned long long* maxima;   // 2 per run: |v|max and |a|max as float64 bit patterns
    int32_t* active;
    // why a run stopped by itself: 0 none, 1 its state stopped being finite, 2 its time step collapsed
    // (set by the step's maxima; a failed run takes no further device-clocked step until its clock
    // is set again)
    int32_t* failed;
    double* speed;                // per sorted slot: |v| of a fluid particle at the gather, else 0
    double* acceleration;         // per sorted slot: |a| of a fluid particle, else 0
};

// Whether run r takes the Shepard filter this step: `mode` 0 never, 1 every active run (a filtered
// step asked from the host), 2 a run whose interval divides its step count above zero
// (sphSolver.py:1979-1981), the count before this step's.
__device__ __forceinline__ bool filterDue(const RunDev& run, const Clock& clock, int r, int mode) {
    if (mode == 0 || !clock.active[r]) return false;
    if (mode == 1) return true;
    const int64_t steps = clock.steps[r];
    return run.filterInterval > 0 && steps > 0 && steps % run.filterInterval == 0;
}

template <typename Real>
struct alignas(16) Quad {
    Real x, y, z, w;
};

// One sorted particle as the pair sweeps read it: four quads, 64 bytes in float32 (one cache
// line, so a pair reads one line) and 128 in float64. For float32 pairs each float64 position is
// carried as a float32 point and a float32 remainder (sphGpu.py:1010-1019); in float64 the
// remainder is zero and never read. The fields sit so that a pair's reads take whole quads:
//
//   3D: q0 point xyz, signed length   q1 remainder xyz, p / rho^2   q2 velocity xyz, density
//       q3 mass, m / rho, rho - rho0, pressure
//   2D: q0 point xy, remainder xy     q1 velocity xy, density, signed length
//       q2 mass, m / rho, rho - rho0, p / rho^2                      q3 pressure
//
// The smoothing length is negated on boundary particles; rho - rho0 is taken in float64
// (sphGpu.py:1031-1038).
template <typename Real>
struct Record {
    Quad<Real> q0, q1, q2, q3;
};

// Every field of a record, unpacked.
template <typename Real>
struct Fields {
    Real p[3];
    float l[3];
    Real sh;      // signed smoothing length
    Real v[3];
    Real rho, mass, vol, dev, pr2, press;
};

template <int D, typename Real>
__host__ __device__ __forceinline__ Record<Real> pack(const Fields<Real>& f) {
    Record<Real> r;
    if constexpr (D == 2) {
        r.q0 = {f.p[0], f.p[1], (Real)f.l[0], (Real)f.l[1]};
        r.q1 = {f.v[0], f.v[1], f.rho, f.sh};
        r.q2 = {f.mass, f.vol, f.dev, f.pr2};
        r.q3 = {f.press, Real(0), Real(0), Real(0)};
    } else {
        r.q0 = {f.p[0], f.p[1], f.p[2], f.sh};
        r.q1 = {(Real)f.l[0], (Real)f.l[1], (Real)f.l[2], f.pr2};
        r.q2 = {f.v[0], f.v[1], f.v[2], f.rho};
        r.q3 = {f.mass, f.vol, f.dev, f.press};
    }
    return r;
}

// Field readers of one record, by dimension.
template <int D, typename Real>
__device__ __forceinline__ Real signedLength(const Record<Real>& r) { return D == 2 ? r.q1.w : r.q0.w; }
template <int D, typename Real>
__device__ __forceinline__ Real densityOf(const Record<Real>& r) { return D == 2 ? r.q1.z : r.q2.w; }
template <int D, typename Real>
__device__ __forceinline__ Real massOf(const Record<Real>& r) { return D == 2 ? r.q2.x : r.q3.x; }
template <int D, typename Real>
__device__ __forceinline__ Real volumeOf(const Record<Real>& r) { return D == 2 ? r.q2.y : r.q3.y; }
template <int D, typename Real>
__device__ __forceinline__ Real deviationOf(const Record<Real>& r) { return D == 2 ? r.q2.z : r.q3.z; }
template <int D, typename Real>
__device__ __forceinline__ Real pressureRatioOf(const Record<Real>& r) { return D == 2 ? r.q2.w : r.q1.w; }
template <int D, typename Real>
__device__ __forceinline__ Real pressureOf(const Record<Real>& r) { return D == 2 ? r.q3.x : r.q3.w; }
template <int D, typename Real>
__device__ __forceinline__ void velocityOf(const Record<Real>& r, Real (&v)[3]) {
    if constexpr (D == 2) {
        const Quad<Real> q = r.q1;
        v[0] = q.x;
        v[1] = q.y;
        v[2] = Real(0);
    } else {
        const Quad<Real> q = r.q2;
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
    }
}

// Field writers of one record: the wall pressure's results on a boundary particle (with rho - rho0
// of the new density, taken in float64, as the reference's deviations are taken after the wall
// pressure, sphGpu.py:2287), and the kicked velocity the continuity sweep reads.
template <int D, typename Real>
__device__ __forceinline__ void setWall(Record<Real>& r, Real pressure, Real rho, Real deviation) {
    const Real pr2 = pressure / (rho * rho);
    if constexpr (D == 2) {
        r.q1.z = rho;
        r.q2.y = r.q2.x / rho;
        r.q2.z = deviation;
        r.q2.w = pr2;
        r.q3.x = pressure;
    } else {
        r.q1.w = pr2;
        r.q2.w = rho;
        r.q3.y = r.q3.x / rho;
        r.q3.z = deviation;
        r.q3.w = pressure;
    }
}
template <int D, typename Real>
__device__ __forceinline__ void setVelocity(Record<Real>& r, const Real (&v)[3]) {
    if constexpr (D == 2) {
        r.q1.x = v[0];
        r.q1.y = v[1];
    } else {
        r.q2.x = v[0];
        r.q2.y = v[1];
        r.q2.z = v[2];
    }
}

// The sorted records, with the particle's cell and the per-particle matrices each sweep writes
// for the next, D*D values per particle, row-major.
template <typename Real>
struct Records {
    Record<Real>* r;
    int32_t* key;    // the particle's cell, in the context's cell array
    Real* G;         // velocity gradient; in a host step of the renormalized diffusion, the density gradient (D values)
    Real* L;         // kernel gradient correction
    Real* A;         // acceleration, D values per particle (zero on the boundary), for the kick
};

__device__ __forceinline__ int wrapIndex(int64_t a, int32_t n) {
    int64_t r = a % n;
    return (int)(r < 0 ? r + n : r);
}

// c wrapped onto [0, n) for -n <= c < 2n, as a neighbor cell one step from a cell in range is:
// a select instead of a division.
__device__ __forceinline__ int wrapNear(int c, int32_t n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// The distinct cell offsets along an axis of n cells, b = 0 .. stencilCount(n) - 1: -1, 0, 1, or
// 0, 1 for two cells and 0 for one (fewer than three cells would repeat one). Computed rather than
// kept in an array, which the compiler would place in scratch memory.
__device__ __forceinline__ int stencilCount(int32_t n) { return n >= 3 ? 3 : n; }
__device__ __forceinline__ int stencilOffset(int32_t n, int b) { return n >= 3 ? b - 1 : b; }

// floor(v / cell) as an integer cell coordinate; a non-finite coordinate reads as cell 0.
__device__ __forceinline__ int64_t cellCoordinate(double v, double cell) {
    const double c = floor(v / cell);
    if (!(c == c) || c > 4.0e18 || c < -4.0e18) return 0;
    return (int64_t)c;
}

// Calls g(begin, end, middle, right, rowY, rowZ, wrapped) for the sorted ranges of the 3x3 (2D) or
// 3x3x3 (3D) cells around cell (cx, cy, cz), x fastest, the three cells of an unwrapped row as one
// range: [begin, middle) its left cell, [middle, right) its middle one and [right, end) its right
// one, each ascending in x. A cell taken alone comes as g(begin, end, begin, end, ...). (rowY,
// rowZ) are the row's cell coordinates before wrapping, and `wrapped` whether either wrapped.
template <int D, class G>
__device__ __forceinline__ void forStencil(const RunDev& run, const int32_t* __restrict__ start, int cx, int cy, int cz,
                                           G&& g) {
    const int32_t nx = run.nx, ny = run.ny, nz = run.nz;
    const int ky = stencilCount(ny);
    const int kz = D == 3 ? stencilCount(nz) : 1;
    for (int a = 0; a < kz; ++a) {
        const int rowZ = D == 3 ? cz + stencilOffset(nz, a) : 0;
        const int64_t plane = D == 3 ? (int64_t)wrapNear(rowZ, nz) : 0;
        for (int b = 0; b < ky; ++b) {
            const int rowY = cy + stencilOffset(ny, b);
            const int wrappedY = wrapNear(rowY, ny);
            const bool wrapped = wrappedY != rowY || (D == 3 && plane != rowZ);
            const int64_t row = run.cellFirst + (plane * ny + wrappedY) * nx;
            if (nx >= 3 && cx >= 1 && cx <= nx - 2) {
                g(start[row + cx - 1], start[row + cx + 2], start[row + cx], start[row + cx + 1], rowY, rowZ, wrapped);
            } else {
                const int kx = stencilCount(nx);
                for (int c = 0; c < kx; ++c) {
                    const int64_t cellIndex = row + wrapNear(cx + stencilOffset(nx, c), nx);
                    const int32_t begin = start[cellIndex], end = start[cellIndex + 1];
                    g(begin, end, begin, end, rowY, rowZ, wrapped);
                }
            }
        }
    }
}

// The candidates of a 3D row that can be neighbors: the left cell's from the first whose point x is
// at least x_i - reach, the right cell's up to the last at most x_i + reach (each cell ascending in
// x), the middle cell whole. `reach` is 2 h_max with a margin over the points' rounding, so every
// pair with r < h_i + h_j stays.
template <typename Real>
__device__ __forceinline__ Real rowReach(Real twoHmax, Real px) {
    return twoHmax * Real(1.0 + 1.0 / 1024.0) + fabs(px) * Real(1.0 / 2097152.0);
}

template <typename Real, class X>
__device__ __forceinline__ int32_t firstAtLeast(int32_t lo, int32_t hi, Real bound, X&& x) {
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (x(mid) < bound) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <typename Real, class X>
__device__ __forceinline__ int32_t firstAbove(int32_t lo, int32_t hi, Real bound, X&& x) {
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (x(mid) <= bound) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The lower edge of a particle's own cell along one axis: its cell coordinate c is the unwrapped
// floor(p / cell) taken modulo the n cells of the axis, so the unwrapped one is c + k n with k the
// nearest integer to (p / cell - c - 1/2) / n (within 1/2 - 1/(2n) of k for n >= 3).
template <typename Real>
__device__ __forceinline__ Real cellLowerEdge(Real p, int c, int32_t n, Real cell) {
    const Real turns = rint((p / cell - (Real)c - Real(0.5)) / (Real)n);
    return ((Real)c + turns * (Real)n) * cell;
}

// The reach along x left in a row of cells at cell coordinates (rowY, rowZ), for a particle at
// point (py, pz) of cell (cy, cz): the particle's distance across to the row's cells takes its share,
// reach^2 - gap_y^2 - gap_z^2, each gap less a margin over the float32 points' rounding (so no pair
// within reach is lost); below zero no candidate of the row can be a neighbor, and it returns -1. A
// wrapped row, a periodic run or an axis of fewer than three cells (where one row of cells lies on
// both sides of the particle) keeps the whole reach. In 2D rows differ in y alone.
template <int D, typename Real>
__device__ __forceinline__ Real rowReachAcross(const RunDev& run, Real reach, Real py, Real pz, int cy, int cz, int rowY,
                                               int rowZ, bool wrapped) {
    if (wrapped || run.period > 0.0 || run.ny < 3 || (D == 3 && run.nz < 3) || (rowY == cy && rowZ == cz)) return reach;
    const Real cell = (Real)run.cell;
    const Real margin = cell * Real(1.0 / 4096.0) + (fabs(py) + fabs(pz)) * Real(1.0 / 1048576.0);
    Real gy = Real(0), gz = Real(0);
    if (rowY != cy) {
        const Real low = cellLowerEdge(py, cy, run.ny, cell);
        gy = rowY < cy ? py - low : low + cell - py;
    }
    if (D == 3 && rowZ != cz) {
        const Real low = cellLowerEdge(pz, cz, run.nz, cell);
        gz = rowZ < cz ? pz - low : low + cell - pz;
    }
    gy = fmax(gy - margin, Real(0));
    gz = fmax(gz - margin, Real(0));
    const Real left = reach * reach - gy * gy - gz * gz;
    return left > Real(0) ? sqrt(left) : Real(-1);
}

// A 3D row's candidates [from, to) within `across` of x: the left cell's from the first whose point
// x is at least x_i - across and the right cell's up to the last at most x_i + across (each cell
// ascending in x), the middle cell whole; with a reduced reach the whole row is searched when it
// ascends across its cells too, as it does while the particles lie within the grid's extent.
template <typename Real, class X>
__device__ __forceinline__ void rowCandidates(int32_t begin, int32_t end, int32_t middle, int32_t right, Real px,
                                              Real reach, Real across, X&& pointX, int32_t& from, int32_t& to) {
    if (across < Real(0)) {
        from = to = begin;
        return;
    }
    const bool ascending = across < reach && (middle == begin || middle == end || pointX(middle - 1) <= pointX(middle)) &&
                           (right == begin || right == end || pointX(right - 1) <= pointX(right));
    if (ascending) {
        from = firstAtLeast(begin, end, px - across, pointX);
        to = firstAbove(from, end, px + across, pointX);
    } else {
        from = firstAtLeast(begin, middle, px - across, pointX);
        to = firstAbove(right, end, px + across, pointX);
    }
}

// EighthTable: each cell's slots, ordered by x, split by the eighth of the cell's width along x
// their x lies in. Byte k (k = 1 .. 7) of a cell's table is how many of its slots lie in its first
// k eighths, so eighth k's slots are [start + byte k, start + byte k+1) with byte 0 = 0 and byte 8
// the cell's count; kEighthsUnknown (byte 0 nonzero) marks a cell whose table does not hold: one of
// its slots lies outside the grid's cells along x (its eighth is not ordered with the others') or
// it holds more than 255. Built by adding kEighthsAbove << 8e for each slot of eighth e.
constexpr uint8_t kOutsideX = 0x80;
constexpr unsigned long long kEighthsAbove = 0x0101010101010100ull;
constexpr unsigned long long kEighthsUnknown = 0xffull;

// The sorted position where eighth k (0 .. 8) of a cell [begin, end) starts; an unknown table
// answers the cell's begin (`late` false) or its end (`late` true), the safe side for each use.
__device__ __forceinline__ int32_t eighthStart(unsigned long long table, int32_t begin, int32_t end, int k, bool late) {
    if (k <= 0) return begin;
    if (k >= 8) return end;
    if (table & 0xffull) return late ? end : begin;
    const int32_t at = begin + (int32_t)((table >> (8 * k)) & 0xffull);
    return at < end ? at : end;
}

// Where eighth e (0 .. 24) of a row of three cells starts: eighths 0-7 are the left cell's
// [begin, middle), 8-15 the middle cell's [middle, right), 16-23 the right cell's [right, end), and
// 24 is the row's end; `tables` holds the three cells' tables in that order.
__device__ __forceinline__ int32_t rowEighthStart(int e, bool late, int32_t begin, int32_t middle, int32_t right,
                                                  int32_t end, const unsigned long long (&tables)[3]) {
    if (e >= 24) return end;
    const int c = e >> 3;
    const int32_t b = c == 0 ? begin : (c == 1 ? middle : right);
    const int32_t f = c == 0 ? middle : (c == 1 ? right : end);
    const unsigned long long table = c == 0 ? tables[0] : (c == 1 ? tables[1] : tables[2]);
    return eighthStart(table, b, f, e & 7, late);
}

// The candidates of an unwrapped row of three cells (left cell [begin, middle), middle
// [middle, right), right [right, end), x ascending within each) that can lie within `across` of x
// along x, from whole eighths of the cells: every slot of an eighth that ends before px - across
// or starts after px + across is left out, and nothing else. Slots of the left cell lie left of
// the middle cell's and those of the right cell right of it, except slots from outside the grid
// along x, which lie at least the grid's width minus two cells away and are no neighbors. `rowLeft`
// is the left cell's lower edge as the particle's own cell coordinate places it, `tables` the three
// cells' tables. With `Refine`, a binary search within the first and last eighths then bounds the
// range exactly as firstAtLeast and firstAbove would over the row (the 3D list build refines).
template <bool Refine, typename Real, class X>
__device__ __forceinline__ void rowEighths(int32_t begin, int32_t middle, int32_t right, int32_t end, Real px,
                                           Real across, Real rowLeft, Real cell, const unsigned long long (&tables)[3],
                                           X&& pointX, int32_t& from, int32_t& to) {
    if (across < Real(0)) {
        from = to = begin;
        return;
    }
    // a sixteenth of an eighth, and the float rounding of x, keep the eighths' cuts on the safe side
    const Real scale = Real(8) / cell;
    const Real margin = Real(1.0 / 16.0) + (fabs(px) + cell) * scale * Real(1.0 / 1048576.0);
    const Real low = px - across, high = px + across;
    const Real lowAt = floor((low - rowLeft) * scale - margin);
    const Real highAt = floor((high - rowLeft) * scale + margin);
    const int first = lowAt < Real(0) ? 0 : (lowAt > Real(24) ? 24 : (int)lowAt);
    const int last = highAt < Real(-1) ? -1 : (highAt > Real(23) ? 23 : (int)highAt);
    int32_t lo = rowEighthStart(first, false, begin, middle, right, end, tables);
    int32_t hi = last < first ? lo : rowEighthStart(last + 1, true, begin, middle, right, end, tables);
    if constexpr (Refine) {
        if (hi > lo) {
            lo = firstAtLeast(lo, rowEighthStart(first + 1, true, begin, middle, right, end, tables), low, pointX);
            const int32_t lastFrom = rowEighthStart(last, false, begin, middle, right, end, tables);
            hi = firstAbove(lastFrom > lo ? lastFrom : lo, hi, high, pointX);
        }
    }
    from = lo;
    to = hi < lo ? lo : hi;
}

template <int D, typename Real>
struct Self {
    Real p[3];
    float l[3];
    Real h;
    Real reach;
    bool fluid;
    int32_t cell;   // the particle's cell within its run's grid
};

// A run-local cell index as (x, y, z) cell coordinates, x fastest.
__device__ __forceinline__ void cellCoordinates(const RunDev& run, int32_t cell, int& cx, int& cy, int& cz) {
    const uint32_t c = (uint32_t)cell, nx = (uint32_t)run.nx, ny = (uint32_t)run.ny;
    const uint32_t row = c / nx;
    cx = (int)(c - row * nx);
    cy = (int)(row % ny);
    cz = (int)(row / ny);
}

// d = r_i - r_j for particle i against record j, with `shift` added along y for j's periodic
// image: (p_i - p_j) + shift + (l_i - l_j) on float32 points and remainders, the plain float64
// difference otherwise.
template <int D, typename Real>
__device__ __forceinline__ void separation(const Self<D, Real>& me, const Record<Real>& r, Real shift, Real& dx, Real& dy,
                                           Real& dz) {
    if constexpr (D == 2) {
        const Quad<Real> q = r.q0;
        if constexpr (sizeof(Real) == 4) {
            dx = (me.p[0] - q.x) + (me.l[0] - q.z);
            dy = (me.p[1] - q.y) + shift + (me.l[1] - q.w);
        } else {
            dx = me.p[0] - q.x;
            dy = (me.p[1] + shift) - q.y;
        }
        dz = Real(0);
    } else {
        const Quad<Real> a = r.q0;
        if constexpr (sizeof(Real) == 4) {
            const Quad<Real> b = r.q1;
            dx = (me.p[0] - a.x) + (me.l[0] - b.x);
            dy = (me.p[1] - a.y) + shift + (me.l[1] - b.y);
            dz = (me.p[2] - a.z) + (me.l[2] - b.z);
        } else {
            dx = me.p[0] - a.x;
            dy = (me.p[1] + shift) - a.y;
            dz = me.p[2] - a.z;
        }
    }
}

// fma(dz, dz, fma(dy, dy, dx * dx)): the squared separation, every fma written out so that every
// caller rounds it alike.
__device__ __forceinline__ float squaredDistance(float dx, float dy, float dz, int dims) {
    float r2 = dx * dx;
    r2 = fmaf(dy, dy, r2);
    if (dims == 3) r2 = fmaf(dz, dz, r2);
    return r2;
}
__device__ __forceinline__ double squaredDistance(double dx, double dy, double dz, int dims) {
    double r2 = dx * dx;
    r2 = fma(dy, dy, r2);
    if (dims == 3) r2 = fma(dz, dz, r2);
    return r2;
}

// The neighbor predicate's distance: the square root of squaredDistance, correctly rounded.
__device__ __forceinline__ float pairDistance(float dx, float dy, float dz, int dims) {
    return sqrtf(squaredDistance(dx, dy, dz, dims));
}
__device__ __forceinline__ double pairDistance(double dx, double dy, double dz, int dims) {
    return sqrt(squaredDistance(dx, dy, dz, dims));
}

// Whether pairDistance < t for a pair of squared distance r2, as the rounded square root decides
// it. Outside a band of t^2 (1 -+ 2^-20) (2^-48 in double) the comparison of r2 with the band's
// edges already decides it: the roundings of t^2, of the edges and of the square root together
// shift the comparison by at most a quarter of 2^-20 (2^-48), so r2 below the band gives a root
// below t and r2 above it a root above t. Inside the band, and for a NaN, the root itself is
// compared.
template <typename Real>
__device__ __forceinline__ bool rootBelow(Real r2, Real t) {
    // an empty volatile asm keeps the rare rounded root in a branch of its own: the compiler would
    // otherwise compute it for every pair and select
    __asm__ volatile("" : "+v"(r2));
    return sqrt(r2) < t;
}
// withinSupport for one support t, its band's edges computed once: holds(r2) == withinSupport(r2, t).
template <typename Real>
struct SupportBand {
    Real t, below, above;
    __device__ __forceinline__ explicit SupportBand(Real support) : t(support) {
        constexpr Real band = sizeof(Real) == 4 ? Real(9.5367431640625e-07) : Real(3.552713678800501e-15);
        const Real t2 = support * support;
        below = t2 * (Real(1) - band);
        above = t2 * (Real(1) + band);
    }
    __device__ __forceinline__ bool holds(Real r2) const {
        if (r2 < below) return true;
        if (r2 > above) return false;
        return rootBelow(r2, t);
    }
};
template <typename Real>
__device__ __forceinline__ bool withinSupport(Real r2, Real t) {
    return SupportBand<Real>(t).holds(r2);
}

// The 3D float32 predicate from the points alone, for candidates of one length: the separation
// without the neighbor's remainder, d' = (p_i - p_j) + shift + l_i, differs from the exact one,
// d = (p_i - p_j) + shift + (l_i - l_j), by |l_j| (at most an ulp of |p_j| / 2, below 2^-23 P with
// P = |p_i|_1 + |shift| + 8 t bounding |p_j| on every axis for a candidate within 4 t) and the
// roundings of the two sums, so for r' = |d'| >= r_t one has |r^2 - r'^2| <= (7 e P / r' + 6 e) r'^2
// with e = 2^-24, and for r' < t, |r^2 - r'^2| <= 7 e P t + 6 e t^2. `slack` (2^-16 + 16 e P / t)
// covers both with room. decide() answers 1 (r^2 is below the band, a neighbor), -1 (above it, not
// one) or 0 (too close to call: the caller reads the remainder and tests exactly), so its answers
// are the exact predicate's wherever it gives one.
template <typename Real>
struct CoarseSupport {
    Real in, outScale, above;
    template <int D>
    __device__ __forceinline__ CoarseSupport(const SupportBand<Real>& band, const Self<D, Real>& me, Real shift) {
        const Real t = band.t;
        const Real bound = fabs(me.p[0]) + fabs(me.p[1]) + fabs(me.p[2]) + fabs(shift) + Real(8) * t;
        const Real slack = Real(1.0 / 65536.0) + Real(16.0 / 16777216.0) * bound / t;
        in = band.below - slack * t * t;
        outScale = Real(1) - slack;
        above = band.above;
    }
    template <int D>
    __device__ __forceinline__ int decide(const Self<D, Real>& me, const Record<Real>& other, Real shift) const {
        const Quad<Real> a = other.q0;
        const Real dx = (me.p[0] - a.x) + (Real)me.l[0];
        const Real dy = ((me.p[1] - a.y) + shift) + (Real)me.l[1];
        const Real dz = (me.p[2] - a.z) + (Real)me.l[2];
        const Real r2 = squaredDistance(dx, dy, dz, 3);
        if (r2 < in) return 1;
        if (r2 * outScale > above) return -1;
        return 0;
    }
};

// The distance a pair sum uses once its pair is known, from its squared distance: in float32 the
// hardware square root (within 1 ulp; the sums are float32 anyway), in float64 the correctly
// rounded one. Every sum takes its pair's distance from here, so the list, the cell walk and the
// list build's fused sweep see the same value; membership is decided apart, by withinSupport.
__device__ __forceinline__ float rootOf(float r2) { return __builtin_amdgcn_sqrtf(r2); }
__device__ __forceinline__ double rootOf(double r2) { return sqrt(r2); }
template <typename Real>
__device__ __forceinline__ Real sumDistance(Real dx, Real dy, Real dz, int dims) {
    return rootOf(squaredDistance(dx, dy, dz, dims));
}

// Entry m's offset in a column-major list of `stride` rows: both are below 2^31, so one 32 x 32 ->
// 64-bit multiply.
__device__ __forceinline__ int64_t listOffset(int32_t m, int64_t stride) {
    return (int64_t)((uint64_t)(uint32_t)m * (uint64_t)(uint32_t)stride);
}

// a / b in the pair sums: in float32 a times the hardware reciprocal of b (within about 2 ulp of
// the rounded quotient; the sums carry float32 rounding throughout), in float64 correctly rounded.
__device__ __forceinline__ float quotient(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double quotient(double a, double b) { return a / b; }

// The smoothing kernels (sphGpu.py:139-186) at one run's constants: W and dW/dr / r, with a fast
// path for runs whose particles share one smoothing length, where every pair is at h = h_max.
template <typename Real, int D>
struct PairKernel {
    int kind;
    bool uniform;
    Real unit;       // sigma at h = 1
    Real h;          // h_max
    Real twoH;       // the support at h_max
    Real invH;
    Real sigma;      // sigma at h_max
    Real slopeC;     // Wendland: -5 sigma / h^2; cubic: sigma / h^2
    Real soft;       // 0.01 h^2

    __device__ __forceinline__ void init(int kernelKind, double sigmaUnit, double hmax, bool uniformLengths) {
        kind = kernelKind;
        uniform = uniformLengths;
        unit = (Real)sigmaUnit;
        h = (Real)hmax;
        twoH = Real(2) * h;
        invH = Real(1) / h;
        sigma = D == 2 ? unit / (h * h) : unit / (h * h * h);
        slopeC = (kind == 1 ? Real(-5) : Real(1)) * sigma / (h * h);
        soft = Real(0.01) * h * h;
    }
    __device__ __forceinline__ Real sigmaAt(Real hij) const {
        return D == 2 ? unit / (hij * hij) : unit / (hij * hij * hij);
    }
    __device__ __forceinline__ Real value(Real r, Real hij) const {
        Real q, s;
        if (uniform) { q = r * invH; s = sigma; } else { q = r / hij; s = sigmaAt(hij); }
        if (kind == 1) {
            const Real rest = Real(1) - Real(0.5) * q, sq = rest * rest;
            return q < Real(2) ? s * sq * sq * (Real(2) * q + Real(1)) : Real(0);
        }
        if (q < Real(1)) return s * (Real(1) - Real(1.5) * q * q + Real(0.75) * q * q * q);
        if (q < Real(2)) {
            const Real rest = Real(2) - q;
            return s * Real(0.25) * rest * rest * rest;
        }
        return Real(0);
    }
    // dW/dr divided by r: grad_i W_ij = slopeOverR * (r_i - r_j).
    __device__ __forceinline__ Real slopeOverR(Real r, Real hij) const {
        Real q, c, inv;
        if (uniform) { q = r * invH; c = slopeC; inv = invH; } else {
            inv = Real(1) / hij;
            q = r * inv;
            c = (kind == 1 ? Real(-5) : Real(1)) * sigmaAt(hij) * inv * inv;
        }
        if (kind == 1) {
            const Real rest = Real(1) - Real(0.5) * q;
            return q < Real(2) ? c * rest * rest * rest : Real(0);
        }
        if (q < Real(1)) return c * (Real(-3) + Real(2.25) * q);
        if (q < Real(2)) {
            const Real rest = Real(2) - q;
            return c * Real(-0.75) * rest * rest * quotient(Real(1), q);
        }
        return Real(0);
    }
};

// Where the neighbor walks read a neighbor's record: straight from the sorted records...
template <typename Real>
struct GlobalRecords {
    const Record<Real>* r;
    __device__ __forceinline__ Record<Real> operator()(int32_t k) const { return r[k]; }
    __device__ __forceinline__ Real pointX(int32_t k) const { return r[k].q0.x; }
};

// A sweep workgroup's dynamic LDS, which holds its NearCells copy, and a record there as an LDS
// address: reads through it stay LDS reads rather than merging with a global read of the same
// record into one generic read.
extern __shared__ __align__(16) unsigned char sweepShared[];
template <typename Real>
using Vec4 = Real __attribute__((ext_vector_type(4)));
template <typename Real>
using LdsVec4 = __attribute__((address_space(3))) Vec4<Real>;

template <typename Real>
__device__ __forceinline__ void storeLds(LdsVec4<Real>* to, const Quad<Real>& q) {
    *to = Vec4<Real>{q.x, q.y, q.z, q.w};
}
template <typename Real>
__device__ __forceinline__ Quad<Real> loadLds(const LdsVec4<Real>* from) {
    const Vec4<Real> v = *from;
    return Quad<Real>{v.x, v.y, v.z, v.w};
}
// The workgroup's copy in LDS is a structure of arrays: quad q of the record at LDS index `index`
// sits at quad position q * stride + index, `stride` being the copy's capacity in records. A wave
// reading one field of many records then reads consecutive 16-byte slots, spread over the banks,
// where 64-byte records would map that field of every record onto a quarter of them.
template <typename Real>
__device__ __forceinline__ Quad<Real> ldsQuad(int32_t index, int q, int32_t stride) {
    return loadLds((const LdsVec4<Real>*)sweepShared + q * stride + index);
}
template <typename Real>
__device__ __forceinline__ Real ldsValue(int32_t index, int q, int c, int32_t stride) {
    using LdsReal = __attribute__((address_space(3))) const Real;
    return ((LdsReal*)sweepShared)[4 * (q * stride + index) + c];
}

// ...or from a workgroup's copy in LDS of the records of the cells around its particles. A tile's
// particles are consecutive sorted particles of one run, so they lie in the run's cells c0..c1
// (cells numbered x fastest), and their neighbors lie in the cells c0 - 1 + o .. c1 + 1 + o for the
// stencil's row offsets o (3 rows in 2D, 9 in 3D): contiguous ranges of sorted records, merged
// where they overlap. The workgroup copies them in one pass with all its loads in flight, and a
// walk then reads a neighbor from LDS when it lies in a copied range, from the records otherwise
// (a periodic image, a wrapped cell, or a range past the copy's capacity).
template <int D, typename Real>
struct NearCells {
    static constexpr int kRows = D == 2 ? 3 : 9;
    int32_t lo[kRows], hi[kRows];   // sorted ranges [lo, hi), ascending and disjoint, some empty
    int32_t at[kRows];              // LDS index of lo
    int32_t held;                   // records copied, at most the capacity
    int32_t stride;                 // the capacity: quads per field array of the copy
    const Record<Real>* r;
    // With `kicked`, every record read takes the kick v + a dt on a fluid particle's velocity, as the
    // continuity sums read the velocities after the kick: `A` is each sorted particle's
    // acceleration and `dt` the step.
    bool kicked;
    const Real* A;
    Real dt;
    // With `withG` (2D), the copy also holds each record's velocity gradient G (4 values, from `G`)
    // as a fifth quad, which the acceleration sweep's limiter reads with the record.
    bool withG = false;
    const Real* G = nullptr;

    __device__ __forceinline__ void kick(Record<Real>& record, int32_t k) const {
        if (!kicked || !(signedLength<D, Real>(record) > Real(0))) return;
        Real v[3];
        velocityOf<D, Real>(record, v);
#pragma unroll
        for (int a = 0; a < D; ++a) v[a] = v[a] + A[(int64_t)k * D + a] * dt;
        setVelocity<D, Real>(record, v);
    }

    // The ranges around sorted particles s0..s1 of `run`, and the first `capacity` of their
    // records to copy. Every value is the same across the workgroup.
    __device__ __forceinline__ void plan(const RunDev& run, const int32_t* __restrict__ start, const int32_t* key,
                                         int64_t s0, int64_t s1, int32_t capacity) {
        const int64_t cells = (int64_t)run.nx * run.ny * run.nz;
        const int64_t c0 = (int64_t)key[s0] - run.cellFirst, c1 = (int64_t)key[s1] - run.cellFirst;
        const int64_t plane = (int64_t)run.nx * run.ny;
        int64_t covered = -1;   // the last cell an earlier range holds
        int32_t total = 0;
#pragma unroll
        for (int row = 0; row < kRows; ++row) {
            const int64_t offset = D == 2 ? (row - 1) * (int64_t)run.nx : (row / 3 - 1) * plane + (row % 3 - 1) * (int64_t)run.nx;
            int64_t low = c0 + offset - 1, high = c1 + offset + 1;
            low = low <= covered ? covered + 1 : low;
            low = low < 0 ? 0 : low;
            high = high > cells - 1 ? cells - 1 : high;
            if (low > high) {
                lo[row] = hi[row] = 0;
            } else {
                lo[row] = start[run.cellFirst + low];
                hi[row] = start[run.cellFirst + high + 1];
                covered = high;
            }
            at[row] = total;
            total += hi[row] - lo[row];
        }
        held = total < capacity ? total : capacity;
        stride = capacity;
    }

    // The sorted position of the record at LDS index `index`.
    __device__ __forceinline__ int32_t sourceOf(int32_t index) const {
        int32_t from = lo[0] + index;
#pragma unroll
        for (int j = 1; j < kRows; ++j)
            if (index >= at[j]) from = lo[j] + (index - at[j]);
        return from;
    }

    // Copies the held records into the workgroup's LDS, two records per thread in flight; every
    // thread of the workgroup calls it, and it ends at a barrier.
    __device__ __forceinline__ void fill() const {
        LdsVec4<Real>* shared = (LdsVec4<Real>*)sweepShared;
        const int32_t step = (int32_t)blockDim.x;
        for (int32_t first = (int32_t)threadIdx.x; first < held; first += 2 * step) {
            const int32_t second = first + step;
            const bool both = second < held;
            const int32_t fromA = sourceOf(first), fromB = sourceOf(both ? second : first);
            Record<Real> a = r[fromA];
            Record<Real> b = r[fromB];
            kick(a, fromA);
            kick(b, fromB);
            Quad<Real> ga{}, gb{};
            if (withG) {
                ga = *(const Quad<Real>*)(G + (int64_t)fromA * 4);
                gb = *(const Quad<Real>*)(G + (int64_t)fromB * 4);
            }
            LdsVec4<Real>* to = shared + first;
            storeLds(to, a.q0);
            storeLds(to + stride, a.q1);
            storeLds(to + 2 * stride, a.q2);
            storeLds(to + 3 * stride, a.q3);
            if (withG) storeLds(to + 4 * stride, ga);
            if (both) {
                to = shared + second;
                storeLds(to, b.q0);
                storeLds(to + stride, b.q1);
                storeLds(to + 2 * stride, b.q2);
                storeLds(to + 3 * stride, b.q3);
                if (withG) storeLds(to + 4 * stride, gb);
            }
        }
        __syncthreads();
    }

    // The LDS index of `begin` when all of [begin, end) is copied, else -1.
    __device__ __forceinline__ int32_t rangeBase(int32_t begin, int32_t end) const {
        int32_t base = -1;
#pragma unroll
        for (int i = 0; i < kRows; ++i)
            if (begin >= lo[i] && end <= hi[i]) base = at[i] + (begin - lo[i]);
        return base >= 0 && base + (end - begin) <= held ? base : -1;
    }

    // The point x of sorted particle k, from LDS when copied there.
    __device__ __forceinline__ Real pointX(int32_t k) const {
        int32_t index = -1;
#pragma unroll
        for (int i = 0; i < kRows; ++i)
            if (k >= lo[i] && k < hi[i]) index = at[i] + (k - lo[i]);
        return index >= 0 && index < held ? ldsValue<Real>(index, 0, 0, stride) : r[k].q0.x;
    }

    // The 2D velocity gradient G of sorted particle k, from the copy when it holds it.
    __device__ __forceinline__ void gradient(int32_t k, Real (&g)[4]) const {
        int32_t index = -1;
#pragma unroll
        for (int i = 0; i < kRows; ++i)
            if (k >= lo[i] && k < hi[i]) index = at[i] + (k - lo[i]);
        Quad<Real> q;
        if (withG && index >= 0 && index < held) q = ldsQuad<Real>(index, 4, stride);
        else q = *(const Quad<Real>*)(G + (int64_t)k * 4);
        g[0] = q.x;
        g[1] = q.y;
        g[2] = q.z;
        g[3] = q.w;
    }

    // The record of sorted particle k, from LDS when copied there.
    __device__ __forceinline__ Record<Real> operator()(int32_t k) const {
        const LdsVec4<Real>* shared = (const LdsVec4<Real>*)sweepShared;
        int32_t index = -1;
#pragma unroll
        for (int i = 0; i < kRows; ++i)
            if (k >= lo[i] && k < hi[i]) index = at[i] + (k - lo[i]);
        Record<Real> out;
        if (index >= 0 && index < held) {
            const LdsVec4<Real>* from = shared + index;
            out.q0 = loadLds(from);
            out.q1 = loadLds(from + stride);
            out.q2 = loadLds(from + 2 * stride);
            out.q3 = loadLds(from + 3 * stride);
        } else {
            out = r[k];
            kick(out, k);
        }
        return out;
    }
};

// Walks every neighbor j != i of sorted particle s with r < 2 h_ij (sphGpu.py:352-363), calling
// f(k, other, dx, dy, dz, r, h) with k the neighbor's sorted position, other its record as `src`
// reads it, d = r_i - r_j at the nearest periodic image, r = |d| and h = h_ij = (h_i + h_j) / 2.
// Each cell range's candidates are dealt to `Lanes` lanes in turn, lane `lane` taking the range's
// candidates lane, lane + Lanes, ..., so a group of lanes shares one particle's walk. With
// `SkipCoincident` pairs closer than 1e-12 are skipped too, as every sum but the wall pressure skips
// them. `y64` is the particle's float64 y, read only for a search from its periodic image.
template <int D, typename Real, int Lanes, bool SkipCoincident, class Src, class F>
__device__ __forceinline__ void forNeighbors(const RunDev& run, const int32_t* __restrict__ start, const Src& src,
                                             int64_t s, const Self<D, Real>& me, double y64, int lane, bool uniform,
                                             F&& f) {
    int cx, cy, cz;
    cellCoordinates(run, me.cell, cx, cy, cz);
    const Real period = (Real)run.period;
    const Real twoHmax = Real(2) * (Real)run.hmax;
    int images = 1;
    if (run.period > 0.0 && (me.p[1] < me.reach || me.p[1] > period - me.reach)) images = 2;
    for (int image = 0; image < images; ++image) {
        Real shift = Real(0);
        int oy = cy;
        if (image == 1) {
            shift = me.p[1] < Real(0.5) * period ? period : -period;
            oy = wrapIndex(cellCoordinate(y64 + (double)shift, run.cell), run.ny);
        }
        const Real reach = rowReach(twoHmax, me.p[0]);
        forStencil<D>(run, start, cx, oy, cz, [&](int32_t begin, int32_t end, int32_t middle, int32_t right, int rowY,
                                                   int rowZ, bool wrapped) {
            auto pointX = [&](int32_t k) { return src.pointX(k); };
            int32_t from = begin, to = end;
            if (D == 3)
                rowCandidates(begin, end, middle, right, me.p[0], reach,
                              rowReachAcross<D>(run, reach, me.p[1], me.p[2], cy, cz, rowY, rowZ, wrapped || image == 1),
                              pointX, from, to);
            for (int32_t k = from + lane; k < to; k += Lanes) {
                if (k == s) continue;
                const Record<Real> other = src(k);
                Real dx, dy, dz;
                separation<D, Real>(me, other, shift, dx, dy, dz);
                const Real r2 = squaredDistance(dx, dy, dz, D);
                Real h, support;
                if (uniform) {
                    h = (Real)run.hmax;
                    support = twoHmax;
                } else {
                    h = Real(0.5) * (me.h + fabs(signedLength<D, Real>(other)));
                    support = Real(2) * h;
                }
                if (!withinSupport(r2, support)) continue;
                const Real r = rootOf(r2);
                if (SkipCoincident && r < Real(1.0e-12)) continue;
                f(k, other, dx, dy, dz, r, h);
            }
        });
    }
}

// Each sorted particle's neighbors for one step, column-major: entry m of particle s at
// index[m * stride + s], the neighbor's sorted position with bit 31 set when it is reached through
// the periodic image. count[s] is the particle's whole count, which may exceed the capacity; such
// a particle's sums walk its cells instead.
struct NeighborList {
    int32_t* index;
    int32_t* count;
    int64_t stride;
    int32_t capacity;
};

constexpr int32_t kImageBit = (int32_t)0x80000000u;

// List entries each lane loads ahead in the sweeps, so their gathers overlap.
#ifndef SPH_LIST_AHEAD
#define SPH_LIST_AHEAD 2
#endif
constexpr int kAhead = SPH_LIST_AHEAD;

// Calls f(k, other, dx, dy, dz, r, h), r from sumDistance, for the neighbors of sorted particle s
// that lane `lane` of its group takes: neighbor m, in the order forNeighbors visits them, goes to
// lane m % Lanes. From the list when the particle's `count` neighbors fit it, from its cells
// otherwise; the same pairs reach the same lanes in the same order either way.
template <int D, typename Real, int Lanes, bool SkipCoincident, class Src, class F>
__device__ __forceinline__ void forListedNeighbors(const RunDev& run, const int32_t* __restrict__ start, const Src& src,
                                                   const NeighborList& list, int64_t s, int32_t count,
                                                   const Self<D, Real>& me, double y64, int lane, bool uniform, F&& f) {
    if (count >= 0 && count <= list.capacity) {
        const Real period = (Real)run.period;
        const Real shiftImage = me.p[1] < Real(0.5) * period ? period : -period;
        const int32_t* column = list.index + s;
        auto visit = [&](int32_t entry) {
            const int32_t k = entry & 0x7fffffff;
            const Real shift = entry < 0 ? shiftImage : Real(0);
            const Record<Real> other = src(k);
            Real dx, dy, dz;
            separation<D, Real>(me, other, shift, dx, dy, dz);
            const Real r = sumDistance(dx, dy, dz, D);
            if (SkipCoincident && r < Real(1.0e-12)) return;
            const Real h = uniform ? (Real)run.hmax : Real(0.5) * (me.h + fabs(signedLength<D, Real>(other)));
            f(k, other, dx, dy, dz, r, h);
        };
        // kAhead entries of a lane in flight at once; each lane still takes its entries in order
        int32_t m = lane;
        for (; m + (kAhead - 1) * Lanes < count; m += kAhead * Lanes) {
            int32_t entry[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) entry[u] = column[listOffset(m + u * Lanes, list.stride)];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) visit(entry[u]);
        }
        for (; m < count; m += Lanes) visit(column[listOffset(m, list.stride)]);
        return;
    }
    int32_t m = 0;
    forNeighbors<D, Real, 1, false>(run, start, src, s, me, y64, 0, uniform,
                                    [&](int32_t k, const Record<Real>& other, Real dx, Real dy, Real dz, Real, Real h) {
                                        const bool mine = (m % Lanes) == lane;
                                        ++m;
                                        if (!mine) return;
                                        const Real r = sumDistance(dx, dy, dz, D);
                                        if (SkipCoincident && r < Real(1.0e-12)) return;
                                        f(k, other, dx, dy, dz, r, h);
                                    });
}

// forListedNeighbors with the count the list holds for the particle; without a list every sum
// walks the cells.
template <int D, typename Real, int Lanes, bool SkipCoincident, class Src, class F>
__device__ __forceinline__ void forEachNeighbor(const RunDev& run, const int32_t* __restrict__ start, const Src& src,
                                                const NeighborList& list, int64_t s, const Self<D, Real>& me,
                                                double y64, int lane, bool uniform, F&& f) {
    const int32_t count = list.capacity > 0 ? list.count[s] : -1;
    forListedNeighbors<D, Real, Lanes, SkipCoincident>(run, start, src, list, s, count, me, y64, lane, uniform, f);
}

// A workgroup's per-particle sums in LDS, after its NearCells copy of `cached` records: each group
// of lanes leaves its particle's sums there, and after a barrier thread p of the workgroup finishes
// particle p. Closing arithmetic that is long against the walk (the continuity sweep's float64
// integration and boundary enforcement) then runs on the fewest waves, rather than on one lane of
// every group while the group's other lanes wait.
template <typename Real>
using LdsReal = __attribute__((address_space(3))) Real;
template <typename Real>
__device__ __forceinline__ LdsReal<Real>* packedSums(int32_t cached) {
    return (LdsReal<Real>*)(sweepShared + (size_t)cached * sizeof(Record<Real>));
}

// The list build's hits kept in LDS after its NearCells copy of `cached` records: bit 30 marks an
// entry that is the LDS index of the neighbor's record in the copy, else it is the neighbor's
// sorted position; bit 31 marks the periodic image, as in the list.
constexpr int32_t kLdsBit = 0x40000000;
using LdsHits = __attribute__((address_space(3))) int32_t*;
template <typename Real>
__device__ __forceinline__ LdsHits ldsHits(int32_t cached) {
    return (LdsHits)(sweepShared + (size_t)cached * sizeof(Record<Real>));
}

// The pair sweeps' workgroups: each covers up to blockDim / Lanes consecutive sorted particles of
// one run, so the run and everything read from its description are the same across the
// workgroup and live in scalar registers.
struct Tiles {
    const int32_t* run;     // run of each workgroup
    const int64_t* first;   // first sorted particle of each workgroup
    int64_t count;
};

// The workgroup's place in the launch with the blocks of each XCD made contiguous: an MI355X deals
// workgroups to its 8 XCDs in turn, each XCD with its own L2, while neighboring blocks of sorted
// particles read the same neighbor records. Renumbering so XCD x runs the x-th contiguous slice
// keeps those shared reads in one L2.
__device__ __forceinline__ int64_t xcdBlock() {
    constexpr unsigned kXcds = 8;
    const unsigned blocks = gridDim.x, b = blockIdx.x;
    const unsigned xcd = b % kXcds, index = b / kXcds;
    const unsigned per = blocks / kXcds, extra = blocks % kXcds;
    return (int64_t)xcd * per + (xcd < extra ? xcd : extra) + index;
}

// The sum of v over a group of `Lanes` consecutive lanes; every lane of the wave must call it.
template <int Lanes, typename T>
__device__ __forceinline__ T groupSum(T v) {
#pragma unroll
    for (int o = Lanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace sph

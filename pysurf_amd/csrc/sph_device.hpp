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
};

// The float32 (or float64) records of the sorted particles the sweeps gather from. For float32
// pairs each float64 position is carried as a float32 point and a float32 remainder
// (sphGpu.py:1010-1019); for float64, `pos` is the position itself and `lo` is unused.
template <typename Real>
struct Records {
    Real* pos[3];
    float* lo[3];
    Real* vel[3];
    Real* rho;       // density copy
    Real* press;     // pressure
    Real* dev;       // rho - rho0, taken in float64 (sphGpu.py:1031-1038)
    Real* mass;
    Real* len;       // smoothing length
    uint8_t* fluid;
    int32_t* key;    // the particle's cell, in the context's cell array
    Real* G;         // velocity gradient, D*D rows
    Real* L;         // kernel gradient correction, D*D rows
    int64_t stride;
};

__device__ __forceinline__ int wrapIndex(int64_t a, int32_t n) {
    int64_t r = a % n;
    return (int)(r < 0 ? r + n : r);
}

// The distinct cell offsets along an axis of n cells; fewer than three cells would repeat one.
__device__ __forceinline__ int stencilOffsets(int32_t n, int* out) {
    if (n >= 3) { out[0] = -1; out[1] = 0; out[2] = 1; return 3; }
    if (n == 2) { out[0] = 0; out[1] = 1; return 2; }
    out[0] = 0;
    return 1;
}

// floor(v / cell) as an integer cell coordinate; a non-finite coordinate reads as cell 0.
__device__ __forceinline__ int64_t cellCoordinate(double v, double cell) {
    const double c = floor(v / cell);
    if (!(c == c) || c > 4.0e18 || c < -4.0e18) return 0;
    return (int64_t)c;
}

// Calls g(begin, end) for the sorted ranges of the 3x3 (2D) or 3x3x3 (3D) cells around cell
// (cx, cy, cz), x fastest, the three cells of an unwrapped row as one range.
template <int D, class G>
__device__ __forceinline__ void forStencil(const RunDev& run, const int32_t* __restrict__ start, int cx, int cy, int cz,
                                           G&& g) {
    const int32_t nx = run.nx, ny = run.ny, nz = run.nz;
    int oy[3], oz[3];
    const int ky = stencilOffsets(ny, oy);
    int kz = 1;
    oz[0] = 0;
    if (D == 3) kz = stencilOffsets(nz, oz);
    for (int a = 0; a < kz; ++a) {
        const int64_t plane = D == 3 ? (int64_t)wrapIndex(cz + oz[a], nz) : 0;
        for (int b = 0; b < ky; ++b) {
            const int64_t row = run.cellFirst + (plane * ny + wrapIndex(cy + oy[b], ny)) * nx;
            if (nx >= 3 && cx >= 1 && cx <= nx - 2) {
                g(start[row + cx - 1], start[row + cx + 2]);
            } else {
                int ox[3];
                const int kx = stencilOffsets(nx, ox);
                for (int c = 0; c < kx; ++c) {
                    const int64_t cellIndex = row + wrapIndex(cx + ox[c], nx);
                    g(start[cellIndex], start[cellIndex + 1]);
                }
            }
        }
    }
}

// The smoothing kernels (sphGpu.py:139-186): W and dW/dr / r at smoothing length h.
template <typename Real, int D>
struct Smoothing {
    int kind;
    Real unit;      // sigma at h = 1

    __device__ __forceinline__ Real sigma(Real h) const {
        return D == 2 ? unit / (h * h) : unit / (h * h * h);
    }
    __device__ __forceinline__ Real value(Real r, Real h) const {
        const Real q = r / h, s = sigma(h);
        if (kind == 1) {
            if (q < Real(2)) {
                const Real rest = Real(1) - Real(0.5) * q, sq = rest * rest;
                return s * sq * sq * (Real(2) * q + Real(1));
            }
            return Real(0);
        }
        if (q < Real(1)) return s * (Real(1) - Real(1.5) * q * q + Real(0.75) * q * q * q);
        if (q < Real(2)) {
            const Real rest = Real(2) - q;
            return s * Real(0.25) * rest * rest * rest;
        }
        return Real(0);
    }
    // dW/dr divided by r: the factor grad_i W_ij = factor * (r_i - r_j) takes.
    __device__ __forceinline__ Real slopeOverR(Real r, Real h) const {
        const Real q = r / h, s = sigma(h);
        if (kind == 1) {
            if (q < Real(2)) {
                const Real rest = Real(1) - Real(0.5) * q;
                return Real(-5) * s * rest * rest * rest / (h * h);
            }
            return Real(0);
        }
        if (q < Real(1)) return s * (Real(-3) + Real(2.25) * q) / (h * h);
        if (q < Real(2)) {
            const Real rest = Real(2) - q;
            return s * (Real(-0.75) * rest * rest) / (h * r);
        }
        return Real(0);
    }
};

template <int D, typename Real>
struct Self {
    Real p[3];
    float l[3];
    Real h;
    Real reach;
};

// sqrt(fma(dz, dz, fma(dy, dy, dx * dx))), the neighbor predicate's distance.
__device__ __forceinline__ float pairDistance(const float* d, int dims) {
    float r2 = d[0] * d[0];
    r2 = fmaf(d[1], d[1], r2);
    if (dims == 3) r2 = fmaf(d[2], d[2], r2);
    return sqrtf(r2);
}
__device__ __forceinline__ double pairDistance(const double* d, int dims) {
    double r2 = d[0] * d[0];
    r2 = fma(d[1], d[1], r2);
    if (dims == 3) r2 = fma(d[2], d[2], r2);
    return sqrt(r2);
}

// Walks every neighbor j != i of sorted particle s with r < 2 h_ij (sphGpu.py:352-363), calling
// f(k, d, r, h) with k the neighbor's sorted position, d = r_i - r_j at the nearest periodic
// image, r = |d| and h = h_ij = (h_i + h_j) / 2. With `skipCoincident` pairs closer than 1e-12 are
// skipped too, as every sum but the wall pressure skips them. `y64` is the particle's float64 y,
// read only for a search from its periodic image.
template <int D, typename Real, bool SkipCoincident, class F>
__device__ __forceinline__ void forNeighbors(const RunDev& run, const int32_t* __restrict__ start,
                                             const Records<Real>& rec, int64_t s, const Self<D, Real>& me,
                                             int cx, int cy, int cz, double y64, F&& f) {
    constexpr bool single = sizeof(Real) == 4;
    const Real period = (Real)run.period;
    int images = 1;
    if (run.period > 0.0 && (me.p[1] < me.reach || me.p[1] > period - me.reach)) images = 2;
    for (int image = 0; image < images; ++image) {
        Real shift = Real(0);
        int oy = cy;
        if (image == 1) {
            shift = me.p[1] < Real(0.5) * period ? period : -period;
            oy = wrapIndex(cellCoordinate(y64 + (double)shift, run.cell), run.ny);
        }
        forStencil<D>(run, start, cx, oy, cz, [&](int32_t begin, int32_t end) {
            for (int32_t k = begin; k < end; ++k) {
                if (k == s) continue;
                Real d[3];
                if constexpr (single) {
                    d[0] = (me.p[0] - rec.pos[0][k]) + (me.l[0] - rec.lo[0][k]);
                    d[1] = (me.p[1] - rec.pos[1][k]) + shift + (me.l[1] - rec.lo[1][k]);
                    if (D == 3) d[2] = (me.p[2] - rec.pos[2][k]) + (me.l[2] - rec.lo[2][k]);
                } else {
                    d[0] = me.p[0] - rec.pos[0][k];
                    d[1] = (me.p[1] + shift) - rec.pos[1][k];
                    if (D == 3) d[2] = me.p[2] - rec.pos[2][k];
                }
                if (D == 2) d[2] = Real(0);
                const Real r = pairDistance(d, D);
                const Real h = Real(0.5) * (me.h + rec.len[k]);
                if (!(r < Real(2) * h)) continue;
                if (SkipCoincident && r < Real(1.0e-12)) continue;
                f(k, d, r, h);
            }
        });
    }
}

}  // namespace sph

// The float64 global sums of a run's fluid particles, shared by the context's sph_global_sums and
// the context-free sph_state_sums: one workgroup of kSumThreads threads sums one chunk of
// kSumChunk slots, each thread over a fixed stride and then a fixed tree, and a run's sums add its
// chunks' partials in chunk order. Both entry points run this same arithmetic in the same order,
// so they agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_hip.h"

namespace sph {

constexpr int kSums = 7;              // fluid mass, kinetic, potential, elastic energy, momentum x, y, z
constexpr int64_t kSumChunk = 8192;   // slots per partial sum and per partial maximum
constexpr int kSumThreads = 256;      // threads of a chunk's workgroup

// The elastic energy per unit mass the Tait equation stores at density rho (sphEnergy.py:120-160,
// sphGpu.py:271-282): e = (c0^2 / 7) x^2 (21 + 35x + 35x^2 + 21x^3 + 7x^4 + x^5) / (6 (1 + x)) with
// x = rho / rho0 - 1, zero below rho0 unless the fluid carries tension.
__device__ __forceinline__ double elasticEnergy(double rho, double rho0, double scale, int tension) {
    const double x = rho / rho0 - 1.0;
    const double polynomial = 21.0 + x * (35.0 + x * (35.0 + x * (21.0 + x * (7.0 + x))));
    const double energy = scale * x * x * polynomial / (6.0 * (1.0 + x));
    return (!tension && x < 0.0) ? 0.0 : energy;
}

// The secant slope of e between two densities, (e(after) - e(before)) / (after - before), or, for a
// change of at most 1e-10 rho0 where the secant would lose its digits, p / rho^2 at the midpoint
// (sphEnergy.py:162-183, sphGpu.py:284-301): a density change times it is the change of e.
__device__ __forceinline__ double elasticSlope(double before, double after, double rho0, double scale, int tension) {
    const double change = after - before;
    if (fabs(change) > 1e-10 * rho0)
        return (elasticEnergy(after, rho0, scale, tension) - elasticEnergy(before, rho0, scale, tension)) / change;
    const double middle = 0.5 * (before + after), ratio = middle / rho0, square = ratio * ratio;
    double pressure = scale * rho0 * (square * square * square * ratio - 1.0);
    if (!tension) pressure = fmax(pressure, 0.0);
    return pressure / (middle * middle);
}

// The physics one chunk's sums read.
struct SumPhysics {
    double rho0, c0;
    double fall;        // g, the downward gravity along the last axis
    int32_t tension;    // negative pressure allowed: elastic energy below rho0 counts
};

// The partials of slots [first, end) into partial[0..kSums): fluid only. K = m |v|^2 / 2 and
// P = m g z (sphParticles.py:236-255), masses as the device stores them. Every thread of the
// workgroup (kSumThreads of them) calls it; `tree` is the workgroup's kSums x kSumThreads LDS.
template <int D, typename Real>
__device__ __forceinline__ void chunkSums(const SumPhysics& run, int64_t first, int64_t end, const sph_state& state,
                                          double (*tree)[kSumThreads], double* __restrict__ partial) {
    const double scale = run.c0 * run.c0 / 7.0;
    const Real* mass = static_cast<const Real*>(state.mass);
    double acc[kSums] = {};
    for (int64_t i = first + threadIdx.x; i < end; i += kSumThreads) {
        if (!state.fluid[i]) continue;
        const double m = (double)mass[i];
        double v2 = 0.0;
        for (int a = 0; a < D; ++a) {
            const double v = state.velocity[a * state.stride + i];
            v2 += v * v;
            acc[4 + a] += m * v;
        }
        acc[0] += m;
        acc[1] += 0.5 * m * v2;
        acc[2] += m * run.fall * state.position[(D - 1) * state.stride + i];
        acc[3] += m * elasticEnergy(state.density[i], run.rho0, scale, run.tension);
    }
    for (int e = 0; e < kSums; ++e) tree[e][threadIdx.x] = acc[e];
    __syncthreads();
    for (int width = kSumThreads / 2; width > 0; width >>= 1) {
        if ((int)threadIdx.x < width)
            for (int e = 0; e < kSums; ++e) tree[e][threadIdx.x] += tree[e][threadIdx.x + width];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int e = 0; e < kSums; ++e) partial[e] = tree[e][0];
}

// A run's sums from its chunks' partials [chunk0, chunk1), added in chunk order.
__device__ __forceinline__ void addChunks(const double* __restrict__ partial, int64_t chunk0, int64_t chunk1,
                                          double* __restrict__ out) {
    double acc[kSums] = {};
    for (int64_t b = chunk0; b < chunk1; ++b)
        for (int e = 0; e < kSums; ++e) acc[e] += partial[b * kSums + e];
    for (int e = 0; e < kSums; ++e) out[e] = acc[e];
}

}  // namespace sph

// The MI355X delta-SPH engine: kernels of the device-resident step and the C ABI of
// include/sph_hip.h.
//
// A step of every active run (sphGpu.py:2237-2315, reordered for one device pass per phase):
//
//   clock      each run's active flag (time < until, steps < maxSteps), maxima cleared
//   cells      cellOrderKey per slot (sphSolver.py:1379-1408), a counting sort by cell whose
//              cells are then ordered by slot, so the order is the stable sort's and repeatable
//   gather     float32 records in cell order: point + remainder (sphGpu.py:1010-1019), velocity,
//              density, rho - rho0, Tait pressure (sphGpu.py:1058-1070); |v|max per run
//   sweep 1    boundary lanes: extrapolated wall pressure (sphGpu.py:804-854);
//              fluid lanes: velocity gradient G (sphGpu.py:643-697)
//   sweep 2    fluid lanes: accelerations with the limited viscosity (sphGpu.py:303-414) and the
//              correction matrix L (sphGpu.py:593-641, positions and densities only, so it can
//              share this sweep); |a|max per run
//   kick       dt from the run's maxima (sphSolver.py:1410-1424), v += a dt (sphGpu.py:1084-1109)
//   sweep 3    continuity with the delta-SPH term (sphGpu.py:416-545), then, per particle, the
//              density advance with its floor (:1111-1140), the drift (:1142-1186) and the
//              boundary enforcement (:1197-1229, :1246-1310, :1332-1419)
//
// Every per-particle result is a gather over the particle's own neighbors, written by its own
// thread: no atomics touch a sum, so a run's results do not depend on the other runs it is
// batched with or on the launch shape. The maxima use atomicMax on float64 bit patterns, which
// is order-independent.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "sph_device.hpp"
#include "sph_sums.hpp"

using namespace sph;

// Waves each SIMD keeps resident in the pair sweeps: the register budget the compiler plans for
// (512 VGPRs / waves). The sweeps wait on dependent gathers, so more waves hide more latency.
#ifndef SPH_WAVES_PER_EU
#define SPH_WAVES_PER_EU 4
#endif
// The acceleration sweep, the longest-lived waves, gains from six per SIMD (measured on MI355X).
#ifndef SPH_WAVES_ACCELERATIONS
#define SPH_WAVES_ACCELERATIONS 6
#endif

// The float32 2D sweeps fit those waves without spilling; the float64 and 3D sweeps need more
// registers than that leaves, so they plan for one wave and take what they need.
template <int D, typename Real>
constexpr int sweepWaves() { return (D == 2 && sizeof(Real) == 4) ? SPH_WAVES_PER_EU : 1; }
template <int D, typename Real>
constexpr int accelerationWaves() { return (D == 2 && sizeof(Real) == 4) ? SPH_WAVES_ACCELERATIONS : 1; }
#ifndef SPH_NEAR_BYTES_3D
#define SPH_NEAR_BYTES_3D 32768
#endif
#ifndef SPH_NEAR_PER_3D
#define SPH_NEAR_PER_3D 12
#endif
#define SWEEP_BOUNDS(waves) __attribute__((amdgpu_flat_work_group_size(1, 256), amdgpu_waves_per_eu(waves)))

// Whether a 3D row's range from whole eighths is then cut exactly by a search within its first
// and last eighth (rowEighths).
#ifndef SPH_REFINE_3D
#define SPH_REFINE_3D 1
#endif

// Whether the sweeps copy the records around their particles into LDS (NearCells).
#ifndef SPH_NEAR_CELLS
#define SPH_NEAR_CELLS 1
#endif

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// ---------------------------------------------------------------------------------------------
// Wave and block scans
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ int32_t waveInclusiveScan(int32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan over a block of kScanThreads threads; `total` receives the block's sum.
__device__ __forceinline__ int32_t blockExclusiveScan(int32_t v, int32_t* waves, int32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kWaves = kScanThreads / 64;
    const int32_t inclusive = waveInclusiveScan(v);
    if (lane == 63) waves[wave] = inclusive;
    __syncthreads();
    if (wave == 0) {
        int32_t w = lane < kWaves ? waves[lane] : 0;
        w = waveInclusiveScan(w);
        if (lane < kWaves) waves[lane] = w;
    }
    __syncthreads();
    const int32_t prefix = wave > 0 ? waves[wave - 1] : 0;
    total = waves[kWaves - 1];
    __syncthreads();
    return prefix + inclusive - v;
}

__global__ void __launch_bounds__(kScanThreads) scanTiles(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                          int32_t* __restrict__ partial, int64_t n) {
    __shared__ int32_t waves[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        sum += v[k];
    }
    int32_t total;
    int32_t running = blockExclusiveScan(sum, waves, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = running;
        running += v[k];
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One block: exclusive scan of the tile totals in place; out[n] takes the grand total.
__global__ void __launch_bounds__(kScanThreads) scanPartials(int32_t* __restrict__ partial, int64_t tiles,
                                                             int32_t* __restrict__ out, int64_t n) {
    __shared__ int32_t waves[kScanThreads / 64];
    int32_t carry = 0;
    for (int64_t base = 0; base < tiles; base += kScanThreads) {
        const int64_t i = base + threadIdx.x;
        const int32_t v = i < tiles ? partial[i] : 0;
        int32_t total;
        const int32_t prefix = blockExclusiveScan(v, waves, total);
        if (i < tiles) partial[i] = carry + prefix;
        carry += total;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ void __launch_bounds__(kScanThreads) scanAdd(int32_t* __restrict__ out, const int32_t* __restrict__ partial,
                                                        int64_t n) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    const int32_t add = partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] += add;
}

// scanPartials and scanAdd in one launch, for at most kScanThreads tiles: each block sums the tile
// totals before its own (one per thread, the block's reduction), adds that to its tile, and the last
// block writes the grand total to out[n]. Integer sums, so the scan is the three-kernel one's.
__global__ void __launch_bounds__(kScanThreads) scanAddFew(int32_t* __restrict__ out, const int32_t* __restrict__ partial,
                                                           int64_t tiles, int64_t n) {
    __shared__ int32_t waves[kScanThreads / 64];
    const int64_t t = threadIdx.x;
    const int32_t before = t < tiles && t < (int64_t)blockIdx.x ? partial[t] : 0;
    int32_t add;
    (void)blockExclusiveScan(before, waves, add);
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] += add;
    if ((int64_t)blockIdx.x == tiles - 1 && threadIdx.x == 0) out[n] = add + partial[blockIdx.x];
}

// ---------------------------------------------------------------------------------------------
// Clock
// ---------------------------------------------------------------------------------------------

__global__ void clockBegin(Clock clock, int32_t runs, int32_t useClock) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    clock.active[r] =
        useClock ? (clock.time[r] < clock.until[r] && clock.steps[r] < clock.maxSteps[r] && !clock.failed[r]) : 1;
    clock.maxima[2 * r] = 0ull;
    clock.maxima[2 * r + 1] = 0ull;
}

// cflTimeStep (sphSolver.py:1410-1424) from the run's maxima.
__device__ __forceinline__ double runTimeStep(const RunDev& run, const Clock& clock, int r) {
    const double vmax = __longlong_as_double((long long)clock.maxima[2 * r]);
    const double amax = __longlong_as_double((long long)clock.maxima[2 * r + 1]);
    double dt = run.cfl * run.hmin / (run.c0 + vmax);
    if (amax > 1e-12) dt = fmin(dt, run.cfl * sqrt(run.hmin / amax));
    return fmin(dt, run.maxTimeStep);
}


// Each run's maxima from the per-slot magnitudes the step wrote (|v| at the gather, |a| in the
// acceleration sweep; zero on boundary particles): a maximum over each chunk of a run's slots,
// then over the run's chunks, written as the run's float64 maxima. No atomics, so no waves queue on
// one address, and a maximum does not depend on the order it is taken in. A chunk takes a workgroup
// of kMaximaThreads, so each thread reads 8 slots of the chunk's 8192: a small run has few chunks,
// and the loop's length is the kernel's.
constexpr int kMaximaThreads = 1024;

__global__ void __launch_bounds__(kMaximaThreads)
    chunkMaxima(const int32_t* __restrict__ chunkRun, const int64_t* __restrict__ chunkFirst, const RunDev* __restrict__ runs,
                const double* __restrict__ speed, const double* __restrict__ acceleration, Clock clock,
                double* __restrict__ partial) {
    __shared__ double top[2][kMaximaThreads];
    const int64_t b = blockIdx.x;
    const int r = chunkRun[b];
    if (!clock.active[r]) return;
    const RunDev& run = runs[r];
    const int64_t first = chunkFirst[b];
    const int64_t end = min(first + kSumChunk, run.first + run.count);
    double v = 0.0, a = 0.0;
    for (int64_t i = first + threadIdx.x; i < end; i += blockDim.x) {
        v = fmax(v, speed[i]);
        a = fmax(a, acceleration[i]);
    }
    top[0][threadIdx.x] = v;
    top[1][threadIdx.x] = a;
    __syncthreads();
    for (int width = blockDim.x / 2; width > 0; width >>= 1) {
        if ((int)threadIdx.x < width) {
            top[0][threadIdx.x] = fmax(top[0][threadIdx.x], top[0][threadIdx.x + width]);
            top[1][threadIdx.x] = fmax(top[1][threadIdx.x], top[1][threadIdx.x + width]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * b] = top[0][0];
        partial[2 * b + 1] = top[1][0];
    }
}

// Each run's maxima; then, as the reference's run loops stop a run (BreakingWave.py:892-919,
// BreakingWave3D.py:748-769), its failure: a maximum that is not finite (a non-finite speed,
// acceleration or position reads as infinite, since fmax passes NaN by), or a time step below a
// hundredth of the acoustic one, cfl hmin / (100 c0) (the reference's step limit: a hundred times
// the steps an interval takes at the acoustic step). With `useClock` a failed run stops at once: the
// step's integration and every later device-clocked step leave it alone.
__global__ void runMaxima(const RunDev* __restrict__ runsDev, const int64_t* __restrict__ runChunk,
                          const double* __restrict__ partial, Clock clock, int32_t runs, int32_t useClock) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs || !clock.active[r]) return;
    double v = 0.0, a = 0.0;
    for (int64_t b = runChunk[r]; b < runChunk[r + 1]; ++b) {
        v = fmax(v, partial[2 * b]);
        a = fmax(a, partial[2 * b + 1]);
    }
    clock.maxima[2 * r] = (unsigned long long)__double_as_longlong(v);
    clock.maxima[2 * r + 1] = (unsigned long long)__double_as_longlong(a);
    const RunDev& run = runsDev[r];
    int32_t failure = 0;
    if (!(v < INFINITY) || !(a < INFINITY)) failure = 1;
    else if (runTimeStep(run, clock, r) < run.cfl * run.hmin / (100.0 * run.c0)) failure = 2;
    if (failure && !clock.failed[r]) clock.failed[r] = failure;
    if (failure && useClock) clock.active[r] = 0;
}

// ---------------------------------------------------------------------------------------------
// Cell list
// ---------------------------------------------------------------------------------------------

// Runs of equal keys in consecutive lanes of a wave (a lane with `valid` false stands alone): the
// lane's run's first lane, the run's length and the lane's rank in it. Slots sorted into cell order
// put whole cells in one run, which then takes one atomic for the cell instead of one per slot.
// Every lane of the wave must call it; workgroups are whole waves.
__device__ __forceinline__ void equalRuns(int32_t key, bool valid, int& head, int& length, int& rank) {
    const int lane = (int)(threadIdx.x & 63);
    const int32_t previous = __shfl_up(key, 1, 64);
    const int previousValid = __shfl_up(valid ? 1 : 0, 1, 64);
    const bool starts = !valid || lane == 0 || !previousValid || previous != key;
    const unsigned long long heads = __ballot(starts);
    const unsigned long long upTo = 2ull << lane;   // wraps to 0 for lane 63, so upTo - 1 keeps every lane
    head = 63 - __clzll(heads & (upTo - 1));
    const unsigned long long after = heads & ~(upTo - 1);
    length = (after ? __ffsll(after) - 1 : 64) - head;
    rank = lane - head;
}

// Each slot's cell and the cells' counts, and with `eighth` the slot's eighth of its cell along x
// (0 .. 7, with kOutsideX set when x lies outside the grid's cells or is not finite). With `clockMode` the first
// slot of each run also opens the run's step: 1 sets its active flag from its clock, 2 makes it
// active whatever the clock; both clear its maxima.
template <int D>
__global__ void cellKeys(const RunDev* __restrict__ runs, const int32_t* __restrict__ runOf, const double* __restrict__ x,
                         const double* __restrict__ y, const double* __restrict__ z, int64_t n, int32_t* __restrict__ key,
                         int32_t* __restrict__ count, uint8_t* __restrict__ eighth, Clock clock, int32_t clockMode) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int32_t k = 0;
    if (valid) {
        const int r = runOf[i];
        const RunDev& run = runs[r];
        if (clockMode && i == run.first) {
            clock.active[r] = clockMode == 2 ||
                              (clock.time[r] < clock.until[r] && clock.steps[r] < clock.maxSteps[r] && !clock.failed[r]);
            clock.maxima[2 * r] = 0ull;
            clock.maxima[2 * r + 1] = 0ull;
        }
        const int64_t ux = cellCoordinate(x[i], run.cell);
        const int cx = wrapIndex(ux, run.nx);
        const int cy = wrapIndex(cellCoordinate(y[i], run.cell), run.ny);
        const int cz = D == 3 ? wrapIndex(cellCoordinate(z[i], run.cell), run.nz) : 0;
        k = (int32_t)(run.cellFirst + ((int64_t)cz * run.ny + cy) * run.nx + cx);
        key[i] = k;
        if (eighth) {
            const double within = x[i] / run.cell - (double)ux;   // in [0, 1) for a finite x
            const bool inside = ux == (int64_t)cx && within >= 0.0 && within < 1.0;
            eighth[i] = inside ? (uint8_t)min((int)(within * 8.0), 7) : kOutsideX;
        }
    }
    int head, length, rank;
    equalRuns(k, valid, head, length, rank);
    if (valid && rank == 0) atomicAdd(count + k, length);
}

__global__ void cellScatter(const int32_t* __restrict__ key, const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                            int32_t* __restrict__ order, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int32_t k = valid ? key[i] : 0;
    int head, length, rank;
    equalRuns(k, valid, head, length, rank);
    int32_t base = 0;
    if (valid && rank == 0) base = atomicAdd(fill + k, length);
    base = __shfl(base, head, 64);
    if (valid) order[start[k] + base + rank] = (int32_t)i;
}

// Orders each cell's slots by x, ties by slot, or by slot alone when `x` is null: the slot at each
// scattered place goes to its cell's start plus the number of the cell's slots before it, one
// thread per place, so a cell of k slots costs k reads per slot rather than a serial sort. With
// `eighths`, the cell's first slot in that order also writes the cell's eighths table (EighthTable):
// how many of the cell's slots lie in its first k eighths along x, k = 1 .. 7, one byte each, which
// bounds a row's candidates along x before a search within one eighth. 3D cells are ordered by x and
// take tables; a 2D cell holds about 7 slots, too few for bounds along x to repay their loads, and
// keeps slot order. The first `cells` threads also clear the cells' counters for the next build.
// The sorted place of the slot at scattered place a (its cell's start plus the cell's slots before
// it), writing the cell's eighths table when asked (see cellRank); `slot` receives the slot.
__device__ __forceinline__ int32_t rankInCell(const int32_t* __restrict__ start, const int32_t* __restrict__ key,
                                              const double* __restrict__ x, const uint8_t* __restrict__ eighth,
                                              const int32_t* __restrict__ scattered,
                                              unsigned long long* __restrict__ eighths, int64_t a, int32_t& slot) {
    const int32_t v = scattered[a];
    const int32_t c = key[v];
    const int32_t begin = start[c], end = start[c + 1];
    int32_t below = 0;
    if (x) {
        const double xv = x[v];
        unsigned long long table = 0ull;
        uint32_t outside = 0u;
        for (int32_t b = begin; b < end; ++b) {
            const int32_t u = scattered[b];
            const double xu = x[u];
            below += xu < xv || (xu == xv && u < v);
            if (eighths) {
                const uint32_t e = eighth[u];
                outside |= e & kOutsideX;
                table += kEighthsAbove << (8u * (e & 7u));
            }
        }
        if (eighths && below == 0) eighths[c] = outside || end - begin > 255 ? kEighthsUnknown : table;
    } else {
        for (int32_t b = begin; b < end; ++b) below += scattered[b] < v;
    }
    slot = v;
    return begin + below;
}

__global__ void cellRank(const int32_t* __restrict__ start, const int32_t* __restrict__ key, const double* __restrict__ x,
                         const uint8_t* __restrict__ eighth, const int32_t* __restrict__ scattered,
                         int32_t* __restrict__ order, unsigned long long* __restrict__ eighths,
                         int32_t* __restrict__ count, int32_t* __restrict__ fill, int64_t n, int64_t cells) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) {
        int32_t v;
        const int32_t at = rankInCell(start, key, x, eighth, scattered, eighths, a, v);
        order[at] = v;
    }
    if (a < cells) {
        count[a] = 0;
        fill[a] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// Records
// ---------------------------------------------------------------------------------------------

// Slot-order arrays a gather reads; `pressure` null takes the Tait pressure of `density`.
template <typename Real>
struct Sources {
    const double* pos[3];
    const double* vel[3];
    const double* density;
    const Real* pressure;
    const Real* mass;
    const Real* len;
    const uint8_t* fluid;
    const int32_t* key;
    Real* pressureOut;       // the Tait pressure per slot, written when non-null
};

// What a gather reads of slot i's state, loaded before the record is made, so that a kernel with
// other work to do first (cellRankGather) has the loads in flight meanwhile.
template <typename Real>
struct GatherIn {
    double x[3], u[3], rho;
    Real pressure, mass, length;
    int32_t key;
    bool fluid;
};

template <int D, typename Real>
__device__ __forceinline__ GatherIn<Real> gatherLoad(int32_t i, const Sources<Real>& src) {
    GatherIn<Real> in;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        in.x[a] = a < D ? src.pos[a][i] : 0.0;
        in.u[a] = a < D ? src.vel[a][i] : 0.0;
    }
    in.rho = src.density[i];
    in.pressure = src.pressure ? src.pressure[i] : Real(0);
    in.mass = src.mass[i];
    in.length = src.len[i];
    in.key = src.key[i];
    in.fluid = src.fluid[i] != 0;
    return in;
}

// Sorted particle s's record from slot i's state as gatherLoad read it: the float32 point and
// remainder of its float64 position (sphGpu.py:1010-1019), its velocity and density as the pair
// sums read them (sphGpu.py:1021-1029), rho - rho0 taken in float64 (sphGpu.py:1031-1038), the
// Tait pressure (sphGpu.py:1058-1070), p / rho^2 and m / rho; returns |v| of a fluid particle
// (infinite when its velocity or position is not finite), 0 on the boundary.
template <int D, typename Real>
__device__ __forceinline__ double gatherStore(const RunDev& run, const GatherIn<Real>& in, int32_t i, int64_t s,
                                              const Sources<Real>& src, Records<Real>& rec) {
    Fields<Real> f;
    Real p[3] = {Real(0), Real(0), Real(0)}, v[3] = {Real(0), Real(0), Real(0)};
    float l[3] = {0.0f, 0.0f, 0.0f};
    double v2 = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double x = in.x[a];
        if constexpr (sizeof(Real) == 4) {
            const float point = (float)x;
            p[a] = point;
            l[a] = (float)(x - (double)point);
        } else {
            p[a] = x;
        }
        const double u = in.u[a];
        v[a] = (Real)u;
        v2 += u * u;
    }
    const double rho = in.rho;
    const bool isFluid = in.fluid;
    Real pressure;
    if (src.pressure) {
        pressure = in.pressure;
    } else {
        const double ratio = rho / run.rho0, sq = ratio * ratio;
        double value = run.stiffness * (sq * sq * sq * ratio - 1.0);
        if (!run.negativePressure || !isFluid) value = fmax(value, 0.0);
        pressure = (Real)value;
        if (src.pressureOut) src.pressureOut[i] = pressure;
    }
    const Real rhoR = (Real)rho, massR = in.mass, h = in.length;
    for (int a = 0; a < 3; ++a) {
        f.p[a] = p[a];
        f.l[a] = l[a];
        f.v[a] = v[a];
    }
    f.sh = isFluid ? h : -h;
    f.rho = rhoR;
    f.mass = massR;
    f.vol = massR / rhoR;
    f.dev = (Real)(rho - run.rho0);
    f.pr2 = pressure / (rhoR * rhoR);
    f.press = pressure;
    rec.r[s] = pack<D, Real>(f);
    rec.key[s] = in.key;
    bool finite = v2 < INFINITY;
#pragma unroll
    for (int a = 0; a < D; ++a) finite = finite && fabs(in.x[a]) < INFINITY;
    return isFluid ? (finite ? sqrt(v2) : INFINITY) : 0.0;
}

template <int D, typename Real>
__device__ __forceinline__ double gatherOne(const RunDev& run, int32_t i, int64_t s, const Sources<Real>& src,
                                            Records<Real>& rec) {
    return gatherStore<D, Real>(run, gatherLoad<D, Real>(i, src), i, s, src, rec);
}

// Each sorted particle's record from its slot's state (gatherOne); resident, each slot's |v| for
// the run's |v|max.
template <int D, typename Real>
__global__ void gatherRecords(const RunDev* __restrict__ runs, const int32_t* __restrict__ runOf,
                              const int32_t* __restrict__ order, Sources<Real> src, Records<Real> rec, Clock clock,
                              int32_t resident, int64_t n, int32_t onlyFiltered = 0) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool inRange = s < n;
    const int r = inRange ? runOf[s] : 0;
    // again after the Shepard filter: only the runs it filtered this step
    if (onlyFiltered && !(inRange && filterDue(runs[r], clock, r, onlyFiltered))) return;
    const bool valid = inRange && (!resident || clock.active[r]);
    double speed = 0.0;
    if (valid) speed = gatherOne<D, Real>(runs[r], order[s], s, src, rec);
    if (resident && inRange && clock.active[r]) clock.speed[s] = speed;
}

// cellRank and the resident gatherRecords in one launch: the thread that places a slot in cell order
// also writes the slot's record there, so the step's gather neither waits for a launch of its own
// nor reads the order back.
template <int D, typename Real>
__global__ void cellRankGather(const int32_t* __restrict__ start, const int32_t* __restrict__ key,
                               const double* __restrict__ x, const uint8_t* __restrict__ eighth,
                               const int32_t* __restrict__ scattered, int32_t* __restrict__ order,
                               unsigned long long* __restrict__ eighths, int32_t* __restrict__ count,
                               int32_t* __restrict__ fill, int64_t n, int64_t cells, const RunDev* __restrict__ runs,
                               const int32_t* __restrict__ runOf, Sources<Real> src, Records<Real> rec, Clock clock) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) {
        // the slot's state is read while its place in the cell is counted
        const int32_t slot = scattered[a];
        const int r = runOf[slot];
        const bool active = clock.active[r] != 0;
        GatherIn<Real> in{};
        if (active) in = gatherLoad<D, Real>(slot, src);
        int32_t v;
        const int32_t at = rankInCell(start, key, x, eighth, scattered, eighths, a, v);
        order[at] = v;
        if (active) clock.speed[at] = gatherStore<D, Real>(runs[r], in, v, at, src, rec);
    }
    if (a < cells) {
        count[a] = 0;
        fill[a] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// Pair sweeps
// ---------------------------------------------------------------------------------------------

template <int D, typename Real>
__device__ __forceinline__ Self<D, Real> loadSelf(const RunDev& run, const Records<Real>& rec, int64_t s) {
    Self<D, Real> me;
    const Record<Real>& own = rec.r[s];
    const Quad<Real> a = own.q0;
    me.l[0] = me.l[1] = me.l[2] = 0.0f;
    if constexpr (D == 2) {
        me.p[0] = a.x;
        me.p[1] = a.y;
        me.p[2] = Real(0);
        if constexpr (sizeof(Real) == 4) {
            me.l[0] = a.z;
            me.l[1] = a.w;
        }
    } else {
        me.p[0] = a.x;
        me.p[1] = a.y;
        me.p[2] = a.z;
        if constexpr (sizeof(Real) == 4) {
            const Quad<Real> b = own.q1;
            me.l[0] = b.x;
            me.l[1] = b.y;
            me.l[2] = b.z;
        }
    }
    const Real sh = signedLength<D, Real>(own);
    me.fluid = sh > Real(0);
    me.h = fabs(sh);
    me.reach = me.h + (Real)run.hmax;
    me.cell = (int32_t)((int64_t)rec.key[s] - run.cellFirst);
    return me;
}

// Determinant and inverse of the leading D x D block (2D: the reference's third axis carries a one
// on the diagonal, sphGpu.py:636-641, which leaves the 2 x 2 block's).
template <int D, typename Real>
__device__ __forceinline__ Real determinant(const Real (&m)[3][3]) {
    if (D == 2) return m[0][0] * m[1][1] - m[0][1] * m[1][0];
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
template <int D, typename Real>
__device__ __forceinline__ void inverse(const Real (&m)[3][3], Real det, Real (&out)[3][3]) {
    const Real inv = Real(1) / det;
    if (D == 2) {
        out[0][0] = m[1][1] * inv;
        out[0][1] = -m[0][1] * inv;
        out[1][0] = -m[1][0] * inv;
        out[1][1] = m[0][0] * inv;
        return;
    }
    out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * inv;
    out[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
    out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    out[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
    out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    out[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    out[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
    out[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
}

template <int D>
__host__ __device__ constexpr double correctionFloorPower() {
    return D == 2 ? 0.0625 : 0.015625;   // 0.25^D, sphSolver.py:262
}

template <typename Real>
struct SweepOut {
    Real* pressure;        // sweep 1: every slot's pressure (boundary from the fluid)
    double* density;       // sweep 1: boundary densities that carry it
    Real* gradients;       // sweep 1: G per slot, D*D rows (optional)
    Real* acceleration;    // sweep 2: D rows per slot
    Real* corrections;     // sweep 2: L per slot, D*D rows (optional)
    Real* rate;            // sweep 3: d rho / dt per slot (host steps)
    double* pos[3];        // sweep 3 resident epilogue: the float64 state
    double* vel[3];
    double* density64;
    int64_t stride;
    // steps with the delta-LES-SPH closure: alpha_i at eddy[i] and delta_i at eddy[stride + i], in
    // slot order, the closure's c0 and rho0, and K nu / c0 (null eddy: no closure)
    const Real* eddy;
    Real eddyC0, eddyRho0, eddyWater;
    // resident steps that keep the energy budget (null ledger: none): sweep 2's parts of a, row
    // (part D + axis) of `parts`, the parts being the pressure and the viscosity from fluid, then from
    // boundary neighbors (sphGpu.py:396-414); the ledger's 14 rows (sphEnergy.py deviceTerms)
    Real* parts;
    double* ledger;
    int64_t ledgerStride;
    // sweep 3's parts of the rate, four rows of `stride` (plain over fluid, plain over boundary, the
    // correction's share, the diffusion), from its parts-only launch before the step's
    Real* rateParts;
};

// Sweep 1's sums: the extrapolated wall pressure of a boundary particle (sphGpu.py:804-854, no
// coincident skip) or the velocity gradient of a fluid particle (sphGpu.py:643-697), each over
// fluid neighbors. One accumulator serves the standalone sweep and the one fused with the list
// build, so both add the same pairs in the same order.
template <int D, typename Real>
struct Sweep1 {
    static constexpr int K = 2 * D * D;   // wall: weight, total; gradients: rises then offsets
    Real acc[K];
    int mode;                             // 0 none, 1 wall pressure, 2 gradients
    Real vi[3];
    Real g[3];

    __device__ __forceinline__ void begin(const RunDev& run, const Records<Real>& rec, int64_t s, int chosen) {
#pragma unroll
        for (int e = 0; e < K; ++e) acc[e] = Real(0);
        mode = chosen;
        velocityOf<D, Real>(rec.r[s], vi);
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a] = (Real)run.gravity[a];
    }
    // Adds neighbor `other`, of which it reads the signed length, and the pressure and density
    // (wall) or the volume and velocity (gradients).
    __device__ __forceinline__ void add(const Record<Real>& other, const PairKernel<Real, D>& kernel, Real dx, Real dy,
                                        Real dz, Real rr, Real h) {
        if (mode == 0) return;
        if (!(signedLength<D, Real>(other) > Real(0))) return;
        if (mode == 1) {
            const Real w = kernel.value(rr, h);
            Real gd = g[0] * dx + g[1] * dy;
            if (D == 3) gd += g[2] * dz;
            acc[0] += w;
            acc[1] += w * (pressureOf<D, Real>(other) + densityOf<D, Real>(other) * gd);
            return;
        }
        if (rr < Real(1.0e-12)) return;
        const Real w = volumeOf<D, Real>(other) * kernel.slopeOverR(rr, h);
        const Real d[3] = {dx, dy, dz};
        Real vj[3];
        velocityOf<D, Real>(other, vj);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const Real dv = w * (vj[a] - vi[a]);
            const Real da = w * d[a];
#pragma unroll
            for (int b = 0; b < D; ++b) {
                acc[a * D + b] += dv * d[b];
                acc[D * D + a * D + b] -= da * d[b];
            }
        }
    }
};

// The mode of sweep 1 for a particle.
__device__ __forceinline__ int sweep1Mode(const RunDev& run, bool fluidI, int32_t resident, int32_t doWall,
                                          int32_t doGradients) {
    if (!fluidI && doWall && (!resident || run.extrapolatedWalls)) return 1;
    if (fluidI && doGradients && (!resident || run.limiter || run.closure)) return 2;
    return 0;
}

// Sweep 1's results for sorted particle s: a boundary particle's pressure and the density that
// carries it, or a fluid particle's G; writes the pass-through outputs of a particle with neither.
template <int D, typename Real>
__device__ __forceinline__ void sweep1Finish(const RunDev& run, const Sweep1<D, Real>& sweep, Records<Real>& rec,
                                             const SweepOut<Real>& out, int64_t s, int32_t i, bool fluidI,
                                             int32_t resident, int32_t doWall, int32_t doGradients) {
    if (sweep.mode == 0) {
        if (doWall && out.pressure && (!fluidI || !resident)) out.pressure[i] = pressureOf<D, Real>(rec.r[s]);
        if (doGradients && out.gradients && !fluidI)
            for (int e = 0; e < D * D; ++e) out.gradients[e * out.stride + i] = Real(0);
        return;
    }
    if (sweep.mode == 1) {
        Real pressure = Real(0);
        if (sweep.acc[0] > Real(0)) pressure = fmax(sweep.acc[1] / sweep.acc[0], Real(0));
        const double rho = run.rho0 * pow(1.0 + (double)pressure / run.stiffness, 1.0 / 7.0);
        setWall<D, Real>(rec.r[s], pressure, (Real)rho, (Real)(rho - run.rho0));
        if (out.pressure) out.pressure[i] = pressure;
        if (out.density) out.density[i] = rho;
        return;
    }
    Real offsets[3][3] = {}, rises[3][3] = {};
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            rises[a][b] = sweep.acc[a * D + b];
            offsets[a][b] = sweep.acc[D * D + a * D + b];
        }
    Real G[3][3] = {};
    const Real det = determinant<D, Real>(offsets);
    if (det >= (Real)correctionFloorPower<D>()) {
        Real inv[3][3];
        inverse<D, Real>(offsets, det, inv);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                Real sum = Real(0);
#pragma unroll
                for (int c = 0; c < D; ++c) sum += rises[a][c] * inv[c][b];
                G[a][b] = sum;
            }
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            rec.G[s * (D * D) + a * D + b] = G[a][b];
            if (out.gradients) out.gradients[(a * D + b) * out.stride + i] = G[a][b];
        }
}

// The step's neighbor list: each particle's neighbors {j != i : r_ij < h_i + h_j} in the order the
// cell walk visits them. A group of `Lanes` lanes tests a particle's candidates `Lanes` at a time
// and compacts the hits with a ballot, so the list keeps the walk's order. The test reads only the
// fields the predicate needs. With `WithSweep1` a second pass over the hits adds sweep 1's sums,
// neighbor m on lane m % Lanes as every list sweep deals them: the sums' arithmetic then runs once
// per neighbor, not once per batch of candidates of which about a third are neighbors. The first
// `hitCapacity` hits of each particle are also kept in LDS after the NearCells copy, as the LDS index
// of the neighbor's record (kLdsBit set) where the copy holds it, so the second pass reads LDS only;
// later hits are read back from the list, and a particle whose neighbors overflow the list walks its
// cells for the sums, dealing them alike.
template <int D, typename Real, int Lanes, bool WithSweep1>
__global__ void SWEEP_BOUNDS((sweepWaves<D, Real>()))
    buildNeighbors(const RunDev* __restrict__ runs, Tiles tiles, const int32_t* __restrict__ order, const int32_t* __restrict__ start,
                   const unsigned long long* __restrict__ eighths, Records<Real> rec,
                   NeighborList list, SweepOut<Real> out, Clock clock, const double* __restrict__ y64,
                   int32_t resident, int32_t cached, int32_t hitCapacity, int64_t n, int32_t filterMode = 0) {
    const int64_t tile = xcdBlock();
    const int r = tiles.run[tile];
    const int64_t s = tiles.first[tile] + threadIdx.x / Lanes;
    const int lane = (int)(threadIdx.x % Lanes);
    const int groupBase = (int)(threadIdx.x & 63) & ~(Lanes - 1);
    const bool runSteps = !resident || clock.active[r];
    bool valid = s < runs[r].first + runs[r].count && runSteps;
    (void)n;
    NearCells<D, Real> cells;
    cells.r = rec.r;
    cells.held = 0;
    cells.stride = cached;
    cells.kicked = false;
#pragma unroll
    for (int i = 0; i < NearCells<D, Real>::kRows; ++i) cells.lo[i] = cells.hi[i] = cells.at[i] = 0;
    if (runSteps && cached > 0) {
        const RunDev& run = runs[r];
        const int64_t first = tiles.first[tile];
        const int64_t last = min(first + (int64_t)(blockDim.x / Lanes), run.first + run.count) - 1;
        if (last >= first) {
            cells.plan(run, start, rec.key, first, last, cached);
            cells.fill();
        }
    }
    // each particle's kept hits at an odd stride, so the groups of a wave read apart in the banks
    const int32_t keep = WithSweep1 ? hitCapacity : 0;
    LdsHits hits = ldsHits<Real>(cached) + (threadIdx.x / Lanes) * (keep | 1);
    Self<D, Real> me{};
    double yy = 0.0;
    int32_t count = 0;
    const bool uniform = runs[r].uniform != 0;
    if (valid) {
        const RunDev& run = runs[r];
        me = loadSelf<D, Real>(run, rec, s);
        int cx, cy, cz;
        cellCoordinates(run, me.cell, cx, cy, cz);
        yy = run.period > 0.0 ? y64[order[s]] : 0.0;
        const Real period = (Real)run.period;
        const Real twoHmax = Real(2) * (Real)run.hmax;
        const uint32_t groupMask = Lanes >= 32 ? 0xffffffffu : ((1u << Lanes) - 1u);
        const uint32_t below = (1u << lane) - 1u;
        const int32_t self = (int32_t)s;
        int images = 1;
        if (run.period > 0.0 && (me.p[1] < me.reach || me.p[1] > period - me.reach)) images = 2;
        for (int image = 0; image < images; ++image) {
            Real shift = Real(0);
            int oy = cy;
            if (image == 1) {
                shift = me.p[1] < Real(0.5) * period ? period : -period;
                oy = wrapIndex(cellCoordinate(yy + (double)shift, run.cell), run.ny);
            }
            const int32_t tag = image ? kImageBit : 0;
            const Real reach = rowReach(twoHmax, me.p[0]);
            // 3D rows of three cells take their candidates from the cells' eighths tables, refined
            // by a search within the first and last eighth; the single cells at the grid's ends
            // along x by the search alone; 2D rows are walked whole
            const bool threeCells = D == 3 && run.nx >= 3 && cx >= 1 && cx <= run.nx - 2;
            const Real cell = (Real)run.cell;
            const Real rowLeft = threeCells ? cellLowerEdge(me.p[0], cx, run.nx, cell) - cell : Real(0);
            forStencil<D>(run, start, cx, oy, cz, [&](int32_t begin, int32_t end, int32_t middle, int32_t right,
                                                       int rowY, int rowZ, bool wrapped) {
                const int32_t copied = cells.rangeBase(begin, end);
                auto pointX = [&](int32_t k) {
                    return copied >= 0 ? ldsValue<Real>(copied + (k - begin), 0, 0, cached) : rec.r[k].q0.x;
                };
                int32_t from = begin, to = end;
                const Real across = D == 3 ? rowReachAcross<D>(run, reach, me.p[1], me.p[2], cy, cz, rowY, rowZ,
                                                               wrapped || image == 1)
                                           : reach;
                if (threeCells) {
                    const int64_t plane = D == 3 ? (int64_t)wrapNear(rowZ, run.nz) : 0;
                    const int64_t left = run.cellFirst + (plane * run.ny + wrapNear(rowY, run.ny)) * run.nx + cx - 1;
                    const unsigned long long tables[3] = {eighths[left], eighths[left + 1], eighths[left + 2]};
                    rowEighths<SPH_REFINE_3D != 0>(begin, middle, right, end, me.p[0], across, rowLeft, cell, tables,
                                                   pointX, from, to);
                } else if (D == 3) {
                    rowCandidates(begin, end, middle, right, me.p[0], reach, across, pointX, from, to);
                }
                // the predicate's fields only: the point and remainder, and the length when lengths
                // differ; a lane past the range reads the range's last record and never hits, so
                // the loop body has no branch but the predicate's rare rounded root. One loop per
                // source of the records (LDS or the sorted records) and per kind of run (one
                // smoothing length or several), so neither is decided per candidate.
                auto scan = [&](auto fetch, auto oneLength) {
                    constexpr bool Same = decltype(oneLength)::value;
                    const SupportBand<Real> band(twoHmax);
                    // 3D float32 with one length: the point alone decides all but a thin shell of
                    // candidates (CoarseSupport), and only those read the remainder
                    constexpr bool Coarse = D == 3 && Same && sizeof(Real) == 4;
                    const CoarseSupport<Real> coarse(band, me, shift);
                    for (int32_t base = from; base < to; base += Lanes) {
                        const int32_t k = base + lane;
                        const bool inside = k < to;
                        const int32_t at = inside ? k : to - 1;
                        Record<Real> other;
                        fetch(at, other, !Coarse && (D == 3 || !Same));
                        bool near;
                        if constexpr (Coarse) {
                            const int decided = coarse.decide(me, other, shift);
                            near = decided > 0;
                            if (decided == 0) {
                                fetch(at, other, true);
                                Real dx, dy, dz;
                                separation<D, Real>(me, other, shift, dx, dy, dz);
                                near = band.holds(squaredDistance(dx, dy, dz, D));
                            }
                        } else {
                            Real dx, dy, dz;
                            separation<D, Real>(me, other, shift, dx, dy, dz);
                            const Real r2 = squaredDistance(dx, dy, dz, D);
                            // 2 h_ij = h_i + h_j exactly, as 2 (0.5 (h_i + h_j)) is
                            near = Same ? band.holds(r2)
                                        : withinSupport<Real>(r2, me.h + fabs(signedLength<D, Real>(other)));
                        }
                        const bool hit = near && inside && k != self;
                        const uint32_t mine = (uint32_t)(__ballot(hit) >> groupBase) & groupMask;
                        if (hit) {
                            const int32_t slot = count + __popc(mine & below);
                            if (slot < list.capacity) list.index[listOffset(slot, list.stride) + s] = k | tag;
                            if (slot < keep) hits[slot] = (copied >= 0 ? (copied + (k - begin)) | kLdsBit : k) | tag;
                        }
                        count += __popc(mine);
                    }
                };
                auto fromLds = [&](int32_t k, Record<Real>& other, bool both) {
                    const int32_t index = copied + (k - begin);
                    other.q0 = ldsQuad<Real>(index, 0, cached);
                    if (both) other.q1 = ldsQuad<Real>(index, 1, cached);
                };
                auto fromRecords = [&](int32_t k, Record<Real>& other, bool both) {
                    other.q0 = rec.r[k].q0;
                    if (both) other.q1 = rec.r[k].q1;
                };
                using One = std::integral_constant<bool, true>;
                using Several = std::integral_constant<bool, false>;
                if (copied >= 0) {
                    if (uniform) scan(fromLds, One{}); else scan(fromLds, Several{});
                } else {
                    if (uniform) scan(fromRecords, One{}); else scan(fromRecords, Several{});
                }
            });
        }
        if (lane == 0) list.count[s] = count;
    }
    if (!WithSweep1) return;
    // a run the Shepard filter takes this step has its sweep 1 after the filter, from its new
    // densities (the workgroup is one run's, so all of it leaves here or none)
    if (resident && filterDue(runs[r], clock, r, filterMode)) return;
    // every thread arrives here: the hits and list entries this workgroup wrote are visible to it
    // after the barrier
    __syncthreads();
    Sweep1<D, Real> sweep;
    sweep.mode = 0;
    const bool fluidI = me.fluid;
    if (valid) {
        const RunDev& run = runs[r];
        sweep.begin(run, rec, s, sweep1Mode(run, fluidI, resident, 1, 1));
        if (sweep.mode) {
            PairKernel<Real, D> kernel;
            kernel.init(run.kernel, run.sigmaUnit, run.hmax, uniform);
            auto add = [&](int32_t, const Record<Real>& other, Real dx, Real dy, Real dz, Real rr, Real h) {
                sweep.add(other, kernel, dx, dy, dz, rr, h);
            };
            if (count <= list.capacity) {
                const Real shiftImage = me.p[1] < Real(0.5) * (Real)run.period ? (Real)run.period : -(Real)run.period;
                auto visit = [&](int32_t entry) {
                    Record<Real> other;
                    if (entry & kLdsBit) {
                        const int32_t index = entry & 0x3fffffff;
                        other.q0 = ldsQuad<Real>(index, 0, cached);
                        other.q1 = ldsQuad<Real>(index, 1, cached);
                        other.q2 = ldsQuad<Real>(index, 2, cached);
                        other.q3 = ldsQuad<Real>(index, 3, cached);
                    } else {
                        other = cells(entry & 0x7fffffff);
                    }
                    Real dx, dy, dz;
                    separation<D, Real>(me, other, entry < 0 ? shiftImage : Real(0), dx, dy, dz);
                    const Real rr = sumDistance(dx, dy, dz, D);
                    const Real h = uniform ? (Real)run.hmax : Real(0.5) * (me.h + fabs(signedLength<D, Real>(other)));
                    add(-1, other, dx, dy, dz, rr, h);
                };
                // the kept hits from LDS, the rest read back from the list
                const int32_t kept = min(count, keep);
                int32_t m = lane;
                for (; m < kept; m += Lanes) visit(hits[m]);
                for (; m < count; m += Lanes) visit(list.index[listOffset(m, list.stride) + s]);
            } else {
                forListedNeighbors<D, Real, Lanes, false>(run, start, cells, list, s, count, me, yy, lane, uniform, add);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < Sweep1<D, Real>::K; ++e) sweep.acc[e] = groupSum<Lanes>(sweep.acc[e]);
    if (!valid || lane != 0) return;
    sweep1Finish<D, Real>(runs[r], sweep, rec, out, s, order[s], fluidI, resident, 1, 1);
}

// Sweep 1 on its own: wall pressure on boundary particles, G on fluid particles, over the list when
// there is one and the cells otherwise, each neighbor dealt as the fused sweep deals it.
template <int D, typename Real, int Lanes>
__global__ void SWEEP_BOUNDS((sweepWaves<D, Real>()))
    sweepWallAndGradients(const RunDev* __restrict__ runs, Tiles tiles, const int32_t* __restrict__ order, const int32_t* __restrict__ start, Records<Real> rec,
                          NeighborList list, SweepOut<Real> out, Clock clock, const double* __restrict__ y64,
                          int32_t resident, int32_t doWall, int32_t doGradients, int64_t n,
                          int32_t onlyFiltered = 0) {
    const int64_t tile = xcdBlock();
    const int r = tiles.run[tile];
    if (onlyFiltered && !filterDue(runs[r], clock, r, onlyFiltered)) return;
    const int64_t s = tiles.first[tile] + threadIdx.x / Lanes;
    const int lane = (int)(threadIdx.x % Lanes);
    Sweep1<D, Real> sweep;
    sweep.mode = 0;
    bool fluidI = false;
    const bool valid = s < runs[r].first + runs[r].count && (!resident || clock.active[r]);
    (void)n;
    if (valid) {
        const RunDev& run = runs[r];
        const Self<D, Real> me = loadSelf<D, Real>(run, rec, s);
        fluidI = me.fluid;
        sweep.begin(run, rec, s, sweep1Mode(run, fluidI, resident, doWall, doGradients));
        if (sweep.mode) {
            PairKernel<Real, D> kernel;
            kernel.init(run.kernel, run.sigmaUnit, run.hmax, run.uniform != 0);
            const double yy = run.period > 0.0 ? y64[order[s]] : 0.0;
            forEachNeighbor<D, Real, Lanes, false>(run, start, GlobalRecords<Real>{rec.r}, list, s, me, yy, lane,
                                                   kernel.uniform,
                                                   [&](int32_t, const Record<Real>& other, Real dx, Real dy, Real dz,
                                                       Real rr, Real h) { sweep.add(other, kernel, dx, dy, dz, rr, h); });
        }
    }
#pragma unroll
    for (int e = 0; e < Sweep1<D, Real>::K; ++e) sweep.acc[e] = groupSum<Lanes>(sweep.acc[e]);
    if (!valid || lane != 0) return;
    sweep1Finish<D, Real>(runs[r], sweep, rec, out, s, order[s], fluidI, resident, doWall, doGradients);
}

// The options off under delta-SPH, host steps only. `which` 0: the Shepard-filtered density,
// sum_j m_j W_ij / sum_j (m_j / rho_j) W_ij over every neighbor and the particle itself at W(0, h_i)
// (sphGpu.py:768-802); 1: the XSPH transport velocity, v_i + epsilon sum_j m_j W_ij / ((rho_i +
// rho_j) / 2) (v_j - v_i) over the fluid neighbors of a fluid particle, the boundary's velocity
// unchanged (sphGpu.py:725-766); neither skips coincident pairs. 2: the renormalized density
// gradient, L_i sum_j V_j (rho_j - rho_i) grad_i W_ij over every neighbor but the coincident, from
// the deviations rho - rho0 and the L in `rec.L`, zero on the boundary (sphGpu.py:548-591), also
// kept per sorted slot in `rec.G` (D values each) for the continuity sum. One result per slot:
// `out[i]`, or `out[a * stride + i]` for axis a; `out` may be null for the gradient.
template <int D, typename Real, int Lanes>
__global__ void SWEEP_BOUNDS((sweepWaves<D, Real>()))
    sweepFilters(const RunDev* __restrict__ runs, Tiles tiles, const int32_t* __restrict__ order,
                 const int32_t* __restrict__ start, Records<Real> rec, NeighborList list,
                 const double* __restrict__ y64, int32_t which, Real epsilon, Real* __restrict__ out, int64_t stride,
                 Clock clock = Clock{}, int32_t onlyFiltered = 0) {
    const int64_t tile = xcdBlock();
    const int r = tiles.run[tile];
    if (onlyFiltered && !filterDue(runs[r], clock, r, onlyFiltered)) return;
    const int64_t s = tiles.first[tile] + threadIdx.x / Lanes;
    const int lane = (int)(threadIdx.x % Lanes);
    const bool valid = s < runs[r].first + runs[r].count;
    Real acc[3] = {Real(0), Real(0), Real(0)};
    Real vi[3] = {Real(0), Real(0), Real(0)};
    Real own = Real(0);
    if (valid) {
        const RunDev& run = runs[r];
        const Self<D, Real> me = loadSelf<D, Real>(run, rec, s);
        const Real rhoI = densityOf<D, Real>(rec.r[s]);
        velocityOf<D, Real>(rec.r[s], vi);
        if (which == 0 || me.fluid) {
            PairKernel<Real, D> kernel;
            kernel.init(run.kernel, run.sigmaUnit, run.hmax, run.uniform != 0);
            own = kernel.sigmaAt(me.h);
            const double yy = run.period > 0.0 ? y64[order[s]] : 0.0;
            const Real devI = deviationOf<D, Real>(rec.r[s]);
            forEachNeighbor<D, Real, Lanes, false>(
                run, start, GlobalRecords<Real>{rec.r}, list, s, me, yy, lane, kernel.uniform,
                [&](int32_t, const Record<Real>& other, Real dx, Real dy, Real dz, Real rr, Real h) {
                    if (which == 2) {
                        if (rr < Real(1.0e-12)) return;
                        const Real w = volumeOf<D, Real>(other) * (deviationOf<D, Real>(other) - devI) *
                                       kernel.slopeOverR(rr, h);
                        const Real d[3] = {dx, dy, dz};
#pragma unroll
                        for (int a = 0; a < D; ++a) acc[a] += w * d[a];
                        return;
                    }
                    const Real w = kernel.value(rr, h);
                    if (which == 0) {
                        acc[0] += massOf<D, Real>(other) * w;
                        acc[1] += volumeOf<D, Real>(other) * w;
                        return;
                    }
                    if (!(signedLength<D, Real>(other) > Real(0))) return;
                    const Real weight = massOf<D, Real>(other) * w / (Real(0.5) * (rhoI + densityOf<D, Real>(other)));
                    Real vj[3];
                    velocityOf<D, Real>(other, vj);
#pragma unroll
                    for (int a = 0; a < D; ++a) acc[a] += weight * (vj[a] - vi[a]);
                });
        }
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) acc[e] = groupSum<Lanes>(acc[e]);
    if (!valid || lane != 0) return;
    const int32_t i = order[s];
    if (which == 0) {
        const Record<Real>& self = rec.r[s];
        out[i] = (massOf<D, Real>(self) * own + acc[0]) / (volumeOf<D, Real>(self) * own + acc[1]);
        return;
    }
    if (which == 2) {
        Real g[3] = {Real(0), Real(0), Real(0)};
        if (signedLength<D, Real>(rec.r[s]) > Real(0)) {
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) g[a] += rec.L[s * (D * D) + a * D + b] * acc[b];
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            rec.G[s * D + a] = g[a];
            if (out) out[a * stride + i] = g[a];
        }
        return;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) out[a * stride + i] = vi[a] + epsilon * acc[a];
}

// Particle shifting after Lind et al., host steps only, on one smoothing length h = hmax
// (sphGpu.py:856-1004). Per sorted slot s, `fields` holds 3 D + 1 values from s (3 D + 1): grad C
// (D), grad C0 (D), div r, grad rho (D).
//   pass 0, on the fluid over every neighbor but the coincident: grad C with the anti-clumping
//     weight 1 + 0.2 (W / W(dx))^4, grad C0 without it, div r = -sum V_j r W'(r), set to -1 where
//     h |grad C0| exceeds 0.3 (cubic spline) or 0.4 (Wendland C2), and grad rho from the
//     deviations; a boundary particle's div r is 1e6, so it never reads as the surface;
//   pass 1, on a fluid particle with div r at least 1.5 (2D) or 2.4 (3D): the surface neighbors'
//     -grad C0_j / |grad C0_j| weighted by W; the shift -size grad C_i (size |v_i| when `local`),
//     zero with a surface neighbor and h |grad C_i| above 0.03, else projected off the normal; the
//     density change shift . grad rho_i; the velocity change sum_j V_j (v_j - v_i) (grad W_ij .
//     shift) over fluid neighbors but the coincident. Shift and velocity change at
//     out[a * stride + i], the density change at density[i]; zero elsewhere.
template <int D>
__host__ __device__ constexpr int shiftWidth() { return 3 * D + 1; }

template <int D, typename Real, int Lanes>
__global__ void SWEEP_BOUNDS((sweepWaves<D, Real>()))
    sweepShifting(const RunDev* __restrict__ runs, Tiles tiles, const int32_t* __restrict__ order,
                  const int32_t* __restrict__ start, Records<Real> rec, NeighborList list,
                  const double* __restrict__ y64, int32_t pass, Real size, int32_t local, Real spacingKernel,
                  Real* __restrict__ fields, Real* __restrict__ shiftOut, Real* __restrict__ densityOut,
                  Real* __restrict__ velocityOut, int64_t stride) {
    constexpr int W = shiftWidth<D>();
    const int64_t tile = xcdBlock();
    const int r = tiles.run[tile];
    const int64_t s = tiles.first[tile] + threadIdx.x / Lanes;
    const int lane = (int)(threadIdx.x % Lanes);
    const bool valid = s < runs[r].first + runs[r].count;
    const RunDev& run = runs[r];
    const Real h = (Real)run.hmax;
    const Real threshold = D == 2 ? Real(1.5) : Real(2.4);
    PairKernel<Real, D> kernel;
    kernel.init(run.kernel, run.sigmaUnit, run.hmax, true);
    bool fluidI = false;
    Self<D, Real> me{};
    double yy = 0.0;
    if (valid) {
        me = loadSelf<D, Real>(run, rec, s);
        fluidI = me.fluid;
        yy = run.period > 0.0 ? y64[order[s]] : 0.0;
    }
    if (pass == 0) {
        Real acc[W];
#pragma unroll
        for (int e = 0; e < W; ++e) acc[e] = Real(0);
        if (valid && fluidI) {
            const Real devI = deviationOf<D, Real>(rec.r[s]);
            forEachNeighbor<D, Real, Lanes, true>(
                run, start, GlobalRecords<Real>{rec.r}, list, s, me, yy, lane, true,
                [&](int32_t, const Record<Real>& other, Real dx, Real dy, Real dz, Real rr, Real) {
                    const Real volume = volumeOf<D, Real>(other);
                    const Real F = kernel.slopeOverR(rr, h);
                    const Real ratio = kernel.value(rr, h) / spacingKernel;
                    const Real square = ratio * ratio;
                    const Real anti = Real(1) + Real(0.2) * square * square;
                    const Real change = deviationOf<D, Real>(other) - devI;
                    const Real d[3] = {dx, dy, dz};
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        const Real weighted = volume * F * d[a];
                        acc[a] += anti * weighted;
                        acc[D + a] += weighted;
                        acc[2 * D + 1 + a] += change * weighted;
                    }
                    acc[2 * D] -= volume * F * rr * rr;
                });
        }
#pragma unroll
        for (int e = 0; e < W; ++e) acc[e] = groupSum<Lanes>(acc[e]);
        if (!valid || lane != 0) return;
        Real* own = fields + s * W;
        if (!fluidI) {
#pragma unroll
            for (int e = 0; e < W; ++e) own[e] = Real(0);
            own[2 * D] = Real(1.0e6);
            return;
        }
        Real plain2 = Real(0);
#pragma unroll
        for (int a = 0; a < D; ++a) plain2 += acc[D + a] * acc[D + a];
        const Real lopsided = run.kernel == 1 ? Real(0.4) : Real(0.3);
        if (h * sqrt(plain2) > lopsided) acc[2 * D] = Real(-1);
#pragma unroll
        for (int e = 0; e < W; ++e) own[e] = acc[e];
        return;
    }
    const Real* own = fields + (valid ? s : 0) * W;
    const bool active = valid && fluidI && own[2 * D] >= threshold;
    Real normal[3] = {Real(0), Real(0), Real(0)};
    Real vicinity = Real(0);
    if (active) {
        forEachNeighbor<D, Real, Lanes, false>(
            run, start, GlobalRecords<Real>{rec.r}, list, s, me, yy, lane, true,
            [&](int32_t k, const Record<Real>&, Real, Real, Real, Real rr, Real) {
                const Real* other = fields + (int64_t)k * W;
                if (other[2 * D] >= threshold) return;
                vicinity += Real(1);
                Real length2 = Real(0);
#pragma unroll
                for (int a = 0; a < D; ++a) length2 += other[D + a] * other[D + a];
                if (length2 > Real(0)) {
                    const Real w = kernel.value(rr, h) / sqrt(length2);
#pragma unroll
                    for (int a = 0; a < D; ++a) normal[a] -= w * other[D + a];
                }
            });
    }
    vicinity = groupSum<Lanes>(vicinity);
#pragma unroll
    for (int a = 0; a < 3; ++a) normal[a] = groupSum<Lanes>(normal[a]);
    Real shift[3] = {Real(0), Real(0), Real(0)};
    Real vi[3] = {Real(0), Real(0), Real(0)};
    if (active) {
        velocityOf<D, Real>(rec.r[s], vi);
        Real scale = size;
        if (local) {
            Real v2 = Real(0);
#pragma unroll
            for (int a = 0; a < D; ++a) v2 += vi[a] * vi[a];
            scale = size * sqrt(v2);
        }
        Real grad2 = Real(0);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            shift[a] = -scale * own[a];
            grad2 += own[a] * own[a];
        }
        if (vicinity > Real(0)) {
            if (h * sqrt(grad2) > Real(0.03))
#pragma unroll
                for (int a = 0; a < D; ++a) shift[a] = Real(0);
            Real weight2 = Real(0);
#pragma unroll
            for (int a = 0; a < D; ++a) weight2 += normal[a] * normal[a];
            if (weight2 > Real(0)) {
                const Real weight = sqrt(weight2);
                Real along = Real(0);
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    normal[a] = normal[a] / weight;
                    along += shift[a] * normal[a];
                }
#pragma unroll
                for (int a = 0; a < D; ++a) shift[a] -= along * normal[a];
            }
        }
    }
    Real shift2 = Real(0);
#pragma unroll
    for (int a = 0; a < D; ++a) shift2 += shift[a] * shift[a];
    Real change[3] = {Real(0), Real(0), Real(0)};
    if (active && shift2 > Real(0)) {
        forEachNeighbor<D, Real, Lanes, true>(
            run, start, GlobalRecords<Real>{rec.r}, list, s, me, yy, lane, true,
            [&](int32_t, const Record<Real>& other, Real dx, Real dy, Real dz, Real rr, Real) {
                if (!(signedLength<D, Real>(other) > Real(0))) return;
                const Real d[3] = {dx, dy, dz};
                Real along = Real(0);
#pragma unroll
                for (int a = 0; a < D; ++a) along += d[a] * shift[a];
                const Real w = volumeOf<D, Real>(other) * kernel.slopeOverR(rr, h) * along;
                Real vj[3];
                velocityOf<D, Real>(other, vj);
#pragma unroll
                for (int a = 0; a < D; ++a) change[a] += w * (vj[a] - vi[a]);
            });
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) change[a] = groupSum<Lanes>(change[a]);
    if (!valid || lane != 0) return;
    const int32_t i = order[s];
    Real density = Real(0);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        density += shift[a] * own[2 * D + 1 + a];
        shiftOut[a * stride + i] = shift[a];
        velocityOut[a * stride + i] = change[a];
    }
    densityOut[i] = density;
}

// The delta-LES-SPH closure's coefficients of each sorted slot from its G in `rec.G` (sphGpu.py:
// 699-723): alpha_i = 2 (D + 2) (C_S l / h)^2 h ||D|| / c0 and delta_i = (C_delta l / h)^2 h ||D|| / c0
// with C_S = 0.12, C_delta = 1.5, l / h = 2 and ||D|| = sqrt(2 D:D) of G's symmetric part, each at
// most 0.2, zero on the boundary; alpha at eddy[i], delta at eddy[stride + i].
template <int D, typename Real>
__global__ void closureCoefficients(const int32_t* __restrict__ order, Records<Real> rec, Real* __restrict__ eddy,
                                    int64_t stride, Real soundSpeed, int64_t n) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const Real sh = signedLength<D, Real>(rec.r[s]);
    Real alpha = Real(0), delta = Real(0);
    if (sh > Real(0)) {
        Real dd = Real(0);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                const Real strain = Real(0.5) * (rec.G[s * (D * D) + a * D + b] + rec.G[s * (D * D) + b * D + a]);
                dd += strain * strain;
            }
        const Real scale = sh * sqrt(Real(2) * dd) / soundSpeed;
        const Real smagorinsky = (Real)((0.12 * 2.0) * (0.12 * 2.0)), diffusion = (Real)((1.5 * 2.0) * (1.5 * 2.0));
        alpha = fmin(Real(2 * (D + 2)) * smagorinsky * scale, Real(0.2));
        delta = fmin(diffusion * scale, Real(0.2));
    }
    const int32_t i = order[s];
    eddy[i] = alpha;
    eddy[stride + i] = delta;
}

// Sweep 2: accelerations with the limited viscosity (sphGpu.py:303-414) and, when asked, the
// correction matrix L (sphGpu.py:593-641) on fluid particles.
template <int D, typename Real, int Lanes, bool Options = false, bool Parts = false>
__global__ void SWEEP_BOUNDS((accelerationWaves<D, Real>()))
    sweepAccelerations(const RunDev* __restrict__ runs, Tiles tiles, const int32_t* __restrict__ order, const int32_t* __restrict__ start, Records<Real> rec,
                       NeighborList list, SweepOut<Real> out, Clock clock, const double* __restrict__ y64,
                       int32_t resident, int32_t doAcceleration, int32_t doCorrection, int32_t hostLimiter,
                       int32_t cached, int64_t n) {
    const int64_t tile = xcdBlock();
    const int r = tiles.run[tile];
    const int64_t s = tiles.first[tile] + threadIdx.x / Lanes;
    const int lane = (int)(threadIdx.x % Lanes);
    constexpr int K = D + D * D;          // acceleration, then M
    Real acc[K];
#pragma unroll
    for (int e = 0; e < K; ++e) acc[e] = Real(0);
    // with Parts, the energy budget's parts of a instead of a itself
    Real parts[4][3] = {};
    bool fluidI = false, correction = false;
    const bool runSteps = !resident || clock.active[r];
    const bool activeRun = s < runs[r].first + runs[r].count && runSteps;
    (void)n;
    NearCells<D, Real> cells;
    cells.r = rec.r;
    cells.held = 0;
    cells.stride = cached;
    cells.kicked = false;
    // 2D: the copy also holds each record's G for the limiter (the launch's LDS has room for it)
    cells.G = rec.G;
    cells.withG = D == 2 && sizeof(Real) == 4 && doAcceleration && (resident ? runs[r].limiter != 0 : hostLimiter != 0);
#pragma unroll
    for (int i = 0; i < NearCells<D, Real>::kRows; ++i) cells.lo[i] = cells.hi[i] = cells.at[i] = 0;
    if (runSteps && cached > 0) {
        const RunDev& run = runs[r];
        const int64_t first = tiles.first[tile];
        const int64_t last = min(first + (int64_t)(blockDim.x / Lanes), run.first + run.count) - 1;
        if (last >= first) {
            cells.plan(run, start, rec.key, first, last, cached);
            cells.fill();
        }
    }
    if (activeRun) {
        const RunDev& run = runs[r];
        // resident, L also feeds the renormalized diffusion's density gradients (sphGpu.py:2288-2291)
        correction = doCorrection && (!resident || run.correction || run.diffusion == 3);
        const bool limiter = resident ? run.limiter != 0 : hostLimiter != 0;
        const Self<D, Real> me = loadSelf<D, Real>(run, rec, s);
        fluidI = me.fluid;
        if (fluidI) {
            PairKernel<Real, D> kernel;
            kernel.init(run.kernel, run.sigmaUnit, run.hmax, run.uniform != 0);
            const double yy = run.period > 0.0 ? y64[order[s]] : 0.0;
            Real vi[3];
            velocityOf<D, Real>(rec.r[s], vi);
            const Real rhoI = densityOf<D, Real>(rec.r[s]);
            const Real own = pressureRatioOf<D, Real>(rec.r[s]);
            Real Gi[3][3] = {};
            if (limiter && doAcceleration)
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = 0; b < D; ++b) Gi[a][b] = rec.G[s * (D * D) + a * D + b];
            const Real alphaC0 = (Real)(run.alpha * run.c0), bf = (Real)run.boundaryFactor;
            Real eddyI = Real(0);
            if constexpr (Options) {
                if (out.eddy) eddyI = out.eddy[order[s]];
            }
            forEachNeighbor<D, Real, Lanes, true>(
                run, start, cells, list, s, me, yy, lane, kernel.uniform,
                [&](int32_t k, const Record<Real>& other, Real dx, Real dy, Real dz, Real rr, Real h) {
                    // G_j with the record: from the copy in 2D, else before the sums need it
                    Real Gj[D * D];
                    if (limiter && doAcceleration) {
                        if constexpr (D == 2) {
                            cells.gradient(k, Gj);
                        } else {
#pragma unroll
                            for (int e = 0; e < D * D; ++e) Gj[e] = rec.G[(int64_t)k * (D * D) + e];
                        }
                    }
                    const Real F = kernel.slopeOverR(rr, h);
                    const Real d[3] = {dx, dy, dz};
                    if (correction) {
                        const Real w = -volumeOf<D, Real>(other) * F;
#pragma unroll
                        for (int a = 0; a < D; ++a) {
                            const Real wa = w * d[a];
#pragma unroll
                            for (int b = 0; b < D; ++b) acc[D + a * D + b] += wa * d[b];
                        }
                    }
                    if (!doAcceleration) return;
                    Real vj[3];
                    velocityOf<D, Real>(other, vj);
                    const bool fluidJ = signedLength<D, Real>(other) > Real(0);
                    Real dv[3] = {};
                    Real approach = Real(0);
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        dv[a] = vi[a] - vj[a];
                        approach += dv[a] * d[a];
                    }
                    const Real rhoJ = densityOf<D, Real>(other);
                    Real viscosity = Real(0);
                    if (approach < Real(0) && limiter && fluidJ) {
                        Real rampI[3] = {}, rampJ[3] = {};
                        Real slopeI = Real(0), slopeJ = Real(0);
#pragma unroll
                        for (int a = 0; a < D; ++a) {
                            Real ti = Real(0), tj = Real(0);
#pragma unroll
                            for (int b = 0; b < D; ++b) {
                                ti += Gi[a][b] * d[b];
                                tj += Gj[a * D + b] * d[b];
                            }
                            rampI[a] = ti;
                            rampJ[a] = tj;
                            slopeI += d[a] * ti;
                            slopeJ += d[a] * tj;
                        }
                        Real limit = Real(0);
                        if (slopeI * slopeJ > Real(0)) {
                            const Real ratio = quotient(slopeI, slopeJ);
                            limit = fmin(Real(1), quotient(Real(4) * ratio, (Real(1) + ratio) * (Real(1) + ratio)));
                            const Real eta = kernel.uniform ? rr * kernel.invH : quotient(rr, h);
                            if (eta < Real(0.3)) {
                                const Real fade = (eta - Real(0.3)) * Real(5);
                                limit = limit * exp(-fade * fade);
                            }
                        }
                        approach = Real(0);
#pragma unroll
                        for (int a = 0; a < D; ++a)
                            approach += (dv[a] - Real(0.5) * limit * (rampI[a] + rampJ[a])) * d[a];
                    }
                    if (approach < Real(0)) {
                        const Real soft = kernel.uniform ? kernel.soft : Real(0.01) * h * h;
                        const Real mu = quotient(h * approach, rr * rr + soft);
                        viscosity = quotient(-alphaC0 * mu, Real(0.5) * (rhoI + rhoJ));
                        if (!fluidJ) viscosity *= bf;
                    }
                    if constexpr (Options) {
                        if (out.eddy && fluidJ) {
                            // the closure on every fluid pair: alpha_ij = K nu / (c0 h_ij) + the harmonic
                            // mean of the eddy alphas, with the unlimited v_ij . r_ij (sphGpu.py:388-393)
                            const Real eddyJ = out.eddy[order[k]];
                            const Real both = eddyI + eddyJ;
                            const Real mean = both > Real(0) ? Real(2) * eddyI * eddyJ / both : Real(0);
                            Real plain = Real(0);
#pragma unroll
                            for (int a = 0; a < D; ++a) plain += dv[a] * d[a];
                            viscosity += -(out.eddyWater / h + mean) * out.eddyC0 * h * out.eddyRho0 * plain /
                                         (rr * rr * rhoI * rhoJ);
                        }
                    }
                    const Real coefficient = -(own + pressureRatioOf<D, Real>(other) + viscosity) * massOf<D, Real>(other) * F;
#pragma unroll
                    for (int a = 0; a < D; ++a) acc[a] += coefficient * d[a];
                    if constexpr (Parts) {
                        const Real pressurePart = -(own + pressureRatioOf<D, Real>(other)) * massOf<D, Real>(other) * F;
                        const Real viscousPart = -viscosity * massOf<D, Real>(other) * F;
#pragma unroll
                        for (int a = 0; a < D; ++a) {
                            if (fluidJ) {
                                parts[0][a] += pressurePart * d[a];
                                parts[1][a] += viscousPart * d[a];
                            } else {
                                parts[2][a] += pressurePart * d[a];
                                parts[3][a] += viscousPart * d[a];
                            }
                        }
                    }
                });
        }
    }
    if constexpr (Parts) {
        // the energy budget's parts of a, sphGpu.py:396-414: this launch writes nothing else
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int a = 0; a < D; ++a) parts[k][a] = groupSum<Lanes>(parts[k][a]);
        if (activeRun && lane == 0) {
            const int32_t i = order[s];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int a = 0; a < D; ++a) out.parts[(k * D + a) * out.stride + i] = fluidI ? parts[k][a] : Real(0);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < K; ++e) acc[e] = groupSum<Lanes>(acc[e]);
    double magnitude = 0.0;
    if (activeRun && lane == 0) {
        const RunDev& run = runs[r];
        const int32_t i = order[s];
        if (!fluidI) {
            if (doAcceleration)
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    out.acceleration[a * out.stride + i] = Real(0);
                    if (resident) rec.A[s * D + a] = Real(0);
                }
            if (correction)
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = 0; b < D; ++b) {
                        const Real e = a == b ? Real(1) : Real(0);
                        rec.L[s * (D * D) + a * D + b] = e;
                        if (out.corrections) out.corrections[(a * D + b) * out.stride + i] = e;
                    }
        } else {
            if (doAcceleration) {
                Real a2 = Real(0);
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const Real v = acc[a] + (Real)run.gravity[a];
                    out.acceleration[a * out.stride + i] = v;
                    if (resident) rec.A[s * D + a] = v;
                    a2 += v * v;
                }
                magnitude = a2 < Real(INFINITY) ? (double)sqrt(a2) : INFINITY;
            }
            if (correction) {
                Real M[3][3] = {}, Lm[3][3] = {};
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = 0; b < D; ++b) M[a][b] = acc[D + a * D + b];
                const Real det = determinant<D, Real>(M);
                if (det >= (Real)correctionFloorPower<D>()) {
                    inverse<D, Real>(M, det, Lm);
                } else {
#pragma unroll
                    for (int a = 0; a < D; ++a) Lm[a][a] = Real(1);
                }
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = 0; b < D; ++b) {
                        rec.L[s * (D * D) + a * D + b] = Lm[a][b];
                        if (out.corrections) out.corrections[(a * D + b) * out.stride + i] = Lm[a][b];
                    }
            }
        }
    }
    if (resident && doAcceleration && activeRun && lane == 0) clock.acceleration[s] = fluidI ? magnitude : 0.0;
}

// The bed's boundary enforcement on one fluid particle (sphGpu.py:1197-1229, 1246-1310,
// 1332-1419; sphBoundaries.py:295-344, 597-640).
template <int D>
__device__ __forceinline__ void enforceBoundary(const sph_boundary& b, double* x, double* v) {
    const double clearance = b.clearance;
    if (b.kind == 1) {
#pragma unroll
        for (int axis = 0; axis < D; ++axis) {
            if (x[axis] < b.lower[axis]) { x[axis] = b.lower[axis]; v[axis] = fmax(v[axis], 0.0); }
            if (axis < D - 1 || b.openTop == 0)
                if (x[axis] > b.upper[axis]) { x[axis] = b.upper[axis]; v[axis] = fmin(v[axis], 0.0); }
        }
    } else if (b.kind == 2) {
        constexpr int vertical = D - 1;
        int low = 0, high = b.nBed - 1;
        while (high - low > 1) {
            const int middle = (low + high) / 2;
            if (b.bedX[middle] <= x[0]) low = middle; else high = middle;
        }
        const double run = b.bedX[low + 1] - b.bedX[low], rise = b.bedZ[low + 1] - b.bedZ[low];
        const double share = fmin(fmax((x[0] - b.bedX[low]) / run, 0.0), 1.0);
        const double bed = b.bedZ[low] + share * rise;
        if (x[vertical] < bed + clearance) {
            x[vertical] = bed + clearance;
            const double slope = rise / run, length = sqrt(1.0 + slope * slope);
            const double nx = -slope / length, nz = 1.0 / length;
            const double inward = v[0] * nx + v[vertical] * nz;
            if (inward < 0.0) { v[0] -= inward * nx; v[vertical] -= inward * nz; }
        }
        if (D == 3 && b.sideWalls) {
            if (x[1] < clearance) { x[1] = clearance; v[1] = fmax(v[1], 0.0); }
            if (x[1] > b.width - clearance) { x[1] = b.width - clearance; v[1] = fmin(v[1], 0.0); }
        }
        if (b.offshoreWall && x[0] < b.wallX + clearance) { x[0] = b.wallX + clearance; v[0] = fmax(v[0], b.pistonSpeed); }
        if (x[0] > b.endLimit) { x[0] = b.endLimit; v[0] = fmin(v[0], 0.0); }
    } else if (b.kind == 3 && D == 3) {
        const int nx = b.fieldNx, ny = b.fieldNy;
        const double rawX = (x[0] - b.originX) / b.stepX, rawY = (x[1] - b.originY) / b.stepY;
        const double fx = fmin(fmax(rawX, 0.0), (double)(nx - 1)), fy = fmin(fmax(rawY, 0.0), (double)(ny - 1));
        const int ix = (int)floor(fx), iy = (int)floor(fy);
        const double ux = fx - ix, uy = fy - iy;
        const double ux2 = ux * ux, ux3 = ux2 * ux, uy2 = uy * uy, uy3 = uy2 * uy;
        const double wx[4] = {-0.5 * ux3 + ux2 - 0.5 * ux, 1.5 * ux3 - 2.5 * ux2 + 1.0, -1.5 * ux3 + 2.0 * ux2 + 0.5 * ux,
                              0.5 * ux3 - 0.5 * ux2};
        const double wy[4] = {-0.5 * uy3 + uy2 - 0.5 * uy, 1.5 * uy3 - 2.5 * uy2 + 1.0, -1.5 * uy3 + 2.0 * uy2 + 0.5 * uy,
                              0.5 * uy3 - 0.5 * uy2};
        const double sx[4] = {-1.5 * ux2 + 2.0 * ux - 0.5, 4.5 * ux2 - 5.0 * ux, -4.5 * ux2 + 4.0 * ux + 0.5, 1.5 * ux2 - ux};
        const double sy[4] = {-1.5 * uy2 + 2.0 * uy - 0.5, 4.5 * uy2 - 5.0 * uy, -4.5 * uy2 + 4.0 * uy + 0.5, 1.5 * uy2 - uy};
        double bed = 0.0, slopeX = 0.0, slopeY = 0.0;
        for (int a = 0; a < 4; ++a) {
            const int row = min(max(iy - 1 + a, 0), ny - 1);
            double along = 0.0, alongSlope = 0.0;
            for (int c = 0; c < 4; ++c) {
                const double value = b.table[(int64_t)row * nx + min(max(ix - 1 + c, 0), nx - 1)];
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
            const double length = sqrt(slopeX * slopeX + slopeY * slopeY + 1.0);
            const double nrm[3] = {-slopeX / length, -slopeY / length, 1.0 / length};
            const double inward = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
            if (inward < 0.0)
                for (int a = 0; a < 3; ++a) v[a] -= inward * nrm[a];
        }
        if (x[1] < b.lowY) { x[1] = b.lowY; v[1] = fmax(v[1], 0.0); }
        if (x[1] > b.highY) { x[1] = b.highY; v[1] = fmin(v[1], 0.0); }
        if (b.offshoreWall && x[0] < b.lowX) { x[0] = b.lowX; v[0] = fmax(v[0], 0.0); }
        if (x[0] > b.highX) { x[0] = b.highX; v[0] = fmin(v[0], 0.0); }
    }
}

// Sweep 3: continuity with the delta-SPH term (sphGpu.py:416-545). Resident, it is also the
// step's integration: the sums read every velocity after the kick v += a dt (:1084-1109, taken in
// float32 on the records as they are read), and each particle's float64 kick, density advance
// with its floor (:1111-1140), drift (:1142-1186) and boundary enforcement follow; the run's first
// particle advances its clock.
template <int D, typename Real, int Lanes, bool Options = false, bool Parts = false>
__global__ void SWEEP_BOUNDS((sweepWaves<D, Real>()))
    sweepContinuity(const RunDev* __restrict__ runs, Tiles tiles, const int32_t* __restrict__ order, const int32_t* __restrict__ start, Records<Real> rec,
                    NeighborList list, SweepOut<Real> out, Clock clock, int32_t resident, int32_t hostCorrection,
                    int32_t hostDiffusion, double forcedDt, int32_t cached, int64_t n) {
    const int64_t tile = xcdBlock();
    const int r = tiles.run[tile];
    const int64_t s = tiles.first[tile] + threadIdx.x / Lanes;
    const int lane = (int)(threadIdx.x % Lanes);
    Real total = Real(0);
    // with Parts, the energy budget's parts of the rate of a fluid particle instead of the step: the
    // plain sum over fluid and over boundary neighbors, the whole continuity sum, and the diffusion
    Real plainFluid = Real(0), plainWall = Real(0), continuitySum = Real(0), diffusive = Real(0);
    bool boundaryI = false;
    const bool runSteps = !resident || clock.active[r];
    const bool activeRun = s < runs[r].first + runs[r].count && runSteps;
    (void)n;
    NearCells<D, Real> cells;
    cells.r = rec.r;
    cells.held = 0;
    cells.stride = cached;
    // resident, the step's kick v += a dt (sphGpu.py:1084-1109) reaches every velocity the sums read
    const double dt = resident && runSteps ? (forcedDt > 0.0 ? forcedDt : runTimeStep(runs[r], clock, r)) : 0.0;
    cells.kicked = resident != 0;
    cells.A = rec.A;
    cells.dt = (Real)dt;
#pragma unroll
    for (int i = 0; i < NearCells<D, Real>::kRows; ++i) cells.lo[i] = cells.hi[i] = cells.at[i] = 0;
    if (runSteps && cached > 0) {
        const RunDev& run = runs[r];
        const int64_t first = tiles.first[tile];
        const int64_t last = min(first + (int64_t)(blockDim.x / Lanes), run.first + run.count) - 1;
        if (last >= first) {
            cells.plan(run, start, rec.key, first, last, cached);
            cells.fill();
        }
    }
    if (activeRun) {
        const RunDev& run = runs[r];
        const Self<D, Real> me = loadSelf<D, Real>(run, rec, s);
        boundaryI = !me.fluid;
        const int diffusion = resident ? run.diffusion : hostDiffusion;
        const bool correction = (resident ? run.correction != 0 : hostCorrection != 0) && !boundaryI;
        const bool diffuses = diffusion != 0 && !boundaryI;
        PairKernel<Real, D> kernel;
        kernel.init(run.kernel, run.sigmaUnit, run.hmax, run.uniform != 0);
        const double yy = run.period > 0.0 ? out.pos[1][order[s]] : 0.0;
        Real vi[3];
        velocityOf<D, Real>(cells((int32_t)s), vi);
        Real Li[3][3] = {};
        if (correction) {
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) Li[a][b] = rec.L[s * (D * D) + a * D + b];
        }
        const Real devI = deviationOf<D, Real>(rec.r[s]);
        Real gradI[3] = {Real(0), Real(0), Real(0)};
        Real eddyI = Real(0);
        const bool tracked = Parts && !boundaryI;
        if constexpr (Options) {
            if (diffusion == 3)
#pragma unroll
                for (int a = 0; a < D; ++a) gradI[a] = rec.G[s * D + a];
            if (out.eddy) eddyI = out.eddy[out.stride + order[s]];
        }
        const Real factor = (Real)(2.0 * run.delta * run.c0);
        const Real invRho0 = (Real)(1.0 / run.rho0);
        const Real head = (Real)(run.rho0 * fabs(run.gravity[D - 1]) / (run.c0 * run.c0));
        forEachNeighbor<D, Real, Lanes, true>(
            run, start, cells, list, s, me, yy, lane, kernel.uniform,
            [&](int32_t k, const Record<Real>& other, Real dx, Real dy, Real dz, Real rr, Real h) {
                const bool fluidJ = signedLength<D, Real>(other) > Real(0);
                if (boundaryI && !fluidJ) return;
                const Real F = kernel.slopeOverR(rr, h);
                const Real d[3] = {dx, dy, dz};
                Real vj[3];
                velocityOf<D, Real>(other, vj);
                Real dot = Real(0);
                if (correction) {
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        Real ld = Real(0);
#pragma unroll
                        for (int b = 0; b < D; ++b) ld += Li[a][b] * d[b];
                        dot += (vi[a] - vj[a]) * ld;
                    }
                } else {
#pragma unroll
                    for (int a = 0; a < D; ++a) dot += (vi[a] - vj[a]) * d[a];
                }
                total += massOf<D, Real>(other) * dot * F;
                if constexpr (Parts) {
                    if (tracked) {
                        continuitySum += massOf<D, Real>(other) * dot * F;
                        Real plain = Real(0);
#pragma unroll
                        for (int a = 0; a < D; ++a) plain += (vi[a] - vj[a]) * d[a];
                        if (fluidJ) plainFluid += massOf<D, Real>(other) * plain * F;
                        else plainWall += massOf<D, Real>(other) * plain * F;
                    }
                }
                if (diffuses && fluidJ) {
                    const Real devJ = deviationOf<D, Real>(other);
                    Real difference = devJ - devI;
                    if (diffusion == 2) {
                        const Real ratio = Real(1) + Real(0.5) * (devI + devJ) * invRho0;
                        const Real sq = ratio * ratio;
                        difference -= quotient(head * d[D - 1], sq * sq * ratio);
                    }
                    Real pairFactor = factor;
                    if constexpr (Options) {
                        if (diffusion == 3) {
                            // less 1/2 (grad rho_i + grad rho_j) . (r_j - r_i), d = r_i - r_j (sphGpu.py:521-524)
                            Real projection = Real(0);
#pragma unroll
                            for (int a = 0; a < D; ++a) projection += (gradI[a] + rec.G[(int64_t)k * D + a]) * d[a];
                            difference += Real(0.5) * projection;
                        }
                        if (out.eddy) {
                            // 2 c0 times the harmonic mean of the two deltas (sphGpu.py:526-531)
                            const Real eddyJ = out.eddy[out.stride + order[k]];
                            const Real both = eddyI + eddyJ;
                            pairFactor = both > Real(0) ? Real(2) * out.eddyC0 * Real(2) * eddyI * eddyJ / both : Real(0);
                        }
                    }
                    const Real soft = kernel.uniform ? kernel.soft : Real(0.01) * h * h;
                    const Real r2 = rr * rr;
                    total += pairFactor * h * volumeOf<D, Real>(other) * difference * quotient(-F * r2, r2 + soft);
                    if constexpr (Parts) {
                        if (tracked)
                            diffusive += pairFactor * h * volumeOf<D, Real>(other) * difference *
                                         quotient(-F * r2, r2 + soft);
                    }
                }
            });
    }
    if constexpr (Parts) {
        // the energy budget's parts of the rate (sphGpu.py:535-545): this launch writes nothing else
        plainFluid = groupSum<Lanes>(plainFluid);
        plainWall = groupSum<Lanes>(plainWall);
        continuitySum = groupSum<Lanes>(continuitySum);
        diffusive = groupSum<Lanes>(diffusive);
        if (activeRun && lane == 0) {
            const int32_t i = order[s];
            const bool tracked = !boundaryI;
            const bool corrected = resident ? runs[r].correction != 0 : hostCorrection != 0;
            out.rateParts[i] = tracked ? plainFluid : Real(0);
            out.rateParts[out.stride + i] = tracked ? plainWall : Real(0);
            out.rateParts[2 * out.stride + i] = tracked && corrected ? continuitySum - plainFluid - plainWall : Real(0);
            out.rateParts[3 * out.stride + i] = tracked ? diffusive : Real(0);
        }
        return;
    }
    total = groupSum<Lanes>(total);
    LdsReal<Real>* packed = packedSums<Real>(cached);
    if (lane == 0) packed[threadIdx.x / Lanes] = total;
    __syncthreads();
    const int64_t own = tiles.first[tile] + threadIdx.x;
    if (threadIdx.x >= blockDim.x / Lanes || !(own < runs[r].first + runs[r].count && runSteps)) return;
    total = packed[threadIdx.x];
    const RunDev& run = runs[r];
    const int32_t i = order[own];
    if (!resident) {
        out.rate[i] = total;
        return;
    }
    if (own == run.first) {
        clock.dt[r] = dt;
        clock.time[r] += dt;
        clock.steps[r] += 1;
    }
    const bool boundaryOwn = !(signedLength<D, Real>(rec.r[own]) > Real(0));
    const double before = out.density64[i];
    double rho = before + (double)total * dt;
    const double unfloored = rho;
    if (run.densityFloor > 0.0 && !boundaryOwn) rho = fmax(rho, run.densityFloor);
    out.density64[i] = rho;
    if (boundaryOwn) return;
    // the kick in float64, then the drift
    double x[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; ++a) {
        v[a] = out.vel[a][i] + (double)out.acceleration[a * out.stride + i] * dt;
        x[a] = out.pos[a][i] + v[a] * dt;
    }
    // the energy budget: each term's change of K + P + E_C over this step (sphGpu.py:1084-1186,
    // 1191-1419), summed per fluid particle
    // (it reads the step's results, never its intermediate products, so keeping it leaves the
    // step's arithmetic alone; the pair parts come from the parts-only launches before this one)
    double* L = nullptr;
    double startV[3] = {0.0, 0.0, 0.0}, kicked[3] = {0.0, 0.0, 0.0}, g[3] = {0.0, 0.0, 0.0};
    double m = 0.0;
    {
        if (out.ledger) {
            L = out.ledger + i;
            const int64_t ls = out.ledgerStride;
            m = (double)massOf<D, Real>(rec.r[own]);
            const double scale = run.c0 * run.c0 / 7.0;
            const double slope = elasticSlope(before, unfloored, run.rho0, scale, run.negativePressure);
#pragma unroll
            for (int k = 0; k < 4; ++k) L[(5 + k) * ls] += m * slope * (double)out.rateParts[k * out.stride + i] * dt;
            L[9 * ls] += m * (elasticEnergy(rho, run.rho0, scale, run.negativePressure) -
                              elasticEnergy(unfloored, run.rho0, scale, run.negativePressure));
            double mean[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int a = 0; a < D; ++a) {
                startV[a] = out.vel[a][i];
                mean[a] = 0.5 * (startV[a] + v[a]);
                g[a] = run.gravity[a];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double work = 0.0;
#pragma unroll
                for (int a = 0; a < D; ++a) work += (double)out.parts[(k * D + a) * out.stride + i] * mean[a];
                L[k * ls] += m * work * dt;
            }
            double fall = 0.0, rise = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                fall += g[a] * mean[a];
                rise += g[a] * (x[a] - out.pos[a][i]);
            }
            L[4 * ls] += m * fall * dt;
            L[10 * ls] -= m * rise;
        }
    }
    if (run.period > 0.0) x[1] = x[1] - run.period * floor(x[1] / run.period);
    if (L)
#pragma unroll
        for (int a = 0; a < D; ++a) {
            startV[a] = x[a];
            kicked[a] = v[a];
        }
    enforceBoundary<D>(run.boundary, x, v);
    {
        if (L) {
            double kinetic = 0.0, potential = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                kinetic += v[a] * v[a] - kicked[a] * kicked[a];
                potential += g[a] * (x[a] - startV[a]);
            }
            L[11 * out.ledgerStride] += m * (0.5 * kinetic - potential);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        out.pos[a][i] = x[a];
        out.vel[a][i] = v[a];
    }
}

template <int D>
__global__ void enforceKernel(const RunDev* __restrict__ runs, const int32_t* __restrict__ runOf,
                              const uint8_t* __restrict__ fluid, double* px, double* py, double* pz, double* vx, double* vy,
                              double* vz, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !fluid[i]) return;
    const RunDev& run = runs[runOf[i]];
    double* P[3] = {px, py, pz};
    double* V[3] = {vx, vy, vz};
    double x[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < D; ++a) { x[a] = P[a][i]; v[a] = V[a][i]; }
    enforceBoundary<D>(run.boundary, x, v);
    for (int a = 0; a < D; ++a) { P[a][i] = x[a]; V[a][i] = v[a]; }
}

// Neighbor counts (rows null) or rows as slots, in walk order: from the list where the particle's
// neighbors fit it, from its cells otherwise.
template <int D, typename Real>
__global__ void SWEEP_BOUNDS((sweepWaves<D, Real>()))
    neighborExport(const RunDev* __restrict__ runs, const int32_t* __restrict__ runOf, const int32_t* __restrict__ order,
                   const int32_t* __restrict__ start, Records<Real> rec, NeighborList list,
                   const double* __restrict__ y64, int32_t* counts, const int64_t* offsets, int32_t* rows, int64_t n) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const RunDev& run = runs[runOf[s]];
    const Self<D, Real> me = loadSelf<D, Real>(run, rec, s);
    const int32_t i = order[s];
    const double yy = run.period > 0.0 ? y64[i] : 0.0;
    int32_t count = 0;
    const int64_t base = rows ? offsets[i] : 0;
    forEachNeighbor<D, Real, 1, false>(run, start, GlobalRecords<Real>{rec.r}, list, s, me, yy, 0, run.uniform != 0,
                                       [&](int32_t k, const Record<Real>&, Real, Real, Real, Real, Real) {
                                           if (rows) rows[base + count] = order[k];
                                           ++count;
                                       });
    if (counts) counts[i] = count;
}

// ---------------------------------------------------------------------------------------------
// The breaking read-out
// ---------------------------------------------------------------------------------------------

// A float64's bits as an unsigned key ordered as the values are; 0 lies below every key.
__device__ __forceinline__ unsigned long long orderedKey(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Whether slot i is one of its run's tracked particles (fluid; in 3D within halfWidth of the strip's
// centre line), and its (x, z).
template <int D>
__device__ __forceinline__ bool trackedPoint(const sph_state& st, int64_t i, double centre, double halfWidth, double& x,
                                             double& z) {
    if (!st.fluid[i]) return false;
    x = st.position[i];
    z = st.position[(D - 1) * st.stride + i];
    return D == 2 || fabs(st.position[st.stride + i] - centre) < halfWidth;
}

// Each column's highest tracked particle: floor((x - x0) / width) as freeSurfaceProfile takes it
// (waveBreaking.py:129-162), the maximum by an atomic on ordered keys, so in any order the same.
template <int D>
__global__ void surfaceTops(const int32_t* __restrict__ runOf, sph_state st, const sph_profile_spec* __restrict__ specs,
                            unsigned long long* __restrict__ tops, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sph_profile_spec& spec = specs[runOf[i]];
    if (spec.columns <= 0) return;
    double x, z;
    if (!trackedPoint<D>(st, i, spec.stripCentre, spec.stripHalfWidth, x, z)) return;
    const double column = floor((x - spec.x0) / spec.width);
    if (!(column >= 0.0 && column < (double)spec.columns)) return;
    atomicMax(&tops[spec.offset + (int64_t)column], orderedKey(z));
}

// The face window's cells, one spacing square, as frontFaceLean marks them (waveBreaking.py:
// 559-634): each tracked particle in the window marks the 2 x 2 cells its own cell overlaps.
template <int D>
__global__ void faceCells(const int32_t* __restrict__ runOf, sph_state st, const sph_face_spec* __restrict__ specs,
                          uint32_t* __restrict__ occupancy, int32_t* __restrict__ overflow, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    const sph_face_spec& spec = specs[r];
    if (!spec.active) return;
    double x, z;
    if (!trackedPoint<D>(st, i, spec.stripCentre, spec.stripHalfWidth, x, z)) return;
    const double s = spec.spacing;
    if (!(x >= spec.windowStart - 0.5 * s && x < spec.windowEnd + 0.5 * s && z > spec.base - 0.5 * s)) return;
    const double firstColumn = floor((x - 0.5 * s - spec.windowStart) / s);
    const double firstBand = floor((z - 0.5 * s - spec.base) / s);
    const int64_t words = (spec.columns + 31) / 32;
    for (int columnStep = 0; columnStep < 2; ++columnStep)
        for (int bandStep = 0; bandStep < 2; ++bandStep) {
            const double column = firstColumn + columnStep, band = firstBand + bandStep;
            if (!(column >= 0.0 && column < (double)spec.columns && band >= 0.0)) continue;
            if (band >= (double)spec.bands) {
                overflow[r] = 1;
                continue;
            }
            const int64_t c = (int64_t)column;
            atomicOr(&occupancy[spec.offset + (int64_t)band * words + c / 32], 1u << (c % 32));
        }
}

// z of the bed's piecewise-linear profile at x, constant beyond its ends.
__device__ __forceinline__ double bedAt(const sph_face_spec& spec, double x) {
    const int32_t m = spec.bedPoints;
    if (x <= spec.bedX[0]) return spec.bedZ[0];
    if (x >= spec.bedX[m - 1]) return spec.bedZ[m - 1];
    int32_t low = 0, high = m - 1;
    while (high - low > 1) {
        const int32_t middle = (low + high) / 2;
        if (spec.bedX[middle] <= x) low = middle; else high = middle;
    }
    const double share = (x - spec.bedX[low]) / (spec.bedX[high] - spec.bedX[low]);
    return spec.bedZ[low] + share * (spec.bedZ[high] - spec.bedZ[low]);
}

constexpr int kLandward = 8;

// Inserts (x, z) into a list of kLandward points kept by descending x.
__device__ __forceinline__ void keepLandward(double (&px)[kLandward], double (&pz)[kLandward], double x, double z) {
    if (!(x > px[kLandward - 1])) return;
#pragma unroll
    for (int m = 0; m < kLandward; ++m) {
        if (x > px[m]) {
            const double tx = px[m], tz = pz[m];
            px[m] = x;
            pz[m] = z;
            x = tx;
            z = tz;
        }
    }
}

// The kLandward most landward tracked particles near the bed among `chunk` slots of a run, one
// workgroup per (run, slice): each thread keeps its own, and the workgroup merges them pairwise into
// the slice's list at partial[(r * slices + slice) * kLandward].
template <int D>
__global__ void __launch_bounds__(256)
    landwardPoints(const RunDev* __restrict__ runs, sph_state st, const sph_face_spec* __restrict__ specs,
                   int64_t chunk, double2* __restrict__ partial) {
    __shared__ double sx[256][kLandward], sz[256][kLandward];
    const int r = blockIdx.x;
    const sph_face_spec& spec = specs[r];
    const RunDev& run = runs[r];
    double px[kLandward], pz[kLandward];
#pragma unroll
    for (int m = 0; m < kLandward; ++m) px[m] = pz[m] = -INFINITY;
    if (spec.active && spec.bedPoints > 0) {
        const int64_t begin = run.first + (int64_t)blockIdx.y * chunk;
        const int64_t end = min(run.first + run.count, begin + chunk);
        for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
            double x, z;
            if (!trackedPoint<D>(st, i, spec.stripCentre, spec.stripHalfWidth, x, z)) continue;
            if (z - bedAt(spec, x) < spec.wetHeight + 1e-9) keepLandward(px, pz, x, z);
        }
    }
#pragma unroll
    for (int m = 0; m < kLandward; ++m) {
        sx[threadIdx.x][m] = px[m];
        sz[threadIdx.x][m] = pz[m];
    }
    __syncthreads();
    for (unsigned half = blockDim.x / 2; half > 0; half /= 2) {
        if (threadIdx.x < half) {
#pragma unroll
            for (int m = 0; m < kLandward; ++m) keepLandward(px, pz, sx[threadIdx.x + half][m], sz[threadIdx.x + half][m]);
#pragma unroll
            for (int m = 0; m < kLandward; ++m) {
                sx[threadIdx.x][m] = px[m];
                sz[threadIdx.x][m] = pz[m];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < kLandward; ++m)
            partial[((int64_t)r * gridDim.y + blockIdx.y) * kLandward + m] = make_double2(px[m], pz[m]);
}

// Each run's slices merged into its kLandward points, (x, z) by descending x, NaN past the last: one
// 64-thread workgroup per run, each thread merging every 64th slice, then the threads pairwise.
__global__ void __launch_bounds__(64)
    mergeLandward(const double2* __restrict__ partial, int32_t slices, double* __restrict__ landward) {
    __shared__ double sx[64][kLandward], sz[64][kLandward];
    const int32_t r = blockIdx.x;
    double px[kLandward], pz[kLandward];
#pragma unroll
    for (int m = 0; m < kLandward; ++m) px[m] = pz[m] = -INFINITY;
    for (int64_t k = threadIdx.x; k < (int64_t)slices * kLandward; k += blockDim.x) {
        const double2 p = partial[(int64_t)r * slices * kLandward + k];
        keepLandward(px, pz, p.x, p.y);
    }
#pragma unroll
    for (int m = 0; m < kLandward; ++m) {
        sx[threadIdx.x][m] = px[m];
        sz[threadIdx.x][m] = pz[m];
    }
    __syncthreads();
    for (unsigned half = blockDim.x / 2; half > 0; half /= 2) {
        if (threadIdx.x < half) {
#pragma unroll
            for (int m = 0; m < kLandward; ++m) keepLandward(px, pz, sx[threadIdx.x + half][m], sz[threadIdx.x + half][m]);
#pragma unroll
            for (int m = 0; m < kLandward; ++m) {
                sx[threadIdx.x][m] = px[m];
                sz[threadIdx.x][m] = pz[m];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < kLandward; ++m) {
            const bool kept = px[m] > -INFINITY;
            landward[(int64_t)r * 2 * kLandward + 2 * m] = kept ? px[m] : NAN;
            landward[(int64_t)r * 2 * kLandward + 2 * m + 1] = kept ? pz[m] : NAN;
        }
}

// The fluid fraction. Each run's tracked particles inside its window are sorted into bins of half a
// spacing: a bin grid reaching `reach` + 1 bins past the nodes on every side, (nx + 2 reach + 2) x
// (ny + 2 reach + 2), holds every particle that can reach a node.

// Each slot's bin (binOffset[run] + local bin), or -1; the bins' counts.
template <int D>
__global__ void fractionKeys(const int32_t* __restrict__ runOf, sph_state st, const sph_fraction_spec* __restrict__ specs,
                             const int64_t* __restrict__ binOffset, int32_t* __restrict__ key, int32_t* __restrict__ counts,
                             int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    const sph_fraction_spec& spec = specs[r];
    int32_t k = -1;
    double x, z;
    if (spec.nx > 0 && trackedPoint<D>(st, i, spec.stripCentre, spec.stripHalfWidth, x, z) && x > spec.lowX &&
        x < spec.highX && z > spec.lowZ && z < spec.highZ) {
        const double bx = floor((x - spec.x0) / spec.binWidth) + spec.reach + 1;
        const double bz = floor((z - spec.y0) / spec.binWidth) + spec.reach + 1;
        const int32_t width = spec.nx + 2 * spec.reach + 2, height = spec.ny + 2 * spec.reach + 2;
        if (bx >= 0.0 && bx < (double)width && bz >= 0.0 && bz < (double)height) {
            k = (int32_t)(binOffset[r] + (int64_t)bz * width + (int64_t)bx);
            atomicAdd(&counts[k], 1);
        }
    }
    key[i] = k;
}

// Slots into their bins, in any order (fill counts down from each bin's count).
__global__ void fractionScatter(const int32_t* __restrict__ key, const int32_t* __restrict__ start,
                                int32_t* __restrict__ fill, int32_t* __restrict__ members, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t k = key[i];
    if (k < 0) return;
    const int32_t at = atomicSub(&fill[k], 1) - 1;
    members[start[k] + at] = (int32_t)i;
}

// Each bin's slots in ascending order (a bin holds a particle or so).
__global__ void fractionOrder(const int32_t* __restrict__ start, int32_t* __restrict__ members, int64_t bins) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bins) return;
    const int32_t first = start[b], last = start[b + 1];
    for (int32_t k = first + 1; k < last; ++k) {
        const int32_t v = members[k];
        int32_t m = k - 1;
        while (m >= first && members[m] > v) {
            members[m + 1] = members[m];
            --m;
        }
        members[m + 1] = v;
    }
}

constexpr int kFractionTile = 16;

// The LDS a fractionNodes workgroup takes: the bin starts of its tile's reach.
inline size_t fractionLds(int32_t reach) {
    return (size_t)(kFractionTile + 2 * reach + 1) * (kFractionTile + 2 * reach + 2) * sizeof(int32_t);
}

// phi at each node of one run: the bins from `reach` + 1 below to `reach` above the node's own,
// rows outer; a bin's particles summed apart, then added. A workgroup takes a tile of 16 x 16 nodes
// and first copies the starts of the bins they reach into LDS, so a node's 196 bin lookups stay on
// the chip.
template <int D, typename Real>
__global__ void __launch_bounds__(256)
    fractionNodes(sph_state st, const sph_fraction_spec* __restrict__ specs, int32_t run, int64_t binOffset,
                  const int32_t* __restrict__ start, const int32_t* __restrict__ members, double* __restrict__ fraction) {
#pragma clang fp contract(off)
    extern __shared__ int32_t binStart[];
    const sph_fraction_spec& spec = specs[run];
    const int32_t R = spec.reach;
    const int32_t width = spec.nx + 2 * R + 2, height = spec.ny + 2 * R + 2;
    const int32_t tilesX = (spec.nx + kFractionTile - 1) / kFractionTile;
    const int32_t i0 = (int32_t)(blockIdx.x % tilesX) * kFractionTile, j0 = (int32_t)(blockIdx.x / tilesX) * kFractionTile;
    const int32_t rows = kFractionTile + 2 * R + 1, cols = kFractionTile + 2 * R + 2;
    // Bin (row, column) of the run's grid at local (r, c); a column one past a row's last bin is the
    // next row's first, whose start ends the row
    for (int32_t k = threadIdx.x; k < rows * cols; k += blockDim.x) {
        const int32_t row = j0 + k / cols, column = i0 + k % cols;
        binStart[k] = row < height && column <= width ? start[binOffset + (int64_t)row * width + column] : 0;
    }
    __syncthreads();
    const int32_t i = i0 + (int32_t)(threadIdx.x % kFractionTile), j = j0 + (int32_t)(threadIdx.x / kFractionTile);
    if (i >= spec.nx || j >= spec.ny) return;
    const double gx = spec.x0 + (double)i * spec.stepX, gz = spec.y0 + (double)j * spec.stepY;
    const double h = spec.h, sigma = 10.0 / (7.0 * M_PI * h * h);
    const Real* mass = static_cast<const Real*>(st.mass);
    double phi = 0.0;
    for (int32_t oy = -R; oy <= R + 1; ++oy)
        for (int32_t ox = -R; ox <= R + 1; ++ox) {
            // Particles whose bin is (i - ox, j - oy), shifted into the bin grid, then into the tile's
            const int32_t local = (j - oy + R + 1 - j0) * cols + (i - ox + R + 1 - i0);
            const int32_t first = binStart[local], last = binStart[local + 1];
            if (first == last) continue;
            double partial = 0.0;
            for (int32_t k = first; k < last; ++k) {
                const int32_t p = members[k];
                const double dx = gx - st.position[p];
                const double dz = gz - st.position[(D - 1) * st.stride + p];
                const double q = hypot(dx, dz) / h;
                double w = 0.0;
                if (q < 1.0)
                    w = sigma * (1.0 - 1.5 * (q * q) + 0.75 * (q * q * q));
                else if (q < 2.0) {
                    const double t = 2.0 - q;
                    w = sigma * 0.25 * (t * t * t);
                }
                const double area = spec.areaPerMass > 0.0 ? (double)mass[p] * spec.areaPerMass : spec.area;
                partial += area * w;
            }
            phi += partial;
        }
    fraction[spec.offset + (int64_t)j * spec.nx + i] = phi;
}

// ---------------------------------------------------------------------------------------------
// Global float64 sums
// ---------------------------------------------------------------------------------------------


// One partial of every sum over chunk b's slots (chunkSums, sph_sums.hpp).
template <int D, typename Real>
__global__ void __launch_bounds__(kSumThreads)
    sumChunks(const RunDev* __restrict__ runs, const int32_t* __restrict__ chunkRun,
              const int64_t* __restrict__ chunkFirst, sph_state state, double* __restrict__ partial) {
    __shared__ double tree[kSums][kSumThreads];
    const int64_t b = blockIdx.x;
    const RunDev& run = runs[chunkRun[b]];
    const int64_t first = chunkFirst[b];
    const int64_t end = min(first + kSumChunk, run.first + run.count);
    const SumPhysics physics{run.rho0, run.c0, -run.gravity[D - 1], run.negativePressure};
    chunkSums<D, Real>(physics, first, end, state, tree, partial + b * kSums);
}

// Each run's sums from its chunks' partials, in chunk order.
__global__ void sumRuns(const int64_t* __restrict__ runChunk, const double* __restrict__ partial, double* out,
                        int32_t runs) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    addChunks(partial, runChunk[r], runChunk[r + 1], out + (int64_t)r * kSums);
}

// ---------------------------------------------------------------------------------------------
// Shepard filter and the energy budget's ledger
// ---------------------------------------------------------------------------------------------

// A fluid particle of a run the filter takes this step gets its Shepard-filtered density (slot
// order, `filtered`), the boundary keeps its own (sphGpu.py:1040-1056); with a ledger the change of
// elastic energy goes to its filter row, 13.
template <typename Real>
__global__ void applyShepard(const RunDev* __restrict__ runs, const int32_t* __restrict__ runOf, sph_state state,
                             const Real* __restrict__ filtered, double* __restrict__ ledger, int64_t ledgerStride,
                             Clock clock, int32_t mode, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = runOf[i];
    const RunDev& run = runs[r];
    if (!filterDue(run, clock, r, mode) || !state.fluid[i]) return;
    const double before = state.density[i], after = (double)filtered[i];
    state.density[i] = after;
    if (ledger) {
        const double scale = run.c0 * run.c0 / 7.0;
        const double m = (double)static_cast<const Real*>(state.mass)[i];
        ledger[13 * ledgerStride + i] += m * (elasticEnergy(after, run.rho0, scale, run.negativePressure) -
                                              elasticEnergy(before, run.rho0, scale, run.negativePressure));
    }
}

constexpr int kLedgerTerms = 14;   // sphEnergy.py deviceTerms

// One partial of each ledger row over chunk b's slots, each thread over a fixed stride and then a
// fixed tree, so a run's totals do not depend on the other runs or on the launch.
__global__ void __launch_bounds__(kSumThreads)
    ledgerChunks(const RunDev* __restrict__ runs, const int32_t* __restrict__ chunkRun,
                 const int64_t* __restrict__ chunkFirst, const double* __restrict__ ledger, int64_t ledgerStride,
                 double* __restrict__ partial) {
    __shared__ double tree[kLedgerTerms][kSumThreads];
    const int64_t b = blockIdx.x;
    const RunDev& run = runs[chunkRun[b]];
    const int64_t first = chunkFirst[b];
    const int64_t end = min(first + kSumChunk, run.first + run.count);
    double acc[kLedgerTerms] = {};
    for (int64_t i = first + threadIdx.x; i < end; i += kSumThreads)
#pragma unroll
        for (int t = 0; t < kLedgerTerms; ++t) acc[t] += ledger[t * ledgerStride + i];
#pragma unroll
    for (int t = 0; t < kLedgerTerms; ++t) tree[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int width = kSumThreads / 2; width > 0; width >>= 1) {
        if ((int)threadIdx.x < width)
            for (int t = 0; t < kLedgerTerms; ++t) tree[t][threadIdx.x] += tree[t][threadIdx.x + width];
        __syncthreads();
    }
    if (threadIdx.x < kLedgerTerms) partial[b * kLedgerTerms + threadIdx.x] = tree[threadIdx.x][0];
}

// Each run's ledger totals from its chunks' partials, in chunk order.
__global__ void ledgerRuns(const int64_t* __restrict__ runChunk, const double* __restrict__ partial,
                           double* __restrict__ out, int32_t runs) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    double acc[kLedgerTerms] = {};
    for (int64_t b = runChunk[r]; b < runChunk[r + 1]; ++b)
        for (int t = 0; t < kLedgerTerms; ++t) acc[t] += partial[b * kLedgerTerms + t];
    for (int t = 0; t < kLedgerTerms; ++t) out[(int64_t)r * kLedgerTerms + t] = acc[t];
}

// ---------------------------------------------------------------------------------------------
// State order
// ---------------------------------------------------------------------------------------------

// Slot s of the bound state takes what slot order[s] held (read from the copy `from`), in every
// field and in `identity`: each run's particles in the cell order of the last cell build, within the
// run's own slots (WcsphSolver.sortParticles, sphSolver.py:2093-2126, on the device).
template <typename Real>
__global__ void permuteState(const int32_t* __restrict__ order, sph_state from, sph_state to,
                             const int32_t* __restrict__ identityFrom, int32_t* identityTo, int32_t dims, int64_t n) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t i = order[s], stride = to.stride;
    for (int a = 0; a < dims; ++a) {
        to.position[a * stride + s] = from.position[a * stride + i];
        to.velocity[a * stride + s] = from.velocity[a * stride + i];
        static_cast<Real*>(to.acceleration)[a * stride + s] = static_cast<const Real*>(from.acceleration)[a * stride + i];
    }
    to.density[s] = from.density[i];
    static_cast<Real*>(to.pressure)[s] = static_cast<const Real*>(from.pressure)[i];
    static_cast<Real*>(to.mass)[s] = static_cast<const Real*>(from.mass)[i];
    static_cast<Real*>(to.length)[s] = static_cast<const Real*>(from.length)[i];
    to.fluid[s] = from.fluid[i];
    if (identityTo) identityTo[s] = identityFrom[i];
}

// -- Splitting on the device -- //

// 1 where slot i splits: fluid, in its run's zone, heavy enough, and in the caller's mask if any.
__global__ void splitMarks(const int32_t* __restrict__ runOf, sph_state st, const sph_split_spec* __restrict__ specs,
                           const uint8_t* __restrict__ mask, int32_t precision, int32_t* __restrict__ marks, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sph_split_spec& spec = specs[runOf[i]];
    const double x = st.position[i];
    const double m = precision == 0 ? (double)static_cast<const float*>(st.mass)[i] : static_cast<const double*>(st.mass)[i];
    marks[i] = spec.active && st.fluid[i] && x >= spec.lowX && x < spec.highX && m >= spec.minMass &&
               (mask == nullptr || mask[i] != 0);
}

// Each run's rank of its first slot, and the total, from the exclusive scan of the marks.
__global__ void runRanks(const RunDev* __restrict__ runs, const int32_t* __restrict__ ranks, int32_t count, int64_t n,
                         int64_t* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > count) return;
    out[r] = r < count ? ranks[runs[r].first] : ranks[n];
}

// Slot i of the old layout into the new: every particle to its run's new place, a mother's first
// daughter in her slot and the others after the run's own particles (splitParticles,
// sphRefinement.py:135-186; the corners ordered as itertools.product((-1/4, 1/4), repeat = dims)).
template <typename Real>
__global__ void splitScatter(const int32_t* __restrict__ runOf, const RunDev* __restrict__ runs,
                             const int64_t* __restrict__ newFirst, const int64_t* __restrict__ firstRank,
                             const int32_t* __restrict__ marks, const int32_t* __restrict__ ranks, sph_state from,
                             sph_state to, int32_t dims, double referenceDensity, double lengthRatio, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    const int64_t base = newFirst[r], local = i - runs[r].first;
    const int64_t fs = from.stride, ts = to.stride;
    const Real* massFrom = static_cast<const Real*>(from.mass);
    const Real* lengthFrom = static_cast<const Real*>(from.length);
    const Real* pressureFrom = static_cast<const Real*>(from.pressure);
    const Real* accelerationFrom = static_cast<const Real*>(from.acceleration);
    Real* massTo = static_cast<Real*>(to.mass);
    Real* lengthTo = static_cast<Real*>(to.length);
    Real* pressureTo = static_cast<Real*>(to.pressure);
    Real* accelerationTo = static_cast<Real*>(to.acceleration);
    const int daughters = 1 << dims;
    const bool splits = marks[i] != 0;
    const Real mass = massFrom[i];
    double spacing = 0.0;
    if (splits) {
        const double volume = (double)mass / referenceDensity;
        spacing = dims == 2 ? sqrt(volume) : cbrt(volume);
    }
    for (int k = 0; k < (splits ? daughters : 1); ++k) {
        const int64_t d = k == 0 ? base + local
                                 : base + runs[r].count + (int64_t)(ranks[i] - firstRank[r]) * (daughters - 1) + (k - 1);
        for (int a = 0; a < dims; ++a) {
            // Bit (dims - 1 - a) of k picks the corner along axis a: the last axis varies fastest
            const double corner = ((k >> (dims - 1 - a)) & 1) ? 0.25 : -0.25;
            to.position[a * ts + d] = from.position[a * fs + i] + (splits ? spacing * corner : 0.0);
            to.velocity[a * ts + d] = from.velocity[a * fs + i];
            accelerationTo[a * ts + d] = accelerationFrom[a * fs + i];
        }
        to.density[d] = from.density[i];
        pressureTo[d] = pressureFrom[i];
        massTo[d] = splits ? mass / (Real)daughters : mass;
        lengthTo[d] = splits ? (Real)((double)lengthFrom[i] * lengthRatio) : lengthFrom[i];
        to.fluid[d] = from.fluid[i];
    }
}

// -- Merging on the device -- //

constexpr int kMergeLevels = 8;

// One pass of a run's merge: the level it groups, the cells' size and shift, and the cell grid's
// lowest integer coordinates, extent and place in the pass's cell array.
struct MergePass {
    int32_t active, level;
    double cell, offset;
    long long low[3], extent[3];
    long long cellBase;
};

__device__ __forceinline__ double massOf(const sph_state& st, int32_t precision, int64_t i) {
    return precision == 0 ? (double)static_cast<const float*>(st.mass)[i] : static_cast<const double*>(st.mass)[i];
}

// Each slot's level if it may merge (0 if not), each run's count per level and its last boundary slot.
__global__ void mergeLevels(const int32_t* __restrict__ runOf, sph_state st, const sph_merge_spec* __restrict__ specs,
                            const uint8_t* __restrict__ mask, int32_t precision, int32_t dims, int32_t* __restrict__ levelOf,
                            int32_t* __restrict__ histogram, long long* __restrict__ lastBoundary, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    levelOf[i] = 0;
    if (!st.fluid[i]) {
        atomicMax(&lastBoundary[r], (long long)i);
        return;
    }
    const sph_merge_spec& spec = specs[r];
    const double x = st.position[i];
    if (!spec.active || !(x >= spec.lowX && x < spec.highX) || (mask && !mask[i])) return;
    const int level = (int)rint(log2(spec.referenceMass / massOf(st, precision, i)) / dims);
    if (level <= 0 || level >= kMergeLevels) return;
    levelOf[i] = level;
    atomicAdd(&histogram[r * kMergeLevels + level], 1);
}

// The scaled coordinate of slot i along axis a in its run's pass grid.
__device__ __forceinline__ double mergeScaled(const sph_state& st, const MergePass& pass, int a, int64_t i) {
    return st.position[a * st.stride + i] / pass.cell + pass.offset;
}

// The lowest and highest integer cell coordinates of each run's free particles of the pass's level.
__global__ void mergeBounds(const int32_t* __restrict__ runOf, sph_state st, const MergePass* __restrict__ passes,
                            int32_t dims, const int32_t* __restrict__ levelOf, long long* __restrict__ low,
                            long long* __restrict__ high, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    const MergePass& pass = passes[r];
    if (!pass.active || levelOf[i] != pass.level) return;
    for (int a = 0; a < dims; ++a) {
        const long long k = (long long)floor(mergeScaled(st, pass, a, i));
        atomicMin(&low[r * 3 + a], k);
        atomicMax(&high[r * 3 + a], k);
    }
}

// Each free particle's cell (row-major over its run's grid, after the run's cells) and squared
// distance from the cell's centre in cells, and the cells' counts.
__global__ void mergeKeys(const int32_t* __restrict__ runOf, sph_state st, const MergePass* __restrict__ passes,
                          int32_t dims, const int32_t* __restrict__ levelOf, int32_t* __restrict__ key,
                          double* __restrict__ distance, int32_t* __restrict__ counts, int64_t n) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    const MergePass& pass = passes[r];
    key[i] = -1;
    if (!pass.active || levelOf[i] != pass.level) return;
    long long flat = 0;
    double sum = 0.0;
    for (int a = 0; a < dims; ++a) {
        const double scaled = mergeScaled(st, pass, a, i);
        const double corner = floor(scaled);
        flat = flat * pass.extent[a] + ((long long)corner - pass.low[a]);
        const double t = scaled - corner - 0.5;
        sum = sum + t * t;
    }
    const int32_t k = (int32_t)(pass.cellBase + flat);
    key[i] = k;
    distance[i] = sum;
    atomicAdd(&counts[k], 1);
}

// Each cell's particles by distance from its centre, ties by slot.
__global__ void mergeOrder(const int32_t* __restrict__ start, int32_t* __restrict__ members,
                           const double* __restrict__ distance, int64_t cells) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int32_t first = start[c], last = start[c + 1];
    for (int32_t k = first + 1; k < last; ++k) {
        const int32_t v = members[k];
        const double dv = distance[v];
        int32_t m = k - 1;
        while (m >= first && (distance[members[m]] > dv || (distance[members[m]] == dv && members[m] > v))) {
            members[m + 1] = members[m];
            --m;
        }
        members[m + 1] = v;
    }
}

// A group of the `size` nearest in every cell holding that many: its members by slot, no longer free.
__global__ void mergeTake(const int32_t* __restrict__ start, const int32_t* __restrict__ members, int32_t size,
                          const int32_t* __restrict__ runOf, int32_t* __restrict__ groups, int32_t* __restrict__ groupCount,
                          int32_t* __restrict__ levelOf, int32_t* __restrict__ taken, int64_t cells) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int32_t first = start[c];
    if (start[c + 1] - first < size) return;
    int32_t chosen[8];
    for (int k = 0; k < size; ++k) {
        int32_t v = members[first + k];
        int m = k - 1;
        while (m >= 0 && chosen[m] > v) {
            chosen[m + 1] = chosen[m];
            --m;
        }
        chosen[m + 1] = v;
    }
    const int32_t g = atomicAdd(groupCount, 1);
    for (int k = 0; k < size; ++k) {
        groups[(int64_t)g * size + k] = chosen[k];
        levelOf[chosen[k]] = 0;
    }
    atomicAdd(&taken[runOf[chosen[0]]], size);
}

// Each group into its first member's slot; the others marked to leave; the energy the merge removes.
template <typename Real>
__global__ void mergeGroups(const int32_t* __restrict__ groups, int32_t count, int32_t size, const int32_t* __restrict__ runOf,
                            const sph_merge_spec* __restrict__ specs, sph_state st, int32_t dims,
                            int32_t* __restrict__ removed, double* __restrict__ energy) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    const int32_t* member = groups + (int64_t)g * size;
    const int64_t stride = st.stride;
    Real* mass = static_cast<Real*>(st.mass);
    Real* length = static_cast<Real*>(st.length);
    Real* pressure = static_cast<Real*>(st.pressure);
    Real* acceleration = static_cast<Real*>(st.acceleration);
    double total = 0.0;
    for (int k = 0; k < size; ++k) total += (double)mass[member[k]];
    double position[3] = {0.0, 0.0, 0.0}, velocity[3] = {0.0, 0.0, 0.0}, accel[3] = {0.0, 0.0, 0.0};
    double p = 0.0, volume = 0.0, h = 0.0;
    for (int k = 0; k < size; ++k) {
        const int32_t j = member[k];
        const double m = (double)mass[j], w = m / total;
        for (int a = 0; a < dims; ++a) {
            position[a] += w * st.position[a * stride + j];
            velocity[a] += w * st.velocity[a * stride + j];
            accel[a] += w * (double)acceleration[a * stride + j];
        }
        p += w * (double)pressure[j];
        volume += m / st.density[j];
        h += (double)length[j];
    }
    double removedEnergy = 0.0;
    for (int k = 0; k < size; ++k) {
        const int32_t j = member[k];
        double speed2 = 0.0;
        for (int a = 0; a < dims; ++a) {
            const double du = st.velocity[a * stride + j] - velocity[a];
            speed2 += du * du;
        }
        removedEnergy += (double)mass[j] * speed2;
    }
    const int32_t first = member[0];
    for (int a = 0; a < dims; ++a) {
        st.position[a * stride + first] = position[a];
        st.velocity[a * stride + first] = velocity[a];
        acceleration[a * stride + first] = (Real)accel[a];
    }
    pressure[first] = (Real)p;
    st.density[first] = total / volume;
    mass[first] = (Real)total;
    length[first] = (Real)(h / size / specs[runOf[first]].lengthRatio);
    atomicAdd(&energy[runOf[first]], 0.5 * removedEnergy);
}

// Every member of a group but its first leaves.
__global__ void markRemoved(const int32_t* __restrict__ groups, int32_t count, int32_t size, int32_t* __restrict__ removed) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    for (int k = 1; k < size; ++k) removed[groups[(int64_t)g * size + k]] = 1;
}

// Which slots are holes at or before their run's last boundary slot, and which remain after it.
__global__ void compactFlags(const int32_t* __restrict__ runOf, const long long* __restrict__ lastBoundary,
                             const int32_t* __restrict__ removed, int32_t* __restrict__ hole, int32_t* __restrict__ tail,
                             int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long last = lastBoundary[runOf[i]];
    hole[i] = removed[i] && i <= last;
    tail[i] = !removed[i] && i > last;
}

// The last as many remaining particles after a run's last boundary slot as it has holes: movers,
// the k-th of them filling the run's k-th hole; and which slots keep a place.
__global__ void compactMovers(const int32_t* __restrict__ runOf, const long long* __restrict__ lastBoundary,
                              const int32_t* __restrict__ removed, const int32_t* __restrict__ tail,
                              const int32_t* __restrict__ tailRank, const long long* __restrict__ tailBase,
                              const long long* __restrict__ holeBase, int32_t* __restrict__ movers,
                              int32_t* __restrict__ keep, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = runOf[i];
    if (i <= lastBoundary[r]) {
        keep[i] = 1;
        return;
    }
    bool mover = false;
    if (tail[i]) {
        const long long holes = holeBase[r + 1] - holeBase[r], tails = tailBase[r + 1] - tailBase[r];
        const long long k = (tailRank[i] - tailBase[r]) - (tails - holes);
        if (k >= 0) {
            movers[holeBase[r] + k] = (int32_t)i;
            mover = true;
        }
    }
    keep[i] = !removed[i] && !mover;
}

// Each kept place takes its own particle, or a hole its mover, at its packed new slot.
template <typename Real>
__global__ void compactPermute(const int32_t* __restrict__ keep, const int32_t* __restrict__ newIndex,
                               const int32_t* __restrict__ hole, const int32_t* __restrict__ holeRank,
                               const int32_t* __restrict__ movers, sph_state from, sph_state to, int32_t dims, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const int64_t src = hole[i] ? movers[holeRank[i]] : i, d = newIndex[i];
    const int64_t fs = from.stride, ts = to.stride;
    for (int a = 0; a < dims; ++a) {
        to.position[a * ts + d] = from.position[a * fs + src];
        to.velocity[a * ts + d] = from.velocity[a * fs + src];
        static_cast<Real*>(to.acceleration)[a * ts + d] = static_cast<const Real*>(from.acceleration)[a * fs + src];
    }
    to.density[d] = from.density[src];
    static_cast<Real*>(to.pressure)[d] = static_cast<const Real*>(from.pressure)[src];
    static_cast<Real*>(to.mass)[d] = static_cast<const Real*>(from.mass)[src];
    static_cast<Real*>(to.length)[d] = static_cast<const Real*>(from.length)[src];
    to.fluid[d] = from.fluid[src];
}

// Run r's value of an exclusive scan at its first slot, and the total after the last.
__global__ void runBases(const RunDev* __restrict__ runs, const int32_t* __restrict__ ranks, int32_t count, int64_t n,
                         long long* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > count) return;
    out[r] = r < count ? ranks[runs[r].first] : ranks[n];
}

inline unsigned gridFor(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Runs BODY with the compile-time constant L set to the context's lanes per particle.
#define WITH_LANES(ctx, BODY)                            \
    switch (lanesOf(ctx)) {                              \
        case 1: { constexpr int L = 1; BODY; } break;    \
        case 2: { constexpr int L = 2; BODY; } break;    \
        case 4: { constexpr int L = 4; BODY; } break;    \
        case 8: { constexpr int L = 8; BODY; } break;    \
        default: { constexpr int L = 16; BODY; } break;  \
    }

}  // namespace

// -------------------------------------------------------------------------------------------------
// Context
// -------------------------------------------------------------------------------------------------

struct sph_context {
    int32_t dims = 2, precision = 0;
    hipStream_t stream = nullptr;
    bool ownsStream = false;
    std::string error;
    int block = 256;
    int lanes = 0;           // lanes per particle in the sweeps; 0 picks by dimension
    int32_t nearBytes = -1;  // the cap of a sweep workgroup's LDS copy of records; -1 picks by dimension
    int32_t nearBytesBuild = -1;   // the same for the step's list build; -1 takes nearBytes
    int listCapacity = -1;   // neighbor-list entries per particle; -1 picks by dimension, 0 walks the cells
    NeighborList list{};
    // A copy of the bound state the state sort permutes from, allocated on first use
    void* sortCopy = nullptr;
    size_t sortBytes = 0;
    void* readScratch = nullptr;   // the breaking read-out's buffers
    size_t readBytes = 0;
    // The merge's per-slot and per-cell buffers
    void* mergeScratch = nullptr;
    size_t mergeBytes = 0;
    // the delta-LES-SPH closure of the host steps: on, c0, rho0, K nu / c0, and the eddy alphas and
    // deltas of the last sum that built them, zero until one has
    int32_t closureOn = 0;
    double closureC0 = 1.0, closureRho0 = 1000.0, closureWater = 0.0;
    void* eddy = nullptr;
    size_t eddyBytes = 0;
    int64_t eddyStride = 0;
    // The energy budget's ledger (sphEnergy.py deviceTerms): 14 float64 rows of the state's stride,
    // each fluid particle's running total of each term; the acceleration sweep's four parts of a (4 D
    // Real rows); the per-chunk partials and per-run totals of sph_energy_sums
    int32_t ledgerOn = 0;
    double* ledger = nullptr;
    void* ledgerParts = nullptr;
    double* ledgerPartial = nullptr;
    double* ledgerOut = nullptr;
    int64_t ledgerStride = 0;
    // The Shepard filter's densities, one Real per slot; a filtered step asked from the host
    void* filtered = nullptr;
    int64_t filteredStride = 0;
    int32_t filterRequest = 0;
    // particle shifting's per-slot fields of the host step, 3 D + 1 values per sorted slot
    void* shiftFields = nullptr;
    size_t shiftBytes = 0;
    void* mergeCells = nullptr;
    size_t mergeCellBytes = 0;
    // The split's marks and ranks between sph_split_count and sph_split_apply
    int32_t* splitMarks = nullptr;
    int64_t splitSlots = 0;
    std::vector<int64_t> splitChosen;
    // The sweeps' tiles for the current workgroup size and lanes (tilePer particles each)
    int32_t* tileRun = nullptr;
    int64_t* tileFirst = nullptr;
    int64_t tileCount = 0;
    int tilePer = 0;

    std::vector<sph_run> runs;
    std::vector<RunDev> runsHost;
    int64_t capacity = 0, total = 0, cells = 0;

    RunDev* dRuns = nullptr;
    int32_t* runOf = nullptr;
    int32_t* key = nullptr;
    int32_t* order = nullptr;
    int32_t* scattered = nullptr;   // the slots as the counting sort places them, before ordering
    int32_t* cellCount = nullptr;
    int32_t* cellFill = nullptr;
    int32_t* cellStart = nullptr;
    uint8_t* eighth = nullptr;                     // per slot: its eighth of its cell along x
    unsigned long long* cellEighths = nullptr;     // per cell: its eighths table (EighthTable)
    int32_t* scanPartial = nullptr;
    int64_t scanTiles = 0;

    std::vector<void*> allocations;
    void* rec = nullptr;   // Records<float> or Records<double>, host copy of device pointers
    Records<float> rec32{};
    Records<double> rec64{};
    double** dVel = nullptr;   // device array of the bound velocity row pointers

    // The global sums' chunks: each run's slots in chunks of kSumChunk, in run order
    int64_t sumChunkCount = 0;
    int32_t* sumChunkRun = nullptr;
    int64_t* sumChunkFirst = nullptr;
    int64_t* sumRunChunk = nullptr;   // runs + 1 offsets into the chunks
    double* sumPartial = nullptr;
    double* sumOut = nullptr;

    Clock clock{};
    sph_state state{};
    bool bound = false;

    hipGraphExec_t graphExec = nullptr;
    hipGraph_t graph = nullptr;
    int32_t graphSteps = 0;
    int graphKey = 0;   // the step options the graph's launches carry

    // Per-kernel timing with HIP events on the context's stream, while profiling
    bool profiling = false;
    std::vector<hipEvent_t> events;            // pool; pairs (start, stop) in use up to eventsUsed
    size_t eventsUsed = 0;
    std::vector<int32_t> eventKernel;          // kernel id of each pair in use
    double kernelMs[16] = {};
    int64_t kernelLaunches[16] = {};
};

namespace {

int lanesOf(const sph_context* ctx) {
    if (ctx->lanes > 0) return ctx->lanes;
    return ctx->dims == 2 ? 8 : 16;
}

int listCapacityOf(const sph_context* ctx) {
    if (ctx->listCapacity >= 0) return ctx->listCapacity;
    return ctx->dims == 2 ? 64 : 160;
}

int fail(sph_context* ctx, int code, const std::string& message) {
    if (ctx) ctx->error = message;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t status_ = (expr);                                                                   \
        if (status_ != hipSuccess)                                                                     \
            return fail(ctx, SPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(status_));    \
    } while (0)

template <typename T>
int allocate(sph_context* ctx, T** out, size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc(&p, count * sizeof(T)));
    ctx->allocations.push_back(p);
    *out = static_cast<T*>(p);
    return SPH_OK;
}

void dropGraph(sph_context* ctx) {
    if (ctx->graphExec) (void)hipGraphExecDestroy(ctx->graphExec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graphExec = nullptr;
    ctx->graph = nullptr;
    ctx->graphSteps = 0;
}

void releaseLedger(sph_context* ctx) {
    if (ctx->ledger) (void)hipFree(ctx->ledger);
    if (ctx->ledgerParts) (void)hipFree(ctx->ledgerParts);
    if (ctx->ledgerPartial) (void)hipFree(ctx->ledgerPartial);
    if (ctx->ledgerOut) (void)hipFree(ctx->ledgerOut);
    ctx->ledger = nullptr;
    ctx->ledgerParts = nullptr;
    ctx->ledgerPartial = nullptr;
    ctx->ledgerOut = nullptr;
    ctx->ledgerStride = 0;
}

void releaseScratch(sph_context* ctx) {
    dropGraph(ctx);
    if (ctx->list.index) (void)hipFree(ctx->list.index);
    if (ctx->list.count) (void)hipFree(ctx->list.count);
    ctx->list = NeighborList{};
    if (ctx->tileRun) (void)hipFree(ctx->tileRun);
    if (ctx->tileFirst) (void)hipFree(ctx->tileFirst);
    ctx->tileRun = nullptr;
    ctx->tileFirst = nullptr;
    ctx->tileCount = 0;
    ctx->tilePer = 0;
    if (ctx->sortCopy) (void)hipFree(ctx->sortCopy);
    ctx->sortCopy = nullptr;
    ctx->sortBytes = 0;
    if (ctx->readScratch) (void)hipFree(ctx->readScratch);
    ctx->readScratch = nullptr;
    ctx->readBytes = 0;
    if (ctx->splitMarks) (void)hipFree(ctx->splitMarks);
    ctx->splitMarks = nullptr;
    ctx->splitSlots = 0;
    ctx->splitChosen.clear();
    if (ctx->mergeScratch) (void)hipFree(ctx->mergeScratch);
    ctx->mergeScratch = nullptr;
    ctx->mergeBytes = 0;
    if (ctx->eddy) (void)hipFree(ctx->eddy);
    ctx->eddy = nullptr;
    ctx->eddyBytes = 0;
    ctx->eddyStride = 0;
    if (ctx->shiftFields) (void)hipFree(ctx->shiftFields);
    ctx->shiftFields = nullptr;
    ctx->shiftBytes = 0;
    if (ctx->mergeCells) (void)hipFree(ctx->mergeCells);
    ctx->mergeCells = nullptr;
    ctx->mergeCellBytes = 0;
    releaseLedger(ctx);
    if (ctx->filtered) (void)hipFree(ctx->filtered);
    ctx->filtered = nullptr;
    ctx->filteredStride = 0;
    for (void* p : ctx->allocations) (void)hipFree(p);
    ctx->allocations.clear();
    ctx->dRuns = nullptr;
}

double sigmaUnitOf(int kernel, int dims) {
    const double pi = 3.14159265358979323846;
    if (kernel == 1) return dims == 2 ? 7.0 / (4.0 * pi) : 21.0 / (16.0 * pi);
    return dims == 2 ? 10.0 / (7.0 * pi) : 1.0 / pi;
}

template <typename Real>
int allocateRecords(sph_context* ctx, Records<Real>& rec, int64_t cap) {
    const int D = ctx->dims;
    int status = allocate(ctx, &rec.r, cap);
    if (status == SPH_OK) status = allocate(ctx, &rec.key, cap);
    if (status == SPH_OK) status = allocate(ctx, &rec.G, cap * D * D);
    if (status == SPH_OK) status = allocate(ctx, &rec.L, cap * D * D);
    if (status == SPH_OK) status = allocate(ctx, &rec.A, cap * D);
    return status;
}

int checkRuns(sph_context* ctx, const sph_run* runs, int32_t count, int64_t capacity) {
    int64_t next = 0;
    for (int32_t r = 0; r < count; ++r) {
        const sph_run& run = runs[r];
        if (run.first != next) return fail(ctx, SPH_ERR_VALUE, "Runs must cover consecutive slots from 0.");
        if (run.count < 0) return fail(ctx, SPH_ERR_VALUE, "A run's particle count cannot be negative.");
        next += run.count;
        if (run.kernel != 0 && run.kernel != 1) return fail(ctx, SPH_ERR_VALUE, "The kernel is not 0 (cubic spline) or 1 (Wendland C2).");
        if (run.grid[0] < 1 || run.grid[1] < 1 || run.grid[2] < 1) return fail(ctx, SPH_ERR_VALUE, "Grid dimensions must be positive.");
        if (ctx->dims == 2 && run.grid[2] != 1) return fail(ctx, SPH_ERR_VALUE, "A 2D grid has one cell along z.");
        if (!(run.cell > 0.0) || !(run.hmax > 0.0) || run.cell < 2.0 * run.hmax * (1.0 - 1e-12))
            return fail(ctx, SPH_ERR_VALUE, "The cell width must be positive and at least twice the longest smoothing length.");
        if (run.period > 0.0 && ctx->dims != 3) return fail(ctx, SPH_ERR_VALUE, "A periodic axis needs three dimensions: it is y.");
        if (run.period > 0.0 && run.period <= 4.0 * run.hmax)
            return fail(ctx, SPH_ERR_VALUE, "The periodic width must exceed twice the kernel support.");
        if (run.diffusion < 0 || run.diffusion > 3)
            return fail(ctx, SPH_ERR_VALUE,
                        "Density diffusion is not 0 (none), 1 (molteni), 2 (hydrostatic) or 3 (antuono).");
        if (run.filterInterval < 0) return fail(ctx, SPH_ERR_VALUE, "The filter interval cannot be negative.");
        const sph_boundary& b = run.boundary;
        if (b.kind < 0 || b.kind > 3) return fail(ctx, SPH_ERR_VALUE, "Unknown boundary kind.");
        if (b.kind == 2 && (b.nBed < 2 || !b.bedX || !b.bedZ)) return fail(ctx, SPH_ERR_VALUE, "A bed needs at least two vertices.");
        if (b.kind == 3 && (ctx->dims != 3 || b.fieldNx < 1 || b.fieldNy < 1 || !b.table))
            return fail(ctx, SPH_ERR_VALUE, "A height field needs three dimensions and a table.");
    }
    if (next > capacity) return fail(ctx, SPH_ERR_VALUE, "The runs cover more slots than the capacity.");
    if (capacity >= (int64_t)1 << 31) return fail(ctx, SPH_ERR_VALUE, "At most 2^31 - 1 slots per context.");
    return SPH_OK;
}

// The neighbor list for the context's capacity of slots; its entries are allocated apart from the
// rest of the scratch so a new list capacity does not reallocate the state's scratch.
int allocateList(sph_context* ctx) {
    if (ctx->list.index) (void)hipFree(ctx->list.index);
    if (ctx->list.count) (void)hipFree(ctx->list.count);
    ctx->list = NeighborList{};
    const int32_t capacity = listCapacityOf(ctx);
    const int64_t cap = std::max<int64_t>(ctx->capacity, 1);
    ctx->list.stride = cap;
    ctx->list.capacity = capacity;
    HIP_TRY(hipMalloc(&ctx->list.count, cap * sizeof(int32_t)));
    if (capacity > 0) HIP_TRY(hipMalloc(&ctx->list.index, (size_t)capacity * cap * sizeof(int32_t)));
    return SPH_OK;
}

// The sweeps' tiles: every run's sorted particles cut into workgroups of block / lanes particles.
// Rebuilt when the workgroup size or the lanes change; never during graph capture, because every
// entry point that launches sweeps calls it first.
int ensureTiles(sph_context* ctx) {
    const int per = std::max(1, ctx->block / lanesOf(ctx));
    if (per == ctx->tilePer && ctx->tileRun) return SPH_OK;
    std::vector<int32_t> run;
    std::vector<int64_t> first;
    for (size_t r = 0; r < ctx->runs.size(); ++r)
        for (int64_t f = ctx->runs[r].first; f < ctx->runs[r].first + ctx->runs[r].count; f += per) {
            run.push_back((int32_t)r);
            first.push_back(f);
        }
    if (run.empty()) {
        run.push_back(0);
        first.push_back(0);
    }
    if (ctx->tileRun) (void)hipFree(ctx->tileRun);
    if (ctx->tileFirst) (void)hipFree(ctx->tileFirst);
    ctx->tileRun = nullptr;
    ctx->tileFirst = nullptr;
    HIP_TRY(hipMalloc(&ctx->tileRun, run.size() * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&ctx->tileFirst, first.size() * sizeof(int64_t)));
    HIP_TRY(hipMemcpyAsync(ctx->tileRun, run.data(), run.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->tileFirst, first.data(), first.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->tileCount = (int64_t)run.size();
    ctx->tilePer = per;
    return SPH_OK;
}

inline Tiles tilesOf(const sph_context* ctx) { return Tiles{ctx->tileRun, ctx->tileFirst, ctx->tileCount}; }
inline unsigned tileGrid(const sph_context* ctx) { return (unsigned)ctx->tileCount; }

// Records a sweep workgroup copies into LDS: the cells around its particles hold about 3 records
// per particle in 2D and 9 in 3D plus the ends of the ranges; the copy is capped (32 KiB unless
// sph_set_near_cells says otherwise) so that LDS leaves the waves' occupancy alone.
template <int D, typename Real>
int32_t nearRecords(const sph_context* ctx, bool build = false) {
    if (!SPH_NEAR_CELLS) return 0;
    const int32_t per = (int32_t)ctx->tilePer;
    const int32_t want = D == 2 ? 4 * per + 64 : SPH_NEAR_PER_3D * per + 128;
    const int32_t chosen = build && ctx->nearBytesBuild >= 0 ? ctx->nearBytesBuild : ctx->nearBytes;
    const int32_t bytes = chosen >= 0 ? chosen : (D == 2 ? 32768 : SPH_NEAR_BYTES_3D);
    const int32_t most = (int32_t)(bytes / sizeof(Record<Real>));
    return want < most ? want : most;
}
// Hits per particle the fused list build keeps in LDS after its NearCells copy: a 6 KiB (2D) or
// 16 KiB (3D) share of the workgroup's LDS, and no more than the list holds. A 2D particle has about
// 20 neighbors at h = 1.3 dx, so at 4 lanes (64 particles a workgroup) nearly all of them fit.
template <int D>
int32_t hitCapacity(const sph_context* ctx);
// A sweep's dynamic LDS: the NearCells copy, then `sums` values per particle of the workgroup.
template <int D, typename Real>
size_t sweepLds(const sph_context* ctx, int sums) {
    return (size_t)nearRecords<D, Real>(ctx) * sizeof(Record<Real>) + (size_t)ctx->tilePer * sums * sizeof(Real);
}
// The acceleration sweep's: the NearCells copy with, in 2D float32, each record's G (a fifth quad).
template <int D, typename Real>
size_t accelerationLds(const sph_context* ctx) {
    const bool withG = D == 2 && sizeof(Real) == 4;
    return (size_t)nearRecords<D, Real>(ctx) * (sizeof(Record<Real>) + (withG ? sizeof(Quad<Real>) : 0));
}

template <int D>
int32_t hitCapacity(const sph_context* ctx) {
    const int32_t per = (int32_t)ctx->tilePer;
    if (per <= 0) return 0;
    const int32_t keep = (D == 2 ? 6144 : 16384) / (int32_t)sizeof(int32_t) / per;
    return std::max(0, std::min(keep, ctx->list.capacity));
}
// The fused list build's dynamic LDS: the NearCells copy (at the list build's own cap), then the
// hits it keeps.
template <int D, typename Real>
size_t buildLds(const sph_context* ctx) {
    return (size_t)nearRecords<D, Real>(ctx, true) * sizeof(Record<Real>) +
           (size_t)ctx->tilePer * (hitCapacity<D>(ctx) | 1) * sizeof(int32_t);
}

template <typename Real>
Records<Real>& recordsOf(sph_context* ctx);
template <>
Records<float>& recordsOf<float>(sph_context* ctx) { return ctx->rec32; }
template <>
Records<double>& recordsOf<double>(sph_context* ctx) { return ctx->rec64; }

int launchCheck(sph_context* ctx, const char* what) {
    const hipError_t status = hipGetLastError();
    if (status != hipSuccess) return fail(ctx, SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(status));
    return SPH_OK;
}

// Kernel ids of the per-kernel timing: include/sph_hip.h lists them for sph_kernel_times.
enum KernelId { kClock = 0, kCellKeys, kScan, kCellScatter, kCellOrder, kGather, kSweepWall, kSweepAcceleration,
                kKick, kSweepContinuity, kNeighbors, kEnforce, kMaxima, kKernelKinds };

// Records a start event before a timed launch; returns its slot, or -1 while not profiling.
int timeBegin(sph_context* ctx, int kernel) {
    if (!ctx->profiling) return -1;
    if (ctx->eventsUsed + 2 > ctx->events.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            ctx->events.push_back(e);
        }
    }
    const size_t slot = ctx->eventsUsed;
    ctx->eventsUsed += 2;
    ctx->eventKernel.push_back(kernel);
    (void)hipEventRecord(ctx->events[slot], ctx->stream);
    return (int)slot;
}

void timeEnd(sph_context* ctx, int slot) {
    if (slot >= 0) (void)hipEventRecord(ctx->events[slot + 1], ctx->stream);
}

// Folds the recorded pairs into the per-kernel totals; synchronizes.
int collectTimes(sph_context* ctx) {
    if (ctx->eventsUsed == 0) return SPH_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < ctx->eventKernel.size(); ++k) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]));
        ctx->kernelMs[ctx->eventKernel[k]] += ms;
        ctx->kernelLaunches[ctx->eventKernel[k]] += 1;
    }
    ctx->eventsUsed = 0;
    ctx->eventKernel.clear();
    return SPH_OK;
}

#define TIMED(ctx, id, launch)                    \
    do {                                          \
        const int slot_ = timeBegin((ctx), (id)); \
        launch;                                   \
        timeEnd((ctx), slot_);                    \
    } while (0)

// The cell list over the bound positions. `rank` launches the ordering within cells on the
// scattered slots (cellRank, or the step's cellRankGather), over max(n, cells) threads.
template <int D, class Rank>
int buildCellsWith(sph_context* ctx, int clockMode, Rank&& rank) {
    const int64_t n = ctx->total;
    const int b = 256;
    double* P = ctx->state.position;
    const int64_t st = ctx->state.stride;
    if (n > 0)
        TIMED(ctx, kCellKeys, (cellKeys<D><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, P, P + st,
                                                                                 D == 3 ? P + 2 * st : P, n, ctx->key,
                                                                                 ctx->cellCount,
                                                                                 D == 3 ? ctx->eighth : nullptr,
                                                                                 ctx->clock, clockMode)));
    const int64_t m = ctx->cells;
    {
        const int slot = timeBegin(ctx, kScan);
        scanTiles<<<(unsigned)ctx->scanTiles, kScanThreads, 0, ctx->stream>>>(ctx->cellCount, ctx->cellStart,
                                                                               ctx->scanPartial, m);
        if (ctx->scanTiles > 0 && ctx->scanTiles <= kScanThreads) {
            scanAddFew<<<(unsigned)ctx->scanTiles, kScanThreads, 0, ctx->stream>>>(ctx->cellStart, ctx->scanPartial,
                                                                                   ctx->scanTiles, m);
        } else {
            scanPartials<<<1, kScanThreads, 0, ctx->stream>>>(ctx->scanPartial, ctx->scanTiles, ctx->cellStart, m);
            scanAdd<<<(unsigned)ctx->scanTiles, kScanThreads, 0, ctx->stream>>>(ctx->cellStart, ctx->scanPartial, m);
        }
        timeEnd(ctx, slot);
    }
    if (n > 0)
        TIMED(ctx, kCellScatter, (cellScatter<<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->key, ctx->cellStart,
                                                                                    ctx->cellFill, ctx->scattered, n)));
    TIMED(ctx, kCellOrder, (rank(gridFor(std::max<int64_t>(n, m), b), b, D == 3 ? P : nullptr,
                                 D == 3 ? ctx->cellEighths : nullptr, n, m)));
    return launchCheck(ctx, "cell build");
}

template <int D>
int buildCells(sph_context* ctx, int clockMode = 0) {
    return buildCellsWith<D>(ctx, clockMode, [&](unsigned grid, int block, const double* x,
                                                 unsigned long long* eighths, int64_t n, int64_t m) {
        cellRank<<<grid, block, 0, ctx->stream>>>(ctx->cellStart, ctx->key, x, ctx->eighth, ctx->scattered, ctx->order,
                                                 eighths, ctx->cellCount, ctx->cellFill, n, m);
    });
}

template <typename Real>
Sources<Real> sourcesOf(sph_context* ctx, const double* velocity, const double* density, const Real* pressure,
                        Real* pressureOut) {
    Sources<Real> src{};
    const int64_t st = ctx->state.stride;
    for (int a = 0; a < 3; ++a) {
        src.pos[a] = a < ctx->dims ? ctx->state.position + a * st : nullptr;
        src.vel[a] = a < ctx->dims ? velocity + a * st : nullptr;
    }
    src.density = density;
    src.pressure = pressure;
    src.mass = static_cast<const Real*>(ctx->state.mass);
    src.len = static_cast<const Real*>(ctx->state.length);
    src.fluid = ctx->state.fluid;
    src.key = ctx->key;
    src.pressureOut = pressureOut;
    return src;
}

template <typename Real>
SweepOut<Real> emptyOut(sph_context* ctx) {
    SweepOut<Real> out{};
    out.stride = ctx->state.stride;
    return out;
}

int requireBound(sph_context* ctx) {
    if (!ctx) return SPH_ERR_VALUE;
    if (!ctx->bound) return fail(ctx, SPH_ERR_VALUE, "No state is bound; call sph_configure and sph_bind first.");
    return ensureTiles(ctx);
}

// The options a context's device step carries besides the delta-SPH scheme, which pick its
// launches: the Shepard filter (some run has an interval, or the host asked for a filtered step),
// the renormalized 'antuono' diffusion (some run has it), the delta-LES-SPH closure and the energy
// budget's ledger. Without any of them the step runs the lean sweeps.
struct StepOptions {
    bool filter = false, antuono = false, closure = false, ledger = false;
    bool any() const { return antuono || closure || ledger; }
};
StepOptions stepOptions(const sph_context* ctx, int filterMode) {
    StepOptions o;
    for (const RunDev& run : ctx->runsHost) {
        if (run.diffusion == 3) o.antuono = true;
        if (run.filterInterval > 0 && filterMode == 2) o.filter = true;
    }
    if (filterMode == 1) o.filter = true;
    o.closure = ctx->closureOn != 0;
    o.ledger = ctx->ledgerOn != 0 && ctx->ledger != nullptr;
    return o;
}

// The option buffers a step reads: the closure's eddy coefficients, the filter's densities and the
// ledger's parts. Allocated here, before any launch or graph capture.
template <typename Real>
int ensureEddy(sph_context* ctx);
int ensureStepBuffers(sph_context* ctx, int filterMode) {
    const StepOptions o = stepOptions(ctx, filterMode);
    const size_t real = ctx->precision == 0 ? sizeof(float) : sizeof(double);
    if (o.closure) {
        const int status = ctx->precision == 0 ? ensureEddy<float>(ctx) : ensureEddy<double>(ctx);
        if (status != SPH_OK) return status;
    }
    if (o.filter && (!ctx->filtered || ctx->filteredStride != ctx->capacity)) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (ctx->filtered) (void)hipFree(ctx->filtered);
        ctx->filtered = nullptr;
        HIP_TRY(hipMalloc(&ctx->filtered, (size_t)ctx->capacity * real));
        ctx->filteredStride = ctx->capacity;
    }
    return SPH_OK;
}

// The two halves of a device-resident step of every active run. `phase` 1 is the rates
// (residentRates), 2 the integration (residentIntegrate) with `forcedDt` when positive and the
// run's own time step otherwise, 3 both. `useClock` 0 steps every run whatever its clock.
// `filterMode` picks the runs the Shepard filter takes (filterDue).
template <int D, typename Real>
int stepOnce(sph_context* ctx, int phase, int useClock, double forcedDt, int filterMode) {
    const int64_t n = ctx->total;
    const int b = ctx->block;
    const int32_t nruns = (int32_t)ctx->runs.size();
    Records<Real>& rec = recordsOf<Real>(ctx);
    const int64_t st = ctx->state.stride;
    const StepOptions o = stepOptions(ctx, filterMode);
    Sources<Real> src = sourcesOf<Real>(ctx, ctx->state.velocity, ctx->state.density, nullptr,
                                        static_cast<Real*>(ctx->state.pressure));
    if (phase & 1) {
        // A run without particles has no slot to open its step; the kernels skip it either way. The
        // ordering within cells also gathers the records (cellRankGather).
        int status = buildCellsWith<D>(ctx, useClock ? 1 : 2, [&](unsigned grid, int block, const double* x,
                                                                   unsigned long long* eighths, int64_t count,
                                                                   int64_t cells) {
            cellRankGather<D, Real><<<grid, block, 0, ctx->stream>>>(
                ctx->cellStart, ctx->key, x, ctx->eighth, ctx->scattered, ctx->order, eighths, ctx->cellCount,
                ctx->cellFill, count, cells, ctx->dRuns, ctx->runOf, src, rec, ctx->clock);
        });
        if (status != SPH_OK) return status;
    }
    (void)nruns;
    if (n == 0) return launchCheck(ctx, "device step");
    SweepOut<Real> out = emptyOut<Real>(ctx);
    out.pressure = static_cast<Real*>(ctx->state.pressure);
    out.density = ctx->state.density;
    out.acceleration = static_cast<Real*>(ctx->state.acceleration);
    for (int a = 0; a < 3; ++a) {
        out.pos[a] = a < D ? ctx->state.position + a * st : nullptr;
        out.vel[a] = a < D ? ctx->state.velocity + a * st : nullptr;
    }
    out.density64 = ctx->state.density;
    if (o.closure) {
        out.eddy = static_cast<Real*>(ctx->eddy);
        out.eddyC0 = (Real)ctx->closureC0;
        out.eddyRho0 = (Real)ctx->closureRho0;
        out.eddyWater = (Real)ctx->closureWater;
    }
    if (o.ledger) {
        out.ledger = ctx->ledger;
        out.ledgerStride = ctx->ledgerStride;
        out.parts = static_cast<Real*>(ctx->ledgerParts);
        out.rateParts = out.parts + (int64_t)4 * D * st;
    }
    // the sweeps' options instantiation carries the closure and the renormalized diffusion; the
    // ledger adds launches of its own and leaves the step's kernels as they are
    const bool physicsOptions = o.antuono || o.closure;
    const int mode = o.filter ? filterMode : 0;
    const double* y64 = ctx->state.position + st;
    if (phase & 1) {
        if (ctx->list.capacity > 0) {
            WITH_LANES(ctx, TIMED(ctx, kNeighbors,
                                  (buildNeighbors<D, Real, L, true><<<tileGrid(ctx), b, buildLds<D, Real>(ctx), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, ctx->cellEighths, rec, ctx->list, out,
                                      ctx->clock, y64, 1, nearRecords<D, Real>(ctx, true), hitCapacity<D>(ctx), n, mode))));
        } else {
            WITH_LANES(ctx, TIMED(ctx, kSweepWall,
                                  (sweepWallAndGradients<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out,
                                      ctx->clock, y64, 1, 1, 1, n))));
        }
        if (o.filter) {
            // the runs the filter takes this step: their fluid takes the Shepard density of the
            // gathered records, then their records and sweep 1 are taken again (sphGpu.py:2251-2254)
            Real* filtered = static_cast<Real*>(ctx->filtered);
            WITH_LANES(ctx, (sweepFilters<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, y64, 0, Real(0),
                            filtered, st, ctx->clock, mode)));
            applyShepard<Real><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->state, filtered,
                                                                     o.ledger ? ctx->ledger : nullptr,
                                                                     ctx->ledgerStride, ctx->clock, mode, n);
            gatherRecords<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->order, src, rec,
                                                                         ctx->clock, 1, n, mode);
            WITH_LANES(ctx, (sweepWallAndGradients<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                            y64, 1, 1, 1, n, mode)));
        }
        if (o.closure)
            closureCoefficients<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(
                ctx->order, rec, static_cast<Real*>(ctx->eddy), st, (Real)ctx->closureC0, n);
        if (o.ledger)
            WITH_LANES(ctx, (sweepAccelerations<D, Real, L, true, true><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx),
                                                                          ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                            y64, 1, 1, 1, 0, nearRecords<D, Real>(ctx), n)));
        if (physicsOptions) {
            WITH_LANES(ctx, TIMED(ctx, kSweepAcceleration,
                                  (sweepAccelerations<D, Real, L, true><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx),
                                                                          ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                                      y64, 1, 1, 1, 0, nearRecords<D, Real>(ctx), n))));
        } else {
            WITH_LANES(ctx, TIMED(ctx, kSweepAcceleration,
                                  (sweepAccelerations<D, Real, L><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                                      y64, 1, 1, 1, 0, nearRecords<D, Real>(ctx), n))));
        }
        if (ctx->sumChunkCount > 0) {
            const int slot = timeBegin(ctx, kMaxima);
            chunkMaxima<<<(unsigned)ctx->sumChunkCount, kMaximaThreads, 0, ctx->stream>>>(
                ctx->sumChunkRun, ctx->sumChunkFirst, ctx->dRuns, ctx->clock.speed, ctx->clock.acceleration, ctx->clock,
                ctx->sumPartial);
            runMaxima<<<gridFor(nruns, 64), 64, 0, ctx->stream>>>(ctx->dRuns, ctx->sumRunChunk, ctx->sumPartial,
                                                                  ctx->clock, nruns, useClock);
            timeEnd(ctx, slot);
        }
    }
    if (phase & 2) {
        if (o.antuono)
            // the renormalized density gradients from the records' deviations and sweep 2's L, kept
            // in the records' G rows for the continuity sums (sphGpu.py:2288-2291)
            WITH_LANES(ctx, (sweepFilters<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, y64, 2, Real(0),
                            nullptr, st)));
        if (o.ledger)
            WITH_LANES(ctx, (sweepContinuity<D, Real, L, true, true><<<tileGrid(ctx), b, sweepLds<D, Real>(ctx, 1),
                                                                       ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                            1, 0, 0, forcedDt, nearRecords<D, Real>(ctx), n)));
        if (physicsOptions) {
            WITH_LANES(ctx, TIMED(ctx, kSweepContinuity,
                                  (sweepContinuity<D, Real, L, true><<<tileGrid(ctx), b, sweepLds<D, Real>(ctx, 1),
                                                                       ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                                  1, 0, 0, forcedDt, nearRecords<D, Real>(ctx), n))));
        } else {
            WITH_LANES(ctx, TIMED(ctx, kSweepContinuity,
                                  (sweepContinuity<D, Real, L><<<tileGrid(ctx), b,
                                                                  sweepLds<D, Real>(ctx, 1), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, ctx->list, out, ctx->clock,
                                  1, 0, 0, forcedDt, nearRecords<D, Real>(ctx), n))));
        }
    }
    return launchCheck(ctx, "device step");
}

template <int D, typename Real>
int advanceSteps(sph_context* ctx, int32_t steps) {
    for (int32_t k = 0; k < steps; ++k) {
        const int status = stepOnce<D, Real>(ctx, 3, 1, 0.0, 2);
        if (status != SPH_OK) return status;
    }
    return SPH_OK;
}

int dispatchAdvance(sph_context* ctx, int32_t steps) {
    if (ctx->dims == 2) return ctx->precision == 0 ? advanceSteps<2, float>(ctx, steps) : advanceSteps<2, double>(ctx, steps);
    return ctx->precision == 0 ? advanceSteps<3, float>(ctx, steps) : advanceSteps<3, double>(ctx, steps);
}

int dispatchPhase(sph_context* ctx, int phase, double forcedDt, int filterMode) {
    const int status = ensureStepBuffers(ctx, filterMode);
    if (status != SPH_OK) return status;
    if (ctx->dims == 2)
        return ctx->precision == 0 ? stepOnce<2, float>(ctx, phase, 0, forcedDt, filterMode)
                                   : stepOnce<2, double>(ctx, phase, 0, forcedDt, filterMode);
    return ctx->precision == 0 ? stepOnce<3, float>(ctx, phase, 0, forcedDt, filterMode)
                               : stepOnce<3, double>(ctx, phase, 0, forcedDt, filterMode);
}

// The closure's eddy alphas and deltas, two rows of the state's stride, zeroed when first made or
// when the stride changes.
template <typename Real>
int ensureEddy(sph_context* ctx) {
    const int64_t stride = ctx->state.stride;
    const size_t bytes = 2 * (size_t)stride * sizeof(Real);
    if (ctx->eddy && ctx->eddyStride == stride && bytes <= ctx->eddyBytes) return SPH_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->eddy) (void)hipFree(ctx->eddy);
    ctx->eddy = nullptr;
    ctx->eddyBytes = 0;
    HIP_TRY(hipMalloc(&ctx->eddy, bytes));
    ctx->eddyBytes = bytes;
    ctx->eddyStride = stride;
    HIP_TRY(hipMemsetAsync(ctx->eddy, 0, bytes, ctx->stream));
    return SPH_OK;
}

// Host-step sums: gather records from caller arrays, then one sweep.
template <int D, typename Real>
int hostSweep(sph_context* ctx, int which, const double* velocity, const double* density, const void* pressure,
              void* out0, double* out1, double scalar) {
    const int64_t n = ctx->total;
    if (n == 0) return SPH_OK;
    // the host steps pick their launches by the first run's options, which every run must then share
    const RunDev& run0 = ctx->runsHost[0];
    for (const RunDev& run : ctx->runsHost)
        if (run.limiter != run0.limiter || run.correction != run0.correction || run.diffusion != run0.diffusion)
            return fail(ctx, SPH_ERR_VALUE,
                        "The host-step sums need every run of the context to share its viscosity limiter, continuity "
                        "correction and density diffusion.");
    const int b = ctx->block;
    const int32_t nruns = (int32_t)ctx->runs.size();
    Records<Real>& rec = recordsOf<Real>(ctx);
    clockBegin<<<gridFor(nruns, 64), 64, 0, ctx->stream>>>(ctx->clock, nruns, 0);
    const double* vel = velocity ? velocity : ctx->state.velocity;
    Sources<Real> src = sourcesOf<Real>(ctx, vel, density, static_cast<const Real*>(pressure), nullptr);
    gatherRecords<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->order, src, rec, ctx->clock, 0, n);
    SweepOut<Real> out = emptyOut<Real>(ctx);
    const bool closure = ctx->closureOn != 0;
    if (closure) {
        const int status = ensureEddy<Real>(ctx);
        if (status != SPH_OK) return status;
        out.eddy = static_cast<Real*>(ctx->eddy);
        out.eddyC0 = (Real)ctx->closureC0;
        out.eddyRho0 = (Real)ctx->closureRho0;
        out.eddyWater = (Real)ctx->closureWater;
    }
    const double* y64 = ctx->state.position + ctx->state.stride;
    const int32_t* order = ctx->order;
    const int32_t* start = ctx->cellStart;
    const NeighborList list = ctx->list;
    if (list.capacity > 0)
        WITH_LANES(ctx, (buildNeighbors<D, Real, L, false><<<tileGrid(ctx), b, sweepLds<D, Real>(ctx, 0), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, ctx->cellEighths, rec, list, emptyOut<Real>(ctx), ctx->clock, y64, 0,
                            nearRecords<D, Real>(ctx), 0, n)));
    switch (which) {
        case 0:   // wall pressure
            out.pressure = static_cast<Real*>(out0);
            out.density = out1;
            WITH_LANES(ctx, (sweepWallAndGradients<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 1, 0, n)));
            break;
        case 1:   // velocity gradients
            out.gradients = static_cast<Real*>(out0);
            WITH_LANES(ctx, (sweepWallAndGradients<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 0, 1, n)));
            break;
        case 2:   // correction matrices
            out.corrections = static_cast<Real*>(out0);
            WITH_LANES(ctx, (sweepAccelerations<D, Real, L><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 0, 1, 0, nearRecords<D, Real>(ctx), n)));
            break;
        case 3:    // accelerations, with G first when the limiter or the closure needs it
        case 8: {  // the closure's coefficients alone
            const int limiter = run0.limiter;
            if (which == 8 && !closure) return fail(ctx, SPH_ERR_VALUE, "The closure is off; see sph_set_closure.");
            if (limiter || closure)
                WITH_LANES(ctx, (sweepWallAndGradients<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 0, 1, n)));
            if (closure)
                closureCoefficients<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(
                    order, rec, static_cast<Real*>(ctx->eddy), ctx->state.stride, (Real)ctx->closureC0, n);
            if (which == 8) break;
            out.acceleration = static_cast<Real*>(out0);
            if (closure) {
                WITH_LANES(ctx, (sweepAccelerations<D, Real, L, true><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx),
                                                                        ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 1, 0, limiter, nearRecords<D, Real>(ctx), n)));
            } else {
                WITH_LANES(ctx, (sweepAccelerations<D, Real, L><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 1, 0, limiter, nearRecords<D, Real>(ctx), n)));
            }
            break;
        }
        case 4: {  // continuity, with the correction matrices first when the run has them or the
                   // renormalized diffusion, then that diffusion's density gradients; with the
                   // closure, the deltas of the last sum that built them
            const int correction = run0.correction;
            const bool renormalized = run0.diffusion == 3;
            if (correction || renormalized)
                WITH_LANES(ctx, (sweepAccelerations<D, Real, L><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 0, 1, 0, nearRecords<D, Real>(ctx), n)));
            if (renormalized)
                WITH_LANES(ctx, (sweepFilters<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, y64, 2, Real(0), nullptr,
                                ctx->state.stride)));
            out.rate = static_cast<Real*>(out0);
            for (int a = 0; a < 3; ++a) out.pos[a] = a < D ? ctx->state.position + a * ctx->state.stride : nullptr;
            if (renormalized || closure) {
                WITH_LANES(ctx, (sweepContinuity<D, Real, L, true><<<tileGrid(ctx), b,
                                                                    sweepLds<D, Real>(ctx, 1), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, 0, correction,
                                run0.diffusion, 0.0, nearRecords<D, Real>(ctx), n)));
            } else {
                WITH_LANES(ctx, (sweepContinuity<D, Real, L><<<tileGrid(ctx), b,
                                                              sweepLds<D, Real>(ctx, 1), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, 0, correction,
                                run0.diffusion, 0.0, nearRecords<D, Real>(ctx), n)));
            }
            break;
        }
        case 7:   // the renormalized density gradients, after the correction matrices
            WITH_LANES(ctx, (sweepAccelerations<D, Real, L><<<tileGrid(ctx), b, accelerationLds<D, Real>(ctx), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, out, ctx->clock, y64, 0, 0, 1, 0, nearRecords<D, Real>(ctx), n)));
            WITH_LANES(ctx, (sweepFilters<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, y64, 2, Real(0),
                                static_cast<Real*>(out0), ctx->state.stride)));
            break;
        case 5:   // the Shepard-filtered density
        case 6:   // the XSPH transport velocity, epsilon in `scalar`
            WITH_LANES(ctx, (sweepFilters<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), order, start, rec, list, y64, which - 5, (Real)scalar,
                                static_cast<Real*>(out0), ctx->state.stride)));
            break;
        default:
            return fail(ctx, SPH_ERR_VALUE, "Unknown sum.");
    }
    return launchCheck(ctx, "host-step sum");
}

// Particle shifting as a host step: the records from the caller's velocities and densities, the
// list, then sweepShifting's two passes.
template <int D, typename Real>
int shiftSweep(sph_context* ctx, const double* velocity, const double* density, double size, int32_t local,
               double spacingKernel, void* shifts, void* densityChanges, void* velocityChanges) {
    const int64_t n = ctx->total;
    if (n == 0) return SPH_OK;
    for (const RunDev& run : ctx->runsHost)
        if (!run.uniform) return fail(ctx, SPH_ERR_VALUE, "Particle shifting needs one smoothing length for every particle.");
    const size_t bytes = (size_t)ctx->state.stride * shiftWidth<D>() * sizeof(Real);
    if (bytes > ctx->shiftBytes) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (ctx->shiftFields) (void)hipFree(ctx->shiftFields);
        ctx->shiftFields = nullptr;
        ctx->shiftBytes = 0;
        HIP_TRY(hipMalloc(&ctx->shiftFields, bytes));
        ctx->shiftBytes = bytes;
    }
    const int b = ctx->block;
    const int32_t nruns = (int32_t)ctx->runs.size();
    Records<Real>& rec = recordsOf<Real>(ctx);
    clockBegin<<<gridFor(nruns, 64), 64, 0, ctx->stream>>>(ctx->clock, nruns, 0);
    Sources<Real> src = sourcesOf<Real>(ctx, velocity, density, nullptr, nullptr);
    gatherRecords<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->order, src, rec, ctx->clock, 0, n);
    const double* y64 = ctx->state.position + ctx->state.stride;
    const NeighborList list = ctx->list;
    if (list.capacity > 0)
        WITH_LANES(ctx, (buildNeighbors<D, Real, L, false><<<tileGrid(ctx), b, sweepLds<D, Real>(ctx, 0), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, ctx->cellEighths, rec, ctx->list, emptyOut<Real>(ctx),
                                ctx->clock, y64, 0, nearRecords<D, Real>(ctx), 0, n)));
    Real* fields = static_cast<Real*>(ctx->shiftFields);
    for (int32_t pass = 0; pass < 2; ++pass)
        WITH_LANES(ctx, (sweepShifting<D, Real, L><<<tileGrid(ctx), b, 0, ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, rec, list, y64, pass, (Real)size,
                                local, (Real)spacingKernel, fields, static_cast<Real*>(shifts),
                                static_cast<Real*>(densityChanges), static_cast<Real*>(velocityChanges),
                                ctx->state.stride)));
    return launchCheck(ctx, "particle shifting");
}

int dispatchHostSweep(sph_context* ctx, int which, const double* velocity, const double* density, const void* pressure,
                      void* out0, double* out1, double scalar = 0.0) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!density) return fail(ctx, SPH_ERR_VALUE, "The sums need the densities.");
    if (ctx->dims == 2)
        return ctx->precision == 0 ? hostSweep<2, float>(ctx, which, velocity, density, pressure, out0, out1, scalar)
                                   : hostSweep<2, double>(ctx, which, velocity, density, pressure, out0, out1, scalar);
    return ctx->precision == 0 ? hostSweep<3, float>(ctx, which, velocity, density, pressure, out0, out1, scalar)
                               : hostSweep<3, double>(ctx, which, velocity, density, pressure, out0, out1, scalar);
}

template <int D, typename Real>
int neighborsImpl(sph_context* ctx, int32_t* counts, const int64_t* offsets, int32_t* rows) {
    const int64_t n = ctx->total;
    if (n == 0) return SPH_OK;
    const int b = ctx->block;
    Records<Real>& rec = recordsOf<Real>(ctx);
    const int32_t nruns = (int32_t)ctx->runs.size();
    clockBegin<<<gridFor(nruns, 64), 64, 0, ctx->stream>>>(ctx->clock, nruns, 0);
    Sources<Real> src = sourcesOf<Real>(ctx, ctx->state.velocity, ctx->state.density, nullptr, nullptr);
    gatherRecords<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->order, src, rec, ctx->clock, 0, n);
    const double* y64 = ctx->state.position + ctx->state.stride;
    if (ctx->list.capacity > 0)
        WITH_LANES(ctx, (buildNeighbors<D, Real, L, false><<<tileGrid(ctx), b, sweepLds<D, Real>(ctx, 0), ctx->stream>>>(
                            ctx->dRuns, tilesOf(ctx), ctx->order, ctx->cellStart, ctx->cellEighths, rec, ctx->list, emptyOut<Real>(ctx),
                            ctx->clock, y64, 0, nearRecords<D, Real>(ctx), 0, n)));
    neighborExport<D, Real><<<gridFor(n, b), b, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->order, ctx->cellStart, rec,
                                                                   ctx->list, y64, counts, offsets, rows, n);
    int status = launchCheck(ctx, "neighbor list");
    if (status != SPH_OK) return status;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

}  // namespace

// -------------------------------------------------------------------------------------------------
// C ABI
// -------------------------------------------------------------------------------------------------

namespace {

// A device buffer of at least `bytes`, grown when short (its contents are not kept).
int growBuffer(sph_context* ctx, void*& buffer, size_t& size, size_t bytes) {
    if (bytes <= size) return SPH_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (buffer) (void)hipFree(buffer);
    buffer = nullptr;
    size = 0;
    HIP_TRY(hipMalloc(&buffer, bytes));
    size = bytes;
    return SPH_OK;
}

// Pieces of one buffer, each 256-byte aligned.
struct Carver {
    char* base;
    size_t used = 0;
    template <typename T>
    T* take(int64_t count) {
        T* p = reinterpret_cast<T*>(base + used);
        used += ((size_t)std::max<int64_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

// An exclusive scan of n int32 values into out (n + 1 entries), with `tiles` partial sums.
void scanInto(sph_context* ctx, const int32_t* in, int32_t* out, int32_t* tiles, int64_t n) {
    const int64_t count = (n + kScanTile - 1) / kScanTile + 1;
    scanTiles<<<(unsigned)count, kScanThreads, 0, ctx->stream>>>(in, out, tiles, n);
    scanPartials<<<1, kScanThreads, 0, ctx->stream>>>(tiles, count, out, n);
    scanAdd<<<(unsigned)count, kScanThreads, 0, ctx->stream>>>(out, tiles, n);
}

}  // namespace

extern "C" {

int sph_create(sph_context** out, int32_t dims, int32_t precision, void* stream) {
    if (!out) return SPH_ERR_VALUE;
    *out = nullptr;
    if (dims != 2 && dims != 3) return SPH_ERR_VALUE;
    if (precision != 0 && precision != 1) return SPH_ERR_VALUE;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return SPH_ERR_HIP;
    sph_context* ctx = new sph_context();
    ctx->dims = dims;
    ctx->precision = precision;
    if (stream) {
        ctx->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return SPH_ERR_HIP;
        }
        ctx->ownsStream = true;
    }
    *out = ctx;
    return SPH_OK;
}

int sph_destroy(sph_context* ctx) {
    if (!ctx) return SPH_OK;
    (void)hipStreamSynchronize(ctx->stream);
    releaseScratch(ctx);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->ownsStream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SPH_OK;
}

const char* sph_last_error(const sph_context* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

int sph_configure(sph_context* ctx, const sph_run* runs, int32_t count, int64_t capacity) {
    if (!ctx) return SPH_ERR_VALUE;
    if (count < 1 || !runs) return fail(ctx, SPH_ERR_VALUE, "At least one run is needed.");
    int status = checkRuns(ctx, runs, count, capacity);
    if (status != SPH_OK) return status;
    // The same slots and grids: only the physics changed, so the scratch, the bound state and a
    // captured graph (which reads the descriptors through their device copy) all stay
    bool sameLayout = ctx->dRuns && (int32_t)ctx->runs.size() == count && ctx->capacity == std::max<int64_t>(capacity, 1);
    for (int32_t r = 0; sameLayout && r < count; ++r) {
        const sph_run& a = ctx->runs[r];
        const sph_run& b = runs[r];
        sameLayout = a.first == b.first && a.count == b.count && a.grid[0] == b.grid[0] && a.grid[1] == b.grid[1] &&
                     a.grid[2] == b.grid[2];
    }
    if (!sameLayout) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        releaseScratch(ctx);
        ctx->bound = false;
    }
    ctx->runs.assign(runs, runs + count);
    ctx->runsHost.clear();
    int64_t cells = 0, total = 0;
    for (const sph_run& run : ctx->runs) {
        RunDev d{};
        d.first = run.first;
        d.count = run.count;
        d.cellFirst = cells;
        d.nx = run.grid[0];
        d.ny = run.grid[1];
        d.nz = ctx->dims == 3 ? run.grid[2] : 1;
        d.kernel = run.kernel;
        d.diffusion = run.diffusion;
        d.extrapolatedWalls = run.extrapolatedWalls;
        d.negativePressure = run.negativePressure;
        d.correction = run.correction;
        d.limiter = run.limiter;
        d.uniform = run.hmin == run.hmax;
        d.filterInterval = run.filterInterval;
        d.closure = ctx->closureOn;
        d.cell = run.cell;
        d.hmax = run.hmax;
        d.hmin = run.hmin;
        d.period = run.period;
        d.rho0 = run.rho0;
        d.c0 = run.c0;
        d.alpha = run.alpha;
        d.boundaryFactor = run.boundaryFactor;
        for (int a = 0; a < 3; ++a) d.gravity[a] = run.gravity[a];
        d.delta = run.delta;
        d.densityFloor = run.densityFloor;
        d.cfl = run.cfl;
        d.maxTimeStep = run.maxTimeStep;
        d.sigmaUnit = sigmaUnitOf(run.kernel, ctx->dims);
        d.stiffness = run.rho0 * run.c0 * run.c0 / 7.0;
        d.boundary = run.boundary;
        ctx->runsHost.push_back(d);
        cells += (int64_t)d.nx * d.ny * d.nz;
        total += run.count;
    }
    if (cells >= ((int64_t)1 << 31) - 1) return fail(ctx, SPH_ERR_VALUE, "Too many cells across the runs.");
    if (sameLayout) {
        HIP_TRY(hipMemcpyAsync(ctx->dRuns, ctx->runsHost.data(), count * sizeof(RunDev), hipMemcpyHostToDevice,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return SPH_OK;
    }
    ctx->cells = cells;
    ctx->total = total;
    ctx->capacity = std::max<int64_t>(capacity, 1);
    ctx->scanTiles = (cells + kScanTile - 1) / kScanTile;
    const int64_t cap = ctx->capacity;

    status = allocate(ctx, &ctx->dRuns, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->runOf, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->key, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->order, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->scattered, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->cellCount, cells);
    if (status == SPH_OK) status = allocate(ctx, &ctx->cellFill, cells);
    if (status == SPH_OK) status = allocate(ctx, &ctx->cellStart, cells + 1);
    if (status == SPH_OK) status = allocate(ctx, &ctx->eighth, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->cellEighths, cells);
    if (status == SPH_OK) status = allocate(ctx, &ctx->scanPartial, ctx->scanTiles);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.time, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.until, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.dt, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.steps, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.maxSteps, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.maxima, 2 * (size_t)count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.active, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.failed, count);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.speed, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->clock.acceleration, cap);
    if (status == SPH_OK) status = allocate(ctx, &ctx->dVel, 3);
    std::vector<int32_t> chunkRun;
    std::vector<int64_t> chunkFirst, runChunk(1, 0);
    for (int32_t r = 0; r < count; ++r) {
        for (int64_t first = runs[r].first; first < runs[r].first + runs[r].count; first += kSumChunk) {
            chunkRun.push_back(r);
            chunkFirst.push_back(first);
        }
        runChunk.push_back((int64_t)chunkRun.size());
    }
    ctx->sumChunkCount = (int64_t)chunkRun.size();
    if (status == SPH_OK) status = allocate(ctx, &ctx->sumChunkRun, chunkRun.size());
    if (status == SPH_OK) status = allocate(ctx, &ctx->sumChunkFirst, chunkFirst.size());
    if (status == SPH_OK) status = allocate(ctx, &ctx->sumRunChunk, runChunk.size());
    if (status == SPH_OK) status = allocate(ctx, &ctx->sumPartial, chunkRun.size() * kSums);
    if (status == SPH_OK) status = allocate(ctx, &ctx->sumOut, (size_t)count * kSums);
    if (status == SPH_OK)
        status = ctx->precision == 0 ? allocateRecords(ctx, ctx->rec32, cap) : allocateRecords(ctx, ctx->rec64, cap);
    if (status == SPH_OK) status = allocateList(ctx);
    if (status != SPH_OK) {
        releaseScratch(ctx);
        return status;
    }
    std::vector<int32_t> runOf(cap, count - 1);
    for (int32_t r = 0; r < count; ++r)
        std::fill(runOf.begin() + runs[r].first, runOf.begin() + runs[r].first + runs[r].count, r);
    HIP_TRY(hipMemcpyAsync(ctx->dRuns, ctx->runsHost.data(), count * sizeof(RunDev), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->runOf, runOf.data(), cap * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    if (!chunkRun.empty()) {
        HIP_TRY(hipMemcpyAsync(ctx->sumChunkRun, chunkRun.data(), chunkRun.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->sumChunkFirst, chunkFirst.data(), chunkFirst.size() * sizeof(int64_t),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(ctx->sumRunChunk, runChunk.data(), runChunk.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->cellCount, 0, std::max<int64_t>(cells, 1) * sizeof(int32_t), ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->cellFill, 0, std::max<int64_t>(cells, 1) * sizeof(int32_t), ctx->stream));
    std::vector<double> zeros(count, 0.0);
    std::vector<int64_t> zeroSteps(count, 0);
    HIP_TRY(hipMemcpyAsync(ctx->clock.time, zeros.data(), count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.until, zeros.data(), count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.dt, zeros.data(), count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.steps, zeroSteps.data(), count * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.maxSteps, zeroSteps.data(), count * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->clock.maxima, 0, 2 * count * sizeof(unsigned long long), ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->clock.failed, 0, count * sizeof(int32_t), ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_bind(sph_context* ctx, const sph_state* state) {
    if (!ctx) return SPH_ERR_VALUE;
    if (ctx->runs.empty()) return fail(ctx, SPH_ERR_VALUE, "Configure the runs before binding a state.");
    if (!state || !state->position || !state->velocity || !state->density || !state->pressure || !state->acceleration ||
        !state->mass || !state->length || !state->fluid)
        return fail(ctx, SPH_ERR_VALUE, "Every state array is needed.");
    if (state->stride < ctx->total) return fail(ctx, SPH_ERR_VALUE, "The state's row stride is shorter than the runs.");
    if (state->stride > ctx->capacity) return fail(ctx, SPH_ERR_VALUE, "The state's row stride exceeds the capacity.");
    ctx->state = *state;
    double* rows[3] = {state->velocity, state->velocity + state->stride,
                       ctx->dims == 3 ? state->velocity + 2 * state->stride : nullptr};
    HIP_TRY(hipMemcpyAsync(ctx->dVel, rows, sizeof rows, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    dropGraph(ctx);
    ctx->bound = true;
    return SPH_OK;
}

int sph_build_cells(sph_context* ctx) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    return ctx->dims == 2 ? buildCells<2>(ctx) : buildCells<3>(ctx);
}

int sph_wall_pressure(sph_context* ctx, const double* density, const void* pressure, void* pressureOut,
                      double* densityOut) {
    if (!pressure || !pressureOut || !densityOut) return fail(ctx, SPH_ERR_VALUE, "The wall pressure needs its arrays.");
    return dispatchHostSweep(ctx, 0, nullptr, density, pressure, pressureOut, densityOut);
}

int sph_velocity_gradients(sph_context* ctx, const double* velocity, const double* density, void* out) {
    if (!velocity || !out) return fail(ctx, SPH_ERR_VALUE, "The velocity gradients need velocities and an output.");
    return dispatchHostSweep(ctx, 1, velocity, density, nullptr, out, nullptr);
}

int sph_correction_matrices(sph_context* ctx, const double* density, void* out) {
    if (!out) return fail(ctx, SPH_ERR_VALUE, "The correction matrices need an output.");
    return dispatchHostSweep(ctx, 2, nullptr, density, nullptr, out, nullptr);
}

int sph_pair_accelerations(sph_context* ctx, const double* velocity, const double* density, const void* pressure,
                           void* out) {
    if (!velocity || !pressure || !out) return fail(ctx, SPH_ERR_VALUE, "The accelerations need velocities, pressures and an output.");
    return dispatchHostSweep(ctx, 3, velocity, density, pressure, out, nullptr);
}

int sph_continuity_rate(sph_context* ctx, const double* velocity, const double* density, void* out) {
    if (!velocity || !out) return fail(ctx, SPH_ERR_VALUE, "The continuity rate needs velocities and an output.");
    return dispatchHostSweep(ctx, 4, velocity, density, nullptr, out, nullptr);
}

int sph_shifting_displacement(sph_context* ctx, const double* velocity, const double* density, double size,
                              int32_t local, double spacingKernel, void* shifts, void* densityChanges,
                              void* velocityChanges) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!velocity || !density || !shifts || !densityChanges || !velocityChanges)
        return fail(ctx, SPH_ERR_VALUE, "Shifting needs velocities, densities and three outputs.");
    if (!(size >= 0.0) || !(spacingKernel > 0.0))
        return fail(ctx, SPH_ERR_VALUE, "The shift's size must be at least zero and W(dx) positive.");
    if (ctx->dims == 2)
        return ctx->precision == 0
                   ? shiftSweep<2, float>(ctx, velocity, density, size, local, spacingKernel, shifts, densityChanges, velocityChanges)
                   : shiftSweep<2, double>(ctx, velocity, density, size, local, spacingKernel, shifts, densityChanges, velocityChanges);
    return ctx->precision == 0
               ? shiftSweep<3, float>(ctx, velocity, density, size, local, spacingKernel, shifts, densityChanges, velocityChanges)
               : shiftSweep<3, double>(ctx, velocity, density, size, local, spacingKernel, shifts, densityChanges, velocityChanges);
}

int sph_set_closure(sph_context* ctx, int32_t enabled, double soundSpeed, double referenceDensity,
                    double kinematicViscosity) {
    if (!ctx) return SPH_ERR_VALUE;
    if (enabled && !(soundSpeed > 0.0 && referenceDensity > 0.0 && kinematicViscosity >= 0.0))
        return fail(ctx, SPH_ERR_VALUE, "The closure needs a positive c0 and rho0 and a viscosity of at least zero.");
    ctx->closureOn = enabled ? 1 : 0;
    if (enabled) {
        ctx->closureC0 = soundSpeed;
        ctx->closureRho0 = referenceDensity;
        ctx->closureWater = 2.0 * (ctx->dims + 2) * kinematicViscosity / soundSpeed;
    }
    // the run descriptors carry the flag, so sweep 1 builds G for the closure
    for (RunDev& run : ctx->runsHost) run.closure = ctx->closureOn;
    if (ctx->dRuns) {
        HIP_TRY(hipMemcpyAsync(ctx->dRuns, ctx->runsHost.data(), ctx->runsHost.size() * sizeof(RunDev),
                               hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return SPH_OK;
}

int sph_closure_coefficients(sph_context* ctx, const double* velocity, const double* density, void* alphas,
                             void* deltas) {
    if (!velocity || !alphas || !deltas) return fail(ctx, SPH_ERR_VALUE, "The closure needs velocities and two outputs.");
    int status = dispatchHostSweep(ctx, 8, velocity, density, nullptr, nullptr, nullptr);
    if (status != SPH_OK || ctx->total == 0) return status;
    const size_t real = ctx->precision == 0 ? sizeof(float) : sizeof(double);
    const size_t bytes = (size_t)ctx->total * real;
    const char* eddy = static_cast<const char*>(ctx->eddy);
    HIP_TRY(hipMemcpyAsync(alphas, eddy, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(deltas, eddy + (size_t)ctx->state.stride * real, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SPH_OK;
}

int sph_density_gradients(sph_context* ctx, const double* density, void* out) {
    if (!out) return fail(ctx, SPH_ERR_VALUE, "The density gradients need an output.");
    return dispatchHostSweep(ctx, 7, nullptr, density, nullptr, out, nullptr);
}

int sph_shepard_density(sph_context* ctx, const double* density, void* out) {
    if (!out) return fail(ctx, SPH_ERR_VALUE, "The Shepard filter needs an output.");
    return dispatchHostSweep(ctx, 5, nullptr, density, nullptr, out, nullptr);
}

int sph_xsph_velocity(sph_context* ctx, const double* velocity, const double* density, double epsilon, void* out) {
    if (!velocity || !out) return fail(ctx, SPH_ERR_VALUE, "XSPH needs velocities and an output.");
    if (!(epsilon >= 0.0 && epsilon <= 1.0)) return fail(ctx, SPH_ERR_VALUE, "XSPH's epsilon lies between 0 and 1.");
    return dispatchHostSweep(ctx, 6, velocity, density, nullptr, out, nullptr, epsilon);
}

int sph_neighbor_counts(sph_context* ctx, int32_t* counts) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!counts) return fail(ctx, SPH_ERR_VALUE, "No counts array.");
    if (ctx->dims == 2)
        return ctx->precision == 0 ? neighborsImpl<2, float>(ctx, counts, nullptr, nullptr)
                                   : neighborsImpl<2, double>(ctx, counts, nullptr, nullptr);
    return ctx->precision == 0 ? neighborsImpl<3, float>(ctx, counts, nullptr, nullptr)
                               : neighborsImpl<3, double>(ctx, counts, nullptr, nullptr);
}

int sph_neighbor_rows(sph_context* ctx, const int64_t* offsets, int32_t* rows) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!offsets || !rows) return fail(ctx, SPH_ERR_VALUE, "No offsets or rows array.");
    if (ctx->dims == 2)
        return ctx->precision == 0 ? neighborsImpl<2, float>(ctx, nullptr, offsets, rows)
                                   : neighborsImpl<2, double>(ctx, nullptr, offsets, rows);
    return ctx->precision == 0 ? neighborsImpl<3, float>(ctx, nullptr, offsets, rows)
                               : neighborsImpl<3, double>(ctx, nullptr, offsets, rows);
}

int sph_enforce(sph_context* ctx) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    const int64_t n = ctx->total, st = ctx->state.stride;
    if (n == 0) return SPH_OK;
    double* P = ctx->state.position;
    double* V = ctx->state.velocity;
    if (ctx->dims == 2)
        enforceKernel<2><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->state.fluid, P, P + st, P,
                                                                    V, V + st, V, n);
    else
        enforceKernel<3><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->dRuns, ctx->runOf, ctx->state.fluid, P, P + st,
                                                                    P + 2 * st, V, V + st, V + 2 * st, n);
    return launchCheck(ctx, "enforce");
}

int sph_set_clock(sph_context* ctx, int32_t run, double time, double untilTime, int64_t steps, int64_t maxSteps) {
    if (!ctx) return SPH_ERR_VALUE;
    if (run < 0 || run >= (int32_t)ctx->runs.size()) return fail(ctx, SPH_ERR_VALUE, "No such run.");
    HIP_TRY(hipMemcpyAsync(ctx->clock.time + run, &time, sizeof time, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.until + run, &untilTime, sizeof untilTime, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.steps + run, &steps, sizeof steps, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->clock.maxSteps + run, &maxSteps, sizeof maxSteps, hipMemcpyHostToDevice, ctx->stream));
    const int32_t cleared = 0;
    HIP_TRY(hipMemcpyAsync(ctx->clock.failed + run, &cleared, sizeof cleared, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_run_failures(sph_context* ctx, int32_t* failures) {
    if (!ctx) return SPH_ERR_VALUE;
    if (!failures) return fail(ctx, SPH_ERR_VALUE, "No failures array.");
    const size_t count = ctx->runs.size();
    if (count == 0) return SPH_OK;
    HIP_TRY(hipMemcpyAsync(failures, ctx->clock.failed, count * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_get_clock(sph_context* ctx, double* time, double* lastDt, int64_t* steps, double* maxima) {
    if (!ctx) return SPH_ERR_VALUE;
    const size_t count = ctx->runs.size();
    if (count == 0) return SPH_OK;
    if (time) HIP_TRY(hipMemcpyAsync(time, ctx->clock.time, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (lastDt) HIP_TRY(hipMemcpyAsync(lastDt, ctx->clock.dt, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (steps) HIP_TRY(hipMemcpyAsync(steps, ctx->clock.steps, count * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    if (maxima) HIP_TRY(hipMemcpyAsync(maxima, ctx->clock.maxima, 2 * count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_advance(sph_context* ctx, int32_t steps, int32_t useGraph) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (steps < 0) return fail(ctx, SPH_ERR_VALUE, "The step count cannot be negative.");
    if (steps == 0) return SPH_OK;
    status = ensureStepBuffers(ctx, 2);
    if (status != SPH_OK) return status;
    if (!useGraph || ctx->profiling) return dispatchAdvance(ctx, steps);
    if (ctx->ownsStream == false && ctx->stream == nullptr)
        return fail(ctx, SPH_ERR_VALUE, "A graph needs a stream other than the null stream.");
    // a graph holds the launches of the options it was captured with
    const StepOptions o = stepOptions(ctx, 2);
    const int key = (o.filter ? 1 : 0) | (o.antuono ? 2 : 0) | (o.closure ? 4 : 0) | (o.ledger ? 8 : 0);
    if (ctx->graphSteps == steps && ctx->graphKey != key) dropGraph(ctx);
    ctx->graphKey = key;
    if (ctx->graphSteps != steps) {
        dropGraph(ctx);
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        status = dispatchAdvance(ctx, steps);
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(ctx->stream, &graph);
        if (status != SPH_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return status;
        }
        if (end != hipSuccess) return fail(ctx, SPH_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(end));
        ctx->graph = graph;
        HIP_TRY(hipGraphInstantiate(&ctx->graphExec, graph, nullptr, nullptr, 0));
        ctx->graphSteps = steps;
    }
    HIP_TRY(hipGraphLaunch(ctx->graphExec, ctx->stream));
    return SPH_OK;
}

int sph_rates(sph_context* ctx, double* maxima) { return sph_resident_rates(ctx, 0, maxima); }

int sph_resident_rates(sph_context* ctx, int32_t filterDensity, double* maxima) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    status = dispatchPhase(ctx, 1, 0.0, filterDensity ? 1 : 0);
    if (status != SPH_OK) return status;
    if (maxima) {
        HIP_TRY(hipMemcpyAsync(maxima, ctx->clock.maxima, 2 * ctx->runs.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return SPH_OK;
}

namespace {

// A copy of the bound state (and `identity`) in the context's copy buffer, as a state to read from.
int snapshotState(sph_context* ctx, int32_t* identity, sph_state& from, const int32_t*& identityFrom) {
    const int D = ctx->dims;
    const size_t real = ctx->precision == 0 ? sizeof(float) : sizeof(double);
    const size_t stride = (size_t)ctx->state.stride;
    const size_t sizes[9] = {8 * D * stride, 8 * D * stride, 8 * stride, real * stride, real * D * stride, real * stride,
                             real * stride, stride, identity ? 4 * stride : 0};
    size_t offsets[9], total = 0;
    for (int f = 0; f < 9; ++f) {
        offsets[f] = total;
        total += (sizes[f] + 255) / 256 * 256;
    }
    if (total > ctx->sortBytes) {
        if (ctx->sortCopy) (void)hipFree(ctx->sortCopy);
        ctx->sortCopy = nullptr;
        ctx->sortBytes = 0;
        HIP_TRY(hipMalloc(&ctx->sortCopy, total));
        ctx->sortBytes = total;
    }
    char* base = static_cast<char*>(ctx->sortCopy);
    from = ctx->state;
    void* fields[9] = {ctx->state.position, ctx->state.velocity, ctx->state.density, ctx->state.pressure,
                       ctx->state.acceleration, ctx->state.mass, ctx->state.length, ctx->state.fluid, identity};
    for (int f = 0; f < 9; ++f)
        if (sizes[f]) HIP_TRY(hipMemcpyAsync(base + offsets[f], fields[f], sizes[f], hipMemcpyDeviceToDevice, ctx->stream));
    from.position = reinterpret_cast<double*>(base + offsets[0]);
    from.velocity = reinterpret_cast<double*>(base + offsets[1]);
    from.density = reinterpret_cast<double*>(base + offsets[2]);
    from.pressure = base + offsets[3];
    from.acceleration = base + offsets[4];
    from.mass = base + offsets[5];
    from.length = base + offsets[6];
    from.fluid = reinterpret_cast<uint8_t*>(base + offsets[7]);
    identityFrom = identity ? reinterpret_cast<const int32_t*>(base + offsets[8]) : nullptr;
    return SPH_OK;
}

}  // namespace

int sph_sort_state(sph_context* ctx, int32_t* identity) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    const int64_t n = ctx->total;
    if (n == 0) return SPH_OK;
    const int D = ctx->dims;
    sph_state from{};
    const int32_t* identityFrom = nullptr;
    status = snapshotState(ctx, identity, from, identityFrom);
    if (status != SPH_OK) return status;
    status = D == 2 ? buildCells<2>(ctx) : buildCells<3>(ctx);
    if (status != SPH_OK) return status;
    if (ctx->precision == 0)
        permuteState<float><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->order, from, ctx->state, identityFrom, identity, D, n);
    else
        permuteState<double><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->order, from, ctx->state, identityFrom, identity, D, n);
    return launchCheck(ctx, "state sort");
}

int sph_split_count(sph_context* ctx, const sph_split_spec* specs, const uint8_t* mask, int64_t* chosen) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!specs || !chosen) return fail(ctx, SPH_ERR_VALUE, "The split needs its specifications and an output.");
    const int32_t runs = (int32_t)ctx->runs.size();
    for (int32_t r = 0; r < runs; ++r)
        if (ctx->runs[r].first != (r == 0 ? 0 : ctx->runs[r - 1].first + ctx->runs[r - 1].count))
            return fail(ctx, SPH_ERR_VALUE, "The split needs the runs packed from slot 0.");
    const int64_t n = ctx->total;
    // Marks and their exclusive scan (n + 1 entries), the scan's tiles, the runs' specs and ranks
    const int64_t tiles = (n + 1 + kScanTile - 1) / kScanTile;
    const int64_t slots = 2 * (n + 1) + tiles + 2 * (runs + 1) + (int64_t)(runs * sizeof(sph_split_spec) + 3) / 4 + 256;
    if (slots > ctx->splitSlots) {
        if (ctx->splitMarks) (void)hipFree(ctx->splitMarks);
        ctx->splitMarks = nullptr;
        ctx->splitSlots = 0;
        HIP_TRY(hipMalloc(&ctx->splitMarks, slots * sizeof(int32_t)));
        ctx->splitSlots = slots;
    }
    int32_t* marks = ctx->splitMarks;
    int32_t* ranks = marks + (n + 1);
    int32_t* tilePartial = ranks + (n + 1);
    int64_t* runFirst = reinterpret_cast<int64_t*>(tilePartial + tiles + (tiles & 1));
    sph_split_spec* dSpecs = reinterpret_cast<sph_split_spec*>(runFirst + (runs + 1));
    HIP_TRY(hipMemcpyAsync(dSpecs, specs, runs * sizeof(sph_split_spec), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(marks, 0, (n + 1) * sizeof(int32_t), ctx->stream));
    if (n > 0)
        splitMarks<<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, mask, ctx->precision, marks, n);
    scanTiles<<<(unsigned)tiles, kScanThreads, 0, ctx->stream>>>(marks, ranks, tilePartial, n);
    scanPartials<<<1, kScanThreads, 0, ctx->stream>>>(tilePartial, tiles, ranks, n);
    scanAdd<<<(unsigned)tiles, kScanThreads, 0, ctx->stream>>>(ranks, tilePartial, n);
    runRanks<<<(runs + 64) / 64, 64, 0, ctx->stream>>>(ctx->dRuns, ranks, runs, n, runFirst);
    status = launchCheck(ctx, "split marks");
    if (status != SPH_OK) return status;
    std::vector<int64_t> first(runs + 1);
    HIP_TRY(hipMemcpyAsync(first.data(), runFirst, (runs + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->splitChosen.assign(runs, 0);
    for (int32_t r = 0; r < runs; ++r) chosen[r] = ctx->splitChosen[r] = first[r + 1] - first[r];
    return SPH_OK;
}

int sph_split_apply(sph_context* ctx, const sph_state* into, double referenceDensity, double lengthRatio) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    const int32_t runs = (int32_t)ctx->runs.size();
    if ((int32_t)ctx->splitChosen.size() != runs || !ctx->splitMarks)
        return fail(ctx, SPH_ERR_VALUE, "sph_split_count must come right before sph_split_apply.");
    if (!into || !into->position || !into->velocity || !into->density || !into->pressure || !into->acceleration ||
        !into->mass || !into->length || !into->fluid)
        return fail(ctx, SPH_ERR_VALUE, "The split needs every array of the state it writes.");
    if (!(referenceDensity > 0.0) || !(lengthRatio > 0.0))
        return fail(ctx, SPH_ERR_VALUE, "The split needs a positive reference density and length ratio.");
    const int daughters = 1 << ctx->dims;
    std::vector<int64_t> newFirst(runs);
    int64_t total = 0;
    for (int32_t r = 0; r < runs; ++r) {
        newFirst[r] = total;
        total += ctx->runs[r].count + ctx->splitChosen[r] * (daughters - 1);
    }
    if (into->stride < total) return fail(ctx, SPH_ERR_VALUE, "The state the split writes is too small for the grown runs.");
    const int64_t n = ctx->total;
    int32_t* marks = ctx->splitMarks;
    int32_t* ranks = marks + (n + 1);
    const int64_t tiles = (n + 1 + kScanTile - 1) / kScanTile;
    int64_t* runFirst = reinterpret_cast<int64_t*>(ranks + (n + 1) + tiles + (tiles & 1));
    // The new runs' first slots after the ranks, in the specs' place
    int64_t* dNewFirst = runFirst + (runs + 1);
    HIP_TRY(hipMemcpyAsync(dNewFirst, newFirst.data(), runs * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    sph_state from{};
    const int32_t* identityFrom = nullptr;
    status = snapshotState(ctx, nullptr, from, identityFrom);
    if (status != SPH_OK) return status;
    if (n > 0) {
        if (ctx->precision == 0)
            splitScatter<float><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->dRuns, dNewFirst, runFirst, marks,
                                                                          ranks, from, *into, ctx->dims, referenceDensity,
                                                                          lengthRatio, n);
        else
            splitScatter<double><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->dRuns, dNewFirst, runFirst,
                                                                           marks, ranks, from, *into, ctx->dims,
                                                                           referenceDensity, lengthRatio, n);
    }
    status = launchCheck(ctx, "split");
    if (status != SPH_OK) return status;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->splitChosen.clear();
    return SPH_OK;
}


int sph_merge(sph_context* ctx, const sph_merge_spec* specs, const uint8_t* mask, int64_t* groups, double* removedEnergy) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!specs || !groups || !removedEnergy) return fail(ctx, SPH_ERR_VALUE, "The merge needs its specifications and outputs.");
    const int32_t runs = (int32_t)ctx->runs.size();
    for (int32_t r = 0; r < runs; ++r) {
        if (ctx->runs[r].first != (r == 0 ? 0 : ctx->runs[r - 1].first + ctx->runs[r - 1].count))
            return fail(ctx, SPH_ERR_VALUE, "The merge needs the runs packed from slot 0.");
        if (specs[r].active && !(specs[r].referenceMass > 0.0 && specs[r].referenceDensity > 0.0 && specs[r].lengthRatio > 0.0))
            return fail(ctx, SPH_ERR_VALUE, "A merge needs a positive reference mass, density and length ratio.");
        groups[r] = 0;
        removedEnergy[r] = 0.0;
    }
    const int64_t n = ctx->total;
    if (n == 0) return SPH_OK;
    const int D = ctx->dims, size = 1 << D;
    const int64_t tileCount = (n + 1 + kScanTile - 1) / kScanTile + 1;
    // Sizes first, then the pieces
    Carver sizing{nullptr};
    for (int k = 0; k < 12; ++k) sizing.take<int32_t>(n + 1);
    sizing.take<double>(n);
    sizing.take<int32_t>(tileCount);
    sizing.take<sph_merge_spec>(runs);
    sizing.take<MergePass>(runs);
    sizing.take<int32_t>((int64_t)runs * kMergeLevels);
    for (int k = 0; k < 3; ++k) sizing.take<long long>(runs);
    sizing.take<long long>((int64_t)runs * 3);
    sizing.take<long long>((int64_t)runs * 3);
    sizing.take<int32_t>(runs);
    sizing.take<int32_t>(1);
    sizing.take<double>(runs);
    sizing.take<long long>(runs + 1);
    sizing.take<long long>(runs + 1);
    status = growBuffer(ctx, ctx->mergeScratch, ctx->mergeBytes, sizing.used);
    if (status != SPH_OK) return status;
    Carver carve{static_cast<char*>(ctx->mergeScratch)};
    int32_t* levelOf = carve.take<int32_t>(n + 1);
    int32_t* key = carve.take<int32_t>(n + 1);
    int32_t* removed = carve.take<int32_t>(n + 1);
    int32_t* hole = carve.take<int32_t>(n + 1);
    int32_t* tail = carve.take<int32_t>(n + 1);
    int32_t* keep = carve.take<int32_t>(n + 1);
    int32_t* holeRank = carve.take<int32_t>(n + 1);
    int32_t* tailRank = carve.take<int32_t>(n + 1);
    int32_t* newIndex = carve.take<int32_t>(n + 1);
    int32_t* members = carve.take<int32_t>(n + 1);
    int32_t* groupMembers = carve.take<int32_t>(n + 1);
    int32_t* movers = carve.take<int32_t>(n + 1);
    double* distance = carve.take<double>(n);
    int32_t* tiles = carve.take<int32_t>(tileCount);
    sph_merge_spec* dSpecs = carve.take<sph_merge_spec>(runs);
    MergePass* dPasses = carve.take<MergePass>(runs);
    int32_t* histogram = carve.take<int32_t>((int64_t)runs * kMergeLevels);
    long long* lastBoundary = carve.take<long long>(runs);
    carve.take<long long>(runs);
    carve.take<long long>(runs);
    long long* low = carve.take<long long>((int64_t)runs * 3);
    long long* high = carve.take<long long>((int64_t)runs * 3);
    int32_t* taken = carve.take<int32_t>(runs);
    int32_t* groupCount = carve.take<int32_t>(1);
    double* energy = carve.take<double>(runs);
    long long* holeBase = carve.take<long long>(runs + 1);
    long long* tailBase = carve.take<long long>(runs + 1);

    HIP_TRY(hipMemcpyAsync(dSpecs, specs, runs * sizeof(sph_merge_spec), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(histogram, 0, (size_t)runs * kMergeLevels * sizeof(int32_t), ctx->stream));
    HIP_TRY(hipMemsetAsync(lastBoundary, 0xff, runs * sizeof(long long), ctx->stream));
    HIP_TRY(hipMemsetAsync(removed, 0, (n + 1) * sizeof(int32_t), ctx->stream));
    HIP_TRY(hipMemsetAsync(taken, 0, runs * sizeof(int32_t), ctx->stream));
    HIP_TRY(hipMemsetAsync(groupCount, 0, sizeof(int32_t), ctx->stream));
    HIP_TRY(hipMemsetAsync(energy, 0, runs * sizeof(double), ctx->stream));
    mergeLevels<<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, mask, ctx->precision, D, levelOf,
                                                         histogram, lastBoundary, n);
    status = launchCheck(ctx, "merge levels");
    if (status != SPH_OK) return status;
    std::vector<int32_t> freeCount((size_t)runs * kMergeLevels);
    std::vector<long long> lastHost(runs);
    HIP_TRY(hipMemcpyAsync(freeCount.data(), histogram, freeCount.size() * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(lastHost.data(), lastBoundary, runs * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));

    std::vector<MergePass> passes(runs);
    std::vector<long long> lowHost((size_t)runs * 3), highHost((size_t)runs * 3);
    std::vector<int32_t> takenHost(runs);
    std::vector<int64_t> membersTaken(runs, 0);
    const double shifts[2] = {0.0, 0.5};
    for (int level = 1; level < kMergeLevels; ++level)
        for (double shift : shifts) {
            bool any = false;
            for (int32_t r = 0; r < runs; ++r) {
                MergePass& pass = passes[r];
                pass = MergePass{};
                pass.level = level;
                pass.offset = shift;
                pass.active = specs[r].active && freeCount[(size_t)r * kMergeLevels + level] >= size;
                if (!pass.active) continue;
                any = true;
                pass.cell = std::pow(specs[r].referenceMass / std::pow((double)size, level - 1) / specs[r].referenceDensity,
                                     1.0 / D);
                for (int a = 0; a < 3; ++a) {
                    lowHost[r * 3 + a] = std::numeric_limits<long long>::max();
                    highHost[r * 3 + a] = std::numeric_limits<long long>::min();
                }
            }
            if (!any) continue;
            HIP_TRY(hipMemcpyAsync(dPasses, passes.data(), runs * sizeof(MergePass), hipMemcpyHostToDevice, ctx->stream));
            HIP_TRY(hipMemcpyAsync(low, lowHost.data(), lowHost.size() * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
            HIP_TRY(hipMemcpyAsync(high, highHost.data(), highHost.size() * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
            mergeBounds<<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dPasses, D, levelOf, low, high, n);
            HIP_TRY(hipMemcpyAsync(lowHost.data(), low, lowHost.size() * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipMemcpyAsync(highHost.data(), high, highHost.size() * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            long long cells = 0;
            for (int32_t r = 0; r < runs; ++r) {
                MergePass& pass = passes[r];
                if (!pass.active) continue;
                long long extent = 1;
                for (int a = 0; a < D; ++a) {
                    pass.low[a] = lowHost[r * 3 + a];
                    pass.extent[a] = highHost[r * 3 + a] - lowHost[r * 3 + a] + 1;
                    extent *= pass.extent[a];
                    if (extent > (1ll << 30)) return fail(ctx, SPH_ERR_VALUE, "The merge's cells span too wide a grid.");
                }
                pass.cellBase = cells;
                cells += extent;
            }
            if (cells >= (1ll << 30)) return fail(ctx, SPH_ERR_VALUE, "The merge's cells span too wide a grid.");
            const int64_t cellTiles = (cells + 1 + kScanTile - 1) / kScanTile + 1;
            status = growBuffer(ctx, ctx->mergeCells, ctx->mergeCellBytes,
                                (size_t)(2 * (cells + 1) + cellTiles) * sizeof(int32_t) + 1024);
            if (status != SPH_OK) return status;
            Carver cellCarve{static_cast<char*>(ctx->mergeCells)};
            int32_t* counts = cellCarve.take<int32_t>(cells + 1);
            int32_t* start = cellCarve.take<int32_t>(cells + 1);
            int32_t* cellTileSums = cellCarve.take<int32_t>(cellTiles);
            HIP_TRY(hipMemcpyAsync(dPasses, passes.data(), runs * sizeof(MergePass), hipMemcpyHostToDevice, ctx->stream));
            HIP_TRY(hipMemsetAsync(counts, 0, (cells + 1) * sizeof(int32_t), ctx->stream));
            mergeKeys<<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dPasses, D, levelOf, key, distance,
                                                               counts, n);
            scanInto(ctx, counts, start, cellTileSums, cells);
            fractionScatter<<<gridFor(n, 256), 256, 0, ctx->stream>>>(key, start, counts, members, n);
            mergeOrder<<<gridFor(cells, 256), 256, 0, ctx->stream>>>(start, members, distance, cells);
            mergeTake<<<gridFor(cells, 256), 256, 0, ctx->stream>>>(start, members, size, ctx->runOf, groupMembers,
                                                                    groupCount, levelOf, taken, cells);
            status = launchCheck(ctx, "merge pass");
            if (status != SPH_OK) return status;
            HIP_TRY(hipMemcpyAsync(takenHost.data(), taken, runs * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipMemsetAsync(taken, 0, runs * sizeof(int32_t), ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            for (int32_t r = 0; r < runs; ++r) {
                freeCount[(size_t)r * kMergeLevels + level] -= takenHost[r];
                membersTaken[r] += takenHost[r];
            }
        }
    int32_t groupTotal = 0;
    HIP_TRY(hipMemcpyAsync(&groupTotal, groupCount, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (groupTotal == 0) return SPH_OK;

    // Who leaves, and whether the holes among each run's boundary can be filled, before any change
    markRemoved<<<gridFor(groupTotal, 256), 256, 0, ctx->stream>>>(groupMembers, groupTotal, size, removed);
    compactFlags<<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, lastBoundary, removed, hole, tail, n);
    scanInto(ctx, hole, holeRank, tiles, n);
    scanInto(ctx, tail, tailRank, tiles, n);
    runBases<<<(runs + 64) / 64, 64, 0, ctx->stream>>>(ctx->dRuns, holeRank, runs, n, holeBase);
    runBases<<<(runs + 64) / 64, 64, 0, ctx->stream>>>(ctx->dRuns, tailRank, runs, n, tailBase);
    std::vector<long long> holesHost(runs + 1), tailsHost(runs + 1);
    HIP_TRY(hipMemcpyAsync(holesHost.data(), holeBase, (runs + 1) * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(tailsHost.data(), tailBase, (runs + 1) * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int32_t r = 0; r < runs; ++r)
        if (tailsHost[r + 1] - tailsHost[r] < holesHost[r + 1] - holesHost[r])
            return fail(ctx, SPH_ERR_VALUE, "Too few fluid particles follow the boundary to fill the places a merge leaves among it.");
    if (ctx->precision == 0)
        mergeGroups<float><<<gridFor(groupTotal, 256), 256, 0, ctx->stream>>>(groupMembers, groupTotal, size, ctx->runOf,
                                                                              dSpecs, ctx->state, D, removed, energy);
    else
        mergeGroups<double><<<gridFor(groupTotal, 256), 256, 0, ctx->stream>>>(groupMembers, groupTotal, size, ctx->runOf,
                                                                               dSpecs, ctx->state, D, removed, energy);
    compactMovers<<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, lastBoundary, removed, tail, tailRank, tailBase,
                                                            holeBase, movers, keep, n);
    scanInto(ctx, keep, newIndex, tiles, n);
    sph_state from{};
    const int32_t* identityFrom = nullptr;
    status = snapshotState(ctx, nullptr, from, identityFrom);
    if (status != SPH_OK) return status;
    if (ctx->precision == 0)
        compactPermute<float><<<gridFor(n, 256), 256, 0, ctx->stream>>>(keep, newIndex, hole, holeRank, movers, from,
                                                                        ctx->state, D, n);
    else
        compactPermute<double><<<gridFor(n, 256), 256, 0, ctx->stream>>>(keep, newIndex, hole, holeRank, movers, from,
                                                                         ctx->state, D, n);
    status = launchCheck(ctx, "merge");
    if (status != SPH_OK) return status;
    std::vector<double> energyHost(runs);
    HIP_TRY(hipMemcpyAsync(energyHost.data(), energy, runs * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int32_t r = 0; r < runs; ++r) {
        groups[r] = membersTaken[r] / size;
        removedEnergy[r] = energyHost[r];
    }
    return SPH_OK;
}

int sph_set_energy_budget(sph_context* ctx, int32_t enabled) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    ctx->ledgerOn = enabled ? 1 : 0;
    if (!enabled) return SPH_OK;
    const int64_t stride = ctx->capacity;
    const size_t real = ctx->precision == 0 ? sizeof(float) : sizeof(double);
    if (!ctx->ledger || ctx->ledgerStride != stride) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        releaseLedger(ctx);
        HIP_TRY(hipMalloc(&ctx->ledger, (size_t)kLedgerTerms * stride * sizeof(double)));
        HIP_TRY(hipMalloc(&ctx->ledgerParts, (size_t)(4 * ctx->dims + 4) * stride * real));
        HIP_TRY(hipMalloc(&ctx->ledgerPartial, (size_t)std::max<int64_t>(ctx->sumChunkCount, 1) * kLedgerTerms *
                                                   sizeof(double)));
        HIP_TRY(hipMalloc(&ctx->ledgerOut, ctx->runs.size() * kLedgerTerms * sizeof(double)));
        ctx->ledgerStride = stride;
    }
    // every term starts from zero, as the reference's push clears its ledger (sphGpu.py:2218-2219)
    HIP_TRY(hipMemsetAsync(ctx->ledger, 0, (size_t)kLedgerTerms * stride * sizeof(double), ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->ledgerParts, 0, (size_t)(4 * ctx->dims + 4) * stride * real, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_energy_sums(sph_context* ctx, double* sums) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!sums) return fail(ctx, SPH_ERR_VALUE, "The energy sums need an output array.");
    if (!ctx->ledgerOn || !ctx->ledger) return fail(ctx, SPH_ERR_VALUE, "The energy budget is off; see sph_set_energy_budget.");
    const int32_t runs = (int32_t)ctx->runs.size();
    if (ctx->sumChunkCount > 0)
        ledgerChunks<<<(unsigned)ctx->sumChunkCount, kSumThreads, 0, ctx->stream>>>(
            ctx->dRuns, ctx->sumChunkRun, ctx->sumChunkFirst, ctx->ledger, ctx->ledgerStride, ctx->ledgerPartial);
    ledgerRuns<<<gridFor(runs, 64), 64, 0, ctx->stream>>>(ctx->sumRunChunk, ctx->ledgerPartial, ctx->ledgerOut, runs);
    status = launchCheck(ctx, "energy sums");
    if (status != SPH_OK) return status;
    HIP_TRY(hipMemcpyAsync(sums, ctx->ledgerOut, (size_t)runs * kLedgerTerms * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_global_sums(sph_context* ctx, double* sums) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!sums) return fail(ctx, SPH_ERR_VALUE, "The sums need an output array.");
    const int32_t runs = (int32_t)ctx->runs.size();
    if (ctx->sumChunkCount > 0) {
        const unsigned chunks = (unsigned)ctx->sumChunkCount;
        if (ctx->dims == 2) {
            if (ctx->precision == 0)
                sumChunks<2, float><<<chunks, kSumThreads, 0, ctx->stream>>>(ctx->dRuns, ctx->sumChunkRun, ctx->sumChunkFirst,
                                                                       ctx->state, ctx->sumPartial);
            else
                sumChunks<2, double><<<chunks, kSumThreads, 0, ctx->stream>>>(ctx->dRuns, ctx->sumChunkRun, ctx->sumChunkFirst,
                                                                        ctx->state, ctx->sumPartial);
        } else {
            if (ctx->precision == 0)
                sumChunks<3, float><<<chunks, kSumThreads, 0, ctx->stream>>>(ctx->dRuns, ctx->sumChunkRun, ctx->sumChunkFirst,
                                                                       ctx->state, ctx->sumPartial);
            else
                sumChunks<3, double><<<chunks, kSumThreads, 0, ctx->stream>>>(ctx->dRuns, ctx->sumChunkRun, ctx->sumChunkFirst,
                                                                        ctx->state, ctx->sumPartial);
        }
    }
    sumRuns<<<gridFor(runs, 64), 64, 0, ctx->stream>>>(ctx->sumRunChunk, ctx->sumPartial, ctx->sumOut, runs);
    status = launchCheck(ctx, "global sums");
    if (status != SPH_OK) return status;
    HIP_TRY(hipMemcpyAsync(sums, ctx->sumOut, (size_t)runs * kSums * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

namespace {

// The read-out's device buffers, grown to `bytes` (allocates outside graph capture only).
int readBuffers(sph_context* ctx, size_t bytes) {
    if (bytes <= ctx->readBytes) return SPH_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->readScratch) (void)hipFree(ctx->readScratch);
    ctx->readScratch = nullptr;
    ctx->readBytes = 0;
    HIP_TRY(hipMalloc(&ctx->readScratch, bytes));
    ctx->readBytes = bytes;
    return SPH_OK;
}

size_t aligned(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

}  // namespace

int sph_surface_profile(sph_context* ctx, const sph_profile_spec* specs, double* tops) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!specs || !tops) return fail(ctx, SPH_ERR_VALUE, "The profile needs its specifications and an output.");
    const int32_t runs = (int32_t)ctx->runs.size();
    int64_t columns = 0;
    for (int32_t r = 0; r < runs; ++r) {
        if (specs[r].columns <= 0) continue;
        if (specs[r].offset < 0 || !(specs[r].width > 0.0))
            return fail(ctx, SPH_ERR_VALUE, "A profile needs a positive column width and a place in the output.");
        columns = std::max<int64_t>(columns, (int64_t)specs[r].offset + specs[r].columns);
    }
    const size_t specBytes = aligned(runs * sizeof(sph_profile_spec));
    status = readBuffers(ctx, specBytes + aligned(std::max<int64_t>(columns, 1) * sizeof(unsigned long long)));
    if (status != SPH_OK) return status;
    sph_profile_spec* dSpecs = static_cast<sph_profile_spec*>(ctx->readScratch);
    unsigned long long* dTops = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->readScratch) + specBytes);
    HIP_TRY(hipMemcpyAsync(dSpecs, specs, runs * sizeof(sph_profile_spec), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(dTops, 0, std::max<int64_t>(columns, 1) * sizeof(unsigned long long), ctx->stream));
    const int64_t n = ctx->total;
    if (n > 0) {
        if (ctx->dims == 2)
            surfaceTops<2><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, dTops, n);
        else
            surfaceTops<3><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, dTops, n);
    }
    status = launchCheck(ctx, "surface profile");
    if (status != SPH_OK) return status;
    std::vector<unsigned long long> keys(std::max<int64_t>(columns, 1));
    HIP_TRY(hipMemcpyAsync(keys.data(), dTops, keys.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int32_t r = 0; r < runs; ++r)
        for (int32_t c = 0; c < specs[r].columns; ++c) {
            const unsigned long long key = keys[specs[r].offset + c];
            double value = -std::numeric_limits<double>::infinity();
            if (key) {
                const unsigned long long bits = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
                std::memcpy(&value, &bits, sizeof value);
            }
            tops[specs[r].offset + c] = value;
        }
    return SPH_OK;
}

int sph_face_cells(sph_context* ctx, const sph_face_spec* specs, uint32_t* occupancy, int32_t* overflow,
                   double* landward) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!specs || !occupancy || !overflow || !landward) return fail(ctx, SPH_ERR_VALUE, "The face cells need their arrays.");
    const int32_t runs = (int32_t)ctx->runs.size();
    int64_t words = 0;
    for (int32_t r = 0; r < runs; ++r) {
        const sph_face_spec& spec = specs[r];
        if (!spec.active) continue;
        if (spec.columns < 0 || spec.bands < 0 || spec.offset < 0 || !(spec.spacing > 0.0) || spec.bedPoints < 0 ||
            (spec.bedPoints > 0 && (!spec.bedX || !spec.bedZ)))
            return fail(ctx, SPH_ERR_VALUE, "A face window needs columns, bands, a spacing, a place in the output and its bed.");
        words = std::max<int64_t>(words, (int64_t)spec.offset + (int64_t)spec.bands * ((spec.columns + 31) / 32));
    }
    const size_t specBytes = aligned(runs * sizeof(sph_face_spec));
    const size_t wordBytes = aligned(std::max<int64_t>(words, 1) * sizeof(uint32_t));
    const size_t flagBytes = aligned(runs * sizeof(int32_t));
    const size_t pointBytes = aligned((size_t)runs * 2 * kLandward * sizeof(double));
    // Slices of at most 2048 slots per workgroup, at most 512 per run
    int64_t longest = 0;
    for (int32_t r = 0; r < runs; ++r) longest = std::max<int64_t>(longest, ctx->runs[r].count);
    const int32_t slices = (int32_t)std::min<int64_t>(512, std::max<int64_t>(1, (longest + 2047) / 2048));
    const int64_t chunk = std::max<int64_t>(1, (longest + slices - 1) / slices);
    const size_t partialBytes = aligned((size_t)runs * slices * kLandward * sizeof(double2));
    status = readBuffers(ctx, specBytes + wordBytes + flagBytes + pointBytes + partialBytes);
    if (status != SPH_OK) return status;
    char* base = static_cast<char*>(ctx->readScratch);
    sph_face_spec* dSpecs = reinterpret_cast<sph_face_spec*>(base);
    uint32_t* dWords = reinterpret_cast<uint32_t*>(base + specBytes);
    int32_t* dFlags = reinterpret_cast<int32_t*>(base + specBytes + wordBytes);
    double* dPoints = reinterpret_cast<double*>(base + specBytes + wordBytes + flagBytes);
    double2* dPartial = reinterpret_cast<double2*>(base + specBytes + wordBytes + flagBytes + pointBytes);
    HIP_TRY(hipMemcpyAsync(dSpecs, specs, runs * sizeof(sph_face_spec), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(dWords, 0, std::max<int64_t>(words, 1) * sizeof(uint32_t), ctx->stream));
    HIP_TRY(hipMemsetAsync(dFlags, 0, runs * sizeof(int32_t), ctx->stream));
    const int64_t n = ctx->total;
    if (n > 0) {
        if (ctx->dims == 2)
            faceCells<2><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, dWords, dFlags, n);
        else
            faceCells<3><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, dWords, dFlags, n);
    }
    const dim3 grid((unsigned)runs, (unsigned)slices);
    if (ctx->dims == 2)
        landwardPoints<2><<<grid, 256, 0, ctx->stream>>>(ctx->dRuns, ctx->state, dSpecs, chunk, dPartial);
    else
        landwardPoints<3><<<grid, 256, 0, ctx->stream>>>(ctx->dRuns, ctx->state, dSpecs, chunk, dPartial);
    mergeLandward<<<(unsigned)runs, 64, 0, ctx->stream>>>(dPartial, slices, dPoints);
    status = launchCheck(ctx, "face cells");
    if (status != SPH_OK) return status;
    if (words > 0)
        HIP_TRY(hipMemcpyAsync(occupancy, dWords, words * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(overflow, dFlags, runs * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(landward, dPoints, (size_t)runs * 2 * kLandward * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_fluid_fraction(sph_context* ctx, const sph_fraction_spec* specs, double* fraction) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!specs || !fraction) return fail(ctx, SPH_ERR_VALUE, "The fluid fraction needs its specifications and an output.");
    const int32_t runs = (int32_t)ctx->runs.size();
    std::vector<int64_t> binOffset(runs, 0);
    int64_t bins = 0, nodes = 0;
    for (int32_t r = 0; r < runs; ++r) {
        const sph_fraction_spec& spec = specs[r];
        binOffset[r] = bins;
        if (spec.nx <= 0) continue;
        if (spec.ny <= 0 || spec.reach < 0 || spec.reach > 16 || spec.offset < 0 || !(spec.binWidth > 0.0) ||
            !(spec.h > 0.0) || (int64_t)spec.nx * spec.ny > (1ll << 28))
            return fail(ctx, SPH_ERR_VALUE, "A fraction grid needs nodes, a bin width, a smoothing length and a reach of at most 16 bins.");
        bins += (int64_t)(spec.nx + 2 * spec.reach + 2) * (spec.ny + 2 * spec.reach + 2);
        nodes = std::max<int64_t>(nodes, spec.offset + (int64_t)spec.nx * spec.ny);
    }
    if (bins >= (1ll << 31) - 1) return fail(ctx, SPH_ERR_VALUE, "The fraction grids hold too many bins.");
    if (nodes == 0) return SPH_OK;
    const int64_t n = ctx->total;
    const int64_t tiles = (bins + 1 + kScanTile - 1) / kScanTile;
    const size_t specBytes = aligned(runs * sizeof(sph_fraction_spec));
    const size_t offsetBytes = aligned(runs * sizeof(int64_t));
    const size_t keyBytes = aligned(std::max<int64_t>(n, 1) * sizeof(int32_t));
    const size_t binBytes = aligned((bins + 1) * sizeof(int32_t));
    const size_t tileBytes = aligned(std::max<int64_t>(tiles, 1) * sizeof(int32_t));
    const size_t nodeBytes = aligned(nodes * sizeof(double));
    status = readBuffers(ctx, specBytes + offsetBytes + 2 * keyBytes + 2 * binBytes + tileBytes + nodeBytes);
    if (status != SPH_OK) return status;
    char* base = static_cast<char*>(ctx->readScratch);
    sph_fraction_spec* dSpecs = reinterpret_cast<sph_fraction_spec*>(base);
    int64_t* dOffsets = reinterpret_cast<int64_t*>(base + specBytes);
    int32_t* dKey = reinterpret_cast<int32_t*>(base + specBytes + offsetBytes);
    int32_t* dMembers = reinterpret_cast<int32_t*>(base + specBytes + offsetBytes + keyBytes);
    int32_t* dCounts = reinterpret_cast<int32_t*>(base + specBytes + offsetBytes + 2 * keyBytes);
    int32_t* dStart = reinterpret_cast<int32_t*>(base + specBytes + offsetBytes + 2 * keyBytes + binBytes);
    int32_t* dTiles = reinterpret_cast<int32_t*>(base + specBytes + offsetBytes + 2 * keyBytes + 2 * binBytes);
    double* dFraction = reinterpret_cast<double*>(base + specBytes + offsetBytes + 2 * keyBytes + 2 * binBytes + tileBytes);
    HIP_TRY(hipMemcpyAsync(dSpecs, specs, runs * sizeof(sph_fraction_spec), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dOffsets, binOffset.data(), runs * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(dCounts, 0, (bins + 1) * sizeof(int32_t), ctx->stream));
    if (n > 0) {
        if (ctx->dims == 2)
            fractionKeys<2><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, dOffsets, dKey,
                                                                       dCounts, n);
        else
            fractionKeys<3><<<gridFor(n, 256), 256, 0, ctx->stream>>>(ctx->runOf, ctx->state, dSpecs, dOffsets, dKey,
                                                                       dCounts, n);
    }
    scanTiles<<<(unsigned)tiles, kScanThreads, 0, ctx->stream>>>(dCounts, dStart, dTiles, bins);
    scanPartials<<<1, kScanThreads, 0, ctx->stream>>>(dTiles, tiles, dStart, bins);
    scanAdd<<<(unsigned)tiles, kScanThreads, 0, ctx->stream>>>(dStart, dTiles, bins);
    if (n > 0) fractionScatter<<<gridFor(n, 256), 256, 0, ctx->stream>>>(dKey, dStart, dCounts, dMembers, n);
    if (bins > 0) fractionOrder<<<gridFor(bins, 256), 256, 0, ctx->stream>>>(dStart, dMembers, bins);
    for (int32_t r = 0; r < runs; ++r) {
        const sph_fraction_spec& spec = specs[r];
        if (spec.nx <= 0) continue;
        const unsigned grid = (unsigned)(((spec.nx + kFractionTile - 1) / kFractionTile) *
                                         ((spec.ny + kFractionTile - 1) / kFractionTile));
        const size_t lds = fractionLds(spec.reach);
        if (ctx->dims == 2) {
            if (ctx->precision == 0)
                fractionNodes<2, float><<<grid, 256, lds, ctx->stream>>>(ctx->state, dSpecs, r, binOffset[r], dStart,
                                                                          dMembers, dFraction);
            else
                fractionNodes<2, double><<<grid, 256, lds, ctx->stream>>>(ctx->state, dSpecs, r, binOffset[r], dStart,
                                                                           dMembers, dFraction);
        } else {
            if (ctx->precision == 0)
                fractionNodes<3, float><<<grid, 256, lds, ctx->stream>>>(ctx->state, dSpecs, r, binOffset[r], dStart,
                                                                          dMembers, dFraction);
            else
                fractionNodes<3, double><<<grid, 256, lds, ctx->stream>>>(ctx->state, dSpecs, r, binOffset[r], dStart,
                                                                           dMembers, dFraction);
        }
    }
    status = launchCheck(ctx, "fluid fraction");
    if (status != SPH_OK) return status;
    // Only the runs' own node ranges are defined; the host copies those
    std::vector<double> host(nodes);
    HIP_TRY(hipMemcpyAsync(host.data(), dFraction, nodes * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int32_t r = 0; r < runs; ++r) {
        const sph_fraction_spec& spec = specs[r];
        if (spec.nx <= 0) continue;
        std::memcpy(fraction + spec.offset, host.data() + spec.offset, (size_t)spec.nx * spec.ny * sizeof(double));
    }
    return SPH_OK;
}

int sph_integrate(sph_context* ctx, double timeStep) {
    int status = requireBound(ctx);
    if (status != SPH_OK) return status;
    if (!(timeStep > 0.0)) return fail(ctx, SPH_ERR_VALUE, "The time step must be positive.");
    return dispatchPhase(ctx, 2, timeStep, 0);
}

int sph_synchronize(sph_context* ctx) {
    if (!ctx) return SPH_ERR_VALUE;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

int sph_profile(sph_context* ctx, int32_t enable) {
    if (!ctx) return SPH_ERR_VALUE;
    int status = collectTimes(ctx);
    if (status != SPH_OK) return status;
    ctx->profiling = enable != 0;
    for (int k = 0; k < 16; ++k) {
        ctx->kernelMs[k] = 0.0;
        ctx->kernelLaunches[k] = 0;
    }
    return SPH_OK;
}

int sph_kernel_times(sph_context* ctx, double* milliseconds, int64_t* launches, int32_t count) {
    if (!ctx) return SPH_ERR_VALUE;
    int status = collectTimes(ctx);
    if (status != SPH_OK) return status;
    for (int32_t k = 0; k < count && k < kKernelKinds; ++k) {
        if (milliseconds) milliseconds[k] = ctx->kernelMs[k];
        if (launches) launches[k] = ctx->kernelLaunches[k];
    }
    return SPH_OK;
}

int sph_set_launch(sph_context* ctx, int32_t blockSize, int32_t lanes, int32_t listCapacity) {
    if (!ctx) return SPH_ERR_VALUE;
    if (listCapacity < -1 || listCapacity > 4096)
        return fail(ctx, SPH_ERR_VALUE, "The neighbor-list capacity must be from 0 to 4096 (-1 picks by dimension).");
    if (blockSize < 64 || blockSize > 256 || blockSize % 64 != 0)
        return fail(ctx, SPH_ERR_VALUE, "The workgroup size must be 64, 128, 192 or 256.");
    if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16)
        return fail(ctx, SPH_ERR_VALUE, "Lanes per particle must be 1, 2, 4, 8 or 16 (0 picks by dimension).");
    if (blockSize != ctx->block || lanes != ctx->lanes) {
        ctx->block = blockSize;
        ctx->lanes = lanes;
        dropGraph(ctx);
    }
    if (listCapacity != ctx->listCapacity) {
        ctx->listCapacity = listCapacity;
        dropGraph(ctx);
        if (ctx->dRuns) {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            return allocateList(ctx);
        }
    }
    return SPH_OK;
}

int sph_set_near_cells(sph_context* ctx, int32_t bytes, int32_t buildBytes) {
    if (!ctx) return SPH_ERR_VALUE;
    if (bytes < -1 || bytes > 98304 || buildBytes < -1 || buildBytes > 98304)
        return fail(ctx, SPH_ERR_VALUE, "The LDS copy's cap must be from 0 to 98304 bytes (-1 picks by dimension).");
    if (bytes != ctx->nearBytes || buildBytes != ctx->nearBytesBuild) {
        ctx->nearBytes = bytes;
        ctx->nearBytesBuild = buildBytes;
        dropGraph(ctx);
    }
    return SPH_OK;
}

}  // extern "C"

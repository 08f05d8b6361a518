// sph_state_sums: sph_global_sums without a context, for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops that forbid hipMalloc). Every device buffer
// it needs (the run table, the chunks' partials and the per-run sums) lives in `scratch`, one
// file-scope array of kScratchBytes per GPU, which the process holds until it ends.
//
// Each GPU has its own copy of `scratch` at its own address. The host never passes the array's
// name to a kernel or a copy: the first call on a device, with that device current, looks the
// address up once with hipGetSymbolAddress and keeps it per device, and every later use goes
// through that pointer.
//
// Calls share their GPU's scratch: two calls that overlap on one GPU (on different streams, or
// from different host threads) race, and the scratch keeps the last call's contents. Everything a
// call reads from scratch it has written first on its own stream, so nothing needs clearing.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "sph_device.hpp"
#include "sph_sums.hpp"

using namespace sph;

namespace {

// One run as the sums read it, with its first chunk among all the runs' chunks.
struct StaticRun {
    int64_t first, count;
    int64_t chunk0;         // first chunk; chunks of run r are [chunk0, chunk0 of run r + 1)
    SumPhysics physics;
};

constexpr int32_t kStaticRuns = 1024;           // runs per call at most
constexpr int64_t kStaticChunks = 16384;        // chunks of kSumChunk slots per call at most (134 M slots)
constexpr size_t kRunBytes = sizeof(StaticRun) * (kStaticRuns + 1);
constexpr size_t kPartialBytes = sizeof(double) * kSums * kStaticChunks;
constexpr size_t kOutBytes = sizeof(double) * kSums * kStaticRuns;
constexpr size_t kAlign = 256;
constexpr size_t roundUp(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kRunOffset = 0;
constexpr size_t kPartialOffset = roundUp(kRunOffset + kRunBytes);
constexpr size_t kOutOffset = roundUp(kPartialOffset + kPartialBytes);
constexpr size_t kScratchBytes = roundUp(kOutOffset + kOutBytes);   // about 1.1 MB

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratchLock;
void* scratchAddress[kMaxDevices] = {};   // scratch's device address, per device, once looked up

thread_local std::string staticError;

int failStatic(int code, const std::string& message) {
    staticError = message;
    return code;
}

// The current device's scratch address, looked up on that device's first call.
int scratchOf(void** out) {
    int device = 0;
    hipError_t status = hipGetDevice(&device);
    if (status != hipSuccess) return failStatic(SPH_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(status));
    if (device < 0 || device >= kMaxDevices) return failStatic(SPH_ERR_VALUE, "The device index is out of range.");
    std::lock_guard<std::mutex> guard(scratchLock);
    if (!scratchAddress[device]) {
        void* address = nullptr;
        status = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
        if (status != hipSuccess || !address)
            return failStatic(SPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(status));
        scratchAddress[device] = address;
    }
    *out = scratchAddress[device];
    return SPH_OK;
}

// The run of chunk b: the last run whose first chunk is at most b (runs with no slots own no chunk).
__device__ __forceinline__ int32_t runOfChunk(const StaticRun* __restrict__ runs, int32_t count, int64_t b) {
    int32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (runs[mid].chunk0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int D, typename Real>
__global__ void __launch_bounds__(kSumThreads)
    staticChunks(const StaticRun* __restrict__ runs, int32_t count, sph_state state, double* __restrict__ partial) {
    __shared__ double tree[kSums][kSumThreads];
    const int64_t b = blockIdx.x;
    const StaticRun& run = runs[runOfChunk(runs, count, b)];
    const int64_t first = run.first + (b - run.chunk0) * kSumChunk;
    const int64_t end = min(first + kSumChunk, run.first + run.count);
    chunkSums<D, Real>(run.physics, first, end, state, tree, partial + b * kSums);
}

__global__ void staticRuns(const StaticRun* __restrict__ runs, int32_t count, const double* __restrict__ partial,
                           double* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    addChunks(partial, runs[r].chunk0, runs[r + 1].chunk0, out + (int64_t)r * kSums);
}

}  // namespace

extern "C" {

int sph_state_sums(int32_t dims, int32_t precision, const sph_run* runs, int32_t count, const sph_state* state,
                   double* sums, void* stream) {
    if (dims != 2 && dims != 3) return failStatic(SPH_ERR_VALUE, "The sums need dims 2 or 3.");
    if (precision != 0 && precision != 1) return failStatic(SPH_ERR_VALUE, "The precision is 0 (float32) or 1 (float64).");
    if (!runs || !state || !sums) return failStatic(SPH_ERR_VALUE, "The sums need runs, a state and an output array.");
    if (count < 1 || count > kStaticRuns)
        return failStatic(SPH_ERR_VALUE, "The sums without a context take 1 to " + std::to_string(kStaticRuns) +
                                             " runs; " + std::to_string(count) + " were given.");
    if (!state->position || !state->velocity || !state->density || !state->mass || !state->fluid)
        return failStatic(SPH_ERR_VALUE, "The state needs positions, velocities, densities, masses and fluid flags.");
    // the run table on the host stack: no allocation anywhere in the call
    StaticRun table[kStaticRuns + 1];
    int64_t chunks = 0;
    for (int32_t r = 0; r < count; ++r) {
        const sph_run& run = runs[r];
        if (run.first < 0 || run.count < 0 || run.first + run.count > state->stride)
            return failStatic(SPH_ERR_VALUE, "Run " + std::to_string(r) + "'s slots lie outside the state's stride.");
        table[r].first = run.first;
        table[r].count = run.count;
        table[r].chunk0 = chunks;
        table[r].physics = SumPhysics{run.rho0, run.c0, -run.gravity[dims - 1], run.negativePressure};
        chunks += (run.count + kSumChunk - 1) / kSumChunk;
        if (chunks > kStaticChunks)
            return failStatic(SPH_ERR_VALUE, "The runs need more than " + std::to_string(kStaticChunks) + " chunks of " +
                                                 std::to_string(kSumChunk) + " slots, the most the sums without a context hold.");
    }
    table[count] = StaticRun{0, 0, chunks, SumPhysics{}};
    void* base = nullptr;
    int status = scratchOf(&base);
    if (status != SPH_OK) return status;
    unsigned char* bytes = static_cast<unsigned char*>(base);
    StaticRun* dRuns = reinterpret_cast<StaticRun*>(bytes + kRunOffset);
    double* partial = reinterpret_cast<double*>(bytes + kPartialOffset);
    double* out = reinterpret_cast<double*>(bytes + kOutOffset);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const sph_state st = *state;
    hipError_t hs = hipMemcpyAsync(dRuns, table, sizeof(StaticRun) * (count + 1), hipMemcpyHostToDevice, s);
    if (hs != hipSuccess) return failStatic(SPH_ERR_HIP, std::string("run table copy: ") + hipGetErrorString(hs));
    if (chunks > 0) {
        const unsigned grid = (unsigned)chunks;
        if (dims == 2) {
            if (precision == 0) staticChunks<2, float><<<grid, kSumThreads, 0, s>>>(dRuns, count, st, partial);
            else staticChunks<2, double><<<grid, kSumThreads, 0, s>>>(dRuns, count, st, partial);
        } else {
            if (precision == 0) staticChunks<3, float><<<grid, kSumThreads, 0, s>>>(dRuns, count, st, partial);
            else staticChunks<3, double><<<grid, kSumThreads, 0, s>>>(dRuns, count, st, partial);
        }
    }
    staticRuns<<<(unsigned)((count + 63) / 64), 64, 0, s>>>(dRuns, count, partial, out);
    hs = hipGetLastError();
    if (hs != hipSuccess) return failStatic(SPH_ERR_HIP, std::string("state sums launch: ") + hipGetErrorString(hs));
    hs = hipMemcpyAsync(sums, out, sizeof(double) * kSums * count, hipMemcpyDeviceToHost, s);
    if (hs == hipSuccess) hs = hipStreamSynchronize(s);
    if (hs != hipSuccess) return failStatic(SPH_ERR_HIP, std::string("state sums: ") + hipGetErrorString(hs));
    return SPH_OK;
}

const char* sph_static_error(void) { return staticError.c_str(); }

}  // extern "C"

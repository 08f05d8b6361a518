/*
 * C ABI of the MI355X delta-SPH engine (libsph_hip.so).
 *
 * The engine replaces the device side of pySurf's SPH backend, sphGpu.PairSums
 * (src/pysurf/sphGpu.py:1492-2323), which WcsphSolver selects through SimulationConfig.backend
 * (src/pysurf/sphSolver.py:425, 1499-1518). Every entry point below names the PairSums method it
 * replaces. The Python mirror of PairSums (pysurf_amd/pair_sums.py) binds these through ctypes.
 *
 * Conventions:
 *  - Every pointer into particle data is a device pointer that the caller owns (PyTorch-ROCm
 *    tensors in the Python mirror). The library never allocates or frees caller memory; its own
 *    scratch (cell lists, sorted float32 records, per-run clocks) lives in the context.
 *  - Particle arrays are structure-of-arrays: a vector field is `dims` rows of `stride` elements.
 *  - "Real" arrays are float for precision 0 (float32 pair sums over float64 state, the
 *    reference's own split, sphGpu.py:28-38) and double for precision 1 (float64 throughout).
 *  - All work is stream-ordered on the context's stream; nothing synchronizes except
 *    sph_get_clock, sph_synchronize and the neighbor-list export.
 *  - Every entry point returns SPH_OK or an error status; sph_last_error gives the message.
 *    SPH_ERR_VALUE corresponds to the reference's ValueError, SPH_ERR_HIP to a HIP runtime error.
 */
#ifndef SPH_HIP_H
#define SPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_ERR_VALUE 1
#define SPH_ERR_HIP 2

typedef struct sph_context sph_context;

/* The boundary a run's device step enforces: PairSums.setBoundary (sphGpu.py:2121-2171). Table
 * pointers are device pointers the caller keeps alive while the run is configured. */
typedef struct {
    int32_t kind;          /* 0 none, 1 container, 2 bed, 3 height field */
    double clearance;      /* a quarter spacing [m] */
    double lower[3], upper[3];
    int32_t openTop;
    int32_t nBed;
    const double* bedX;
    const double* bedZ;
    int32_t sideWalls;
    double width;
    int32_t offshoreWall;
    double wallX, pistonSpeed, endLimit;
    int32_t fieldNx, fieldNy;
    const double* table;   /* fieldNy rows of fieldNx bed elevations */
    double originX, originY, stepX, stepY;
    double lowY, highY, lowX, highX;
} sph_boundary;

/* One run: its slots, grid and physics. Replaces the PairSums constructor and its set* calls
 * (sphGpu.py:1537-1570, setDiffusion :1879, setWallPressure :1906, setNegativePressure :1916,
 * setContinuityCorrection :1696, setViscosityLimiter :1734, setPhysics :2173). */
typedef struct {
    int64_t first;         /* first slot of the run's particles */
    int64_t count;         /* the run's particles, fluid and boundary */
    int32_t kernel;        /* 0 cubic spline, 1 Wendland C2 */
    int32_t grid[3];       /* cells along x, y, z (z = 1 in 2D): gridDimensions, sphGpu.py:1466-1478 */
    double cell;           /* cell width, 2 h_max [m] */
    double hmax, hmin;     /* longest and shortest smoothing length [m] */
    double period;         /* periodic y length [m], 0 for none (3D only) */
    double rho0, c0, alpha, boundaryFactor;
    double gravity[3];
    int32_t diffusion;     /* 0 none, 1 molteni, 2 hydrostatic, 3 antuono */
    double delta;
    int32_t extrapolatedWalls, negativePressure, correction, limiter;
    double densityFloor;   /* smallest fluid density [kg/m^3], 0 for none */
    double cfl, maxTimeStep;
    sph_boundary boundary;
    /* the Shepard filter of the fluid density every this many device-clocked steps, at a step count
       above zero that it divides (densityFilterInterval, sphSolver.py:1979-1981); 0 never */
    int32_t filterInterval;
} sph_run;

/* The state of every run, slot order. Replaces PairSums' device arrays (sphGpu.py:111-116). */
typedef struct {
    double* position;      /* dims rows */
    double* velocity;      /* dims rows */
    double* density;       /* 1 row */
    void* pressure;        /* Real, 1 row */
    void* acceleration;    /* Real, dims rows */
    void* mass;            /* Real; read by the steps, permuted by sph_sort_state */
    void* length;          /* Real, each particle's smoothing length; likewise */
    uint8_t* fluid;        /* 1 for fluid, 0 for boundary; likewise */
    int64_t stride;        /* elements between rows */
} sph_state;

/* A context for `dims` (2 or 3) and `precision` (0 float32 pairs, 1 float64) on `stream`
 * (a hipStream_t; null for the default stream). Replaces PairSums.__init__ (sphGpu.py:1537). */
int sph_create(sph_context** out, int32_t dims, int32_t precision, void* stream);
int sph_destroy(sph_context* ctx);
const char* sph_last_error(const sph_context* ctx);

/* The runs the context steps together, and the slots they cover: PairSums._reserve and the set*
 * calls (sphGpu.py:1580-1594). Allocates the context's scratch for `capacity` slots. */
int sph_configure(sph_context* ctx, const sph_run* runs, int32_t count, int64_t capacity);

/* Binds the state arrays: PairSums.push/update (sphGpu.py:1998-2042, 2207-2221). */
int sph_bind(sph_context* ctx, const sph_state* state);

/* The cell list over the bound positions: PairSums._rebuild (sphGpu.py:1635-1641). */
int sph_build_cells(sph_context* ctx);

/* Host-step sums over the cells of the last sph_build_cells, each reading caller arrays in slot
 * order and writing one result per slot. Every run of the context must share its viscosity limiter,
 * continuity correction and density diffusion (the reference's PairSums holds one run); otherwise
 * they return SPH_ERR_VALUE. */
int sph_wall_pressure(sph_context* ctx, const double* density, const void* pressure, void* pressureOut,
                      double* densityOut);                                        /* sphGpu.py:2090 */
int sph_velocity_gradients(sph_context* ctx, const double* velocity, const double* density,
                           void* out);                                            /* sphGpu.py:1827 */
int sph_correction_matrices(sph_context* ctx, const double* density, void* out);  /* sphGpu.py:1855 */
int sph_pair_accelerations(sph_context* ctx, const double* velocity, const double* density,
                           const void* pressure, void* out);                      /* sphGpu.py:2044 */
int sph_continuity_rate(sph_context* ctx, const double* velocity, const double* density,
                        void* out);                                               /* sphGpu.py:2061 */
/* The two options the delta-SPH scheme leaves off, as host-step sums: the Shepard-filtered density
 * of every slot, sum_j m_j W_ij / sum_j (m_j / rho_j) W_ij with the particle itself included, and
 * the XSPH transport velocity (D rows of `stride`), the fluid's blended toward its fluid neighbors'
 * by `epsilon` in [0, 1], the boundary's unchanged. */
int sph_shepard_density(sph_context* ctx, const double* density, void* out);     /* sphGpu.py:2104 */
/* The renormalized density gradient of every slot (D rows of `stride`), L_i sum_j V_j (rho_j -
 * rho_i) grad_i W_ij, zero on the boundary: the 'antuono' diffusion's (sph_run.diffusion 3), which
 * sph_continuity_rate and the device steps subtract. */
int sph_density_gradients(sph_context* ctx, const double* density, void* out);   /* sphGpu.py:1868 */
/* The delta-LES-SPH closure of the host-step sums (sphGpu.py:1754-1778): while on,
 * sph_pair_accelerations builds G and each slot's eddy alpha and delta first and adds the closure's
 * viscosity on every fluid pair, and sph_continuity_rate's diffusion takes each fluid pair's delta
 * from the deltas the last such sum built (zero before any). The device steps carry it too: each
 * step builds G, the coefficients and then the sums with the closure (sphGpu.py:2256-2262). */
int sph_set_closure(sph_context* ctx, int32_t enabled, double soundSpeed, double referenceDensity,
                    double kinematicViscosity);                                   /* sphGpu.py:1754 */
/* Particle shifting as a host step (sphGpu.py:1966-1979, runs of one smoothing length): each slot's
 * shift (D rows of `stride`), zero on the boundary and the free surface, the density change that
 * goes with it and the velocity change (D rows), all in the pair precision. `size` is the shift's
 * size, CFL Ma (2h)^2 or, with `local`, A h dt to be scaled by each |v_i|; `spacingKernel` is W at
 * the particle spacing. The device steps do not shift. */
int sph_shifting_displacement(sph_context* ctx, const double* velocity, const double* density, double size,
                              int32_t local, double spacingKernel, void* shifts, void* densityChanges,
                              void* velocityChanges);                             /* sphGpu.py:1966 */
/* Each slot's eddy alpha and delta at the bound positions, into two device arrays of the pair
 * precision: PairSums.closureCoefficients (sphGpu.py:1841). */
int sph_closure_coefficients(sph_context* ctx, const double* velocity, const double* density, void* alphas,
                             void* deltas);                                      /* sphGpu.py:1841 */
int sph_xsph_velocity(sph_context* ctx, const double* velocity, const double* density, double epsilon,
                      void* out);                                                 /* sphGpu.py:2080 */

/* Each slot's neighbor count, then (offsets from the caller's prefix sum of the counts) its
 * neighbors' slots in visit order; the predicate is the sums' own. Synchronizes. */
int sph_neighbor_counts(sph_context* ctx, int32_t* counts);
int sph_neighbor_rows(sph_context* ctx, const int64_t* offsets, int32_t* rows);

/* The bound boundary enforcement on the bound state: PairSums.enforceBoundary (sphGpu.py:2317). */
int sph_enforce(sph_context* ctx);

/* Each run's clock: WcsphSolver.advance's loop condition (sphSolver.py:1983). A run steps while
 * time < untilTime and steps < maxSteps and it has not failed (sph_run_failures); setting the clock
 * clears a failure. */
int sph_set_clock(sph_context* ctx, int32_t run, double time, double untilTime, int64_t steps, int64_t maxSteps);
/* Copies every run's time, last step size, steps taken and maxima (vmax, amax) to host arrays
 * (any may be null). Synchronizes. */
int sph_get_clock(sph_context* ctx, double* time, double* lastDt, int64_t* steps, double* maxima);
/* Each run's failure, one int32 per run into a host array: 0 none, 1 its state stopped being finite
 * (a speed, acceleration or position), 2 its time step fell below a hundredth of the acoustic step,
 * cfl hmin / (100 c0). The reference's run loops raise NumericalInstabilityError for both
 * (BreakingWave.py:892-919, BreakingWave3D.py:748-769); here a failed run's clock stops on the device
 * at the step that found it, the other runs of the context step on, and sph_set_clock clears it.
 * Synchronizes. */
int sph_run_failures(sph_context* ctx, int32_t* failures);

/* `steps` device-resident steps of every active run: PairSums.residentRates + residentIntegrate
 * (sphGpu.py:2237-2315) with the time step reduced on the device instead of cflTimeStep on the
 * host (sphSolver.py:1980-1986). Each run takes the Shepard filter at the steps its filterInterval
 * names (WcsphSolver.advance, sphSolver.py:1979-1981), its diffusion mode ('antuono' included) and
 * the context's closure; the energy ledger keeps its terms when on. With `useGraph` the steps are
 * captured once into a HIP graph and replayed (captured again when the options change).
 * Asynchronous. */
int sph_advance(sph_context* ctx, int32_t steps, int32_t useGraph);

/* The two halves of one step with the time step from the host, for step-for-step parity with the
 * reference's loop: sph_rates is PairSums.residentRates (sphGpu.py:2237-2273) and copies every
 * run's (|v|max, |a|max) to `maxima` (may be null; then it does not synchronize);
 * sph_integrate is PairSums.residentIntegrate(dt) (sphGpu.py:2275-2315). Both step every run
 * whatever its clock, as the host loop decides when to stop; sph_integrate advances the clocks. */
int sph_rates(sph_context* ctx, double* maxima);
/* sph_rates with PairSums.residentRates' filterDensity argument (sphGpu.py:2237-2254): nonzero
 * filters every run's fluid density by Shepard's sum first. sph_rates(ctx, m) is
 * sph_resident_rates(ctx, 0, m). */
int sph_resident_rates(sph_context* ctx, int32_t filterDensity, double* maxima);
int sph_integrate(sph_context* ctx, double timeStep);

/* Each run's float64 sums over its fluid particles of the bound state, 7 per run in run order:
 * mass, kinetic energy sum m |v|^2 / 2 and potential energy sum m g z (ParticleSystem.kineticEnergy
 * and potentialEnergy, sphParticles.py:236-255, g the downward gravity along the last axis), the
 * Tait elastic energy sum m e(rho) (sphEnergy.elasticEnergy, sphEnergy.py:120-160; the device form
 * sphGpu.py:271-282) and momentum sum m v along x, y, z (z zero in 2D). The summation order is fixed,
 * so the sums repeat bit for bit. Synchronizes. */
int sph_global_sums(sph_context* ctx, double* sums);

/* The energy budget of the device steps (PairSums.setEnergyBudget, sphGpu.py:1780-1793): on, each
 * fluid particle keeps a float64 running total of every term of sphEnergy.deviceTerms
 * (sphEnergy.py:92-98; 14 terms: the pressure and viscosity work from fluid and from boundary
 * neighbors, gravity's work, the plain continuity over fluid and over boundary neighbors, the
 * correction's share, the diffusion, the floor, the drift, the enforcement, shifting (zero: no
 * device step shifts) and the filter), cleared to zero by every call that switches it on. Needs a
 * bound state. Synchronizes. */
int sph_set_energy_budget(sph_context* ctx, int32_t enabled);
/* Each run's total of every term since the budget was switched on [J] (per unit depth in 2D),
 * 14 per run in run order, summed in a fixed order (PairSums.energySums, sphGpu.py:1805-1825).
 * SPH_ERR_VALUE while the budget is off. Synchronizes. */
int sph_energy_sums(sph_context* ctx, double* sums);

/* Permutes the bound state, every field, into the cell order of the current positions, each run
 * within its own slots, carrying `identity` (one int32 per slot; may be null) along:
 * WcsphSolver.sortParticles (sphSolver.py:2093-2126) on the device, with the boundary sorted too.
 * Gathers through the sorted order then read neighboring slots. Asynchronous; must not be called
 * during graph capture (it may allocate its copy buffer on first use). */
int sph_sort_state(sph_context* ctx, int32_t* identity);

/* The breaking read-out's measures, taken on the device from the bound state. A run's tracked
 * particles are its fluid particles (in 3D those with |y - stripCentre| < stripHalfWidth), read as
 * (x, z). They replace the host arrays the reference's BreakingTracker.update reads every
 * 0.01 sqrt(d/g) (waveBreaking.py:810-906; the cells face measure :559-655): no reference entry
 * point computes them. Each call reads the runs with `columns` > 0 (`active` for the face cells)
 * and synchronizes. */
typedef struct {
    double x0, width;                     /* columns [x0 + c width, x0 + (c + 1) width) */
    int32_t columns;                      /* 0 skips the run */
    int32_t offset;                       /* the run's first column in `tops` */
    double stripCentre, stripHalfWidth;   /* 3D only */
} sph_profile_spec;

/* tops[offset + c]: the highest z of the run's tracked particles in column c, -inf where none
 * (freeSurfaceProfile before its half-column lift, waveBreaking.py:129-162). */
int sph_surface_profile(sph_context* ctx, const sph_profile_spec* specs, double* tops);

typedef struct {
    double windowStart, windowEnd, base, spacing;   /* the face window (waveBreaking.py:559-634) */
    int32_t columns, bands;               /* the occupancy kept: bands x columns cells (0: landward points only) */
    int32_t offset;                       /* the run's first 32-bit word of `occupancy` */
    int32_t active;                       /* 0 skips the run */
    double wetHeight;                     /* the waterline's reach above the bed (:636-655) */
    const double* bedX;                   /* the tracker's bed, bedPoints device values each */
    const double* bedZ;
    int32_t bedPoints;
    double stripCentre, stripHalfWidth;   /* 3D only */
} sph_face_spec;

/* For each active run: `occupancy`, bands rows of ceil(columns / 32) words from `offset`, bit c of
 * row b set where frontFaceLean marks cell (band b, column c); overflow[r] = 1 where a tracked
 * particle marks a band at or above `bands`; landward[r * 16 + 2 m + (0, 1)], the (x, z) of the
 * 8 most landward tracked particles within wetHeight (widened by 1e-9 of a metre) of the bed's
 * linear interpolation, by descending x, NaN past the last. */
int sph_face_cells(sph_context* ctx, const sph_face_spec* specs, uint32_t* occupancy, int32_t* overflow,
                   double* landward);

typedef struct {
    double x0, y0;                        /* node (0, 0) */
    double stepX, stepY;                  /* node (i, j) stands at (x0 + i stepX, y0 + j stepY) */
    double binWidth;                      /* the bins particles are sorted into: half a spacing */
    int32_t nx, ny;                       /* nodes; nx = 0 skips the run */
    int32_t reach;                        /* ceil(2 h / binWidth) */
    int32_t pad;
    int64_t offset;                       /* the run's first node in `fraction`, ny rows of nx */
    double h;                             /* the cubic spline's smoothing length */
    double area;                          /* each particle's area where areaPerMass is 0 */
    double areaPerMass;                   /* > 0: a particle's area is its mass times this */
    double lowX, lowZ, highX, highZ;      /* the particles taken: lowX < x < highX and lowZ < z < highZ */
    double stripCentre, stripHalfWidth;   /* 3D only */
} sph_fraction_spec;

/* fraction[offset + j nx + i]: the kernel-smoothed fluid fraction sum_p A_p W(|node - p|, h) of the
 * run's tracked particles at node (i, j), with the 2D cubic spline (fluidFraction,
 * waveBreaking.py:206-262). Each node sums the particles bin by bin, rows of bins outer, and in slot
 * order within a bin, so the result does not depend on the run's neighbours or the launch. */
int sph_fluid_fraction(sph_context* ctx, const sph_fraction_spec* specs, double* fraction);

/* Splitting fluid particles into 2^dims daughters on the device (splitParticles,
 * sphRefinement.py:135-186): a square (cube) s/2 across around each mother, s = (m / rho0)^(1/dims);
 * each daughter takes m / 2^dims, the mother's length times lengthRatio and her velocity, density,
 * pressure and acceleration; the first takes the mother's slot, the others follow the run's own
 * particles, mother by mother in slot order. Two calls, with the caller's new arrays between them:
 *
 * sph_split_count marks the particles of each run with `active` that are fluid, lowX <= x < highX,
 * of mass >= minMass and, when `mask` (a device array of one byte per slot) is given, set there;
 * chosen[r] receives run r's count. Synchronizes.
 *
 * sph_split_apply writes the split state into `into` (its stride at least the new total): run r
 * starts at the sum of the runs before it, each grown by chosen (2^dims - 1); the runs must lie
 * packed from slot 0, and the context's state stays bound and unchanged. The caller then
 * configures the grown runs (sph_configure) and binds `into` (sph_bind), and sets the runs'
 * clocks again. No other call may come between the two. */
typedef struct {
    int32_t active;
    int32_t pad;
    double lowX, highX;
    double minMass;
} sph_split_spec;

int sph_split_count(sph_context* ctx, const sph_split_spec* specs, const uint8_t* mask, int64_t* chosen);
int sph_split_apply(sph_context* ctx, const sph_state* into, double referenceDensity, double lengthRatio);

/* Merging split fluid particles on the device (mergeParticles, sphRefinement.py:212-311), in the
 * bound state. In each run with `active`, the fluid particles with lowX <= x < highX (and set in
 * `mask`, one byte per slot, when given) that have split, level = rint(log2(referenceMass / m) /
 * dims) > 0, merge 2^dims of one level at a time: level by level, a pass with the grid unshifted
 * and one shifted by half a cell groups the level's free particles in cells the size of a particle
 * one level up, (referenceMass / 2^(dims (level - 1)) / referenceDensity)^(1/dims) across; in each
 * cell holding 2^dims or more, the 2^dims nearest its centre (ties by slot) form a group. A group
 * becomes one particle in its first member's slot: mass-weighted position, velocity, acceleration
 * and pressure, the total mass, the total mass over the members' volumes as density, and the
 * members' mean length over lengthRatio. The others leave; the holes they leave at or before a run's
 * last boundary slot take, in order, the run's last as many remaining particles after it, and the
 * runs close up, packed from slot 0 (which they must be). groups[r] and removedEnergy[r] (the
 * kinetic energy of the members' motion about their groups' centres of mass) receive each run's.
 * The caller then configures the runs with counts less groups (2^dims - 1), binds the same state
 * and sets the clocks again. Synchronizes. */
typedef struct {
    int32_t active;
    int32_t pad;
    double lowX, highX;
    double referenceMass;      /* the mass of a fluid particle that has not split */
    double referenceDensity;
    double lengthRatio;        /* the ratio the daughters' lengths took at their split */
} sph_merge_spec;

int sph_merge(sph_context* ctx, const sph_merge_spec* specs, const uint8_t* mask, int64_t* groups, double* removedEnergy);

int sph_synchronize(sph_context* ctx);

/* sph_global_sums without a context, for callers that cannot allocate device memory: code under
 * HIP graph capture, serving loops and harnesses that forbid hipMalloc. No counterpart in the
 * reference (PairSums allocates through Warp). The same 7 float64 sums per run, in the same order
 * and bit for bit the same, over `state` (device pointers, slot order) for the `count` runs of
 * `runs` (only first, count, rho0, c0, gravity and negativePressure are read), on `stream`
 * (a hipStream_t; null for the default stream); `sums` is a host array of 7 * count doubles.
 * It allocates nothing: its device buffers live in one file-scope array of about 1.1 MB per GPU,
 * held until the process ends. At most 1024 runs and 16384 chunks of 8192 slots (each run's slots
 * rounded up to whole chunks); larger calls return SPH_ERR_VALUE. Calls on one GPU share that
 * array, so calls that overlap on one GPU race. Synchronizes the stream. Errors are read with
 * sph_static_error, thread by thread. */
int sph_state_sums(int32_t dims, int32_t precision, const sph_run* runs, int32_t count, const sph_state* state,
                   double* sums, void* stream);
const char* sph_static_error(void);

/* Per-kernel timing with HIP events recorded on the context's stream around every launch of a
 * step (no counterpart in the reference, which times whole runs only). sph_profile(1) clears the
 * totals and starts recording (steps then run as stream launches, not graphs); sph_kernel_times
 * synchronizes and copies, per kernel id, the accumulated milliseconds and launch counts:
 * 0 clock, 1 cell keys, 2 cell-count scan, 3 cell scatter, 4 cell order (the rank within cells), 5 record gather,
 * 6 sweep 1 (wall pressure, G), 7 sweep 2 (accelerations, L), 8 unused (the kick is part of
 * sweep 3), 9 sweep 3 (kick, continuity, density advance, drift, enforcement), 10 neighbor list
 * (with sweep 1), 11 enforcement alone, 12 the runs' maxima. */
int sph_profile(sph_context* ctx, int32_t enable);
int sph_kernel_times(sph_context* ctx, double* milliseconds, int64_t* launches, int32_t count);

/* Launch shape of the pair sweeps: threads per workgroup (64, 128, 192 or 256), lanes per
 * particle (1, 2, 4, 8 or 16; 0 picks 8 in 2D, 16 in 3D), the lanes of a particle sharing its
 * neighbors, and the neighbor list's entries per particle (0 walks the cells in every sweep; -1
 * picks 64 in 2D, 160 in 3D; a particle with more neighbors walks its cells). The summation order
 * depends on the lanes, so runs compared bit for bit are stepped with the same lanes. */
int sph_set_launch(sph_context* ctx, int32_t blockSize, int32_t lanes, int32_t listCapacity);

/* The cap of the records a pair-sweep workgroup copies into LDS around its particles, in bytes
 * (0 to 98304; 0 reads every neighbor from HBM; -1 picks 32 KiB), and the same for the device
 * step's list build (-1 takes `bytes`). No counterpart in the reference: launch-shape parameters for
 * the autotuner (pysurf_amd/autotune.py). A neighbor outside the copy is read from HBM with the same
 * value, so results do not depend on them. */
int sph_set_near_cells(sph_context* ctx, int32_t bytes, int32_t buildBytes);

#ifdef __cplusplus
}
#endif

#endif /* SPH_HIP_H */

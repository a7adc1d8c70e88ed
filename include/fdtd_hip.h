/* fdtd_hip.h - C ABI of the MI355X (gfx950) per-timestep Yee update.
 *
 * Drop-in boundary.  One handle replaces, for one simulation, the calls the
 * reference's serial step body makes on host vectors
 * (3D_FDTD_version008 src/3D_FDTD_version008.cpp:269-349):
 *
 *   reference call (file:line)                                   replaced by
 *   ------------------------------------------------------------ -----------------------
 *   std::fill(J..)+ISource::inject_half_step  src/3D_FDTD_version008.cpp:270-275,
 *                      src/sources/isource.hpp:28-31                fdtd_set_sources + fdtd_step values
 *   update_polarization_multi_region      src/two_level_system.hpp:574-609   fdtd_step (fused, E kernel)
 *   fdtd_update_H                         src/fdtd_stepper.hpp:8-47          fdtd_step (H kernel)
 *   IBoundary::apply_after_H (CPML)       src/boundary.hpp:150-225           fdtd_step (fused, H kernel)
 *   fdtd_update_E / _with_gain            src/fdtd_stepper.hpp:49-89,
 *                                         src/two_level_system.hpp:299-342   fdtd_step (E kernel)
 *   IBoundary::apply_after_E (CPML / PEC) src/boundary.hpp:227-300, 51-79    fdtd_step (fused / face kernel)
 *   PointFieldDetector::record_after_E    src/detectors/point_field_detector.hpp:111-124
 *                                                                            fdtd_add_probe + fdtd_read_probe
 *   update_populations_multi_region       src/two_level_system.hpp:612-666   fdtd_step (fused, E kernel)
 *   StabilityMonitor::check_stability     src/stability_monitor.hpp:42-144   fdtd_step monitor + fdtd_read_check
 *
 * Conventions: all host arrays are the reference's own layout (FP64, idx3
 * order (i*NyT + j)*NzT + k, src/global_function.hpp:21-24, unpadded).  The
 * library copies them to device memory it owns; the caller keeps ownership of
 * its buffers.  Every entry point returns 0 on success or a negative FDTD_E*
 * code; fdtd_last_error() returns the message of the last failure on the
 * calling thread.  One handle binds one HIP stream and is not thread-safe;
 * distinct handles may be driven from distinct host threads.
 */
#ifndef FDTD_HIP_H
#define FDTD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FDTD_OK = 0,
    FDTD_EINVAL = -1,     /* bad argument / unsupported configuration */
    FDTD_EDEVICE = -2,    /* HIP runtime error or no device */
    FDTD_ENOMEM = -3,     /* device allocation failed */
    FDTD_EUNSTABLE = -4,  /* stability monitor verdict: simulation unstable */
};

/* Storage/compute precision of fields and CPML state.
 * FP64: FP64 storage and arithmetic (parity path).
 * FP32: FP32 storage and arithmetic; populations and polarization in FP64.
 * FP16: FP16 storage with per-component power-of-two scales, FP32 arithmetic,
 *       FP64 populations, polarization and DFT accumulators. */
enum { FDTD_FP64 = 0, FDTD_FP32 = 1, FDTD_FP16 = 2 };

typedef struct fdtd_sim* fdtd_handle;

/* GridSpacing (src/global_function.hpp:92-105) and dt (src/params.hpp:265-266). */
typedef struct {
    int nx, ny, nz;                          /* NxT, NyT, NzT: field dimensions */
    const double *inv_dx, *inv_dy, *inv_dz;  /* length >= nx, ny, nz */
    double dt;
} fdtd_grid_desc;

/* One structure of StructureScene::items (src/structure_material.hpp:
 * 96-100, shapes :44-93) for the device bake: shape type (0 box, 1 sphere,
 * 2 z-cylinder, as Box / Sphere / CylinderZ), its parameters in total-grid
 * physical coordinates (box: x0 x1 y0 y1 z0 z1 with half-open extents;
 * sphere: cx cy cz r; cylinder: cx cy r z0 z1) and its Material
 * (:17-22). */
typedef struct {
    int type;
    double p[6];
    double eps_r, mu_r, sigma, sigma_m;
} fdtd_bake_shape;

/* Inputs of StructureScene::bake (src/structure_material.hpp:211-288) run on
 * the device: GridSpacing bounds (n+1 values per axis) and spacings, the
 * background Material, the structures in insertion order (later ones
 * override earlier ones), subpixel_samples (rounded to s^3 samples per Yee
 * component, as the reference does), dt and the eps0 / mu0 the host uses. */
typedef struct {
    const double *x_bounds, *y_bounds, *z_bounds;
    const double *dx, *dy, *dz;
    double bg_eps_r, bg_mu_r, bg_sigma, bg_sigma_m;
    int n_shapes;
    const fdtd_bake_shape* shapes;
    int subpixel_samples;
    double dt, eps0, mu0;
} fdtd_bake_desc;

/* MaterialGrids, src/global_function.hpp:70-90 (as filled by
 * StructureScene::bake, src/structure_material.hpp:211-288): 12 arrays of
 * nx*ny*nz values in idx3 order.  Alternatively, with the 12 arrays NULL and
 * `bake` set, fdtd_create runs the bake itself on the device (every value
 * bitwise the host bake's) and no N-sized material array exists on the host:
 * cells whose subpixel samples all fall in one material take that
 * material's coefficients; the others are computed per cell. */
typedef struct {
    const double *aEx, *bEx, *aEy, *bEy, *aEz, *bEz;
    const double *aHx, *bHx, *aHy, *bHy, *aHz, *bHz;
    const fdtd_bake_desc* bake;
} fdtd_material_desc;

/* Boundary, src/boundary.hpp:16-136.  type FDTD_BC_PEC (0) or FDTD_BC_CPML (1).
 * For CPML the 1D profiles of CPMLBoundary (kx_/ky_/kz_, bE*_/cE*_,
 * bH*_/cH*_, :87-91, filled by build_profiles_and_coeffs :303-373),
 * lengths nx / ny / nz. */
typedef struct {
    int type;
    int npml;
    const double *kx, *ky, *kz;
    const double *bEx, *cEx, *bEy, *cEy, *bEz, *cEz;
    const double *bHx, *cHx, *bHy, *cHy, *bHz, *cHz;
} fdtd_boundary_desc;

/* Per-region TwoLevelParams after finalize / finalize_with_dt
 * (src/two_level_system.hpp:17-110). */
typedef struct {
    double kapa, pa2, pa3, inv_dt;   /* kapa_coeff, pa2, pa3, inv_dt */
    double tau, hbar, omega_a;       /* for inv_tau = 1/tau, inv_hbar_omega = 1/(hbar*omega_a) */
    double N0;                       /* N0_total: Ndip = Ng0 = N0 on every cell of the region */
    double dt;
    int clamp;                       /* enable_population_clamp */
} fdtd_tls_region;

/* TwoLevelState + TLSRegionManager::cell_region_map
 * (src/two_level_system.hpp:112-197, 488-571).  All arrays nx*ny*nz, idx3. */
typedef struct {
    int n_regions;
    const fdtd_tls_region* regions;
    const int* cell_region_map;           /* -1 = no TLS */
    const double *Ndip, *Ng0;             /* validated against regions[].N0 */
    const double *Nu, *Ng, *Pz, *Pz_prev; /* initial state */
} fdtd_tls_desc;

/* A current-source footprint: component (0 Jx, 1 Jy, 2 Jz) and an inclusive
 * index box.  DipoleSource is a 1-cell box (src/sources/dipole_source.hpp:60-85),
 * PlaneWaveSource a plane rectangle (src/sources/plane_wave_source.hpp:102-179).
 * Per-step values are supplied to fdtd_step in source order. */
typedef struct {
    int component;
    int i0, i1, j0, j1, k0, k1;
} fdtd_source_desc;

/* Stability reduction at one monitor step (StabilityMonitor inputs). */
typedef struct {
    long step;
    double max_E, max_H, max_P;
    int nan_E, nan_H, nan_P, nan_N;
    double pop_total, pop_initial;  /* sums over Ndip > 1e10 cells */
} fdtd_check_record;

/* ---- lifecycle ---- */
int fdtd_create(const fdtd_grid_desc* grid, const fdtd_material_desc* mat, const fdtd_boundary_desc* bc,
                const fdtd_tls_desc* tls /* NULL = no gain */, int precision, fdtd_handle* out);
int fdtd_destroy(fdtd_handle h);
const char* fdtd_last_error(void);
/* Device bytes held by the handle (admission control). */
int fdtd_device_bytes(fdtd_handle h, size_t* bytes);
/* MaterialGrids array q (0..11: aEx bEx aEy bEy aEz bEz aHx bHx aHy bHy aHz
 * bHz) as the handle holds it, expanded from the per-cell material ids and
 * the exact FP64 coefficient table into nx*ny*nz values (idx3 order).  The
 * counterpart of reading mg.<array> after bake_and_bind
 * (src/params.hpp:512-516). */
int fdtd_download_material(fdtd_handle h, int q, double* out);

/* ---- sources, probes, monitors ---- */
int fdtd_set_sources(fdtd_handle h, int n, const fdtd_source_desc* src);
/* Returns the probe id (>= 0) or a negative error.  component 0..11 as
 * Detectors::FieldComponent (src/detectors/idetector.hpp:19-29).  Values are
 * recorded after the E step whenever n % save_every == 0 (n = global step),
 * up to `capacity` records. */
int fdtd_add_probe(fdtd_handle h, int i, int j, int k, int component, int save_every, long capacity);
/* Running DFT over a half-open index box (no reference counterpart).  Returns
 * monitor id.  X(f) = sum_n F(t_n) exp(-i 2 pi f t_n) dt, t_n = (n+1) dt for E,
 * (n+1/2) dt for H components. */
int fdtd_add_dft(fdtd_handle h, int component, const int box[6], int n_freqs, const double* freqs);
/* FP16 storage scales (value = stored * scale), 6 fields Ex..Hz then the 12
 * psi arrays (Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y Hx_y Hx_z Hy_z Hy_x Hz_x Hz_y);
 * powers of two.  They start from an estimate (E 1, H 2^-8, psi from the
 * finest spacing; fdtd_set_fp16_scales replaces it) and adapt on the device
 * every 10 steps so the largest stored magnitude stays within [2^2, 2^11],
 * never dropping below 2^-6 of the value last set.  Stores saturate at the
 * FP16 maximum.  No-ops for FP64/FP32 handles.  No counterpart in the
 * reference (FP64 only). */
int fdtd_set_fp16_scales(fdtd_handle h, const double* scales18);
int fdtd_get_fp16_scales(fdtd_handle h, double* scales18);
/* Kernel schedule: mode 0 = separate H and E kernels (in place, default),
 * `param` = the k-width of the cells one wavefront updates (16, 32 or 64
 * consecutive k of 64/param rows; 32 by default; 0 keeps the current
 * value), for both kernels, or
 * width_H + 256 * width_E to set them apart (plus 131072: 64-bit element
 * offsets even where 32-bit byte offsets, the default with FP16 storage
 * when every array is below 4 GiB, would do; plus 1048576: the E kernel
 * walks its tiles in the H kernel's order instead of backwards, which by
 * default lets each kernel start on the tiles the previous one finished;
 * plus 262144: the stability monitor reads the six fields in a pass of its
 * own instead of taking their maxima in the next step's H launch; plus
 * N << 21: bands of N j-tiles in the tile order);
 * mode 1 = one fused step kernel: a work-group owns a block of rows over
 * the whole k extent (plus one halo row whose H it recomputes) and marches
 * along x, reading E^n and H^{n-1/2} once and writing E^{n+1} and
 * H^{n+1/2} once (the second field and psi_H buffer is allocated on the
 * first switch); `param` = planes per work-group (0..4095) + 4096 * cells
 * per lane per row chunk (1..4) + 65536 * rows per block (1..15, above 4
 * taken as 4), each 0 for the library's choice (8 planes, 2 cells, 2 rows).
 * Returns FDTD_EINVAL (and keeps the schedule) for a grid the fused kernel
 * does not take: per-cell H coefficients, an array of 4 GiB or more, or a
 * row that does not fit 8 wavefronts of that many cells;
 * mode 2 = the dependency-ordered step: one launch runs every H tile and
 * every E tile but those of the first plane of each of 8 plane queues (one
 * per XCD), each E tile after the four H tiles it reads and overwrites for
 * have stored their flags; a small launch then updates E on those 8 planes.
 * `param` = the plane lag of a queue's E tiles behind its H tiles (1..64;
 * 0 for 2).  A call in which some tile did not run or a wait timed out
 * returns FDTD_EDEVICE and the handle returns to mode 0.  Results are
 * identical in every schedule; only instruction count and memory traffic
 * differ.  No reference counterpart. */
int fdtd_set_kernel_mode(fdtd_handle h, int mode, int param);
/* Stability monitor cadence (STABILITY_MONITOR_INTERVAL, default 10; 0 = off). */
int fdtd_set_monitor(fdtd_handle h, int interval);
/* Process-wide cache of device memory: with max_cached_bytes > 0, the
 * blocks of a destroyed handle (and the device bake's temporary buffers) are
 * kept, up to that many bytes per process, and a later allocation on the same
 * device reuses a cached block at most 1/8 larger than it asks for; a
 * handle's pinned host buffers (upload staging, record and result copies)
 * are kept the same way, in power-of-two sizes.  So a serving loop that
 * creates and destroys handles while others run makes neither hipMalloc nor
 * hipFree (which waits for the whole device), nor the slow pinned host
 * allocations, on its common path.  0 (the default) turns it off and frees what it holds.  A
 * handle's fdtd_device_bytes stays the bytes it asked for.  No reference
 * counterpart (the reference allocates once per run). */
int fdtd_set_device_pool(long max_cached_bytes);
/* Bytes the cache holds now, and its cap (either pointer may be null). */
int fdtd_device_pool_bytes(long* cached, long* max_cached_bytes);
/* on != 0: each stepping call on the two-kernel schedule records its
 * launches once in a HIP graph and replays it for later calls of the same
 * shape (step count, step phase in the monitor and FP16 cadences, schedule
 * and device buffers), so a call costs the host one graph launch and two
 * small uploads instead of a few launches per step.  The DFT phase factors
 * and probe record slots of each call's steps come from a device table
 * written before the launch; results are those of the direct launches, bit
 * for bit.  Calls with timing events, TLS diagnostics or the fused schedule
 * launch directly.  Off (the default) frees the recorded graphs.  No
 * reference counterpart (for serving loops that keep many simulations
 * stepping). */
int fdtd_set_step_graphs(fdtd_handle h, int on);

/* ---- time stepping ----
 * Advances nsteps steps starting at global step n0, in the reference's step
 * order.  src_values[s * n_sources + q] is the J value source q contributes at
 * step n0 + s (the host-side waveform value divided by the cell area for a
 * dipole, or 2 E(t)/(eta0 d) * sign for a plane, exactly as the reference
 * computes it).  Returns FDTD_EUNSTABLE if a monitor step in this call fails
 * the reference's stability policy (NaN/Inf or more than 10 threshold
 * warnings); fdtd_read_check gives the records. */
int fdtd_step(fdtd_handle h, long n0, int nsteps, const double* src_values);
/* fdtd_step in two halves, so that a caller can keep many handles' steps in
 * flight at once (each handle has its own stream): fdtd_step_async enqueues
 * the steps and returns without waiting for the device (src_values is read
 * before it returns); fdtd_step_wait waits for them and applies the
 * stability policy to their monitor records (FDTD_EUNSTABLE as fdtd_step).
 * An enqueue on a handle with steps still pending first completes those.
 * Results are those of fdtd_step.  No reference counterpart (the reference
 * steps one simulation synchronously). */
int fdtd_step_async(fdtd_handle h, long n0, int nsteps, const double* src_values);
int fdtd_step_wait(fdtd_handle h);
/* 1 when the handle's enqueued steps have all run on the device (a
 * following fdtd_step_wait returns without waiting), 0 while some are still
 * in flight; never blocks. */
int fdtd_step_query(fdtd_handle h);
int fdtd_synchronize(fdtd_handle h);
/* fdtd_step that also times, with HIP events on the handle's stream, every
 * H-kernel and E-kernel launch: ms[0] = sum of H kernel times, ms[1] = sum of
 * E kernel times, ms[2] = elapsed time of the whole call (all in ms).
 * Synchronizes.  No counterpart in the reference (its timing is the
 * per-10-step wall clock of src/3D_FDTD_version008.cpp:351-353). */
int fdtd_step_timed(fdtd_handle h, long n0, int nsteps, const double* src_values, double* ms);
/* fdtd_step_timed with events around the H and E launches of every
 * `every`-th step only (steps 0, every, 2*every, ... of the call): ms[0] and
 * ms[1] are the sums over those steps, ms[2] the elapsed time of the whole
 * call, ms[3] the number of steps timed.  An event pair costs the stream a
 * few microseconds, so sampling keeps the call's elapsed time that of
 * fdtd_step while the kernel averages stay live measurements.  No
 * counterpart in the reference. */
int fdtd_step_timed_every(fdtd_handle h, long n0, int nsteps, const double* src_values, int every, double* ms);
/* fdtd_step that allocates no device memory, for callers that cannot call
 * hipMalloc (code under graph capture, tight serving loops, harnesses that
 * forbid allocation).  No counterpart in the reference.  fdtd_step grows
 * two per-handle buffers on demand (the call's source values and its
 * monitor records); here both live in one static device array per GPU of
 * FDTD_NOALLOC_SCRATCH_BYTES, records first (48 B per monitor step, rounded
 * up to 256 B) then nsteps * n_sources doubles.  A call needing more returns
 * FDTD_EINVAL without stepping.  Results and return codes are those of
 * fdtd_step; synchronizes.  Every call on one GPU shares that GPU's array,
 * so calls that overlap on one GPU (from several threads) race. */
#define FDTD_NOALLOC_SCRATCH_BYTES 1048576
int fdtd_step_noalloc(fdtd_handle h, long n0, int nsteps, const double* src_values);

/* ---- data movement (reference layout, FP64 on the host side) ----
 * which: 0..5 = Ex,Ey,Ez,Hx,Hy,Hz; 6 = Nu, 7 = Ng, 8 = Pz, 9 = Pz_prev
 * (TLS arrays are full-grid, zero outside gain cells). */
int fdtd_download(fdtd_handle h, int which, double* out);
int fdtd_upload(fdtd_handle h, int which, const double* in);
/* Copies up to max records of probe p into out; *count receives the number
 * recorded so far. */
int fdtd_read_probe(fdtd_handle h, int probe, double* out, long max, long* count);
/* Accumulators of DFT monitor m, layout [freq][cell][re, im], cell in box
 * order (i, j, k with k fastest). */
int fdtd_read_dft(fdtd_handle h, int m, double* out);
/* Enqueues, on the handle's stream after the steps enqueued so far, a copy
 * of every probe's records and every DFT monitor's accumulators into pinned
 * host memory; until the next stepping call, fdtd_read_probe and
 * fdtd_read_dft read that copy (after the handle's own stream) instead of
 * queueing a device-to-host copy, which waits behind other streams' kernels
 * on the shared hardware queues.  For a request loop: stage with a
 * request's last steps, read at retirement.  Not for batch members.  No
 * reference counterpart. */
int fdtd_stage_results(fdtd_handle h);
/* TLS diagnostics as device reductions: every save_every steps (after the
 * population update, where the reference records its TLS detectors,
 * src/3D_FDTD_version008.cpp:330-336) one record of 5 values is kept, up to
 * `capacity` records:
 *   [0] total Nu  = sum Nu dV     (PopulationTimeSeriesDetector::record,
 *   [1] total Ng  = sum Ng dV      src/two_level_detectors.hpp:70-103)
 *   [2] U_int     = sum -0.5 Pz Ez dV   (PolarizationDiagnostics::record,
 *   [3] Pz at cell (i, j, k)             src/two_level_detectors.hpp:230-257)
 *   [4] dPz/dt at cell (i, j, k)
 * over the TLS region cells, dV = dx[i]*dy[j]*dz[k] from the GridSpacing
 * arrays passed here.  The reference sums sequentially; the device sums in
 * double-double and rounds once, so a total is the correctly rounded sum of
 * the same terms (deterministic, within a few ulps per term of the
 * sequential sum).  Inversion, inversion fraction and the TLSGlobalHistory
 * percentages follow from [0] and [1] on the host.  Single handles only. */
int fdtd_add_tls_monitor(fdtd_handle h, const double* dx, const double* dy, const double* dz, int i, int j, int k,
                         int save_every, long capacity);
/* Copies up to max records (5 values each) into out; *count = records taken. */
int fdtd_read_tls_monitor(fdtd_handle h, double* out, long max, long* count);
/* Monitor records of the last fdtd_step call (up to max); *count = number. */
int fdtd_read_check(fdtd_handle h, fdtd_check_record* out, int max, int* count);
/* 1 while the stability policy has passed every monitor step so far, 0 once
 * it has failed (StabilityMonitor::simulation_stable,
 * src/stability_monitor.hpp:28, after which the reference's main returns 1,
 * src/3D_FDTD_version008.cpp:339-349).  Tells a batch's failed members from
 * the rest after fdtd_batch_step returns FDTD_EUNSTABLE. */
int fdtd_is_stable(fdtd_handle h);

/* ---- introspection ---- */
/* info[0..7] = nx, ny, nz, precision, n_materials (LUT entries), material id
 * bytes per cell, gain cells, npml. */
int fdtd_info(fdtd_handle h, long* info);

/* ---- batches: many independent domains of one precision stepped together ----
 * No reference counterpart (the reference runs one domain per process).  A
 * batch advances every member domain with one persistent H launch and one
 * persistent E launch per step (work-groups take tiles of all domains from a
 * device counter) plus one launch each for probes, DFT monitors and the
 * stability monitor of all domains.  Results of a domain are identical
 * whatever batch it runs in.  A handle belongs to at most one batch and must
 * not be stepped on its own while it is a member. */
typedef struct fdtd_batch* fdtd_batch_handle;
int fdtd_batch_create(int precision, fdtd_batch_handle* out);
int fdtd_batch_destroy(fdtd_batch_handle b);
/* Adds a domain; returns its slot (>= 0) or an error. */
int fdtd_batch_add(fdtd_batch_handle b, fdtd_handle h);
int fdtd_batch_remove(fdtd_batch_handle b, fdtd_handle h);
int fdtd_batch_count(fdtd_batch_handle b);
/* Tiling, tuning[4] = {rows per tile (1, 2 or 4), persistent work-groups per
 * CU (1..16), tiles taken per counter increment (1..64; 0: no persistent
 * work-groups, one work-group per tile, the second entry unused), x-planes
 * per tile (1..16)}; a tile is 64 k-values by rows by planes. */
int fdtd_batch_set_tuning(fdtd_batch_handle b, const int* tuning);
int fdtd_batch_get_tuning(fdtd_batch_handle b, int* tuning);
/* on != 0: capture the launches of an fdtd_batch_step call in a HIP graph and
 * replay it for later calls with the same step count and membership. */
int fdtd_batch_set_graph(fdtd_batch_handle b, int on);
/* Runs a block of nsteps steps.  n0[d] is member d's global step index (slot
 * order); member_steps[d] (0..nsteps; NULL = all nsteps) is how many of the
 * block's steps member d takes (it idles for the rest, so members can finish
 * mid-block); src_values[d] its source values [member_steps[d]][n_sources] as
 * for fdtd_step (may be NULL for a domain without sources).  ms (may be NULL)
 * receives the device time of the call.  Returns FDTD_EUNSTABLE if any member
 * fails the stability policy (fdtd_read_check on its handle). */
int fdtd_batch_step(fdtd_batch_handle b, int nsteps, const long* n0, const int* member_steps,
                    const double* const* src_values, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_HIP_H */

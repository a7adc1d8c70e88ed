/* fdtd_scene.h - scenario description for the per-timestep Yee update.
 *
 * This is the runtime form of the reference's compile-time configuration
 * (3D_FDTD_version008 src/user_config.hpp:11-211, src/tls_materials.hpp:23-25)
 * as consumed by Config::sim / make_context / bake_and_bind
 * (src/params.hpp:41-124, 180-516, 557-821).  Every field keeps the meaning,
 * unit and integer encoding of the UserConfig entry it replaces.
 *
 * The scene layer (fdtd_scene_*) turns a spec into the low-level device
 * descriptors of fdtd_hip.h exactly as the reference's setup does on the host
 * (grid spacing, dt rule, CPML profiles, 2x2x2 subpixel bake, TLS regions,
 * source placement and probe placement), then owns one device simulation.
 * Plain C, no allocation by the caller beyond the spec itself.
 */
#ifndef FDTD_SCENE_H
#define FDTD_SCENE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shape types: UserConfig::StructureDef::type "box" / "sphere" / "cylinder"
 * (src/user_config.hpp:162-173).  params[] as documented there, physical
 * coordinates in metres (the PML offset is added by the scene layer,
 * src/params.hpp:560-588). */
enum { FDTD_SHAPE_BOX = 0, FDTD_SHAPE_SPHERE = 1, FDTD_SHAPE_CYLINDER_Z = 2 };

/* Waveforms, src/sources/isource.hpp:16-21 and src/params.hpp:598-606. */
enum { FDTD_WF_RICKER = 0, FDTD_WF_GAUSS_MOD_SINE = 1, FDTD_WF_RICKER_GAUSS2 = 2, FDTD_WF_CW = 3 };

/* Polarizations Ex..Hz = 0..5, src/sources/isource.hpp:23. */
/* Plane-wave directions +X,-X,+Y,-Y,+Z,-Z = 0..5, src/params.hpp:655-663. */
/* Field components 0..11 (Ex,Ey,Ez,Hx,Hy,Hz,|E|,|H|,Sx,Sy,Sz,|S|),
 * src/detectors/idetector.hpp:19-29. */

/* Boundary types, src/user_config.hpp:53 (0 = PEC, 1 = CPML). */
enum { FDTD_BC_PEC = 0, FDTD_BC_CPML = 1 };

typedef struct {
    int    type;        /* FDTD_SHAPE_* */
    double n;           /* refractive index, eps_r = n*n (src/structure_material.hpp:24-30) */
    double params[6];
    int    tls_material;/* index into fdtd_scene_spec::tls_materials, -1 = none */
} fdtd_structure_def;

/* TLSMaterials::TLSMaterial, src/tls_materials.hpp:13-20. */
typedef struct {
    double lambda0, gamma, tau, N0, inversion_fraction;
} fdtd_tls_material;

/* UserConfig::DipoleSourceDef, src/user_config.hpp:68-79 (enabled implied). */
typedef struct {
    double x, y, z, amplitude, frequency, wavelength, tau, df_fwhm, t0_factor;
    int    waveform, polarization;
} fdtd_dipole_def;

/* UserConfig::PlaneWaveSourceDef, src/user_config.hpp:85-98 (enabled implied). */
typedef struct {
    double injection_position, amplitude, frequency, wavelength, tau, df_fwhm, t0_factor;
    int    waveform, direction, polarization;
    double x_min, x_max, y_min, y_max, z_min, z_max;
} fdtd_plane_wave_def;

/* UserConfig::PointFieldDetectorDef, src/user_config.hpp:131-136. */
typedef struct {
    double x, y, z;
    int    n_components;
    int    components[12];
} fdtd_probe_def;

/* Frequency-domain (running DFT) monitor.  No counterpart in the reference
 * C++ (its only spectra are Python FFTs of probe traces,
 * Py_Visualization/plot_point_field.py:139-168).  Accumulates
 * X(f) = sum_n F(t_n) exp(-i 2 pi f t_n) dt over a half-open index box of the
 * total grid; t_n = (n+1) dt for E components, (n+1/2) dt for H components. */
typedef struct {
    int    component;               /* 0..5 */
    int    i0, i1, j0, j1, k0, k1;  /* half-open, total-grid indices */
    int    n_freqs;                 /* <= 8 */
    double freqs[8];                /* Hz */
} fdtd_dft_def;

typedef struct {
    /* Mesh: UNIFORM mode (src/params.hpp:219-234) unless dx/dy/dz are given,
     * in which case they are the full total-grid spacing arrays (core plus
     * 2*npml cells per axis), e.g. a graded nonuniform mesh. */
    int    nx_core, ny_core, nz_core;
    int    npml;                      /* CPMLConfig::npml */
    double dx_uniform, dy_uniform, dz_uniform;
    const double *dx, *dy, *dz;       /* optional, length n*_core + 2*npml */
    double domain_x_min, domain_y_min, domain_z_min;
    double S;                         /* CFL factor */

    int    boundary_type;             /* FDTD_BC_* */
    double cpml_m, cpml_rerr, cpml_alpha0, cpml_kappa_max;
    int    cpml_alpha_linear;

    double background_n;
    int    n_structures;
    const fdtd_structure_def* structures;

    int    tls_enabled, tls_clamp;    /* UserConfig::TLS_ENABLED / TLS_ENABLE_CLAMP */
    int    n_tls_materials;
    const fdtd_tls_material* tls_materials;

    int    n_dipoles;
    const fdtd_dipole_def* dipoles;
    int    n_planes;
    const fdtd_plane_wave_def* planes;

    int    n_probes;
    const fdtd_probe_def* probes;
    int    save_every;                /* SimulationParams::saveEvery */

    int    n_dfts;
    const fdtd_dft_def* dfts;

    int    subpixel_samples;          /* bake_and_bind argument, 8 in the reference */
    int    stability_interval;        /* STABILITY_MONITOR_INTERVAL, 10 */

    /* 1: StructureScene::bake runs on the device when a scene is created
     * (fdtd_create's bake path, bitwise the host bake's values); the setup
     * then keeps no N-sized material array and fdtd_setup_array returns -1
     * for "aEx" ... "bHz" (fdtd_download_material reads them back). */
    int    device_bake;

    /* Optional grid lines (length n*_core + 2*npml + 1 per axis), used with
     * dx/dy/dz: the lines of an auto-meshed grid as make_context keeps them
     * (GridSpacing::*_bounds from AutoMeshGenerator::generate,
     * src/params.hpp:209-211, see fdtd_mesh.h).  Each width is the difference
     * of two neighbouring lines, so running sums of the widths rebuild the
     * lines except where a line is below half its successor (the first cells
     * of an axis); passing them keeps the grid exact there too.  NULL: running
     * sums (build_cumulative_bounds, src/global_function.hpp:160-167). */
    const double *x_bounds, *y_bounds, *z_bounds;
} fdtd_scene_spec;

/* ---- host setup (no device needed) ----
 * fdtd_setup_build runs, on the host, the reference's setup for a spec:
 * make_context grid/dt (src/params.hpp:180-296), make_boundary + CPML
 * profiles (src/boundary.hpp:105-136, 303-373), the structure package and
 * bake_and_bind (src/params.hpp:512-592, src/structure_material.hpp:211-288),
 * main's TLS block (src/3D_FDTD_version008.cpp:100-172), the source package
 * (src/params.hpp:594-678) and the point-probe placement
 * (src/detectors/point_field_detector.hpp:57-70). */
typedef struct fdtd_setup fdtd_setup;
int    fdtd_setup_build(const fdtd_scene_spec* spec, fdtd_setup** out);
void   fdtd_setup_free(fdtd_setup* s);
/* info[0..9] = NxT, NyT, NzT, npml, n_sources, n_probe_channels, n_dfts,
 * tls_on, n_regions, boundary type */
int    fdtd_setup_info(const fdtd_setup* s, long* info);
double fdtd_setup_dt(const fdtd_setup* s);
/* Copies a named setup array (same names as the MaterialGrids / GridSpacing /
 * CPMLBoundary / TwoLevelState members: "bEx", "inv_dx", "kx", "cpml_bHy",
 * "Nu", "Ndip", ...) into out when non-NULL; returns its length or -1. */
long   fdtd_setup_array(const fdtd_setup* s, const char* name, double* out);
/* J value of every source at steps n0 .. n0+nsteps-1, out[s*n_sources + q]. */
int    fdtd_setup_source_values(const fdtd_setup* s, long n0, int nsteps, double* out);
/* Probe channel c (one per probe component): cell and component. */
int    fdtd_setup_probe(const fdtd_setup* s, int c, long* ijk_comp);
/* The initial FP16 storage scales a scene of this setup starts from (fields
 * Ex..Hz, then psi Ex_y..Hz_y; powers of two estimated from the source
 * amplitudes, see fdtd_set_fp16_scales).  Returns 1 and fills out18, or 0
 * when the setup has no source.  No reference counterpart. */
int    fdtd_setup_fp16_scales(const fdtd_setup* s, double* out18);

/* ---- scene: setup + one device simulation ---- */
typedef struct fdtd_scene fdtd_scene;
int  fdtd_scene_create(const fdtd_setup* setup, int precision, long probe_capacity, fdtd_scene** out);
int  fdtd_scene_destroy(fdtd_scene* sc);
/* Advances nsteps steps (source values computed on the host per step). */
int  fdtd_scene_step(fdtd_scene* sc, int nsteps);
/* fdtd_scene_step through fdtd_step_timed (per-kernel HIP-event times, ms[3]). */
int  fdtd_scene_step_timed(fdtd_scene* sc, int nsteps, double* ms);
/* fdtd_step_timed_every on the scene's handle (ms[0..3]).  No reference
 * counterpart. */
int  fdtd_scene_step_timed_every(fdtd_scene* sc, int nsteps, int every, double* ms);
/* fdtd_scene_step through fdtd_step_async / fdtd_step_wait: enqueue, and
 * wait for (and check) what was enqueued.  steps_done counts enqueued steps. */
int  fdtd_scene_step_async(fdtd_scene* sc, int nsteps);
int  fdtd_scene_step_wait(fdtd_scene* sc);
/* fdtd_step_query of the scene's simulation (1 idle, 0 busy). */
int  fdtd_scene_step_query(fdtd_scene* sc);
long fdtd_scene_steps_done(const fdtd_scene* sc);
/* Batch steps never advance the scene past step n_stop (-1: no limit): a
 * batch call in which it reaches n_stop leaves it idle for the rest. */
int  fdtd_scene_set_stop(fdtd_scene* sc, long n_stop);
/* Device handle of the scene (fdtd_hip.h entry points). */
struct fdtd_sim* fdtd_scene_handle(fdtd_scene* sc);

/* ---- batches of scenes (fdtd_batch_* with host-computed source values) ----
 * Member scenes keep their own step counters; a batch step advances every
 * member by nsteps.  Destroying the batch leaves the scenes alive. */
typedef struct fdtd_scene_batch fdtd_scene_batch;
int  fdtd_scene_batch_create(int precision, fdtd_scene_batch** out);
int  fdtd_scene_batch_destroy(fdtd_scene_batch* sb);
int  fdtd_scene_batch_add(fdtd_scene_batch* sb, fdtd_scene* sc);
int  fdtd_scene_batch_remove(fdtd_scene_batch* sb, fdtd_scene* sc);
int  fdtd_scene_batch_count(const fdtd_scene_batch* sb);
/* ms (may be NULL) receives the device time of the call. */
int  fdtd_scene_batch_step(fdtd_scene_batch* sb, int nsteps, double* ms);
/* Autotuner: runs trial_steps real steps under each candidate tiling (rows
 * 1, 4 x work-groups per CU 4, 8 x tiles per counter increment 4, 16 x
 * planes per tile 1, 4, 8; after one warm-up trial), keeps the fastest and
 * writes it to tuning[4] (fdtd_batch_set_tuning order).  A winner is
 * remembered per (precision, member count, total cells rounded to a power of
 * two) for the life of the process; a later call with the same key applies it
 * without trials.  Returns the number of steps taken. */
int  fdtd_scene_batch_autotune(fdtd_scene_batch* sb, int trial_steps, int* tuning);
struct fdtd_batch* fdtd_scene_batch_handle(fdtd_scene_batch* sb);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_SCENE_H */

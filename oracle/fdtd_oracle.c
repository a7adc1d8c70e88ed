/* fdtd_oracle.c - CPU FP64 restatement of the reference's per-timestep Yee
 * update and of the host setup that feeds it.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed
 * "reference OpenMP path" CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links, loads or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to 3D_FDTD_version008/).  Loop ranges, operation order, the OpenMP
 * placement (parallel over i for H, E and TLS; serial CPML passes) and the
 * J zero-fill follow the reference so that, built with the same flags
 * (-O3 -fopenmp -ffp-contract=off), results are bitwise identical to the
 * compiled reference and the timing is representative of it.
 *
 * Parity pins (see tests/): bitwise against golden vectors written by the
 * compiled reference (oracle/_ref, tests/golden/); closed forms for the
 * waveform, CPML profile and TLS coefficient formulas; free-decay of the
 * population (Nu_n = Nu_0 (1 - dt/tau)^n); discrete div H = 0; energy in a
 * lossless PEC cavity.  The DFT monitor has no reference counterpart and is
 * pinned by a direct O(N) sum of a probe trace and by the closed-form DFT of a
 * sampled sinusoid.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/fdtd_scene.h"

/* PhysConst, src/global_function.hpp:12-19 */
static const double PI_ = 3.14159265358979323846;
static const double C0_ = 299792458.0;
static const double EPS0_ = 8.854187817e-12;
#define MU0_ (4.0 * PI_ * 1e-7)
static const double HBAR_ = 1.054571817e-34;
static const double ETA0_ = 376.730313668; /* src/sources/plane_wave_source.hpp:42 */

typedef struct { double x0, x1, y0, y1, z0, z1; } aabb_t;

typedef struct {
    int type;
    double p[6];       /* total-coordinate shape parameters */
    double eps_r, mu_r, sigma, sigma_m;
    int tls;           /* material index or -1 */
} shape_t;

typedef struct {
    /* TwoLevelParams after finalize/finalize_with_dt, src/two_level_system.hpp:17-110 */
    double lambda0, omega_a, gamma, tau, mu_z, N0, hbar;
    double kapa, pa2, pa3, inv_dt;
    int clamp, init;
    size_t i0, i1, j0, j1, k0, k1;
    double inversion_fraction;
    int shape;         /* index into shapes */
} tls_region_t;

typedef struct {
    int kind;          /* 0 dipole, 1 plane */
    int waveform;
    double I0, f0, tau, t0, t_shift;
    int polarization;
    /* dipole */
    size_t i, j, k;
    double cell_area;
    /* plane */
    int direction;
    size_t plane_index, i_min, i_max, j_min, j_max, k_min, k_max;
} source_t;

typedef struct {
    size_t i0, j0, k0;
    int nc;
    int comps[12];
    double* rec;       /* [record][nc] */
    size_t nrec, cap;
} probe_t;

typedef struct {
    fdtd_dft_def def;
    double* acc;       /* [freq][cell][2] */
    size_t ncell;
} dft_t;

typedef struct oracle_sim oracle_sim;
void oracle_destroy(oracle_sim* S);

struct oracle_sim {
    size_t NxT, NyT, NzT, N, npml;
    size_t nx_grid, ny_grid, nz_grid; /* spacing-array lengths */
    double dt;
    double *dx, *dy, *dz, *inv_dx, *inv_dy, *inv_dz, *xb, *yb, *zb;
    int bc_type;
    /* CPML, src/boundary.hpp:81-136 */
    double *kx, *ky, *kz, *sx, *sy, *sz, *ax, *ay, *az;
    double *cbEx, *ccEx, *cbEy, *ccEy, *cbEz, *ccEz, *cbHx, *ccHx, *cbHy, *ccHy, *cbHz, *ccHz;
    double *psi[12]; /* Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y Hx_y Hx_z Hy_z Hy_x Hz_x Hz_y */
    /* MaterialGrids, src/global_function.hpp:70-90 */
    double *aEx, *bEx, *aEy, *bEy, *aEz, *bEz, *aHx, *bHx, *aHy, *bHy, *aHz, *bHz;
    double *Ex, *Ey, *Ez, *Hx, *Hy, *Hz, *Jx, *Jy, *Jz;
    /* scene */
    double bg_eps, bg_mu, bg_sig, bg_sigm;
    int nshapes;
    shape_t* shapes;
    /* TLS, src/two_level_system.hpp:112-197, 488-571 */
    int tls_on, nreg;
    tls_region_t* regs;
    int* cell_region_map;
    double *Ng, *Nu, *Ng0, *Pz, *Pz_prev, *dPz_dt, *Ez_old, *Ndip;
    /* sources, probes, monitors */
    int nsrc;
    source_t* src;
    int nprobe;
    probe_t* probes;
    size_t save_every;
    int ndft;
    dft_t* dfts;
    size_t stab_interval;
    /* StabilityMonitor, src/stability_monitor.hpp:19-40 */
    double st_prev_max_E, st_last_max_E, st_last_cons;
    size_t st_count;
    int st_stable;
    long st_fail_step;
    size_t n;          /* steps done */
    /* storage-format emulation (oracle_set_storage): 0 = FP64 (none) */
    int emu, emu_bits, emu_psi_bits, emu_stoch, emu_block;
    double emu_sc[18], emu_floor[18];
    uint64_t emu_rng;
};

static inline size_t idx3(size_t i, size_t j, size_t k, size_t Ny, size_t Nz) {
    return (i * Ny + j) * Nz + k; /* src/global_function.hpp:21-24 */
}

static double* zalloc(size_t n) {
    double* p = (double*)calloc(n ? n : 1, sizeof(double));
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu doubles)\n", n); abort(); }
    return p;
}
static double* filled(size_t n, double v) {
    double* p = zalloc(n);
    for (size_t i = 0; i < n; ++i) p[i] = v;
    return p;
}

/* build_cumulative_bounds, src/global_function.hpp:160-167 */
static double* cumulative_bounds(const double* s, size_t n) {
    double* b = zalloc(n + 1);
    b[0] = 0.0;
    for (size_t i = 0; i < n; ++i) b[i + 1] = b[i] + s[i];
    return b;
}

/* GridSpacing::physical_to_index_*, src/global_function.hpp:137-157 */
static size_t phys_to_index(const double* bounds, size_t nb, size_t npml, double phys) {
    double off = (npml < nb) ? bounds[npml] : 0.0; /* pml_offset_*: :111-113 */
    double total = phys + off;
    if (total <= bounds[0]) return 0;
    if (total >= bounds[nb - 1]) return nb - 2;
    /* std::upper_bound: first element > total */
    size_t lo = 0, hi = nb;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (!(total < bounds[mid])) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

/* Shape::contains, src/structure_material.hpp:50-85 */
static int shape_contains(const shape_t* s, double x, double y, double z) {
    switch (s->type) {
    case FDTD_SHAPE_BOX:
        return (x >= s->p[0] && x < s->p[1] && y >= s->p[2] && y < s->p[3] && z >= s->p[4] && z < s->p[5]);
    case FDTD_SHAPE_SPHERE: {
        double dx = x - s->p[0], dy = y - s->p[1], dz = z - s->p[2];
        return (dx * dx + dy * dy + dz * dz) <= s->p[3] * s->p[3];
    }
    case FDTD_SHAPE_CYLINDER_Z: {
        if (z < s->p[3] || z >= s->p[4]) return 0;
        double dx = x - s->p[0], dy = y - s->p[1];
        return (dx * dx + dy * dy) <= s->p[2] * s->p[2];
    }
    }
    return 0;
}

/* Shape::bounding_box, src/structure_material.hpp:58, 68-70, 82-84 */
static aabb_t shape_bbox(const shape_t* s) {
    aabb_t b;
    switch (s->type) {
    case FDTD_SHAPE_BOX:
        b.x0 = s->p[0]; b.x1 = s->p[1]; b.y0 = s->p[2]; b.y1 = s->p[3]; b.z0 = s->p[4]; b.z1 = s->p[5];
        break;
    case FDTD_SHAPE_SPHERE:
        b.x0 = s->p[0] - s->p[3]; b.x1 = s->p[0] + s->p[3];
        b.y0 = s->p[1] - s->p[3]; b.y1 = s->p[1] + s->p[3];
        b.z0 = s->p[2] - s->p[3]; b.z1 = s->p[2] + s->p[3];
        break;
    default:
        b.x0 = s->p[0] - s->p[2]; b.x1 = s->p[0] + s->p[2];
        b.y0 = s->p[1] - s->p[2]; b.y1 = s->p[1] + s->p[2];
        b.z0 = s->p[3]; b.z1 = s->p[4];
        break;
    }
    return b;
}

/* CPMLBoundary::build_profiles_and_coeffs, src/boundary.hpp:303-373 */
static double avg_pml_spacing(const double* sp, int N, int npml) {
    double sum = 0;
    int count = 0;
    for (int i = 0; i < npml; ++i) {
        sum += sp[i] + sp[N - 1 - i];
        count += 2;
    }
    return sum / count;
}

static void fill_1d(int N, int npml, double smax, const double* sp, double m, double kmax,
                    double alpha0, int alpha_linear, double* s, double* k, double* a) {
    for (int n = 0; n < N; ++n) {
        int d = n < N - 1 - n ? n : N - 1 - n;
        if (d >= npml) { s[n] = 0.0; k[n] = 1.0; a[n] = 0.0; continue; }
        double r = (double)(npml - (d + 0.5)) / (double)npml;
        double rm = pow(r, m);
        double kap = 1.0 + (kmax - 1.0) * rm;
        double sig = smax * kap * rm;
        double ds_local = sp[n];
        double alf = alpha_linear ? (alpha0 * (C0_ / ds_local) * (1.0 - r))
                                  : (alpha0 * (C0_ / ds_local) * pow(1.0 - r, m));
        s[n] = sig; k[n] = kap; a[n] = alf;
    }
}

static void build_bc(const double* s, const double* k, const double* a, int N, double norm, double dt,
                     double* b, double* c) {
    for (int n = 0; n < N; ++n) {
        double sig_n = s[n] / norm;
        double kap = k[n];
        double alf = a[n];
        double be = exp(-(sig_n / kap + alf) * dt);
        be = be < 1e-30 ? 1e-30 : be; /* std::max<real>(be, 1e-30) */
        double ce = 0.0;
        if (fabs(sig_n) > 1e-30 || fabs(alf) > 1e-30) ce = (sig_n * (be - 1.0)) / (kap * (sig_n + kap * alf));
        b[n] = be; c[n] = ce;
    }
}

/* StructureScene::bake, src/structure_material.hpp:211-288.  Each cell is
 * independent, so it is parallel over i here; the per-cell sample order and
 * summation order are the reference's, so the arrays are identical. */
static void sample_mat(const oracle_sim* S, double x, double y, double z, double* eps_r, double* mu_r,
                       double* sig, double* sigm) {
    double e = S->bg_eps, m = S->bg_mu, s = S->bg_sig, sm = S->bg_sigm;
    for (int q = 0; q < S->nshapes; ++q) {
        const shape_t* sh = &S->shapes[q];
        if (shape_contains(sh, x, y, z)) { e = sh->eps_r; m = sh->mu_r; s = sh->sigma; sm = sh->sigma_m; }
    }
    *eps_r = e; *mu_r = m; *sig = s; *sigm = sm;
}

static void avg_at(const oracle_sim* S, int ns, size_t i, size_t j, size_t k, double ux, double uy, double uz,
                   double* eps, double* mu, double* sg, double* sgm) {
    double dxl = S->dx[i], dyl = S->dy[j], dzl = S->dz[k];
    double x0 = S->xb[i], y0 = S->yb[j], z0 = S->zb[k];
    double eps_sum = 0, mu_sum = 0, sg_sum = 0, sgm_sum = 0;
    int cnt = 0;
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b < ns; ++b)
            for (int c = 0; c < ns; ++c) {
                double rx = (a + 0.5) / ns, ry = (b + 0.5) / ns, rz = (c + 0.5) / ns;
                double xs = x0 + (ux + rx) * dxl;
                double ys = y0 + (uy + ry) * dyl;
                double zs = z0 + (uz + rz) * dzl;
                double er, mr, s, sm;
                sample_mat(S, xs, ys, zs, &er, &mr, &s, &sm);
                eps_sum += er * EPS0_;
                mu_sum += mr * MU0_;
                sg_sum += s;
                sgm_sum += sm;
                ++cnt;
            }
    *eps = eps_sum / cnt; *mu = mu_sum / cnt; *sg = sg_sum / cnt; *sgm = sgm_sum / cnt;
}

static void bake(oracle_sim* S, int subpixel_samples) {
    const size_t N = S->N;
    S->aEx = filled(N, 1); S->bEx = zalloc(N); S->aEy = filled(N, 1); S->bEy = zalloc(N);
    S->aEz = filled(N, 1); S->bEz = zalloc(N); S->aHx = filled(N, 1); S->bHx = zalloc(N);
    S->aHy = filled(N, 1); S->bHy = zalloc(N); S->aHz = filled(N, 1); S->bHz = zalloc(N);
    int ns = (int)round(cbrt((double)subpixel_samples));
    if (ns < 1) ns = 1;
    const double dt = S->dt;
    const long NxT = (long)S->NxT;
#pragma omp parallel for schedule(dynamic, 1)
    for (long il = 0; il < NxT; ++il) {
        size_t i = (size_t)il;
        for (size_t j = 0; j < S->NyT; ++j)
            for (size_t k = 0; k < S->NzT; ++k) {
                size_t id = idx3(i, j, k, S->NyT, S->NzT);
                static const double off[6][3] = {{0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0},
                                                 {0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}};
                double* A[6] = {S->aEx, S->aEy, S->aEz, S->aHx, S->aHy, S->aHz};
                double* B[6] = {S->bEx, S->bEy, S->bEz, S->bHx, S->bHy, S->bHz};
                for (int c = 0; c < 6; ++c) {
                    double eps, mu, sigma, sigm;
                    avg_at(S, ns, i, j, k, off[c][0], off[c][1], off[c][2], &eps, &mu, &sigma, &sigm);
                    if (c < 3) {
                        double c1 = sigma * dt / (2 * eps);
                        A[c][id] = (1 - c1) / (1 + c1);
                        B[c][id] = (dt / eps) / (1 + c1);
                    } else {
                        double c2 = sigm * dt / (2 * mu);
                        A[c][id] = (1 - c2) / (1 + c2);
                        B[c][id] = (dt / mu) / (1 + c2);
                    }
                }
            }
    }
}

/* compute_waveform, src/sources/isource.hpp:41-68 */
static double waveform(int type, double t, double I0, double f0, double tau, double t0, double t_shift) {
    const double pi = PI_;
    switch (type) {
    case FDTD_WF_RICKER: {
        const double a = pi * f0 * (t - t0);
        const double a2 = a * a;
        return I0 * (1 - 2 * a2) * exp(-a2);
    }
    case FDTD_WF_GAUSS_MOD_SINE: {
        const double x = (t - t0);
        return I0 * sin(2 * pi * f0 * (t - t_shift)) * exp(-(x * x) / (tau * tau));
    }
    case FDTD_WF_RICKER_GAUSS2: {
        const double x = (t - t0) / tau;
        const double x2 = x * x;
        return I0 * (1 - 2 * x2) * exp(-x2);
    }
    case FDTD_WF_CW: {
        const double envelope = (t < tau) ? (0.5 * (1 - cos(pi * t / tau))) : 1.0;
        return I0 * envelope * sin(2 * pi * f0 * t);
    }
    }
    return 0;
}

/* SourceConfig::get_frequency/get_tau/get_t0, src/sources/isource.hpp:90-105;
 * tau_from_bandwidth :70-73 */
static double src_freq(double frequency, double wavelength) {
    if (frequency > 0.0) return frequency;
    if (wavelength > 0.0) return C0_ / wavelength;
    return C0_ / 500e-9;
}
static double src_tau(double tau, double df, double dt) {
    if (tau > 0.0) return tau;
    if (df > 0.0) return 2.0 * sqrt(log(2.0)) / (PI_ * df);
    return 20.0 * dt;
}

/* One source's J contribution at step n (the value added to each covered
 * cell): DipoleSource::inject_half_step (src/sources/dipole_source.hpp:60-85)
 * and PlaneWaveSource::inject_*_plane (src/sources/plane_wave_source.hpp:102-179). */
static double source_value(const source_t* s, size_t n, double dt) {
    const double t_half = ((double)n + 0.5) * dt;
    const double v = waveform(s->waveform, t_half, s->I0, s->f0, s->tau, s->t0, s->t_shift);
    if (s->kind == 0) return v / s->cell_area;
    return v; /* plane: scaled per plane below */
}

static void inject(oracle_sim* S, size_t n) {
    const size_t NyT = S->NyT, NzT = S->NzT;
    for (int q = 0; q < S->nsrc; ++q) {
        const source_t* s = &S->src[q];
        if (s->kind == 0) {
            const size_t id = idx3(s->i, s->j, s->k, NyT, NzT);
            const double Jval = source_value(s, n, S->dt);
            switch (s->polarization) {
            case 0: case 3: S->Jx[id] += Jval; break;
            case 1: case 4: S->Jy[id] += Jval; break;
            default: S->Jz[id] += Jval; break;
            }
        } else {
            const double Et = source_value(s, n, S->dt);
            const int dir = s->direction;
            if (dir == 4 || dir == 5) {
                const double sign = (dir == 4) ? 1.0 : -1.0;
                for (size_t i = s->i_min; i <= s->i_max; ++i)
                    for (size_t j = s->j_min; j <= s->j_max; ++j) {
                        const size_t id = idx3(i, j, s->plane_index, NyT, NzT);
                        const double J = 2.0 * Et / (ETA0_ * S->dz[s->plane_index]) * sign;
                        if (s->polarization == 0 || s->polarization == 3) S->Jx[id] += J;
                        else if (s->polarization == 1 || s->polarization == 4) S->Jy[id] += J;
                    }
            } else if (dir == 0 || dir == 1) {
                const double sign = (dir == 0) ? 1.0 : -1.0;
                for (size_t j = s->j_min; j <= s->j_max; ++j)
                    for (size_t k = s->k_min; k <= s->k_max; ++k) {
                        const size_t id = idx3(s->plane_index, j, k, NyT, NzT);
                        const double J = 2.0 * Et / (ETA0_ * S->dx[s->plane_index]) * sign;
                        if (s->polarization == 1 || s->polarization == 4) S->Jy[id] += J;
                        else if (s->polarization == 2 || s->polarization == 5) S->Jz[id] += J;
                    }
            } else {
                const double sign = (dir == 2) ? 1.0 : -1.0;
                for (size_t i = s->i_min; i <= s->i_max; ++i)
                    for (size_t k = s->k_min; k <= s->k_max; ++k) {
                        const size_t id = idx3(i, s->plane_index, k, NyT, NzT);
                        const double J = 2.0 * Et / (ETA0_ * S->dy[s->plane_index]) * sign;
                        if (s->polarization == 0 || s->polarization == 3) S->Jx[id] += J;
                        else if (s->polarization == 2 || s->polarization == 5) S->Jz[id] += J;
                    }
            }
        }
    }
}

/* fdtd_update_H, src/fdtd_stepper.hpp:8-47 */
static void update_H(oracle_sim* S) {
    const size_t NyT = S->NyT, NzT = S->NzT, sI = NyT * NzT, sJ = NzT, sK = 1;
    const double *Ex = S->Ex, *Ey = S->Ey, *Ez = S->Ez;
    double *Hx = S->Hx, *Hy = S->Hy, *Hz = S->Hz;
    const int ni = (int)(S->NxT - 1), nj = (int)(NyT - 1), nk = (int)(NzT - 1);
#pragma omp parallel for
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                size_t id = idx3((size_t)i, (size_t)j, (size_t)k, NyT, NzT);
                const double inv_dx = S->inv_dx[i], inv_dy = S->inv_dy[j], inv_dz = S->inv_dz[k];
                double curlEx = (Ez[id + sJ] - Ez[id]) * inv_dy - (Ey[id + sK] - Ey[id]) * inv_dz;
                double curlEy = (Ex[id + sK] - Ex[id]) * inv_dz - (Ez[id + sI] - Ez[id]) * inv_dx;
                double curlEz = (Ey[id + sI] - Ey[id]) * inv_dx - (Ex[id + sJ] - Ex[id]) * inv_dy;
                Hx[id] = S->aHx[id] * Hx[id] - S->bHx[id] * curlEx;
                Hy[id] = S->aHy[id] * Hy[id] - S->bHy[id] * curlEy;
                Hz[id] = S->aHz[id] * Hz[id] - S->bHz[id] * curlEz;
            }
}

/* fdtd_update_E (src/fdtd_stepper.hpp:49-89) and fdtd_update_E_with_gain
 * (src/two_level_system.hpp:299-342) */
static void update_E(oracle_sim* S, int with_gain) {
    const size_t NyT = S->NyT, NzT = S->NzT, sI = NyT * NzT, sJ = NzT, sK = 1;
    double *Ex = S->Ex, *Ey = S->Ey, *Ez = S->Ez;
    const double *Hx = S->Hx, *Hy = S->Hy, *Hz = S->Hz;
    const int ni = (int)S->NxT, nj = (int)NyT, nk = (int)NzT;
#pragma omp parallel for
    for (int i = 1; i < ni; ++i)
        for (int j = 1; j < nj; ++j)
            for (int k = 1; k < nk; ++k) {
                size_t id = idx3((size_t)i, (size_t)j, (size_t)k, NyT, NzT);
                const double inv_dx = S->inv_dx[i], inv_dy = S->inv_dy[j], inv_dz = S->inv_dz[k];
                double curlHx = (Hz[id] - Hz[id - sJ]) * inv_dy - (Hy[id] - Hy[id - sK]) * inv_dz;
                double curlHy = (Hx[id] - Hx[id - sK]) * inv_dz - (Hz[id] - Hz[id - sI]) * inv_dx;
                double curlHz = (Hy[id] - Hy[id - sI]) * inv_dx - (Hx[id] - Hx[id - sJ]) * inv_dy;
                Ex[id] = S->aEx[id] * Ex[id] + S->bEx[id] * (curlHx - S->Jx[id]);
                Ey[id] = S->aEy[id] * Ey[id] + S->bEy[id] * (curlHy - S->Jy[id]);
                if (with_gain) {
                    S->Ez_old[id] = Ez[id];
                    double Pz_source = S->dPz_dt[id];
                    Ez[id] = S->aEz[id] * Ez[id] + S->bEz[id] * (curlHz - S->Jz[id] - Pz_source);
                } else {
                    Ez[id] = S->aEz[id] * Ez[id] + S->bEz[id] * (curlHz - S->Jz[id]);
                }
            }
}

static inline int in_pml(int i, int N, int npml) {
    int d = i < N - 1 - i ? i : N - 1 - i;
    return d < npml;
}

/* CPMLBoundary::apply_after_H, src/boundary.hpp:150-225 (serial) */
static void cpml_after_H(oracle_sim* S) {
    const size_t NxT = S->NxT, NyT = S->NyT, NzT = S->NzT, sI = NyT * NzT, sJ = NzT, sK = 1;
    const int np = (int)S->npml;
    const double *Ex = S->Ex, *Ey = S->Ey, *Ez = S->Ez;
    double *Hx = S->Hx, *Hy = S->Hy, *Hz = S->Hz;
    for (size_t i = 0; i < NxT - 1; ++i)
        for (size_t j = 0; j < NyT - 1; ++j)
            for (size_t k = 0; k < NzT - 1; ++k) {
                size_t id = idx3(i, j, k, NyT, NzT);
                const double inv_dx = 1.0 / S->dx[i], inv_dy = 1.0 / S->dy[j], inv_dz = 1.0 / S->dz[k];
                if (in_pml((int)j, (int)NyT, np)) {
                    double d = (Ez[id + sJ] - Ez[id]) * inv_dy;
                    double corr = d / S->ky[j];
                    double* p = &S->psi[6][id]; *p = S->cbHy[j] * *p + S->ccHy[j] * d; corr += *p;
                    Hx[id] -= S->bHx[id] * (corr - d);
                }
                if (in_pml((int)k, (int)NzT, np)) {
                    double d = (Ey[id + sK] - Ey[id]) * inv_dz;
                    double corr = d / S->kz[k];
                    double* p = &S->psi[7][id]; *p = S->cbHz[k] * *p + S->ccHz[k] * d; corr += *p;
                    Hx[id] -= S->bHx[id] * (-(corr - d));
                }
                if (in_pml((int)k, (int)NzT, np)) {
                    double d = (Ex[id + sK] - Ex[id]) * inv_dz;
                    double corr = d / S->kz[k];
                    double* p = &S->psi[8][id]; *p = S->cbHz[k] * *p + S->ccHz[k] * d; corr += *p;
                    Hy[id] -= S->bHy[id] * (corr - d);
                }
                if (in_pml((int)i, (int)NxT, np)) {
                    double d = (Ez[id + sI] - Ez[id]) * inv_dx;
                    double corr = d / S->kx[i];
                    double* p = &S->psi[9][id]; *p = S->cbHx[i] * *p + S->ccHx[i] * d; corr += *p;
                    Hy[id] -= S->bHy[id] * (-(corr - d));
                }
                if (in_pml((int)i, (int)NxT, np)) {
                    double d = (Ey[id + sI] - Ey[id]) * inv_dx;
                    double corr = d / S->kx[i];
                    double* p = &S->psi[10][id]; *p = S->cbHx[i] * *p + S->ccHx[i] * d; corr += *p;
                    Hz[id] -= S->bHz[id] * (corr - d);
                }
                if (in_pml((int)j, (int)NyT, np)) {
                    double d = (Ex[id + sJ] - Ex[id]) * inv_dy;
                    double corr = d / S->ky[j];
                    double* p = &S->psi[11][id]; *p = S->cbHy[j] * *p + S->ccHy[j] * d; corr += *p;
                    Hz[id] -= S->bHz[id] * (-(corr - d));
                }
            }
}

/* CPMLBoundary::apply_after_E, src/boundary.hpp:227-300 (serial) */
static void cpml_after_E(oracle_sim* S) {
    const size_t NxT = S->NxT, NyT = S->NyT, NzT = S->NzT, sI = NyT * NzT, sJ = NzT, sK = 1;
    const int np = (int)S->npml;
    double *Ex = S->Ex, *Ey = S->Ey, *Ez = S->Ez;
    const double *Hx = S->Hx, *Hy = S->Hy, *Hz = S->Hz;
    for (size_t i = 1; i < NxT; ++i)
        for (size_t j = 1; j < NyT; ++j)
            for (size_t k = 1; k < NzT; ++k) {
                size_t id = idx3(i, j, k, NyT, NzT);
                const double inv_dx = 1.0 / S->dx[i], inv_dy = 1.0 / S->dy[j], inv_dz = 1.0 / S->dz[k];
                if (in_pml((int)j, (int)NyT, np)) {
                    double d = (Hz[id] - Hz[id - sJ]) * inv_dy;
                    double corr = d / S->ky[j];
                    double* p = &S->psi[0][id]; *p = S->cbEy[j] * *p + S->ccEy[j] * d; corr += *p;
                    Ex[id] += S->bEx[id] * (corr - d);
                }
                if (in_pml((int)k, (int)NzT, np)) {
                    double d = (Hy[id] - Hy[id - sK]) * inv_dz;
                    double corr = d / S->kz[k];
                    double* p = &S->psi[1][id]; *p = S->cbEz[k] * *p + S->ccEz[k] * d; corr += *p;
                    Ex[id] += S->bEx[id] * (-(corr - d));
                }
                if (in_pml((int)k, (int)NzT, np)) {
                    double d = (Hx[id] - Hx[id - sK]) * inv_dz;
                    double corr = d / S->kz[k];
                    double* p = &S->psi[2][id]; *p = S->cbEz[k] * *p + S->ccEz[k] * d; corr += *p;
                    Ey[id] += S->bEy[id] * (corr - d);
                }
                if (in_pml((int)i, (int)NxT, np)) {
                    double d = (Hz[id] - Hz[id - sI]) * inv_dx;
                    double corr = d / S->kx[i];
                    double* p = &S->psi[3][id]; *p = S->cbEx[i] * *p + S->ccEx[i] * d; corr += *p;
                    Ey[id] += S->bEy[id] * (-(corr - d));
                }
                if (in_pml((int)i, (int)NxT, np)) {
                    double d = (Hy[id] - Hy[id - sI]) * inv_dx;
                    double corr = d / S->kx[i];
                    double* p = &S->psi[4][id]; *p = S->cbEx[i] * *p + S->ccEx[i] * d; corr += *p;
                    Ez[id] += S->bEz[id] * (corr - d);
                }
                if (in_pml((int)j, (int)NyT, np)) {
                    double d = (Hx[id] - Hx[id - sJ]) * inv_dy;
                    double corr = d / S->ky[j];
                    double* p = &S->psi[5][id]; *p = S->cbEy[j] * *p + S->ccEy[j] * d; corr += *p;
                    Ez[id] += S->bEz[id] * (-(corr - d));
                }
            }
}

/* PECBoundary::apply_after_E, src/boundary.hpp:51-79 */
static void pec_after_E(oracle_sim* S) {
    const size_t Nx = S->NxT, Ny = S->NyT, Nz = S->NzT;
    for (size_t j = 0; j < Ny; ++j)
        for (size_t k = 0; k < Nz; ++k) {
            size_t a = idx3(0, j, k, Ny, Nz), b = idx3(Nx - 1, j, k, Ny, Nz);
            S->Ex[a] = S->Ey[a] = S->Ez[a] = 0.0;
            S->Ex[b] = S->Ey[b] = S->Ez[b] = 0.0;
        }
    for (size_t i = 0; i < Nx; ++i)
        for (size_t k = 0; k < Nz; ++k) {
            size_t a = idx3(i, 0, k, Ny, Nz), b = idx3(i, Ny - 1, k, Ny, Nz);
            S->Ex[a] = S->Ey[a] = S->Ez[a] = 0.0;
            S->Ex[b] = S->Ey[b] = S->Ez[b] = 0.0;
        }
    for (size_t i = 0; i < Nx; ++i)
        for (size_t j = 0; j < Ny; ++j) {
            size_t a = idx3(i, j, 0, Ny, Nz), b = idx3(i, j, Nz - 1, Ny, Nz);
            S->Ex[a] = S->Ey[a] = S->Ez[a] = 0.0;
            S->Ex[b] = S->Ey[b] = S->Ez[b] = 0.0;
        }
}

/* update_polarization_multi_region, src/two_level_system.hpp:574-609 */
static void update_polarization(oracle_sim* S) {
    const size_t NyT = S->NyT, NzT = S->NzT;
    const int ni = (int)S->NxT, nj = (int)NyT, nk = (int)NzT;
#pragma omp parallel for
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                size_t id = idx3((size_t)i, (size_t)j, (size_t)k, NyT, NzT);
                if (S->Ndip[id] <= 0) continue;
                int r = S->cell_region_map[id];
                if (r < 0 || r >= S->nreg) continue;
                const tls_region_t* p = &S->regs[r];
                if (!p->init) continue;
                double Ng_frac = (S->Ng0[id] > 0) ? (S->Ng[id] - S->Nu[id]) / S->Ng0[id] : 0.0;
                double driving = S->Ndip[id] * Ng_frac * S->Ez[id];
                double Pz_new = p->kapa * driving + p->pa2 * S->Pz[id] + p->pa3 * S->Pz_prev[id];
                double Pz_old = S->Pz[id];
                S->Pz_prev[id] = Pz_old;
                S->Pz[id] = Pz_new;
                S->dPz_dt[id] = (Pz_new - Pz_old) * p->inv_dt;
            }
}

/* update_populations_multi_region, src/two_level_system.hpp:612-666 */
static void update_populations(oracle_sim* S) {
    const size_t NyT = S->NyT, NzT = S->NzT;
    const double dt = S->dt;
    const int ni = (int)S->NxT, nj = (int)NyT, nk = (int)NzT;
#pragma omp parallel for
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                size_t id = idx3((size_t)i, (size_t)j, (size_t)k, NyT, NzT);
                if (S->Ndip[id] <= 0) continue;
                int r = S->cell_region_map[id];
                if (r < 0 || r >= S->nreg) continue;
                const tls_region_t* p = &S->regs[r];
                double Nu_curr = S->Nu[id];
                double Ntotal = S->Ng0[id];
                if (Ntotal <= 0) continue;
                double inv_tau = 1.0 / p->tau;
                double inv_hbar_omega = 1.0 / (p->hbar * p->omega_a);
                double E_avg = 0.5 * (S->Ez[id] + S->Ez_old[id]);
                double delta_P = S->Pz[id] - S->Pz_prev[id];
                double stim_rate = E_avg * delta_P * p->inv_dt * inv_hbar_omega;
                double dNu_dt = -inv_tau * Nu_curr + stim_rate;
                double dNg_dt = +inv_tau * Nu_curr - stim_rate;
                S->Nu[id] += dt * dNu_dt;
                S->Ng[id] += dt * dNg_dt;
                S->Nu[id] = (0.0 < S->Nu[id]) ? S->Nu[id] : 0.0; /* std::max(0.0, x) */
                S->Ng[id] = (0.0 < S->Ng[id]) ? S->Ng[id] : 0.0;
                if (p->clamp) {
                    double total = S->Nu[id] + S->Ng[id];
                    if (total > 1e-30) {
                        double scale = Ntotal / total;
                        S->Nu[id] *= scale;
                        S->Ng[id] *= scale;
                    }
                    S->Nu[id] = (Ntotal < S->Nu[id]) ? Ntotal : S->Nu[id]; /* std::min */
                    S->Ng[id] = Ntotal - S->Nu[id];
                }
            }
}

/* get_field_value, src/detectors/idetector.hpp:79-109 */
static double field_value(const oracle_sim* S, int c, size_t id) {
    const double ex = S->Ex[id], ey = S->Ey[id], ez = S->Ez[id];
    const double hx = S->Hx[id], hy = S->Hy[id], hz = S->Hz[id];
    switch (c) {
    case 0: return ex;
    case 1: return ey;
    case 2: return ez;
    case 3: return hx;
    case 4: return hy;
    case 5: return hz;
    case 6: return sqrt(ex * ex + ey * ey + ez * ez);
    case 7: return sqrt(hx * hx + hy * hy + hz * hz);
    case 8: return ey * hz - ez * hy;
    case 9: return ez * hx - ex * hz;
    case 10: return ex * hy - ey * hx;
    case 11: {
        double sx = ey * hz - ez * hy, sy = ez * hx - ex * hz, sz = ex * hy - ey * hx;
        return sqrt(sx * sx + sy * sy + sz * sz);
    }
    }
    return 0.0;
}

/* PointFieldDetector::record_after_E, src/detectors/point_field_detector.hpp:111-124 */
static void record_probes(oracle_sim* S, size_t n) {
    if (n % S->save_every != 0) return;
    for (int p = 0; p < S->nprobe; ++p) {
        probe_t* pr = &S->probes[p];
        size_t id = idx3(pr->i0, pr->j0, pr->k0, S->NyT, S->NzT);
        if (pr->nrec == pr->cap) {
            pr->cap = pr->cap ? 2 * pr->cap : 1024;
            pr->rec = (double*)realloc(pr->rec, pr->cap * (size_t)pr->nc * sizeof(double));
        }
        for (int c = 0; c < pr->nc; ++c) pr->rec[pr->nrec * (size_t)pr->nc + (size_t)c] = field_value(S, pr->comps[c], id);
        pr->nrec++;
    }
}

/* Running DFT (no reference counterpart).  Weights cos(w t) dt and
 * -sin(w t) dt with w = 2 pi f. */
static void accumulate_dft(oracle_sim* S, size_t n) {
    for (int m = 0; m < S->ndft; ++m) {
        dft_t* d = &S->dfts[m];
        const fdtd_dft_def* D = &d->def;
        const double t = (D->component < 3) ? ((double)n + 1.0) * S->dt : ((double)n + 0.5) * S->dt;
        for (int f = 0; f < D->n_freqs; ++f) {
            const double w = 2.0 * PI_ * D->freqs[f];
            const double wr = cos(w * t) * S->dt, wi = -sin(w * t) * S->dt;
            double* acc = d->acc + (size_t)f * d->ncell * 2;
            size_t c = 0;
            for (int i = D->i0; i < D->i1; ++i)
                for (int j = D->j0; j < D->j1; ++j)
                    for (int k = D->k0; k < D->k1; ++k, ++c) {
                        size_t id = idx3((size_t)i, (size_t)j, (size_t)k, S->NyT, S->NzT);
                        const double F = field_value(S, D->component, id);
                        acc[2 * c] += F * wr;
                        acc[2 * c + 1] += F * wi;
                    }
        }
    }
}

/* NumericUtils::has_nan_or_inf / max_abs, src/global_function.hpp:169-186 */
static int has_nan_or_inf(const double* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (isnan(v[i]) || isinf(v[i])) return 1;
    return 0;
}
static double max_abs(const double* v, size_t n) {
    double m = 0;
    for (size_t i = 0; i < n; ++i) {
        double a = fabs(v[i]);
        m = (m < a) ? a : m; /* std::max(m, |v|) */
    }
    return m;
}

/* StabilityMonitor::calculate_conservation_error, src/stability_monitor.hpp:155-170 */
static double conservation_error(const oracle_sim* S) {
    if (!S->tls_on) return 0.0;
    double total = 0.0, init = 0.0;
    for (size_t id = 0; id < S->N; ++id)
        if (S->Ndip[id] > 1e10) { total += (S->Nu[id] + S->Ng[id]); init += S->Ng0[id]; }
    if (init < 1e-30) return 0.0;
    return fabs(total - init) / init;
}

/* StabilityMonitor::check_stability, src/stability_monitor.hpp:42-144.
 * Thresholds: src/user_config.hpp:156-160. */
static int check_stability(oracle_sim* S, size_t step, double cons) {
    const size_t N = S->N;
    if (has_nan_or_inf(S->Ex, N) || has_nan_or_inf(S->Ey, N) || has_nan_or_inf(S->Ez, N)) { S->st_stable = 0; return 0; }
    if (has_nan_or_inf(S->Hx, N) || has_nan_or_inf(S->Hy, N) || has_nan_or_inf(S->Hz, N)) { S->st_stable = 0; return 0; }
    if (S->tls_on) {
        if (has_nan_or_inf(S->Pz, N) || has_nan_or_inf(S->dPz_dt, N)) { S->st_stable = 0; return 0; }
        if (has_nan_or_inf(S->Nu, N) || has_nan_or_inf(S->Ng, N)) { S->st_stable = 0; return 0; }
    }
    double a = max_abs(S->Ex, N), b = max_abs(S->Ey, N), c = max_abs(S->Ez, N);
    double max_E = a;
    if (max_E < b) max_E = b;
    if (max_E < c) max_E = c;
    S->st_last_max_E = max_E;
    if (max_E > 1e10) S->st_count++;
    if (S->tls_on) {
        double max_P = max_abs(S->Pz, N);
        if (max_P > 1e5) S->st_count++;
    }
    (void)cons; /* conservation / negative-population warnings do not change the verdict */
    if (step > 0 && step % 100 == 0 && S->st_prev_max_E > 1e-20) {
        S->st_prev_max_E = max_E;
    } else if (step == 0 || S->st_prev_max_E < 1e-20) {
        S->st_prev_max_E = max_E > 1e-20 ? max_E : 1e-20;
    }
    if (S->st_count > 10) { S->st_stable = 0; return 0; }
    return S->st_stable;
}

/* ---------------------------------------------------------------------- */
/* Setup: make_context (src/params.hpp:180-296), register_default_structures
 * (:557-592), bake_and_bind (:512-516), main's TLS block
 * (src/3D_FDTD_version008.cpp:100-172), register_default_sources
 * (src/params.hpp:594-678), PointFieldDetector::initialize
 * (src/detectors/point_field_detector.hpp:57-124). */
oracle_sim* oracle_create(const fdtd_scene_spec* sp) {
    oracle_sim* S = (oracle_sim*)calloc(1, sizeof(oracle_sim));
    const size_t npml_spec = (size_t)sp->npml;
    size_t nxg, nyg, nzg;
    if (sp->dx && sp->dy && sp->dz) {
        nxg = (size_t)sp->nx_core + 2 * npml_spec;
        nyg = (size_t)sp->ny_core + 2 * npml_spec;
        nzg = (size_t)sp->nz_core + 2 * npml_spec;
        S->dx = zalloc(nxg); memcpy(S->dx, sp->dx, nxg * sizeof(double));
        S->dy = zalloc(nyg); memcpy(S->dy, sp->dy, nyg * sizeof(double));
        S->dz = zalloc(nzg); memcpy(S->dz, sp->dz, nzg * sizeof(double));
    } else {
        nxg = (size_t)sp->nx_core + 2 * npml_spec;
        nyg = (size_t)sp->ny_core + 2 * npml_spec;
        nzg = (size_t)sp->nz_core + 2 * npml_spec;
        S->dx = filled(nxg, sp->dx_uniform);
        S->dy = filled(nyg, sp->dy_uniform);
        S->dz = filled(nzg, sp->dz_uniform);
    }
    S->nx_grid = nxg; S->ny_grid = nyg; S->nz_grid = nzg;
    if (sp->dx && sp->dy && sp->dz && sp->x_bounds && sp->y_bounds && sp->z_bounds) {
        /* auto-meshed lines as given (GridSpacing after AutoMeshGenerator::generate) */
        S->xb = zalloc(nxg + 1); memcpy(S->xb, sp->x_bounds, (nxg + 1) * sizeof(double));
        S->yb = zalloc(nyg + 1); memcpy(S->yb, sp->y_bounds, (nyg + 1) * sizeof(double));
        S->zb = zalloc(nzg + 1); memcpy(S->zb, sp->z_bounds, (nzg + 1) * sizeof(double));
    } else {
        S->xb = cumulative_bounds(S->dx, nxg);
        S->yb = cumulative_bounds(S->dy, nyg);
        S->zb = cumulative_bounds(S->dz, nzg);
    }
    /* GridSpacing::compute_inverses, src/global_function.hpp:98-105 */
    S->inv_dx = zalloc(nxg); S->inv_dy = zalloc(nyg); S->inv_dz = zalloc(nzg);
    for (size_t i = 0; i < nxg; ++i) S->inv_dx[i] = 1.0 / S->dx[i];
    for (size_t i = 0; i < nyg; ++i) S->inv_dy[i] = 1.0 / S->dy[i];
    for (size_t i = 0; i < nzg; ++i) S->inv_dz[i] = 1.0 / S->dz[i];

    /* dt rule, src/params.hpp:243-268 */
    double dxm = S->dx[0], dym = S->dy[0], dzm = S->dz[0];
    for (size_t i = 1; i < nxg; ++i) if (S->dx[i] < dxm) dxm = S->dx[i];
    for (size_t i = 1; i < nyg; ++i) if (S->dy[i] < dym) dym = S->dy[i];
    for (size_t i = 1; i < nzg; ++i) if (S->dz[i] < dzm) dzm = S->dz[i];
    const double sum_inv2 = (1.0 / (dxm * dxm)) + (1.0 / (dym * dym)) + (1.0 / (dzm * dzm));
    S->dt = sp->S / (C0_ * sqrt(sum_inv2));

    S->bc_type = sp->boundary_type;
    if (sp->boundary_type == FDTD_BC_PEC) {
        /* PECBoundary(NxCore,...) reports the core size and npml 0 */
        S->NxT = (size_t)sp->nx_core; S->NyT = (size_t)sp->ny_core; S->NzT = (size_t)sp->nz_core;
        S->npml = 0;
    } else {
        S->NxT = nxg; S->NyT = nyg; S->NzT = nzg;
        S->npml = npml_spec;
    }
    const size_t NxT = S->NxT, NyT = S->NyT, NzT = S->NzT;
    S->N = NxT * NyT * NzT;
    const size_t N = S->N;

    /* CPML profiles and coefficients */
    S->kx = filled(NxT, 1.0); S->ky = filled(NyT, 1.0); S->kz = filled(NzT, 1.0);
    S->sx = zalloc(NxT); S->sy = zalloc(NyT); S->sz = zalloc(NzT);
    S->ax = zalloc(NxT); S->ay = zalloc(NyT); S->az = zalloc(NzT);
    S->cbEx = filled(NxT, 1.0); S->ccEx = zalloc(NxT); S->cbEy = filled(NyT, 1.0); S->ccEy = zalloc(NyT);
    S->cbEz = filled(NzT, 1.0); S->ccEz = zalloc(NzT); S->cbHx = filled(NxT, 1.0); S->ccHx = zalloc(NxT);
    S->cbHy = filled(NyT, 1.0); S->ccHy = zalloc(NyT); S->cbHz = filled(NzT, 1.0); S->ccHz = zalloc(NzT);
    if (sp->boundary_type == FDTD_BC_CPML) {
        const int np = (int)S->npml;
        for (int q = 0; q < 12; ++q) S->psi[q] = zalloc(N);
        double dxa = avg_pml_spacing(S->dx, (int)NxT, np);
        double dya = avg_pml_spacing(S->dy, (int)NyT, np);
        double dza = avg_pml_spacing(S->dz, (int)NzT, np);
        double smx, smy, smz;
        {
            double d;
            d = np * dxa; smx = ((sp->cpml_m + 1.0) * log(1.0 / sp->cpml_rerr) * EPS0_ * C0_) / (2.0 * d);
            d = np * dya; smy = ((sp->cpml_m + 1.0) * log(1.0 / sp->cpml_rerr) * EPS0_ * C0_) / (2.0 * d);
            d = np * dza; smz = ((sp->cpml_m + 1.0) * log(1.0 / sp->cpml_rerr) * EPS0_ * C0_) / (2.0 * d);
        }
        fill_1d((int)NxT, np, smx, S->dx, sp->cpml_m, sp->cpml_kappa_max, sp->cpml_alpha0, sp->cpml_alpha_linear, S->sx, S->kx, S->ax);
        fill_1d((int)NyT, np, smy, S->dy, sp->cpml_m, sp->cpml_kappa_max, sp->cpml_alpha0, sp->cpml_alpha_linear, S->sy, S->ky, S->ay);
        fill_1d((int)NzT, np, smz, S->dz, sp->cpml_m, sp->cpml_kappa_max, sp->cpml_alpha0, sp->cpml_alpha_linear, S->sz, S->kz, S->az);
        build_bc(S->sx, S->kx, S->ax, (int)NxT, EPS0_, S->dt, S->cbEx, S->ccEx);
        build_bc(S->sy, S->ky, S->ay, (int)NyT, EPS0_, S->dt, S->cbEy, S->ccEy);
        build_bc(S->sz, S->kz, S->az, (int)NzT, EPS0_, S->dt, S->cbEz, S->ccEz);
        build_bc(S->sx, S->kx, S->ax, (int)NxT, MU0_, S->dt, S->cbHx, S->ccHx);
        build_bc(S->sy, S->ky, S->ay, (int)NyT, MU0_, S->dt, S->cbHy, S->ccHy);
        build_bc(S->sz, S->kz, S->az, (int)NzT, MU0_, S->dt, S->cbHz, S->ccHz);
    }

    /* scene: make_nk(n) => eps_r = n*n, mu_r 1, sigma 0 */
    S->bg_eps = sp->background_n * sp->background_n; S->bg_mu = 1.0; S->bg_sig = 0.0; S->bg_sigm = 0.0;
    const double pml_x = S->xb[S->npml], pml_y = S->yb[S->npml], pml_z = S->zb[S->npml];
    S->nshapes = sp->n_structures;
    S->shapes = (shape_t*)calloc((size_t)(sp->n_structures ? sp->n_structures : 1), sizeof(shape_t));
    for (int q = 0; q < sp->n_structures; ++q) {
        const fdtd_structure_def* d = &sp->structures[q];
        shape_t* sh = &S->shapes[q];
        sh->type = d->type;
        sh->eps_r = d->n * d->n; sh->mu_r = 1.0; sh->sigma = 0.0; sh->sigma_m = 0.0;
        sh->tls = (sp->tls_enabled && d->tls_material >= 0 && d->tls_material < sp->n_tls_materials) ? d->tls_material : -1;
        if (d->type == FDTD_SHAPE_BOX) {
            sh->p[0] = pml_x + d->params[0]; sh->p[1] = pml_x + d->params[1];
            sh->p[2] = pml_y + d->params[2]; sh->p[3] = pml_y + d->params[3];
            sh->p[4] = pml_z + d->params[4]; sh->p[5] = pml_z + d->params[5];
        } else if (d->type == FDTD_SHAPE_SPHERE) {
            sh->p[0] = pml_x + d->params[0]; sh->p[1] = pml_y + d->params[1];
            sh->p[2] = pml_z + d->params[2]; sh->p[3] = d->params[3];
        } else {
            sh->p[0] = pml_x + d->params[0]; sh->p[1] = pml_y + d->params[1]; sh->p[2] = d->params[2];
            sh->p[3] = pml_z + d->params[3]; sh->p[4] = pml_z + d->params[4];
        }
    }
    bake(S, sp->subpixel_samples > 0 ? sp->subpixel_samples : 8);

    S->Ex = zalloc(N); S->Ey = zalloc(N); S->Ez = zalloc(N);
    S->Hx = zalloc(N); S->Hy = zalloc(N); S->Hz = zalloc(N);
    S->Jx = zalloc(N); S->Jy = zalloc(N); S->Jz = zalloc(N);

    /* TLS: src/3D_FDTD_version008.cpp:100-172 */
    int ntls = 0;
    for (int q = 0; q < S->nshapes; ++q) if (S->shapes[q].tls >= 0) ++ntls;
    S->tls_on = sp->tls_enabled && ntls > 0;
    if (S->tls_on) {
        S->Ng = zalloc(N); S->Nu = zalloc(N); S->Ng0 = zalloc(N); S->Pz = zalloc(N);
        S->Pz_prev = zalloc(N); S->dPz_dt = zalloc(N); S->Ez_old = zalloc(N); S->Ndip = zalloc(N);
        S->regs = (tls_region_t*)calloc((size_t)ntls, sizeof(tls_region_t));
        S->nreg = 0;
        for (int q = 0; q < S->nshapes; ++q) {
            const shape_t* sh = &S->shapes[q];
            if (sh->tls < 0) continue;
            const fdtd_tls_material* m = &sp->tls_materials[sh->tls];
            aabb_t bb = shape_bbox(sh);
            tls_region_t* R = &S->regs[S->nreg++];
            /* TLSRegionManager::add_region_from_structure, src/two_level_system.hpp:495-521 */
            R->i0 = phys_to_index(S->xb, nxg + 1, S->npml, bb.x0 - pml_x);
            R->i1 = phys_to_index(S->xb, nxg + 1, S->npml, bb.x1 - pml_x) + 1;
            R->j0 = phys_to_index(S->yb, nyg + 1, S->npml, bb.y0 - pml_y);
            R->j1 = phys_to_index(S->yb, nyg + 1, S->npml, bb.y1 - pml_y) + 1;
            R->k0 = phys_to_index(S->zb, nzg + 1, S->npml, bb.z0 - pml_z);
            R->k1 = phys_to_index(S->zb, nzg + 1, S->npml, bb.z1 - pml_z) + 1;
            R->lambda0 = m->lambda0; R->gamma = m->gamma; R->tau = m->tau; R->N0 = m->N0;
            R->inversion_fraction = m->inversion_fraction;
            R->clamp = sp->tls_clamp;
            R->shape = q;
        }
        S->cell_region_map = (int*)malloc(N * sizeof(int));
        for (size_t id = 0; id < N; ++id) S->cell_region_map[id] = -1;
        for (int r = 0; r < S->nreg; ++r) {
            tls_region_t* R = &S->regs[r];
            /* TwoLevelParams::finalize, src/two_level_system.hpp:38-69 */
            R->hbar = HBAR_;
            R->omega_a = 2.0 * PI_ * C0_ / R->lambda0;
            double numerator = 3.0 * PI_ * EPS0_ * R->hbar * C0_ * C0_ * C0_;
            double denominator = R->omega_a * R->omega_a * R->omega_a * R->tau;
            R->mu_z = sqrt(numerator / denominator);
            /* TLSRegion::initialize + map, :454-485, :533-546 */
            if (R->i0 >= R->i1 || R->j0 >= R->j1 || R->k0 >= R->k1) continue;
            const shape_t* sh = &S->shapes[R->shape];
            for (size_t i = R->i0; i < R->i1; ++i)
                for (size_t j = R->j0; j < R->j1; ++j)
                    for (size_t k = R->k0; k < R->k1; ++k) {
                        double x = 0.5 * (S->xb[i] + S->xb[i + 1]);
                        double y = 0.5 * (S->yb[j] + S->yb[j + 1]);
                        double z = 0.5 * (S->zb[k] + S->zb[k + 1]);
                        if (!shape_contains(sh, x, y, z)) continue;
                        size_t id = idx3(i, j, k, NyT, NzT);
                        S->Ndip[id] = R->N0;
                        S->Nu[id] = R->inversion_fraction * R->N0;
                        S->Ng[id] = (1.0 - R->inversion_fraction) * R->N0;
                        S->Ng0[id] = R->N0;
                        S->cell_region_map[id] = r;
                    }
        }
        /* finalize_with_dt, :71-109 */
        for (int r = 0; r < S->nreg; ++r) {
            tls_region_t* R = &S->regs[r];
            if (R->omega_a <= 0.0) continue;
            const double dt = S->dt;
            const double dt2 = dt * dt;
            const double omega_sq = R->omega_a * R->omega_a;
            const double beta = 2.0 + R->gamma * dt;
            const double two_omega_over_hbar = 2.0 * R->omega_a / R->hbar;
            const double mu_z_sq = R->mu_z * R->mu_z;
            R->kapa = (2.0 * dt2 / beta) * two_omega_over_hbar * mu_z_sq;
            R->pa2 = 2.0 * (2.0 - omega_sq * dt2) / beta;
            R->pa3 = (R->gamma * dt - 2.0) / beta;
            R->inv_dt = 1.0 / dt;
            R->init = 1;
        }
    }

    /* sources, src/params.hpp:620-674 (dipoles, then plane waves) */
    S->nsrc = sp->n_dipoles + sp->n_planes;
    S->src = (source_t*)calloc((size_t)(S->nsrc ? S->nsrc : 1), sizeof(source_t));
    int q = 0;
    for (int d = 0; d < sp->n_dipoles; ++d, ++q) {
        const fdtd_dipole_def* D = &sp->dipoles[d];
        source_t* s = &S->src[q];
        /* make_dipole_source, src/sources/dipole_source.hpp:90-153 */
        s->kind = 0;
        s->i = phys_to_index(S->xb, nxg + 1, S->npml, D->x - sp->domain_x_min);
        s->j = phys_to_index(S->yb, nyg + 1, S->npml, D->y - sp->domain_y_min);
        s->k = phys_to_index(S->zb, nzg + 1, S->npml, D->z - sp->domain_z_min);
        s->f0 = src_freq(D->frequency, D->wavelength);
        s->tau = src_tau(D->tau, D->df_fwhm, S->dt);
        s->t0 = (D->t0_factor > 0.0) ? D->t0_factor * s->tau : 3.0 / s->f0;
        s->t_shift = s->t0;
        s->I0 = D->amplitude;
        s->waveform = D->waveform;
        s->polarization = D->polarization;
        switch (D->polarization) {
        case 0: case 3: s->cell_area = S->dy[s->j] * S->dz[s->k]; break;
        case 1: case 4: s->cell_area = S->dx[s->i] * S->dz[s->k]; break;
        default: s->cell_area = S->dx[s->i] * S->dy[s->j]; break;
        }
    }
    for (int d = 0; d < sp->n_planes; ++d, ++q) {
        const fdtd_plane_wave_def* D = &sp->planes[d];
        source_t* s = &S->src[q];
        /* make_plane_wave_source, src/sources/plane_wave_source.hpp:213-310 */
        s->kind = 1;
        s->direction = D->direction;
        s->polarization = D->polarization;
        s->waveform = D->waveform;
        s->f0 = src_freq(D->frequency, D->wavelength);
        s->tau = src_tau(D->tau, D->df_fwhm, S->dt);
        s->t0 = (D->t0_factor > 0.0) ? (D->t0_factor * s->tau) : (3.0 / s->f0);
        s->t_shift = s->t0;
        s->I0 = D->amplitude;
        const size_t np = S->npml;
        size_t pi;
        size_t lim;
        if (D->direction >= 4) { pi = phys_to_index(S->zb, nzg + 1, np, D->injection_position - sp->domain_z_min); lim = NzT - np - 2; }
        else if (D->direction <= 1) { pi = phys_to_index(S->xb, nxg + 1, np, D->injection_position - sp->domain_x_min); lim = NxT - np - 2; }
        else { pi = phys_to_index(S->yb, nyg + 1, np, D->injection_position - sp->domain_y_min); lim = NyT - np - 2; }
        pi = pi < lim ? pi : lim;
        pi = (np + 1 > pi) ? np + 1 : pi;
        s->plane_index = pi;
        /* compute_bounds lambda, :266-290 */
        {
            double mn = D->x_min - sp->domain_x_min, mx = D->x_max - sp->domain_x_min;
            if (mn == 0.0 && mx == 0.0) { s->i_min = np; s->i_max = NxT - np - 1; }
            else {
                s->i_min = phys_to_index(S->xb, nxg + 1, np, mn); s->i_max = phys_to_index(S->xb, nxg + 1, np, mx);
                s->i_min = np > s->i_min ? np : s->i_min; s->i_max = (NxT - np - 1) < s->i_max ? (NxT - np - 1) : s->i_max;
            }
            mn = D->y_min - sp->domain_y_min; mx = D->y_max - sp->domain_y_min;
            if (mn == 0.0 && mx == 0.0) { s->j_min = np; s->j_max = NyT - np - 1; }
            else {
                s->j_min = phys_to_index(S->yb, nyg + 1, np, mn); s->j_max = phys_to_index(S->yb, nyg + 1, np, mx);
                s->j_min = np > s->j_min ? np : s->j_min; s->j_max = (NyT - np - 1) < s->j_max ? (NyT - np - 1) : s->j_max;
            }
            mn = D->z_min - sp->domain_z_min; mx = D->z_max - sp->domain_z_min;
            if (mn == 0.0 && mx == 0.0) { s->k_min = np; s->k_max = NzT - np - 1; }
            else {
                s->k_min = phys_to_index(S->zb, nzg + 1, np, mn); s->k_max = phys_to_index(S->zb, nzg + 1, np, mx);
                s->k_min = np > s->k_min ? np : s->k_min; s->k_max = (NzT - np - 1) < s->k_max ? (NzT - np - 1) : s->k_max;
            }
        }
    }

    /* probes, src/detectors/point_field_detector.hpp:57-70 */
    S->nprobe = sp->n_probes;
    S->probes = (probe_t*)calloc((size_t)(S->nprobe ? S->nprobe : 1), sizeof(probe_t));
    for (int p = 0; p < sp->n_probes; ++p) {
        const fdtd_probe_def* D = &sp->probes[p];
        probe_t* pr = &S->probes[p];
        size_t i0 = phys_to_index(S->xb, nxg + 1, S->npml, D->x - sp->domain_x_min);
        size_t j0 = phys_to_index(S->yb, nyg + 1, S->npml, D->y - sp->domain_y_min);
        size_t k0 = phys_to_index(S->zb, nzg + 1, S->npml, D->z - sp->domain_z_min);
        i0 = i0 < NxT - 2 ? i0 : NxT - 2; i0 = i0 > 1 ? i0 : 1;
        j0 = j0 < NyT - 2 ? j0 : NyT - 2; j0 = j0 > 1 ? j0 : 1;
        k0 = k0 < NzT - 2 ? k0 : NzT - 2; k0 = k0 > 1 ? k0 : 1;
        pr->i0 = i0; pr->j0 = j0; pr->k0 = k0;
        pr->nc = D->n_components;
        for (int c = 0; c < D->n_components; ++c) pr->comps[c] = D->components[c];
    }
    S->save_every = sp->save_every > 0 ? (size_t)sp->save_every : 1;

    for (int m = 0; m < sp->n_dfts; ++m) { /* reject monitors outside the grid */
        const fdtd_dft_def* D = &sp->dfts[m];
        if (D->i0 < 0 || D->i1 > (int)NxT || D->i0 >= D->i1 || D->j0 < 0 || D->j1 > (int)NyT || D->j0 >= D->j1 ||
            D->k0 < 0 || D->k1 > (int)NzT || D->k0 >= D->k1 || D->n_freqs < 1 || D->n_freqs > 8 ||
            D->component < 0 || D->component > 5) {
            oracle_destroy(S);
            return NULL;
        }
    }
    S->ndft = sp->n_dfts;
    S->dfts = (dft_t*)calloc((size_t)(S->ndft ? S->ndft : 1), sizeof(dft_t));
    for (int m = 0; m < S->ndft; ++m) {
        S->dfts[m].def = sp->dfts[m];
        const fdtd_dft_def* D = &sp->dfts[m];
        S->dfts[m].ncell = (size_t)(D->i1 - D->i0) * (size_t)(D->j1 - D->j0) * (size_t)(D->k1 - D->k0);
        S->dfts[m].acc = zalloc(S->dfts[m].ncell * (size_t)D->n_freqs * 2);
    }
    S->stab_interval = sp->stability_interval > 0 ? (size_t)sp->stability_interval : 10;
    S->st_stable = 1;
    S->st_fail_step = -1;
    return S;
}

/* One serial-mode step, src/3D_FDTD_version008.cpp:269-349 */
/* ---- Storage-format emulation (test infrastructure, no reference
 * counterpart).  The device's FP16 form stores every field and psi value as
 * an IEEE binary16 number times a per-array power-of-two scale and computes
 * in FP32; its populations and polarization stay FP64.  This emulates that
 * storage on the FP64 oracle: after the H update with its CPML correction,
 * and after the E update with its CPML correction (where the device stores
 * each array once per step), every H/psi_H (then E/psi_E) value is rounded
 * as the store would round it.  The compute stays FP64 (FP32 rounding,
 * 2^-24, is far below the 2^-12 of the store).  Formats:
 *   1: binary16 with scale (11 significant bits, round to nearest even,
 *      subnormal below 2^-14 x scale, saturation at 65504 x scale);
 *   2: the same with stochastic rounding (seeded xorshift);
 *   3: binary16 plus a residual byte (19 significant bits, same range);
 *   4: binary16 fields, psi rounded to binary32;
 *   5, 6, 7: block floating point: a 16-bit signed mantissa per value and
 *      one power-of-two exponent per block of 32, 64 or 8 consecutive k
 *      cells of a row (the exponent puts the block's largest magnitude in
 *      [2^14, 2^15); round to nearest even; no per-array scale).  A value
 *      a factor f below its block's largest keeps about 15 - log2(f)
 *      significant bits.
 * Scales start from `scales18` (fields Ex..Hz, psi Ex_y..Hz_y, the values the
 * device starts from) and follow the device's plan every 10 steps: when the
 * largest stored magnitude leaves [2^2, 2^11] the scale moves so that it
 * lands in [2^7, 2^8), never below 2^-6 of its start. */
static inline uint64_t dbits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }
static inline double bitsd(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }
static inline uint64_t xorshift64(uint64_t* s) {
    uint64_t x = *s;
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    *s = x;
    return x;
}
/* y rounded to `bits` significant bits (normal range of binary16 and below:
 * a fixed quantum 2^-24 x 2^-(bits-11)), saturated at the binary16 maximum. */
static double round_store(double y, int bits, int stoch, uint64_t* rng) {
    if (!(y == y) || y == 0.0) return y;
    const double ay = fabs(y);
    if (ay >= 65504.0 * 2.0) return copysign(65504.0, y); /* far beyond: saturate */
    if (ay < 6.103515625e-05) { /* below 2^-14: binary16 subnormal spacing */
        const double q = ldexp(1.0, -24 - (bits - 11));
        double t = y / q, r;
        if (stoch) {
            const double fl = floor(t);
            r = ((double)(xorshift64(rng) >> 11) * 0x1.0p-53 < t - fl) ? fl + 1.0 : fl;
        } else {
            r = nearbyint(t);
        }
        return r * q;
    }
    const int drop = 52 - (bits - 1);
    uint64_t b = dbits(y);
    const uint64_t mask = ((uint64_t)1 << drop) - 1;
    if (stoch) {
        b += xorshift64(rng) & mask;
    } else {
        const uint64_t lsb = (b >> drop) & 1;
        b += ((uint64_t)1 << (drop - 1)) - 1 + lsb;
    }
    b &= ~mask;
    double r = bitsd(b);
    if (fabs(r) > 65504.0) r = copysign(65504.0, r);
    return r;
}
/* Block floating point (formats 5-7): each block of B consecutive k of a
 * row shares the exponent that puts its largest magnitude in [2^14, 2^15)
 * (the next one up where that largest value would round to 2^15, so every
 * mantissa fits a signed 16-bit integer); values become round-to-nearest-
 * even multiples of the block's quantum.  rows x nz values, row-major. */
static void round_blocks(double* a, long rows, int nz, int B) {
#pragma omp parallel for
    for (long r = 0; r < rows; ++r) {
        double* row = a + r * nz;
        for (int k0 = 0; k0 < nz; k0 += B) {
            const int k1 = k0 + B < nz ? k0 + B : nz;
            double m = 0.0;
            for (int k = k0; k < k1; ++k) m = fmax(m, fabs(row[k]));
            if (!(m > 0.0) || !(m == m) || m > 1e300) continue;  /* zeros, NaN/Inf: unchanged */
            int e;
            frexp(m, &e);                        /* m in [2^(e-1), 2^e) */
            double q = ldexp(1.0, e - 15);       /* m / q in [2^14, 2^15) */
            if (nearbyint(m / q) > 32767.0) q *= 2.0;
            for (int k = k0; k < k1; ++k) row[k] = nearbyint(row[k] / q) * q;
        }
    }
}
static void emu_round_blocks(oracle_sim* S, double* a) {
    round_blocks(a, (long)(S->N / S->NzT), (int)S->NzT, S->emu_block);
}
/* The block rounding on its own (for tests). */
void oracle_round_blocks(double* a, long rows, int nz, int B) { round_blocks(a, rows, nz, B); }
static void emu_round(oracle_sim* S, double* a, int arr) {
    if (S->emu_block) {
        emu_round_blocks(S, a);
        return;
    }
    const double sc = S->emu_sc[arr];
    const int bits = arr >= 6 ? S->emu_psi_bits : S->emu_bits;
    const long n = (long)S->N;
    if (bits >= 24) { /* binary32 */
#pragma omp parallel for
        for (long i = 0; i < n; ++i) a[i] = (double)(float)a[i];
        return;
    }
    if (S->emu_stoch) { /* one random stream, in cell order */
        for (long i = 0; i < n; ++i)
            if (a[i] != 0.0) a[i] = round_store(a[i] / sc, bits, 1, &S->emu_rng) * sc;
        return;
    }
#pragma omp parallel for
    for (long i = 0; i < n; ++i)
        if (a[i] != 0.0) a[i] = round_store(a[i] / sc, bits, 0, NULL) * sc;
}
/* psi index (0..11) of the device's scale table for oracle psi array q */
static void emu_store_H(oracle_sim* S) {
    emu_round(S, S->Hx, 3); emu_round(S, S->Hy, 4); emu_round(S, S->Hz, 5);
    if (S->bc_type == FDTD_BC_CPML)
        for (int q = 6; q < 12; ++q) emu_round(S, S->psi[q], 6 + q);
}
static void emu_store_E(oracle_sim* S) {
    emu_round(S, S->Ex, 0); emu_round(S, S->Ey, 1); emu_round(S, S->Ez, 2);
    if (S->bc_type == FDTD_BC_CPML)
        for (int q = 0; q < 6; ++q) emu_round(S, S->psi[q], 6 + q);
}
static void emu_plan(oracle_sim* S) {
    double* arr[18] = {S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz};
    for (int q = 0; q < 12; ++q) arr[6 + q] = S->psi[q];
    for (int c = 0; c < 18; ++c) {
        if (!arr[c] || (c >= 6 && S->bc_type != FDTD_BC_CPML)) continue;
        double m = 0.0;
        for (size_t i = 0; i < S->N; ++i) m = fmax(m, fabs(arr[c][i]));
        m /= S->emu_sc[c];
        if (m > 2048.0 || (m > 0.0 && m < 4.0)) {
            int e;
            frexp(m, &e);
            int shift = e - 8;
            while (shift < 0 && ldexp(S->emu_sc[c], shift) < S->emu_floor[c]) ++shift;
            S->emu_sc[c] = ldexp(S->emu_sc[c], shift);
        }
    }
}

/* One emulated store of a value already divided by its scale (for tests). */
double oracle_round_store(double y, int bits, int stochastic) {
    uint64_t rng = 0x9e3779b97f4a7c15ull;
    return round_store(y, bits, stochastic, &rng);
}

/* fmt 0 turns the emulation off. */
int oracle_set_storage(oracle_sim* S, int fmt, const double* scales18, unsigned long long seed) {
    if (fmt < 0 || fmt > 7 || (fmt && !scales18)) return -1;
    S->emu = fmt;
    S->emu_block = fmt == 5 ? 32 : fmt == 6 ? 64 : fmt == 7 ? 8 : 0;
    S->emu_bits = fmt == 3 ? 19 : 11;
    S->emu_psi_bits = fmt == 3 ? 19 : fmt == 4 ? 24 : 11;
    S->emu_stoch = fmt == 2;
    S->emu_rng = seed ? seed : 0x9e3779b97f4a7c15ull;
    for (int q = 0; fmt && q < 18; ++q) {
        S->emu_sc[q] = scales18[q];
        S->emu_floor[q] = ldexp(scales18[q], -6);
    }
    return 0;
}

static int one_step(oracle_sim* S) {
    const size_t n = S->n;
    memset(S->Jx, 0, S->N * sizeof(double));
    memset(S->Jy, 0, S->N * sizeof(double));
    memset(S->Jz, 0, S->N * sizeof(double));
    inject(S, n);
    if (S->tls_on) update_polarization(S);
    update_H(S);
    if (S->bc_type == FDTD_BC_CPML) cpml_after_H(S);
    if (S->emu) emu_store_H(S);
    update_E(S, S->tls_on);
    if (S->bc_type == FDTD_BC_CPML) cpml_after_E(S); else pec_after_E(S);
    if (S->emu) emu_store_E(S);
    record_probes(S, n);
    accumulate_dft(S, n);
    if (S->tls_on) update_populations(S);
    S->n++;
    if (S->emu && n % 10 == 0) emu_plan(S);
    if (n % S->stab_interval == 0) {
        double cons = conservation_error(S);
        S->st_last_cons = cons;
        if (!check_stability(S, n, cons)) { S->st_fail_step = (long)n; return 1; }
    }
    return 0;
}

int oracle_step(oracle_sim* S, int nsteps) {
    for (int s = 0; s < nsteps; ++s)
        if (one_step(S)) return 1;
    return 0;
}

/* Time `nsteps` steps of the loop body (no setup), returns seconds. */
double oracle_time_steps(oracle_sim* S, int nsteps) {
#ifdef _OPENMP
    double t0 = omp_get_wtime();
    oracle_step(S, nsteps);
    return omp_get_wtime() - t0;
#else
    (void)S; (void)nsteps;
    return -1.0;
#endif
}

int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_destroy(oracle_sim* S) {
    if (!S) return;
    double* arrs[] = {S->dx, S->dy, S->dz, S->inv_dx, S->inv_dy, S->inv_dz, S->xb, S->yb, S->zb,
                      S->kx, S->ky, S->kz, S->sx, S->sy, S->sz, S->ax, S->ay, S->az,
                      S->cbEx, S->ccEx, S->cbEy, S->ccEy, S->cbEz, S->ccEz, S->cbHx, S->ccHx, S->cbHy, S->ccHy, S->cbHz, S->ccHz,
                      S->aEx, S->bEx, S->aEy, S->bEy, S->aEz, S->bEz, S->aHx, S->bHx, S->aHy, S->bHy, S->aHz, S->bHz,
                      S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz, S->Jx, S->Jy, S->Jz,
                      S->Ng, S->Nu, S->Ng0, S->Pz, S->Pz_prev, S->dPz_dt, S->Ez_old, S->Ndip};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); ++q) free(arrs[q]);
    for (int q = 0; q < 12; ++q) free(S->psi[q]);
    for (int p = 0; p < S->nprobe; ++p) free(S->probes[p].rec);
    for (int m = 0; m < S->ndft; ++m) free(S->dfts[m].acc);
    free(S->probes); free(S->dfts); free(S->src); free(S->regs); free(S->shapes); free(S->cell_region_map);
    free(S);
}

/* ---- accessors used by the tests ---- */
void oracle_info(const oracle_sim* S, long* out) {
    out[0] = (long)S->NxT; out[1] = (long)S->NyT; out[2] = (long)S->NzT; out[3] = (long)S->npml;
    out[4] = (long)S->n; out[5] = S->tls_on; out[6] = S->nreg; out[7] = S->nsrc;
    out[8] = S->st_stable; out[9] = S->st_fail_step; out[10] = (long)S->st_count;
    out[11] = (long)S->nx_grid; out[12] = (long)S->ny_grid; out[13] = (long)S->nz_grid;
}
double oracle_dt(const oracle_sim* S) { return S->dt; }
double oracle_last_max_E(const oracle_sim* S) { return S->st_last_max_E; }
double oracle_last_conservation(const oracle_sim* S) { return S->st_last_cons; }

static const char* const k_names[] = {
    "Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Jx", "Jy", "Jz",
    "aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz", "bHz",
    "Ng", "Nu", "Ng0", "Pz", "Pz_prev", "dPz_dt", "Ez_old", "Ndip",
    "psi_Ex_y", "psi_Ex_z", "psi_Ey_z", "psi_Ey_x", "psi_Ez_x", "psi_Ez_y",
    "psi_Hx_y", "psi_Hx_z", "psi_Hy_z", "psi_Hy_x", "psi_Hz_x", "psi_Hz_y",
    "dx", "dy", "dz", "inv_dx", "inv_dy", "inv_dz", "x_bounds", "y_bounds", "z_bounds",
    "kx", "ky", "kz", "cpml_bEx", "cpml_cEx", "cpml_bEy", "cpml_cEy", "cpml_bEz", "cpml_cEz",
    "cpml_bHx", "cpml_cHx", "cpml_bHy", "cpml_cHy", "cpml_bHz", "cpml_cHz", NULL};

/* Copies the named array into out (if non-NULL) and returns its length, or -1. */
long oracle_get_array(const oracle_sim* S, const char* name, double* out) {
    const double* p = NULL;
    size_t n = S->N;
    int q = -1;
    for (int t = 0; k_names[t]; ++t) if (strcmp(k_names[t], name) == 0) { q = t; break; }
    if (q < 0) return -1;
    const double* tab[] = {S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz, S->Jx, S->Jy, S->Jz,
                           S->aEx, S->bEx, S->aEy, S->bEy, S->aEz, S->bEz, S->aHx, S->bHx, S->aHy, S->bHy, S->aHz, S->bHz,
                           S->Ng, S->Nu, S->Ng0, S->Pz, S->Pz_prev, S->dPz_dt, S->Ez_old, S->Ndip};
    if (q < 29) { p = tab[q]; }
    else if (q < 41) { p = S->psi[q - 29]; }
    else {
        switch (q) {
        case 41: p = S->dx; n = S->nx_grid; break;
        case 42: p = S->dy; n = S->ny_grid; break;
        case 43: p = S->dz; n = S->nz_grid; break;
        case 44: p = S->inv_dx; n = S->nx_grid; break;
        case 45: p = S->inv_dy; n = S->ny_grid; break;
        case 46: p = S->inv_dz; n = S->nz_grid; break;
        case 47: p = S->xb; n = S->nx_grid + 1; break;
        case 48: p = S->yb; n = S->ny_grid + 1; break;
        case 49: p = S->zb; n = S->nz_grid + 1; break;
        case 50: p = S->kx; n = S->NxT; break;
        case 51: p = S->ky; n = S->NyT; break;
        case 52: p = S->kz; n = S->NzT; break;
        case 53: p = S->cbEx; n = S->NxT; break;
        case 54: p = S->ccEx; n = S->NxT; break;
        case 55: p = S->cbEy; n = S->NyT; break;
        case 56: p = S->ccEy; n = S->NyT; break;
        case 57: p = S->cbEz; n = S->NzT; break;
        case 58: p = S->ccEz; n = S->NzT; break;
        case 59: p = S->cbHx; n = S->NxT; break;
        case 60: p = S->ccHx; n = S->NxT; break;
        case 61: p = S->cbHy; n = S->NyT; break;
        case 62: p = S->ccHy; n = S->NyT; break;
        case 63: p = S->cbHz; n = S->NzT; break;
        case 64: p = S->ccHz; n = S->NzT; break;
        }
    }
    if (!p) return -1;
    if (out) memcpy(out, p, n * sizeof(double));
    return (long)n;
}

/* Region map as doubles (-1 outside gain). */
long oracle_get_region_map(const oracle_sim* S, int* out) {
    if (!S->tls_on) return 0;
    if (out) memcpy(out, S->cell_region_map, S->N * sizeof(int));
    return (long)S->N;
}

/* TLS coefficients of region r: lambda0 omega_a gamma tau mu_z N0 hbar kapa pa2 pa3 inv_dt i0 i1 j0 j1 k0 k1 */
int oracle_get_region(const oracle_sim* S, int r, double* out) {
    if (r < 0 || r >= S->nreg) return -1;
    const tls_region_t* R = &S->regs[r];
    double v[] = {R->lambda0, R->omega_a, R->gamma, R->tau, R->mu_z, R->N0, R->hbar, R->kapa, R->pa2, R->pa3, R->inv_dt,
                  (double)R->i0, (double)R->i1, (double)R->j0, (double)R->j1, (double)R->k0, (double)R->k1};
    memcpy(out, v, sizeof(v));
    return 0;
}

/* Source q: kind comp i j k (or plane box) cell_area and its J value at step n. */
int oracle_get_source(const oracle_sim* S, int q, long n, double* out) {
    if (q < 0 || q >= S->nsrc) return -1;
    const source_t* s = &S->src[q];
    out[0] = s->kind; out[1] = s->polarization; out[2] = s->direction;
    out[3] = (double)s->i; out[4] = (double)s->j; out[5] = (double)s->k; out[6] = s->cell_area;
    out[7] = (double)s->plane_index; out[8] = (double)s->i_min; out[9] = (double)s->i_max;
    out[10] = (double)s->j_min; out[11] = (double)s->j_max; out[12] = (double)s->k_min; out[13] = (double)s->k_max;
    out[14] = source_value(s, (size_t)n, S->dt);
    if (s->kind == 1) { /* plane: the per-cell J of inject_*_plane */
        const int dir = s->direction;
        const double d = (dir >= 4) ? S->dz[s->plane_index] : ((dir <= 1) ? S->dx[s->plane_index] : S->dy[s->plane_index]);
        const double sign = (dir == 0 || dir == 2 || dir == 4) ? 1.0 : -1.0;
        out[14] = 2.0 * out[14] / (ETA0_ * d) * sign;
    }
    out[15] = s->f0; out[16] = s->tau; out[17] = s->t0; out[18] = s->I0;
    return 0;
}

long oracle_probe_records(const oracle_sim* S, int p) { return (p >= 0 && p < S->nprobe) ? (long)S->probes[p].nrec : -1; }
int oracle_probe_index(const oracle_sim* S, int p, long* ijk) {
    if (p < 0 || p >= S->nprobe) return -1;
    ijk[0] = (long)S->probes[p].i0; ijk[1] = (long)S->probes[p].j0; ijk[2] = (long)S->probes[p].k0;
    return S->probes[p].nc;
}
/* Copies component c of probe p (all records) into out. */
long oracle_get_probe(const oracle_sim* S, int p, int c, double* out) {
    if (p < 0 || p >= S->nprobe) return -1;
    const probe_t* pr = &S->probes[p];
    if (c < 0 || c >= pr->nc) return -1;
    for (size_t r = 0; r < pr->nrec; ++r) out[r] = pr->rec[r * (size_t)pr->nc + (size_t)c];
    return (long)pr->nrec;
}
/* DFT accumulators of monitor m, layout [freq][cell][re,im]. */
long oracle_get_dft(const oracle_sim* S, int m, double* out) {
    if (m < 0 || m >= S->ndft) return -1;
    const dft_t* d = &S->dfts[m];
    size_t n = d->ncell * (size_t)d->def.n_freqs * 2;
    if (out) memcpy(out, d->acc, n * sizeof(double));
    return (long)n;
}

/* Overwrite a field (for tests that seed initial conditions). */
int oracle_set_array(oracle_sim* S, const char* name, const double* in) {
    long n = oracle_get_array(S, name, NULL);
    if (n < 0) return -1;
    const char* fields[] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    double* ptrs[] = {S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz};
    for (int q = 0; q < 6; ++q)
        if (strcmp(fields[q], name) == 0) { memcpy(ptrs[q], in, (size_t)n * sizeof(double)); return 0; }
    return -1;
}

// fdtd_setup.cpp - host setup of the Yee-step engine (include/fdtd_scene.h).
//
// Restates, on the host, the reference's setup path that produces the inputs
// of the per-timestep update: grid spacing and dt (src/params.hpp:180-296),
// CPML profiles (src/boundary.hpp:105-136, 303-373), the structure package
// and 2x2x2 subpixel bake (src/params.hpp:557-592,
// src/structure_material.hpp:211-288), the TLS regions and coefficients
// (src/3D_FDTD_version008.cpp:100-172, src/two_level_system.hpp:17-110,
// 421-571), the source package (src/params.hpp:594-678,
// src/sources/*.hpp) and point-probe placement
// (src/detectors/point_field_detector.hpp:57-70).  The arithmetic keeps the
// reference's operation order (host libm, no contraction) so the arrays are
// bitwise the reference's; the bake is split over host threads by i-plane and
// skips shapes whose (slightly widened) bounding box excludes a sample, which
// cannot change which shape last contains it.
#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fdtd_hip.h"
#include "../../include/fdtd_scene.h"

namespace fdtd_host {

constexpr double PI = 3.14159265358979323846;
constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.854187817e-12;
constexpr double MU0 = 4.0 * PI * 1e-7;   // PhysConst::MU0, src/global_function.hpp:16
constexpr double HBAR = 1.054571817e-34;
constexpr double ETA0 = 376.730313668;    // src/sources/plane_wave_source.hpp:42

// GridSpacing, src/global_function.hpp:92-158
struct GridSpacing {
    std::vector<double> dx, dy, dz, inv_dx, inv_dy, inv_dz, x_bounds, y_bounds, z_bounds;
    size_t npml = 0;
    static std::vector<double> bounds(const std::vector<double>& s) {
        std::vector<double> b(s.size() + 1);
        b[0] = 0.0;
        for (size_t i = 0; i < s.size(); ++i) b[i + 1] = b[i] + s[i];
        return b;
    }
    static size_t to_index(const std::vector<double>& b, size_t npml, double phys) {
        const double off = (npml < b.size()) ? b[npml] : 0.0;
        const double total = phys + off;
        if (total <= b.front()) return 0;
        if (total >= b.back()) return b.size() - 2;
        return (size_t)(std::upper_bound(b.begin(), b.end(), total) - b.begin()) - 1;
    }
    size_t physical_to_index_x(double v) const { return to_index(x_bounds, npml, v); }
    size_t physical_to_index_y(double v) const { return to_index(y_bounds, npml, v); }
    size_t physical_to_index_z(double v) const { return to_index(z_bounds, npml, v); }
};

struct Shape {
    int type;
    double p[6];
    double eps_r, mu_r, sigma, sigma_m;
    int tls;
    double bb[6];  // widened bounding box for sample culling
    bool contains(double x, double y, double z) const {
        switch (type) {
        case FDTD_SHAPE_BOX:
            return (x >= p[0] && x < p[1] && y >= p[2] && y < p[3] && z >= p[4] && z < p[5]);
        case FDTD_SHAPE_SPHERE: {
            const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
            return (dx * dx + dy * dy + dz * dz) <= p[3] * p[3];
        }
        default: {
            if (z < p[3] || z >= p[4]) return false;
            const double dx = x - p[0], dy = y - p[1];
            return (dx * dx + dy * dy) <= p[2] * p[2];
        }
        }
    }
    void bounding_box(double* b) const {
        if (type == FDTD_SHAPE_BOX) {
            for (int q = 0; q < 6; ++q) b[q] = p[q];
        } else if (type == FDTD_SHAPE_SPHERE) {
            b[0] = p[0] - p[3]; b[1] = p[0] + p[3]; b[2] = p[1] - p[3];
            b[3] = p[1] + p[3]; b[4] = p[2] - p[3]; b[5] = p[2] + p[3];
        } else {
            b[0] = p[0] - p[2]; b[1] = p[0] + p[2]; b[2] = p[1] - p[2];
            b[3] = p[1] + p[2]; b[4] = p[3]; b[5] = p[4];
        }
    }
};

struct Region {
    double lambda0, omega_a, gamma, tau, mu_z, N0, hbar, kapa, pa2, pa3, inv_dt, inversion_fraction;
    int clamp;
    size_t i0, i1, j0, j1, k0, k1;
    int shape;
};

struct Source {
    int kind;  // 0 dipole, 1 plane
    int waveform, polarization, direction;
    double I0, f0, tau, t0, t_shift, cell_area, plane_d;
    size_t i, j, k, plane_index, i_min, i_max, j_min, j_max, k_min, k_max;
};

struct ProbeChan {
    size_t i, j, k;
    int comp;
};

// compute_waveform, src/sources/isource.hpp:41-68
double compute_waveform(int type, double t, double I0, double f0, double tau, double t0, double t_shift) {
    const double pi = PI;
    switch (type) {
    case FDTD_WF_RICKER: {
        const double a = pi * f0 * (t - t0);
        const double a2 = a * a;
        return I0 * (1 - 2 * a2) * std::exp(-a2);
    }
    case FDTD_WF_GAUSS_MOD_SINE: {
        const double x = (t - t0);
        return I0 * std::sin(2 * pi * f0 * (t - t_shift)) * std::exp(-(x * x) / (tau * tau));
    }
    case FDTD_WF_RICKER_GAUSS2: {
        const double x = (t - t0) / tau;
        const double x2 = x * x;
        return I0 * (1 - 2 * x2) * std::exp(-x2);
    }
    case FDTD_WF_CW: {
        const double envelope = (t < tau) ? (0.5 * (1 - std::cos(pi * t / tau))) : 1.0;
        return I0 * envelope * std::sin(2 * pi * f0 * t);
    }
    }
    return 0;
}

}  // namespace fdtd_host

using namespace fdtd_host;

struct fdtd_setup {
    size_t NxT = 0, NyT = 0, NzT = 0, N = 0, npml = 0;
    int bc_type = 1;
    double dt = 0;
    GridSpacing gs;
    // CPMLBoundary 1D profiles
    std::vector<double> kx, ky, kz, sx, sy, sz, ax, ay, az;
    std::vector<double> bEx, cEx, bEy, cEy, bEz, cEz, bHx, cHx, bHy, cHy, bHz, cHz;
    // MaterialGrids
    std::vector<double> m[12];  // aEx bEx aEy bEy aEz bEz aHx bHx aHy bHy aHz bHz
    std::vector<Shape> shapes;
    double bg_eps = 1, bg_mu = 1, bg_sig = 0, bg_sigm = 0;
    int subpixel = 8;
    bool device_bake = false;  // bake on the device at scene creation
    // TLS
    bool tls_on = false;
    std::vector<Region> regs;
    std::vector<int> region_map;
    std::vector<double> Ndip, Ng0, Nu, Ng, Pz, Pz_prev;
    // sources, probes, monitors
    std::vector<Source> src;
    std::vector<ProbeChan> probes;
    int save_every = 1;
    std::vector<fdtd_dft_def> dfts;
    int stab_interval = 10;
};

namespace {

// CPMLBoundary::build_profiles_and_coeffs, src/boundary.hpp:303-373
void build_cpml(fdtd_setup* S, const fdtd_scene_spec* sp) {
    const int np = (int)S->npml;
    auto avg = [np](const std::vector<double>& s, int N) {
        double sum = 0;
        int count = 0;
        for (int i = 0; i < np; ++i) {
            sum += s[i] + s[N - 1 - i];
            count += 2;
        }
        return sum / count;
    };
    auto sigma_max = [&](double ds) {
        const double d = np * ds;
        return ((sp->cpml_m + 1.0) * std::log(1.0 / sp->cpml_rerr) * EPS0 * C0) / (2.0 * d);
    };
    auto fill_1d = [&](int N, double smax, const std::vector<double>& sp_arr, std::vector<double>& s,
                       std::vector<double>& k, std::vector<double>& a) {
        for (int n = 0; n < N; ++n) {
            const int d = std::min(n, N - 1 - n);
            if (d >= np) { s[n] = 0.0; k[n] = 1.0; a[n] = 0.0; continue; }
            const double r = double(np - (d + 0.5)) / double(np);
            const double rm = std::pow(r, sp->cpml_m);
            const double kap = 1.0 + (sp->cpml_kappa_max - 1.0) * rm;
            const double sig = smax * kap * rm;
            const double ds_local = sp_arr[n];
            const double alf = sp->cpml_alpha_linear ? (sp->cpml_alpha0 * (C0 / ds_local) * (1.0 - r))
                                                     : (sp->cpml_alpha0 * (C0 / ds_local) * std::pow(1.0 - r, sp->cpml_m));
            s[n] = sig; k[n] = kap; a[n] = alf;
        }
    };
    auto build_bc = [&](const std::vector<double>& s, const std::vector<double>& k, const std::vector<double>& a,
                        double norm, std::vector<double>& b, std::vector<double>& c) {
        for (size_t n = 0; n < s.size(); ++n) {
            const double sig_n = s[n] / norm;
            const double kap = k[n];
            const double alf = a[n];
            double be = std::exp(-(sig_n / kap + alf) * S->dt);
            be = std::max<double>(be, 1e-30);
            double ce = 0.0;
            if (std::abs(sig_n) > 1e-30 || std::abs(alf) > 1e-30) ce = (sig_n * (be - 1.0)) / (kap * (sig_n + kap * alf));
            b[n] = be; c[n] = ce;
        }
    };
    const double dxa = avg(S->gs.dx, (int)S->NxT), dya = avg(S->gs.dy, (int)S->NyT), dza = avg(S->gs.dz, (int)S->NzT);
    fill_1d((int)S->NxT, sigma_max(dxa), S->gs.dx, S->sx, S->kx, S->ax);
    fill_1d((int)S->NyT, sigma_max(dya), S->gs.dy, S->sy, S->ky, S->ay);
    fill_1d((int)S->NzT, sigma_max(dza), S->gs.dz, S->sz, S->kz, S->az);
    build_bc(S->sx, S->kx, S->ax, EPS0, S->bEx, S->cEx);
    build_bc(S->sy, S->ky, S->ay, EPS0, S->bEy, S->cEy);
    build_bc(S->sz, S->kz, S->az, EPS0, S->bEz, S->cEz);
    build_bc(S->sx, S->kx, S->ax, MU0, S->bHx, S->cHx);
    build_bc(S->sy, S->ky, S->ay, MU0, S->bHy, S->cHy);
    build_bc(S->sz, S->kz, S->az, MU0, S->bHz, S->cHz);
}

// StructureScene::bake, src/structure_material.hpp:211-288
void bake(fdtd_setup* S, int subpixel_samples) {
    const size_t N = S->N;
    for (int q = 0; q < 12; ++q) S->m[q].assign(N, (q % 2 == 0) ? 1.0 : 0.0);
    int ns = (int)std::round(std::cbrt((double)subpixel_samples));
    if (ns < 1) ns = 1;
    const double dt = S->dt;
    const auto& g = S->gs;
    static const double off[6][3] = {{0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0},
                                     {0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}};
    auto worker = [&](size_t ib, size_t ie) {
        std::vector<int> cx, cxy;
        for (size_t i = ib; i < ie; ++i) {
            // shapes whose widened x-extent meets the samples of plane i
            const double xlo = g.x_bounds[i], xhi = g.x_bounds[i] + 2.0 * g.dx[i];
            cx.clear();
            for (int q = 0; q < (int)S->shapes.size(); ++q)
                if (S->shapes[q].bb[1] >= xlo && S->shapes[q].bb[0] <= xhi) cx.push_back(q);
            for (size_t j = 0; j < S->NyT; ++j) {
                const double ylo = g.y_bounds[j], yhi = g.y_bounds[j] + 2.0 * g.dy[j];
                cxy.clear();
                for (int q : cx)
                    if (S->shapes[q].bb[3] >= ylo && S->shapes[q].bb[2] <= yhi) cxy.push_back(q);
                for (size_t k = 0; k < S->NzT; ++k) {
                    const size_t id = (i * S->NyT + j) * S->NzT + k;
                    for (int c = 0; c < 6; ++c) {
                        const double dxl = g.dx[i], dyl = g.dy[j], dzl = g.dz[k];
                        const double x0 = g.x_bounds[i], y0 = g.y_bounds[j], z0 = g.z_bounds[k];
                        double eps_sum = 0, mu_sum = 0, sg_sum = 0, sgm_sum = 0;
                        int cnt = 0;
                        for (int a = 0; a < ns; ++a)
                            for (int b = 0; b < ns; ++b)
                                for (int cc = 0; cc < ns; ++cc) {
                                    const double xs = x0 + (off[c][0] + (a + 0.5) / ns) * dxl;
                                    const double ys = y0 + (off[c][1] + (b + 0.5) / ns) * dyl;
                                    const double zs = z0 + (off[c][2] + (cc + 0.5) / ns) * dzl;
                                    double e = S->bg_eps, mu = S->bg_mu, sg = S->bg_sig, sgm = S->bg_sigm;
                                    for (int q : cxy) {
                                        const Shape& sh = S->shapes[q];
                                        if (zs < sh.bb[4] || zs > sh.bb[5]) continue;
                                        if (sh.contains(xs, ys, zs)) { e = sh.eps_r; mu = sh.mu_r; sg = sh.sigma; sgm = sh.sigma_m; }
                                    }
                                    eps_sum += e * EPS0;
                                    mu_sum += mu * MU0;
                                    sg_sum += sg;
                                    sgm_sum += sgm;
                                    ++cnt;
                                }
                        const double eps = eps_sum / cnt, mmu = mu_sum / cnt, sigma = sg_sum / cnt, sigm = sgm_sum / cnt;
                        if (c < 3) {
                            const double c1 = sigma * dt / (2 * eps);
                            S->m[2 * c][id] = (1 - c1) / (1 + c1);
                            S->m[2 * c + 1][id] = (dt / eps) / (1 + c1);
                        } else {
                            const double c2 = sigm * dt / (2 * mmu);
                            S->m[2 * c][id] = (1 - c2) / (1 + c2);
                            S->m[2 * c + 1][id] = (dt / mmu) / (1 + c2);
                        }
                    }
                }
            }
        }
    };
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (S->NxT < 2 * nt) nt = 1;
    std::vector<std::thread> th;
    const size_t chunk = (S->NxT + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t b = t * chunk, e = std::min(S->NxT, b + chunk);
        if (b < e) th.emplace_back(worker, b, e);
    }
    for (auto& t : th) t.join();
}

double src_freq(double frequency, double wavelength) {
    if (frequency > 0.0) return frequency;
    if (wavelength > 0.0) return C0 / wavelength;
    return C0 / 500e-9;
}
double src_tau(double tau, double df, double dt) {
    if (tau > 0.0) return tau;
    if (df > 0.0) return 2.0 * std::sqrt(std::log(2.0)) / (PI * df);  // tau_from_bandwidth, isource.hpp:70-73
    return 20.0 * dt;
}

}  // namespace

extern "C" {

int fdtd_setup_build(const fdtd_scene_spec* sp, fdtd_setup** out) {
    if (!sp || !out) return FDTD_EINVAL;
    *out = nullptr;
    if (sp->nx_core < 1 || sp->ny_core < 1 || sp->nz_core < 1 || sp->npml < 0) return FDTD_EINVAL;
    auto* S = new fdtd_setup();
    const size_t np_spec = (size_t)sp->npml;
    const size_t nxg = (size_t)sp->nx_core + 2 * np_spec, nyg = (size_t)sp->ny_core + 2 * np_spec,
                 nzg = (size_t)sp->nz_core + 2 * np_spec;
    auto& g = S->gs;
    if (sp->dx && sp->dy && sp->dz) {
        g.dx.assign(sp->dx, sp->dx + nxg);
        g.dy.assign(sp->dy, sp->dy + nyg);
        g.dz.assign(sp->dz, sp->dz + nzg);
    } else {
        g.dx.assign(nxg, sp->dx_uniform);
        g.dy.assign(nyg, sp->dy_uniform);
        g.dz.assign(nzg, sp->dz_uniform);
    }
    const bool lines = sp->dx && sp->dy && sp->dz && sp->x_bounds && sp->y_bounds && sp->z_bounds;
    if (lines) {
        g.x_bounds.assign(sp->x_bounds, sp->x_bounds + nxg + 1);
        g.y_bounds.assign(sp->y_bounds, sp->y_bounds + nyg + 1);
        g.z_bounds.assign(sp->z_bounds, sp->z_bounds + nzg + 1);
    } else {
        g.x_bounds = GridSpacing::bounds(g.dx);
        g.y_bounds = GridSpacing::bounds(g.dy);
        g.z_bounds = GridSpacing::bounds(g.dz);
    }
    auto inv = [](const std::vector<double>& d) {
        std::vector<double> v(d.size());
        for (size_t i = 0; i < d.size(); ++i) v[i] = 1.0 / d[i];
        return v;
    };
    g.inv_dx = inv(g.dx); g.inv_dy = inv(g.dy); g.inv_dz = inv(g.dz);
    // dt rule, src/params.hpp:243-268
    const double dx_min = *std::min_element(g.dx.begin(), g.dx.end());
    const double dy_min = *std::min_element(g.dy.begin(), g.dy.end());
    const double dz_min = *std::min_element(g.dz.begin(), g.dz.end());
    const double sum_inv2 = (1.0 / (dx_min * dx_min)) + (1.0 / (dy_min * dy_min)) + (1.0 / (dz_min * dz_min));
    S->dt = sp->S / (C0 * std::sqrt(sum_inv2));

    S->bc_type = sp->boundary_type == FDTD_BC_PEC ? 0 : 1;
    if (S->bc_type == 0) {  // PECBoundary reports the core size and npml 0
        S->NxT = (size_t)sp->nx_core; S->NyT = (size_t)sp->ny_core; S->NzT = (size_t)sp->nz_core;
        S->npml = 0;
    } else {
        S->NxT = nxg; S->NyT = nyg; S->NzT = nzg;
        S->npml = np_spec;
    }
    g.npml = S->npml;
    S->N = S->NxT * S->NyT * S->NzT;
    S->kx.assign(S->NxT, 1.0); S->ky.assign(S->NyT, 1.0); S->kz.assign(S->NzT, 1.0);
    S->sx.assign(S->NxT, 0.0); S->sy.assign(S->NyT, 0.0); S->sz.assign(S->NzT, 0.0);
    S->ax.assign(S->NxT, 0.0); S->ay.assign(S->NyT, 0.0); S->az.assign(S->NzT, 0.0);
    S->bEx.assign(S->NxT, 1.0); S->cEx.assign(S->NxT, 0.0); S->bEy.assign(S->NyT, 1.0); S->cEy.assign(S->NyT, 0.0);
    S->bEz.assign(S->NzT, 1.0); S->cEz.assign(S->NzT, 0.0); S->bHx.assign(S->NxT, 1.0); S->cHx.assign(S->NxT, 0.0);
    S->bHy.assign(S->NyT, 1.0); S->cHy.assign(S->NyT, 0.0); S->bHz.assign(S->NzT, 1.0); S->cHz.assign(S->NzT, 0.0);
    if (S->bc_type == 1) {
        if (S->npml < 1 || 2 * S->npml > std::min({S->NxT, S->NyT, S->NzT})) { delete S; return FDTD_EINVAL; }
        build_cpml(S, sp);
    }

    // structure package, src/params.hpp:557-592 (make_nk: eps_r = n*n)
    S->bg_eps = sp->background_n * sp->background_n;
    const double pml_x = g.x_bounds[S->npml], pml_y = g.y_bounds[S->npml], pml_z = g.z_bounds[S->npml];
    for (int q = 0; q < sp->n_structures; ++q) {
        const fdtd_structure_def& d = sp->structures[q];
        Shape sh{};
        sh.type = d.type;
        sh.eps_r = d.n * d.n; sh.mu_r = 1.0; sh.sigma = 0.0; sh.sigma_m = 0.0;
        sh.tls = (sp->tls_enabled && d.tls_material >= 0 && d.tls_material < sp->n_tls_materials) ? d.tls_material : -1;
        if (d.type == FDTD_SHAPE_BOX) {
            sh.p[0] = pml_x + d.params[0]; sh.p[1] = pml_x + d.params[1];
            sh.p[2] = pml_y + d.params[2]; sh.p[3] = pml_y + d.params[3];
            sh.p[4] = pml_z + d.params[4]; sh.p[5] = pml_z + d.params[5];
        } else if (d.type == FDTD_SHAPE_SPHERE) {
            sh.p[0] = pml_x + d.params[0]; sh.p[1] = pml_y + d.params[1];
            sh.p[2] = pml_z + d.params[2]; sh.p[3] = d.params[3];
        } else {
            sh.p[0] = pml_x + d.params[0]; sh.p[1] = pml_y + d.params[1]; sh.p[2] = d.params[2];
            sh.p[3] = pml_z + d.params[3]; sh.p[4] = pml_z + d.params[4];
        }
        double b[6];
        sh.bounding_box(b);
        for (int a = 0; a < 3; ++a) {  // widen by a relative margin so culling stays conservative
            const double w = 1e-9 * (std::fabs(b[2 * a]) + std::fabs(b[2 * a + 1]) + 1e-12);
            sh.bb[2 * a] = b[2 * a] - w;
            sh.bb[2 * a + 1] = b[2 * a + 1] + w;
        }
        S->shapes.push_back(sh);
    }
    S->subpixel = sp->subpixel_samples > 0 ? sp->subpixel_samples : 8;
    S->device_bake = sp->device_bake != 0;
    if (!S->device_bake) bake(S, S->subpixel);

    // TLS, src/3D_FDTD_version008.cpp:100-172
    for (const Shape& sh : S->shapes)
        if (sh.tls >= 0) S->tls_on = true;
    S->tls_on = S->tls_on && sp->tls_enabled;
    if (S->tls_on) {
        const size_t N = S->N;
        S->Ndip.assign(N, 0.0); S->Ng0.assign(N, 0.0); S->Nu.assign(N, 0.0); S->Ng.assign(N, 0.0);
        S->Pz.assign(N, 0.0); S->Pz_prev.assign(N, 0.0);
        S->region_map.assign(N, -1);
        for (int q = 0; q < (int)S->shapes.size(); ++q) {
            const Shape& sh = S->shapes[q];
            if (sh.tls < 0) continue;
            const fdtd_tls_material& m = sp->tls_materials[sh.tls];
            double bb[6];
            sh.bounding_box(bb);
            Region R{};
            // TLSRegionManager::add_region_from_structure, src/two_level_system.hpp:495-521
            R.i0 = g.physical_to_index_x(bb[0] - pml_x);
            R.i1 = g.physical_to_index_x(bb[1] - pml_x) + 1;
            R.j0 = g.physical_to_index_y(bb[2] - pml_y);
            R.j1 = g.physical_to_index_y(bb[3] - pml_y) + 1;
            R.k0 = g.physical_to_index_z(bb[4] - pml_z);
            R.k1 = g.physical_to_index_z(bb[5] - pml_z) + 1;
            R.lambda0 = m.lambda0; R.gamma = m.gamma; R.tau = m.tau; R.N0 = m.N0;
            R.inversion_fraction = m.inversion_fraction;
            R.clamp = sp->tls_clamp;
            R.shape = q;
            S->regs.push_back(R);
        }
        for (size_t r = 0; r < S->regs.size(); ++r) {
            Region& R = S->regs[r];
            // TwoLevelParams::finalize, src/two_level_system.hpp:38-69
            R.hbar = HBAR;
            R.omega_a = 2.0 * PI * C0 / R.lambda0;
            const double numerator = 3.0 * PI * EPS0 * R.hbar * C0 * C0 * C0;
            const double denominator = R.omega_a * R.omega_a * R.omega_a * R.tau;
            R.mu_z = std::sqrt(numerator / denominator);
            // TLSRegion::initialize + cell map, :454-485, :533-546
            if (R.i0 >= R.i1 || R.j0 >= R.j1 || R.k0 >= R.k1) continue;
            const Shape& sh = S->shapes[(size_t)R.shape];
            for (size_t i = R.i0; i < R.i1; ++i)
                for (size_t j = R.j0; j < R.j1; ++j)
                    for (size_t k = R.k0; k < R.k1; ++k) {
                        const double x = 0.5 * (g.x_bounds[i] + g.x_bounds[i + 1]);
                        const double y = 0.5 * (g.y_bounds[j] + g.y_bounds[j + 1]);
                        const double z = 0.5 * (g.z_bounds[k] + g.z_bounds[k + 1]);
                        if (!sh.contains(x, y, z)) continue;
                        if (i >= S->NxT || j >= S->NyT || k >= S->NzT) continue;
                        const size_t id = (i * S->NyT + j) * S->NzT + k;
                        S->Ndip[id] = R.N0;
                        S->Nu[id] = R.inversion_fraction * R.N0;
                        S->Ng[id] = (1.0 - R.inversion_fraction) * R.N0;
                        S->Ng0[id] = R.N0;
                        S->region_map[id] = (int)r;
                    }
        }
        // TwoLevelParams::finalize_with_dt, :71-109
        for (Region& R : S->regs) {
            const double dt = S->dt;
            const double dt2 = dt * dt;
            const double omega_sq = R.omega_a * R.omega_a;
            const double beta = 2.0 + R.gamma * dt;
            const double two_omega_over_hbar = 2.0 * R.omega_a / R.hbar;
            const double mu_z_sq = R.mu_z * R.mu_z;
            R.kapa = (2.0 * dt2 / beta) * two_omega_over_hbar * mu_z_sq;
            R.pa2 = 2.0 * (2.0 - omega_sq * dt2) / beta;
            R.pa3 = (R.gamma * dt - 2.0) / beta;
            R.inv_dt = 1.0 / dt;
        }
    }

    // sources, src/params.hpp:620-674 (dipoles first, then plane waves)
    for (int d = 0; d < sp->n_dipoles; ++d) {
        const fdtd_dipole_def& D = sp->dipoles[d];
        Source s{};
        s.kind = 0;  // make_dipole_source, src/sources/dipole_source.hpp:90-153
        s.i = g.physical_to_index_x(D.x - sp->domain_x_min);
        s.j = g.physical_to_index_y(D.y - sp->domain_y_min);
        s.k = g.physical_to_index_z(D.z - sp->domain_z_min);
        s.f0 = src_freq(D.frequency, D.wavelength);
        s.tau = src_tau(D.tau, D.df_fwhm, S->dt);
        s.t0 = (D.t0_factor > 0.0) ? D.t0_factor * s.tau : 3.0 / s.f0;
        s.t_shift = s.t0;
        s.I0 = D.amplitude;
        s.waveform = D.waveform;
        s.polarization = D.polarization;
        switch (D.polarization) {
        case 0: case 3: s.cell_area = g.dy[s.j] * g.dz[s.k]; break;
        case 1: case 4: s.cell_area = g.dx[s.i] * g.dz[s.k]; break;
        default: s.cell_area = g.dx[s.i] * g.dy[s.j]; break;
        }
        if (s.i >= S->NxT || s.j >= S->NyT || s.k >= S->NzT) { delete S; return FDTD_EINVAL; }
        S->src.push_back(s);
    }
    for (int d = 0; d < sp->n_planes; ++d) {
        const fdtd_plane_wave_def& D = sp->planes[d];
        Source s{};
        s.kind = 1;  // make_plane_wave_source, src/sources/plane_wave_source.hpp:213-310
        s.direction = D.direction;
        s.polarization = D.polarization;
        s.waveform = D.waveform;
        s.f0 = src_freq(D.frequency, D.wavelength);
        s.tau = src_tau(D.tau, D.df_fwhm, S->dt);
        s.t0 = (D.t0_factor > 0.0) ? (D.t0_factor * s.tau) : (3.0 / s.f0);
        s.t_shift = s.t0;
        s.I0 = D.amplitude;
        const size_t np = S->npml;
        size_t pi, lim;
        if (D.direction >= 4) { pi = g.physical_to_index_z(D.injection_position - sp->domain_z_min); lim = S->NzT - np - 2; }
        else if (D.direction <= 1) { pi = g.physical_to_index_x(D.injection_position - sp->domain_x_min); lim = S->NxT - np - 2; }
        else { pi = g.physical_to_index_y(D.injection_position - sp->domain_y_min); lim = S->NyT - np - 2; }
        s.plane_index = std::max(np + 1, std::min(pi, lim));
        auto bounds = [&](double mn, double mx, size_t Nax, size_t& lo, size_t& hi, const std::vector<double>& b) {
            if (mn == 0.0 && mx == 0.0) { lo = np; hi = Nax - np - 1; return; }
            lo = GridSpacing::to_index(b, np, mn);
            hi = GridSpacing::to_index(b, np, mx);
            lo = std::max(np, lo);
            hi = std::min(Nax - np - 1, hi);
        };
        bounds(D.x_min - sp->domain_x_min, D.x_max - sp->domain_x_min, S->NxT, s.i_min, s.i_max, g.x_bounds);
        bounds(D.y_min - sp->domain_y_min, D.y_max - sp->domain_y_min, S->NyT, s.j_min, s.j_max, g.y_bounds);
        bounds(D.z_min - sp->domain_z_min, D.z_max - sp->domain_z_min, S->NzT, s.k_min, s.k_max, g.z_bounds);
        s.plane_d = D.direction >= 4 ? g.dz[s.plane_index] : (D.direction <= 1 ? g.dx[s.plane_index] : g.dy[s.plane_index]);
        S->src.push_back(s);
    }
    // probes, src/detectors/point_field_detector.hpp:57-70
    for (int p = 0; p < sp->n_probes; ++p) {
        const fdtd_probe_def& D = sp->probes[p];
        size_t i0 = g.physical_to_index_x(D.x - sp->domain_x_min);
        size_t j0 = g.physical_to_index_y(D.y - sp->domain_y_min);
        size_t k0 = g.physical_to_index_z(D.z - sp->domain_z_min);
        i0 = std::max(size_t(1), std::min(i0, S->NxT - 2));
        j0 = std::max(size_t(1), std::min(j0, S->NyT - 2));
        k0 = std::max(size_t(1), std::min(k0, S->NzT - 2));
        for (int c = 0; c < D.n_components; ++c) S->probes.push_back(ProbeChan{i0, j0, k0, D.components[c]});
    }
    S->save_every = sp->save_every > 0 ? sp->save_every : 1;
    for (int m = 0; m < sp->n_dfts; ++m) S->dfts.push_back(sp->dfts[m]);
    S->stab_interval = sp->stability_interval > 0 ? sp->stability_interval : 10;
    *out = S;
    return FDTD_OK;
}

void fdtd_setup_free(fdtd_setup* s) { delete s; }

int fdtd_setup_info(const fdtd_setup* s, long* info) {
    if (!s || !info) return FDTD_EINVAL;
    info[0] = (long)s->NxT; info[1] = (long)s->NyT; info[2] = (long)s->NzT; info[3] = (long)s->npml;
    info[4] = (long)s->src.size(); info[5] = (long)s->probes.size(); info[6] = (long)s->dfts.size();
    info[7] = s->tls_on; info[8] = (long)s->regs.size(); info[9] = s->bc_type;
    return FDTD_OK;
}

double fdtd_setup_dt(const fdtd_setup* s) { return s ? s->dt : 0.0; }

long fdtd_setup_array(const fdtd_setup* s, const char* name, double* out) {
    if (!s || !name) return -1;
    const std::string n(name);
    const std::vector<double>* v = nullptr;
    static const char* mn[12] = {"aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz", "bHz"};
    for (int q = 0; q < 12; ++q)
        if (n == mn[q]) {
            if (s->m[q].empty()) return -1;  // baked on the device
            v = &s->m[q];
        }
    const struct { const char* n; const std::vector<double>* v; } tab[] = {
        {"dx", &s->gs.dx}, {"dy", &s->gs.dy}, {"dz", &s->gs.dz}, {"inv_dx", &s->gs.inv_dx}, {"inv_dy", &s->gs.inv_dy},
        {"inv_dz", &s->gs.inv_dz}, {"x_bounds", &s->gs.x_bounds}, {"y_bounds", &s->gs.y_bounds},
        {"z_bounds", &s->gs.z_bounds}, {"kx", &s->kx}, {"ky", &s->ky}, {"kz", &s->kz},
        {"cpml_bEx", &s->bEx}, {"cpml_cEx", &s->cEx}, {"cpml_bEy", &s->bEy}, {"cpml_cEy", &s->cEy},
        {"cpml_bEz", &s->bEz}, {"cpml_cEz", &s->cEz}, {"cpml_bHx", &s->bHx}, {"cpml_cHx", &s->cHx},
        {"cpml_bHy", &s->bHy}, {"cpml_cHy", &s->cHy}, {"cpml_bHz", &s->bHz}, {"cpml_cHz", &s->cHz},
        {"Ndip", &s->Ndip}, {"Ng0", &s->Ng0}, {"Nu", &s->Nu}, {"Ng", &s->Ng}, {"Pz", &s->Pz}, {"Pz_prev", &s->Pz_prev}};
    for (const auto& t : tab)
        if (n == t.n) v = t.v;
    if (n == "n_grid") {
        // make_n_grid_from_scene (src/structure_material.hpp:291-310): the
        // index sqrt(eps_r mu_r) of the last structure containing each cell
        // centre, the background's elsewhere (what the mesh detector slices)
        const auto& xb = s->gs.x_bounds;
        const auto& yb = s->gs.y_bounds;
        const auto& zb = s->gs.z_bounds;
        if (out)
            for (size_t i = 0; i < s->NxT; ++i)
                for (size_t j = 0; j < s->NyT; ++j)
                    for (size_t k = 0; k < s->NzT; ++k) {
                        const double x = 0.5 * (xb[i] + xb[i + 1]), y = 0.5 * (yb[j] + yb[j + 1]),
                                     z = 0.5 * (zb[k] + zb[k + 1]);
                        double eps = s->bg_eps, mu = s->bg_mu;
                        for (const Shape& sh : s->shapes)
                            if (sh.contains(x, y, z)) {
                                eps = sh.eps_r;
                                mu = sh.mu_r;
                            }
                        out[(i * s->NyT + j) * s->NzT + k] = std::sqrt(eps * mu);
                    }
        return (long)s->N;
    }
    if (n == "region_map" && s->tls_on) {
        if (out)
            for (size_t i = 0; i < s->region_map.size(); ++i) out[i] = (double)s->region_map[i];
        return (long)s->region_map.size();
    }
    if (n == "tls_coeffs") {  // per region: kapa pa2 pa3 inv_dt omega_a mu_z N0 i0 i1 j0 j1 k0 k1
        if (out)
            for (size_t r = 0; r < s->regs.size(); ++r) {
                const Region& R = s->regs[r];
                const double v13[13] = {R.kapa, R.pa2, R.pa3, R.inv_dt, R.omega_a, R.mu_z, R.N0,
                                        (double)R.i0, (double)R.i1, (double)R.j0, (double)R.j1, (double)R.k0, (double)R.k1};
                std::memcpy(out + 13 * r, v13, sizeof(v13));
            }
        return (long)(13 * s->regs.size());
    }
    if (!v) return -1;
    if (out && !v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (long)v->size();
}

// J value of source q at step n: DipoleSource::inject_half_step
// (src/sources/dipole_source.hpp:60-85) and PlaneWaveSource::inject_*_plane
// (src/sources/plane_wave_source.hpp:102-179).
int fdtd_setup_source_values(const fdtd_setup* s, long n0, int nsteps, double* out) {
    if (!s || !out || nsteps < 0) return FDTD_EINVAL;
    const size_t ns = s->src.size();
    for (int t = 0; t < nsteps; ++t) {
        const double t_half = (static_cast<double>(n0 + t) + 0.5) * s->dt;
        for (size_t q = 0; q < ns; ++q) {
            const Source& src = s->src[q];
            const double w = compute_waveform(src.waveform, t_half, src.I0, src.f0, src.tau, src.t0, src.t_shift);
            double v;
            if (src.kind == 0) {
                v = w / src.cell_area;
            } else {
                const bool plus = src.direction == 0 || src.direction == 2 || src.direction == 4;
                const double sign = plus ? 1.0 : -1.0;
                v = 2.0 * w / (ETA0 * src.plane_d) * sign;
            }
            out[(size_t)t * ns + q] = v;
        }
    }
    return FDTD_OK;
}

int fdtd_setup_probe(const fdtd_setup* s, int c, long* v) {
    if (!s || c < 0 || c >= (int)s->probes.size() || !v) return FDTD_EINVAL;
    v[0] = (long)s->probes[(size_t)c].i; v[1] = (long)s->probes[(size_t)c].j; v[2] = (long)s->probes[(size_t)c].k;
    v[3] = s->probes[(size_t)c].comp;
    return FDTD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scene: one device simulation built from a setup.

struct fdtd_scene {
    const fdtd_setup* setup = nullptr;
    fdtd_handle h = nullptr;
    long n = 0;
    std::vector<double> vals;
    struct fdtd_scene_batch* batch = nullptr;  // the scene batch it belongs to
    long stop = -1;                            // batch steps never take n past this (-1: no limit)
};

extern "C" {

// Initial FP16 scales from the field the sources drive: a current sheet J
// radiates |E| ~ J eta0 d / 2 (a dipole: d the square root of its cell
// area); H ~ E / eta0; psi ~ field / d.  The device adapts them from there
// (never below 2^-6 of these).  Returns 1 with the 18 scales (powers of two)
// in out18, 0 when the scene has no source (the library's defaults stay).
int fdtd_setup_fp16_scales(const fdtd_setup* S, double* out18) {
    if (!S || !out18) return FDTD_EINVAL;
    if (S->src.empty()) return 0;
    double e_est = 0.0;
    const int nprobe = 4096;
    std::vector<double> v((size_t)nprobe * S->src.size());
    for (long n0 = 0; n0 < 40000; n0 += nprobe) {
        fdtd_setup_source_values(S, n0, nprobe, v.data());
        for (int t = 0; t < nprobe; ++t)
            for (size_t q = 0; q < S->src.size(); ++q) {
                const Source& src = S->src[q];
                const double d = src.kind == 1 ? src.plane_d : std::sqrt(src.cell_area);
                e_est = std::max(e_est, std::fabs(v[(size_t)t * S->src.size() + q]) * ETA0 * d / 2.0);
            }
    }
    if (!(e_est > 0)) return 0;
    double imax = 0;
    for (double x : S->gs.inv_dx) imax = std::max(imax, x);
    for (double x : S->gs.inv_dy) imax = std::max(imax, x);
    for (double x : S->gs.inv_dz) imax = std::max(imax, x);
    auto p2 = [](double x) { int e; std::frexp(x, &e); return std::ldexp(1.0, e - 8); };
    const double h_est = e_est / ETA0;
    for (int q = 0; q < 3; ++q) out18[q] = p2(e_est);
    for (int q = 3; q < 6; ++q) out18[q] = p2(h_est);
    for (int q = 6; q < 12; ++q) out18[q] = p2(h_est * imax);   // psi_E ~ dH/dx
    for (int q = 12; q < 18; ++q) out18[q] = p2(e_est * imax);  // psi_H ~ dE/dx
    return 1;
}

int fdtd_scene_create(const fdtd_setup* S, int precision, long probe_capacity, fdtd_scene** out) {
    if (!S || !out) return FDTD_EINVAL;
    *out = nullptr;
    fdtd_grid_desc g{};
    g.nx = (int)S->NxT; g.ny = (int)S->NyT; g.nz = (int)S->NzT;
    g.inv_dx = S->gs.inv_dx.data(); g.inv_dy = S->gs.inv_dy.data(); g.inv_dz = S->gs.inv_dz.data();
    g.dt = S->dt;
    fdtd_material_desc m{};
    fdtd_bake_desc bd{};
    std::vector<fdtd_bake_shape> bsh;
    if (S->device_bake) {
        bd.x_bounds = S->gs.x_bounds.data(); bd.y_bounds = S->gs.y_bounds.data(); bd.z_bounds = S->gs.z_bounds.data();
        bd.dx = S->gs.dx.data(); bd.dy = S->gs.dy.data(); bd.dz = S->gs.dz.data();
        bd.bg_eps_r = S->bg_eps; bd.bg_mu_r = S->bg_mu; bd.bg_sigma = S->bg_sig; bd.bg_sigma_m = S->bg_sigm;
        for (const Shape& sh : S->shapes) {
            fdtd_bake_shape b{};
            b.type = sh.type;
            for (int r = 0; r < 6; ++r) b.p[r] = sh.p[r];
            b.eps_r = sh.eps_r; b.mu_r = sh.mu_r; b.sigma = sh.sigma; b.sigma_m = sh.sigma_m;
            bsh.push_back(b);
        }
        bd.n_shapes = (int)bsh.size();
        bd.shapes = bsh.data();
        bd.subpixel_samples = S->subpixel;
        bd.dt = S->dt; bd.eps0 = EPS0; bd.mu0 = MU0;
        m.bake = &bd;
    } else {
        const double** mp[12] = {&m.aEx, &m.bEx, &m.aEy, &m.bEy, &m.aEz, &m.bEz, &m.aHx, &m.bHx, &m.aHy, &m.bHy, &m.aHz, &m.bHz};
        for (int q = 0; q < 12; ++q) *mp[q] = S->m[q].data();
    }
    fdtd_boundary_desc b{};
    b.type = S->bc_type;
    b.npml = (int)S->npml;
    b.kx = S->kx.data(); b.ky = S->ky.data(); b.kz = S->kz.data();
    b.bEx = S->bEx.data(); b.cEx = S->cEx.data(); b.bEy = S->bEy.data(); b.cEy = S->cEy.data();
    b.bEz = S->bEz.data(); b.cEz = S->cEz.data(); b.bHx = S->bHx.data(); b.cHx = S->cHx.data();
    b.bHy = S->bHy.data(); b.cHy = S->cHy.data(); b.bHz = S->bHz.data(); b.cHz = S->cHz.data();
    std::vector<fdtd_tls_region> regs;
    fdtd_tls_desc t{};
    const fdtd_tls_desc* tp = nullptr;
    if (S->tls_on) {
        for (const Region& R : S->regs) {
            fdtd_tls_region r{};
            r.kapa = R.kapa; r.pa2 = R.pa2; r.pa3 = R.pa3; r.inv_dt = R.inv_dt;
            r.tau = R.tau; r.hbar = R.hbar; r.omega_a = R.omega_a; r.N0 = R.N0; r.dt = S->dt; r.clamp = R.clamp;
            regs.push_back(r);
        }
        t.n_regions = (int)regs.size();
        t.regions = regs.data();
        t.cell_region_map = S->region_map.data();
        t.Ndip = S->Ndip.data(); t.Ng0 = S->Ng0.data(); t.Nu = S->Nu.data(); t.Ng = S->Ng.data();
        t.Pz = S->Pz.data(); t.Pz_prev = S->Pz_prev.data();
        tp = &t;
    }
    fdtd_handle h = nullptr;
    int rc = fdtd_create(&g, &m, &b, tp, precision, &h);
    if (rc) return rc;
    std::vector<fdtd_source_desc> sd;
    for (const Source& s : S->src) {
        fdtd_source_desc d{};
        if (s.kind == 0) {
            d.component = (s.polarization == 0 || s.polarization == 3) ? 0 : ((s.polarization == 1 || s.polarization == 4) ? 1 : 2);
            d.i0 = d.i1 = (int)s.i; d.j0 = d.j1 = (int)s.j; d.k0 = d.k1 = (int)s.k;
        } else {
            int comp = -1;
            const int p = s.polarization;
            if (s.direction >= 4) comp = (p == 0 || p == 3) ? 0 : ((p == 1 || p == 4) ? 1 : -1);
            else if (s.direction <= 1) comp = (p == 1 || p == 4) ? 1 : ((p == 2 || p == 5) ? 2 : -1);
            else comp = (p == 0 || p == 3) ? 0 : ((p == 2 || p == 5) ? 2 : -1);
            d.component = comp < 0 ? 0 : comp;
            if (s.direction >= 4) {
                d.i0 = (int)s.i_min; d.i1 = (int)s.i_max; d.j0 = (int)s.j_min; d.j1 = (int)s.j_max;
                d.k0 = d.k1 = (int)s.plane_index;
            } else if (s.direction <= 1) {
                d.i0 = d.i1 = (int)s.plane_index; d.j0 = (int)s.j_min; d.j1 = (int)s.j_max;
                d.k0 = (int)s.k_min; d.k1 = (int)s.k_max;
            } else {
                d.i0 = (int)s.i_min; d.i1 = (int)s.i_max; d.j0 = d.j1 = (int)s.plane_index;
                d.k0 = (int)s.k_min; d.k1 = (int)s.k_max;
            }
            if (comp < 0) d.i1 = d.i0 - 1;  // polarization along the propagation axis injects nothing
        }
        sd.push_back(d);
    }
    rc = fdtd_set_sources(h, (int)sd.size(), sd.data());
    for (const ProbeChan& p : S->probes) {
        if (rc < 0) break;
        int id = fdtd_add_probe(h, (int)p.i, (int)p.j, (int)p.k, p.comp, S->save_every, probe_capacity);
        if (id < 0) rc = id;
    }
    for (const fdtd_dft_def& d : S->dfts) {
        if (rc < 0) break;
        const int box[6] = {d.i0, d.i1, d.j0, d.j1, d.k0, d.k1};
        int id = fdtd_add_dft(h, d.component, box, d.n_freqs, d.freqs);
        if (id < 0) rc = id;
    }
    if (rc >= 0) rc = fdtd_set_monitor(h, S->stab_interval);
    if (rc >= 0 && precision == FDTD_FP16) {
        double sc[18];
        if (fdtd_setup_fp16_scales(S, sc) == 1) rc = fdtd_set_fp16_scales(h, sc);
    }
    if (rc < 0) {
        fdtd_destroy(h);
        return rc;
    }
    auto* sc = new fdtd_scene();
    sc->setup = S;
    sc->h = h;
    *out = sc;
    return FDTD_OK;
}

int fdtd_scene_destroy(fdtd_scene* sc) {
    if (!sc) return FDTD_OK;
    if (sc->batch) fdtd_scene_batch_remove(sc->batch, sc);
    fdtd_destroy(sc->h);
    delete sc;
    return FDTD_OK;
}

int fdtd_scene_step(fdtd_scene* sc, int nsteps) {
    if (!sc || nsteps < 0) return FDTD_EINVAL;
    const size_t ns = sc->setup->src.size();
    sc->vals.resize((size_t)nsteps * std::max<size_t>(ns, 1));
    if (ns) fdtd_setup_source_values(sc->setup, sc->n, nsteps, sc->vals.data());
    const int rc = fdtd_step(sc->h, sc->n, nsteps, ns ? sc->vals.data() : nullptr);
    if (rc == FDTD_OK || rc == FDTD_EUNSTABLE) sc->n += nsteps;
    return rc;
}

int fdtd_scene_step_timed(fdtd_scene* sc, int nsteps, double* ms) {
    if (!sc || nsteps < 0 || !ms) return FDTD_EINVAL;
    const size_t ns = sc->setup->src.size();
    sc->vals.resize((size_t)nsteps * std::max<size_t>(ns, 1));
    if (ns) fdtd_setup_source_values(sc->setup, sc->n, nsteps, sc->vals.data());
    const int rc = fdtd_step_timed(sc->h, sc->n, nsteps, ns ? sc->vals.data() : nullptr, ms);
    if (rc == FDTD_OK || rc == FDTD_EUNSTABLE) sc->n += nsteps;
    return rc;
}

int fdtd_scene_step_timed_every(fdtd_scene* sc, int nsteps, int every, double* ms) {
    if (!sc || nsteps < 0 || !ms || every < 1) return FDTD_EINVAL;
    const size_t ns = sc->setup->src.size();
    sc->vals.resize((size_t)nsteps * std::max<size_t>(ns, 1));
    if (ns) fdtd_setup_source_values(sc->setup, sc->n, nsteps, sc->vals.data());
    const int rc = fdtd_step_timed_every(sc->h, sc->n, nsteps, ns ? sc->vals.data() : nullptr, every, ms);
    if (rc == FDTD_OK || rc == FDTD_EUNSTABLE) sc->n += nsteps;
    return rc;
}

int fdtd_scene_step_async(fdtd_scene* sc, int nsteps) {
    if (!sc || nsteps < 0) return FDTD_EINVAL;
    const size_t ns = sc->setup->src.size();
    sc->vals.resize((size_t)nsteps * std::max<size_t>(ns, 1));
    if (ns) fdtd_setup_source_values(sc->setup, sc->n, nsteps, sc->vals.data());
    const int rc = fdtd_step_async(sc->h, sc->n, nsteps, ns ? sc->vals.data() : nullptr);
    if (rc == FDTD_OK) sc->n += nsteps;
    return rc;
}

int fdtd_scene_step_wait(fdtd_scene* sc) {
    if (!sc) return FDTD_EINVAL;
    return fdtd_step_wait(sc->h);
}

int fdtd_scene_step_query(fdtd_scene* sc) {
    if (!sc) return FDTD_EINVAL;
    return fdtd_step_query(sc->h);
}

long fdtd_scene_steps_done(const fdtd_scene* sc) { return sc ? sc->n : -1; }

int fdtd_scene_set_stop(fdtd_scene* sc, long n_stop) {
    if (!sc || n_stop < -1) return FDTD_EINVAL;
    sc->stop = n_stop;
    return FDTD_OK;
}

fdtd_sim* fdtd_scene_handle(fdtd_scene* sc) { return sc ? sc->h : nullptr; }

}  // extern "C"

struct fdtd_scene_batch {
    fdtd_batch_handle b = nullptr;
    int prec = 0;
    std::vector<fdtd_scene*> scenes;
    std::vector<std::vector<double>> vals;
};

namespace {
struct TuneKey {
    int prec, count, log2cells;
    bool operator<(const TuneKey& o) const {
        return std::tie(prec, count, log2cells) < std::tie(o.prec, o.count, o.log2cells);
    }
};
std::map<TuneKey, std::array<int, 4>>& tune_cache() {
    static std::map<TuneKey, std::array<int, 4>> m;
    return m;
}
std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}
}  // namespace

extern "C" {

int fdtd_scene_batch_create(int precision, fdtd_scene_batch** out) {
    if (!out) return FDTD_EINVAL;
    *out = nullptr;
    fdtd_batch_handle b = nullptr;
    if (int rc = fdtd_batch_create(precision, &b)) return rc;
    auto* sb = new fdtd_scene_batch();
    sb->b = b;
    sb->prec = precision;
    *out = sb;
    return FDTD_OK;
}

int fdtd_scene_batch_destroy(fdtd_scene_batch* sb) {
    if (!sb) return FDTD_OK;
    for (fdtd_scene* sc : sb->scenes) sc->batch = nullptr;
    fdtd_batch_destroy(sb->b);
    delete sb;
    return FDTD_OK;
}

int fdtd_scene_batch_add(fdtd_scene_batch* sb, fdtd_scene* sc) {
    if (!sb || !sc) return FDTD_EINVAL;
    const int rc = fdtd_batch_add(sb->b, sc->h);
    if (rc < 0) return rc;
    sb->scenes.push_back(sc);
    sc->batch = sb;
    return rc;
}

int fdtd_scene_batch_remove(fdtd_scene_batch* sb, fdtd_scene* sc) {
    if (!sb || !sc) return FDTD_EINVAL;
    auto it = std::find(sb->scenes.begin(), sb->scenes.end(), sc);
    if (it == sb->scenes.end()) return FDTD_EINVAL;
    const int rc = fdtd_batch_remove(sb->b, sc->h);
    if (rc) return rc;
    sb->scenes.erase(it);
    sc->batch = nullptr;
    return FDTD_OK;
}

int fdtd_scene_batch_count(const fdtd_scene_batch* sb) { return sb ? (int)sb->scenes.size() : FDTD_EINVAL; }

int fdtd_scene_batch_step(fdtd_scene_batch* sb, int nsteps, double* ms) {
    if (!sb || nsteps < 0) return FDTD_EINVAL;
    const size_t nd = sb->scenes.size();
    std::vector<long> n0(nd);
    std::vector<int> steps(nd);
    std::vector<const double*> sv(nd, nullptr);
    sb->vals.resize(nd);
    for (size_t d = 0; d < nd; ++d) {
        fdtd_scene* sc = sb->scenes[d];
        n0[d] = sc->n;
        steps[d] = sc->stop < 0 ? nsteps : (int)std::max<long>(0, std::min<long>(nsteps, sc->stop - sc->n));
        const size_t ns = sc->setup->src.size();
        if (ns) {
            sb->vals[d].resize((size_t)std::max(steps[d], 1) * ns);
            if (steps[d]) fdtd_setup_source_values(sc->setup, sc->n, steps[d], sb->vals[d].data());
            sv[d] = sb->vals[d].data();
        }
    }
    const int rc = fdtd_batch_step(sb->b, nsteps, n0.data(), steps.data(), sv.data(), ms);
    if (rc == FDTD_OK || rc == FDTD_EUNSTABLE)
        for (size_t d = 0; d < nd; ++d) sb->scenes[d]->n += steps[d];
    return rc;
}

int fdtd_scene_batch_autotune(fdtd_scene_batch* sb, int trial_steps, int* tuning) {
    if (!sb || trial_steps < 1 || !tuning) return FDTD_EINVAL;
    long cells = 0;
    for (fdtd_scene* sc : sb->scenes) cells += (long)sc->setup->NxT * sc->setup->NyT * sc->setup->NzT;
    int l2 = 0;
    while ((1L << l2) < cells) ++l2;
    const TuneKey key{sb->prec, (int)sb->scenes.size(), l2};
    {
        std::lock_guard<std::mutex> g(tune_mutex());
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
            for (int q = 0; q < 4; ++q) tuning[q] = it->second[q];
            return fdtd_batch_set_tuning(sb->b, tuning);
        }
    }
    int taken = 0;
    double ms = 0;
    // warm-up (table uploads, first-touch costs) under the current tiling
    int rc = fdtd_scene_batch_step(sb, trial_steps, &ms);
    if (rc) return rc;
    taken += trial_steps;
    double best = 1e300;
    std::array<int, 4> win{4, 4, 4, 1};
    for (int r : {1, 4})
        for (int w : {4, 8})
            for (int c : {4, 16})
                for (int p : {1, 4, 8}) {
                    const int cand[4] = {r, w, c, p};
                    if ((rc = fdtd_batch_set_tuning(sb->b, cand))) return rc;
                    if ((rc = fdtd_scene_batch_step(sb, trial_steps, &ms))) return rc;
                    taken += trial_steps;
                    if (ms < best) {
                        best = ms;
                        win = {r, w, c, p};
                    }
                }
    // one work-group per tile (chunk 0; work-groups per CU unused)
    for (int r : {1, 4})
        for (int p : {1, 4, 8}) {
            const int cand[4] = {r, 4, 0, p};
            if ((rc = fdtd_batch_set_tuning(sb->b, cand))) return rc;
            if ((rc = fdtd_scene_batch_step(sb, trial_steps, &ms))) return rc;
            taken += trial_steps;
            if (ms < best) {
                best = ms;
                win = {r, 4, 0, p};
            }
        }
    if ((rc = fdtd_batch_set_tuning(sb->b, win.data()))) return rc;
    {
        std::lock_guard<std::mutex> g(tune_mutex());
        tune_cache()[key] = win;
    }
    for (int q = 0; q < 4; ++q) tuning[q] = win[q];
    return taken;
}

fdtd_batch_handle fdtd_scene_batch_handle(fdtd_scene_batch* sb) { return sb ? sb->b : nullptr; }

}  // extern "C"

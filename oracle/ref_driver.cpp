// ref_driver.cpp - harness that runs the compiled reference (its own headers,
// read in place from the reference tree via -I) behind the same C ABI as the
// oracle.  TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into
// oracle/_ref/libref.so when the reference tree is present; used to write and
// check golden vectors and to cross-check the oracle bitwise.
//
// It replaces only the reference's compile-time configuration and main():
// the setup calls are the reference's (make_boundary, StructureScene::bake,
// make_dipole_source, make_plane_wave_source, TLSRegionManager, TwoLevelParams,
// StabilityMonitor) and the step order is the serial-mode body of
// src/3D_FDTD_version008.cpp:269-349.  make_context (src/params.hpp:180-296)
// is restated for the grid / dt part because it only supports the UNIFORM
// and auto-mesh modes and writes files.
#include <sstream>
#include <iostream>
#include <vector>
#include <memory>
#include <cstring>
#include <streambuf>
#include <chrono>
#include <omp.h>

#include "global_function.hpp"
#include "fdtd_stepper.hpp"
#include "boundary.hpp"
#include "structure_material.hpp"
#include "two_level_system.hpp"
#include "sources/dipole_source.hpp"
#include "sources/plane_wave_source.hpp"
#include "detectors/idetector.hpp"
#include "detectors/point_field_detector.hpp"
#include "detectors/field_movie_2d.hpp"
#include "detectors/mesh_detector.hpp"
#include "mesh_1d_profile_export.hpp"
#include "two_level_detectors.hpp"
#include "stability_monitor.hpp"

#include "../include/fdtd_scene.h"
#ifdef REF_DROPIN
#include <cmath>
#include "../include/fdtd_hip.h"
#endif

namespace {

struct NullBuf : std::streambuf {
    int overflow(int c) override { return c; }
};

struct Probe {
    size_t i0, j0, k0;
    std::vector<int> comps;
    std::vector<double> rec;
};

}  // namespace

struct ref_sim {
    size_t NxT{}, NyT{}, NzT{}, N{}, npml{};
    real dt{};
    GridSpacing gs;
    std::unique_ptr<IBoundary> bc;
    StructureScene scene;
    MaterialGrids mats;
    std::vector<std::unique_ptr<Sources::ISource>> sources;
    std::vector<real> Ex, Ey, Ez, Hx, Hy, Hz, Jx, Jy, Jz;
    bool tls_on{false};
    TwoLevelState tls_state;
    TLSRegionManager tls_manager;
    std::vector<Probe> probes;
    size_t save_every{1};
    size_t stab_interval{10};
    StabilityMonitor mon;
    long fail_step{-1};
    double last_max_E{0};
    size_t n{0};
    // the reference's own file-writing detectors, recorded after the E update
    std::vector<std::unique_ptr<Detectors::IDetector>> dets;
    // main's TLS detectors (src/3D_FDTD_version008.cpp:182-199), recorded
    // after the population update (:330-337)
    std::unique_ptr<PopulationTimeSeriesDetector> pop_det;
    std::unique_ptr<PolarizationDiagnostics> pol_det;
    std::unique_ptr<TLSGlobalHistory> tls_hist;
};

extern "C" {

static ref_sim* create_impl(const fdtd_scene_spec* sp, const double* const* mats12);

ref_sim* ref_create(const fdtd_scene_spec* sp) { return create_impl(sp, nullptr); }

// ref_create with the 12 MaterialGrids arrays given (idx3 order, NxT*NyT*NzT
// each: aEx bEx aEy bEy aEz bEz aHx bHx aHy bHy aHz bHz) in place of
// StructureScene::bake, whose serial subpixel loop over every structure takes
// hours on 1e8-cell grids with many structures.  The arrays a bake would
// produce are the input; everything else is the reference's own setup.  Used
// to time the reference's step on such grids (the CPU baseline).
ref_sim* ref_create_prebaked(const fdtd_scene_spec* sp, const double* const* mats12) {
    return mats12 ? create_impl(sp, mats12) : nullptr;
}

// OpenMP threads of the reference's parallel loops (omp_set_num_threads).
void ref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

static ref_sim* create_impl(const fdtd_scene_spec* sp, const double* const* mats12) {
    NullBuf nb;
    std::streambuf* old_out = std::cout.rdbuf(&nb);
    std::streambuf* old_err = std::cerr.rdbuf(&nb);
    auto* S = new ref_sim();
    const size_t npml_spec = (size_t)sp->npml;
    const size_t nxg = (size_t)sp->nx_core + 2 * npml_spec;
    const size_t nyg = (size_t)sp->ny_core + 2 * npml_spec;
    const size_t nzg = (size_t)sp->nz_core + 2 * npml_spec;
    GridSpacing g;
    if (sp->dx && sp->dy && sp->dz) {
        g.dx.assign(sp->dx, sp->dx + nxg);
        g.dy.assign(sp->dy, sp->dy + nyg);
        g.dz.assign(sp->dz, sp->dz + nzg);
    } else {
        g.dx.assign(nxg, sp->dx_uniform);
        g.dy.assign(nyg, sp->dy_uniform);
        g.dz.assign(nzg, sp->dz_uniform);
    }
    if (sp->dx && sp->dy && sp->dz && sp->x_bounds && sp->y_bounds && sp->z_bounds) {
        g.x_bounds.assign(sp->x_bounds, sp->x_bounds + nxg + 1);
        g.y_bounds.assign(sp->y_bounds, sp->y_bounds + nyg + 1);
        g.z_bounds.assign(sp->z_bounds, sp->z_bounds + nzg + 1);
    } else {
        build_cumulative_bounds(g.dx, g.x_bounds);
        build_cumulative_bounds(g.dy, g.y_bounds);
        build_cumulative_bounds(g.dz, g.z_bounds);
    }
    g.compute_inverses();
    real dx_min = *std::min_element(g.dx.begin(), g.dx.end());
    real dy_min = *std::min_element(g.dy.begin(), g.dy.end());
    real dz_min = *std::min_element(g.dz.begin(), g.dz.end());
    const real sum_inv2 = (1.0 / (dx_min * dx_min)) + (1.0 / (dy_min * dy_min)) + (1.0 / (dz_min * dz_min));
    S->dt = sp->S / (PhysConst::C0 * std::sqrt(sum_inv2));

    BoundaryParams bp;
    bp.type = sp->boundary_type == FDTD_BC_PEC ? BcType::PEC : BcType::CPML_RC;
    bp.cpml.npml = sp->npml;
    bp.cpml.m = sp->cpml_m;
    bp.cpml.Rerr = sp->cpml_rerr;
    bp.cpml.alpha0 = sp->cpml_alpha0;
    bp.cpml.kappa_max = sp->cpml_kappa_max;
    bp.cpml.alpha_linear = sp->cpml_alpha_linear != 0;
    S->bc = make_boundary(bp, sp->nx_core, sp->ny_core, sp->nz_core, g, S->dt, PhysConst::EPS0, PhysConst::MU0, PhysConst::C0);
    S->NxT = S->bc->NxT(); S->NyT = S->bc->NyT(); S->NzT = S->bc->NzT();
    S->npml = S->bc->npml();
    S->N = S->NxT * S->NyT * S->NzT;
    S->gs = g;
    S->gs.npml = S->npml;

    StructureScene& sc = S->scene;
    sc.NxT = S->NxT; sc.NyT = S->NyT; sc.NzT = S->NzT;
    sc.Nx_core = sp->nx_core; sc.Ny_core = sp->ny_core; sc.Nz_core = sp->nz_core;
    sc.npml_cells = S->npml;
    sc.grid_spacing = S->gs;
    sc.bg = make_nk(sp->background_n);
    const real pml_x = sc.grid_spacing.x_bounds[sc.npml_cells];
    const real pml_y = sc.grid_spacing.y_bounds[sc.npml_cells];
    const real pml_z = sc.grid_spacing.z_bounds[sc.npml_cells];
    for (int q = 0; q < sp->n_structures; ++q) {
        const fdtd_structure_def& s = sp->structures[q];
        Material mat = make_nk(s.n);
        StructureTLSConfig tls_cfg;
        if (sp->tls_enabled && s.tls_material >= 0 && s.tls_material < sp->n_tls_materials) {
            const fdtd_tls_material& m = sp->tls_materials[s.tls_material];
            tls_cfg.enabled = true;
            tls_cfg.lambda0 = m.lambda0; tls_cfg.gamma = m.gamma; tls_cfg.tau = m.tau;
            tls_cfg.N0 = m.N0; tls_cfg.inversion_fraction = m.inversion_fraction;
        }
        if (s.type == FDTD_SHAPE_BOX)
            sc.add_box(AABB{pml_x + s.params[0], pml_x + s.params[1], pml_y + s.params[2], pml_y + s.params[3],
                            pml_z + s.params[4], pml_z + s.params[5]}, mat, tls_cfg);
        else if (s.type == FDTD_SHAPE_SPHERE)
            sc.add_sphere(pml_x + s.params[0], pml_y + s.params[1], pml_z + s.params[2], s.params[3], mat, tls_cfg);
        else
            sc.add_cylinder_z(pml_x + s.params[0], pml_y + s.params[1], s.params[2], pml_z + s.params[3],
                              pml_z + s.params[4], mat, tls_cfg);
    }
    if (mats12) {
        S->mats.allocate(S->NxT, S->NyT, S->NzT);
        std::vector<real>* mv[12] = {&S->mats.aEx, &S->mats.bEx, &S->mats.aEy, &S->mats.bEy, &S->mats.aEz, &S->mats.bEz,
                                     &S->mats.aHx, &S->mats.bHx, &S->mats.aHy, &S->mats.bHy, &S->mats.aHz, &S->mats.bHz};
        for (int q = 0; q < 12; ++q) mv[q]->assign(mats12[q], mats12[q] + S->N);
    } else {
        sc.bake(S->mats, S->dt, PhysConst::EPS0, PhysConst::MU0, sp->subpixel_samples > 0 ? sp->subpixel_samples : 8);
    }
    S->bc->bind_material_coeffs(&S->mats.bEx, &S->mats.bEy, &S->mats.bEz, &S->mats.bHx, &S->mats.bHy, &S->mats.bHz);

    for (int d = 0; d < sp->n_dipoles; ++d) {
        const fdtd_dipole_def& D = sp->dipoles[d];
        Sources::SourceConfig cfg;
        cfg.x = D.x; cfg.y = D.y; cfg.z = D.z;
        cfg.amplitude = D.amplitude; cfg.frequency = D.frequency; cfg.wavelength = D.wavelength;
        cfg.tau = D.tau; cfg.df_fwhm = D.df_fwhm; cfg.t0_factor = D.t0_factor;
        cfg.waveform = static_cast<Sources::Waveform>(D.waveform);
        cfg.polarization = static_cast<Sources::Polarization>(D.polarization);
        auto src = Sources::make_dipole_source(cfg, S->gs, S->NyT, S->NzT, S->dt, sp->domain_x_min,
                                               sp->domain_y_min, sp->domain_z_min);
        if (src) S->sources.push_back(std::move(src));
    }
    for (int d = 0; d < sp->n_planes; ++d) {
        const fdtd_plane_wave_def& D = sp->planes[d];
        Sources::PlaneWaveConfig cfg;
        cfg.injection_position = D.injection_position;
        cfg.amplitude = D.amplitude; cfg.frequency = D.frequency; cfg.wavelength = D.wavelength;
        cfg.tau = D.tau; cfg.df_fwhm = D.df_fwhm; cfg.t0_factor = D.t0_factor;
        cfg.waveform = static_cast<Sources::Waveform>(D.waveform);
        cfg.direction = static_cast<Sources::PlaneWaveDirection>(D.direction);
        cfg.polarization = static_cast<Sources::Polarization>(D.polarization);
        cfg.x_min = D.x_min; cfg.x_max = D.x_max; cfg.y_min = D.y_min; cfg.y_max = D.y_max;
        cfg.z_min = D.z_min; cfg.z_max = D.z_max;
        auto src = Sources::make_plane_wave_source(cfg, S->gs, S->NxT, S->NyT, S->NzT, S->npml, S->dt,
                                                   sp->domain_x_min, sp->domain_y_min, sp->domain_z_min);
        if (src) S->sources.push_back(std::move(src));
    }

    S->Ex.assign(S->N, 0); S->Ey.assign(S->N, 0); S->Ez.assign(S->N, 0);
    S->Hx.assign(S->N, 0); S->Hy.assign(S->N, 0); S->Hz.assign(S->N, 0);
    S->Jx.assign(S->N, 0); S->Jy.assign(S->N, 0); S->Jz.assign(S->N, 0);

    // TLS block of main(), src/3D_FDTD_version008.cpp:100-172
    S->tls_on = sp->tls_enabled && sc.has_any_tls();
    if (S->tls_on) {
        S->tls_state.allocate(S->NxT, S->NyT, S->NzT);
        auto items = sc.get_tls_structures();
        for (size_t idx = 0; idx < items.size(); ++idx) {
            const auto* item = items[idx];
            const auto& tc = item->tls_config;
            AABB bb = item->shape->bounding_box();
            TwoLevelParams rp;
            rp.lambda0 = tc.lambda0; rp.gamma = tc.gamma; rp.tau = tc.tau; rp.N0_total = tc.N0;
            rp.enable_population_clamp = sp->tls_clamp != 0;
            S->tls_manager.add_region_from_structure(S->gs, bb.x0 - pml_x, bb.x1 - pml_x, bb.y0 - pml_y,
                                                     bb.y1 - pml_y, bb.z0 - pml_z, bb.z1 - pml_z, rp,
                                                     tc.inversion_fraction, item->shape.get(),
                                                     "structure_" + std::to_string(idx));
        }
        S->tls_manager.initialize(S->tls_state, S->gs);
        S->tls_manager.finalize_with_dt(S->dt);
        // the union box of the regions (src/3D_FDTD_version008.cpp:146-163),
        // which TLSGlobalHistory integrates over
        if (S->tls_manager.num_regions() > 0) {
            auto& st = S->tls_state;
            const auto& r0 = S->tls_manager.regions[0];
            st.gain_i0 = r0.i0; st.gain_i1 = r0.i1; st.gain_j0 = r0.j0;
            st.gain_j1 = r0.j1; st.gain_k0 = r0.k0; st.gain_k1 = r0.k1;
            for (size_t r = 1; r < S->tls_manager.num_regions(); ++r) {
                const auto& rr = S->tls_manager.regions[r];
                st.gain_i0 = std::min(st.gain_i0, rr.i0); st.gain_i1 = std::max(st.gain_i1, rr.i1);
                st.gain_j0 = std::min(st.gain_j0, rr.j0); st.gain_j1 = std::max(st.gain_j1, rr.j1);
                st.gain_k0 = std::min(st.gain_k0, rr.k0); st.gain_k1 = std::max(st.gain_k1, rr.k1);
            }
        }
    }

    // Probe placement of PointFieldDetector::initialize, point_field_detector.hpp:57-70
    for (int p = 0; p < sp->n_probes; ++p) {
        const fdtd_probe_def& D = sp->probes[p];
        Probe pr;
        pr.i0 = S->gs.physical_to_index_x(D.x - sp->domain_x_min);
        pr.j0 = S->gs.physical_to_index_y(D.y - sp->domain_y_min);
        pr.k0 = S->gs.physical_to_index_z(D.z - sp->domain_z_min);
        pr.i0 = std::max(std::size_t(1), std::min(pr.i0, S->NxT - 2));
        pr.j0 = std::max(std::size_t(1), std::min(pr.j0, S->NyT - 2));
        pr.k0 = std::max(std::size_t(1), std::min(pr.k0, S->NzT - 2));
        for (int c = 0; c < D.n_components; ++c) pr.comps.push_back(D.components[c]);
        S->probes.push_back(pr);
    }
    S->save_every = sp->save_every > 0 ? (size_t)sp->save_every : 1;
    S->stab_interval = sp->stability_interval > 0 ? (size_t)sp->stability_interval : 10;
    std::cout.rdbuf(old_out);
    std::cerr.rdbuf(old_err);
    return S;
}

int ref_step(ref_sim* S, int nsteps) {
    NullBuf nb;
    std::streambuf* old_err = std::cerr.rdbuf(&nb);
    const size_t NxT = S->NxT, NyT = S->NyT, NzT = S->NzT;
    auto& g = S->gs;
    auto& m = S->mats;
    int rc = 0;
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        const size_t n = S->n;
        std::fill(S->Jx.begin(), S->Jx.end(), 0.0);
        std::fill(S->Jy.begin(), S->Jy.end(), 0.0);
        std::fill(S->Jz.begin(), S->Jz.end(), 0.0);
        for (auto& src : S->sources) src->inject_half_step(n, S->dt, S->Jx, S->Jy, S->Jz);
        if (S->tls_on) update_polarization_multi_region(NxT, NyT, NzT, S->tls_manager, S->Ez, S->tls_state);
        fdtd_update_H(NxT, NyT, NzT, g.dx.data(), g.dy.data(), g.dz.data(), g.inv_dx.data(), g.inv_dy.data(),
                      g.inv_dz.data(), m.aHx.data(), m.bHx.data(), m.aHy.data(), m.bHy.data(), m.aHz.data(),
                      m.bHz.data(), S->Ex.data(), S->Ey.data(), S->Ez.data(), S->Hx.data(), S->Hy.data(), S->Hz.data());
        S->bc->apply_after_H(S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz);
        if (S->tls_on)
            fdtd_update_E_with_gain(NxT, NyT, NzT, g.dx.data(), g.dy.data(), g.dz.data(), g.inv_dx.data(),
                                    g.inv_dy.data(), g.inv_dz.data(), m.aEx.data(), m.bEx.data(), m.aEy.data(),
                                    m.bEy.data(), m.aEz.data(), m.bEz.data(), S->Ex.data(), S->Ey.data(), S->Ez.data(),
                                    S->Hx.data(), S->Hy.data(), S->Hz.data(), S->Jx.data(), S->Jy.data(), S->Jz.data(),
                                    S->tls_state);
        else
            fdtd_update_E(NxT, NyT, NzT, g.dx.data(), g.dy.data(), g.dz.data(), g.inv_dx.data(), g.inv_dy.data(),
                          g.inv_dz.data(), m.aEx.data(), m.bEx.data(), m.aEy.data(), m.bEy.data(), m.aEz.data(),
                          m.bEz.data(), S->Ex.data(), S->Ey.data(), S->Ez.data(), S->Hx.data(), S->Hy.data(),
                          S->Hz.data(), S->Jx.data(), S->Jy.data(), S->Jz.data());
        S->bc->apply_after_E(S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz);
        if (n % S->save_every == 0) {
            for (auto& pr : S->probes) {
                size_t id = idx3(pr.i0, pr.j0, pr.k0, NyT, NzT);
                for (int c : pr.comps)
                    pr.rec.push_back(Detectors::get_field_value(static_cast<Detectors::FieldComponent>(c), id, S->Ex,
                                                                S->Ey, S->Ez, S->Hx, S->Hy, S->Hz));
            }
        }
        for (auto& d : S->dets) d->record_after_E(n, S->dt, S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz);
        if (S->tls_on) update_populations_multi_region(NxT, NyT, NzT, S->dt, S->tls_manager, S->Ez, S->tls_state);
        if (S->pop_det) {
            NullBuf nb2;
            std::streambuf* o2 = std::cout.rdbuf(&nb2);
            S->pop_det->record(n, S->tls_state);
            S->pol_det->record(n, S->tls_state, S->Ez, S->gs);
            S->tls_hist->record(n, S->tls_state);
            std::cout.rdbuf(o2);
        }
        S->n++;
        if (n % S->stab_interval == 0) {
            real cons = S->mon.calculate_conservation_error(S->tls_state);
            bool ok = S->mon.check_stability(n, S->Ex, S->Ey, S->Ez, S->Hx, S->Hy, S->Hz, S->tls_state, cons);
            S->last_max_E = S->mon.get_max_E(S->Ex, S->Ey, S->Ez);
            if (!ok) { S->fail_step = (long)n; rc = 1; }
        }
    }
    std::cerr.rdbuf(old_err);
    return rc;
}

void ref_destroy(ref_sim* S) {
    for (auto& d : S->dets) d->finalize();
    if (S->pop_det) {  // main's end of run (src/3D_FDTD_version008.cpp:384-389)
        NullBuf nb;
        std::streambuf* o = std::cout.rdbuf(&nb);
        S->pop_det->finalize(S->tls_state);
        S->tls_hist->finalize();
        std::cout.rdbuf(o);
    }
    delete S;
}

// main's three TLS detectors (src/3D_FDTD_version008.cpp:182-199) under
// out_root, with the polarization monitor at cell (ic, jc, kc).
int ref_add_tls_detectors(ref_sim* S, const char* out_root, long ic, long jc, long kc, long save_every, long n_steps) {
    if (!S->tls_on) return -1;
    NullBuf nb;
    std::streambuf* o = std::cout.rdbuf(&nb);
    const fs::path root(out_root);
    S->pop_det = std::make_unique<PopulationTimeSeriesDetector>(root, "populations", S->NxT, S->NyT, S->NzT,
                                                                (size_t)save_every, (size_t)n_steps, S->dt, S->gs);
    S->pol_det = std::make_unique<PolarizationDiagnostics>(root, "polarization", S->NxT, S->NyT, S->NzT, (size_t)ic,
                                                           (size_t)jc, (size_t)kc, (size_t)save_every, S->dt);
    S->tls_hist = std::make_unique<TLSGlobalHistory>(root, S->tls_state, S->gs, (size_t)save_every, S->dt, false);
    std::cout.rdbuf(o);
    return 0;
}

// PointFieldDetector (src/detectors/point_field_detector.hpp:183-196)
// writing <out_root>/<name>/{<comp>_ts.bin, metadata.json}.
int ref_add_point_detector(ref_sim* S, const char* out_root, const char* name, double x, double y, double z,
                           int ncomp, const int* comps, int save_every, long n_steps, int float64, double dmx,
                           double dmy, double dmz) {
    NullBuf nb;
    std::streambuf* old_out = std::cout.rdbuf(&nb);
    Detectors::PointFieldDetectorConfig c;
    c.name = name;
    c.x = x; c.y = y; c.z = z;
    c.components.clear();
    for (int q = 0; q < ncomp; ++q) c.components.push_back(static_cast<Detectors::FieldComponent>(comps[q]));
    c.save_every = (std::size_t)save_every;
    c.n_steps = (std::size_t)n_steps;
    c.write_float64 = float64 != 0;
    S->dets.push_back(Detectors::make_point_field_detector(out_root, c, S->NxT, S->NyT, S->NzT, S->npml, S->gs, S->dt,
                                                           dmx, dmy, dmz));
    std::cout.rdbuf(old_out);
    return 0;
}

// FieldMovie2D (src/detectors/field_movie_2d.hpp:216-229) writing
// <out_root>/<name>/{frame_NNNN.raw, metadata.json}.
int ref_add_movie(ref_sim* S, const char* out_root, const char* name, int comp, int plane, double pos,
                  int save_every, long n_steps, int float64, long nx_phys, long ny_phys, long nz_phys, double Lx,
                  double Ly, double Lz, double dmx, double dmy, double dmz) {
    NullBuf nb;
    std::streambuf* old_out = std::cout.rdbuf(&nb);
    Detectors::FieldMovie2DConfig c;
    c.name = name;
    c.component = static_cast<Detectors::FieldComponent>(comp);
    c.slice_plane = plane;
    c.slice_position = pos;
    c.save_every = (std::size_t)save_every;
    c.n_steps = (std::size_t)n_steps;
    c.write_float64 = float64 != 0;
    S->dets.push_back(Detectors::make_field_movie_2d(out_root, c, S->NxT, S->NyT, S->NzT, (std::size_t)nx_phys,
                                                     (std::size_t)ny_phys, (std::size_t)nz_phys, S->npml, S->gs, Lx, Ly,
                                                     Lz, dmx, dmy, dmz));
    std::cout.rdbuf(old_out);
    return 0;
}

// MeshDetector (src/detectors/mesh_detector.hpp:265-280) as the detector
// package runs it (src/params.hpp:695-739): the refractive-index slice of
// make_n_grid_from_scene, then metadata, mesh lines and spacing arrays.
int ref_add_mesh_detector(ref_sim* S, const char* out_root, const char* name, int plane, double pos, int float64,
                          int lines, int spacing, long nx_phys, long ny_phys, long nz_phys, double Lx, double Ly,
                          double Lz, double dmx, double dmy, double dmz) {
    NullBuf nb;
    std::streambuf* old_out = std::cout.rdbuf(&nb);
    Detectors::MeshDetectorConfig c;
    c.name = name;
    c.slice_plane = plane;
    c.slice_position = pos;
    c.write_float64 = float64 != 0;
    c.export_mesh_lines = lines != 0;
    c.export_spacing_arrays = spacing != 0;
    std::vector<real> n_grid;
    make_n_grid_from_scene(S->scene, n_grid);
    auto det = Detectors::make_mesh_detector(out_root, c, S->NxT, S->NyT, S->NzT, (std::size_t)nx_phys,
                                             (std::size_t)ny_phys, (std::size_t)nz_phys, S->npml, S->gs, Lx, Ly, Lz,
                                             dmx, dmy, dmz);
    det->save_index_slice(n_grid);
    det->save_mesh_info();
    std::cout.rdbuf(old_out);
    return 0;
}

// mesh_export::dump_mesh_1d_profiles (src/mesh_1d_profile_export.hpp:14-66):
// <out_dir>/mesh_dx_profile.csv and mesh_dy_profile.csv.
int ref_dump_mesh_profiles(ref_sim* S, const char* out_dir) {
    NullBuf nb;
    std::streambuf* old_out = std::cout.rdbuf(&nb);
    const bool ok = mesh_export::dump_mesh_1d_profiles(S->gs, S->npml, out_dir);
    std::cout.rdbuf(old_out);
    return ok ? 0 : -1;
}

// make_n_grid_from_scene (src/structure_material.hpp:291-310): the
// refractive index at every cell centre (out: NxT*NyT*NzT values).
long ref_n_grid(ref_sim* S, double* out) {
    std::vector<real> n_grid;
    make_n_grid_from_scene(S->scene, n_grid);
    if (out) std::memcpy(out, n_grid.data(), n_grid.size() * sizeof(double));
    return (long)n_grid.size();
}

void ref_info(const ref_sim* S, long* out) {
    out[0] = (long)S->NxT; out[1] = (long)S->NyT; out[2] = (long)S->NzT; out[3] = (long)S->npml;
    out[4] = (long)S->n; out[5] = S->tls_on; out[6] = (long)S->tls_manager.num_regions(); out[7] = (long)S->sources.size();
    out[8] = S->mon.simulation_stable; out[9] = S->fail_step; out[10] = (long)S->mon.instability_count;
    out[11] = (long)S->gs.dx.size(); out[12] = (long)S->gs.dy.size(); out[13] = (long)S->gs.dz.size();
}
double ref_dt(const ref_sim* S) { return S->dt; }
double ref_last_max_E(const ref_sim* S) { return S->last_max_E; }

long ref_get_array(const ref_sim* S, const char* name, double* out) {
    const std::string nm(name);
    const std::vector<real>* v = nullptr;
    auto* cp = dynamic_cast<const CPMLBoundary*>(S->bc.get());
    const auto& m = S->mats;
    const auto& t = S->tls_state;
    static const char* fn[] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Jx", "Jy", "Jz"};
    const std::vector<real>* fv[] = {&S->Ex, &S->Ey, &S->Ez, &S->Hx, &S->Hy, &S->Hz, &S->Jx, &S->Jy, &S->Jz};
    for (int q = 0; q < 9; ++q) if (nm == fn[q]) v = fv[q];
    static const char* mn[] = {"aEx", "bEx", "aEy", "bEy", "aEz", "bEz", "aHx", "bHx", "aHy", "bHy", "aHz", "bHz"};
    const std::vector<real>* mv[] = {&m.aEx, &m.bEx, &m.aEy, &m.bEy, &m.aEz, &m.bEz, &m.aHx, &m.bHx, &m.aHy, &m.bHy, &m.aHz, &m.bHz};
    for (int q = 0; q < 12; ++q) if (nm == mn[q]) v = mv[q];
    if (S->tls_on) {
        static const char* tn[] = {"Ng", "Nu", "Ng0", "Pz", "Pz_prev", "dPz_dt", "Ez_old", "Ndip"};
        const std::vector<real>* tv[] = {&t.Ng, &t.Nu, &t.Ng0, &t.Pz, &t.Pz_prev, &t.dPz_dt, &t.Ez_old, &t.Ndip};
        for (int q = 0; q < 8; ++q) if (nm == tn[q]) v = tv[q];
    }
    if (cp) {
        static const char* pn[] = {"psi_Ex_y", "psi_Ex_z", "psi_Ey_z", "psi_Ey_x", "psi_Ez_x", "psi_Ez_y",
                                   "psi_Hx_y", "psi_Hx_z", "psi_Hy_z", "psi_Hy_x", "psi_Hz_x", "psi_Hz_y",
                                   "kx", "ky", "kz", "cpml_bEx", "cpml_cEx", "cpml_bEy", "cpml_cEy", "cpml_bEz",
                                   "cpml_cEz", "cpml_bHx", "cpml_cHx", "cpml_bHy", "cpml_cHy", "cpml_bHz", "cpml_cHz"};
        const std::vector<real>* pv[] = {&cp->psi_Ex_y_, &cp->psi_Ex_z_, &cp->psi_Ey_z_, &cp->psi_Ey_x_, &cp->psi_Ez_x_,
                                         &cp->psi_Ez_y_, &cp->psi_Hx_y_, &cp->psi_Hx_z_, &cp->psi_Hy_z_, &cp->psi_Hy_x_,
                                         &cp->psi_Hz_x_, &cp->psi_Hz_y_, &cp->kx_, &cp->ky_, &cp->kz_, &cp->bEx_,
                                         &cp->cEx_, &cp->bEy_, &cp->cEy_, &cp->bEz_, &cp->cEz_, &cp->bHx_, &cp->cHx_,
                                         &cp->bHy_, &cp->cHy_, &cp->bHz_, &cp->cHz_};
        for (int q = 0; q < 27; ++q) if (nm == pn[q]) v = pv[q];
    }
    const auto& g = S->gs;
    if (nm == "dx") v = &g.dx; if (nm == "dy") v = &g.dy; if (nm == "dz") v = &g.dz;
    if (nm == "inv_dx") v = &g.inv_dx; if (nm == "inv_dy") v = &g.inv_dy; if (nm == "inv_dz") v = &g.inv_dz;
    if (nm == "x_bounds") v = &g.x_bounds; if (nm == "y_bounds") v = &g.y_bounds; if (nm == "z_bounds") v = &g.z_bounds;
    if (!v) return -1;
    if (out) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (long)v->size();
}

long ref_get_region_map(const ref_sim* S, int* out) {
    if (!S->tls_on) return 0;
    const auto& m = S->tls_manager.cell_region_map;
    if (out) std::memcpy(out, m.data(), m.size() * sizeof(int));
    return (long)m.size();
}

long ref_probe_records(const ref_sim* S, int p) {
    if (p < 0 || p >= (int)S->probes.size()) return -1;
    return (long)(S->probes[p].rec.size() / std::max<size_t>(1, S->probes[p].comps.size()));
}
int ref_probe_index(const ref_sim* S, int p, long* ijk) {
    if (p < 0 || p >= (int)S->probes.size()) return -1;
    ijk[0] = (long)S->probes[p].i0; ijk[1] = (long)S->probes[p].j0; ijk[2] = (long)S->probes[p].k0;
    return (int)S->probes[p].comps.size();
}
long ref_get_probe(const ref_sim* S, int p, int c, double* out) {
    if (p < 0 || p >= (int)S->probes.size()) return -1;
    const auto& pr = S->probes[p];
    const size_t nc = pr.comps.size();
    if (c < 0 || (size_t)c >= nc) return -1;
    const size_t nr = pr.rec.size() / nc;
    for (size_t r = 0; r < nr; ++r) out[r] = pr.rec[r * nc + (size_t)c];
    return (long)nr;
}

int ref_set_array(ref_sim* S, const char* name, const double* in) {
    const std::string nm(name);
    static const char* fn[] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    std::vector<real>* fv[] = {&S->Ex, &S->Ey, &S->Ez, &S->Hx, &S->Hy, &S->Hz};
    for (int q = 0; q < 6; ++q)
        if (nm == fn[q]) { std::memcpy(fv[q]->data(), in, S->N * sizeof(double)); return 0; }
    return -1;
}

// Wall-clock seconds of nsteps steps of the reference's own step functions
// (their OpenMP loops over omp_get_max_threads() threads): the CPU baseline.
double ref_time_steps(ref_sim* S, int nsteps) {
    const auto t0 = std::chrono::steady_clock::now();
    ref_step(S, nsteps);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int ref_threads(void) { return omp_get_max_threads(); }

#ifdef REF_DROPIN
}  // extern "C"
#include "../include/fdtd_gpu_stepper.hpp"
extern "C" {
// The binding of the device library into the reference's main(), compiled:
// the reference's own state (its setup code above) goes to the adaptor of
// include/fdtd_gpu_stepper.hpp exactly as a maintainer's main() hands it
// over (GpuStepper from GridSpacing, MaterialGrids or the structure list,
// the boundary and the TLS objects; its sources; its probes), the device
// advances it in blocks of 10 steps (the stability monitor's cadence), and
// the reference advances the same state with its own step on the CPU.  rel
// receives the relative L2 difference of Ex Ey Ez Hx Hy Hz Nu Ng Pz Pz_prev
// (0..9), the largest one over the probe traces (10) and over the 12
// material arrays the handle holds against the reference's bake (11);
// returns the number of probe channels, or < 0 on error.
int ref_dropin_compare(const fdtd_scene_spec* sp, int nsteps, int precision, int device_bake, double* rel) {
    ref_sim* S = ref_create(sp);
    MaterialGrids& mg = S->mats;
    TwoLevelState& ts = S->tls_state;
    const TwoLevelState* tls = S->tls_on ? &ts : nullptr;
    const TLSRegionManager* regs = S->tls_on ? &S->tls_manager : nullptr;
    int rc = 0;
    try {
        std::unique_ptr<fdtd_gpu::GpuStepper> gpu;
        if (device_bake)
            gpu.reset(new fdtd_gpu::GpuStepper(S->gs, S->dt, S->scene, sp->subpixel_samples, PhysConst::EPS0,
                                               PhysConst::MU0, *S->bc, tls, regs, precision));
        else
            gpu.reset(new fdtd_gpu::GpuStepper(S->gs, S->dt, mg, *S->bc, tls, regs, precision));
        gpu->bind_sources(S->sources);
        // the harness's probes are PointFieldDetectors placed and clamped by
        // the reference's own code
        std::vector<std::vector<int>> chans;
        for (const Probe& pr : S->probes) {
            Detectors::PointFieldDetector d;
            d.i0 = pr.i0; d.j0 = pr.j0; d.k0 = pr.k0; d.save_every = S->save_every;
            for (int c : pr.comps) d.components.push_back(static_cast<Detectors::FieldComponent>(c));
            chans.push_back(gpu->bind_point_detector(d, (size_t)nsteps));
        }
        auto rl2 = [](const std::vector<double>& a, const std::vector<double>& r) {
            double num = 0, den = 0;
            for (size_t q = 0; q < a.size(); ++q) {
                num += (a[q] - r[q]) * (a[q] - r[q]);
                den += r[q] * r[q];
            }
            return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
        };
        std::vector<real> f[6];
        for (auto& v : f) v.resize(S->N);
        const std::vector<real>* ref[6] = {&S->Ex, &S->Ey, &S->Ez, &S->Hx, &S->Hy, &S->Hz};
        double snap[6] = {0, 0, 0, 0, 0, 0};
        // host detectors' steps (a movie every 7 steps, offset 3): the
        // reference's CPU step is brought to the same step and the whole
        // fields are compared there (an unstable verdict does not stop the
        // comparison)
        gpu->run(0, (size_t)nsteps, S->dt, [](size_t n) { return n % 7 == 3; }, [&](size_t n) {
            gpu->download_fields(f[0], f[1], f[2], f[3], f[4], f[5]);
            ref_step(S, (int)(n + 1 - S->n));
            for (int q = 0; q < 6; ++q) snap[q] = std::max(snap[q], rl2(f[q], *ref[q]));
        });
        if (S->n < (size_t)nsteps) ref_step(S, (int)((size_t)nsteps - S->n));
        gpu->download_fields(f[0], f[1], f[2], f[3], f[4], f[5]);
        for (int q = 0; q < 6; ++q) rel[q] = std::max(snap[q], rl2(f[q], *ref[q]));
        for (int q = 6; q < 10; ++q) rel[q] = 0.0;
        if (S->tls_on) {
            TwoLevelState dev = ts;
            gpu->download_tls(dev);
            rel[6] = rl2(dev.Nu, ts.Nu); rel[7] = rl2(dev.Ng, ts.Ng);
            rel[8] = rl2(dev.Pz, ts.Pz); rel[9] = rl2(dev.Pz_prev, ts.Pz_prev);
        }
        rel[11] = 0.0;
        std::vector<double> dev(S->N);
        const std::vector<real>* mref[12] = {&mg.aEx, &mg.bEx, &mg.aEy, &mg.bEy, &mg.aEz, &mg.bEz,
                                             &mg.aHx, &mg.bHx, &mg.aHy, &mg.bHy, &mg.aHz, &mg.bHz};
        for (int q = 0; q < 12 && rc == 0; ++q) {
            rc = fdtd_download_material(gpu->handle(), q, dev.data());
            rel[11] = std::max(rel[11], rl2(dev, *mref[q]));
        }
        rel[10] = 0.0;
        size_t nch = 0;
        for (size_t p = 0; p < S->probes.size(); ++p) {
            const Probe& pr = S->probes[p];
            const size_t nc = pr.comps.size();
            for (size_t c = 0; c < nc; ++c, ++nch) {
                std::vector<double> r;
                for (size_t t = c; t < pr.rec.size(); t += nc) r.push_back(pr.rec[t]);
                const std::vector<double> d = gpu->probe_trace(chans[p][c]);
                if (d.size() != r.size()) rc = -2;
                else rel[10] = std::max(rel[10], rl2(d, r));
            }
        }
        if (rc == 0) rc = (int)nch;
    } catch (const std::exception&) {
        rc = -1;
    }
    ref_destroy(S);
    return rc;
}
#endif

}  // extern "C"

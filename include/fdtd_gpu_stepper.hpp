/* fdtd_gpu_stepper.hpp - C++ adaptor that puts the MI355X step (fdtd_hip.h)
 * behind the reference's own simulation objects.
 *
 * Header-only, compiled inside the reference's build: include it after the
 * reference's headers (global_function.hpp, boundary.hpp,
 * two_level_system.hpp, structure_material.hpp, sources/*.hpp,
 * detectors/point_field_detector.hpp) and link lib3dfdtd_hip.so.  It replaces
 * the body of the serial step loop (src/3D_FDTD_version008.cpp:269-349):
 *
 *   reference (per step)                                   adaptor
 *   ------------------------------------------------------ ---------------------------
 *   std::fill(J..); src->inject_half_step(n, dt, J..)       GpuStepper::step: the value
 *     src/3D_FDTD_version008.cpp:270-275                      each source's waveform gives
 *   update_polarization_multi_region, fdtd_update_H,        (one fdtd_step call per block
 *   bc->apply_after_H, fdtd_update_E(_with_gain),             of steps; no N-sized J array)
 *   bc->apply_after_E, update_populations_multi_region
 *   det->record_after_E (PointFieldDetector)                bind_point_detector +
 *                                                             probe_trace
 *   monitor.check_stability every 10 steps                  step() returns false where the
 *     src/3D_FDTD_version008.cpp:339-349                      reference's policy fails
 *
 * The reference's setup (make_context, bake_and_bind, the TLS region and
 * state set-up, make_dipole_source / make_plane_wave_source, the detectors'
 * initialize) is unchanged and runs first; the adaptor reads the objects it
 * built.  Host-side source values keep the reference's libm and operation
 * order, so the FP64 device path is bitwise the reference's step.
 *
 * Errors: the constructor throws std::runtime_error with the library's
 * message; step() returns false on an unstable run (check last_error() for a
 * device error), as main returns 1.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fdtd_hip.h"

namespace fdtd_gpu {

// The J component a source polarization writes (Dipole/PlaneWaveSource:
// the "H" polarizations inject into J too, src/sources/dipole_source.hpp:71-84).
inline int j_component(Sources::Polarization p) {
    using P = Sources::Polarization;
    return (p == P::Ex || p == P::Hx) ? 0 : (p == P::Ey || p == P::Hy) ? 1 : 2;
}

class GpuStepper {
public:
    // From the reference's objects after bake_and_bind: GridSpacing, dt,
    // MaterialGrids, the boundary (PEC or CPML), and for gain the
    // TwoLevelState with its TLSRegionManager (nullptrs without gain).
    // precision: FDTD_FP64 (parity), FDTD_FP32 or FDTD_FP16.
    GpuStepper(const GridSpacing& gs, real dt, const MaterialGrids& mg, const IBoundary& bc,
               const TwoLevelState* tls, const TLSRegionManager* regions, int precision = FDTD_FP64) {
        fdtd_material_desc m{mg.aEx.data(), mg.bEx.data(), mg.aEy.data(), mg.bEy.data(), mg.aEz.data(), mg.bEz.data(),
                             mg.aHx.data(), mg.bHx.data(), mg.aHy.data(), mg.bHy.data(), mg.aHz.data(),
                             mg.bHz.data(), nullptr};
        create(gs, dt, m, bc, tls, regions, precision);
    }
    // The same with the reference's structure list baked on the device
    // (StructureScene::bake, src/structure_material.hpp:211-288, every
    // coefficient bitwise the host bake's) instead of host MaterialGrids.
    GpuStepper(const GridSpacing& gs, real dt, const StructureScene& scene, int subpixel_samples, real eps0,
               real mu0, const IBoundary& bc, const TwoLevelState* tls, const TLSRegionManager* regions,
               int precision = FDTD_FP64) {
        for (const auto& it : scene.items) {
            fdtd_bake_shape q{};
            if (auto* bx = dynamic_cast<const Box*>(it.shape.get())) {
                q.type = 0;
                q.p[0] = bx->bb.x0; q.p[1] = bx->bb.x1; q.p[2] = bx->bb.y0;
                q.p[3] = bx->bb.y1; q.p[4] = bx->bb.z0; q.p[5] = bx->bb.z1;
            } else if (auto* sh = dynamic_cast<const Sphere*>(it.shape.get())) {
                q.type = 1;
                q.p[0] = sh->cx; q.p[1] = sh->cy; q.p[2] = sh->cz; q.p[3] = sh->r;
            } else if (auto* cy = dynamic_cast<const CylinderZ*>(it.shape.get())) {
                q.type = 2;
                q.p[0] = cy->cx; q.p[1] = cy->cy; q.p[2] = cy->r; q.p[3] = cy->z0; q.p[4] = cy->z1;
            } else {
                throw std::runtime_error("fdtd_gpu: structure shape without a device bake");
            }
            q.eps_r = it.mat.eps_r; q.mu_r = it.mat.mu_r; q.sigma = it.mat.sigma; q.sigma_m = it.mat.sigma_m;
            shapes_.push_back(q);
        }
        const Material& bg = scene.bg;
        fdtd_bake_desc bd{gs.x_bounds.data(), gs.y_bounds.data(), gs.z_bounds.data(), gs.dx.data(), gs.dy.data(),
                          gs.dz.data(), bg.eps_r, bg.mu_r, bg.sigma, bg.sigma_m, (int)shapes_.size(), shapes_.data(),
                          subpixel_samples > 0 ? subpixel_samples : 8, dt, eps0, mu0};
        fdtd_material_desc m{};
        m.bake = &bd;
        create(gs, dt, m, bc, tls, regions, precision);
    }
    ~GpuStepper() {
        if (h_) fdtd_destroy(h_);
    }
    GpuStepper(const GpuStepper&) = delete;
    GpuStepper& operator=(const GpuStepper&) = delete;

    // Source footprints of the reference's sources in their order: a
    // DipoleSource is its cell, a PlaneWaveSource the rectangle of its plane
    // (src/sources/plane_wave_source.hpp:102-179).  Per-step values are
    // computed in step() from the sources' own waveforms.
    void bind_sources(const std::vector<std::unique_ptr<Sources::ISource>>& sources) {
        srcs_.clear();
        std::vector<fdtd_source_desc> fp;
        for (const auto& s : sources) {
            Src q{};
            if (auto* d = dynamic_cast<const Sources::DipoleSource*>(s.get())) {
                q.dipole = d;
                fp.push_back({j_component(d->polarization), (int)d->i, (int)d->i, (int)d->j, (int)d->j, (int)d->k,
                              (int)d->k});
            } else if (auto* pw = dynamic_cast<const Sources::PlaneWaveSource*>(s.get())) {
                using D = Sources::PlaneWaveDirection;
                using P = Sources::Polarization;
                q.plane = pw;
                fdtd_source_desc f{j_component(pw->polarization), (int)pw->i_min, (int)pw->i_max, (int)pw->j_min,
                                   (int)pw->j_max, (int)pw->k_min, (int)pw->k_max};
                const P p = pw->polarization;
                if (pw->direction == D::PlusZ || pw->direction == D::MinusZ) {
                    f.k0 = f.k1 = (int)pw->plane_index;
                    q.d = pw->dz_array[pw->plane_index];
                    q.active = p == P::Ex || p == P::Hx || p == P::Ey || p == P::Hy;
                    q.sign = pw->direction == D::PlusZ ? 1.0 : -1.0;
                } else if (pw->direction == D::PlusX || pw->direction == D::MinusX) {
                    f.i0 = f.i1 = (int)pw->plane_index;
                    q.d = pw->dx_array[pw->plane_index];
                    q.active = p == P::Ey || p == P::Hy || p == P::Ez || p == P::Hz;
                    q.sign = pw->direction == D::PlusX ? 1.0 : -1.0;
                } else {
                    f.j0 = f.j1 = (int)pw->plane_index;
                    q.d = pw->dy_array[pw->plane_index];
                    q.active = p == P::Ex || p == P::Hx || p == P::Ez || p == P::Hz;
                    q.sign = pw->direction == D::PlusY ? 1.0 : -1.0;
                }
                fp.push_back(f);
            } else {
                throw std::runtime_error("fdtd_gpu: source type without a device footprint: " + s->name());
            }
            srcs_.push_back(q);
        }
        check(fdtd_set_sources(h_, (int)fp.size(), fp.data()));
    }

    // The J value source q adds at global step n (what its inject_half_step
    // adds to each footprint cell, with the reference's operations).
    double source_value(size_t q, size_t n, real dt) const {
        const real t_half = (static_cast<real>(n) + real(0.5)) * dt;
        const Src& s = srcs_[q];
        if (s.dipole) return s.dipole->waveform_value(t_half) / s.dipole->cell_area;
        if (!s.active) return 0.0;
        const real Et = s.plane->waveform_value(t_half);
        return 2.0 * Et / (Sources::PlaneWaveSource::eta0 * s.d) * s.sign;
    }

    // The probe channels of a PointFieldDetector (its clamped cell, its
    // components, its save_every): channel ids in component order.
    std::vector<int> bind_point_detector(const Detectors::PointFieldDetector& d, size_t n_steps) {
        std::vector<int> ids;
        for (auto c : d.components) {
            const int id = fdtd_add_probe(h_, (int)d.i0, (int)d.j0, (int)d.k0, (int)c, (int)d.save_every,
                                          (long)(n_steps / d.save_every + 1));
            check(id);
            ids.push_back(id);
        }
        return ids;
    }

    // Advances nsteps steps from global step n0 in the reference's order,
    // with the reference's stability policy at every 10th step; one
    // fdtd_step call for the block.  false where the reference's
    // check_stability fails (main returns 1).
    bool step(size_t n0, int nsteps, real dt) {
        vals_.resize((size_t)nsteps * srcs_.size());
        for (int s = 0; s < nsteps; ++s)
            for (size_t q = 0; q < srcs_.size(); ++q) vals_[(size_t)s * srcs_.size() + q] = source_value(q, n0 + s, dt);
        const int rc = fdtd_step(h_, (long)n0, nsteps, srcs_.empty() ? nullptr : vals_.data());
        if (rc == FDTD_EUNSTABLE) return false;
        check(rc);
        return true;
    }

    // Steps [n0, n0 + nsteps) for a main() that also runs host detectors
    // (FieldMovie2D, MeshDetector: they record from whole fields, as
    // rt.after_E hands them, src/params.hpp:793-810).  Blocks end at every
    // 10th step (the stability cadence) and after each step n with
    // wants(n); after such a step, after(n) runs with the device state at
    // the end of step n (download_fields gives the fields the reference's
    // record_after_E would see).  false where the reference's policy stops.
    template <typename Wants, typename After>
    bool run(size_t n0, size_t nsteps, real dt, Wants wants, After after) {
        const size_t end = n0 + nsteps;
        for (size_t n = n0; n < end;) {
            size_t stop = std::min(end, (n / 10 + 1) * 10);
            for (size_t s = n; s < stop; ++s)
                if (wants(s)) { stop = s + 1; break; }
            if (!step(n, (int)(stop - n), dt)) return false;
            if (wants(stop - 1)) after(stop - 1);
            n = stop;
        }
        return true;
    }

    // Probe channel trace (records so far).
    std::vector<double> probe_trace(int channel) const {
        long cnt = 0;
        check(fdtd_read_probe(h_, channel, nullptr, 0, &cnt));
        std::vector<double> out((size_t)cnt);
        if (cnt) check(fdtd_read_probe(h_, channel, out.data(), cnt, &cnt));
        return out;
    }

    // Fields and TLS state back into the reference's vectors (idx3 order),
    // for its detectors, snapshots or a later CPU step.
    void download_fields(std::vector<real>& Ex, std::vector<real>& Ey, std::vector<real>& Ez, std::vector<real>& Hx,
                         std::vector<real>& Hy, std::vector<real>& Hz) const {
        std::vector<real>* f[6] = {&Ex, &Ey, &Ez, &Hx, &Hy, &Hz};
        for (int q = 0; q < 6; ++q) check(fdtd_download(h_, q, f[q]->data()));
    }
    void download_tls(TwoLevelState& ts) const {
        std::vector<real>* f[4] = {&ts.Nu, &ts.Ng, &ts.Pz, &ts.Pz_prev};
        for (int q = 0; q < 4; ++q) check(fdtd_download(h_, 6 + q, f[q]->data()));
    }

    fdtd_handle handle() const { return h_; }

private:
    struct Src {
        const Sources::DipoleSource* dipole;
        const Sources::PlaneWaveSource* plane;
        real d;
        real sign;
        bool active;
    };

    static void check(int rc) {
        if (rc < 0) throw std::runtime_error(std::string("fdtd_gpu: ") + fdtd_last_error());
    }

    void create(const GridSpacing& gs, real dt, const fdtd_material_desc& m, const IBoundary& bc,
                const TwoLevelState* tls, const TLSRegionManager* regions, int precision) {
        fdtd_grid_desc g{(int)bc.NxT(), (int)bc.NyT(), (int)bc.NzT(), gs.inv_dx.data(), gs.inv_dy.data(),
                         gs.inv_dz.data(), dt};
        fdtd_boundary_desc b{0, (int)bc.npml()};
        if (auto* cp = dynamic_cast<const CPMLBoundary*>(&bc)) {
            b.type = 1;
            b.kx = cp->kx_.data(); b.ky = cp->ky_.data(); b.kz = cp->kz_.data();
            b.bEx = cp->bEx_.data(); b.cEx = cp->cEx_.data(); b.bEy = cp->bEy_.data(); b.cEy = cp->cEy_.data();
            b.bEz = cp->bEz_.data(); b.cEz = cp->cEz_.data(); b.bHx = cp->bHx_.data(); b.cHx = cp->cHx_.data();
            b.bHy = cp->bHy_.data(); b.cHy = cp->cHy_.data(); b.bHz = cp->bHz_.data(); b.cHz = cp->cHz_.data();
        }
        fdtd_tls_desc t{};
        std::vector<fdtd_tls_region> regs;
        if (tls && regions) {
            for (const auto& r : regions->regions)
                regs.push_back({r.params.kapa_coeff, r.params.pa2, r.params.pa3, r.params.inv_dt, r.params.tau,
                                r.params.hbar, r.params.omega_a, r.params.N0_total, dt,
                                r.params.enable_population_clamp ? 1 : 0});
            t = fdtd_tls_desc{(int)regs.size(), regs.data(), regions->cell_region_map.data(), tls->Ndip.data(),
                              tls->Ng0.data(), tls->Nu.data(), tls->Ng.data(), tls->Pz.data(), tls->Pz_prev.data()};
        }
        check(fdtd_create(&g, &m, &b, (tls && regions) ? &t : nullptr, precision, &h_));
    }

    fdtd_handle h_ = nullptr;
    std::vector<Src> srcs_;
    std::vector<double> vals_;
    std::vector<fdtd_bake_shape> shapes_;
};

}  // namespace fdtd_gpu

// fdtd_automesh.cpp - automatic nonuniform mesh (include/fdtd_mesh.h).
//
// Restates AutoMeshGenerator (src/auto_mesh_generator.hpp) for one axis at a
// time.  The grid lines are fixed by four steps, kept here with the
// reference's floating-point operations in the reference's order (compiled
// without contraction), so the lines are bitwise the reference's:
//   1. split [0, domain] at override and structure-interface positions and
//      give each piece a target cell size (:389-453);
//   2. fill each piece: uniform inside overrides, graded between two
//      overrides, uniform inside structures, entry/stable/exit sequences
//      elsewhere (:455-840);
//   3. halve the larger cell of the worst grading violation, up to 300
//      times (:859-897), then drop near-duplicate lines (:846-854);
//   4. add npml PML cells on each side and shift to start at 0 (:279-375).
// What the reference computes only to print (constraint lists, validation
// and statistics, :256-276, :900-1263) has no effect on the lines and is
// not restated.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <set>
#include <vector>

#include "../../include/fdtd_hip.h"
#include "../../include/fdtd_mesh.h"
#include "../../include/fdtd_scene.h"

namespace {

constexpr double kSpeedOfLight = 299792458.0;  // PhysConst::C0, src/global_function.hpp:14

struct Piece {         // one region of step 1
    double lo, hi, target;
    bool fixed;        // inside an override region
};

struct Override1D {    // an override region seen along one axis
    double lo, hi, d;
};

class AxisMesher {
public:
    AxisMesher(const fdtd_auto_mesh_config& c, int axis, double domain, const fdtd_mesh_bounds* s, int ns,
               double background)
        : axis_(axis), L_(domain), st_(s), ns_(ns), bg_(background), dmin_(c.dx_min), dmax_(c.dx_max),
          g_(c.max_grading_ratio), graded_pml_(c.pml_graded != 0) {
        // AutoMeshConfig::get_lambda_min / get_ppw_target (:104-116)
        lambda_ = c.f_max > 0 ? kSpeedOfLight / c.f_max : c.lambda_min;
        if (c.ppw_override > 0) ppw_ = c.ppw_override;
        else if (c.mesh_accuracy <= 1) ppw_ = 6.0;
        else if (c.mesh_accuracy == 2) ppw_ = 10.0;
        else ppw_ = 14.0 + (c.mesh_accuracy - 3) * 4.0;
    }

    void add_override(const fdtd_mesh_override& r) {
        // collect_override_segments_1d (:969-996)
        double lo = axis_ == 0 ? r.x0 : axis_ == 1 ? r.y0 : r.z0;
        double hi = axis_ == 0 ? r.x1 : axis_ == 1 ? r.y1 : r.z1;
        double d = axis_ == 0 ? r.dx : axis_ == 1 ? r.dy : r.dz;
        lo = std::max(0.0, std::min(lo, L_));
        hi = std::max(0.0, std::min(hi, L_));
        if (lo < hi) ov_.push_back({lo, hi, std::clamp(d, dmin_, dmax_)});
    }

    std::vector<double> physical_lines() {
        build_pieces();
        lines_.assign(1, 0.0);
        prev_ = pieces_[0].target;
        for (size_t p = 0; p < pieces_.size(); ++p) fill_piece(p);
        lines_.back() = L_;
        fix_grading();
        std::vector<double> out{lines_[0]};
        for (size_t i = 1; i < lines_.size(); ++i)
            if (lines_[i] > out.back() + dmin_ * 0.001) out.push_back(lines_[i]);
        return out;
    }

    std::vector<double> with_pml(const std::vector<double>& phys, size_t npml) const {
        if (npml == 0 || phys.size() < 2) return phys;
        const double dl = phys[1] - phys[0];
        const double dr = phys.back() - phys[phys.size() - 2];
        std::vector<double> left, right;
        if (!graded_pml_) {
            const double thick = npml * dl;
            for (size_t i = 0; i < npml; ++i) left.push_back(-thick + i * dl);
            const double end = phys.back();
            for (size_t i = 1; i <= npml; ++i) right.push_back(end + i * dr);
        } else {
            double d = dl, x = 0;
            for (size_t i = 0; i < npml; ++i) {
                x -= d;
                left.push_back(x);
                d = std::min(d * g_, dmax_);
            }
            std::reverse(left.begin(), left.end());
            d = dr;
            x = phys.back();
            for (size_t i = 0; i < npml; ++i) {
                x += d;
                right.push_back(x);
                d = std::min(d * g_, dmax_);
            }
        }
        std::vector<double> all;
        all.reserve(left.size() + phys.size() + right.size());
        all.insert(all.end(), left.begin(), left.end());
        all.insert(all.end(), phys.begin(), phys.end());
        all.insert(all.end(), right.begin(), right.end());
        const double shift = -all[0];
        for (double& x : all) x += shift;
        return all;
    }

private:
    void lim(const fdtd_mesh_bounds& s, double& lo, double& hi) const {
        lo = axis_ == 0 ? s.x_min : axis_ == 1 ? s.y_min : s.z_min;
        hi = axis_ == 0 ? s.x_max : axis_ == 1 ? s.y_max : s.z_max;
    }

    // sample_n_at_position (:1078-1098): the first structure containing pos
    double index_at(double pos) const {
        for (int q = 0; q < ns_; ++q) {
            double lo, hi;
            lim(st_[q], lo, hi);
            if (pos >= lo && pos < hi) return st_[q].n;
        }
        return bg_;
    }

    double auto_size(double pos) const {
        const double n = index_at(pos);
        const double lam = lambda_ / n;
        return std::clamp(lam / ppw_, dmin_, dmax_);
    }

    void build_pieces() {
        std::set<double> cuts{0.0, L_};
        for (const auto& o : ov_) {
            cuts.insert(std::max(0.0, std::min(o.lo, L_)));
            cuts.insert(std::max(0.0, std::min(o.hi, L_)));
        }
        for (int q = 0; q < ns_; ++q) {
            double lo, hi;
            lim(st_[q], lo, hi);
            if (lo > 0 && lo < L_) cuts.insert(lo);
            if (hi > 0 && hi < L_) cuts.insert(hi);
        }
        const std::vector<double> c(cuts.begin(), cuts.end());
        pieces_.clear();
        for (size_t i = 0; i + 1 < c.size(); ++i) {
            const double mid = (c[i] + c[i + 1]) / 2;
            const Override1D* hit = nullptr;
            for (const auto& o : ov_)
                if (mid >= o.lo && mid <= o.hi) { hit = &o; break; }
            pieces_.push_back({c[i], c[i + 1], hit ? hit->d : auto_size(mid), hit != nullptr});
        }
    }

    // n equal cells from `from` to `to`, the last line exactly `to`
    double uniform(double from, double to, double len, double d) {
        const size_t n = std::max(size_t(1), (size_t)std::round(len / d));
        const double w = len / n;
        for (size_t i = 1; i <= n; ++i) lines_.push_back(i == n ? to : from + i * w);
        return w;
    }

    // close a sequence: fill `rest` with equal cells near `d`, within grading of `cur`
    double close_out(double& pos, double rest, double d, double cur) {
        int n = std::max(1, (int)std::round(rest / d));
        double w = rest / n;
        if (w / cur > g_ || cur / w > g_) {
            n = std::max(2, (int)std::ceil(rest / (cur * g_)));
            w = rest / n;
        }
        for (int j = 0; j < n; ++j) {
            pos += w;
            lines_.push_back(pos);
        }
        return w;
    }

    void fill_piece(size_t p) {
        const Piece& r = pieces_[p];
        if (r.fixed) {
            prev_ = uniform(r.lo, r.hi, r.hi - r.lo, r.target);
            return;
        }
        const bool after_fixed = p > 0 && pieces_[p - 1].fixed;
        const bool before_fixed = p + 1 < pieces_.size() && pieces_[p + 1].fixed;
        if (after_fixed && before_fixed) return between_overrides(p);
        const double air = lambda_ / bg_ / ppw_;
        const bool in_structure = (r.target < air * 0.6) && (r.target > dmin_ * 2);
        if (in_structure && !before_fixed) return inside_structure(p);
        graded(p, before_fixed);
    }

    // a gap between two override regions: grade up to a peak and back down (:474-563)
    void between_overrides(size_t p) {
        const Piece& r = pieces_[p];
        const double len = r.hi - r.lo;
        const double d0 = prev_, d1 = pieces_[p + 1].target;
        const double half = len / 2;
        double peak = d0, run = 0;
        while (run < half && peak < r.target) {
            peak *= g_;
            run += peak;
        }
        peak = std::min(peak, r.target);
        std::vector<double> seq;
        double d = d0;
        while (d < peak * 0.95) {
            d = std::min(d * g_, peak);
            seq.push_back(d);
        }
        const size_t n_up = seq.size();
        d = peak;
        while (d > d1 * 1.05) {
            d = std::max(d / g_, d1);
            seq.push_back(d);
        }
        double total = 0;
        for (size_t i = 0; i < n_up; ++i) total += seq[i];
        for (size_t i = n_up; i < seq.size(); ++i) total += seq[i];
        if (total > 0 && total <= len * 1.3) {
            const double k = len / total;
            double x = r.lo;
            for (size_t i = 0; i < seq.size(); ++i) {
                x += seq[i] * k;
                lines_.push_back(i + 1 == seq.size() ? r.hi : x);
            }
            prev_ = seq.back() * k;
            return;
        }
        prev_ = uniform(r.lo, r.hi, len, std::min(d0, d1));
    }

    // a structure piece not followed by an override: grade up from finer
    // cells if needed, then equal cells at the material's size (:565-635)
    void inside_structure(size_t p) {
        const Piece& r = pieces_[p];
        const double len = r.hi - r.lo;
        double pos = lines_.back();
        if (prev_ < r.target * 0.7) {
            std::vector<double> up;
            double d = prev_;
            while (d < r.target * 0.95) {
                d *= g_;
                if (d > r.target) d = r.target;
                up.push_back(d);
            }
            double run = 0;
            for (double u : up) run += u;
            if (run < len) {
                for (double u : up) {
                    pos += u;  // advanced also for the cell that is then not placed
                    if (pos > r.hi - dmin_) break;
                    lines_.push_back(pos);
                    prev_ = u;
                }
            }
        }
        const double rest = r.hi - pos;
        if (!(rest > dmin_)) return;
        double d = r.target;
        if (d / prev_ > g_) d = prev_ * g_;
        if (prev_ / d > g_) d = prev_ / g_;
        size_t n = std::max(size_t(1), (size_t)std::round(rest / d));
        double w = rest / n;
        if (w / prev_ > g_ || prev_ / w > g_) {
            w = prev_ < w ? std::min(w, prev_ * g_) : std::max(w, prev_ / g_);
            n = std::max(size_t(1), (size_t)std::round(rest / w));
            w = rest / n;
        }
        for (size_t i = 1; i <= n; ++i) lines_.push_back(i == n ? r.hi : pos + i * w);
        prev_ = w;
    }

    // air, or a piece leading into an override: entry, stable and exit
    // sequences (:637-840)
    void graded(size_t p, bool before_fixed) {
        const Piece& r = pieces_[p];
        const double len = r.hi - r.lo;
        const double end = r.hi;
        double pos = lines_.back();
        const double entry = prev_;
        double exit_to = r.target;
        bool shrink = false;
        if (before_fixed) {
            exit_to = pieces_[p + 1].target;
            shrink = true;
        } else if (p + 1 < pieces_.size() && pieces_[p + 1].target < r.target) {
            exit_to = pieces_[p + 1].target;
            shrink = true;
        }
        std::vector<double> in_seq, out_seq;
        if (entry < r.target * 0.99) {
            double d = entry;
            while (d < r.target * 0.99) {
                d = std::clamp(std::min(d * g_, r.target), dmin_, dmax_);
                in_seq.push_back(d);
            }
        }
        if (shrink && exit_to < r.target * 0.99) {
            double d = r.target;
            while (d > exit_to * 1.01) {
                d = std::clamp(std::max(d / g_, exit_to), dmin_, dmax_);
                out_seq.push_back(d);
            }
            if (out_seq.empty() || out_seq.back() > exit_to * 1.2) out_seq.push_back(exit_to);
        }
        double in_len = 0, out_len = 0;
        for (double d : in_seq) in_len += d;
        for (double d : out_seq) out_len += d;
        double stable_len = len - in_len - out_len;
        if (stable_len < 0) {
            const double k = len / (in_len + out_len) * 0.95;
            for (double& d : in_seq) d *= k;
            for (double& d : out_seq) d *= k;
            in_len *= k;
            out_len *= k;
            stable_len = len - in_len - out_len;
        }
        const double stable = std::clamp(r.target, dmin_, dmax_);
        double cur = entry;
        for (double d : in_seq) {  // entry
            const double w = std::clamp(std::clamp(d, cur / g_, cur * g_), dmin_, dmax_);
            if (pos + w > end - dmin_) break;
            pos += w;
            lines_.push_back(pos);
            cur = w;
            prev_ = w;
        }
        if (stable_len > stable * 0.5 && pos < end - out_len - dmin_) {  // stable run
            const double stop = end - out_len;
            const double rest = stop - pos;
            if (rest > 0) {
                double d = stable;
                if (d / cur > g_) d = cur * g_;
                if (cur / d > g_) d = cur / g_;
                d = std::clamp(d, dmin_, dmax_);
                const size_t n = std::max(size_t(1), (size_t)std::round(rest / d));
                const double w = rest / n;
                for (size_t i = 1; i <= n; ++i) lines_.push_back(i == n ? stop : pos + i * w);
                pos = stop;
                cur = w;
                prev_ = w;
            }
        }
        if (!out_seq.empty() && pos < end - dmin_ * 0.5) {  // exit
            for (size_t i = 0; i < out_seq.size(); ++i) {
                const double w = std::clamp(std::clamp(out_seq[i], cur / g_, cur * g_), dmin_, dmax_);
                const double rest = end - pos;
                if (rest < w * 0.5) break;
                if (i + 1 == out_seq.size() || rest < w * 1.5) {
                    prev_ = close_out(pos, rest, w, cur);
                    break;
                }
                pos += w;
                lines_.push_back(pos);
                cur = w;
                prev_ = w;
            }
        } else if (out_seq.empty() && pos < end - dmin_ * 0.5) {
            while (pos < end - dmin_ * 0.5) {
                const double rest = end - pos;
                const double w = std::clamp(std::clamp(stable, cur / g_, cur * g_), dmin_, dmax_);
                if (rest < w * 1.5) {
                    prev_ = close_out(pos, rest, w, cur);
                    break;
                }
                pos += w;
                lines_.push_back(pos);
                cur = w;
                prev_ = w;
            }
        }
        if (std::abs(lines_.back() - end) > dmin_ * 0.1 && lines_.back() < end) {
            lines_.push_back(end);
            prev_ = end - lines_[lines_.size() - 2];
        }
    }

    // fix_grading_by_subdivision (:859-897)
    void fix_grading() {
        const double tol = dmin_ * 0.01;
        for (int it = 0; it < 300; ++it) {
            double worst = 1.0;
            size_t at = 0;
            for (size_t i = 0; i + 2 < lines_.size(); ++i) {
                const double a = lines_[i + 1] - lines_[i], b = lines_[i + 2] - lines_[i + 1];
                if (a < tol || b < tol) continue;
                const double q = std::max(a / b, b / a);
                if (q > worst) {
                    worst = q;
                    at = i;
                }
            }
            if (worst <= g_ * 1.005) break;
            const double a = lines_[at + 1] - lines_[at], b = lines_[at + 2] - lines_[at + 1];
            const size_t s = a > b ? at : at + 1;
            if (lines_[s + 1] - lines_[s] < dmin_ * 1.5) break;
            lines_.insert(lines_.begin() + s + 1, 0.5 * (lines_[s] + lines_[s + 1]));
        }
    }

    int axis_;
    double L_;
    const fdtd_mesh_bounds* st_;
    int ns_;
    double bg_, dmin_, dmax_, g_, lambda_, ppw_;
    bool graded_pml_;
    std::vector<Override1D> ov_;
    std::vector<Piece> pieces_;
    std::vector<double> lines_;
    double prev_ = 0;
};

}  // namespace

extern "C" int fdtd_auto_mesh_axis(const fdtd_auto_mesh_config* cfg, int axis, double domain, int npml,
                                   int n_structures, const fdtd_mesh_bounds* structures, int n_overrides,
                                   const fdtd_mesh_override* overrides, double background_n, double* lines,
                                   long capacity, long* n_lines) {
    if (n_lines) *n_lines = 0;
    if (!cfg || !n_lines || axis < 0 || axis > 2 || npml < 0 || n_structures < 0 || n_overrides < 0 ||
        (n_structures > 0 && !structures) || (n_overrides > 0 && !overrides) || capacity < 0 ||
        (capacity > 0 && !lines))
        return FDTD_EINVAL;
    // what keeps every loop finite: a positive domain, index and cell-size range
    // with dx_min <= dx_max, a grading ratio above 1 and a positive wavelength
    if (!(domain > 0) || !(background_n > 0) || !(cfg->dx_min > 0) || !(cfg->dx_max >= cfg->dx_min) ||
        !(cfg->max_grading_ratio > 1) || !std::isfinite(domain) || !std::isfinite(cfg->dx_max))
        return FDTD_EINVAL;
    if (!(cfg->f_max > 0) && !(cfg->lambda_min > 0)) return FDTD_EINVAL;
    for (int q = 0; q < n_structures; ++q)
        if (!(structures[q].n > 0)) return FDTD_EINVAL;
    AxisMesher m(*cfg, axis, domain, structures, n_structures, background_n);
    if (cfg->override_enabled)
        for (int q = 0; q < n_overrides; ++q)
            if (overrides[q].enabled) m.add_override(overrides[q]);
    const std::vector<double> out = m.with_pml(m.physical_lines(), (size_t)npml);
    *n_lines = (long)out.size();
    if ((long)out.size() > capacity) return lines ? FDTD_EINVAL : FDTD_OK;
    std::copy(out.begin(), out.end(), lines);
    return FDTD_OK;
}

extern "C" int fdtd_structure_mesh_bounds(int type, const double* p, double n, fdtd_mesh_bounds* out) {
    if (!p || !out) return FDTD_EINVAL;
    switch (type) {
    case FDTD_SHAPE_BOX:
        *out = {p[0], p[1], p[2], p[3], p[4], p[5], n};
        return FDTD_OK;
    case FDTD_SHAPE_SPHERE:
        *out = {p[0] - p[3], p[0] + p[3], p[1] - p[3], p[1] + p[3], p[2] - p[3], p[2] + p[3], n};
        return FDTD_OK;
    case FDTD_SHAPE_CYLINDER_Z:
        *out = {p[0] - p[2], p[0] + p[2], p[1] - p[2], p[1] + p[2], p[3], p[4], n};
        return FDTD_OK;
    default:
        return FDTD_EINVAL;
    }
}

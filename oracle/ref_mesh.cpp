// ref_mesh.cpp - TEST INFRASTRUCTURE ONLY: exposes the reference's own
// AutoMeshGenerator (src/auto_mesh_generator.hpp, compiled from the
// reference tree) through a C entry point, so the library's restatement
// (fdtd_auto_mesh_axis) can be compared with it bit for bit.
#include <iostream>
#include <sstream>
#include <vector>

#include "auto_mesh_generator.hpp"

extern "C" {

// cfg: lambda_min, f_max, mesh_accuracy, ppw_override, dx_min, dx_max,
// max_grading_ratio, pml_graded, override_enabled (as doubles).
// structures: 7 doubles each (x_min, x_max, y_min, y_max, z_min, z_max, n);
// overrides: 11 doubles each (enabled, x0, x1, y0, y1, z0, z1, dx, dy, dz, priority).
// Writes the x, y, z bounds into out[0..cap), out[cap..2cap), out[2cap..3cap)
// and their lengths into counts[3]; returns 0, or -1 if cap is too small.
int ref_auto_mesh(const double* cfg, double dom_x, double dom_y, double dom_z, long npml, int ns,
                  const double* st, int no, const double* ov, double background_n, double* out, long cap,
                  long* counts) {
    AutoMeshConfig c;
    c.lambda_min = cfg[0];
    c.f_max = cfg[1];
    c.mesh_accuracy = (int)cfg[2];
    c.ppw_override = cfg[3];
    c.dx_min = cfg[4];
    c.dx_max = cfg[5];
    c.max_grading_ratio = cfg[6];
    c.pml_strategy = cfg[7] != 0 ? PMLStrategy::GRADED : PMLStrategy::UNIFORM;
    c.mesh_override_enabled = cfg[8] != 0;
    for (int q = 0; q < no; ++q) {
        const double* o = ov + 11 * q;
        MeshOverrideRegion r;
        r.enabled = o[0] != 0;
        r.x0 = o[1]; r.x1 = o[2]; r.y0 = o[3]; r.y1 = o[4]; r.z0 = o[5]; r.z1 = o[6];
        r.dx = o[7]; r.dy = o[8]; r.dz = o[9];
        r.priority = (int)o[10];
        c.override_regions.push_back(r);
    }
    std::vector<AutoMeshStructureBounds> s;
    for (int q = 0; q < ns; ++q) {
        const double* b = st + 7 * q;
        s.push_back({b[0], b[1], b[2], b[3], b[4], b[5], b[6]});
    }
    std::ostringstream sink;
    std::streambuf* old_out = std::cout.rdbuf(sink.rdbuf());
    std::streambuf* old_err = std::cerr.rdbuf(sink.rdbuf());
    AutoMeshGenerator gen(c);
    GridSpacing g = gen.generate(dom_x, dom_y, dom_z, (size_t)npml, s, background_n);
    std::cout.rdbuf(old_out);
    std::cerr.rdbuf(old_err);
    const std::vector<real>* b[3] = {&g.x_bounds, &g.y_bounds, &g.z_bounds};
    int rc = 0;
    for (int a = 0; a < 3; ++a) {
        counts[a] = (long)b[a]->size();
        if (counts[a] > cap) { rc = -1; continue; }
        for (size_t i = 0; i < b[a]->size(); ++i) out[a * cap + (long)i] = (*b[a])[i];
        // GridSpacing::dx is filled from the same lines: check that here once
        for (size_t i = 0; i + 1 < b[a]->size(); ++i) {
            const std::vector<real>& d = a == 0 ? g.dx : a == 1 ? g.dy : g.dz;
            if (d[i] != (*b[a])[i + 1] - (*b[a])[i]) rc = -2;
        }
    }
    return rc;
}
}

// fdtd_bake.hpp - StructureScene::bake (src/structure_material.hpp:211-288)
// on the device.
//
// Every Yee component of a cell averages eps, mu, sigma and sigma_m over s^3
// subpixel samples at its stagger offset (:246-260, offsets :263-265); the
// material at a sample is that of the last structure containing it, the
// background otherwise (:215-221).  The coefficients follow :270-283.  One
// thread per cell evaluates all 6 * s^3 samples with the same FP64
// operations in the same order as the host (the library is built with
// -ffp-contract=off and FP64 division is correctly rounded), so each value
// equals the host bake's bit for bit.
//
// The device holds materials as a per-cell id into a table of distinct
// coefficient tuples.  A cell whose samples all fall in one material (a
// "pure" cell) records that material's index and marks it used; the host
// computes the tuple of each used material once, with the same operation
// sequence.  A cell with mixed samples (an interface cell) appends its index
// and its 12 coefficients to a list.  The host deduplicates the tuples and
// builds the table; k_ids_remap turns material indices into table ids and
// k_bake_scatter writes the ids of the interface cells.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdtdk {

struct BakeShape {
    int type;     // 0 box, 1 sphere, 2 z-cylinder
    double p[6];  // as fdtd_bake_shape::p
};

struct BakeArgs {
    int nx, ny, nz;
    int ns;   // samples per axis (s)
    int nsh;  // structures
    const double *xb, *yb, *zb;  // GridSpacing bounds
    const double *dx, *dy, *dz;
    const BakeShape* sh;
    const double* props;   // [nsh + 1][4]: eps_r, mu_r, sigma, sigma_m; row 0 = background
    double dt, eps0, mu0;
    uint16_t* ids;         // out: per-cell material index (0xffff on interface cells)
    uint8_t* used;         // out: [nsh + 1], 1 where some cell is pure in that material
    unsigned long long* count;
    unsigned long long cap;
    long long* icell;      // [cap] interface cell index
    double* itup;          // [cap][12] its coefficients (aEx bEx ... aHz bHz)
};

// Box / Sphere / CylinderZ::contains, src/structure_material.hpp:47-93.
__device__ __forceinline__ bool bake_contains(const BakeShape& s, double x, double y, double z) {
    if (s.type == 0) return x >= s.p[0] && x < s.p[1] && y >= s.p[2] && y < s.p[3] && z >= s.p[4] && z < s.p[5];
    if (s.type == 1) {
        const double dx = x - s.p[0], dy = y - s.p[1], dz = z - s.p[2];
        return (dx * dx + dy * dy + dz * dz) <= s.p[3] * s.p[3];
    }
    if (z < s.p[3] || z >= s.p[4]) return false;
    const double dx = x - s.p[0], dy = y - s.p[1];
    return (dx * dx + dy * dy) <= s.p[2] * s.p[2];
}

// grid (ceil(nz / 256), ny, nx), 256 threads: one cell per thread.
__global__ void __launch_bounds__(256) k_bake(BakeArgs a) {
    const int k = (int)(blockIdx.x * 256 + threadIdx.x), j = (int)blockIdx.y, i = (int)blockIdx.z;
    if (k >= a.nz) return;
    const long long id = ((long long)i * a.ny + j) * a.nz + k;
    // Ex Ey Ez Hx Hy Hz stagger offsets (src/structure_material.hpp:263-265)
    const double off[6][3] = {{0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0},
                              {0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}};
    const double dxl = a.dx[i], dyl = a.dy[j], dzl = a.dz[k];
    const double x0 = a.xb[i], y0 = a.yb[j], z0 = a.zb[k];
    const int ns = a.ns;
    double tup[12];
    int m0 = -1;
    bool pure = true;
    for (int c = 0; c < 6; ++c) {
        double eps_sum = 0, mu_sum = 0, sg_sum = 0, sgm_sum = 0;
        int cnt = 0;
        for (int sa = 0; sa < ns; ++sa)
            for (int sb = 0; sb < ns; ++sb)
                for (int sc = 0; sc < ns; ++sc) {
                    const double xs = x0 + (off[c][0] + (sa + 0.5) / ns) * dxl;
                    const double ys = y0 + (off[c][1] + (sb + 0.5) / ns) * dyl;
                    const double zs = z0 + (off[c][2] + (sc + 0.5) / ns) * dzl;
                    int m = 0;
                    for (int q = 0; q < a.nsh; ++q)
                        if (bake_contains(a.sh[q], xs, ys, zs)) m = q + 1;
                    if (m0 < 0) m0 = m;
                    pure &= (m == m0);
                    const double* pr = a.props + 4 * m;
                    eps_sum += pr[0] * a.eps0;
                    mu_sum += pr[1] * a.mu0;
                    sg_sum += pr[2];
                    sgm_sum += pr[3];
                    ++cnt;
                }
        const double eps = eps_sum / cnt, mu = mu_sum / cnt, sigma = sg_sum / cnt, sigm = sgm_sum / cnt;
        if (c < 3) {
            const double c1 = sigma * a.dt / (2 * eps);
            tup[2 * c] = (1 - c1) / (1 + c1);
            tup[2 * c + 1] = (a.dt / eps) / (1 + c1);
        } else {
            const double c2 = sigm * a.dt / (2 * mu);
            tup[2 * c] = (1 - c2) / (1 + c2);
            tup[2 * c + 1] = (a.dt / mu) / (1 + c2);
        }
    }
    if (pure) {
        a.ids[id] = (uint16_t)m0;
        a.used[m0] = 1;  // every writer stores the same value
        return;
    }
    a.ids[id] = 0xffff;
    const unsigned long long slot = atomicAdd(a.count, 1ull);
    if (slot < a.cap) {
        a.icell[slot] = id;
        for (int q = 0; q < 12; ++q) a.itup[slot * 12 + q] = tup[q];
    }
}

// Table ids of the pure cells: out[c] = remap[in[c]] (interface cells, 0xffff,
// get 0 until k_bake_scatter writes them).  ID is uint8_t or uint16_t.
template <typename ID>
__global__ void k_ids_remap(const uint16_t* in, const uint16_t* remap, ID* out, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const uint16_t v = in[t];
        out[t] = (ID)(v == 0xffff ? 0 : remap[v]);
    }
}

// ids[icell[t]] = iid[t] for the n interface cells.
template <typename ID>
__global__ void k_bake_scatter(ID* ids, const long long* icell, const uint16_t* iid, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) ids[icell[t]] = (ID)iid[t];
}

}  // namespace fdtdk

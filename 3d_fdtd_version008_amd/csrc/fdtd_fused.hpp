// fdtd_fused.hpp - one-launch Yee step: H^{n+1/2} and E^{n+1} (with CPML,
// J, gain) of the whole grid in a single pass over HBM.
//
// Bytes: the two-kernel step reads E and H twice and writes each once
// (18 s per cell plus psi, ids and gain state); this kernel reads E^n and
// H^{n-1/2} once and writes H^{n+1/2} and E^{n+1} once (12 s).
//
// Shape (k_step_march below).  A work-group is the kc k-chunk wavefronts of
// one block of JW consecutive rows j and a chunk of consecutive x-planes;
// a wavefront holds CW cells per lane of a row (k = kbase + 64 q + lane,
// q < CW), so every load of a cell block is one coalesced row segment.  The
// wavefront marches along x and, inside each plane, walks the halo row
// j0 - 1 (H only, nothing stored) and then the block's rows in order.
//
// All reads come from buffer A (E^n, H^{n-1/2}, psi_H^{n-1/2}) and all field
// writes go to buffer B, so the halo row and the row above are read while
// their owners write, without ordering between work-groups.  psi_E and the
// gain state have one owner each and are updated in place.  Cells the
// reference never updates are never written and hold the same values in
// both buffers.
//
// Per cell the arithmetic is h_compute / e_compute (fdtd_kernels.hpp): the
// reference's operations in the reference's order, so the results are
// identical to the two-kernel step in every storage format.
#pragma once
#include "fdtd_kernels.hpp"

namespace fdtdk {

constexpr int FZ_MAX_CW = 4;  // cells per lane per row chunk

struct FusedGeom {
    int kc;   // k-chunk wavefronts per work-group
    int jw;   // rows per work-group (plus the halo row)
    int ic;   // x-planes per work-group
    int njb;  // j-blocks
    int nwg;  // work-groups
};

// Lane l receives x from lane l - 1 (DPP wave_ror:1; lane 0 from lane 63).
__device__ __forceinline__ int dpp_prev_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x13C, 0xF, 0xF, false); }
// Lane l receives x from lane l + 1 (DPP wave_rol:1; lane 63 from lane 0).
__device__ __forceinline__ int dpp_next_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x134, 0xF, 0xF, false); }
template <bool NEXT, typename R>
__device__ __forceinline__ R lane_shift(R x) {
    if constexpr (sizeof(R) == 8) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(x);
        int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
        lo = NEXT ? dpp_next_i(lo) : dpp_prev_i(lo);
        hi = NEXT ? dpp_next_i(hi) : dpp_prev_i(hi);
        return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    } else {
        const int v = NEXT ? dpp_next_i(__float_as_int(x)) : dpp_prev_i(__float_as_int(x));
        return __int_as_float(v);
    }
}

// Work-group barrier for LDS traffic only: waits for this wavefront's LDS
// (and scalar) operations, then meets the others.  Global loads and stores
// in flight stay in flight (a __syncthreads fence would wait for them).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The E-side inputs of one cell other than the fields and the z inverse
// spacing: material coefficients (id m), psi_E, gain state, axis values,
// sources.  Wave-uniform
// conditions select the loads (the row is uniform; kw is the first k of the
// wavefront's 64-cell block), with clamped indices and opaque defaults as in
// e_gather.
template <typename T, typename R, typename IX>
__device__ __forceinline__ void fz_e_inputs(const StepArgs<T, R>& a, const Tabs<R>& tb, int i, int j, int k, int kw,
                                            int m, EIn<R>& v) {
    const int np = a.np, nx = a.nx, ny = a.ny, nz = a.nz;
    const int kk = k < nz ? k : nz - 1;
    const int gi = i - a.g0x, gj = j - a.g0y, gk = kk - a.g0z;
    const bool growu = a.gain && ((unsigned)gi < (unsigned)a.gnx) & ((unsigned)gj < (unsigned)a.gny) &
                       (kw < a.g0z + a.gnz) & (kw + 63 >= a.g0z);
    const bool gin = growu && ((unsigned)gk < (unsigned)a.gnz);
    const IX g = ((IX)gi * a.gny + gj) * a.gnz + (gin ? gk : 0);
    int r8 = opaque(-1);
    if (growu) r8 = *at(a.greg, g);
    using RawT = typename Raw<T>::type;
    RawT praw[6];
    for (int q = 0; q < 6; ++q) praw[q] = opaque((RawT)0);
    const bool px = a.cpml && in_pml(i, nx, np);
    const bool py = a.cpml && in_pml(j, ny, np);
    const bool pzw = a.cpml && ((kw < np) | (kw + 63 >= nz - np));
    if (py) {
        const IX q = psi_y<IX>(i, j, kk, nx, ny, nz, np);
        praw[0] = *at(a.psi[0], q);
        praw[5] = *at(a.psi[5], q);
    }
    if (pzw) {
        const IX q = psi_z<IX>(i, j, in_pml(kk, nz, np) ? kk : 0, nx, ny, nz, np);
        praw[1] = *at(a.psi[1], q);
        praw[2] = *at(a.psi[2], q);
    }
    if (px) {
        const IX q = psi_x<IX>(i, j, kk, nx, ny, nz, np);
        praw[3] = *at(a.psi[3], q);
        praw[4] = *at(a.psi[4], q);
    }
    double gs[4];
    for (int q = 0; q < 4; ++q) gs[q] = opaque(0.0);
    if (growu) { gs[0] = *at(a.gNu, g); gs[1] = *at(a.gNg, g); gs[2] = *at(a.gPz, g); gs[3] = *at(a.gPp, g); }
    const Axes<R>& ax = tb.ax;
    v.idx = ax.X(0, i);
    v.idy = ax.Y(0, j);
    for (int q = 0; q < 4; ++q) v.kx[q] = v.ky[q] = 0;
    if (px) { v.kx[0] = ax.X(1, i); v.kx[1] = ax.X(6, i); v.kx[2] = ax.X(2, i); v.kx[3] = ax.X(3, i); }
    if (py) { v.ky[0] = ax.Y(1, j); v.ky[1] = ax.Y(6, j); v.ky[2] = ax.Y(2, j); v.ky[3] = ax.Y(3, j); }
    for (int q = 0; q < 4; ++q) v.kz[q] = opaque((R)0);
    if (pzw) { v.kz[0] = ax.Z(1, kk); v.kz[1] = ax.Z(6, kk); v.kz[2] = ax.Z(2, kk); v.kz[3] = ax.Z(3, kk); }
    e_sources(tb, a.nsrc, i, j, k, v.jx, v.jy, v.jz);
    if (a.gain && a.ntls == 1) v.tc = tb.tls[0];
    v.cE[1] = a.lut[1 * a.nmat + m];
    v.cE[3] = a.lut[3 * a.nmat + m];
    v.cE[5] = a.lut[5 * a.nmat + m];
    v.cE[0] = v.cE[2] = v.cE[4] = 0;
    if (!a.lossless) {
        v.cE[0] = a.lut[0 * a.nmat + m];
        v.cE[2] = a.lut[2 * a.nmat + m];
        v.cE[4] = a.lut[4 * a.nmat + m];
    }
    if (a.gain && a.ntls != 1) v.tc = tb.tls[r8 < 0 ? 0 : r8];
    v.reg = gin && k < nz ? r8 : -1;
    v.g = (long)g;
    v.Nu = gs[0]; v.Ng = gs[1]; v.Pz = gs[2]; v.Pp = gs[3];
    for (int q = 0; q < 6; ++q) v.p[q] = cv<T, R>(praw[q], a.sc[6 + q]);
    v.hns[0] = v.hns[1] = v.hns[2] = (R)1;  // the caller's H neighbours are values
}

// psi_H of one cell (zeros outside the slabs), raw loads with clamped
// indices under wave-uniform conditions.
template <typename T, typename R, typename IX>
__device__ __forceinline__ void fz_psi_h(const StepArgs<T, R>& a, int i, int j, int kk, int kw, R* p, R* kz,
                                         const Tabs<R>& tb) {
    const int np = a.np, nx = a.nx, ny = a.ny, nz = a.nz;
    using RawT = typename Raw<T>::type;
    RawT praw[6];
    for (int q = 0; q < 6; ++q) praw[q] = opaque((RawT)0);
    const bool px = a.cpml && in_pml(i, nx, np);
    const bool py = a.cpml && in_pml(j, ny, np);
    const bool pzw = a.cpml && ((kw < np) | (kw + 63 >= nz - np));
    if (py) {
        const IX q = psi_y<IX>(i, j, kk, nx, ny, nz, np);
        praw[0] = *at(a.psi[6], q);
        praw[5] = *at(a.psi[11], q);
    }
    if (pzw) {
        const IX q = psi_z<IX>(i, j, in_pml(kk, nz, np) ? kk : 0, nx, ny, nz, np);
        praw[1] = *at(a.psi[7], q);
        praw[2] = *at(a.psi[8], q);
    }
    if (px) {
        const IX q = psi_x<IX>(i, j, kk, nx, ny, nz, np);
        praw[3] = *at(a.psi[9], q);
        praw[4] = *at(a.psi[10], q);
    }
    for (int q = 0; q < 4; ++q) kz[q] = opaque((R)0);
    if (pzw) {
        const Axes<R>& ax = tb.ax;
        kz[0] = ax.Z(1, kk); kz[1] = ax.Z(6, kk); kz[2] = ax.Z(4, kk); kz[3] = ax.Z(5, kk);
    }
    for (int q = 0; q < 6; ++q) p[q] = cv<T, R>(praw[q], a.sc[12 + q]);
}

// The kernel reads its arguments through the kernarg segment pointer,
// laundered at each phase, instead of a by-value copy: the compiler then
// loads the few fields a phase uses (scalar loads, constant cache) rather
// than holding every array pointer of StepArgs in scalar registers for the
// whole march (which spilled them into vector registers).  Each array
// pointer is marked as global memory (an address-space cast the compiler
// propagates), so accesses through it are global, not flat, loads and stores.
template <typename T, typename R>
using KArgs = const __attribute__((address_space(4))) StepArgs<T, R>*;
template <typename P>
__device__ __forceinline__ P* gptr(P* p) {
    return (P*)(__attribute__((address_space(1))) P*)p;
}
template <typename T, typename R>
__device__ __forceinline__ StepArgs<T, R> phase_args(KArgs<T, R> ka) {
    asm volatile("" : "+s"(ka));
    StepArgs<T, R> a = *(const StepArgs<T, R>*)ka;
    a.Ex = gptr(a.Ex); a.Ey = gptr(a.Ey); a.Ez = gptr(a.Ez);
    a.Hx = gptr(a.Hx); a.Hy = gptr(a.Hy); a.Hz = gptr(a.Hz);
    a.Exw = gptr(a.Exw); a.Eyw = gptr(a.Eyw); a.Ezw = gptr(a.Ezw);
    a.Hxw = gptr(a.Hxw); a.Hyw = gptr(a.Hyw); a.Hzw = gptr(a.Hzw);
    a.cpx = gptr(a.cpx); a.cpy = gptr(a.cpy); a.cpz = gptr(a.cpz);
    for (int q = 0; q < 12; ++q) { a.psi[q] = gptr(a.psi[q]); a.psiw[q] = gptr(a.psiw[q]); }
    a.sc = gptr(a.sc);
    a.matid = gptr(a.matid);
    a.lut = gptr(a.lut);
    a.src = gptr(a.src);
    a.srcv = gptr(a.srcv);
    a.greg = gptr(a.greg);
    a.gPz = gptr(a.gPz); a.gPp = gptr(a.gPp); a.gNu = gptr(a.gNu); a.gNg = gptr(a.gNg);
    a.tls = gptr(a.tls);
    return a;
}

// ---- The fused step kernel (k_step_march) ----
//
// Work-groups do not wait on each other: a work-group is only the kc
// k-chunk wavefronts of one row block (kc = 1 to 3 on most grids), and each
// wavefront walks the block's JW rows in order inside every plane:
//   halo row j0 - 1: H^{n+1/2} (Hx, Hz only used, nothing stored);
//   row j0 + r:      H^{n+1/2}(i) -> store; E^{n+1}(i) with H(j-1) from the
//                    row before it (registers), H(i-1) carried from the
//                    previous plane, H(k-1) by DPP rotation (the chunk's
//                    first lane from the previous chunk's wavefront through
//                    LDS: the only barrier, and only between kc wavefronts).
// Per plane a wavefront keeps E^n(i) and H^{n+1/2}(i-1) of its JW rows in
// registers (carried to the next plane), so the block's own data is read
// once; the halo row is the only recomputation (Hx, Hz of one row in
// JW + 1), and the halo row and the row above the block are read again by
// the neighbouring blocks.  The first plane of each chunk also recomputes
// H of the plane before it (not stored) for the carried H(i0-1).
constexpr int FZ_MARCH_MAX_KC = 8;  // k-chunk wavefronts per work-group

template <typename T, typename R, int CW, int JW>
__global__ void __launch_bounds__(64 * FZ_MARCH_MAX_KC) k_step_march(StepArgs<T, R> args, FusedGeom gm) {
    (void)args;
    using IX = unsigned;
    using SA = StepArgs<T, R>;
    const KArgs<T, R> ka = (KArgs<T, R>)__builtin_amdgcn_kernarg_segment_ptr();
    // the last Hx, Hy of each wavefront's chunk, two alternating slots
    __shared__ R kslot[2][FZ_MARCH_MAX_KC][2];
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // k-chunk
    const int lane = (int)(threadIdx.x & 63);
    // work-groups are dealt round-robin to the 8 XCDs: each XCD gets a
    // contiguous run of (j-block, i-chunk) tiles, so neighbouring row blocks
    // (whose rows are each other's halo and row above) share an L2
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned per = (nb + 7) / 8;
    const unsigned t = (nb % 8 == 0) ? (b % 8) * per + b / 8 : b;
    const int jb = (int)(t % (unsigned)gm.njb), ich = (int)(t / (unsigned)gm.njb);
    int nx, ny, nz;
    nx = ka->nx; ny = ka->ny; nz = ka->nz;
    const int j0 = jb * JW;
    const int nrow = min(JW, ny - j0);  // own rows of the block
    const bool hal = j0 > 0;            // the halo row j0 - 1 exists
    const int jh = hal ? j0 - 1 : 0;
    const int i0 = ich * gm.ic, i1 = min(i0 + gm.ic, nx);
    const int kbase = w * CW * 64;
    int kq[CW], kk[CW];
    for (int q = 0; q < CW; ++q) {
        kq[q] = kbase + 64 * q + lane;
        kk[q] = kq[q] < nz ? kq[q] : nz - 1;
    }
    const int knext = min(kbase + 64 * CW, nz - 1);  // first k of the next chunk

    // carried: E^n(i) and H^{n+1/2}(i-1) (Hy, Hz) of the own rows, E^n(i)
    // of the halo row, the z inverse spacing
    R ex[JW][CW], ey[JW][CW], ez[JW][CW], hpy[JW][CW], hpz[JW][CW];
    R gx[CW], gy[CW], gz[CW], idz[CW];
    for (int r = 0; r < JW; ++r)
        for (int q = 0; q < CW; ++q) ex[r][q] = ey[r][q] = ez[r][q] = hpy[r][q] = hpz[r][q] = 0;
    for (int q = 0; q < CW; ++q) gx[q] = gy[q] = gz[q] = 0;

    // first plane of the march: i0 - 1 (H only, for the carried H(i0-1)) or i0
    const int ps = i0 > 0 ? i0 - 1 : i0;
    {
        const SA a = phase_args<T, R>(ka);
#pragma unroll
        for (int r = 0; r < JW; ++r) {
            if (r >= nrow) continue;
            for (int q = 0; q < CW; ++q) {
                const IX id = ((IX)ps * ny + (j0 + r)) * nz + kk[q];
                ex[r][q] = ldx<T, R>(a.Ex, id, (R)a.sc[0]);
                ey[r][q] = ldx<T, R>(a.Ey, id, (R)a.sc[1]);
                ez[r][q] = ldx<T, R>(a.Ez, id, (R)a.sc[2]);
            }
        }
        if (hal) {
            for (int q = 0; q < CW; ++q) {
                const IX id = ((IX)ps * ny + jh) * nz + kk[q];
                gx[q] = ldx<T, R>(a.Ex, id, (R)a.sc[0]);
                gy[q] = ldx<T, R>(a.Ey, id, (R)a.sc[1]);
                gz[q] = ldx<T, R>(a.Ez, id, (R)a.sc[2]);
            }
        }
        for (int q = 0; q < CW; ++q) idz[q] = a.cpz[kk[q]];
    }

    int slot = 0;
    for (int i = ps; i < i1; ++i) {
        const bool honly = i < i0;  // the plane before the chunk: H only
        const int ip = min(i + 1, nx - 1);
        R hjx[CW], hjz[CW];  // H^{n+1/2}(i) Hx, Hz of the row below the current one
        for (int q = 0; q < CW; ++q) hjx[q] = hjz[q] = 0;
        R gnx[CW], gny[CW], gnz[CW];  // E^n(i+1) of the halo row
        for (int q = 0; q < CW; ++q) { gnx[q] = gx[q]; gny[q] = gy[q]; gnz[q] = gz[q]; }
        if (hal) {
            // ---- halo row: E^n(i+1) for the carry, Hx and Hz of H^{n+1/2}(i) ----
            const SA a = phase_args<T, R>(ka);
            const Tabs<R> tb = global_tabs(a);
            for (int q = 0; q < CW; ++q) {
                const IX idn = ((IX)ip * ny + jh) * nz + kk[q];
                gnx[q] = ldx<T, R>(a.Ex, idn, (R)a.sc[0]);
                gny[q] = ldx<T, R>(a.Ey, idn, (R)a.sc[1]);
                gnz[q] = ldx<T, R>(a.Ez, idn, (R)a.sc[2]);
            }
            if (!honly) {
                R hox[CW], hoy[CW], hoz[CW];
                for (int q = 0; q < CW; ++q) {
                    const IX id = ((IX)i * ny + jh) * nz + kk[q];
                    hox[q] = ldx<T, R>(a.Hx, id, (R)a.sc[3]);
                    hoy[q] = ldx<T, R>(a.Hy, id, (R)a.sc[4]);
                    hoz[q] = ldx<T, R>(a.Hz, id, (R)a.sc[5]);
                }
                const IX idk = ((IX)i * ny + jh) * nz + knext;
                const R ekx = ldx<T, R>(a.Ex, idk, (R)a.sc[0]);
                const R eky = ldx<T, R>(a.Ey, idk, (R)a.sc[1]);
                R ph[CW][6], kzh[CW][4];
                for (int q = 0; q < CW; ++q) fz_psi_h<T, R, IX>(a, i, jh, kk[q], kbase + 64 * q, ph[q], kzh[q], tb);
                R rx[CW], ry[CW];
                for (int q = 0; q < CW; ++q) {
                    rx[q] = lane_shift<true>(gx[q]);
                    ry[q] = lane_shift<true>(gy[q]);
                }
                const R idx = tb.ax.X(0, i), idy = tb.ax.Y(0, jh);
                R kxr[4] = {0, 0, 0, 0}, kyr[4] = {0, 0, 0, 0};
                if (a.cpml && in_pml(i, nx, a.np)) {
                    kxr[0] = tb.ax.X(1, i); kxr[1] = tb.ax.X(6, i); kxr[2] = tb.ax.X(4, i); kxr[3] = tb.ax.X(5, i);
                }
                if (a.cpml && in_pml(jh, ny, a.np)) {
                    kyr[0] = tb.ax.Y(1, jh); kyr[1] = tb.ax.Y(6, jh); kyr[2] = tb.ax.Y(4, jh); kyr[3] = tb.ax.Y(5, jh);
                }
                for (int q = 0; q < CW; ++q) {
                    const bool last = lane == 63;
                    HIn<R> v;
                    v.ex = gx[q]; v.ey = gy[q]; v.ez = gz[q];
                    v.ex_j = ex[0][q]; v.ez_j = ez[0][q];
                    v.ey_i = gny[q]; v.ez_i = gnz[q];
                    v.ex_k = last ? (q + 1 < CW ? rx[q + 1] : ekx) : rx[q];
                    v.ey_k = last ? (q + 1 < CW ? ry[q + 1] : eky) : ry[q];
                    v.hx = hox[q]; v.hy = hoy[q]; v.hz = hoz[q];
                    v.ns[0] = v.ns[1] = v.ns[2] = (R)1;  // neighbours are values
                    for (int u = 0; u < 6; ++u) { v.p[u] = ph[q][u]; v.cH[u] = tb.hco[u * a.nmat]; }
                    v.upd = (i < nx - 1) & (jh < ny - 1) & (kq[q] < nz - 1);
                    v.idx = idx; v.idy = idy; v.idz = idz[q];
                    for (int u = 0; u < 4; ++u) { v.kx[u] = kxr[u]; v.ky[u] = kyr[u]; v.kz[u] = kzh[q][u]; }
                    const H3<R> h = h_compute<T, R, true, IX>(a, tb, i, jh, kk[q], v, false);
                    hjx[q] = h.x; hjz[q] = h.z;
                }
            }
        }
        // ---- the block's rows, in order ----
#pragma unroll
        for (int r = 0; r < JW; ++r) {
            if (r >= nrow) continue;
            const int j = j0 + r;
            R hx[CW], hy[CW], hz[CW], enx[CW], eny[CW], enz[CW];
            int mid[CW];
            {
                const SA a = phase_args<T, R>(ka);
                const Tabs<R> tb = global_tabs(a);
                R hox[CW], hoy[CW], hoz[CW], ejx[CW], ejz[CW];
                const int jn = min(j + 1, ny - 1);
                for (int q = 0; q < CW; ++q) {
                    const IX id = ((IX)i * ny + j) * nz + kk[q];
                    const IX idn = ((IX)ip * ny + j) * nz + kk[q];
                    enx[q] = ldx<T, R>(a.Ex, idn, (R)a.sc[0]);
                    eny[q] = ldx<T, R>(a.Ey, idn, (R)a.sc[1]);
                    enz[q] = ldx<T, R>(a.Ez, idn, (R)a.sc[2]);
                    hox[q] = ldx<T, R>(a.Hx, id, (R)a.sc[3]);
                    hoy[q] = ldx<T, R>(a.Hy, id, (R)a.sc[4]);
                    hoz[q] = ldx<T, R>(a.Hz, id, (R)a.sc[5]);
                    mid[q] = 0;
                    if (!honly) mid[q] = mat_of<R, IX>(a.matid, a.mid16, id);
                }
                if (r + 1 < nrow) {  // the row above is the block's next row: its carried E^n(i)
                    for (int q = 0; q < CW; ++q) {
                        ejx[q] = ex[r + 1 < JW ? r + 1 : r][q];
                        ejz[q] = ez[r + 1 < JW ? r + 1 : r][q];
                    }
                } else {
                    for (int q = 0; q < CW; ++q) {
                        const IX idj = ((IX)i * ny + jn) * nz + kk[q];
                        ejx[q] = ldx<T, R>(a.Ex, idj, (R)a.sc[0]);
                        ejz[q] = ldx<T, R>(a.Ez, idj, (R)a.sc[2]);
                    }
                }
                const IX idk = ((IX)i * ny + j) * nz + knext;
                const R ekx = ldx<T, R>(a.Ex, idk, (R)a.sc[0]);
                const R eky = ldx<T, R>(a.Ey, idk, (R)a.sc[1]);
                R ph[CW][6], kzh[CW][4];
                for (int q = 0; q < CW; ++q) fz_psi_h<T, R, IX>(a, i, j, kk[q], kbase + 64 * q, ph[q], kzh[q], tb);
                R rx[CW], ry[CW];
                for (int q = 0; q < CW; ++q) {
                    rx[q] = lane_shift<true>(ex[r][q]);
                    ry[q] = lane_shift<true>(ey[r][q]);
                }
                const R idx = tb.ax.X(0, i), idy = tb.ax.Y(0, j);
                R kxr[4] = {0, 0, 0, 0}, kyr[4] = {0, 0, 0, 0};
                if (a.cpml && in_pml(i, nx, a.np)) {
                    kxr[0] = tb.ax.X(1, i); kxr[1] = tb.ax.X(6, i); kxr[2] = tb.ax.X(4, i); kxr[3] = tb.ax.X(5, i);
                }
                if (a.cpml && in_pml(j, ny, a.np)) {
                    kyr[0] = tb.ax.Y(1, j); kyr[1] = tb.ax.Y(6, j); kyr[2] = tb.ax.Y(4, j); kyr[3] = tb.ax.Y(5, j);
                }
                for (int q = 0; q < CW; ++q) {
                    const bool last = lane == 63;
                    HIn<R> v;
                    v.ex = ex[r][q]; v.ey = ey[r][q]; v.ez = ez[r][q];
                    v.ex_j = ejx[q]; v.ez_j = ejz[q];
                    v.ey_i = eny[q]; v.ez_i = enz[q];
                    v.ex_k = last ? (q + 1 < CW ? rx[q + 1] : ekx) : rx[q];
                    v.ey_k = last ? (q + 1 < CW ? ry[q + 1] : eky) : ry[q];
                    v.hx = hox[q]; v.hy = hoy[q]; v.hz = hoz[q];
                    v.ns[0] = v.ns[1] = v.ns[2] = (R)1;  // neighbours are values
                    for (int u = 0; u < 6; ++u) { v.p[u] = ph[q][u]; v.cH[u] = tb.hco[u * a.nmat]; }
                    v.upd = (i < nx - 1) & (j < ny - 1) & (kq[q] < nz - 1);
                    v.idx = idx; v.idy = idy; v.idz = idz[q];
                    for (int u = 0; u < 4; ++u) { v.kx[u] = kxr[u]; v.ky[u] = kyr[u]; v.kz[u] = kzh[q][u]; }
                    const H3<R> h = h_compute<T, R, true, IX>(a, tb, i, j, kk[q], v, !honly && kq[q] < nz);
                    hx[q] = h.x; hy[q] = h.y; hz[q] = h.z;
                }
            }
            if (!honly) {
                // ---- E^{n+1}(i) of the row ----
                R kpx = 0, kpy = 0;
                if (gm.kc > 1) {
                    if (lane == 63) { kslot[slot][w][0] = hx[CW - 1]; kslot[slot][w][1] = hy[CW - 1]; }
                    lds_barrier();
                    if (w > 0) { kpx = kslot[slot][w - 1][0]; kpy = kslot[slot][w - 1][1]; }
                    slot ^= 1;
                }
                const SA a = phase_args<T, R>(ka);
                const Tabs<R> tb = global_tabs(a);
                R px_[CW], py_[CW];
                for (int q = 0; q < CW; ++q) {
                    px_[q] = lane_shift<false>(hx[q]);
                    py_[q] = lane_shift<false>(hy[q]);
                }
                for (int q = 0; q < CW; ++q) {
                    const bool first = lane == 0;
                    const R hxk = first ? (q > 0 ? px_[q - 1] : kpx) : px_[q];
                    const R hyk = first ? (q > 0 ? py_[q - 1] : kpy) : py_[q];
                    EIn<R> v;
                    fz_e_inputs<T, R, IX>(a, tb, i, j, kq[q], kbase + 64 * q, mid[q], v);
                    v.ex0 = ex[r][q]; v.ey0 = ey[r][q]; v.ez0 = ez[r][q];
                    v.idz = idz[q];
                    if (kq[q] < nz)
                        e_compute<T, R, true, IX>(a, tb, i, j, kq[q], v, hx[q], hy[q], hz[q], hjx[q], hjz[q], hxk, hyk,
                                                  hpy[r][q], hpz[r][q]);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            // carry to the next row and plane
            for (int q = 0; q < CW; ++q) {
                hjx[q] = hx[q]; hjz[q] = hz[q];
                hpy[r][q] = hy[q]; hpz[r][q] = hz[q];
                ex[r][q] = enx[q]; ey[r][q] = eny[q]; ez[r][q] = enz[q];
            }
        }
        for (int q = 0; q < CW; ++q) { gx[q] = gnx[q]; gy[q] = gny[q]; gz[q] = gnz[q]; }
    }
}

}  // namespace fdtdk

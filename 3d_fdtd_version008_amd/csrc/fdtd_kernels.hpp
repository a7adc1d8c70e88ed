// fdtd_kernels.hpp - gfx950 kernels of the per-timestep Yee update.
//
// Layout in HBM: every field component is its own array in the reference's
// idx3 order ((i*NyT + j)*NzT + k, k fastest), so a 64-lane wavefront reads
// 64 consecutive k values of one (i, j) row: one fully coalesced
// 256 B (FP32) / 512 B (FP64) / 128 B (FP16) request per operand.
// CPML psi arrays exist only on the boundary slabs (x slabs [0,npml) and
// [NxT-npml,NxT) over all j,k; likewise y and z), gain state only inside the
// gain bounding box, and the per-cell material coefficients are a 1- or
// 2-byte material id per cell into an exact lookup table of the distinct
// (aE, bE, aH, bH) tuples of the baked MaterialGrids.
//
// Arithmetic restates, per cell, the reference's sequence of operations:
// fdtd_update_H / fdtd_update_E (src/fdtd_stepper.hpp:8-89), the separate
// CPML correction passes (src/boundary.hpp:150-300, same order of the six
// corrections per cell), fdtd_update_E_with_gain and the two TLS passes
// (src/two_level_system.hpp:299-342, 574-666).  Fusing the passes into one
// thread per cell only removes memory round trips; each value goes through
// the same operations in the same order.  The library is compiled with
// -ffp-contract=off, so every FP64 operation is individually rounded as on
// the host.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <type_traits>
#include "fdtd_fastdiv.hpp"

namespace fdtdk {

// Saturation at the largest finite FP16 value: a finite value that outgrows
// the array's scale between two scale checks is clipped (and the next check
// raises the scale) instead of turning into Inf.  A NaN or an Inf is stored
// as it is, so the stability monitor sees it as the reference's NaN/Inf scan
// would: a finite value goes through the median-of-three clamp, anything
// else (one class test) passes unchanged (fminf/fmaxf alone would turn a NaN
// into a bound).
__device__ __forceinline__ float fp16_sat(float x) {
    return __builtin_isfinite(x) ? __builtin_amdgcn_fmed3f(x, -65504.0f, 65504.0f) : x;
}

// Storage <-> compute conversion.  FP16 storage carries a power-of-two scale
// per component (value = stored * scale), so scaling is exact.
template <typename T, typename R>
__device__ __forceinline__ R ld(const T* p, long i, R scale) {
    if constexpr (sizeof(T) == 2) {
        return (R)__half2float(reinterpret_cast<const __half*>(p)[i]) * scale;
    } else {
        (void)scale;
        return (R)p[i];
    }
}
template <typename T, typename R>
__device__ __forceinline__ void st(T* p, long i, R v, R inv_scale) {
    if constexpr (sizeof(T) == 2) {
        reinterpret_cast<__half*>(p)[i] = __float2half_rn(fp16_sat((float)(v * inv_scale)));
    } else {
        (void)inv_scale;
        p[i] = (T)v;
    }
}
// Cell-indexed access with index type IX: long (64-bit element offsets) or
// unsigned, where the host guarantees every array the kernel touches is below
// 4 GiB: the byte offset is then formed in 32 bits, so an access is a uniform
// base plus one 32-bit offset register (neighbours at +1 become immediate
// offsets) instead of a 64-bit address per access.
template <typename IX>
constexpr bool is_off32() { return sizeof(IX) == 4; }
template <typename P, typename IX>
__device__ __forceinline__ P* at(P* p, IX i) {
    if constexpr (is_off32<IX>()) {
        using C = typename std::conditional<std::is_const<P>::value, const char, char>::type;
        return reinterpret_cast<P*>(reinterpret_cast<C*>(p) + (unsigned)i * (unsigned)sizeof(P));
    } else {
        return p + (long)i;
    }
}
template <typename T, typename R, typename IX>
__device__ __forceinline__ R ldx(const T* p, IX i, R scale) {
    return ld<T, R>(at(p, i), 0L, scale);
}
template <typename T, typename R, typename IX>
__device__ __forceinline__ void stx(T* p, IX i, R v, R inv_scale) {
    st<T, R>(at(p, i), 0L, v, inv_scale);
}
// ldx through a load that does not hit the CU's L1 (an agent-scope relaxed
// atomic load), for a value another work-group stored earlier in the same
// launch: a line this CU cached before that store would otherwise be read
// stale (in an ordinary dispatch an L1 invalidate at work-group scope does
// not reach those lines).  COH = false is ldx.
template <bool COH, typename T, typename R, typename IX>
__device__ __forceinline__ R ldxh(const T* p, IX i, R scale) {
    if constexpr (!COH) {
        return ldx<T, R>(p, i, scale);
    } else if constexpr (sizeof(T) == 2) {
        const unsigned short b = __hip_atomic_load(reinterpret_cast<const unsigned short*>(at(p, i)), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        return (R)__half2float(__ushort_as_half(b)) * scale;
    } else {
        (void)scale;
        return (R)__hip_atomic_load(at(p, i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The compiler cannot see through this (no instruction is emitted): a value
// that passes through it is not a constant, so operations on a variable that
// is either such a default or a loaded value are not folded into the branch
// that loads it (where they would wait for the load).
template <typename X>
__device__ __forceinline__ X opaque(X x) {
    asm("" : "+v"(x));
    return x;
}
// Register type of a raw storage value: FP16 bits are held zero-extended in
// 32 bits (the load extends them; packing two into one register would need
// the loaded value at once).
template <typename T>
struct Raw {
    using type = T;
};
template <>
struct Raw<uint16_t> {
    using type = unsigned;
};
// Storage value to compute type (the conversion half of ld).
template <typename T, typename R>
__device__ __forceinline__ R cv(typename Raw<T>::type x, R scale) {
    if constexpr (sizeof(T) == 2) {
        return (R)__half2float(__ushort_as_half((unsigned short)x)) * scale;
    } else {
        (void)scale;
        return (R)x;
    }
}

// The value a store followed by a load would return (rounding to storage).
template <typename T, typename R>
__device__ __forceinline__ double stored_value(R v, R scale, R inv_scale) {
    if constexpr (sizeof(T) == 2) {
        return (double)(__half2float(__float2half_rn(fp16_sat((float)(v * inv_scale)))) * scale);
    } else {
        (void)scale; (void)inv_scale;
        return (double)(T)v;
    }
}

struct SourceBox {
    int comp, i0, i1, j0, j1, k0, k1;
};

struct TLSCoef {
    double kapa, pa2, pa3, inv_dt, inv_tau, inv_hbar_omega, N0, inv_N0, dt;
    int clamp;
};

template <typename T, typename R>
struct StepArgs {
    int nx, ny, nz, np;
    long sI;
    int sJ;
    T *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;        // read buffers (E^n, H^{n-1/2})
    T *Exw, *Eyw, *Ezw, *Hxw, *Hyw, *Hzw;  // write buffers (the same arrays when updating in place)
    int pec;
    // Per-axis tables [7][n]: inverse spacing, then the CPML 1D profiles
    // kappa, bE, cE, bH, cH (kappa 1, b 1, c 0 outside the slabs), 1/kappa.
    const R *cpx, *cpy, *cpz;
    int cpml;
    // psi slabs: 0 Ex_y 1 Ex_z 2 Ey_z 3 Ey_x 4 Ez_x 5 Ez_y 6 Hx_y 7 Hx_z 8 Hy_z 9 Hy_x 10 Hz_x 11 Hz_y
    T* psi[12];
    T* psiw[12];  // psi_H write buffers (6..11); psi_E (0..5) is updated in place
    // FP16 power-of-two scales (device memory, value = stored * scale):
    // sc[0..5] fields, sc[6..17] psi, sc[18..35] their inverses.
    const float* sc;
    // materials: lut[12][nmat] = aEx bEx aEy bEy aEz bEz aHx bHx aHy bHy aHz bHz
    const void* matid;
    int mid16;
    const R* lut;
    int nmat;
    int lossless;
    int h_uniform;
    // sources
    const SourceBox* src;
    int nsrc;
    const double* srcv;  // values of this step
    // gain box
    int gain;
    int g0x, g0y, g0z, gnx, gny, gnz;
    const int8_t* greg;
    double *gPz, *gPp, *gNu, *gNg;
    const TLSCoef* tls;
    int ntls;
    // tiles of the two-kernel grid: k-tiles of 64 cells and j-tiles of
    // TILE_ROWS rows, as divisors of the tile index
    FastDiv tk, tj;
    // 1: each XCD walks its range of tiles from the end (the launch starts
    // where the previous one, walking forward, finished)
    int rev;
    // Tile order: bands of `bj` j-tiles, each band swept plane by plane
    // (order band, plane, j-tile in band, k-tile).  A tile reads the next
    // plane's E (or the previous plane's H) that the same band of the next
    // (previous) plane loads again one band-plane later, while it is still in
    // the XCD's L2; with whole planes in the sweep (bj = tj) a plane's six
    // arrays outgrow the L2 once a plane is larger than about 4 MB.
    // tbp: tiles per band-plane, tband: tiles per band (full bands), tbpl:
    // tiles per plane of the last, partial band; full_tiles: tiles in full
    // bands.
    int bj;
    unsigned full_tiles;
    FastDiv tbp, tband, tbpl;
};

__device__ __forceinline__ bool in_pml(int i, int n, int np) {
    int d = i < n - 1 - i ? i : n - 1 - i;
    return d < np;
}
// psi slab offsets
template <typename IX = long>
__device__ __forceinline__ IX psi_x(int i, int j, int k, int nx, int ny, int nz, int np) {
    int ii = i < np ? i : i - (nx - 2 * np);
    return ((IX)ii * ny + j) * nz + k;
}
template <typename IX = long>
__device__ __forceinline__ IX psi_y(int i, int j, int k, int nx, int ny, int nz, int np) {
    (void)nx;
    int jj = j < np ? j : j - (ny - 2 * np);
    return ((IX)i * (2 * np) + jj) * nz + k;
}
template <typename IX = long>
__device__ __forceinline__ IX psi_z(int i, int j, int k, int nx, int ny, int nz, int np) {
    (void)nx;
    int kk = k < np ? k : k - (nz - 2 * np);
    return ((IX)i * ny + j) * (2 * np) + kk;
}

template <typename R, typename IX = long>
__device__ __forceinline__ int mat_of(const void* m, int mid16, IX id) {
    return mid16 ? (int)*at(reinterpret_cast<const uint16_t*>(m), id) : (int)*at(reinterpret_cast<const uint8_t*>(m), id);
}

// Where a kernel reads the per-axis tables from: global memory, or the
// work-group's LDS copy of the rows it touches (row r of axis entry n at
// p[r * stride + n - off]).
template <typename R>
struct Axes {
    const R* x;
    int sx, ox;
    const R* y;
    int sy, oy;
    const R* z;
    int sz, oz;
    __device__ __forceinline__ R X(int r, int n) const { return x[r * sx + n - ox]; }
    __device__ __forceinline__ R Y(int r, int n) const { return y[r * sy + n - oy]; }
    __device__ __forceinline__ R Z(int r, int n) const { return z[r * sz + n - oz]; }
};
template <typename T, typename R>
__device__ __forceinline__ Axes<R> global_axes(const StepArgs<T, R>& a) {
    return Axes<R>{a.cpx, a.nx, 0, a.cpy, a.ny, 0, a.cpz, a.nz, 0};
}

// Small read-only tables the compute phase reads: axis tables, the
// cell-invariant H coefficients, source footprints and values, TLS
// coefficients.
template <typename R>
struct Tabs {
    Axes<R> ax;
    const R* hco;           // aHx bHx aHy bHy aHz bHz when cell-invariant
    const SourceBox* src;
    const double* srcv;
    const TLSCoef* tls;
};
template <typename T, typename R>
__device__ __forceinline__ Tabs<R> global_tabs(const StepArgs<T, R>& a) {
    return Tabs<R>{global_axes(a), a.lut + 6 * a.nmat, a.src, a.srcv, a.tls};
}
// psi array q as read (0..11; psi_E is updated in place, so also its write array)
template <typename T, typename R>
__device__ __forceinline__ T* psi_rd(const StepArgs<T, R>& a, const Tabs<R>& tb, int q) {
    (void)tb;
    return a.psi[q];
}
// psi_H write array (q = 6..11)
template <typename T, typename R>
__device__ __forceinline__ T* psi_wr(const StepArgs<T, R>& a, const Tabs<R>& tb, int q) {
    (void)tb;
    return a.psiw[q];
}
// gain state array: 0 Nu, 1 Ng, 2 Pz, 3 Pz_prev
template <typename T, typename R>
__device__ __forceinline__ double* gain_arr(const StepArgs<T, R>& a, const Tabs<R>& tb, int q) {
    (void)tb;
    return q == 0 ? a.gNu : q == 1 ? a.gNg : q == 2 ? a.gPz : a.gPp;
}

constexpr int STAGE_MAX_TLS = 16;  // TLS regions of one simulation

// Arguments kept in a device-side table (batches) store every pointer as its
// byte offset from an anchor address; the kernel adds the anchor (a kernel
// argument, so known to address global memory) back, and the compiler then
// emits global (not flat) accesses through them.  The anchor is not the
// table itself, so stores through these pointers cannot alias the table and
// the table is read through the scalar cache.  The host applies the same map
// with the opposite sign.
template <typename P>
__host__ __device__ __forceinline__ void shift_ptr(P*& p, const char* base, long sign) {
    if (sign > 0) p = (P*)((char*)base + (long)p);   // offset -> pointer (a GEP from base)
    else p = (P*)((long)p - (long)base);             // pointer -> offset
}
template <typename T, typename R>
__host__ __device__ __forceinline__ void shift_args(StepArgs<T, R>& a, const char* base, long sign) {
    shift_ptr(a.Ex, base, sign); shift_ptr(a.Ey, base, sign); shift_ptr(a.Ez, base, sign);
    shift_ptr(a.Hx, base, sign); shift_ptr(a.Hy, base, sign); shift_ptr(a.Hz, base, sign);
    shift_ptr(a.Exw, base, sign); shift_ptr(a.Eyw, base, sign); shift_ptr(a.Ezw, base, sign);
    shift_ptr(a.Hxw, base, sign); shift_ptr(a.Hyw, base, sign); shift_ptr(a.Hzw, base, sign);
    shift_ptr(a.cpx, base, sign); shift_ptr(a.cpy, base, sign); shift_ptr(a.cpz, base, sign);
    shift_ptr(a.sc, base, sign);
    shift_ptr(a.matid, base, sign);
    shift_ptr(a.lut, base, sign);
    shift_ptr(a.src, base, sign);
    shift_ptr(a.srcv, base, sign);
    shift_ptr(a.greg, base, sign);
    shift_ptr(a.gPz, base, sign); shift_ptr(a.gPp, base, sign); shift_ptr(a.gNu, base, sign); shift_ptr(a.gNg, base, sign);
    shift_ptr(a.tls, base, sign);
    for (int q = 0; q < 12; ++q) {
        shift_ptr(a.psi[q], base, sign);
        shift_ptr(a.psiw[q], base, sign);
    }
}

// One CPML correction (src/boundary.hpp:182-223, 257-298) on an already
// loaded psi:  psi = b*psi + c*d;  corr = d/kappa + psi;  returns (corr - d)
// and the new psi in `pnew`.  FP64 divides as the reference does; the FP32
// and FP16 forms multiply by 1/kappa (rounded from the FP64 value).
template <typename R>
__device__ __forceinline__ R cpml_term(R psi_old, R d, R kap, R ikap, R b, R c, R& pnew) {
    R corr;
    if constexpr (sizeof(R) == 8) {
        (void)ikap;
        corr = d / kap;
    } else {
        (void)kap;
        corr = d * ikap;
    }
    const R p = b * psi_old + c * d;
    pnew = p;
    corr += p;
    return corr - d;
}

template <typename R>
struct H3 {
    R x, y, z;
};

// b * s - a for a neighbour b held unscaled with scale s (a power of two, or
// 1): b * s is exact, so the single rounding of the FMA gives the bits of
// the difference of the scaled values (signed zeros included).  With s = 1
// it is the subtraction b - a.
template <typename R>
__device__ __forceinline__ R nb_minus(R b, R s, R a) {
    if constexpr (sizeof(R) == 8) return __builtin_fma(b, s, -a);
    else return __builtin_fmaf(b, s, -a);
}
// a - b * s, the same way (with s = 1 the subtraction a - b).
template <typename R>
__device__ __forceinline__ R own_minus(R a, R b, R s) {
    if constexpr (sizeof(R) == 8) return __builtin_fma(b, -s, a);
    else return __builtin_fmaf(b, -s, a);
}

// Inputs of one cell's H update, gathered by h_load so that a caller can
// issue them ahead of use (the fused kernel loads plane i+1 while it
// computes plane i).
template <typename R>
struct HIn {
    R ex, ex_j, ex_k, ey, ey_i, ey_k, ez, ez_i, ez_j, hx, hy, hz;
    // scale of the neighbour values ex_j/ex_k, ey_i/ey_k, ez_i/ez_j: the FP16
    // gathers hold them unscaled (the scale is applied inside the difference,
    // one mixed-precision FMA instead of a conversion, a multiply and a
    // subtraction); 1 where they are values already
    R ns[3];
    R p[6];  // psi_H: Hx_y, Hx_z, Hy_z, Hy_x, Hz_x, Hz_y (where in a slab)
    R cH[6]; // aHx bHx aHy bHy aHz bHz
    bool upd;
    // axis tables at the cell: inverse spacings, and the CPML H rows
    // (kappa, 1/kappa, bH, cH) at i, j and k where the cell is in that slab
    R idx, idy, idz;
    R kx[4], ky[4], kz[4];
};

// Inputs of the H update of cell (i, j, k), one the reference updates
// (i < NxT-1, j < NyT-1, k < NzT-1), for a wavefront that covers KW
// consecutive k of 64/KW consecutive rows of plane i (i, kw and jw, the k and
// j of its first lane, uniform across it).  As in e_gather, every input is requested before the
// first one is used: the material id first (when H coefficients vary per
// cell), then E^n and H^{n-1/2}, psi_H and the z-axis table, with clamped
// indices under wave-uniform conditions and opaque defaults; X/Y axis values
// and cell-invariant coefficients come through the scalar cache.
template <typename T, typename R, int KW = 64, bool CP = true, typename IX = long>
__device__ __forceinline__ HIn<R> h_gather(const StepArgs<T, R>& a, const Tabs<R>& tb, int i, int j, int k, int kw,
                                           int jw) {
    constexpr int RW = 64 / KW;  // rows per wavefront
    HIn<R> v;
    const IX id = ((IX)i * a.ny + j) * a.nz + k;
    const int np = a.np, nx = a.nx, ny = a.ny, nz = a.nz;
    int m = 0;
    if (!a.h_uniform) m = mat_of<R, IX>(a.matid, a.mid16, id);
    v.upd = true;
    const IX sI = (IX)a.sI, sJ = (IX)a.sJ;
    v.hx = ldx<T, R>(a.Hx, id, a.sc[3]);
    v.hy = ldx<T, R>(a.Hy, id, a.sc[4]);
    v.hz = ldx<T, R>(a.Hz, id, a.sc[5]);
    // neighbours unscaled (FP16; scale 1 otherwise), applied in h_compute
    const R n1 = 1;
    v.ex = ldx<T, R>(a.Ex, id, a.sc[0]); v.ex_j = ldx<T, R>(a.Ex, id + sJ, n1); v.ex_k = ldx<T, R>(a.Ex, id + 1, n1);
    v.ey = ldx<T, R>(a.Ey, id, a.sc[1]); v.ey_i = ldx<T, R>(a.Ey, id + sI, n1); v.ey_k = ldx<T, R>(a.Ey, id + 1, n1);
    v.ez = ldx<T, R>(a.Ez, id, a.sc[2]); v.ez_i = ldx<T, R>(a.Ez, id + sI, n1); v.ez_j = ldx<T, R>(a.Ez, id + sJ, n1);
    for (int q = 0; q < 3; ++q) v.ns[q] = sizeof(T) == 2 ? (R)a.sc[q] : (R)1;
    using RawT = typename Raw<T>::type;
    RawT praw[6];
    for (int q = 0; q < 6; ++q) praw[q] = opaque((RawT)0);
    const bool px = CP && a.cpml && in_pml(i, nx, np);
    const bool pyw = CP && a.cpml && (in_pml(jw, ny, np) | in_pml(jw + RW - 1, ny, np));
    const bool pzw = CP && a.cpml && ((kw < np) | (kw + KW - 1 >= nz - np));
    if (pyw) {
        const IX q = psi_y<IX>(i, in_pml(j, ny, np) ? j : 0, k, nx, ny, nz, np);
        praw[0] = *at(a.psi[6], q);
        praw[5] = *at(a.psi[11], q);
    }
    if (pzw) {
        const IX q = psi_z<IX>(i, j, in_pml(k, nz, np) ? k : 0, nx, ny, nz, np);
        praw[1] = *at(a.psi[7], q);
        praw[2] = *at(a.psi[8], q);
    }
    if (px) {
        const IX q = psi_x<IX>(i, j, k, nx, ny, nz, np);
        praw[3] = *at(a.psi[9], q);
        praw[4] = *at(a.psi[10], q);
    }
    const Axes<R>& ax = tb.ax;
    v.idx = ax.X(0, i);
    v.idy = ax.Y(0, j);
    v.idz = ax.Z(0, k);
    for (int q = 0; q < 4; ++q) v.kx[q] = 0;
    if (px) { v.kx[0] = ax.X(1, i); v.kx[1] = ax.X(6, i); v.kx[2] = ax.X(4, i); v.kx[3] = ax.X(5, i); }
    for (int q = 0; q < 4; ++q) v.ky[q] = opaque((R)0);
    if (pyw) { v.ky[0] = ax.Y(1, j); v.ky[1] = ax.Y(6, j); v.ky[2] = ax.Y(4, j); v.ky[3] = ax.Y(5, j); }
    for (int q = 0; q < 4; ++q) v.kz[q] = opaque((R)0);
    if (pzw) { v.kz[0] = ax.Z(1, k); v.kz[1] = ax.Z(6, k); v.kz[2] = ax.Z(4, k); v.kz[3] = ax.Z(5, k); }
    if (a.h_uniform) {
        for (int q = 0; q < 6; ++q) v.cH[q] = a.lut[(6 + q) * a.nmat];
    } else {
        for (int q = 0; q < 6; ++q) v.cH[q] = a.lut[(6 + q) * a.nmat + m];
    }
    for (int q = 0; q < 6; ++q) v.p[q] = cv<T, R>(praw[q], a.sc[12 + q]);
    return v;
}

// H^{n+1/2} of one cell: fdtd_update_H (src/fdtd_stepper.hpp:31-46) followed
// by the six CPMLBoundary::apply_after_H corrections (src/boundary.hpp:172-224)
// in the reference's order.  Stores H and psi_H to the write buffers when
// `own` (halo recomputation by a neighbouring tile must not store).  Cells the
// reference never updates (i = NxT-1, j = NyT-1 or k = NzT-1) return their
// stored value.  Returns the values a later read of the stored H gives.
template <typename T, typename R, bool CP = true, typename IX = long>
__device__ __forceinline__ H3<R> h_compute(const StepArgs<T, R>& a, const Tabs<R>& tb, int i, int j, int k,
                                           const HIn<R>& v, bool own, bool store_fields = true) {
    if (!v.upd) return H3<R>{v.hx, v.hy, v.hz};
    const IX id = ((IX)i * a.ny + j) * a.nz + k;
    const R idx = v.idx, idy = v.idy, idz = v.idz;
    // (neighbour - own) * inverse spacing, the neighbour's scale inside the FMA
    const R dEz_y = nb_minus(v.ez_j, v.ns[2], v.ez) * idy, dEy_z = nb_minus(v.ey_k, v.ns[1], v.ey) * idz;
    const R dEx_z = nb_minus(v.ex_k, v.ns[0], v.ex) * idz, dEz_x = nb_minus(v.ez_i, v.ns[2], v.ez) * idx;
    const R dEy_x = nb_minus(v.ey_i, v.ns[1], v.ey) * idx, dEx_y = nb_minus(v.ex_j, v.ns[0], v.ex) * idy;
    const R cx = dEz_y - dEy_z, cy = dEx_z - dEz_x, cz = dEy_x - dEx_y;
    const R aHx = v.cH[0], bHx = v.cH[1], aHy = v.cH[2], bHy = v.cH[3], aHz = v.cH[4], bHz = v.cH[5];
    R hx = aHx * v.hx - bHx * cx;
    R hy = aHy * v.hy - bHy * cy;
    R hz = aHz * v.hz - bHz * cz;
    if (CP && a.cpml) {
        const int np = a.np, nx = a.nx, ny = a.ny, nz = a.nz;
        const bool px = in_pml(i, nx, np), py = in_pml(j, ny, np), pz = in_pml(k, nz, np);
        if (px | py | pz) {
            R pn[6];
            if (py) hx -= bHx * cpml_term<R>(v.p[0], dEz_y, v.ky[0], v.ky[1], v.ky[2], v.ky[3], pn[0]);
            if (pz) hx -= bHx * (-cpml_term<R>(v.p[1], dEy_z, v.kz[0], v.kz[1], v.kz[2], v.kz[3], pn[1]));
            if (pz) hy -= bHy * cpml_term<R>(v.p[2], dEx_z, v.kz[0], v.kz[1], v.kz[2], v.kz[3], pn[2]);
            if (px) hy -= bHy * (-cpml_term<R>(v.p[3], dEz_x, v.kx[0], v.kx[1], v.kx[2], v.kx[3], pn[3]));
            if (px) hz -= bHz * cpml_term<R>(v.p[4], dEy_x, v.kx[0], v.kx[1], v.kx[2], v.kx[3], pn[4]);
            if (py) hz -= bHz * (-cpml_term<R>(v.p[5], dEx_y, v.ky[0], v.ky[1], v.ky[2], v.ky[3], pn[5]));
            if (own) {
                if (py) {
                    const IX q = psi_y<IX>(i, j, k, nx, ny, nz, np);
                    stx<T, R>(psi_wr(a, tb, 6), q, pn[0], a.sc[30]);
                    stx<T, R>(psi_wr(a, tb, 11), q, pn[5], a.sc[35]);
                }
                if (pz) {
                    const IX q = psi_z<IX>(i, j, k, nx, ny, nz, np);
                    stx<T, R>(psi_wr(a, tb, 7), q, pn[1], a.sc[31]);
                    stx<T, R>(psi_wr(a, tb, 8), q, pn[2], a.sc[32]);
                }
                if (px) {
                    const IX q = psi_x<IX>(i, j, k, nx, ny, nz, np);
                    stx<T, R>(psi_wr(a, tb, 9), q, pn[3], a.sc[33]);
                    stx<T, R>(psi_wr(a, tb, 10), q, pn[4], a.sc[34]);
                }
            }
        }
    }
    if (own && store_fields) {
        stx<T, R>(a.Hxw, id, hx, a.sc[21]);
        stx<T, R>(a.Hyw, id, hy, a.sc[22]);
        stx<T, R>(a.Hzw, id, hz, a.sc[23]);
    }
    return H3<R>{(R)stored_value<T, R>(hx, a.sc[3], a.sc[21]), (R)stored_value<T, R>(hy, a.sc[4], a.sc[22]),
                 (R)stored_value<T, R>(hz, a.sc[5], a.sc[23])};
}

// The FP16 scales copied into registers at kernel start, before any store:
// afterwards a.sc points at the copy (fully promoted to registers, since every
// index is a constant), so no scale is read from memory after a store.
#define FDTD_SCALES_IN_REGISTERS(a)                       \
    float scl_[54];                                       \
    if constexpr (sizeof(T) == 2) {                       \
        for (int q_ = 0; q_ < 54; ++q_) scl_[q_] = a.sc[q_]; \
        a.sc = scl_;                                      \
    }

// Tile of a two-kernel launch.  The grid is 1-D over (k-block, j-block,
// plane) tiles of 64 x TILE_ROWS x 1 cells, one 256-thread work-group each.
// Work-group ids are dealt round-robin to the 8 XCDs, each with its own L2;
// the remap hands every XCD one contiguous range of tiles (whole runs of
// planes), so the neighbouring rows and the previous plane a tile reads were
// loaded recently by the same XCD.  The tile index is split with precomputed
// divisors (a.tk, a.tj).  Inside the tile each wavefront covers KW
// consecutive k of 64/KW consecutive rows; (kw, jw) is its first cell.
constexpr int TILE_ROWS = 4;
struct Tile {
    int i, j, k, kw, jw;
};
template <int KW, typename T, typename R>
__device__ __forceinline__ Tile tile_of(const StepArgs<T, R>& a) {
    static_assert(KW == 16 || KW == 32 || KW == 64, "wavefront k-width");
    constexpr int G = 64 / KW;  // k-groups per tile = rows per wavefront
    // linear work-group id; grids whose tile count times 256 threads
    // exceeds 32 bits are launched two-dimensional (gridDim.y planes' worth)
    const unsigned nb = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned per = (nb + 7) / 8;
    unsigned bb = (b % 8) * per + (a.rev ? per - 1 - b / 8 : b / 8);
    if (nb % 8 != 0) bb = a.rev ? nb - 1 - b : b;  // the remap is a permutation only for multiples of 8
    unsigned band, i, r2;
    if (bb < a.full_tiles) {
        band = fdiv(bb, a.tband);
        const unsigned rem = bb - band * a.tband.d;
        i = fdiv(rem, a.tbp);
        r2 = rem - i * a.tbp.d;
    } else {
        band = a.full_tiles / a.tband.d;
        const unsigned rem = bb - a.full_tiles;
        i = fdiv(rem, a.tbpl);
        r2 = rem - i * a.tbpl.d;
    }
    const unsigned jr = fdiv(r2, a.tk);
    const unsigned kt = r2 - jr * a.tk.d;
    const unsigned jt = band * (unsigned)a.bj + jr;
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wavefront in the work-group
    const int l = (int)(threadIdx.x & 63);
    Tile t;
    t.kw = (int)kt * 64 + (w % G) * KW;
    t.jw = (int)jt * TILE_ROWS + (w / G) * G;
    t.k = t.kw + l % KW;
    t.j = t.jw + l / KW;
    t.i = (int)i;
    return t;
}

// H update kernel of the two-kernel step: one 256-thread work-group per
// tile, each thread one cell (tile_of).
template <typename T, typename R, int KW, typename IX = long>
__global__ void __launch_bounds__(256) k_update_H(StepArgs<T, R> a) {
    FDTD_SCALES_IN_REGISTERS(a)
    const Tabs<R> tb = global_tabs(a);
    const Tile t = tile_of<KW>(a);
    const int i = t.i, j = t.j, k = t.k;
    if (k >= a.nz - 1 || j >= a.ny - 1 || i >= a.nx - 1) return;
    (void)h_compute<T, R, true, IX>(a, tb, i, j, k, h_gather<T, R, KW, true, IX>(a, tb, i, j, k, t.kw, t.jw), true);
}

// The H update of the step after a monitor step, which also takes the
// monitor's field maxima (StabilityMonitor::check_stability's max |E|,
// max |H| and NaN/Inf scan, src/stability_monitor.hpp:51-121) of E^{n+1}
// and H^{n+1/2} of the monitor step: exactly the E and H values this launch
// reads at each cell before it writes H, so the separate pass over the six
// fields is saved.  Every cell in the grid contributes its own E and H
// (cells the reference does not H-update load them directly).  Each
// wavefront reduces its lanes and takes atomicMax on one of MON_SLOTS slots
// (the maximum of bit patterns is order-free, so the result is exact);
// k_monitor_final merges the slots.
constexpr int MON_SLOTS = 1024;
constexpr int MON_SLOT_WORDS = 8;  // one 64-byte line per slot
template <typename T, typename R>
__device__ __forceinline__ unsigned long long absbits_v(R v) {
    return (unsigned long long)__double_as_longlong(fabs((double)v));
}
template <typename T, typename R, int KW, typename IX = long>
__global__ void __launch_bounds__(256) k_update_H_mon(StepArgs<T, R> a, unsigned long long* fmax) {
    FDTD_SCALES_IN_REGISTERS(a)
    const Tabs<R> tb = global_tabs(a);
    const Tile t = tile_of<KW>(a);
    const int i = t.i, j = t.j, k = t.k;
    unsigned long long mE = 0, mH = 0;
    if (k < a.nz - 1 && j < a.ny - 1 && i < a.nx - 1) {
        const HIn<R> v = h_gather<T, R, KW, true, IX>(a, tb, i, j, k, t.kw, t.jw);
        mE = max(max(absbits_v<T, R>(v.ex), absbits_v<T, R>(v.ey)), absbits_v<T, R>(v.ez));
        mH = max(max(absbits_v<T, R>(v.hx), absbits_v<T, R>(v.hy)), absbits_v<T, R>(v.hz));
        (void)h_compute<T, R, true, IX>(a, tb, i, j, k, v, true);
    } else if (k < a.nz && j < a.ny && i < a.nx) {
        const IX id = ((IX)i * a.ny + j) * a.nz + k;
        mE = max(max(absbits_v<T, R>(ldx<T, R>(a.Ex, id, a.sc[0])), absbits_v<T, R>(ldx<T, R>(a.Ey, id, a.sc[1]))),
                 absbits_v<T, R>(ldx<T, R>(a.Ez, id, a.sc[2])));
        mH = max(max(absbits_v<T, R>(ldx<T, R>(a.Hx, id, a.sc[3])), absbits_v<T, R>(ldx<T, R>(a.Hy, id, a.sc[4]))),
                 absbits_v<T, R>(ldx<T, R>(a.Hz, id, a.sc[5])));
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long e = __shfl_xor(mE, off), h = __shfl_xor(mH, off);
        mE = e > mE ? e : mE;
        mH = h > mH ? h : mH;
    }
    if ((threadIdx.x & 63) == 0) {
        const unsigned b = (blockIdx.y * gridDim.x + blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
        unsigned long long* q = fmax + (size_t)(b % MON_SLOTS) * MON_SLOT_WORDS;
        if (mE) atomicMax(q, mE);
        if (mH) atomicMax(q + 1, mH);
    }
}

// Two-level polarization step (update_polarization_multi_region,
// src/two_level_system.hpp:595-605).  Ndip = Ng0 = N0 on region cells.  The
// FP64 form divides by N0 as the reference does; the FP32/FP16 forms (whose
// populations are still FP64) multiply by 1/N0.
template <bool EXACT>
__device__ __forceinline__ double tls_polarize(const TLSCoef& c, double Nu, double Ng, double ez, double& Pz, double& Pp) {
    const double Ng_frac = EXACT ? (Ng - Nu) / c.N0 : (Ng - Nu) * c.inv_N0;
    const double driving = c.N0 * Ng_frac * ez;
    const double Pn = c.kapa * driving + c.pa2 * Pz + c.pa3 * Pp;
    const double Po = Pz;
    Pp = Po;
    Pz = Pn;
    return (Pn - Po) * c.inv_dt;
}

// Population step (update_populations_multi_region, src/two_level_system.hpp:633-662).
__device__ __forceinline__ void tls_populate(const TLSCoef& c, double ez_new, double ez_old, double Pz, double Pp,
                                             double& Nu, double& Ng) {
    const double Nu_curr = Nu;
    const double Ntotal = c.N0;
    const double E_avg = 0.5 * (ez_new + ez_old);
    const double delta_P = Pz - Pp;
    const double stim = E_avg * delta_P * c.inv_dt * c.inv_hbar_omega;
    const double dNu = -c.inv_tau * Nu_curr + stim;
    const double dNg = +c.inv_tau * Nu_curr - stim;
    Nu += c.dt * dNu;
    Ng += c.dt * dNg;
    Nu = (0.0 < Nu) ? Nu : 0.0;
    Ng = (0.0 < Ng) ? Ng : 0.0;
    if (c.clamp) {
        const double total = Nu + Ng;
        if (total > 1e-30) {
            const double scale = Ntotal / total;
            Nu *= scale;
            Ng *= scale;
        }
        Nu = (Ntotal < Nu) ? Ntotal : Nu;
        Ng = Ntotal - Nu;
    }
}

// Inputs of one cell's E update that do not depend on H^{n+1/2}.
template <typename R>
struct EIn {
    R ex0, ey0, ez0;
    R p[6];  // psi_E: Ex_y, Ex_z, Ey_z, Ey_x, Ez_x, Ez_y (where in a slab)
    R cE[6]; // aEx bEx aEy bEy aEz bEz of the cell's material
    int reg;
    long g;
    double Nu, Ng, Pz, Pp;
    // axis tables at the cell: inverse spacings, and the CPML E rows
    // (kappa, 1/kappa, bE, cE) at i, j and k where the cell is in that slab
    R idx, idy, idz;
    R kx[4], ky[4], kz[4];
    R jx, jy, jz;  // source current J at the cell
    TLSCoef tc;    // TLS coefficients of region `reg` (when reg >= 0)
    // scale of the H neighbour values e_compute takes (hx_j, hx_k: Hx;
    // hy_k, hy_i: Hy; hz_j, hz_i: Hz), as HIn::ns
    R hns[3];
};

// Source current at cell (i, j, k): the sum over the footprints that hold it.
// Footprints and values are read for every source (uniform addresses, so
// through the scalar cache) and the cell's share is selected without a
// branch, so no read waits behind the cell's vector loads.
template <typename R>
__device__ __forceinline__ void e_sources(const Tabs<R>& tb, int nsrc, int i, int j, int k, R& jx, R& jy, R& jz) {
    jx = jy = jz = 0;
    for (int s = 0; s < nsrc; ++s) {
        const SourceBox b = tb.src[s];
        const R val = (R)tb.srcv[s];
        const bool hit = (i >= b.i0) & (i <= b.i1) & (j >= b.j0) & (j <= b.j1) & (k >= b.k0) & (k <= b.k1);
        const R add = hit ? val : (R)0;
        jx += b.comp == 0 ? add : (R)0;
        jy += b.comp == 1 ? add : (R)0;
        jz += b.comp == 2 ? add : (R)0;
    }
}
// Inputs of the E update of cell (i, j, k) for a wavefront that covers KW
// consecutive k of 64/KW consecutive rows of plane i: i, kw and jw (the k
// and j of the wavefront's first lane) are uniform across it.  Every input is requested
// before the first one is used, so a wavefront makes one trip to memory plus
// one to the caches: the material and region ids first (the gathers after
// them wait for these alone; loads complete in issue order), then H^{n+1/2}
// and E^n, psi_E, the gain state and the z-axis table.  Wave-uniform
// conditions select which of these a wavefront loads; within a wavefront
// every lane loads, with indices clamped into bounds where a lane's cell is
// outside the slab or the gain box.  Values loaded under a condition start
// from opaque defaults and are converted after the last load is issued, so
// no arithmetic on them is placed (and waits) inside the branch.  X/Y axis
// values (Y too when a wavefront holds one row), source footprints and the
// coefficients of a single TLS region are uniform and come through the scalar
// cache.  Narrow wavefronts (KW < 64) confine the CPML z slabs and a z-normal
// gain slab to fewer wavefronts: a wavefront pays for the slab code when any
// of its lanes is in the slab.
// Returns H^{n+1/2} at the cell and at its -x, -y, -z neighbours in h[9]:
// hx hy hz hx_j hz_j hx_k hy_k hy_i hz_i.
template <typename T, typename R, int KW = 64, bool CP = true, bool GN = true, typename IX = long, bool COH = false>
__device__ __forceinline__ EIn<R> e_gather(const StepArgs<T, R>& a, const Tabs<R>& tb, int i, int j, int k, int kw,
                                           int jw, R* h) {
    constexpr int RW = 64 / KW;  // rows per wavefront
    EIn<R> v;
    const IX id = ((IX)i * a.ny + j) * a.nz + k;
    const int np = a.np, nx = a.nx, ny = a.ny, nz = a.nz;
    // ids
    const int m = mat_of<R, IX>(a.matid, a.mid16, id);
    const int gi = i - a.g0x, gj = j - a.g0y, gk = k - a.g0z;
    const bool growu = GN && a.gain && ((unsigned)gi < (unsigned)a.gnx) & (jw < a.g0y + a.gny) & (jw + RW - 1 >= a.g0y) &
                       (kw < a.g0z + a.gnz) & (kw + KW - 1 >= a.g0z);
    const bool gin = growu && ((unsigned)gj < (unsigned)a.gny) & ((unsigned)gk < (unsigned)a.gnz);
    const IX g = ((IX)gi * a.gny + (gin ? gj : 0)) * a.gnz + (gin ? gk : 0);
    int r8 = opaque(-1);
    if (growu) r8 = *at(a.greg, g);
    // fields.  On the rows i == 0 or j == 0, which the reference does not
    // update, the -x/-y offsets are taken as 0 and on the first row id - 1 as
    // id, so every address is in bounds; those values are never used (nor
    // is the k == 0 lane's H(k-1)).
    const IX sI = i > 0 ? (IX)a.sI : 0, sJ = j > 0 ? (IX)a.sJ : 0;
    const IX km = id > 0 ? 1 : 0;
    v.ez0 = ldx<T, R>(a.Ez, id, a.sc[2]);
    // neighbours unscaled (FP16; scale 1 otherwise), applied in e_compute
    const R n1 = 1;
    // (COH: H^{n+1/2} stored by other work-groups of the same launch)
    h[0] = ldxh<COH, T, R>(a.Hx, id, a.sc[3]); h[3] = ldxh<COH, T, R>(a.Hx, id - sJ, n1); h[5] = ldxh<COH, T, R>(a.Hx, id - km, n1);
    h[1] = ldxh<COH, T, R>(a.Hy, id, a.sc[4]); h[7] = ldxh<COH, T, R>(a.Hy, id - sI, n1); h[6] = ldxh<COH, T, R>(a.Hy, id - km, n1);
    h[2] = ldxh<COH, T, R>(a.Hz, id, a.sc[5]); h[8] = ldxh<COH, T, R>(a.Hz, id - sI, n1); h[4] = ldxh<COH, T, R>(a.Hz, id - sJ, n1);
    for (int q = 0; q < 3; ++q) v.hns[q] = sizeof(T) == 2 ? (R)a.sc[3 + q] : (R)1;
    v.ex0 = ldx<T, R>(a.Ex, id, a.sc[0]);
    v.ey0 = ldx<T, R>(a.Ey, id, a.sc[1]);
    // psi_E on the slabs
    using RawT = typename Raw<T>::type;
    RawT praw[6];
    for (int q = 0; q < 6; ++q) praw[q] = opaque((RawT)0);
    const bool px = CP && a.cpml && in_pml(i, nx, np);
    const bool pyw = CP && a.cpml && (in_pml(jw, ny, np) | in_pml(jw + RW - 1, ny, np));
    const bool pzw = CP && a.cpml && ((kw < np) | (kw + KW - 1 >= nz - np));
    if (pyw) {
        const IX q = psi_y<IX>(i, in_pml(j, ny, np) ? j : 0, k, nx, ny, nz, np);
        praw[0] = *at(a.psi[0], q);
        praw[5] = *at(a.psi[5], q);
    }
    if (pzw) {
        const IX q = psi_z<IX>(i, j, in_pml(k, nz, np) ? k : 0, nx, ny, nz, np);
        praw[1] = *at(a.psi[1], q);
        praw[2] = *at(a.psi[2], q);
    }
    if (px) {
        const IX q = psi_x<IX>(i, j, k, nx, ny, nz, np);
        praw[3] = *at(a.psi[3], q);
        praw[4] = *at(a.psi[4], q);
    }
    // gain state
    double gs[4];
    for (int q = 0; q < 4; ++q) gs[q] = opaque(0.0);
    if (growu) { gs[0] = *at(a.gNu, g); gs[1] = *at(a.gNg, g); gs[2] = *at(a.gPz, g); gs[3] = *at(a.gPp, g); }
    // axis tables
    const Axes<R>& ax = tb.ax;
    v.idx = ax.X(0, i);
    v.idy = ax.Y(0, j);
    v.idz = ax.Z(0, k);
    for (int q = 0; q < 4; ++q) v.kx[q] = 0;
    if (px) { v.kx[0] = ax.X(1, i); v.kx[1] = ax.X(6, i); v.kx[2] = ax.X(2, i); v.kx[3] = ax.X(3, i); }
    for (int q = 0; q < 4; ++q) v.ky[q] = opaque((R)0);
    if (pyw) { v.ky[0] = ax.Y(1, j); v.ky[1] = ax.Y(6, j); v.ky[2] = ax.Y(2, j); v.ky[3] = ax.Y(3, j); }
    for (int q = 0; q < 4; ++q) v.kz[q] = opaque((R)0);
    if (pzw) { v.kz[0] = ax.Z(1, k); v.kz[1] = ax.Z(6, k); v.kz[2] = ax.Z(2, k); v.kz[3] = ax.Z(3, k); }
    e_sources(tb, a.nsrc, i, j, k, v.jx, v.jy, v.jz);
    if (GN && a.gain && a.ntls == 1) v.tc = tb.tls[0];
    // the gathers that wait for the ids
    v.cE[1] = a.lut[1 * a.nmat + m];
    v.cE[3] = a.lut[3 * a.nmat + m];
    v.cE[5] = a.lut[5 * a.nmat + m];
    v.cE[0] = v.cE[2] = v.cE[4] = 0;
    if (!a.lossless) {
        v.cE[0] = a.lut[0 * a.nmat + m];
        v.cE[2] = a.lut[2 * a.nmat + m];
        v.cE[4] = a.lut[4 * a.nmat + m];
    }
    if (GN && a.gain && a.ntls != 1) v.tc = tb.tls[r8 < 0 ? 0 : r8];
    v.reg = gin ? r8 : -1;
    v.g = (long)g;
    v.Nu = gs[0]; v.Ng = gs[1]; v.Pz = gs[2]; v.Pp = gs[3];
    for (int q = 0; q < 6; ++q) v.p[q] = cv<T, R>(praw[q], a.sc[6 + q]);
    return v;
}

// E^{n+1} of one cell with J, the CPML-E correction, the TLS polarization
// (with Ez^n) and the population update (with the final Ez^{n+1}) fused:
// fdtd_update_E(_with_gain) (src/fdtd_stepper.hpp:73-88,
// src/two_level_system.hpp:320-341), CPMLBoundary::apply_after_E
// (src/boundary.hpp:248-299), update_polarization/populations_multi_region
// (src/two_level_system.hpp:585-608, 623-665).  The caller passes H^{n+1/2} at
// the cell and at its -x, -y, -z neighbours.  Cells with i, j or k == 0 are
// never updated by the reference's E pass; if they are gain cells their
// polarization and populations still advance with Ez_old = 0, as in the
// reference where Ez_old is only written on updated cells.  With a PEC
// boundary the six faces are written as zero (PECBoundary::apply_after_E,
// src/boundary.hpp:55-74).
//
// Returns the cell's E after the step: its new value, zero on a PEC face, and
// for a cell the reference does not update, v.ex0/ey0/ez0 (so a caller that
// stores fields itself, store_fields = false, passes the loaded values
// there).
template <typename T, typename R, bool CP = true, typename IX = long>
__device__ __forceinline__ H3<R> e_compute(const StepArgs<T, R>& a, const Tabs<R>& tb, int i, int j, int k,
                                           const EIn<R>& v, R hx, R hy, R hz, R hx_j, R hz_j, R hx_k, R hy_k, R hy_i,
                                           R hz_i, bool store_fields = true) {
    const IX id = ((IX)i * a.ny + j) * a.nz + k;
    const int reg = v.reg;
    const IX g = (IX)v.g;
    const bool interior = (i > 0) & (j > 0) & (k > 0);
    const bool pec_face = a.pec && (!interior || i == a.nx - 1 || j == a.ny - 1 || k == a.nz - 1);
    if (!interior) {
        if (reg >= 0) {
            const TLSCoef& c = v.tc;
            double Nu = v.Nu, Ng = v.Ng, Pz = v.Pz, Pp = v.Pp;
            const double ez = (double)v.ez0;
            (void)tls_polarize<sizeof(R) == 8>(c, Nu, Ng, ez, Pz, Pp);
            tls_populate(c, ez, 0.0, Pz, Pp, Nu, Ng);
            *at(gain_arr(a, tb, 0), g) = Nu; *at(gain_arr(a, tb, 1), g) = Ng; *at(gain_arr(a, tb, 2), g) = Pz; *at(gain_arr(a, tb, 3), g) = Pp;
        }
        if (pec_face) {
            if (store_fields) {
                stx<T, R>(a.Exw, id, (R)0, a.sc[18]);
                stx<T, R>(a.Eyw, id, (R)0, a.sc[19]);
                stx<T, R>(a.Ezw, id, (R)0, a.sc[20]);
            }
            return H3<R>{0, 0, 0};
        }
        return H3<R>{v.ex0, v.ey0, v.ez0};
    }
    const R idx = v.idx, idy = v.idy, idz = v.idz;
    // (own - neighbour) * inverse spacing, the neighbour's scale inside the FMA
    const R dHz_y = own_minus(hz, hz_j, v.hns[2]) * idy, dHy_z = own_minus(hy, hy_k, v.hns[1]) * idz;
    const R dHx_z = own_minus(hx, hx_k, v.hns[0]) * idz, dHz_x = own_minus(hz, hz_i, v.hns[2]) * idx;
    const R dHy_x = own_minus(hy, hy_i, v.hns[1]) * idx, dHx_y = own_minus(hx, hx_j, v.hns[0]) * idy;
    const R cx = dHz_y - dHy_z, cy = dHx_z - dHz_x, cz = dHy_x - dHx_y;
    const R jx = v.jx, jy = v.jy, jz = v.jz;
    const R bEx = v.cE[1], bEy = v.cE[3], bEz = v.cE[5];
    double Nu = v.Nu, Ng = v.Ng, Pz = v.Pz, Pp = v.Pp;
    R rz = cz - jz;
    const TLSCoef& tc = v.tc;
    if (reg >= 0) {
        const double dP = tls_polarize<sizeof(R) == 8>(tc, Nu, Ng, (double)v.ez0, Pz, Pp);
        rz = rz - (R)dP;
    }
    R ex, ey, ez;
    if (a.lossless) {
        ex = v.ex0 + bEx * (cx - jx);
        ey = v.ey0 + bEy * (cy - jy);
        ez = v.ez0 + bEz * rz;
    } else {
        ex = v.cE[0] * v.ex0 + bEx * (cx - jx);
        ey = v.cE[2] * v.ey0 + bEy * (cy - jy);
        ez = v.cE[4] * v.ez0 + bEz * rz;
    }
    if (CP && a.cpml) {
        const int np = a.np, nx = a.nx, ny = a.ny, nz = a.nz;
        const bool px = in_pml(i, nx, np), py = in_pml(j, ny, np), pz = in_pml(k, nz, np);
        if (px | py | pz) {
            R pn[6];
            if (py) ex += bEx * cpml_term<R>(v.p[0], dHz_y, v.ky[0], v.ky[1], v.ky[2], v.ky[3], pn[0]);
            if (pz) ex += bEx * (-cpml_term<R>(v.p[1], dHy_z, v.kz[0], v.kz[1], v.kz[2], v.kz[3], pn[1]));
            if (pz) ey += bEy * cpml_term<R>(v.p[2], dHx_z, v.kz[0], v.kz[1], v.kz[2], v.kz[3], pn[2]);
            if (px) ey += bEy * (-cpml_term<R>(v.p[3], dHz_x, v.kx[0], v.kx[1], v.kx[2], v.kx[3], pn[3]));
            if (px) ez += bEz * cpml_term<R>(v.p[4], dHy_x, v.kx[0], v.kx[1], v.kx[2], v.kx[3], pn[4]);
            if (py) ez += bEz * (-cpml_term<R>(v.p[5], dHx_y, v.ky[0], v.ky[1], v.ky[2], v.ky[3], pn[5]));
            if (py) {
                const IX q = psi_y<IX>(i, j, k, nx, ny, nz, np);
                stx<T, R>(psi_rd(a, tb, 0), q, pn[0], a.sc[24]);
                stx<T, R>(psi_rd(a, tb, 5), q, pn[5], a.sc[29]);
            }
            if (pz) {
                const IX q = psi_z<IX>(i, j, k, nx, ny, nz, np);
                stx<T, R>(psi_rd(a, tb, 1), q, pn[1], a.sc[25]);
                stx<T, R>(psi_rd(a, tb, 2), q, pn[2], a.sc[26]);
            }
            if (px) {
                const IX q = psi_x<IX>(i, j, k, nx, ny, nz, np);
                stx<T, R>(psi_rd(a, tb, 3), q, pn[3], a.sc[27]);
                stx<T, R>(psi_rd(a, tb, 4), q, pn[4], a.sc[28]);
            }
        }
    }
    if (pec_face) { ex = 0; ey = 0; ez = 0; }
    if (store_fields) {
        stx<T, R>(a.Exw, id, ex, a.sc[18]);
        stx<T, R>(a.Eyw, id, ey, a.sc[19]);
        stx<T, R>(a.Ezw, id, ez, a.sc[20]);
    }
    if (reg >= 0) {
        const TLSCoef& c = tc;
        // The reference stores Ez and reads it back for the population pass.
        const double ez_new = stored_value<T, R>(ez, a.sc[2], a.sc[20]);
        tls_populate(c, ez_new, (double)v.ez0, Pz, Pp, Nu, Ng);
        *at(gain_arr(a, tb, 0), g) = Nu; *at(gain_arr(a, tb, 1), g) = Ng; *at(gain_arr(a, tb, 2), g) = Pz; *at(gain_arr(a, tb, 3), g) = Pp;
    }
    return H3<R>{ex, ey, ez};
}

// E update of one cell (i, kw and jw uniform across the wavefront, as for
// e_gather).  CP and GN compile the CPML and gain code in (the default) or out;
// COH reads H through loads that bypass the CU's L1 (e_gather).
template <typename T, typename R, int KW = 64, bool CP = true, bool GN = true, typename IX = long, bool COH = false>
__device__ __forceinline__ void e_cell(const StepArgs<T, R>& a, const Tabs<R>& tb, int i, int j, int k, int kw, int jw) {
    R h[9];
    const EIn<R> v = e_gather<T, R, KW, CP, GN, IX, COH>(a, tb, i, j, k, kw, jw, h);
    e_compute<T, R, CP, IX>(a, tb, i, j, k, v, h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], h[8]);
}



// E update kernel of the two-kernel step: H^{n+1/2} neighbours from memory.
template <typename T, typename R, int KW, typename IX = long>
__global__ void __launch_bounds__(256) k_update_E(StepArgs<T, R> a) {
    FDTD_SCALES_IN_REGISTERS(a)
    const Tabs<R> tb = global_tabs(a);
    const Tile t = tile_of<KW>(a);
    const int i = t.i, j = t.j, k = t.k;
    if (k >= a.nz || j >= a.ny || i >= a.nx) return;
    e_cell<T, R, KW, true, true, IX>(a, tb, i, j, k, t.kw, t.jw);
}


struct ProbeSlot {
    long id;
    int comp;
    double* out;
};
struct ProbeArgs {
    int n;
    // 0: out is the record's address; 1: out holds the address of a device
    // word that holds it (null: no record this step), so a launch recorded
    // in a graph takes each replay's slots from a table written before it
    int ind;
    ProbeSlot p[16];
};
__device__ __forceinline__ void probe_store(const ProbeArgs& pa, int t, double v) {
    double* dst = pa.ind ? *reinterpret_cast<double* const*>(pa.p[t].out) : pa.p[t].out;
    if (dst) *dst = v;
}

template <typename T, typename R>
__device__ __forceinline__ double field_value(const StepArgs<T, R>& a, int comp, long id) {
    if (comp < 6) {  // a raw component: one load
        const T* p = comp == 0 ? a.Ex : comp == 1 ? a.Ey : comp == 2 ? a.Ez : comp == 3 ? a.Hx : comp == 4 ? a.Hy : a.Hz;
        return (double)ld<T, R>(p, id, a.sc[comp]);
    }
    const double ex = (double)ld<T, R>(a.Ex, id, a.sc[0]), ey = (double)ld<T, R>(a.Ey, id, a.sc[1]),
                 ez = (double)ld<T, R>(a.Ez, id, a.sc[2]);
    const double hx = (double)ld<T, R>(a.Hx, id, a.sc[3]), hy = (double)ld<T, R>(a.Hy, id, a.sc[4]),
                 hz = (double)ld<T, R>(a.Hz, id, a.sc[5]);
    switch (comp) {
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
        const double sx = ey * hz - ez * hy, sy = ez * hx - ex * hz, sz = ex * hy - ey * hx;
        return sqrt(sx * sx + sy * sy + sz * sz);
    }
    }
    return 0.0;
}

// ---- TLS diagnostics (PopulationTimeSeriesDetector::record,
// PolarizationDiagnostics::record, src/two_level_detectors.hpp:70-103,
// 230-257) ----
// The reference sums over the region cells in idx3 order.  A parallel sum
// cannot keep that order, so the terms are accumulated in double-double
// (Knuth's TwoSum: hi + lo carries about 106 bits) and the block partials
// are merged in a fixed order: the result is the correctly rounded sum of
// the terms (short of contrived cancellation), the same on every run, and
// within a few ulps per term of the reference's sequential sum.
struct DD {
    double hi, lo;
};
__device__ __forceinline__ void dd_add(DD& s, double x) {
    const double t = s.hi + x;
    const double bp = t - s.hi;
    const double e = (s.hi - (t - bp)) + (x - bp);
    s.hi = t;
    s.lo += e;
}
__device__ __forceinline__ DD dd_merge(const DD& a, const DD& b) {
    const double t = a.hi + b.hi;
    const double bp = t - a.hi;
    const double e = (a.hi - (t - bp)) + (b.hi - bp);
    const double lo = (a.lo + b.lo) + e;
    const double hi = t + lo;
    return DD{hi, lo - (hi - t)};
}

// Block partials of (sum Nu dV, sum Ng dV, sum -0.5 Pz Ez dV) over the region
// cells of the gain box, dV = dx[i]*dy[j]*dz[k]; part[block][3][2].
template <typename T, typename R>
__global__ void __launch_bounds__(256) k_tls_partial(StepArgs<T, R> a, const double* dx, const double* dy,
                                                     const double* dz, double* part) {
    __shared__ DD sh[3][256];
    DD s[3] = {{0, 0}, {0, 0}, {0, 0}};
    const long total = (long)a.gnx * a.gny * a.gnz;
    for (long g = (long)blockIdx.x * 256 + threadIdx.x; g < total; g += (long)gridDim.x * 256) {
        if (a.greg[g] < 0) continue;
        const long r = g / a.gnz;
        const int k = a.g0z + (int)(g - r * a.gnz);
        const int j = a.g0y + (int)(r % a.gny);
        const int i = a.g0x + (int)(r / a.gny);
        const long id = ((long)i * a.ny + j) * a.nz + k;
        const double dV = dx[i] * dy[j] * dz[k];
        const double ez = (double)ld<T, R>(a.Ez, id, a.sc[2]);
        dd_add(s[0], a.gNu[g] * dV);
        dd_add(s[1], a.gNg[g] * dV);
        dd_add(s[2], -0.5 * a.gPz[g] * ez * dV);
    }
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = dd_merge(sh[q][threadIdx.x], sh[q][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) {
            part[(long)blockIdx.x * 6 + 2 * q] = sh[q][0].hi;
            part[(long)blockIdx.x * 6 + 2 * q + 1] = sh[q][0].lo;
        }
}

// One record: rec[0..4] = total Nu, total Ng, interaction energy, Pz and
// dPz/dt at the monitor cell (gain-box index gmon, -1 outside the box; zero
// where the cell is not a region cell, as the reference's zero-initialised
// state arrays are there).  One thread merges the partials in block order.
__global__ void k_tls_final(const double* part, int nparts, const int8_t* greg, const double* gPz, const double* gPp,
                            const TLSCoef* tls, long gmon, double* rec) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int q = 0; q < 3; ++q) {
        DD s{0, 0};
        for (int b = 0; b < nparts; ++b) s = dd_merge(s, DD{part[(long)b * 6 + 2 * q], part[(long)b * 6 + 2 * q + 1]});
        rec[q] = s.hi + s.lo;
    }
    double pz = 0, dpz = 0;
    if (gmon >= 0) {
        const int reg = greg[gmon];
        if (reg >= 0) {
            pz = gPz[gmon];
            dpz = (gPz[gmon] - gPp[gmon]) * tls[reg].inv_dt;
        }
    }
    rec[3] = pz;
    rec[4] = dpz;
}

// PointFieldDetector::record_after_E / get_field_value
// (src/detectors/point_field_detector.hpp:111-124, src/detectors/idetector.hpp:79-109).
template <typename T, typename R>
__global__ void k_probes(StepArgs<T, R> a, ProbeArgs pa) {
    const int t = threadIdx.x;
    if (t >= pa.n) return;
    probe_store(pa, t, field_value<T, R>(a, pa.p[t].comp, pa.p[t].id));
}

struct DftArgs {
    int comp, i0, j0, k0, ny_b, nz_b;
    long ncell;
    int nf;
    double wr[8], wi[8];
    // where non-null, the phase factors are read from device memory
    // (wtab[f] and wtab[nf + f]) instead of wr/wi: a launch recorded in a
    // graph takes each replay's phases from a table written before it
    const double* wtab;
    double* acc;
};
__device__ __forceinline__ double dft_wr(const DftArgs& d, int f) { return d.wtab ? d.wtab[f] : d.wr[f]; }
__device__ __forceinline__ double dft_wi(const DftArgs& d, int f) { return d.wtab ? d.wtab[d.nf + f] : d.wi[f]; }

template <typename T, typename R>
__device__ __forceinline__ void dft_cell(const StepArgs<T, R>& a, const DftArgs& d, long c) {
    const int kk = (int)(c % d.nz_b);
    const long r = c / d.nz_b;
    const int jj = (int)(r % d.ny_b);
    const int ii = (int)(r / d.ny_b);
    const long id = ((long)(d.i0 + ii) * a.ny + (d.j0 + jj)) * a.nz + (d.k0 + kk);
    const double F = field_value<T, R>(a, d.comp, id);
    for (int f = 0; f < d.nf; ++f) {
        double* p = d.acc + ((long)f * d.ncell + c) * 2;
        p[0] += F * d.wr[f];
        p[1] += F * d.wi[f];
    }
}

// One step's observers on one domain in one launch: DftMonitor accumulation
// and, in the last block when pa.n > 0, the point probes.  Each accumulator
// gets the same += F * w as in dft_cell.  A box with a k extent reads its
// field along k, coalesced: one thread per cell and frequency (blocks
// [f * bpf, (f + 1) * bpf) hold frequency f).  A box one cell thick in k (a
// z-normal plane) reads one value per nz-strided line: one thread per cell
// then reads it once for every frequency (`per_cell`).
struct DftLaunch {
    DftArgs d;
    FastDiv fz, fy;  // divisions by nz_b and ny_b of the box
    unsigned bpf;    // blocks per frequency (per_cell: blocks of the box)
    unsigned per_cell;
};

template <typename T, typename R>
__global__ void __launch_bounds__(256) k_observe(StepArgs<T, R> a, DftLaunch L, ProbeArgs pa) {
    const unsigned nb_dft = L.per_cell ? L.bpf : L.bpf * (unsigned)L.d.nf;
    if (blockIdx.x >= nb_dft) {
        const int t = threadIdx.x;
        if (t < pa.n) probe_store(pa, t, field_value<T, R>(a, pa.p[t].comp, pa.p[t].id));
        return;
    }
    if (L.per_cell) {
        const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= (unsigned)L.d.ncell) return;
        const unsigned r = fdiv(c, L.fz);
        const unsigned ii = fdiv(r, L.fy);
        const int kk = (int)(c - r * (unsigned)L.d.nz_b), jj = (int)(r - ii * (unsigned)L.d.ny_b);
        const long id = ((long)(L.d.i0 + (int)ii) * a.ny + (L.d.j0 + jj)) * a.nz + (L.d.k0 + kk);
        const double F = field_value<T, R>(a, L.d.comp, id);
        for (int f = 0; f < L.d.nf; ++f) {
            double* p = L.d.acc + ((long)f * L.d.ncell + c) * 2;
            const double re = p[0], im = p[1];
            p[0] = re + F * dft_wr(L.d, f);
            p[1] = im + F * dft_wi(L.d, f);
        }
        return;
    }
    const unsigned f = blockIdx.x / L.bpf;
    const unsigned c = (blockIdx.x - f * L.bpf) * blockDim.x + threadIdx.x;
    if (c >= (unsigned)L.d.ncell) return;
    const unsigned r = fdiv(c, L.fz);
    const unsigned ii = fdiv(r, L.fy);
    const int kk = (int)(c - r * (unsigned)L.d.nz_b), jj = (int)(r - ii * (unsigned)L.d.ny_b);
    const long id = ((long)(L.d.i0 + (int)ii) * a.ny + (L.d.j0 + jj)) * a.nz + (L.d.k0 + kk);
    double* p = L.d.acc + ((long)f * L.d.ncell + c) * 2;
    const double re = p[0], im = p[1];
    const double F = field_value<T, R>(a, L.d.comp, id);
    p[0] = re + F * dft_wr(L.d, (int)f);
    p[1] = im + F * dft_wi(L.d, (int)f);
}

// Stability reduction (StabilityMonitor::check_stability inputs,
// src/stability_monitor.hpp:51-121, and calculate_conservation_error
// :155-170).  max|.| is taken over the absolute-value bit patterns: for
// non-negative IEEE values the integer order is the numeric order and every
// NaN/Inf pattern is above every finite one, so one atomicMax yields both the
// maximum and the NaN/Inf flag.
struct MonOut {
    unsigned long long maxE, maxH, maxP, nanN;
    double pop_total, pop_init;
};
static_assert(sizeof(MonOut) == 48, "monitor record layout (48 B, six 64-bit words)");

// One monitor record per work-group: wavefront reduction (64 lanes), then
// across the work-group's wavefronts in LDS in wavefront order.  Every thread
// of the work-group calls it; thread 0 gets the record.
__device__ __forceinline__ MonOut block_record(unsigned long long mE, unsigned long long mH, unsigned long long mP,
                                               unsigned long long nN, double pt, double pi) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o;
        o = __shfl_down(mE, off); mE = o > mE ? o : mE;
        o = __shfl_down(mH, off); mH = o > mH ? o : mH;
        o = __shfl_down(mP, off); mP = o > mP ? o : mP;
        o = __shfl_down(nN, off); nN = o > nN ? o : nN;
        pt += __shfl_down(pt, off);
        pi += __shfl_down(pi, off);
    }
    __shared__ unsigned long long su[4][4];
    __shared__ double sd[2][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        su[0][w] = mE; su[1][w] = mH; su[2][w] = mP; su[3][w] = nN;
        sd[0][w] = pt; sd[1][w] = pi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        for (int q = 1; q < nw; ++q) {
            mE = su[0][q] > mE ? su[0][q] : mE;
            mH = su[1][q] > mH ? su[1][q] : mH;
            mP = su[2][q] > mP ? su[2][q] : mP;
            nN = su[3][q] > nN ? su[3][q] : nN;
            pt += sd[0][q];
            pi += sd[1][q];
        }
    }
    return MonOut{mE, mH, mP, nN, pt, pi};
}

template <typename T, typename R>
__device__ __forceinline__ unsigned long long absbits(const T* p, long i, R s) {
    const double v = fabs((double)ld<T, R>(p, i, s));
    return (unsigned long long)__double_as_longlong(v);
}

// |raw| of a stored element: the sign bit cleared.  For the stored formats
// (IEEE binary64/32/16) the integer order of these patterns is the numeric
// order of |value|, with every NaN/Inf above every finite value, and the
// conversion to the FP64 value (times a positive scale for FP16 storage) is
// non-decreasing; so the largest pattern, converted once, is what absbits
// would give as the maximum over the elements.
template <typename T>
using RawU = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;

template <typename T>
__device__ __forceinline__ RawU<T> raw_absmax16(const uint4 q) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long x = (((unsigned long long)q.y << 32) | q.x) & 0x7fffffffffffffffull;
        const unsigned long long y = (((unsigned long long)q.w << 32) | q.z) & 0x7fffffffffffffffull;
        return x > y ? x : y;
    } else if constexpr (sizeof(T) == 4) {
        return max(max(q.x & 0x7fffffffu, q.y & 0x7fffffffu), max(q.z & 0x7fffffffu, q.w & 0x7fffffffu));
    } else {  // eight binary16 patterns, two per word
        unsigned m = 0;
        const unsigned w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) m = max(m, max(w[t] & 0x7fffu, (w[t] >> 16) & 0x7fffu));
        return m;
    }
}

template <typename T, typename R>
__device__ __forceinline__ unsigned long long raw_to_absbits(RawU<T> m, R scale) {
    T t;
    if constexpr (sizeof(T) == 2) {
        const uint16_t b = (uint16_t)m;
        __builtin_memcpy(&t, &b, 2);
    } else {
        __builtin_memcpy(&t, &m, sizeof(T));
    }
    return absbits<T, R>(&t, 0, scale);
}

template <typename T, typename R, bool VEC = false>
__device__ __forceinline__ void monitor_body(const StepArgs<T, R>& a, long N, long ngain, MonOut* out, int bid,
                                             int nblk) {
    unsigned long long mE = 0, mH = 0, mP = 0, nN = 0;
    double pt = 0.0, pi = 0.0;
    const long stride = (long)nblk * blockDim.x;
    if constexpr (VEC) {
        // 16-byte loads (the arrays are 16-byte aligned), per-array maxima of
        // |raw|, one conversion per array at the end
        constexpr int V = 16 / (int)sizeof(T);
        const long nv = N / V;
        const T* f[6] = {a.Ex, a.Ey, a.Ez, a.Hx, a.Hy, a.Hz};
        RawU<T> m[6] = {0, 0, 0, 0, 0, 0};
        for (long v = (long)bid * blockDim.x + threadIdx.x; v < nv; v += stride) {
            uint4 q[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) q[c] = reinterpret_cast<const uint4*>(f[c])[v];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const RawU<T> x = raw_absmax16<T>(q[c]);
                m[c] = x > m[c] ? x : m[c];
            }
        }
        for (long id = nv * V + (long)bid * blockDim.x + threadIdx.x; id < N; id += stride) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                RawU<T> x;
                if constexpr (sizeof(T) == 2) {
                    x = (RawU<T>)(reinterpret_cast<const uint16_t*>(f[c])[id] & 0x7fffu);
                } else {
                    __builtin_memcpy(&x, f[c] + id, sizeof(T));
                    x &= sizeof(T) == 8 ? (RawU<T>)0x7fffffffffffffffull : (RawU<T>)0x7fffffffu;
                }
                m[c] = x > m[c] ? x : m[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const unsigned long long v = raw_to_absbits<T, R>(m[c], a.sc[c]);
            if (c < 3) mE = v > mE ? v : mE;
            else mH = v > mH ? v : mH;
        }
    } else {
        for (long id = (long)bid * blockDim.x + threadIdx.x; id < N; id += stride) {
            unsigned long long v;
            v = absbits<T, R>(a.Ex, id, a.sc[0]); mE = v > mE ? v : mE;
            v = absbits<T, R>(a.Ey, id, a.sc[1]); mE = v > mE ? v : mE;
            v = absbits<T, R>(a.Ez, id, a.sc[2]); mE = v > mE ? v : mE;
            v = absbits<T, R>(a.Hx, id, a.sc[3]); mH = v > mH ? v : mH;
            v = absbits<T, R>(a.Hy, id, a.sc[4]); mH = v > mH ? v : mH;
            v = absbits<T, R>(a.Hz, id, a.sc[5]); mH = v > mH ? v : mH;
        }
    }
    for (long g = (long)bid * blockDim.x + threadIdx.x; g < ngain; g += stride) {
        const int r = a.greg[g];
        if (r < 0) continue;
        const double pz = a.gPz[g], pp = a.gPp[g], nu = a.gNu[g], ng = a.gNg[g];
        unsigned long long v = (unsigned long long)__double_as_longlong(fabs(pz));
        mP = v > mP ? v : mP;
        v = (unsigned long long)__double_as_longlong(fabs(pp));
        // Pz_prev only contributes to the NaN test (dPz_dt is formed from Pz, Pz_prev)
        if (v >= 0x7ff0000000000000ull) mP = v > mP ? v : mP;
        v = (unsigned long long)__double_as_longlong(fabs(nu));
        nN = v > nN ? v : nN;
        v = (unsigned long long)__double_as_longlong(fabs(ng));
        nN = v > nN ? v : nN;
        if (a.tls[r].N0 > 1e10) {
            pt += (nu + ng);
            pi += a.tls[r].N0;
        }
    }
    const MonOut rec = block_record(mE, mH, mP, nN, pt, pi);
    if (threadIdx.x == 0) {
        if constexpr (VEC) {
            out[bid] = rec;  // one partial record per block; k_monitor_final merges them
        } else {
            if (rec.maxE) atomicMax(&out->maxE, rec.maxE);
            if (rec.maxH) atomicMax(&out->maxH, rec.maxH);
            if (rec.maxP) atomicMax(&out->maxP, rec.maxP);
            if (rec.nanN) atomicMax(&out->nanN, rec.nanN);
            if (rec.pop_total != 0.0) atomicAdd(&out->pop_total, rec.pop_total);
            if (rec.pop_init != 0.0) atomicAdd(&out->pop_init, rec.pop_init);
        }
    }
}

template <typename T, typename R, bool VEC>
__global__ void __launch_bounds__(256) k_monitor(StepArgs<T, R> a, long N, long ngain, MonOut* out) {
    monitor_body<T, R, VEC>(a, N, ngain, out, blockIdx.x, gridDim.x);
}

// Merges the per-block monitor records in a fixed order (so the population
// sums are the same on every run) into the step's record.
// fmax (optional): nf field-maxima slots of MON_SLOT_WORDS words each,
// (max |E|, max |H|) first, filled by k_update_H_mon.
__global__ void __launch_bounds__(64) k_monitor_final(const MonOut* __restrict__ part, int n, MonOut* slot,
                                                      const unsigned long long* __restrict__ fmax, int nf) {
    unsigned long long mE = 0, mH = 0, mP = 0, nN = 0;
    double pt = 0.0, pi = 0.0;
    for (int q = threadIdx.x; q < nf; q += 64) {
        const unsigned long long e = fmax[q * MON_SLOT_WORDS], h = fmax[q * MON_SLOT_WORDS + 1];
        mE = e > mE ? e : mE;
        mH = h > mH ? h : mH;
    }
    for (int q = threadIdx.x; q < n; q += 64) {
        const MonOut r = part[q];
        mE = r.maxE > mE ? r.maxE : mE;
        mH = r.maxH > mH ? r.maxH : mH;
        mP = r.maxP > mP ? r.maxP : mP;
        nN = r.nanN > nN ? r.nanN : nN;
        pt += r.pop_total;
        pi += r.pop_init;
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o;
        o = __shfl_down(mE, off); mE = o > mE ? o : mE;
        o = __shfl_down(mH, off); mH = o > mH ? o : mH;
        o = __shfl_down(mP, off); mP = o > mP ? o : mP;
        o = __shfl_down(nN, off); nN = o > nN ? o : nN;
        pt += __shfl_down(pt, off);
        pi += __shfl_down(pi, off);
    }
    if (threadIdx.x == 0) *slot = MonOut{mE, mH, mP, nN, pt, pi};
}

// Copies n 64-bit words (device records into pinned host memory at the end
// of a call, so the host reads them once the stream is idle instead of
// queueing a copy behind other streams' work).
// Copies n doubles (grid-stride), e.g. into mapped host memory.
__global__ void __launch_bounds__(256) k_copy_f64(const double* __restrict__ src, double* __restrict__ dst, long n) {
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) dst[q] = src[q];
}
__global__ void __launch_bounds__(64) k_copy_words(const unsigned long long* __restrict__ src,
                                                   unsigned long long* __restrict__ dst, long n) {
    for (long q = threadIdx.x; q < n; q += 64) dst[q] = src[q];
}

// Zeroes n 64-bit words: a kernel rather than a memset node, so that the
// clear also holds inside a captured graph.
__global__ void __launch_bounds__(256) k_zero_words(unsigned long long* __restrict__ p, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) p[i] = 0ull;
}

// ---- Multi-domain (batch) step ----
// Many independent domains of one precision advance together: per step one
// persistent H launch and one persistent E launch cover every domain, their
// work-groups taking `chunk` tiles (64 k-values by blockDim.y rows by
// `planes` consecutive x-planes) at a time from a device counter, and one launch each records the
// probes, accumulates the DFT monitors and reduces the stability monitor of
// every domain.  Step-dependent inputs (source values, probe slots, DFT
// phases, monitor slots) are indexed by a device-side step counter, so a block
// of steps can be captured once in a HIP graph and replayed.  Each cell goes
// through the same operations as in the single-domain step, so a domain's
// results do not depend on the batch it runs in.
// The tile list of a batch pass is cut into BATCH_SEGS contiguous segments,
// one per XCD: a work-group takes tiles from its own XCD's segment first, so
// neighbouring tiles (whose rows and planes overlap as halos) run under one
// L2, and moves on to the other segments when its own is done.
constexpr int BATCH_SEGS = 8;
struct BatchCtl {
    int ctr[2][BATCH_SEGS][16];  // tiles taken per pass (H, E) and segment, one 64-byte line each
    int lstep;                   // step index within the current block
    int pad[15];
};
struct BatchDom {
    long n0;         // global step index of the block's first step
    long N, ngain;   // cells, gain-box cells (monitor)
    long mon_first;  // first monitor slot of the block: ceil(n0 / interval)
    int interval;    // monitor cadence (0: off)
    int steps;       // steps this member takes in the block (it idles after)
};
struct BatchProbe {
    int dom, comp;
    long id;
    long save_every, cap;
    double* out;
};
struct BatchDft {
    int dom;
    DftArgs d;       // wr/wi unused: phases are computed from the step index
    double dt;
    double omega[8];
};

__global__ void k_batch_reset(BatchCtl* c) {
    for (int q = threadIdx.x; q < 2 * BATCH_SEGS * 16; q += blockDim.x) (&c->ctr[0][0][0])[q] = 0;
    if (threadIdx.x == 0) c->lstep = -1;
}
__global__ void k_batch_begin(BatchCtl* c) {
    for (int q = threadIdx.x; q < 2 * BATCH_SEGS * 16; q += blockDim.x) (&c->ctr[0][0][0])[q] = 0;
    if (threadIdx.x == 0) c->lstep += 1;
}

// Domain of tile t: cum[d] <= t < cum[d+1].
__device__ __forceinline__ int batch_domain(const int* __restrict__ cum, int ndom, int t) {
    int lo = 0, hi = ndom;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cum[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Persistent H (EPASS false) or E pass over every member of a batch.  A tile
// of a domain is KW consecutive k of blockDim.y * (64 / KW) rows over
// `planes` consecutive planes: each wavefront (one threadIdx.y) covers KW k
// of 64 / KW rows, as in the plain kernels.  The host stores the tile
// divisors of each domain's tiling in its table entry (tk: k-tiles per row
// group, tj: row groups per plane group).  IX is the cell-offset type
// (unsigned where every member's arrays are below 4 GiB).
// One tile of a batch pass: member d's tile t of the pass's tile list (its
// tiles k-block fastest, then rows, then groups of `planes` planes).
template <typename T, typename R, bool EPASS, int KW, typename IX>
__device__ __forceinline__ void batch_tile(const StepArgs<T, R>* __restrict__ doms, const char* anchor,
                                           const int* __restrict__ cum, int d, int t, int lstep, int planes, int ty) {
    constexpr int G = 64 / KW;  // rows per wavefront
    StepArgs<T, R> a = doms[d];
    shift_args(a, anchor, 1);
    a.srcv += (long)lstep * a.nsrc;  // this step's source values
    FDTD_SCALES_IN_REGISTERS(a)
    const Tabs<R> tb = global_tabs(a);
    const unsigned lt = (unsigned)(t - cum[d]);
    const unsigned r = fdiv(lt, a.tk), ig = fdiv(r, a.tj);
    const int kt = (int)(lt - r * a.tk.d), jt = (int)(r - ig * a.tj.d);
    const int kw = kt * KW, jw = (jt * (int)blockDim.y + ty) * G;
    const int k = kw + (int)threadIdx.x % KW, j = jw + (int)threadIdx.x / KW;
    const int i0 = (int)ig * planes, i1 = min(i0 + planes, a.nx);
    for (int i = i0; i < i1; ++i) {
        if (!EPASS) {
            if (k < a.nz - 1 && j < a.ny - 1 && i < a.nx - 1)
                (void)h_compute<T, R, true, IX>(a, tb, i, j, k, h_gather<T, R, KW, true, IX>(a, tb, i, j, k, kw, jw),
                                                true);
        } else if (k < a.nz && j < a.ny) {
            e_cell<T, R, KW, true, true, IX>(a, tb, i, j, k, kw, jw);
        }
    }
}

template <typename T, typename R, bool EPASS, int KW, typename IX>
__global__ void __launch_bounds__(256) k_batch_step(const StepArgs<T, R>* __restrict__ doms, const char* anchor,
                                                    const int* __restrict__ cum, const BatchDom* __restrict__ bd,
                                                    int ndom, int chunk, int planes, BatchCtl* ctl) {
    __shared__ int first_s;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int total = cum[ndom];
    const int lstep = ctl->lstep;
    // work-groups are dealt to the XCDs round-robin by id
    const int home = (int)(blockIdx.x % BATCH_SEGS);
    int seg = home, lo = (int)((long)total * seg / BATCH_SEGS), hi = (int)((long)total * (seg + 1) / BATCH_SEGS);
    for (;;) {
        if (threadIdx.x == 0 && threadIdx.y == 0) first_s = lo + atomicAdd(&ctl->ctr[EPASS ? 1 : 0][seg][0], chunk);
        __syncthreads();
        const int first = __builtin_amdgcn_readfirstlane(first_s);  // one value for the work-group
        __syncthreads();
        if (first >= hi) {
            // this segment is done: the next one, until all are
            seg = (seg + 1) % BATCH_SEGS;
            if (seg == home) return;
            lo = (int)((long)total * seg / BATCH_SEGS);
            hi = (int)((long)total * (seg + 1) / BATCH_SEGS);
            continue;
        }
        const int last = min(first + chunk, hi);
        int d = batch_domain(cum, ndom, first);
        for (int t = first; t < last; ++t) {
            while (t >= cum[d + 1]) ++d;
            if (lstep >= bd[d].steps) continue;
            batch_tile<T, R, EPASS, KW, IX>(doms, anchor, cum, d, t, lstep, planes, ty);
        }
    }
}

// The same pass with one work-group per tile (tuning chunk 0): no counter and
// no loop, so a work-group keeps the plain kernels' short life and register
// footprint.  Ids are dealt to the XCDs round-robin; the remap gives each XCD
// one contiguous eighth of the tile list, as the persistent form's segments
// do, so neighbouring tiles share an L2.
template <typename T, typename R, bool EPASS, int KW, typename IX = long>
__global__ void __launch_bounds__(256) k_batch_tiles(const StepArgs<T, R>* __restrict__ doms, const char* anchor,
                                                     const int* __restrict__ cum, const BatchDom* __restrict__ bd,
                                                     int ndom, int planes, const BatchCtl* ctl) {
    const int total = cum[ndom];
    const int per = (int)(gridDim.x / BATCH_SEGS);
    const int t = (int)(blockIdx.x % BATCH_SEGS) * per + (int)(blockIdx.x / BATCH_SEGS);
    if (t >= total) return;
    const int d = batch_domain(cum, ndom, t);
    const int lstep = ctl->lstep;
    if (lstep >= bd[d].steps) return;
    batch_tile<T, R, EPASS, KW, IX>(doms, anchor, cum, d, t, lstep, planes, __builtin_amdgcn_readfirstlane(threadIdx.y));
}

template <typename T, typename R>
__global__ void k_batch_probes(const StepArgs<T, R>* __restrict__ doms, const char* anchor,
                               const BatchDom* __restrict__ bd, const BatchProbe* __restrict__ pr, int npr,
                               const BatchCtl* ctl) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npr) return;
    const BatchProbe b = pr[p];
    if (ctl->lstep >= bd[b.dom].steps) return;
    const long n = bd[b.dom].n0 + ctl->lstep;
    if (n % b.save_every != 0) return;
    const long slot = n / b.save_every;
    if (slot >= b.cap) return;
    StepArgs<T, R> a = doms[b.dom];
    shift_args(a, anchor, 1);
    b.out[slot] = field_value<T, R>(a, b.comp, b.id);
}

template <typename T, typename R>
__global__ void k_batch_dft(const StepArgs<T, R>* __restrict__ doms, const char* anchor,
                            const BatchDom* __restrict__ bd, const BatchDft* __restrict__ dd, const BatchCtl* ctl) {
    const BatchDft& m = dd[blockIdx.y];
    if (ctl->lstep >= bd[m.dom].steps) return;
    DftArgs d = m.d;
    const long n = bd[m.dom].n0 + ctl->lstep;
    const double t = (d.comp < 3) ? ((double)n + 1.0) * m.dt : ((double)n + 0.5) * m.dt;
    for (int f = 0; f < d.nf; ++f) {
        d.wr[f] = cos(m.omega[f] * t) * m.dt;
        d.wi[f] = -sin(m.omega[f] * t) * m.dt;
    }
    StepArgs<T, R> a = doms[m.dom];
    shift_args(a, anchor, 1);
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < d.ncell; c += (long)gridDim.x * blockDim.x)
        dft_cell<T, R>(a, d, c);
}

template <typename T, typename R>
__global__ void __launch_bounds__(256) k_batch_monitor(const StepArgs<T, R>* __restrict__ doms, const char* anchor,
                                                       const BatchDom* __restrict__ bd, MonOut* out, int slots_per_dom,
                                                       const BatchCtl* ctl) {
    const int d = blockIdx.y;
    const BatchDom m = bd[d];
    if (m.interval <= 0 || ctl->lstep >= m.steps) return;
    const long n = m.n0 + ctl->lstep;
    if (n % m.interval != 0) return;
    const long slot = n / m.interval - m.mon_first;
    if (slot < 0 || slot >= slots_per_dom) return;
    StepArgs<T, R> a = doms[d];
    shift_args(a, anchor, 1);
    monitor_body<T, R>(a, m.N, m.ngain, out + (long)d * slots_per_dom + slot, blockIdx.x, gridDim.x);
}

// ---- FP16 storage: adaptive per-component power-of-two scales ----
// Maxima of |stored value| per array (6 fields, 12 psi slabs), as float bits
// (non-negative floats order like their bit patterns).
struct ArrSet {
    const uint16_t* p[18];
    long n[18];
};
// |x| of an FP16 value is its bit pattern with the sign bit cleared, and for
// non-negative FP16 values the integer order of the patterns is the numeric
// order (NaN/Inf above every finite value), so the maximum is taken on the
// 16-bit patterns, eight per 16-byte load.  blockIdx.y is the array.
__global__ void __launch_bounds__(256) k_fp16_maxima(ArrSet s, unsigned* maxbits) {
    const int c = blockIdx.y;
    const uint16_t* p = s.p[c];
    if (!p) return;
    const long n = s.n[c];
    const long n8 = n / 8;  // hipMalloc'ed arrays are 16-byte aligned
    const uint4* v = reinterpret_cast<const uint4*>(p);
    unsigned m = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    // four independent 16-byte loads per thread and iteration
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += 4 * stride) {
        uint4 w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long iu = i + u * stride;
            w[u] = iu < n8 ? v[iu] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned x[4] = {w[u].x & 0x7fff7fffu, w[u].y & 0x7fff7fffu, w[u].z & 0x7fff7fffu,
                                   w[u].w & 0x7fff7fffu};
            for (int q = 0; q < 4; ++q) {
                const unsigned lo = x[q] & 0xffffu, hi = x[q] >> 16;
                m = lo > m ? lo : m;
                m = hi > m ? hi : m;
            }
        }
    }
    for (long i = n8 * 8 + (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned b = p[i] & 0x7fffu;
        m = b > m ? b : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_down(m, off);
        m = o > m ? o : m;
    }
    // one record per (array, block), no atomics: same-address atomics from
    // every wavefront serialized at the L2 and held the scan at half speed
    __shared__ unsigned wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wm[w] > m ? wm[w] : m;
        // as the float bit pattern of |value| (what the plan reads)
        const float f = __half2float(__ushort_as_half((unsigned short)m));
        maxbits[(long)c * gridDim.x + blockIdx.x] = __float_as_uint(f);
    }
}
// Decide new scales: keep the largest stored magnitude near 2^8 (2^8 x
// headroom below the FP16 maximum for growth between two checks; values down
// to 2^-22 of the maximum stay normal).  A scale moves only when the maximum
// leaves [2^2, 2^11]; the stored values are then multiplied by an exact
// power of two.
// sc[36 + c] is the floor of scale c: a scale never drops below it, so a
// field that is still growing from zero keeps the scale its physical
// magnitude calls for instead of one fitted to its first tiny values.
// One 64-thread block per array c: the maximum over its `nblk` block records
// from k_fp16_maxima, then the plan.  factor[0..17] receives the rescale
// factors; factor[18..19] are the non-finite flags (unsigned words).
__global__ void __launch_bounds__(64) k_fp16_plan(float* sc, float* factor, const unsigned* maxbits, int nblk) {
    const int c = blockIdx.x;
    unsigned mb = 0;
    for (int b = threadIdx.x; b < nblk; b += 64) {
        const unsigned x = maxbits[(long)c * nblk + b];
        mb = x > mb ? x : mb;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_down(mb, off);
        mb = o > mb ? o : mb;
    }
    if (threadIdx.x != 0) return;
    const float m = __uint_as_float(mb);
    float f = 1.0f;
    if (mb >= 0x7f800000u) {
        // a NaN or Inf is stored: keep the scale (frexpf of a non-finite
        // value has no defined exponent) and raise the sticky flag of the
        // array's class (word 18: E fields and psi_E, word 19: H fields and
        // psi_H), which the host turns into a failed stability check
        const int cls = (c < 3 || (c >= 6 && c < 12)) ? 18 : 19;
        atomicOr(reinterpret_cast<unsigned*>(factor) + cls, 1u);
    } else if (m > 2048.0f || (m > 0.0f && m < 4.0f)) {
        int e;
        frexpf(m, &e);      // m = fr * 2^e, fr in [0.5, 1)
        int shift = e - 8;  // new stored max in [2^7, 2^8)
        while (shift < 0 && ldexpf(sc[c], shift) < sc[36 + c]) ++shift;
        if (shift != 0) {
            f = ldexpf(1.0f, -shift);
            sc[c] = ldexpf(sc[c], shift);
            sc[18 + c] = ldexpf(sc[18 + c], -shift);
        }
    }
    factor[c] = f;
}
__global__ void __launch_bounds__(256) k_fp16_rescale(ArrSet s, const float* factor) {
    for (int c = 0; c < 18; ++c) {
        const float f = factor[c];
        if (f == 1.0f || !s.p[c]) continue;
        __half* a = reinterpret_cast<__half*>(const_cast<uint16_t*>(s.p[c]));
        const long stride = (long)gridDim.x * blockDim.x;
        for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < s.n[c]; i += stride)
            a[i] = __float2half_rn(__half2float(a[i]) * f);
    }
}

}  // namespace fdtdk

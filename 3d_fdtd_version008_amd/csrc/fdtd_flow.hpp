// fdtd_flow.hpp - the dependency-ordered step (kernel mode 2): the H and E
// tiles of one step in a single launch, ordered so that each E tile reads
// the H^{n+1/2} and E^n its H neighbours have just moved through the XCD's
// L2, instead of a second pass over the whole grid.
//
// Why: the two-kernel step (k_update_H, then k_update_E) reads E and H twice
// per step (18 s + m bytes per cell against the 12 s + m a step must move).
// Fused kernels that carry H and E from plane to plane in registers or LDS
// lose more to register pressure than they save in bytes.  Here every work-
// group still runs one plain tile of the two-kernel step, with those
// kernels' per-cell code and register footprint; only the order of the tiles
// and a wait on four flags are new.
//
// Order.  Each XCD owns a contiguous run of x-planes (a queue).  The launch
// is persistent: a work-group reads the XCC_ID hardware register and takes
// tickets of that XCD's queue, in order, from a counter in the XCD's own L2.
// So a tile's flags and data pass only between work-groups of one XCD, every
// ticket is taken exactly once, and a ticket a work-group waits on was taken
// earlier by a running work-group that only waits on lower tickets: the
// launch cannot deadlock, whatever order the dispatcher starts work-groups
// in.  A queue walks bands of j-tiles (as the two-kernel order does, so a
// band-plane's halo stays in the L2); within a band it walks slots, one per
// plane: slot s holds the H tiles of plane p0 + s, then the E tiles of plane
// p0 + s - lag.  k_step_flow_edge checks after the launch that every tile
// ran once this step.
//
// Dependencies.  E(p, jt, kt) reads H^{n+1/2} at (p, jt, kt), at k - 1 (tile
// kt - 1), at j - 1 (tile jt - 1) and at plane p - 1; it overwrites E^n(p),
// which exactly the same four H tiles read (H uses E at i + 1, j + 1, k + 1).
// So an E tile waits until the four H tiles have stored their flag for this
// step, and nothing else is ordered.  Every one of them has a lower ticket in
// the same queue, except H(p0 - 1) for a queue's first plane p0: the E tiles
// of those planes are left to k_step_flow_edge, launched after the step
// (H(p0 + P - 1), the last H of a queue, reads E^n(p0 of the next queue),
// which therefore stays E^n for the whole launch).
//
// Visibility.  A flag is stored after the tile's own stores have completed
// (each wavefront waits for its vector memory counter to reach zero, then the
// work-group barrier, then one store of the flag): they are then in the XCD's
// L2, which every work-group of the queue shares.  The waiting work-group
// polls the flags, meets at a barrier, and reads the H values of this launch
// through loads that bypass its CU's L1 (e_cell with COH): a line the CU
// cached for a neighbouring tile before the producer stored into it would
// otherwise be read stale, and in an ordinary dispatch an L1 invalidate at
// work-group scope (buffer_inv sc0) does not reach it (measured: a consumer
// that had read the line before the producer's store read the old value in
// 9-12 of 12 runs of a probe, with the invalidate; 0 of 12 with L1-bypassing
// loads, at 3% of the probe's time).  Every other value an E or H tile reads
// was not stored in this launch before it is read (E^n, psi, gain state:
// their owner tile is the only writer), so plain loads see it.  The agent-
// scope forms (an L2 write-back per release, an L2 invalidate per acquire)
// cost 5-40x the whole probe.
//
// Waits are bounded: after FLOW_SPIN_MAX polls (a legitimate wait ends within
// a few) a work-group counts a time-out and goes on, and once one wait has
// timed out every later one gives up at once, so a launch always drains
// quickly; the host reads the counter after
// the call and reports the call as failed.  Flags hold a step epoch (read
// from device memory, advanced by k_flow_advance after the step), so they
// are never cleared and a replayed graph sees each step's own epoch.
#pragma once
#include "fdtd_kernels.hpp"

namespace fdtdk {

constexpr unsigned FLOW_SPIN_MAX = 1u << 16;
// ctl: [0] epoch, [1] error counts, then one ticket counter per XCD, each in
// its own 128-byte line
constexpr unsigned FLOW_CTR_WORDS = 32;
constexpr unsigned FLOW_CTL_WORDS = FLOW_CTR_WORDS * 9;

struct FlowGeom {
    unsigned* hflags;        // per (plane, tile of the plane): epoch of its finished H tile
    unsigned* eflags;        // the same for the E tile (checked after the launch)
    unsigned* ctl;           // [0] this step's epoch, [1] tiles not run << 16 | wait time-outs
    unsigned pbase, prem;    // queue q: planes [q*pbase + min(q, prem), ... + pbase + (q < prem))
    unsigned lag;            // E tiles of plane p share the slot of the H tiles of plane p + lag
    unsigned tpp;            // tiles per plane (tk * tj)
    unsigned nbf;            // full bands
    unsigned tfull;          // tickets of the full bands of one queue
    unsigned tlast;          // tickets of the last, partial band (0: none)
    FastDiv band_t;          // tickets per full band: 2 * T * (pmax + lag)
    FastDiv slot_t;          // tickets per slot of a full band: 2 * T (T = bj * tk tiles)
    FastDiv slot_l;          // tickets per slot of the partial band: 2 * Tl
    FastDiv tppd;            // tpp, as a divisor (edge launch)
};

__device__ __forceinline__ unsigned flow_plane0(const FlowGeom& g, unsigned q) {
    return q * g.pbase + (q < g.prem ? q : g.prem);
}

template <int KW>
__device__ __forceinline__ Tile tile_at(unsigned plane, unsigned jt, unsigned kt) {
    constexpr int G = 64 / KW;
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = (int)(threadIdx.x & 63);
    Tile t;
    t.kw = (int)kt * 64 + (w % G) * KW;
    t.jw = (int)jt * TILE_ROWS + (w / G) * G;
    t.k = t.kw + l % KW;
    t.j = t.jw + l / KW;
    t.i = (int)plane;
    return t;
}

// H of one tile; with MON also the monitor's field maxima of what it reads
// (the body of k_update_H_mon; `slot` picks the record slot).
template <typename T, typename R, int KW, typename IX, bool MON>
__device__ __forceinline__ void h_tile(const StepArgs<T, R>& a, const Tabs<R>& tb, const Tile& t,
                                       unsigned long long* fmax, unsigned slot) {
    const int i = t.i, j = t.j, k = t.k;
    if constexpr (!MON) {
        (void)fmax; (void)slot;
        if (k < a.nz - 1 && j < a.ny - 1 && i < a.nx - 1)
            (void)h_compute<T, R, true, IX>(a, tb, i, j, k, h_gather<T, R, KW, true, IX>(a, tb, i, j, k, t.kw, t.jw), true);
    } else {
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
            unsigned long long* q = fmax + (size_t)(slot % MON_SLOTS) * MON_SLOT_WORDS;
            if (mE) atomicMax(q, mE);
            if (mH) atomicMax(q + 1, mH);
        }
    }
}

// One ticket of queue q: an H tile, an E tile, or nothing (the empty slots
// at a band's start and end).  Every thread of the work-group calls it with
// the same ticket; a tile's barriers are inside.
template <typename T, typename R, int KW, typename IX, bool MON>
__device__ __forceinline__ void flow_ticket(const StepArgs<T, R>& a, const Tabs<R>& tb, const FlowGeom& g,
                                            unsigned q, unsigned tkt, unsigned ep, unsigned long long* fmax) {
    // ticket -> (band, slot, kind, tile of the band-plane)
    unsigned band, slot, r, tb_;
    if (tkt < g.tfull) {
        band = fdiv(tkt, g.band_t);
        const unsigned r0 = tkt - band * g.band_t.d;
        slot = fdiv(r0, g.slot_t);
        r = r0 - slot * g.slot_t.d;
        tb_ = g.slot_t.d >> 1;
    } else {
        const unsigned r0 = tkt - g.tfull;
        band = g.nbf;
        slot = fdiv(r0, g.slot_l);
        r = r0 - slot * g.slot_l.d;
        tb_ = g.slot_l.d >> 1;
    }
    const bool etile = r >= tb_;
    const unsigned rr = etile ? r - tb_ : r;
    const int np = (int)(g.pbase + (q < g.prem ? 1u : 0u));
    const int s = etile ? (int)slot - (int)g.lag : (int)slot;
    if (etile ? (s < 1 || s >= np) : (s >= np)) return;  // uniform over the work-group
    const unsigned plane = flow_plane0(g, q) + (unsigned)s;
    const unsigned jr = fdiv(rr, a.tk);
    const unsigned kt = rr - jr * a.tk.d;
    const unsigned jt = band * (unsigned)a.bj + jr;
    const unsigned f0 = plane * g.tpp + jt * a.tk.d + kt;
    const Tile t = tile_at<KW>(plane, jt, kt);
    if (!etile) {
        h_tile<T, R, KW, IX, MON>(a, tb, t, fmax, f0 * 4u + (threadIdx.x >> 6));
        // release: this wavefront's stores have reached the L2 ...
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // ... and so have every other wavefront's
        if (threadIdx.x == 0) __hip_atomic_store(g.hflags + f0, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        if (threadIdx.x < 4) {
            const unsigned d = threadIdx.x;
            const bool has = d == 0 || (d == 1 && kt > 0) || (d == 2 && jt > 0) || d == 3;
            if (has) {
                const unsigned f = d == 0 ? f0 : d == 1 ? f0 - 1u : d == 2 ? f0 - a.tk.d : f0 - g.tpp;
                unsigned n = 0;
                while (__hip_atomic_load(g.hflags + f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ep) {
                    // give up after FLOW_SPIN_MAX polls, or at once when a
                    // wait has already timed out (the call is failed anyway)
                    if (++n >= FLOW_SPIN_MAX ||
                        ((n & 255u) == 0 && (__hip_atomic_load(g.ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0xffffu))) {
                        atomicAdd(g.ctl + 1, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) g.eflags[f0] = ep;  // read only after the launch
        // H of this launch through L1-bypassing loads (the header explains why)
        if (t.k < a.nz && t.j < a.ny && t.i < a.nx)
            e_cell<T, R, KW, true, true, IX, true>(a, tb, t.i, t.j, t.k, t.kw, t.jw);
    }
}

// Persistent: each work-group takes tickets of its XCD's queue, in order,
// from that XCD's counter (an atomic in the XCD's own L2, touched by no other
// XCD), the next one while the current tile runs, until the queue is empty.
template <typename T, typename R, int KW, typename IX, bool MON>
__global__ void __launch_bounds__(256) k_step_flow(StepArgs<T, R> a, FlowGeom g, unsigned long long* fmax) {
    FDTD_SCALES_IN_REGISTERS(a)
    const Tabs<R> tb = global_tabs(a);
    unsigned xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    const unsigned q = xcc & 7u;
    unsigned* ctr = g.ctl + FLOW_CTR_WORDS * (q + 1u);
    const unsigned total = g.tfull + g.tlast, ep = g.ctl[0];
    __shared__ unsigned s_tk[2];
    if (threadIdx.x == 0) s_tk[0] = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    unsigned tkt = s_tk[0], par = 1;
    while (tkt < total) {
        if (threadIdx.x == 0) s_tk[par] = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        flow_ticket<T, R, KW, IX, MON>(a, tb, g, q, tkt, ep, fmax);
        __syncthreads();
        tkt = s_tk[par];
        par ^= 1u;
    }
}

// The E tiles of each queue's first plane, after the flow launch (their
// H(p0 - 1) belongs to the previous queue).  Work-group (queue, tile) also
// checks that the flow launch ran every H and E tile of that tile column in
// the queue's planes, and counts those it did not.
template <typename T, typename R, int KW, typename IX>
__global__ void __launch_bounds__(256) k_step_flow_edge(StepArgs<T, R> a, FlowGeom g) {
    FDTD_SCALES_IN_REGISTERS(a)
    const Tabs<R> tb = global_tabs(a);
    const unsigned b = blockIdx.x;
    const unsigned qi = fdiv(b, g.tppd);
    const unsigned rr = b - qi * g.tppd.d;
    const unsigned jt = fdiv(rr, a.tk);
    const unsigned kt = rr - jt * a.tk.d;
    const unsigned p0 = flow_plane0(g, qi), np = g.pbase + (qi < g.prem ? 1u : 0u), ep = g.ctl[0];
    unsigned miss = 0;
    for (unsigned s = threadIdx.x; s < np; s += blockDim.x) {
        const unsigned f = (p0 + s) * g.tpp + rr;
        miss += (g.hflags[f] != ep) + (s > 0 && g.eflags[f] != ep);
    }
    if (miss) atomicAdd(g.ctl + 1, miss << 16);
    const Tile t = tile_at<KW>(p0, jt, kt);
    if (t.k < a.nz && t.j < a.ny && t.i < a.nx) e_cell<T, R, KW, true, true, IX>(a, tb, t.i, t.j, t.k, t.kw, t.jw);
}

// Advances the step epoch and clears the XCD ticket counters after the
// step's flow and edge launches.
__global__ void k_flow_advance(unsigned* ctl) {
    if (threadIdx.x == 0) ctl[0] = ctl[0] + 1u;
    if (threadIdx.x < 8) ctl[FLOW_CTR_WORDS * (threadIdx.x + 1)] = 0u;
}

}  // namespace fdtdk

// fdtd_device.hip - C ABI of the gfx950 Yee-step engine (include/fdtd_hip.h).
//
// Host side: validates and repacks the reference-layout host arrays into the
// device layout (material-id + exact LUT, CPML 1D profiles in the compute
// precision, psi on the boundary slabs only, gain state in its bounding box),
// owns one stream per handle, launches the fused H and E kernels in the
// reference's step order, and applies the reference's stability policy to the
// device reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fdtd_hip.h"
#include "fdtd_kernels.hpp"
#include "fdtd_fused.hpp"
#include "fdtd_flow.hpp"
#include "fdtd_bake.hpp"

using fdtdk::MonOut;
using fdtdk::SourceBox;
using fdtdk::StepArgs;
using fdtdk::TLSCoef;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(FDTD_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

struct Probe {
    long id;
    int comp, save_every;
    long cap, count;
    double* d_buf;
};

struct Dft {
    int comp;
    int box[6];
    int nf;
    double freqs[8];
    long ncell;
    double* d_acc;
};

struct Key12 {
    double v[12];
    bool operator==(const Key12& o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
};
struct Key12Hash {
    size_t operator()(const Key12& k) const {
        uint64_t h = 1469598103934665603ull;
        const uint64_t* w = reinterpret_cast<const uint64_t*>(k.v);
        for (int q = 0; q < 12; ++q) {
            h ^= w[q] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        }
        return (size_t)h;
    }
};

// Device memory cache (fdtd_set_device_pool).  hipFree waits for the whole
// device, so a serving loop that frees finished requests while others run
// stalls every stream at each retirement, and hipMalloc is slow at each
// admission.  With the cache on, a handle's blocks return here when it is
// destroyed (after its own stream has drained) and a later allocation of the
// same device reuses a cached block at most 1/8 larger than it asks for;
// cached bytes above the cap are freed at once.  Blocks are keyed by
// (device, size); a handle still reports the bytes it asked for.
struct DevPool {
    std::mutex mu;
    size_t cap = 0;     // bytes the cache may hold (0: off, every free is a hipFree)
    size_t cached = 0;
    std::multimap<std::pair<int, size_t>, void*> blocks;
};
DevPool& dev_pool() {
    static DevPool p;
    return p;
}
// Allocates at least b bytes on `dev` (the current device); *got = the block's size.
hipError_t pool_alloc(int dev, size_t b, void** out, size_t* got) {
    DevPool& P = dev_pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        if (P.cap) {
            auto it = P.blocks.lower_bound({dev, b});
            if (it != P.blocks.end() && it->first.first == dev && it->first.second <= b + b / 8) {
                *out = it->second;
                *got = it->first.second;
                P.cached -= it->first.second;
                P.blocks.erase(it);
                return hipSuccess;
            }
        }
    }
    *got = b;
    return hipMalloc(out, b);
}
// Pinned host blocks share the cache under negative keys: mapped and
// coherent (records and result copies the device writes), or plain (upload
// staging); their sizes are powers of two, so blocks match across handles.
constexpr int PIN_MAPPED = -1, PIN_PLAIN = -2;
// Evicts cached blocks (largest key first) until at most `keep` bytes remain.
void pool_trim(DevPool& P, size_t keep) {
    while (P.cached > keep && !P.blocks.empty()) {
        auto it = std::prev(P.blocks.end());
        if (it->first.first < 0) {
            (void)hipHostFree(it->second);
        } else {
            int cur = 0;
            (void)hipGetDevice(&cur);
            if (cur != it->first.first) (void)hipSetDevice(it->first.first);
            (void)hipFree(it->second);
            if (cur != it->first.first) (void)hipSetDevice(cur);
        }
        P.cached -= it->first.second;
        P.blocks.erase(it);
    }
}
// Pinned host memory of at least b bytes (kind PIN_MAPPED or PIN_PLAIN); *got = its size.
hipError_t pool_host_alloc(int kind, size_t b, void** out, size_t* got) {
    size_t r = 4096;
    while (r < b) r <<= 1;
    *got = r;
    DevPool& P = dev_pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        if (P.cap) {
            auto it = P.blocks.find({kind, r});
            if (it != P.blocks.end()) {
                *out = it->second;
                P.cached -= r;
                P.blocks.erase(it);
                return hipSuccess;
            }
        }
    }
    return hipHostMalloc(out, r, kind == PIN_MAPPED ? (hipHostMallocMapped | hipHostMallocCoherent) : hipHostMallocDefault);
}
// Returns a pinned block no copy or kernel uses any more.
void pool_host_free(int kind, void* p, size_t size) {
    DevPool& P = dev_pool();
    std::lock_guard<std::mutex> g(P.mu);
    if (P.cap && size <= P.cap) {
        P.blocks.insert({{kind, size}, p});
        P.cached += size;
        pool_trim(P, P.cap);
        return;
    }
    (void)hipHostFree(p);
}
// Returns a block no stream uses any more.
void pool_free(int dev, void* p, size_t size) {
    DevPool& P = dev_pool();
    std::lock_guard<std::mutex> g(P.mu);
    if (P.cap && size <= P.cap) {
        P.blocks.insert({{dev, size}, p});
        P.cached += size;
        pool_trim(P, P.cap);
        return;
    }
    (void)hipFree(p);
}

}  // namespace

struct fdtd_sim {
    int nx = 0, ny = 0, nz = 0, np = 0, prec = 0, bc_type = 0;
    long N = 0;
    double dt = 0;
    hipStream_t stream = nullptr;
    size_t esz = 8;     // bytes per stored field value
    size_t rsz = 8;     // bytes per compute value
    std::vector<void*> allocs;
    std::vector<size_t> alloc_sz;  // size of each block in allocs (the cache may hand out a larger one)
    // stepping calls so far (batch steps included): staged results are valid
    // while this equals res_gen
    unsigned long gen = 0, res_gen = ~0ul;
    // pinned, coherent host copy of the probe records and DFT accumulators
    // (fdtd_stage_results): h_res / d_res, res_words doubles; where each
    // probe's and DFT's copy starts
    double *h_res = nullptr, *d_res = nullptr;
    size_t res_words = 0;
    std::vector<size_t> res_probe_off, res_dft_off;
    std::vector<Key12> h_tuples;  // the coefficient table in FP64, by material id
    // steps enqueued by fdtd_step_async whose monitor records are not read yet
    bool pending = false;
    std::vector<long> pending_mon;
    size_t bytes = 0;
    // Fields and psi_H are double-buffered: the fused step reads buffer `cur`
    // and writes buffer cur^1; the two-kernel step updates buffer `cur` in
    // place.  psi_E lives in buffer 0 only (owner-only, in place).
    void* F2[2][6] = {};
    void* psi2[2][12] = {};
    void* F[6] = {};    // = F2[cur]
    void* psi[12] = {}; // = psi2[cur] (psi_E always psi2[0])
    int cur = 0;
    int mode = 0;       // 0: H and E kernels (default), 1: fused step kernel, 2: dependency-ordered step
    // mode 2 (k_step_flow): per (plane, tile) flags of finished H tiles, then
    // of finished E tiles, and {step epoch, tiles not run << 16 | wait time-outs}
    unsigned* d_flowflags = nullptr;
    unsigned* d_flowctl = nullptr;
    int flow_lag = 0;   // planes between a slot's H tiles and its E tiles (0: 2)
    unsigned flow_wg = 0;  // work-groups of the persistent flow launch (0: not yet sized)
    int ichunk = 0;     // x-planes per work-group of the fused kernel (0: by grid size)
    int fz_cw = 0;      // fused kernel: cells per lane per row chunk (0: by grid size)
    int fz_jw = 0;      // fused kernel: rows per work-group (0: by grid size)
    int kwid_h = 0;     // k-width of a wavefront's cells in the H kernel (16, 32, 64; 0: by precision)
    int kwid_e = 0;     // the same for the E kernel
    unsigned band_tiles = 0;  // H/E tile order: j-tiles per band (0: by plane size)
    bool fits32 = false; // every array a step kernel indexes by cell is below 4 GiB
    bool off32 = false; // the H/E kernels use 32-bit byte offsets (fits32, FP16 storage)
    bool force64 = false;  // use the 64-bit offset kernels anyway
    // The E kernel walks each XCD's range of tiles from the end, so it starts
    // on the tiles the H kernel finished last and the next H kernel starts on
    // the tiles E finished last: both begin on lines still held in the
    // memory-side cache (C2 FP32 step 0.341 -> 0.331 ms).  Off: forward order.
    bool erev = true;
    long psi_len[12] = {};
    void* inv[3] = {};
    void* cp[3] = {};
    void* matid = nullptr;
    int mid16 = 0, nmat = 0;
    void* lut = nullptr;
    bool lossless = false, h_uniform = false;
    // gain
    bool gain = false;
    int g0[3] = {0, 0, 0}, gn[3] = {0, 0, 0};
    long gbox = 0, ngain = 0;
    int8_t* greg = nullptr;
    double *gPz = nullptr, *gPp = nullptr, *gNu = nullptr, *gNg = nullptr;
    TLSCoef* d_tls = nullptr;
    // TLS diagnostics recorder: every tls_every steps (0: off), up to
    // tls_cap records of 5 values at d_tlsrec; FP64 spacings for dV
    int tls_every = 0;
    long tls_cap = 0, tls_count = 0, tls_gmon = -1;
    double *d_tlsrec = nullptr, *d_tlspart = nullptr, *d_dxyz = nullptr;
    std::vector<TLSCoef> h_tls;
    // sources
    SourceBox* d_src = nullptr;
    int nsrc = 0;
    double* d_srcv = nullptr;
    long srcv_cap = 0;
    int device = 0;                // the device the handle was created on
    bool mon_fused_h = true;       // monitor field maxima in the next H launch (else a separate pass)
    // Step graphs (fdtd_set_step_graphs): a call's launches recorded once in
    // a HIP graph and replayed for later calls of the same shape.  The
    // per-step values that differ between calls (DFT phase factors, probe
    // record slots) then come from a device table written before each
    // launch; tab_step is the current step's row of it while launching.
    bool graph_ticks = false;
    const double* tab_step = nullptr;
    std::vector<int> tab_dft_off;  // offsets of each DFT's 2*nf phases in a row
    int tab_probe_off = 0, tab_stride = 0;
    double* d_tab = nullptr;
    long tab_cap = 0;
    // pinned host staging of a call's uploads (source values, then the step
    // table): a copy from pinned memory is queued without the host thread
    // waiting on the copy engine (a pageable copy is staged synchronously)
    double* h_pin = nullptr;
    size_t pin_words = 0;
    // pinned, coherent host copy of a call's monitor records and the FP16
    // non-finite flags, written by a copy kernel at the end of the call
    // (h_rec: host address, d_rec: its device address)
    unsigned long long* h_rec = nullptr;
    unsigned long long* d_rec = nullptr;
    size_t rec_words = 0;
    bool rec_host = false;  // the call in flight copies its records to h_rec
    struct StepGraph {
        std::vector<long> key;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs;
    double* srcv_cur = nullptr;    // source values of the call in flight
    MonOut* mon_cur = nullptr;     // monitor records of the call in flight
    // probes / dft
    std::vector<Probe> probes;
    std::vector<Dft> dfts;
    // monitor
    int mon_interval = 10;
    MonOut* d_mon = nullptr;
    int mon_cap = 0;
    MonOut* d_monpart = nullptr;  // per-block partial records of the monitor reduction
    // Field maxima of a monitor step taken by the next step's H launch
    // (k_update_H_mon): their slots, the record they complete, and the
    // number of gain partial records in d_monpart for it.
    unsigned long long* d_fmax = nullptr;
    MonOut* mon_defer = nullptr;
    int mon_defer_nb = 0;
    std::vector<fdtd_check_record> checks;
    double prev_max_E = 0.0;
    size_t inst_count = 0;
    bool stable = true;
    // FP16 scales (value = stored * scale), device-resident so kernels read
    // the current ones: [0..5] fields, [6..17] psi, [18..35] inverses.
    float* d_sc = nullptr;
    float* d_factor = nullptr;
    unsigned* d_maxbits = nullptr;
    int fp16_interval = 10;
    struct fdtd_batch* batch = nullptr;  // the batch this domain belongs to

    ~fdtd_sim() {
        if (stream) hipStreamSynchronize(stream);
        for (auto& g : graphs) hipGraphExecDestroy(g.exec);
        if (h_pin) pool_host_free(PIN_PLAIN, h_pin, pin_words * sizeof(double));
        if (h_rec) pool_host_free(PIN_MAPPED, h_rec, rec_words * sizeof(unsigned long long));
        if (h_res) pool_host_free(PIN_MAPPED, h_res, res_words * sizeof(double));
        for (size_t q = 0; q < allocs.size(); ++q) pool_free(device, allocs[q], alloc_sz[q]);
        if (stream) hipStreamDestroy(stream);
    }

    int alloc(void** p, size_t b) {
        if (b == 0) b = 8;
        size_t got = b;
        hipError_t e = pool_alloc(device, b, p, &got);
        if (e != hipSuccess) return fail(FDTD_ENOMEM, "hipMalloc(" + std::to_string(b) + " B) failed: " + hipGetErrorString(e));
        allocs.push_back(*p);
        alloc_sz.push_back(got);
        bytes += b;
        return 0;
    }
    template <typename X>
    int upload(X** dst, const X* src, size_t n) {
        int rc = alloc(reinterpret_cast<void**>(dst), n * sizeof(X));
        if (rc) return rc;
        HIPCHK(hipMemcpy(*dst, src, n * sizeof(X), hipMemcpyHostToDevice));
        return 0;
    }
};

namespace {

template <typename R>
std::vector<R> to_r(const double* p, size_t n) {
    std::vector<R> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (R)p[i];
    return v;
}

// j-tiles per band of the H/E tile order (StepArgs::bj)
unsigned tile_band(const fdtd_sim* s, unsigned tj) {
    unsigned bj = s->band_tiles;
    if (bj == 0) bj = 6.0 * s->ny * s->nz * (double)s->esz > 1.0 * (1 << 20) ? 16u : tj;
    return std::min(std::max(bj, 1u), tj);
}

template <typename T, typename R>
StepArgs<T, R> make_args(const fdtd_sim* s) {
    StepArgs<T, R> a{};
    a.nx = s->nx; a.ny = s->ny; a.nz = s->nz; a.np = s->np;
    a.sI = (long)s->ny * s->nz;
    a.sJ = s->nz;
    a.Ex = (T*)s->F[0]; a.Ey = (T*)s->F[1]; a.Ez = (T*)s->F[2];
    a.Hx = (T*)s->F[3]; a.Hy = (T*)s->F[4]; a.Hz = (T*)s->F[5];
    a.Exw = a.Ex; a.Eyw = a.Ey; a.Ezw = a.Ez; a.Hxw = a.Hx; a.Hyw = a.Hy; a.Hzw = a.Hz;
    a.pec = s->bc_type == 0;
    a.cpml = s->bc_type == 1 && s->np > 0;
    a.cpx = (const R*)s->cp[0]; a.cpy = (const R*)s->cp[1]; a.cpz = (const R*)s->cp[2];
    for (int q = 0; q < 12; ++q) a.psi[q] = a.psiw[q] = (T*)s->psi[q];
    a.sc = s->d_sc;
    a.matid = s->matid;
    a.mid16 = s->mid16;
    a.lut = (const R*)s->lut;
    a.nmat = s->nmat;
    a.lossless = s->lossless;
    a.h_uniform = s->h_uniform;
    a.src = s->d_src;
    a.nsrc = s->nsrc;
    a.srcv = s->d_srcv;
    a.gain = s->gain;
    a.g0x = s->g0[0]; a.g0y = s->g0[1]; a.g0z = s->g0[2];
    a.gnx = s->gn[0]; a.gny = s->gn[1]; a.gnz = s->gn[2];
    a.greg = s->greg;
    a.gPz = s->gPz; a.gPp = s->gPp; a.gNu = s->gNu; a.gNg = s->gNg;
    a.tls = s->d_tls;
    a.ntls = (int)s->h_tls.size();
    const unsigned tk = (unsigned)((s->nz + 63) / 64), tj = (unsigned)((s->ny + fdtdk::TILE_ROWS - 1) / fdtdk::TILE_ROWS);
    a.tk = fdtdk::make_fastdiv(tk);
    a.tj = fdtdk::make_fastdiv(tj);
    a.rev = 0;
    // bands of j-tiles: where a plane of the six field arrays outgrows
    // 1 MiB, sweep bands of 64 rows (16 tiles) plane by plane so the next
    // plane's reads meet the lines this band-plane loaded in the L2.
    // Measured (HBM counters, 512x512x384 FP32): whole-plane order reads
    // 1.31x (H) and 1.27x (E) the algorithmic bytes, bands of 16 tiles 1.00x
    const unsigned bj = tile_band(s, tj);
    const unsigned nfull = tj / bj, bjl = tj - nfull * bj;
    a.bj = (int)bj;
    a.tbp = fdtdk::make_fastdiv(bj * tk);
    a.tband = fdtdk::make_fastdiv(bj * tk * (unsigned)s->nx);
    a.tbpl = fdtdk::make_fastdiv(bjl ? bjl * tk : 1u);
    a.full_tiles = nfull * bj * tk * (unsigned)s->nx;
    return a;
}

// Material ids and distinct tuples (first-occurrence order) from the 12
// host MaterialGrids arrays.
int ids_from_arrays(fdtd_sim* s, const fdtd_material_desc* m, std::vector<Key12>& tuples) {
    const double* arrs[12] = {m->aEx, m->bEx, m->aEy, m->bEy, m->aEz, m->bEz, m->aHx, m->bHx, m->aHy, m->bHy, m->aHz, m->bHz};
    for (int q = 0; q < 12; ++q)
        if (!arrs[q]) return fail(FDTD_EINVAL, "material array missing");
    std::unordered_map<Key12, int, Key12Hash> ids;
    std::vector<uint16_t> mid((size_t)s->N);
    Key12 last{};
    int last_id = -1;
    for (long id = 0; id < s->N; ++id) {
        Key12 k;
        for (int q = 0; q < 12; ++q) k.v[q] = arrs[q][id];
        int t;
        if (last_id >= 0 && k == last) {
            t = last_id;
        } else {
            auto it = ids.find(k);
            if (it == ids.end()) {
                t = (int)tuples.size();
                if (t >= 65536) return fail(FDTD_EINVAL, "more than 65536 distinct material coefficient tuples");
                ids.emplace(k, t);
                tuples.push_back(k);
            } else {
                t = it->second;
            }
            last = k;
            last_id = t;
        }
        mid[(size_t)id] = (uint16_t)t;
    }
    s->nmat = (int)tuples.size();
    s->mid16 = s->nmat > 256;
    if (s->mid16) {
        uint16_t* d;
        if (int rc = s->upload(&d, mid.data(), mid.size())) return rc;
        s->matid = d;
    } else {
        std::vector<uint8_t> m8(mid.size());
        for (size_t i = 0; i < mid.size(); ++i) m8[i] = (uint8_t)mid[i];
        uint8_t* d;
        if (int rc = s->upload(&d, m8.data(), m8.size())) return rc;
        s->matid = d;
    }
    return 0;
}

// Device memory that lives for one setup call, used by kernels on `stream`.
struct TempBuf {
    int dev;
    hipStream_t stream;
    std::vector<void*> p;
    std::vector<size_t> sz;
    TempBuf(int d, hipStream_t st) : dev(d), stream(st) {}
    ~TempBuf() {
        if (!p.empty()) (void)hipStreamSynchronize(stream);
        for (size_t q = 0; q < p.size(); ++q) pool_free(dev, p[q], sz[q]);
    }
    template <typename X>
    int get(X** out, size_t n) {
        void* q = nullptr;
        size_t got = 0;
        if (pool_alloc(dev, std::max<size_t>(n * sizeof(X), 8), &q, &got) != hipSuccess)
            return fail(FDTD_ENOMEM, "hipMalloc of a bake buffer failed");
        p.push_back(q);
        sz.push_back(got);
        *out = (X*)q;
        return 0;
    }
    template <typename X>
    int put(X** out, const X* src, size_t n) {
        if (int rc = get(out, n)) return rc;
        if (n && hipMemcpy(*out, src, n * sizeof(X), hipMemcpyHostToDevice) != hipSuccess)
            return fail(FDTD_EDEVICE, "bake upload failed");
        return 0;
    }
};

// StructureScene::bake (src/structure_material.hpp:211-288) on the device:
// per-cell material ids and the distinct coefficient tuples (the pure tuples
// of the materials some cell lies wholly in, in material order, then those of
// interface cells in cell order).
template <typename ID>
int finish_ids(fdtd_sim* s, TempBuf& tmp, const uint16_t* ids16, const std::vector<uint16_t>& remap,
               const std::vector<long long>& icell, const std::vector<uint16_t>& iid) {
    void* d;
    if (int rc = s->alloc(&d, (size_t)s->N * sizeof(ID))) return rc;
    uint16_t* dremap;
    if (int rc = tmp.put(&dremap, remap.data(), remap.size())) return rc;
    hipLaunchKernelGGL(fdtdk::k_ids_remap<ID>, dim3(2048), dim3(256), 0, s->stream, ids16, (const uint16_t*)dremap, (ID*)d,
                       (long long)s->N);
    HIPCHK(hipGetLastError());
    if (!icell.empty()) {
        long long* dc;
        uint16_t* di;
        if (int rc = tmp.put(&dc, icell.data(), icell.size())) return rc;
        if (int rc = tmp.put(&di, iid.data(), iid.size())) return rc;
        hipLaunchKernelGGL(fdtdk::k_bake_scatter<ID>, dim3((unsigned)((icell.size() + 255) / 256)), dim3(256), 0, s->stream,
                           (ID*)d, (const long long*)dc, (const uint16_t*)di, (long long)icell.size());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->matid = d;
    return 0;
}

int bake_on_device(fdtd_sim* s, const fdtd_bake_desc* bd, std::vector<Key12>& tuples) {
    if (!bd->x_bounds || !bd->y_bounds || !bd->z_bounds || !bd->dx || !bd->dy || !bd->dz || bd->n_shapes < 0 ||
        bd->n_shapes > 65533 || (bd->n_shapes > 0 && !bd->shapes))
        return fail(FDTD_EINVAL, "incomplete bake description (or more than 65533 structures)");
    int ns = (int)std::round(std::cbrt((double)bd->subpixel_samples));
    if (ns < 1) ns = 1;
    const int nsh = bd->n_shapes;
    // material properties, row 0 = background, row q + 1 = structure q
    std::vector<double> props((size_t)4 * (nsh + 1));
    props[0] = bd->bg_eps_r; props[1] = bd->bg_mu_r; props[2] = bd->bg_sigma; props[3] = bd->bg_sigma_m;
    std::vector<fdtdk::BakeShape> sh((size_t)nsh);
    for (int q = 0; q < nsh; ++q) {
        const fdtd_bake_shape& b = bd->shapes[q];
        if (b.type < 0 || b.type > 2) return fail(FDTD_EINVAL, "unknown bake shape type");
        sh[q].type = b.type;
        for (int r = 0; r < 6; ++r) sh[q].p[r] = b.p[r];
        double* pr = &props[(size_t)4 * (q + 1)];
        pr[0] = b.eps_r; pr[1] = b.mu_r; pr[2] = b.sigma; pr[3] = b.sigma_m;
    }
    TempBuf tmp(s->device, s->stream);
    fdtdk::BakeArgs a{};
    a.nx = s->nx; a.ny = s->ny; a.nz = s->nz; a.ns = ns; a.nsh = nsh;
    a.dt = bd->dt; a.eps0 = bd->eps0; a.mu0 = bd->mu0;
    int rc = 0;
    double *xb, *yb, *zb, *dx, *dy, *dz, *dprops;
    fdtdk::BakeShape* dsh;
    uint16_t* ids16;
    uint8_t* dused;
    unsigned long long* dcount;
    if ((rc = tmp.put(&xb, bd->x_bounds, (size_t)s->nx + 1)) || (rc = tmp.put(&yb, bd->y_bounds, (size_t)s->ny + 1)) ||
        (rc = tmp.put(&zb, bd->z_bounds, (size_t)s->nz + 1)) || (rc = tmp.put(&dx, bd->dx, (size_t)s->nx)) ||
        (rc = tmp.put(&dy, bd->dy, (size_t)s->ny)) || (rc = tmp.put(&dz, bd->dz, (size_t)s->nz)) ||
        (rc = tmp.put(&dprops, props.data(), props.size())) || (rc = tmp.put(&dsh, sh.data(), sh.size())) ||
        (rc = tmp.get(&ids16, (size_t)s->N)) || (rc = tmp.get(&dused, (size_t)nsh + 1)) || (rc = tmp.get(&dcount, 1)))
        return rc;
    a.xb = xb; a.yb = yb; a.zb = zb; a.dx = dx; a.dy = dy; a.dz = dz;
    a.props = dprops; a.sh = dsh;
    a.ids = ids16; a.used = dused; a.count = dcount;
    // interface list: sized for a sixteenth of the cells, grown once if short
    unsigned long long cap = (unsigned long long)std::max<long>(1L << 20, s->N / 16), count = 0;
    std::vector<long long> hc;
    std::vector<double> ht;
    std::vector<uint8_t> used((size_t)nsh + 1);
    bool done = false;
    for (int pass = 0; pass < 2 && !done; ++pass) {
        TempBuf lst(s->device, s->stream);
        long long* icell;
        double* itup;
        if ((rc = lst.get(&icell, (size_t)cap)) || (rc = lst.get(&itup, (size_t)cap * 12))) return rc;
        a.icell = icell; a.itup = itup; a.cap = cap;
        HIPCHK(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), s->stream));
        HIPCHK(hipMemsetAsync(dused, 0, (size_t)nsh + 1, s->stream));
        const dim3 grid((unsigned)((s->nz + 255) / 256), (unsigned)s->ny, (unsigned)s->nx);
        hipLaunchKernelGGL(fdtdk::k_bake, grid, dim3(256), 0, s->stream, a);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&count, dcount, sizeof(count), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (count > cap) {
            cap = count;
            continue;
        }
        hc.resize((size_t)count);
        ht.resize((size_t)count * 12);
        if (count) {
            HIPCHK(hipMemcpy(hc.data(), icell, count * sizeof(long long), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(ht.data(), itup, count * 12 * sizeof(double), hipMemcpyDeviceToHost));
        }
        HIPCHK(hipMemcpy(used.data(), dused, used.size(), hipMemcpyDeviceToHost));
        done = true;
    }
    if (!done) return fail(FDTD_EDEVICE, "bake interface list overflow");
    // pure tuples of the used materials: the per-component averages of ns^3
    // equal samples
    std::unordered_map<Key12, int, Key12Hash> ids;
    std::vector<uint16_t> remap((size_t)nsh + 1, 0);
    const int cnt = ns * ns * ns;
    for (int mtl = 0; mtl <= nsh; ++mtl) {
        if (!used[(size_t)mtl]) continue;
        const double* pr = &props[(size_t)4 * mtl];
        double eps_sum = 0, mu_sum = 0, sg_sum = 0, sgm_sum = 0;
        for (int c = 0; c < cnt; ++c) {
            eps_sum += pr[0] * bd->eps0;
            mu_sum += pr[1] * bd->mu0;
            sg_sum += pr[2];
            sgm_sum += pr[3];
        }
        const double eps = eps_sum / cnt, mu = mu_sum / cnt, sigma = sg_sum / cnt, sigm = sgm_sum / cnt;
        const double c1 = sigma * bd->dt / (2 * eps), c2 = sigm * bd->dt / (2 * mu);
        const double aE = (1 - c1) / (1 + c1), bE = (bd->dt / eps) / (1 + c1);
        const double aH = (1 - c2) / (1 + c2), bH = (bd->dt / mu) / (1 + c2);
        Key12 k;
        for (int c = 0; c < 3; ++c) {
            k.v[2 * c] = aE; k.v[2 * c + 1] = bE;
            k.v[6 + 2 * c] = aH; k.v[7 + 2 * c] = bH;
        }
        auto it = ids.find(k);
        if (it == ids.end()) {
            it = ids.emplace(k, (int)tuples.size()).first;
            tuples.push_back(k);
        }
        remap[(size_t)mtl] = (uint16_t)it->second;
    }
    // interface tuples, deduplicated in cell order
    std::vector<size_t> ord(hc.size());
    for (size_t t = 0; t < ord.size(); ++t) ord[t] = t;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return hc[x] < hc[y]; });
    std::vector<long long> sc(hc.size());
    std::vector<uint16_t> si(hc.size());
    for (size_t r = 0; r < ord.size(); ++r) {
        Key12 k;
        std::memcpy(k.v, &ht[ord[r] * 12], sizeof(k.v));
        auto it = ids.find(k);
        if (it == ids.end()) {
            if (tuples.size() >= 65536) return fail(FDTD_EINVAL, "more than 65536 distinct material coefficient tuples");
            it = ids.emplace(k, (int)tuples.size()).first;
            tuples.push_back(k);
        }
        sc[r] = hc[ord[r]];
        si[r] = (uint16_t)it->second;
    }
    s->nmat = (int)tuples.size();
    s->mid16 = s->nmat > 256;
    return s->mid16 ? finish_ids<uint16_t>(s, tmp, ids16, remap, sc, si) : finish_ids<uint8_t>(s, tmp, ids16, remap, sc, si);
}

template <typename R>
int setup_coeffs(fdtd_sim* s, const fdtd_grid_desc* g, const fdtd_material_desc* m, const fdtd_boundary_desc* bc) {
    // per-axis tables [7][n]: inverse spacing, kappa, bE, cE, bH, cH, 1/kappa
    const int dims[3] = {s->nx, s->ny, s->nz};
    const double* inv[3] = {g->inv_dx, g->inv_dy, g->inv_dz};
    for (int a = 0; a < 3; ++a) {
        const int n = dims[a];
        std::vector<R> v((size_t)7 * n);
        for (int i = 0; i < n; ++i) v[(size_t)5 * n + i] = 0;
        for (int i = 0; i < n; ++i) {
            if (s->bc_type == 1) {
                const double* K[3] = {bc->kx, bc->ky, bc->kz};
                const double* BE[3] = {bc->bEx, bc->bEy, bc->bEz};
                const double* CE[3] = {bc->cEx, bc->cEy, bc->cEz};
                const double* BH[3] = {bc->bHx, bc->bHy, bc->bHz};
                const double* CH[3] = {bc->cHx, bc->cHy, bc->cHz};
                v[(size_t)n + i] = (R)K[a][i];
                v[(size_t)2 * n + i] = (R)BE[a][i];
                v[(size_t)3 * n + i] = (R)CE[a][i];
                v[(size_t)4 * n + i] = (R)BH[a][i];
                v[(size_t)5 * n + i] = (R)CH[a][i];
            } else {
                v[(size_t)n + i] = 1; v[(size_t)2 * n + i] = 1; v[(size_t)3 * n + i] = 0;
                v[(size_t)4 * n + i] = 1; v[(size_t)5 * n + i] = 0;
            }
            v[i] = (R)inv[a][i];
            const double kap = s->bc_type == 1 ? (a == 0 ? bc->kx : a == 1 ? bc->ky : bc->kz)[i] : 1.0;
            v[(size_t)6 * n + i] = (R)(1.0 / kap);
        }
        R* d;
        if (int rc = s->upload(&d, v.data(), v.size())) return rc;
        s->cp[a] = d;
    }
    // Material tuples -> ids + exact LUT.
    std::vector<Key12> tuples;
    if (m->bake && !m->aEx) {
        if (int rc = bake_on_device(s, m->bake, tuples)) return rc;
    } else if (int rc = ids_from_arrays(s, m, tuples)) {
        return rc;
    }
    std::vector<R> lut((size_t)12 * s->nmat);
    bool lossless = true, hu = true;
    for (int t = 0; t < s->nmat; ++t) {
        for (int q = 0; q < 12; ++q) lut[(size_t)q * s->nmat + t] = (R)tuples[t].v[q];
        if (tuples[t].v[0] != 1.0 || tuples[t].v[2] != 1.0 || tuples[t].v[4] != 1.0) lossless = false;
        for (int q = 6; q < 12; ++q)
            if (tuples[t].v[q] != tuples[0].v[q]) hu = false;
    }
    s->lossless = lossless;
    s->h_uniform = hu;
    R* d;
    if (int rc = s->upload(&d, lut.data(), lut.size())) return rc;
    s->lut = d;
    s->h_tuples = std::move(tuples);
    return 0;
}

int setup_gain(fdtd_sim* s, const fdtd_tls_desc* t) {
    if (!t || t->n_regions <= 0 || !t->cell_region_map) return 0;
    if (t->n_regions > fdtdk::STAGE_MAX_TLS) return fail(FDTD_EINVAL, "at most 16 TLS regions");
    if (!t->Ndip || !t->Ng0 || !t->Nu || !t->Ng || !t->Pz || !t->Pz_prev)
        return fail(FDTD_EINVAL, "TLS state array missing");
    int lo[3] = {s->nx, s->ny, s->nz}, hi[3] = {-1, -1, -1};
    long count = 0;
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                const long id = ((long)i * s->ny + j) * s->nz + k;
                const int r = t->cell_region_map[id];
                if (r < 0 || r >= t->n_regions || !(t->Ndip[id] > 0)) continue;
                const double N0 = t->regions[r].N0;
                if (t->Ndip[id] != N0 || t->Ng0[id] != N0)
                    return fail(FDTD_EINVAL, "TLS cell with Ndip/Ng0 different from its region's N0");
                if (s->bc_type == 0 && (i == 0 || j == 0 || k == 0 || i == s->nx - 1 || j == s->ny - 1 || k == s->nz - 1))
                    return fail(FDTD_EINVAL, "TLS cells on a PEC face are not supported");
                const int c[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], c[a]);
                    hi[a] = std::max(hi[a], c[a]);
                }
                ++count;
            }
    if (count == 0) return 0;
    s->gain = true;
    s->ngain = count;
    for (int a = 0; a < 3; ++a) {
        s->g0[a] = lo[a];
        s->gn[a] = hi[a] - lo[a] + 1;
    }
    s->gbox = (long)s->gn[0] * s->gn[1] * s->gn[2];
    std::vector<int8_t> reg((size_t)s->gbox, -1);
    std::vector<double> Pz((size_t)s->gbox, 0.0), Pp((size_t)s->gbox, 0.0), Nu((size_t)s->gbox, 0.0), Ng((size_t)s->gbox, 0.0);
    for (int gi = 0; gi < s->gn[0]; ++gi)
        for (int gj = 0; gj < s->gn[1]; ++gj)
            for (int gk = 0; gk < s->gn[2]; ++gk) {
                const long g = ((long)gi * s->gn[1] + gj) * s->gn[2] + gk;
                const long id = ((long)(gi + s->g0[0]) * s->ny + (gj + s->g0[1])) * s->nz + (gk + s->g0[2]);
                const int r = t->cell_region_map[id];
                if (r < 0 || r >= t->n_regions || !(t->Ndip[id] > 0)) continue;
                reg[(size_t)g] = (int8_t)r;
                Pz[(size_t)g] = t->Pz[id];
                Pp[(size_t)g] = t->Pz_prev[id];
                Nu[(size_t)g] = t->Nu[id];
                Ng[(size_t)g] = t->Ng[id];
            }
    int rc;
    if ((rc = s->upload(&s->greg, reg.data(), reg.size()))) return rc;
    if ((rc = s->upload(&s->gPz, Pz.data(), Pz.size()))) return rc;
    if ((rc = s->upload(&s->gPp, Pp.data(), Pp.size()))) return rc;
    if ((rc = s->upload(&s->gNu, Nu.data(), Nu.size()))) return rc;
    if ((rc = s->upload(&s->gNg, Ng.data(), Ng.size()))) return rc;
    s->h_tls.resize((size_t)t->n_regions);
    for (int r = 0; r < t->n_regions; ++r) {
        const fdtd_tls_region& R = t->regions[r];
        TLSCoef c{};
        c.kapa = R.kapa; c.pa2 = R.pa2; c.pa3 = R.pa3; c.inv_dt = R.inv_dt;
        c.inv_tau = 1.0 / R.tau;                        // src/two_level_system.hpp:637
        c.inv_hbar_omega = 1.0 / (R.hbar * R.omega_a);  // :638
        c.N0 = R.N0; c.inv_N0 = 1.0 / R.N0; c.dt = R.dt; c.clamp = R.clamp;
        s->h_tls[(size_t)r] = c;
    }
    return s->upload(&s->d_tls, s->h_tls.data(), s->h_tls.size());
}

// Geometry of the fused step kernel for this grid: CW cells per lane and kc
// k-chunk wavefronts per work-group (kc * CW * 64 >= NzT with the least
// unused lanes, then 2, 3, 4, 1 cells per lane in that order: 2 was fastest
// at 512x512x384), jw rows per work-group (2), ic planes per work-group (8).
// Overrides (fdtd_set_kernel_mode): s->fz_cw, s->fz_jw, s->ichunk.
constexpr int MARCH_MAX_JW = 4;
bool march_geom(const fdtd_sim* s, fdtdk::FusedGeom& g, int& cw) {
    const int nseg = (s->nz + 63) / 64;
    static const int pref[4] = {2, 3, 4, 1};
    int best = -1;
    for (int t = 0; t < 4; ++t) {
        const int c = pref[t];
        if (s->fz_cw && c != s->fz_cw) continue;
        const int kc = (nseg + c - 1) / c;
        if (kc > fdtdk::FZ_MARCH_MAX_KC) continue;
        const int score = (100 - (kc * c - nseg)) * 10 + (4 - t);
        if (score > best) { best = score; cw = c; g.kc = kc; }
    }
    if (best < 0) return false;
    g.jw = s->fz_jw ? std::min(s->fz_jw, MARCH_MAX_JW) : 2;
    g.njb = (s->ny + g.jw - 1) / g.jw;
    g.ic = std::min(s->ichunk > 0 ? s->ichunk : 8, s->nx);
    g.nwg = g.njb * ((s->nx + g.ic - 1) / g.ic);
    return true;
}
// The fused kernel addresses cells by 32-bit byte offsets, reads the
// cell-invariant H coefficients, and covers a row with at most
// FZ_MARCH_MAX_KC wavefronts.
bool fused_ok(const fdtd_sim* s) {
    if (!s->h_uniform || !s->fits32) return false;
    fdtdk::FusedGeom g{};
    int cw = 0;
    return march_geom(s, g, cw);
}
template <typename T, typename R, int CW>
void launch_march_cw(fdtd_sim* s, StepArgs<T, R>& a, const fdtdk::FusedGeom& g) {
    const dim3 grd((unsigned)g.nwg), blk((unsigned)(64 * g.kc));
    switch (g.jw) {
    case 1: hipLaunchKernelGGL((fdtdk::k_step_march<T, R, CW, 1>), grd, blk, 0, s->stream, a, g); break;
    case 2: hipLaunchKernelGGL((fdtdk::k_step_march<T, R, CW, 2>), grd, blk, 0, s->stream, a, g); break;
    case 3: hipLaunchKernelGGL((fdtdk::k_step_march<T, R, CW, 3>), grd, blk, 0, s->stream, a, g); break;
    default: hipLaunchKernelGGL((fdtdk::k_step_march<T, R, CW, 4>), grd, blk, 0, s->stream, a, g); break;
    }
}
template <typename T, typename R>
void launch_fused(fdtd_sim* s, StepArgs<T, R>& a) {
    fdtdk::FusedGeom g{};
    int cw = 0;
    march_geom(s, g, cw);
    switch (cw) {
    case 1: launch_march_cw<T, R, 1>(s, a, g); break;
    case 2: launch_march_cw<T, R, 2>(s, a, g); break;
    case 3: launch_march_cw<T, R, 3>(s, a, g); break;
    default: launch_march_cw<T, R, 4>(s, a, g); break;
    }
}
// Geometry of the dependency-ordered step (k_step_flow): queues of planes
// per XCD, bands of the two-kernel tile order, slots of H and E tiles.
// false where a launch would not fit: tickets or flags beyond 2^31.
bool flow_geom(const fdtd_sim* s, fdtdk::FlowGeom& g, unsigned& nwg, unsigned& nedge) {
    const unsigned long tk = (unsigned long)((s->nz + 63) / 64),
                        tj = (unsigned long)((s->ny + fdtdk::TILE_ROWS - 1) / fdtdk::TILE_ROWS);
    const unsigned long bj = tile_band(s, (unsigned)tj);
    const unsigned long nx = (unsigned long)s->nx;
    g = fdtdk::FlowGeom{};
    g.pbase = (unsigned)(nx / 8);
    g.prem = (unsigned)(nx % 8);
    const unsigned long pmax = g.pbase + (g.prem ? 1 : 0);
    g.lag = s->flow_lag > 0 ? (unsigned)s->flow_lag : 2u;
    const unsigned long tpp = tk * tj, nbf = tj / bj, bjl = tj - nbf * bj;
    const unsigned long T = bj * tk, Tl = bjl * tk, slots = pmax + g.lag;
    const unsigned long per_band = 2 * T * slots, tfull = nbf * per_band, tlast = 2 * Tl * slots;
    const unsigned long tickets = tfull + tlast;
    if (tickets >= (1ul << 31) || 2 * nx * tpp >= (1ul << 31)) return false;
    g.tpp = (unsigned)tpp;
    g.nbf = (unsigned)nbf;
    g.tfull = (unsigned)tfull;
    g.tlast = (unsigned)tlast;
    g.band_t = fdtdk::make_fastdiv((unsigned)per_band);
    g.slot_t = fdtdk::make_fastdiv((unsigned)(2 * T));
    g.slot_l = fdtdk::make_fastdiv(Tl ? (unsigned)(2 * Tl) : 2u);
    g.tppd = fdtdk::make_fastdiv((unsigned)tpp);
    g.hflags = s->d_flowflags;
    g.eflags = s->d_flowflags ? s->d_flowflags + nx * tpp : nullptr;
    g.ctl = s->d_flowctl;
    nwg = 0;  // persistent: set by the launch from the occupancy
    nedge = (unsigned)(std::min<unsigned long>(8, nx) * tpp);
    return true;
}
// The flags (zero: no step's epoch) and the control words ({1, 0}) of the
// dependency-ordered step.
int ensure_flow(fdtd_sim* s) {
    if (s->d_flowctl) return FDTD_OK;
    fdtdk::FlowGeom g{};
    unsigned nwg = 0, nedge = 0;
    if (!flow_geom(s, g, nwg, nedge)) return fail(FDTD_EINVAL, "mode 2: the grid is too large for one flow launch");
    const size_t nflags = 2 * (size_t)s->nx * g.tpp;
    unsigned ctl0[fdtdk::FLOW_CTL_WORDS] = {};
    ctl0[0] = 1u;
    void* f = nullptr;
    void* c = nullptr;
    if (int rc = s->alloc(&f, nflags * sizeof(unsigned))) return rc;
    if (int rc = s->alloc(&c, sizeof(ctl0))) return rc;
    HIPCHK(hipMemset(f, 0, nflags * sizeof(unsigned)));
    HIPCHK(hipMemcpy(c, ctl0, sizeof(ctl0), hipMemcpyHostToDevice));
    s->d_flowflags = (unsigned*)f;
    s->d_flowctl = (unsigned*)c;
    return FDTD_OK;
}
template <typename T, typename R, typename IX>
void launch_flow_ix(fdtd_sim* s, StepArgs<T, R>& a, unsigned long long* fm, hipEvent_t* ev) {
    fdtdk::FlowGeom g{};
    unsigned nwg = 0, nedge = 0;
    flow_geom(s, g, nwg, nedge);
    // persistent: as many work-groups as stay resident at once
    if (!s->flow_wg) {
        int per_cu = 0, cus = 0, dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fdtdk::k_step_flow<T, R, 32, IX, true>, 256, 0);
        s->flow_wg = (unsigned)std::max(8, std::max(per_cu, 1) * std::max(cus, 8));
    }
    nwg = s->flow_wg;
    if (ev) (void)hipEventRecord(ev[0], s->stream);
    if (fm) hipLaunchKernelGGL((fdtdk::k_step_flow<T, R, 32, IX, true>), dim3(nwg), dim3(256), 0, s->stream, a, g, fm);
    else hipLaunchKernelGGL((fdtdk::k_step_flow<T, R, 32, IX, false>), dim3(nwg), dim3(256), 0, s->stream, a, g, fm);
    if (ev) (void)hipEventRecord(ev[1], s->stream);
    if (ev) (void)hipEventRecord(ev[2], s->stream);
    hipLaunchKernelGGL((fdtdk::k_step_flow_edge<T, R, 32, IX>), dim3(nedge), dim3(256), 0, s->stream, a, g);
    hipLaunchKernelGGL(fdtdk::k_flow_advance, dim3(1), dim3(64), 0, s->stream, s->d_flowctl);
    if (ev) (void)hipEventRecord(ev[3], s->stream);
}

void set_cur(fdtd_sim* s, int c) {
    s->cur = c;
    for (int q = 0; q < 6; ++q) s->F[q] = s->F2[c][q];
    for (int q = 0; q < 12; ++q) s->psi[q] = (q < 6) ? s->psi2[0][q] : s->psi2[c][q];
}

// The second buffer of the fields and of psi_H, holding a copy of the
// current contents (cells a step does not write keep their values in both).
int ensure_second(fdtd_sim* s) {
    const int o = s->cur ^ 1;
    if (s->F2[o][0]) return FDTD_OK;
    for (int q = 0; q < 6; ++q) {
        if (int rc = s->alloc(&s->F2[o][q], (size_t)s->N * s->esz)) return rc;
        HIPCHK(hipMemcpyAsync(s->F2[o][q], s->F2[s->cur][q], (size_t)s->N * s->esz, hipMemcpyDeviceToDevice, s->stream));
    }
    for (int q = 6; q < 12; ++q) {
        if (!s->psi2[s->cur][q]) continue;
        if (int rc = s->alloc(&s->psi2[o][q], (size_t)s->psi_len[q] * s->esz)) return rc;
        HIPCHK(hipMemcpyAsync(s->psi2[o][q], s->psi2[s->cur][q], (size_t)s->psi_len[q] * s->esz,
                              hipMemcpyDeviceToDevice, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

constexpr int TLS_PARTS = 256;  // blocks of the TLS diagnostic reduction
constexpr long MONPART_MAX = 2048;  // blocks (and partial records) of the monitor reduction

void dispatch_monitor(fdtd_sim* s, MonOut* slot);

// DFT phase factors of step n (the step's field time, times dt), as the
// DFT monitor accumulates them.
void dft_phases(const fdtd_sim* s, const Dft& d, long n, double* wr, double* wi) {
    const double t = (d.comp < 3) ? ((double)n + 1.0) * s->dt : ((double)n + 0.5) * s->dt;
    for (int f = 0; f < d.nf; ++f) {
        const double w = 2.0 * 3.14159265358979323846 * d.freqs[f];
        wr[f] = std::cos(w * t) * s->dt;
        wi[f] = -std::sin(w * t) * s->dt;
    }
}

template <typename T, typename R>
int launch_step(fdtd_sim* s, long n, int sidx, hipEvent_t* ev) {
    StepArgs<T, R> a = make_args<T, R>(s);
    a.srcv = s->srcv_cur + (long)sidx * s->nsrc;
    if (s->mode == 1 && fused_ok(s) && s->F2[s->cur ^ 1][0]) {
        const int o = s->cur ^ 1;
        a.Exw = (T*)s->F2[o][0]; a.Eyw = (T*)s->F2[o][1]; a.Ezw = (T*)s->F2[o][2];
        a.Hxw = (T*)s->F2[o][3]; a.Hyw = (T*)s->F2[o][4]; a.Hzw = (T*)s->F2[o][5];
        for (int q = 6; q < 12; ++q) a.psiw[q] = (T*)s->psi2[o][q];
        if (s->mon_defer) {  // the fused step takes no monitor maxima: the full pass, before it
            dispatch_monitor(s, s->mon_defer);
            s->mon_defer = nullptr;
        }
        if (ev) (void)hipEventRecord(ev[0], s->stream);
        launch_fused<T, R>(s, a);
        if (ev) (void)hipEventRecord(ev[1], s->stream);
        if (ev) (void)hipEventRecord(ev[2], s->stream);
        if (ev) (void)hipEventRecord(ev[3], s->stream);
        set_cur(s, o);
        a = make_args<T, R>(s);
        a.srcv = s->srcv_cur + (long)sidx * s->nsrc;
    } else if (s->mode == 2 && s->d_flowctl) {
        // after a monitor step: the H tiles also take its field maxima
        unsigned long long* fm = s->mon_defer ? s->d_fmax : nullptr;
        if (s->off32 && !s->force64) launch_flow_ix<T, R, unsigned>(s, a, fm, ev);
        else launch_flow_ix<T, R, long>(s, a, fm, ev);
    } else {
        const dim3 blk(64 * fdtdk::TILE_ROWS);
        // narrow wavefronts pay off where instruction issue, not HBM, bounds
        // the kernels (FP16 storage)
        const int kw_def = 32;  // measured best or tied for every storage format
        const int kwid = s->kwid_h ? s->kwid_h : kw_def;
        const int kwid_e = s->kwid_e ? s->kwid_e : kw_def;
        // 32-bit byte offsets where every array the kernels touch is below 4 GiB
        const bool o32 = s->off32 && !s->force64;
        if (ev) (void)hipEventRecord(ev[0], s->stream);
        {
            const int n = s->nx;
            // one work-group per tile; a 1-D launch holds at most 2^32 threads,
            // so larger grids put the planes in gridDim.y (tile_of linearizes)
            const unsigned long ntile = (unsigned long)a.tk.d * a.tj.d * (unsigned long)n;
            const dim3 grd = ntile * blk.x < (1ul << 32) ? dim3((unsigned)ntile) : dim3(a.tk.d * a.tj.d, (unsigned)n);
            a.rev = 0;
            // after a monitor step: the H launch also takes its field maxima
            unsigned long long* fm = s->mon_defer ? s->d_fmax : nullptr;
            if (fm) {
                switch (kwid * 2 + (o32 ? 1 : 0)) {
                case 129: hipLaunchKernelGGL((fdtdk::k_update_H_mon<T, R, 64, unsigned>), grd, blk, 0, s->stream, a, fm); break;
                case 128: hipLaunchKernelGGL((fdtdk::k_update_H_mon<T, R, 64>), grd, blk, 0, s->stream, a, fm); break;
                case 65: hipLaunchKernelGGL((fdtdk::k_update_H_mon<T, R, 32, unsigned>), grd, blk, 0, s->stream, a, fm); break;
                case 64: hipLaunchKernelGGL((fdtdk::k_update_H_mon<T, R, 32>), grd, blk, 0, s->stream, a, fm); break;
                case 33: hipLaunchKernelGGL((fdtdk::k_update_H_mon<T, R, 16, unsigned>), grd, blk, 0, s->stream, a, fm); break;
                default: hipLaunchKernelGGL((fdtdk::k_update_H_mon<T, R, 16>), grd, blk, 0, s->stream, a, fm); break;
                }
            } else switch (kwid * 2 + (o32 ? 1 : 0)) {
            case 129: hipLaunchKernelGGL((fdtdk::k_update_H<T, R, 64, unsigned>), grd, blk, 0, s->stream, a); break;
            case 128: hipLaunchKernelGGL((fdtdk::k_update_H<T, R, 64>), grd, blk, 0, s->stream, a); break;
            case 65: hipLaunchKernelGGL((fdtdk::k_update_H<T, R, 32, unsigned>), grd, blk, 0, s->stream, a); break;
            case 64: hipLaunchKernelGGL((fdtdk::k_update_H<T, R, 32>), grd, blk, 0, s->stream, a); break;
            case 33: hipLaunchKernelGGL((fdtdk::k_update_H<T, R, 16, unsigned>), grd, blk, 0, s->stream, a); break;
            default: hipLaunchKernelGGL((fdtdk::k_update_H<T, R, 16>), grd, blk, 0, s->stream, a); break;
            }
            if (ev) (void)hipEventRecord(ev[1], s->stream);
            if (ev) (void)hipEventRecord(ev[2], s->stream);
            a.rev = s->erev ? 1 : 0;  // E walks back over the tiles H has just walked
            switch (kwid_e * 2 + (o32 ? 1 : 0)) {
            case 129: hipLaunchKernelGGL((fdtdk::k_update_E<T, R, 64, unsigned>), grd, blk, 0, s->stream, a); break;
            case 128: hipLaunchKernelGGL((fdtdk::k_update_E<T, R, 64>), grd, blk, 0, s->stream, a); break;
            case 65: hipLaunchKernelGGL((fdtdk::k_update_E<T, R, 32, unsigned>), grd, blk, 0, s->stream, a); break;
            case 64: hipLaunchKernelGGL((fdtdk::k_update_E<T, R, 32>), grd, blk, 0, s->stream, a); break;
            case 33: hipLaunchKernelGGL((fdtdk::k_update_E<T, R, 16, unsigned>), grd, blk, 0, s->stream, a); break;
            default: hipLaunchKernelGGL((fdtdk::k_update_E<T, R, 16>), grd, blk, 0, s->stream, a); break;
            }
            a.rev = 0;
        }
        if (ev) (void)hipEventRecord(ev[3], s->stream);
    }
    // probes: PointFieldDetector records when n % save_every == 0 (with a
    // step table, every probe rides in every step and the table holds this
    // step's record slot, or null)
    fdtdk::ProbeArgs pa{};
    const bool tab = s->tab_step != nullptr;
    pa.ind = tab ? 1 : 0;
    int pidx = 0;
    for (auto& p : s->probes) {
        const int q = pidx++;
        if (tab) {
            pa.p[pa.n].out = (double*)(s->tab_step + s->tab_probe_off + q);
        } else {
            if (n % p.save_every != 0 || p.count >= p.cap) continue;
            pa.p[pa.n].out = p.d_buf + p.count;
            p.count++;
        }
        pa.p[pa.n].id = p.id;
        pa.p[pa.n].comp = p.comp;
        if (++pa.n == 16) {
            hipLaunchKernelGGL((fdtdk::k_probes<T, R>), dim3(1), dim3(64), 0, s->stream, a, pa);
            pa.n = 0;
        }
    }
    // the remaining probes ride in the first DFT launch
    if (pa.n && s->dfts.empty()) {
        hipLaunchKernelGGL((fdtdk::k_probes<T, R>), dim3(1), dim3(64), 0, s->stream, a, pa);
        pa.n = 0;
    }
    for (size_t m = 0; m < s->dfts.size(); ++m) {
        const Dft& d = s->dfts[m];
        fdtdk::DftLaunch L{};
        fdtdk::DftArgs& da = L.d;
        da.comp = d.comp;
        da.i0 = d.box[0]; da.j0 = d.box[2]; da.k0 = d.box[4];
        da.ny_b = d.box[3] - d.box[2];
        da.nz_b = d.box[5] - d.box[4];
        da.ncell = d.ncell;
        da.nf = d.nf;
        if (tab) da.wtab = s->tab_step + s->tab_dft_off[m];
        else dft_phases(s, d, n, da.wr, da.wi);
        da.acc = d.d_acc;
        L.fz = fdtdk::make_fastdiv((unsigned)da.nz_b);
        L.fy = fdtdk::make_fastdiv((unsigned)da.ny_b);
        L.bpf = (unsigned)((d.ncell + 255) / 256);
        L.per_cell = da.nz_b == 1 && d.nf > 1;
        const unsigned nb = L.bpf * (L.per_cell ? 1u : (unsigned)d.nf) + (pa.n ? 1u : 0u);
        hipLaunchKernelGGL((fdtdk::k_observe<T, R>), dim3(nb), dim3(256), 0, s->stream, a, L, pa);
        pa.n = 0;
    }
    // TLS diagnostics after the populations (src/3D_FDTD_version008.cpp:330-336)
    if (s->tls_every > 0 && n % s->tls_every == 0 && s->tls_count < s->tls_cap) {
        const double* dx = s->d_dxyz;
        hipLaunchKernelGGL((fdtdk::k_tls_partial<T, R>), dim3(TLS_PARTS), dim3(256), 0, s->stream, a, dx, dx + s->nx,
                           dx + s->nx + s->ny, s->d_tlspart);
        hipLaunchKernelGGL(fdtdk::k_tls_final, dim3(1), dim3(64), 0, s->stream, (const double*)s->d_tlspart, TLS_PARTS,
                           (const int8_t*)s->greg, (const double*)s->gPz, (const double*)s->gPp,
                           (const fdtdk::TLSCoef*)s->d_tls, s->tls_gmon, s->d_tlsrec + 5 * s->tls_count);
        s->tls_count++;
    }
    if (s->mon_defer) {  // the previous step's monitor record: gain partials and field-maxima slots
        hipLaunchKernelGGL(fdtdk::k_monitor_final, dim3(1), dim3(64), 0, s->stream, (const MonOut*)s->d_monpart,
                           s->mon_defer_nb, s->mon_defer, (const unsigned long long*)s->d_fmax, fdtdk::MON_SLOTS);
        s->mon_defer = nullptr;
    }
    return 0;
}

template <typename T, typename R>
void launch_monitor(fdtd_sim* s, MonOut* slot) {
    StepArgs<T, R> a = make_args<T, R>(s);
    const void* f[6] = {a.Ex, a.Ey, a.Ez, a.Hx, a.Hy, a.Hz};
    bool aligned = true;
    for (const void* p : f) aligned = aligned && ((uintptr_t)p % 16 == 0);
    const long ng = s->gain ? s->gbox : 0L;
    if (aligned && s->d_monpart) {
        // at least four 16-byte vectors per thread and array, at most
        // MONPART_MAX blocks: block records (no same-address atomics),
        // merged by one small launch
        const long nv = s->N / (16 / (long)sizeof(T));
        const long nb = std::min<long>(std::max<long>((nv + 1023) / 1024, 1), MONPART_MAX);
        hipLaunchKernelGGL((fdtdk::k_monitor<T, R, true>), dim3((unsigned)nb), dim3(256), 0, s->stream, a, s->N, ng,
                           s->d_monpart);
        hipLaunchKernelGGL(fdtdk::k_monitor_final, dim3(1), dim3(64), 0, s->stream, (const MonOut*)s->d_monpart,
                           (int)nb, slot, (const unsigned long long*)nullptr, 0);
    } else {
        const long nb = std::min<long>(std::max<long>((s->N + 2047) / 2048, 1), 2048);
        hipLaunchKernelGGL((fdtdk::k_monitor<T, R, false>), dim3((unsigned)nb), dim3(256), 0, s->stream, a, s->N, ng, slot);
    }
}

// The monitor of a step whose next step runs in the same call on the
// two-kernel path: the gain part now (k_monitor over no field cells, with the
// full pass's block decomposition, so its partial records are the same), the
// field maxima in the next H launch (k_update_H_mon), merged after it.
// false where that does not apply (then the full pass runs).
template <typename T, typename R>
bool launch_monitor_begin(fdtd_sim* s, MonOut* slot) {
    if (s->mode == 1 && fused_ok(s) && s->F2[s->cur ^ 1][0]) return false;
    StepArgs<T, R> a = make_args<T, R>(s);
    const void* f[6] = {a.Ex, a.Ey, a.Ez, a.Hx, a.Hy, a.Hz};
    bool aligned = true;
    for (const void* p : f) aligned = aligned && ((uintptr_t)p % 16 == 0);
    if (!aligned || !s->d_monpart || !s->d_fmax || s->mon_defer) return false;
    const long ng = s->gain ? s->gbox : 0L;
    const long nv = s->N / (16 / (long)sizeof(T));
    const long nb = std::min<long>(std::max<long>((nv + 1023) / 1024, 1), MONPART_MAX);
    const long words = (long)fdtdk::MON_SLOTS * fdtdk::MON_SLOT_WORDS;
    hipLaunchKernelGGL(fdtdk::k_zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s->stream, s->d_fmax,
                       words);
    if (ng > 0)
        hipLaunchKernelGGL((fdtdk::k_monitor<T, R, true>), dim3((unsigned)nb), dim3(256), 0, s->stream, a, 0L, ng,
                           s->d_monpart);
    s->mon_defer = slot;
    s->mon_defer_nb = ng > 0 ? (int)nb : 0;
    return true;
}

int dispatch_step(fdtd_sim* s, long n, int sidx, hipEvent_t* ev) {
    switch (s->prec) {
    case FDTD_FP64: return launch_step<double, double>(s, n, sidx, ev);
    case FDTD_FP32: return launch_step<float, float>(s, n, sidx, ev);
    default: return launch_step<uint16_t, float>(s, n, sidx, ev);
    }
}
void dispatch_monitor(fdtd_sim* s, MonOut* slot) {
    switch (s->prec) {
    case FDTD_FP64: launch_monitor<double, double>(s, slot); break;
    case FDTD_FP32: launch_monitor<float, float>(s, slot); break;
    default: launch_monitor<uint16_t, float>(s, slot); break;
    }
}
bool dispatch_monitor_begin(fdtd_sim* s, MonOut* slot) {
    switch (s->prec) {
    case FDTD_FP64: return launch_monitor_begin<double, double>(s, slot);
    case FDTD_FP32: return launch_monitor_begin<float, float>(s, slot);
    default: return launch_monitor_begin<uint16_t, float>(s, slot);
    }
}

fdtdk::ArrSet fp16_arrays(const fdtd_sim* s) {
    fdtdk::ArrSet a{};
    for (int q = 0; q < 6; ++q) { a.p[q] = (const uint16_t*)s->F[q]; a.n[q] = s->N; }
    for (int q = 0; q < 12; ++q) { a.p[6 + q] = (const uint16_t*)s->psi[q]; a.n[6 + q] = s->psi[q] ? s->psi_len[q] : 0; }
    return a;
}

constexpr int FP16_MAX_BLOCKS = 256;  // blocks per array of the maxima scan (one record each)

// Keeps every FP16 array's largest stored magnitude inside [2^6, 2^13] by
// power-of-two rescaling (exact), entirely on the device.
void fp16_maintain(fdtd_sim* s, hipStream_t st) {
    const fdtdk::ArrSet a = fp16_arrays(s);
    hipLaunchKernelGGL(fdtdk::k_fp16_maxima, dim3(FP16_MAX_BLOCKS, 18), dim3(256), 0, st, a, s->d_maxbits);
    hipLaunchKernelGGL(fdtdk::k_fp16_plan, dim3(18), dim3(64), 0, st, s->d_sc, s->d_factor,
                       (const unsigned*)s->d_maxbits, FP16_MAX_BLOCKS);
    hipLaunchKernelGGL(fdtdk::k_fp16_rescale, dim3(512), dim3(256), 0, st, a, (const float*)s->d_factor);
    if (!s->F2[s->cur ^ 1][0]) return;  // single-buffered
    fdtdk::ArrSet o{};  // the other buffer holds the never-updated cells too
    for (int q = 0; q < 6; ++q) { o.p[q] = (const uint16_t*)s->F2[s->cur ^ 1][q]; o.n[q] = s->N; }
    for (int q = 6; q < 12; ++q) {
        o.p[6 + q] = (const uint16_t*)s->psi2[s->cur ^ 1][q];
        o.n[6 + q] = o.p[6 + q] ? s->psi_len[q] : 0;
    }
    hipLaunchKernelGGL(fdtdk::k_fp16_rescale, dim3(512), dim3(256), 0, st, o, (const float*)s->d_factor);
}

int read_scales(fdtd_sim* s, float* out36) {
    HIPCHK(hipMemcpyAsync(out36, s->d_sc, 36 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int write_scales(fdtd_sim* s, const float* in36) {
    HIPCHK(hipMemcpyAsync(s->d_sc, in36, 36 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

double bits_to_double(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof(d));
    return d;
}

// StabilityMonitor::check_stability policy (src/stability_monitor.hpp:42-144)
// on one device record; returns false when the reference would terminate.
bool apply_policy(fdtd_sim* s, const fdtd_check_record& r) {
    if (r.nan_E || r.nan_H) { s->stable = false; return false; }
    if (s->gain && (r.nan_P || r.nan_N)) { s->stable = false; return false; }
    const double max_E = r.max_E;
    if (max_E > 1e10) s->inst_count++;
    if (s->gain && r.max_P > 1e5) s->inst_count++;
    const size_t step = (size_t)r.step;
    if (step > 0 && step % 100 == 0 && s->prev_max_E > 1e-20) {
        s->prev_max_E = max_E;
    } else if (step == 0 || s->prev_max_E < 1e-20) {
        s->prev_max_E = std::max(max_E, 1e-20);
    }
    if (s->inst_count > 10) { s->stable = false; return false; }
    return s->stable;
}

template <typename T>
int download_raw(fdtd_sim* s, const void* src, size_t n, double scale, double* out) {
    std::vector<T> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < n; ++i) {
        if constexpr (sizeof(T) == 2) {
            __half h;
            std::memcpy(&h, &tmp[i], 2);
            out[i] = (double)__half2float(h) * scale;
        } else {
            out[i] = (double)tmp[i];
        }
    }
    return 0;
}

int download_any(fdtd_sim* s, const void* src, size_t n, double scale, double* out) {
    switch (s->prec) {
    case FDTD_FP64: return download_raw<double>(s, src, n, scale, out);
    case FDTD_FP32: return download_raw<float>(s, src, n, scale, out);
    default: return download_raw<uint16_t>(s, src, n, scale, out);
    }
}

int upload_any(fdtd_sim* s, void* dst, const double* in, size_t n, double scale) {
    std::vector<unsigned char> tmp(n * s->esz);
    for (size_t i = 0; i < n; ++i) {
        if (s->prec == FDTD_FP64) {
            std::memcpy(&tmp[i * 8], &in[i], 8);
        } else if (s->prec == FDTD_FP32) {
            float f = (float)in[i];
            std::memcpy(&tmp[i * 4], &f, 4);
        } else {
            __half h = __float2half_rn((float)(in[i] / scale));
            std::memcpy(&tmp[i * 2], &h, 2);
        }
    }
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), tmp.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// psi slab index of full-grid cell (i,j,k) for psi array q, or -1.
long psi_slab_index(const fdtd_sim* s, int q, int i, int j, int k) {
    const int np = s->np, nx = s->nx, ny = s->ny, nz = s->nz;
    auto in = [np](int v, int n) { int d = v < n - 1 - v ? v : n - 1 - v; return d < np; };
    static const int axis[12] = {1, 2, 2, 0, 0, 1, 1, 2, 2, 0, 0, 1};
    switch (axis[q]) {
    case 0:
        if (!in(i, nx)) return -1;
        return ((long)(i < np ? i : i - (nx - 2 * np)) * ny + j) * nz + k;
    case 1:
        if (!in(j, ny)) return -1;
        return ((long)i * (2 * np) + (j < np ? j : j - (ny - 2 * np))) * nz + k;
    default:
        if (!in(k, nz)) return -1;
        return ((long)i * ny + j) * (2 * np) + (k < np ? k : k - (nz - 2 * np));
    }
}

}  // namespace

extern "C" {

const char* fdtd_last_error(void) { return g_err.c_str(); }

int fdtd_create(const fdtd_grid_desc* g, const fdtd_material_desc* m, const fdtd_boundary_desc* bc,
                const fdtd_tls_desc* tls, int precision, fdtd_handle* out) {
    if (!g || !m || !bc || !out) return fail(FDTD_EINVAL, "null argument");
    *out = nullptr;
    if (g->nx < 3 || g->ny < 3 || g->nz < 3) return fail(FDTD_EINVAL, "grid must be at least 3 cells per axis");
    if (!g->inv_dx || !g->inv_dy || !g->inv_dz) return fail(FDTD_EINVAL, "spacing arrays missing");
    if (precision < FDTD_FP64 || precision > FDTD_FP16) return fail(FDTD_EINVAL, "unknown precision");
    if (bc->type != 0 && bc->type != 1) return fail(FDTD_EINVAL, "unknown boundary type");
    if (bc->type == 1) {
        if (bc->npml < 1 || 2 * bc->npml > g->nx || 2 * bc->npml > g->ny || 2 * bc->npml > g->nz)
            return fail(FDTD_EINVAL, "npml out of range");
        const double* need[] = {bc->kx, bc->ky, bc->kz, bc->bEx, bc->cEx, bc->bEy, bc->cEy, bc->bEz, bc->cEz,
                                bc->bHx, bc->cHx, bc->bHy, bc->cHy, bc->bHz, bc->cHz};
        for (const double* p : need)
            if (!p) return fail(FDTD_EINVAL, "CPML profile array missing");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FDTD_EDEVICE, "no HIP device available");
    fdtd_sim* s = new fdtd_sim();
    s->nx = g->nx; s->ny = g->ny; s->nz = g->nz;
    s->N = (long)g->nx * g->ny * g->nz;
    s->dt = g->dt;
    s->prec = precision;
    s->bc_type = bc->type;
    s->np = bc->type == 1 ? bc->npml : 0;
    s->esz = precision == FDTD_FP64 ? 8 : (precision == FDTD_FP32 ? 4 : 2);
    s->rsz = precision == FDTD_FP64 ? 8 : 4;
    int rc = 0;
    if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(FDTD_EDEVICE, "hipStreamCreate failed");
    }
    // one buffer of fields and psi; the second (read by the fused step
    // kernel while it writes the first) is allocated on the first switch to
    // a schedule that needs it (ensure_second)
    for (int q = 0; q < 6 && !rc; ++q) rc = s->alloc(&s->F2[0][q], (size_t)s->N * s->esz);
    if (!rc && s->bc_type == 1) {
        const long lx = 2L * s->np * s->ny * s->nz, ly = (long)s->nx * 2 * s->np * s->nz, lz = (long)s->nx * s->ny * 2 * s->np;
        static const int axis[12] = {1, 2, 2, 0, 0, 1, 1, 2, 2, 0, 0, 1};
        for (int q = 0; q < 12 && !rc; ++q) {
            s->psi_len[q] = axis[q] == 0 ? lx : (axis[q] == 1 ? ly : lz);
            rc = s->alloc(&s->psi2[0][q], (size_t)s->psi_len[q] * s->esz);
        }
    }
    if (!rc) {
        for (int q = 0; q < 6; ++q) hipMemsetAsync(s->F2[0][q], 0, (size_t)s->N * s->esz, s->stream);
        for (int q = 0; q < 12; ++q)
            if (s->psi2[0][q]) hipMemsetAsync(s->psi2[0][q], 0, (size_t)s->psi_len[q] * s->esz, s->stream);
        set_cur(s, 0);
    }
    if (!rc) rc = precision == FDTD_FP64 ? setup_coeffs<double>(s, g, m, bc) : setup_coeffs<float>(s, g, m, bc);
    if (!rc) rc = setup_gain(s, tls);
    if (!rc) rc = s->alloc(reinterpret_cast<void**>(&s->d_src), sizeof(SourceBox) * 64);
    if (!rc) rc = s->alloc(reinterpret_cast<void**>(&s->d_sc), 54 * sizeof(float));
    if (!rc) rc = s->alloc(reinterpret_cast<void**>(&s->d_factor), 20 * sizeof(float));
    if (!rc && hipMemset(s->d_factor, 0, 20 * sizeof(float)) != hipSuccess) rc = fail(FDTD_EDEVICE, "scale init failed");
    if (!rc) rc = s->alloc(reinterpret_cast<void**>(&s->d_maxbits), 18 * FP16_MAX_BLOCKS * sizeof(unsigned));
    if (!rc) {
        // Initial FP16 scales: E ~ 1, H ~ E/eta0 (2^-8), psi_E ~ H/d, psi_H ~ E/d;
        // they adapt from the first monitor step on.
        float sc[54];
        double imax = 0;
        for (int i = 0; i < g->nx; ++i) imax = std::max(imax, g->inv_dx[i]);
        for (int i = 0; i < g->ny; ++i) imax = std::max(imax, g->inv_dy[i]);
        for (int i = 0; i < g->nz; ++i) imax = std::max(imax, g->inv_dz[i]);
        int e;
        std::frexp(imax, &e);
        for (int q = 0; q < 18; ++q) sc[q] = 1.0f;
        if (precision == FDTD_FP16) {
            for (int q = 3; q < 6; ++q) sc[q] = std::ldexp(1.0f, -8);
            for (int q = 0; q < 6; ++q) sc[6 + q] = std::ldexp(1.0f, e - 8);   // psi_E
            for (int q = 6; q < 12; ++q) sc[6 + q] = std::ldexp(1.0f, e);      // psi_H
        }
        for (int q = 0; q < 18; ++q) sc[18 + q] = 1.0f / sc[q];
        for (int q = 0; q < 18; ++q) sc[36 + q] = std::ldexp(sc[q], -6);  // floors
        if (hipMemcpy(s->d_sc, sc, sizeof(sc), hipMemcpyHostToDevice) != hipSuccess) rc = fail(FDTD_EDEVICE, "scale init failed");
        if (!rc && hipMemset(s->d_maxbits, 0, 18 * FP16_MAX_BLOCKS * sizeof(unsigned)) != hipSuccess)
            rc = fail(FDTD_EDEVICE, "scale init failed");
    }
    if (!rc) {
        // 32-bit byte offsets in the H/E kernels when every array they index
        // by cell (fields, material ids, psi slabs, gain state) is below
        // 4 GiB.  Measured on the 256^3 and 512x512x384 gain-slab grids: 1.2%
        // faster with FP16 storage, 1% slower with FP32 (whose kernels are
        // bound by memory, not issue), so FP16 storage only.
        const double esz = precision == FDTD_FP64 ? 8.0 : precision == FDTD_FP32 ? 4.0 : 2.0;
        const double lim = 4294967295.0;
        s->fits32 = (double)s->N * std::max(esz, 2.0) < lim && (double)s->gbox * 8.0 < lim;
        s->off32 = precision == FDTD_FP16 && s->fits32;
    }
    if (!rc) {  // monitor block partials (allocated here so the footprint is known at creation)
        void* p = nullptr;
        rc = s->alloc(&p, (size_t)MONPART_MAX * sizeof(MonOut));
        s->d_monpart = (MonOut*)p;
    }
    if (!rc) {
        void* p = nullptr;
        rc = s->alloc(&p, (size_t)fdtdk::MON_SLOTS * fdtdk::MON_SLOT_WORDS * sizeof(unsigned long long));
        s->d_fmax = (unsigned long long*)p;
    }
    if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(FDTD_EDEVICE, "initialisation failed");
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return FDTD_OK;
}

int fdtd_destroy(fdtd_handle h) {
    if (h && h->batch) fdtd_batch_remove(h->batch, h);
    delete h;
    return FDTD_OK;
}

int fdtd_device_bytes(fdtd_handle h, size_t* bytes) {
    if (!h || !bytes) return fail(FDTD_EINVAL, "null argument");
    *bytes = h->bytes;
    return FDTD_OK;
}

static int step_finish(fdtd_handle h);

int fdtd_add_tls_monitor(fdtd_handle h, const double* dx, const double* dy, const double* dz, int i, int j, int k,
                         int save_every, long capacity) {
    if (!h || !dx || !dy || !dz) return fail(FDTD_EINVAL, "null argument");
    if (!h->gain) return fail(FDTD_EINVAL, "no TLS region on this handle");
    if (h->batch) return fail(FDTD_EINVAL, "domain is a batch member");
    if (save_every < 1 || capacity < 1) return fail(FDTD_EINVAL, "bad TLS monitor arguments");
    if (i < 0 || j < 0 || k < 0 || i >= h->nx || j >= h->ny || k >= h->nz) return fail(FDTD_EINVAL, "monitor cell outside grid");
    if (int rc = step_finish(h)) return rc;
    std::vector<double> d((size_t)h->nx + h->ny + h->nz);
    std::copy(dx, dx + h->nx, d.begin());
    std::copy(dy, dy + h->ny, d.begin() + h->nx);
    std::copy(dz, dz + h->nz, d.begin() + h->nx + h->ny);
    if (int rc = h->upload(&h->d_dxyz, d.data(), d.size())) return rc;
    void* p;
    if (int rc = h->alloc(&p, (size_t)capacity * 5 * sizeof(double))) return rc;
    h->d_tlsrec = (double*)p;
    if (int rc = h->alloc(&p, (size_t)TLS_PARTS * 6 * sizeof(double))) return rc;
    h->d_tlspart = (double*)p;
    const int gi = i - h->g0[0], gj = j - h->g0[1], gk = k - h->g0[2];
    h->tls_gmon = (gi >= 0 && gj >= 0 && gk >= 0 && gi < h->gn[0] && gj < h->gn[1] && gk < h->gn[2])
                      ? ((long)gi * h->gn[1] + gj) * h->gn[2] + gk
                      : -1;
    h->tls_every = save_every;
    h->tls_cap = capacity;
    h->tls_count = 0;
    return FDTD_OK;
}

int fdtd_read_tls_monitor(fdtd_handle h, double* out, long max, long* count) {
    if (!h || !count || (max > 0 && !out)) return fail(FDTD_EINVAL, "null argument");
    if (int rc = step_finish(h)) return rc;
    const long n = std::min(max, h->tls_count);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(out, h->d_tlsrec, (size_t)n * 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    *count = h->tls_count;
    return FDTD_OK;
}

int fdtd_download_material(fdtd_handle h, int q, double* out) {
    if (!h || !out || q < 0 || q > 11) return fail(FDTD_EINVAL, "bad material array request");
    if (h->batch) return fail(FDTD_EINVAL, "domain is a batch member");
    std::vector<uint8_t> raw((size_t)h->N * (h->mid16 ? 2 : 1));
    HIPCHK(hipMemcpyAsync(raw.data(), h->matid, raw.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const uint16_t* m16 = reinterpret_cast<const uint16_t*>(raw.data());
    for (long id = 0; id < h->N; ++id) {
        const int t = h->mid16 ? m16[id] : raw[(size_t)id];
        if (t >= (int)h->h_tuples.size()) return fail(FDTD_EDEVICE, "material id out of range");
        out[id] = h->h_tuples[(size_t)t].v[q];
    }
    return FDTD_OK;
}

int fdtd_set_sources(fdtd_handle h, int n, const fdtd_source_desc* src) {
    if (!h || n < 0 || n > 64 || (n > 0 && !src)) return fail(FDTD_EINVAL, "bad source list (at most 64)");
    std::vector<SourceBox> b((size_t)n);
    for (int q = 0; q < n; ++q) {
        const fdtd_source_desc& d = src[q];
        if (d.component < 0 || d.component > 2) return fail(FDTD_EINVAL, "source component must be 0..2");
        b[(size_t)q] = SourceBox{d.component, d.i0, d.i1, d.j0, d.j1, d.k0, d.k1};
    }
    if (n) HIPCHK(hipMemcpyAsync(h->d_src, b.data(), b.size() * sizeof(SourceBox), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->nsrc = n;
    return FDTD_OK;
}

int fdtd_add_probe(fdtd_handle h, int i, int j, int k, int component, int save_every, long capacity) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    if (i < 0 || j < 0 || k < 0 || i >= h->nx || j >= h->ny || k >= h->nz) return fail(FDTD_EINVAL, "probe outside grid");
    if (component < 0 || component > 11 || save_every < 1 || capacity < 1) return fail(FDTD_EINVAL, "bad probe arguments");
    Probe p{};
    p.id = ((long)i * h->ny + j) * h->nz + k;
    p.comp = component;
    p.save_every = save_every;
    p.cap = capacity;
    p.count = 0;
    if (int rc = h->alloc(reinterpret_cast<void**>(&p.d_buf), (size_t)capacity * sizeof(double))) return rc;
    h->probes.push_back(p);
    return (int)h->probes.size() - 1;
}

int fdtd_add_dft(fdtd_handle h, int component, const int box[6], int n_freqs, const double* freqs) {
    if (!h || !box || !freqs) return fail(FDTD_EINVAL, "null argument");
    if (component < 0 || component > 5 || n_freqs < 1 || n_freqs > 8) return fail(FDTD_EINVAL, "bad DFT arguments");
    if (box[0] < 0 || box[1] > h->nx || box[0] >= box[1] || box[2] < 0 || box[3] > h->ny || box[2] >= box[3] ||
        box[4] < 0 || box[5] > h->nz || box[4] >= box[5])
        return fail(FDTD_EINVAL, "DFT box outside grid");
    Dft d{};
    d.comp = component;
    std::memcpy(d.box, box, sizeof(d.box));
    d.nf = n_freqs;
    for (int f = 0; f < n_freqs; ++f) d.freqs[f] = freqs[f];
    d.ncell = (long)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
    if (d.ncell >= (1L << 31)) return fail(FDTD_EINVAL, "DFT box of 2^31 cells or more");
    const size_t b = (size_t)d.ncell * n_freqs * 2 * sizeof(double);
    if (int rc = h->alloc(reinterpret_cast<void**>(&d.d_acc), b)) return rc;
    HIPCHK(hipMemsetAsync(d.d_acc, 0, b, h->stream));
    h->dfts.push_back(d);
    return (int)h->dfts.size() - 1;
}

int fdtd_set_fp16_scales(fdtd_handle h, const double* scales18) {
    if (!h || !scales18) return fail(FDTD_EINVAL, "null argument");
    if (h->prec != FDTD_FP16) return FDTD_OK;
    float sc[54];
    for (int q = 0; q < 18; ++q) {
        int e;
        const double fr = std::frexp(scales18[q], &e);
        if (!(scales18[q] > 0) || fr != 0.5) return fail(FDTD_EINVAL, "FP16 scales must be positive powers of two");
        sc[q] = (float)scales18[q];
        sc[18 + q] = (float)(1.0 / scales18[q]);
        sc[36 + q] = std::ldexp(sc[q], -6);
    }
    HIPCHK(hipMemcpyAsync(h->d_sc, sc, sizeof(sc), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

int fdtd_get_fp16_scales(fdtd_handle h, double* scales18) {
    if (!h || !scales18) return fail(FDTD_EINVAL, "null argument");
    float sc[36];
    if (int rc = read_scales(h, sc)) return rc;
    for (int q = 0; q < 18; ++q) scales18[q] = sc[q];
    return FDTD_OK;
}

int fdtd_set_kernel_mode(fdtd_handle h, int mode, int param) {
    if (!h || mode < 0 || mode > 2)
        return fail(FDTD_EINVAL, "mode must be 0 (H and E kernels), 1 (fused step kernel) or 2 (dependency-ordered step)");
    if (mode == 2) {
        if (param < 0 || param > 64) return fail(FDTD_EINVAL, "mode 2: param is the plane lag of the E tiles, 0..64 (0: 2)");
        if (int rc = step_finish(h)) return rc;
        if (int rc = ensure_flow(h)) return rc;
        h->flow_lag = param;
        h->mode = 2;
        return FDTD_OK;
    }
    auto width_ok = [](int w) { return w == 16 || w == 32 || w == 64; };
    if (mode == 0) {
        h->force64 = (param & 0x20000) != 0;  // 64-bit element offsets even where 32-bit ones fit
        h->erev = (param & 0x100000) == 0;    // set: E tiles in forward order, as H walks them
        h->mon_fused_h = (param & 0x40000) == 0;  // set: the monitor's field maxima in a pass of their own
        if (param >> 21) h->band_tiles = (unsigned)((param >> 21) & 0x3ff);  // j-tiles per band of the tile order
        param &= 0x1ffff;
    }
    const int wh = param < 256 ? param : (param & 255), we = param < 256 ? param : (param >> 8);
    if (mode == 0 && param != 0 && !(param < 65536 && width_ok(wh) && width_ok(we)))
        return fail(FDTD_EINVAL, "mode 0: the wavefront k-widths must be 16, 32 or 64 (0 keeps them)");
    const int m1_ic = param & 0xfff, m1_cw = (param >> 12) & 0xf, m1_jw = (param >> 16) & 0xf;
    if (mode == 1 && (param < 0 || param >= (1 << 20) || m1_cw > fdtdk::FZ_MAX_CW))
        return fail(FDTD_EINVAL, "mode 1: planes_per_chunk + 4096 * cells_per_lane + 65536 * rows, "
                                 "planes_per_chunk in [0, 4095], cells_per_lane in [0, 4], rows in [0, 15] (above 4: 4)");
    if (mode == 1) {
        const int keep[3] = {h->ichunk, h->fz_cw, h->fz_jw};
        h->ichunk = m1_ic; h->fz_cw = m1_cw; h->fz_jw = m1_jw;
        if (!fused_ok(h)) {
            h->ichunk = keep[0]; h->fz_cw = keep[1]; h->fz_jw = keep[2];
            return fail(FDTD_EINVAL, "mode 1: the fused step kernel does not take this grid (it needs cell-invariant H "
                                     "coefficients, arrays below 4 GiB and a row in at most 8 wavefronts)");
        }
        if (int rc = ensure_second(h)) return rc;
    }
    h->mode = mode;
    if (param && mode == 0) {
        h->kwid_h = wh;
        h->kwid_e = we;
    }
    return FDTD_OK;
}

int fdtd_set_step_graphs(fdtd_handle h, int on) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    if (int rc = step_finish(h)) return rc;
    h->graph_ticks = on != 0;
    if (!h->graph_ticks) {
        for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
        h->graphs.clear();
    }
    return FDTD_OK;
}

int fdtd_set_device_pool(long max_cached_bytes) {
    if (max_cached_bytes < 0) return fail(FDTD_EINVAL, "negative cache size");
    DevPool& P = dev_pool();
    std::lock_guard<std::mutex> g(P.mu);
    P.cap = (size_t)max_cached_bytes;
    pool_trim(P, P.cap);
    return FDTD_OK;
}

int fdtd_device_pool_bytes(long* cached, long* max_cached_bytes) {
    if (!cached && !max_cached_bytes) return fail(FDTD_EINVAL, "null argument");
    DevPool& P = dev_pool();
    std::lock_guard<std::mutex> g(P.mu);
    if (cached) *cached = (long)P.cached;
    if (max_cached_bytes) *max_cached_bytes = (long)P.cap;
    return FDTD_OK;
}

int fdtd_set_monitor(fdtd_handle h, int interval) {
    if (!h || interval < 0) return fail(FDTD_EINVAL, "bad monitor interval");
    h->mon_interval = interval;
    return FDTD_OK;
}

// Scratch of fdtd_step_noalloc: the monitor records and the per-step source
// values of one call.  Every device holds its own copy at its own address,
// looked up once per device with that device current; the array's name is a
// host-side stand-in, so host code uses only the looked-up address.
static __device__ __align__(256) unsigned char scratch[FDTD_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_addr[64] = {};

static int scratch_address(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= 64) return fail(FDTD_EDEVICE, "device ordinal out of range for the no-alloc scratch");
    if (!g_scratch_addr[dev]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return fail(FDTD_EDEVICE, "hipGetDevice failed");
        if (cur != dev && hipSetDevice(dev) != hipSuccess) return fail(FDTD_EDEVICE, "hipSetDevice failed");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(FDTD_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = (unsigned char*)g_scratch_addr[dev];
    return FDTD_OK;
}

// Reads the monitor records of the steps last enqueued and applies the
// stability policy (StabilityMonitor::check_stability, src/stability_monitor.hpp:42-144).
static int step_finish(fdtd_handle h) {
    if (!h->pending) return FDTD_OK;
    h->pending = false;
    if (h->d_flowctl) {
        // the dependency-ordered step's counter (read after every call of a
        // handle that has used that step): a tile the launch did not run, or
        // a wait that timed out, means the call's results cannot be trusted
        unsigned c = 0;
        HIPCHK(hipMemcpyAsync(&c, h->d_flowctl + 1, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (c) {
            const unsigned zero = 0;
            HIPCHK(hipMemcpy(h->d_flowctl + 1, &zero, sizeof(zero), hipMemcpyHostToDevice));
            h->pending_mon.clear();
            h->mode = 0;
            return fail(FDTD_EDEVICE, "dependency-ordered step: " + std::to_string(c >> 16) +
                                          " tiles were not run (work-groups not dealt round-robin to the XCDs) and " +
                                          std::to_string(c & 0xffff) +
                                          " dependency waits timed out; this call's results are invalid "
                                          "(the handle is back on the two-kernel step)");
        }
    }
    const std::vector<long> mon_steps = std::move(h->pending_mon);
    h->pending_mon.clear();
    h->checks.clear();
    if (mon_steps.empty()) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipGetLastError());
        return FDTD_OK;
    }
    std::vector<MonOut> mo(mon_steps.size());
    // FP16 storage: the scale planner's sticky flags of a NaN/Inf stored in
    // an E-class or H-class array (fields or psi)
    unsigned nonfinite[2] = {0u, 0u};
    if (h->rec_host) {
        // the call's last launches copied both into pinned host memory: no
        // copy to queue now behind other streams' work on the hardware queue
        HIPCHK(hipStreamSynchronize(h->stream));
        std::memcpy(mo.data(), h->h_rec, mo.size() * sizeof(MonOut));
        if (h->prec == FDTD_FP16)
            std::memcpy(nonfinite, h->h_rec + mo.size() * (sizeof(MonOut) / sizeof(unsigned long long)),
                        sizeof(nonfinite));
    } else {
        HIPCHK(hipMemcpyAsync(mo.data(), h->mon_cur, mo.size() * sizeof(MonOut), hipMemcpyDeviceToHost, h->stream));
        if (h->prec == FDTD_FP16)
            HIPCHK(hipMemcpyAsync(nonfinite, h->d_factor + 18, sizeof(nonfinite), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    bool ok = true;
    long fail_step = -1;
    for (size_t q = 0; q < mo.size(); ++q) {
        fdtd_check_record r{};
        const unsigned long long inf = 0x7ff0000000000000ull;
        r.step = mon_steps[q];
        r.nan_E = mo[q].maxE >= inf; r.nan_H = mo[q].maxH >= inf;
        r.nan_P = mo[q].maxP >= inf; r.nan_N = mo[q].nanN >= inf;
        r.max_E = bits_to_double(mo[q].maxE);
        r.max_H = bits_to_double(mo[q].maxH);
        r.max_P = bits_to_double(mo[q].maxP);
        r.pop_total = mo[q].pop_total;
        r.pop_initial = mo[q].pop_init;
        if (q + 1 == mo.size()) {
            r.nan_E |= nonfinite[0] != 0u;
            r.nan_H |= nonfinite[1] != 0u;
        }
        h->checks.push_back(r);
        if (ok && !apply_policy(h, r)) { ok = false; fail_step = r.step; }
    }
    if (!ok) return fail(FDTD_EUNSTABLE, "simulation unstable at step " + std::to_string(fail_step));
    return FDTD_OK;
}

constexpr int REC_FLAG_WORDS = 1;  // the FP16 non-finite flags (two 32-bit words)
// Host copy of a call's monitor records (and FP16 flags): grown before the
// call's launches, never while a call is in flight.
static int ensure_host_records(fdtd_handle h, size_t nrec) {
    const size_t words = nrec * (sizeof(MonOut) / sizeof(unsigned long long)) + REC_FLAG_WORDS;
    if (words <= h->rec_words) return FDTD_OK;
    if (h->h_rec) {
        HIPCHK(hipStreamSynchronize(h->stream));
        pool_host_free(PIN_MAPPED, h->h_rec, h->rec_words * sizeof(unsigned long long));
    }
    h->h_rec = h->d_rec = nullptr;
    h->rec_words = 0;
    const size_t cap = std::max<size_t>(words, 64 * (sizeof(MonOut) / sizeof(unsigned long long)) + REC_FLAG_WORDS);
    size_t got = 0;
    if (pool_host_alloc(PIN_MAPPED, cap * sizeof(unsigned long long), (void**)&h->h_rec, &got) != hipSuccess)
        return fail(FDTD_ENOMEM, "hipHostMalloc of the monitor record copy failed");
    HIPCHK(hipHostGetDevicePointer((void**)&h->d_rec, h->h_rec, 0));
    h->rec_words = got / sizeof(unsigned long long);
    return FDTD_OK;
}
// The copy kernels at the end of a call: records, then the FP16 flags.
static void enqueue_record_copy(fdtd_handle h, size_t nrec) {
    const long rw = (long)(nrec * (sizeof(MonOut) / sizeof(unsigned long long)));
    if (rw > 0)
        hipLaunchKernelGGL(fdtdk::k_copy_words, dim3(1), dim3(64), 0, h->stream, (const unsigned long long*)h->mon_cur,
                           h->d_rec, rw);
    if (h->prec == FDTD_FP16)
        hipLaunchKernelGGL(fdtdk::k_copy_words, dim3(1), dim3(64), 0, h->stream,
                           (const unsigned long long*)(h->d_factor + 18), h->d_rec + rw, (long)REC_FLAG_WORDS);
}

// The launches of one call (steps n0 .. n0 + nsteps - 1) as a replayed HIP
// graph.  The call's table (each step's DFT phase factors and probe record
// slots) is uploaded first; the graph for this call's shape is recorded on
// the first call of that shape: the launches are the direct path's, with
// the monitor records cleared by a fill kernel (no memset node in a graph)
// and the per-step values read from the table.  The shape is everything the
// launch sequence and its arguments depend on: the step count, the step's
// phase in the monitor and FP16 cadences, the kernel schedule and the device
// buffers the launches name.
static int step_graph_launch(fdtd_handle h, long n0, int nsteps, const std::vector<long>& mon_steps,
                             double* host_tab) {
    // the table: per step, each DFT's nf cos and nf sin factors, then one
    // word per probe (its record slot's address, or null)
    h->tab_dft_off.assign(h->dfts.size(), 0);
    int stride = 0;
    for (size_t m = 0; m < h->dfts.size(); ++m) {
        h->tab_dft_off[m] = stride;
        stride += 2 * h->dfts[m].nf;
    }
    h->tab_probe_off = stride;
    stride += (int)h->probes.size();
    h->tab_stride = stride;
    const long words = (long)stride * nsteps;
    if (words > h->tab_cap) {
        void* p;
        if (int rc = h->alloc(&p, (size_t)words * sizeof(double))) return rc;
        h->d_tab = (double*)p;
        h->tab_cap = words;
    }
    for (int st = 0; st < nsteps; ++st) {
        const long n = n0 + st;
        double* row = host_tab + (size_t)st * stride;
        for (size_t m = 0; m < h->dfts.size(); ++m) {
            const Dft& d = h->dfts[m];
            dft_phases(h, d, n, row + h->tab_dft_off[m], row + h->tab_dft_off[m] + d.nf);
        }
        for (size_t q = 0; q < h->probes.size(); ++q) {
            Probe& p = h->probes[q];
            double* slot = nullptr;
            if (n % p.save_every == 0 && p.count < p.cap) slot = p.d_buf + p.count++;
            std::memcpy(row + h->tab_probe_off + q, &slot, sizeof(slot));
        }
    }
    if (words > 0)
        HIPCHK(hipMemcpyAsync(h->d_tab, host_tab, (size_t)words * sizeof(double), hipMemcpyHostToDevice, h->stream));
    std::vector<long> key = {nsteps,
                             h->mon_interval > 0 ? n0 % h->mon_interval : -1,
                             h->prec == FDTD_FP16 ? n0 % h->fp16_interval : -1,
                             h->mon_interval,
                             h->prec == FDTD_FP16 ? h->fp16_interval : 0,
                             h->kwid_h, h->kwid_e, h->force64, h->erev, h->mon_fused_h, (long)h->band_tiles, h->cur, h->mode, h->flow_lag,
                             (long)h->probes.size(), (long)h->dfts.size(), h->nsrc,
                             (long)(uintptr_t)h->srcv_cur, (long)(uintptr_t)h->mon_cur, (long)(uintptr_t)h->d_tab,
                             (long)(uintptr_t)h->F2[h->cur ^ 1][0], (long)(uintptr_t)h->d_rec, h->rec_host};
    hipGraphExec_t exec = nullptr;
    for (auto& g : h->graphs)
        if (g.key == key) exec = g.exec;
    if (!exec) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        if (!mon_steps.empty()) {
            const long mw = (long)(mon_steps.size() * sizeof(MonOut) / sizeof(unsigned long long));
            hipLaunchKernelGGL(fdtdk::k_zero_words, dim3((unsigned)std::min<long>((mw + 255) / 256, 64)), dim3(256), 0,
                               h->stream, (unsigned long long*)h->mon_cur, mw);
        }
        int slot = 0;
        for (int st = 0; st < nsteps; ++st) {
            const long n = n0 + st;
            h->tab_step = h->d_tab + (size_t)st * stride;
            dispatch_step(h, n, st, nullptr);
            if (h->mon_interval > 0 && n % h->mon_interval == 0) {
                MonOut* rec = h->mon_cur + slot++;
                if (!(st + 1 < nsteps && h->mon_fused_h && dispatch_monitor_begin(h, rec))) dispatch_monitor(h, rec);
            }
            if (h->prec == FDTD_FP16 && n % h->fp16_interval == 0) fp16_maintain(h, h->stream);
        }
        h->tab_step = nullptr;
        if (h->rec_host) enqueue_record_copy(h, mon_steps.size());
        const hipError_t e = hipStreamEndCapture(h->stream, &graph);
        if (e != hipSuccess) return fail(FDTD_EDEVICE, std::string("step graph capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(FDTD_EDEVICE, std::string("step graph: ") + hipGetErrorString(ei));
        if (h->graphs.size() >= 8) {  // keep the most recent shapes
            (void)hipGraphExecDestroy(h->graphs.front().exec);
            h->graphs.erase(h->graphs.begin());
        }
        h->graphs.push_back(fdtd_sim::StepGraph{key, exec});
    }
    HIPCHK(hipGraphLaunch(exec, h->stream));
    return FDTD_OK;
}

// Enqueues nsteps steps on the handle's stream (and, with ms, the timing
// events); the monitor records are read by step_finish.
// ev (optional) receives HIP events around the H and E launches of every
// `every`-th step (steps 0, every, 2*every, ... of the call), 4 per timed step,
// plus one at each end of the call.
static int step_launch(fdtd_handle h, long n0, int nsteps, const double* src_values, std::vector<hipEvent_t>* ev,
                       int every = 1, bool noalloc = false) {
    if (h) ++h->gen;
    if (!h || nsteps < 0 || n0 < 0) return fail(FDTD_EINVAL, "bad step arguments");
    if (h->batch) return fail(FDTD_EINVAL, "domain is a batch member: step it through its batch");
    if (int rc = step_finish(h)) return rc;  // the monitor buffer is reused
    if (nsteps == 0) return FDTD_OK;
    if (h->nsrc > 0 && !src_values) return fail(FDTD_EINVAL, "src_values required");
    const long need = (long)nsteps * h->nsrc;
    // a recorded call: the two-kernel step without timing events or TLS
    // diagnostics (whose launches depend on more than the step's phase)
    const bool use_graph = h->graph_ticks && !noalloc && !ev && h->mode == 0 && h->tls_every == 0;
    // monitor slots for this call
    std::vector<long> mon_steps;
    if (h->mon_interval > 0)
        for (int s = 0; s < nsteps; ++s)
            if ((n0 + s) % h->mon_interval == 0) mon_steps.push_back(n0 + s);
    if (noalloc) {
        // records first (256-byte aligned), then the source values
        const size_t rec = ((mon_steps.size() * sizeof(MonOut)) + 255) / 256 * 256;
        const size_t total = rec + (size_t)need * sizeof(double);
        if (total > FDTD_NOALLOC_SCRATCH_BYTES)
            return fail(FDTD_EINVAL, "fdtd_step_noalloc: the call needs " + std::to_string(total) +
                                         " B of scratch, above the " + std::to_string(FDTD_NOALLOC_SCRATCH_BYTES) +
                                         " B maximum: split it into shorter calls");
        unsigned char* base = nullptr;
        if (int rc = scratch_address(h->device, &base)) return rc;
        h->mon_cur = (MonOut*)base;
        h->srcv_cur = (double*)(base + rec);
        if (!mon_steps.empty()) {
            const long words = (long)(mon_steps.size() * sizeof(MonOut) / sizeof(unsigned long long));
            hipLaunchKernelGGL(fdtdk::k_zero_words, dim3((unsigned)std::min<long>((words + 255) / 256, 64)), dim3(256), 0,
                               h->stream, (unsigned long long*)h->mon_cur, words);
        }
    } else {
        if (need > h->srcv_cap) {
            void* p;
            if (int rc = h->alloc(&p, (size_t)need * sizeof(double))) return rc;
            h->d_srcv = (double*)p;
            h->srcv_cap = need;
        }
        if ((int)mon_steps.size() > h->mon_cap) {
            void* p;
            const int cap = std::max<int>((int)mon_steps.size(), 64);
            if (int rc = h->alloc(&p, (size_t)cap * sizeof(MonOut))) return rc;
            h->d_mon = (MonOut*)p;
            h->mon_cap = cap;
        }
        h->srcv_cur = h->d_srcv;
        h->mon_cur = h->d_mon;
        if (!mon_steps.empty() && !use_graph)
            HIPCHK(hipMemsetAsync(h->d_mon, 0, mon_steps.size() * sizeof(MonOut), h->stream));
    }
    // the table's words (graph calls): per step, each DFT's 2 nf phase
    // factors and one word per probe
    long tab_words = 0;
    if (use_graph) {
        long row = (long)h->probes.size();
        for (const Dft& d : h->dfts) row += 2 * d.nf;
        tab_words = row * nsteps;
    }
    const double* up = src_values;
    if (!noalloc && need + tab_words > 0) {
        if ((size_t)(need + tab_words) > h->pin_words) {
            if (h->h_pin) {
                HIPCHK(hipStreamSynchronize(h->stream));  // an earlier upload may still read it
                pool_host_free(PIN_PLAIN, h->h_pin, h->pin_words * sizeof(double));
            }
            h->h_pin = nullptr;
            h->pin_words = 0;
            const size_t words = std::max<size_t>((size_t)(need + tab_words), 4096);
            size_t got = 0;
            if (pool_host_alloc(PIN_PLAIN, words * sizeof(double), (void**)&h->h_pin, &got) != hipSuccess)
                return fail(FDTD_ENOMEM, "hipHostMalloc of the upload staging failed");
            h->pin_words = got / sizeof(double);
        }
        if (need > 0) std::memcpy(h->h_pin, src_values, (size_t)need * sizeof(double));
        up = h->h_pin;
    }
    h->rec_host = !noalloc && !mon_steps.empty();
    if (h->rec_host)
        if (int rc = ensure_host_records(h, mon_steps.size())) return rc;
    if (need > 0)
        HIPCHK(hipMemcpyAsync(h->srcv_cur, up, (size_t)need * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (use_graph) {
        if (int rc = step_graph_launch(h, n0, nsteps, mon_steps, h->h_pin + need)) return rc;
        h->pending = true;
        h->pending_mon = std::move(mon_steps);
        return FDTD_OK;
    }
    int slot = 0;
    if (ev) {
        ev->resize((size_t)((nsteps + every - 1) / every) * 4 + 2);
        for (auto& e : *ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord((*ev)[0], h->stream));
    }
    for (int s = 0; s < nsteps; ++s) {
        const long n = n0 + s;
        dispatch_step(h, n, s, ev && s % every == 0 ? &(*ev)[1 + (size_t)(s / every) * 4] : nullptr);
        if (h->mon_interval > 0 && n % h->mon_interval == 0) {
            // the field maxima ride in the next step's H launch when it is in this call
            MonOut* rec = h->mon_cur + slot++;
            if (!(s + 1 < nsteps && h->mon_fused_h && dispatch_monitor_begin(h, rec))) dispatch_monitor(h, rec);
        }
        if (h->prec == FDTD_FP16 && n % h->fp16_interval == 0) fp16_maintain(h, h->stream);
    }
    if (h->rec_host) enqueue_record_copy(h, mon_steps.size());
    HIPCHK(hipGetLastError());
    if (ev) HIPCHK(hipEventRecord(ev->back(), h->stream));
    h->pending = true;
    h->pending_mon = std::move(mon_steps);
    return FDTD_OK;
}

static int step_impl(fdtd_handle h, long n0, int nsteps, const double* src_values, double* ms, int every = 1,
                     double* n_timed = nullptr) {
    std::vector<hipEvent_t> ev;
    if (int rc = step_launch(h, n0, nsteps, src_values, ms ? &ev : nullptr, every)) return rc;
    if (ms && !ev.empty()) {
        HIPCHK(hipEventSynchronize(ev.back()));
        double th = 0, te = 0;
        const int nt = (nsteps + every - 1) / every;
        for (int s = 0; s < nt; ++s) {
            float a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, ev[1 + (size_t)s * 4], ev[2 + (size_t)s * 4]));
            HIPCHK(hipEventElapsedTime(&b, ev[3 + (size_t)s * 4], ev[4 + (size_t)s * 4]));
            th += a;
            te += b;
        }
        float tot = 0;
        HIPCHK(hipEventElapsedTime(&tot, ev[0], ev.back()));
        ms[0] = th; ms[1] = te; ms[2] = tot;
        if (n_timed) *n_timed = nt;
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    return step_finish(h);
}

int fdtd_step(fdtd_handle h, long n0, int nsteps, const double* src_values) {
    return step_impl(h, n0, nsteps, src_values, nullptr);
}

int fdtd_step_noalloc(fdtd_handle h, long n0, int nsteps, const double* src_values) {
    if (int rc = step_launch(h, n0, nsteps, src_values, nullptr, 1, true)) return rc;
    return step_finish(h);
}

int fdtd_step_async(fdtd_handle h, long n0, int nsteps, const double* src_values) {
    return step_launch(h, n0, nsteps, src_values, nullptr);
}

int fdtd_step_query(fdtd_handle h) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    const hipError_t e = hipStreamQuery(h->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(FDTD_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int fdtd_step_wait(fdtd_handle h) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    if (int rc = step_finish(h)) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

int fdtd_step_timed(fdtd_handle h, long n0, int nsteps, const double* src_values, double* ms) {
    if (!ms) return fail(FDTD_EINVAL, "null timing output");
    return step_impl(h, n0, nsteps, src_values, ms);
}

int fdtd_step_timed_every(fdtd_handle h, long n0, int nsteps, const double* src_values, int every, double* ms) {
    if (!ms || every < 1) return fail(FDTD_EINVAL, "null timing output or stride below 1");
    return step_impl(h, n0, nsteps, src_values, ms, every, ms + 3);
}

int fdtd_synchronize(fdtd_handle h) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return FDTD_OK;
}

int fdtd_download(fdtd_handle h, int which, double* out) {
    if (!h || !out) return fail(FDTD_EINVAL, "null argument");
    float sc[36] = {};
    if (h->prec == FDTD_FP16)
        if (int rc = read_scales(h, sc)) return rc;
    if (which >= 0 && which < 6) return download_any(h, h->F[which], (size_t)h->N, sc[which], out);
    if (which >= 6 && which < 10) {
        std::memset(out, 0, (size_t)h->N * sizeof(double));
        if (!h->gain) return FDTD_OK;
        double* srcs[4] = {h->gNu, h->gNg, h->gPz, h->gPp};
        std::vector<double> box((size_t)h->gbox);
        HIPCHK(hipMemcpyAsync(box.data(), srcs[which - 6], box.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int gi = 0; gi < h->gn[0]; ++gi)
            for (int gj = 0; gj < h->gn[1]; ++gj)
                for (int gk = 0; gk < h->gn[2]; ++gk) {
                    const long g = ((long)gi * h->gn[1] + gj) * h->gn[2] + gk;
                    const long id = ((long)(gi + h->g0[0]) * h->ny + (gj + h->g0[1])) * h->nz + (gk + h->g0[2]);
                    out[id] = box[(size_t)g];
                }
        return FDTD_OK;
    }
    if (which >= 10 && which < 22) {
        const int q = which - 10;
        std::memset(out, 0, (size_t)h->N * sizeof(double));
        if (!h->psi[q]) return FDTD_OK;
        std::vector<double> slab((size_t)h->psi_len[q]);
        if (int rc = download_any(h, h->psi[q], slab.size(), sc[6 + q], slab.data())) return rc;
        for (int i = 0; i < h->nx; ++i)
            for (int j = 0; j < h->ny; ++j)
                for (int k = 0; k < h->nz; ++k) {
                    const long p = psi_slab_index(h, q, i, j, k);
                    if (p >= 0) out[((long)i * h->ny + j) * h->nz + k] = slab[(size_t)p];
                }
        return FDTD_OK;
    }
    return fail(FDTD_EINVAL, "unknown array id");
}

int fdtd_upload(fdtd_handle h, int which, const double* in) {
    if (!h || !in) return fail(FDTD_EINVAL, "null argument");
    if (which >= 0 && which < 6) {
        double scale = 1.0;
        if (h->prec == FDTD_FP16) {  // choose the component's scale from the data
            float sc[36];
            if (int rc = read_scales(h, sc)) return rc;
            double m = 0;
            for (long i = 0; i < h->N; ++i)  // non-finite values are stored as they are
                if (std::isfinite(in[i])) m = std::max(m, std::fabs(in[i]));
            if (m > 0) {
                int e;
                std::frexp(m, &e);
                sc[which] = std::ldexp(1.0f, e - 8);
                sc[18 + which] = std::ldexp(1.0f, 8 - e);
                if (int rc = write_scales(h, sc)) return rc;
            }
            scale = sc[which];
        }
        for (int b = 0; b < 2; ++b) {  // both buffers where the second exists
            if (!h->F2[b][which]) continue;
            if (int rc = upload_any(h, h->F2[b][which], in, (size_t)h->N, scale)) return rc;
        }
        return FDTD_OK;
    }
    if (which >= 6 && which < 10) {
        if (!h->gain) return FDTD_OK;
        double* dsts[4] = {h->gNu, h->gNg, h->gPz, h->gPp};
        std::vector<double> box((size_t)h->gbox);
        for (int gi = 0; gi < h->gn[0]; ++gi)
            for (int gj = 0; gj < h->gn[1]; ++gj)
                for (int gk = 0; gk < h->gn[2]; ++gk) {
                    const long g = ((long)gi * h->gn[1] + gj) * h->gn[2] + gk;
                    const long id = ((long)(gi + h->g0[0]) * h->ny + (gj + h->g0[1])) * h->nz + (gk + h->g0[2]);
                    box[(size_t)g] = in[id];
                }
        HIPCHK(hipMemcpyAsync(dsts[which - 6], box.data(), box.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return FDTD_OK;
    }
    return fail(FDTD_EINVAL, "unknown array id");
}

// The copies fdtd_stage_results enqueued are current (no stepping call since).
static bool staged(fdtd_handle h) {
    return h->h_res && h->res_gen == h->gen && h->res_probe_off.size() == h->probes.size() &&
           h->res_dft_off.size() == h->dfts.size();
}

int fdtd_stage_results(fdtd_handle h) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    if (h->batch) return fail(FDTD_EINVAL, "domain is a batch member");
    size_t words = 0;
    h->res_probe_off.clear();
    h->res_dft_off.clear();
    for (const Probe& p : h->probes) {
        h->res_probe_off.push_back(words);
        words += (size_t)p.count;
    }
    for (const Dft& d : h->dfts) {
        h->res_dft_off.push_back(words);
        words += (size_t)d.ncell * d.nf * 2;
    }
    if (words > h->res_words) {
        if (h->h_res) {
            HIPCHK(hipStreamSynchronize(h->stream));  // an earlier staging may still write the old copy
            pool_host_free(PIN_MAPPED, h->h_res, h->res_words * sizeof(double));
        }
        h->h_res = h->d_res = nullptr;
        h->res_words = 0;
        size_t got = 0;
        if (pool_host_alloc(PIN_MAPPED, words * sizeof(double), (void**)&h->h_res, &got) != hipSuccess)
            return fail(FDTD_ENOMEM, "hipHostMalloc of the result copy failed");
        HIPCHK(hipHostGetDevicePointer((void**)&h->d_res, h->h_res, 0));
        h->res_words = got / sizeof(double);
    }
    auto copy = [&](const double* src, size_t off, long n) {
        if (n <= 0) return;
        const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 64);
        hipLaunchKernelGGL(fdtdk::k_copy_f64, dim3(blocks), dim3(256), 0, h->stream, src, h->d_res + off, n);
    };
    for (size_t q = 0; q < h->probes.size(); ++q) copy(h->probes[q].d_buf, h->res_probe_off[q], h->probes[q].count);
    for (size_t m = 0; m < h->dfts.size(); ++m)
        copy(h->dfts[m].d_acc, h->res_dft_off[m], h->dfts[m].ncell * h->dfts[m].nf * 2);
    HIPCHK(hipGetLastError());
    h->res_gen = h->gen;
    return FDTD_OK;
}

int fdtd_read_probe(fdtd_handle h, int probe, double* out, long max, long* count) {
    if (!h || probe < 0 || probe >= (int)h->probes.size()) return fail(FDTD_EINVAL, "bad probe id");
    const Probe& p = h->probes[(size_t)probe];
    if (count) *count = p.count;
    const long n = std::min(max, p.count);
    if (n > 0 && out) {
        if (staged(h)) {  // the handle's own stream, then host memory: no copy queued behind other streams
            HIPCHK(hipStreamSynchronize(h->stream));
            std::memcpy(out, h->h_res + h->res_probe_off[(size_t)probe], (size_t)n * sizeof(double));
            return FDTD_OK;
        }
        HIPCHK(hipMemcpyAsync(out, p.d_buf, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return FDTD_OK;
}

int fdtd_read_dft(fdtd_handle h, int m, double* out) {
    if (!h || m < 0 || m >= (int)h->dfts.size() || !out) return fail(FDTD_EINVAL, "bad DFT id");
    const Dft& d = h->dfts[(size_t)m];
    const size_t n = (size_t)d.ncell * d.nf * 2;
    if (staged(h)) {
        HIPCHK(hipStreamSynchronize(h->stream));
        std::memcpy(out, h->h_res + h->res_dft_off[(size_t)m], n * sizeof(double));
        return FDTD_OK;
    }
    HIPCHK(hipMemcpyAsync(out, d.d_acc, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

int fdtd_is_stable(fdtd_handle h) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    return h->stable ? 1 : 0;
}

int fdtd_read_check(fdtd_handle h, fdtd_check_record* out, int max, int* count) {
    if (!h) return fail(FDTD_EINVAL, "null handle");
    const int n = (int)h->checks.size();
    if (count) *count = n;
    for (int q = 0; q < std::min(n, max); ++q) out[q] = h->checks[(size_t)q];
    return FDTD_OK;
}

int fdtd_info(fdtd_handle h, long* info) {
    if (!h || !info) return fail(FDTD_EINVAL, "null argument");
    info[0] = h->nx; info[1] = h->ny; info[2] = h->nz; info[3] = h->prec;
    info[4] = h->nmat; info[5] = h->mid16 ? 2 : 1; info[6] = h->ngain; info[7] = h->np;
    return FDTD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches (include/fdtd_hip.h, fdtd_batch_*).

struct fdtd_batch {
    int prec = 0;
    hipStream_t stream = nullptr;
    std::vector<fdtd_sim*> doms;
    // default tiling: one work-group per tile of 4 planes (256-domain batch:
    // 3.84 ms per step against 4.55-4.77 ms for persistent tilings)
    int rows = 4, wg_per_cu = 4, chunk = 0, planes = 4;
    int n_cu = 256;
    bool use_graph = false;
    // device tables (grown, never shrunk)
    std::vector<void*> allocs;
    void* d_args = nullptr;
    size_t args_cap = 0;
    int* d_cum = nullptr;
    size_t cum_cap = 0;
    fdtdk::BatchCtl* d_ctl = nullptr;
    fdtdk::BatchDom* d_dom = nullptr;
    size_t dom_cap = 0;
    double* d_srcv = nullptr;
    size_t srcv_cap = 0;
    fdtdk::BatchProbe* d_probes = nullptr;
    size_t probes_cap = 0;
    fdtdk::BatchDft* d_dfts = nullptr;
    size_t dfts_cap = 0;
    MonOut* d_mon = nullptr;
    size_t mon_cap = 0;
    // host copies
    std::vector<char> h_args;
    std::vector<int> h_cum;
    std::vector<fdtdk::BatchDom> h_dom;
    std::vector<fdtdk::BatchProbe> h_probes;
    std::vector<fdtdk::BatchDft> h_dfts;
    std::vector<double> h_srcv;
    std::vector<long> srcv_off;
    int slots = 0;
    // captured step block
    hipGraphExec_t gexec = nullptr;
    int g_nsteps = -1;
    unsigned long g_version = 0;
    unsigned long version = 1;  // bumped when membership, tuning or device tables change
    std::vector<int> g_phase;

    ~fdtd_batch() {
        if (stream) hipStreamSynchronize(stream);
        if (gexec) hipGraphExecDestroy(gexec);
        for (void* p : allocs) hipFree(p);
        if (stream) hipStreamDestroy(stream);
    }
    // (re)allocates a device table of at least `need` bytes
    int ensure(void** p, size_t* cap, size_t need) {
        if (need <= *cap && *p) return 0;
        size_t n = std::max<size_t>(need, 256);
        n = std::max(n, *cap * 2);
        void* q;
        hipError_t e = hipMalloc(&q, n);
        if (e != hipSuccess) return fail(FDTD_ENOMEM, "batch table allocation failed");
        allocs.push_back(q);
        *p = q;
        *cap = n;
        ++version;
        return 0;
    }
};

namespace {
// Wavefront k-width of the batch passes.  Persistent work-groups: 64
// consecutive k of blockDim.y rows; 32- and 16-wide wavefronts (two and four
// rows each) fill the short rows of small domains better but measured slower
// on the 256-domain workload (26.5-27.1 and 20.8-22.5 Gcell/s against
// 28.3-28.4 at 64; same box, graph blocks).  One work-group per tile: 32
// wide, 3.73 against 3.84 ms per step (same box, 4 rows x 4 planes).
constexpr int BATCH_KW = 64;
constexpr int BATCH_KW_TILES = 32;
}  // namespace

namespace {

template <typename T, typename R>
int batch_tables(fdtd_batch* b, int nsteps, const long* n0, const int* msteps) {
    const int nd = (int)b->doms.size();
    b->h_args.resize(sizeof(StepArgs<T, R>) * nd);
    StepArgs<T, R>* args = reinterpret_cast<StepArgs<T, R>*>(b->h_args.data());
    if (int rc = b->ensure(&b->d_args, &b->args_cap, b->h_args.size())) return rc;
    if (!b->d_ctl) {
        size_t cap = 0;
        if (int rc = b->ensure((void**)&b->d_ctl, &cap, sizeof(fdtdk::BatchCtl))) return rc;
    }
    // source values: domain d at srcv_off[d], [nsteps][nsrc]
    b->srcv_off.assign(nd, 0);
    size_t nsv = 0;
    for (int d = 0; d < nd; ++d) {
        b->srcv_off[d] = (long)nsv;
        nsv += (size_t)nsteps * b->doms[d]->nsrc;
    }
    if (int rc = b->ensure((void**)&b->d_srcv, &b->srcv_cap, std::max<size_t>(nsv, 1) * sizeof(double))) return rc;
    b->h_cum.assign(nd + 1, 0);
    b->h_dom.assign(nd, fdtdk::BatchDom{});
    b->h_probes.clear();
    b->h_dfts.clear();
    int min_iv = 0;
    for (int d = 0; d < nd; ++d) {
        fdtd_sim* s = b->doms[d];
        args[d] = make_args<T, R>(s);
        args[d].srcv = b->d_srcv + b->srcv_off[d];
        // pointers travel as offsets from the table (the kernel adds it back)
        fdtdk::shift_args(args[d], (const char*)b->d_ctl, -1);
        const int kw = b->chunk ? BATCH_KW : BATCH_KW_TILES, rows = b->rows * (64 / kw);
        const unsigned nkt = (unsigned)((s->nz + kw - 1) / kw), njt = (unsigned)((s->ny + rows - 1) / rows);
        args[d].tk = fdtdk::make_fastdiv(nkt);
        args[d].tj = fdtdk::make_fastdiv(njt);
        const long tiles = (long)((s->nx + b->planes - 1) / b->planes) * njt * nkt;
        if ((long)b->h_cum[d] + tiles > 0x7fffffffL) return fail(FDTD_EINVAL, "batch too large (tile count)");
        b->h_cum[d + 1] = b->h_cum[d] + (int)tiles;
        fdtdk::BatchDom& m = b->h_dom[d];
        m.n0 = n0[d];
        m.steps = msteps[d];
        m.N = s->N;
        m.ngain = s->gain ? s->gbox : 0;
        m.interval = s->mon_interval;
        m.mon_first = m.interval > 0 ? (n0[d] + m.interval - 1) / m.interval : 0;
        if (m.interval > 0) min_iv = min_iv ? std::min(min_iv, m.interval) : m.interval;
        for (const Probe& p : s->probes)
            b->h_probes.push_back(fdtdk::BatchProbe{d, p.comp, p.id, (long)p.save_every, p.cap, p.d_buf});
        for (const Dft& q : s->dfts) {
            fdtdk::BatchDft m2{};
            m2.dom = d;
            m2.d.comp = q.comp;
            m2.d.i0 = q.box[0]; m2.d.j0 = q.box[2]; m2.d.k0 = q.box[4];
            m2.d.ny_b = q.box[3] - q.box[2];
            m2.d.nz_b = q.box[5] - q.box[4];
            m2.d.ncell = q.ncell;
            m2.d.nf = q.nf;
            m2.d.acc = q.d_acc;
            m2.dt = s->dt;
            for (int f = 0; f < q.nf; ++f) m2.omega[f] = 2.0 * 3.14159265358979323846 * q.freqs[f];
            b->h_dfts.push_back(m2);
        }
    }
    b->slots = min_iv > 0 ? nsteps / min_iv + 2 : 0;
    int rc;
    if ((rc = b->ensure((void**)&b->d_cum, &b->cum_cap, b->h_cum.size() * sizeof(int)))) return rc;
    if ((rc = b->ensure((void**)&b->d_dom, &b->dom_cap, b->h_dom.size() * sizeof(fdtdk::BatchDom)))) return rc;
    if ((rc = b->ensure((void**)&b->d_probes, &b->probes_cap, b->h_probes.size() * sizeof(fdtdk::BatchProbe)))) return rc;
    if ((rc = b->ensure((void**)&b->d_dfts, &b->dfts_cap, b->h_dfts.size() * sizeof(fdtdk::BatchDft)))) return rc;
    if ((rc = b->ensure((void**)&b->d_mon, &b->mon_cap, (size_t)nd * std::max(b->slots, 1) * sizeof(MonOut)))) return rc;
    return 0;
}

// The launches of one batch step (also what a graph captures).
template <typename T, typename R>
void batch_launch_step(fdtd_batch* b, int s, const long* n0, const int* msteps) {
    const int nd = (int)b->doms.size();
    const StepArgs<T, R>* args = reinterpret_cast<const StepArgs<T, R>*>(b->d_args);
    hipLaunchKernelGGL(fdtdk::k_batch_begin, dim3(1), dim3(64), 0, b->stream, b->d_ctl);
    const dim3 blk(64, (unsigned)b->rows, 1);
    const int total = b->h_cum[nd];
    const int nwg = b->chunk ? std::max(1, std::min(b->n_cu * b->wg_per_cu * (4 / b->rows), (total + b->chunk - 1) / b->chunk))
                             : 0;
    const char* anchor = (const char*)b->d_ctl;
    // 32-bit byte offsets where every member's arrays are below 4 GiB
    // (256-domain workload: 4.13 -> 4.01-4.06 ms per step), unless a member
    // asks for 64-bit offsets (fdtd_set_kernel_mode, mode 0, + 131072)
    bool o32 = true;
    for (int d = 0; d < nd && o32; ++d) o32 = b->doms[d]->fits32 && !b->doms[d]->force64;
    if (b->chunk == 0) {
        // one work-group per tile: the grid is the tile list, padded to eighths
        const unsigned per = (unsigned)((total + fdtdk::BATCH_SEGS - 1) / fdtdk::BATCH_SEGS), grd = per * fdtdk::BATCH_SEGS;
        if (o32) {
            hipLaunchKernelGGL((fdtdk::k_batch_tiles<T, R, false, BATCH_KW_TILES, unsigned>), dim3(grd), blk, 0, b->stream,
                               args, anchor, b->d_cum, b->d_dom, nd, b->planes, b->d_ctl);
            hipLaunchKernelGGL((fdtdk::k_batch_tiles<T, R, true, BATCH_KW_TILES, unsigned>), dim3(grd), blk, 0, b->stream,
                               args, anchor, b->d_cum, b->d_dom, nd, b->planes, b->d_ctl);
        } else {
            hipLaunchKernelGGL((fdtdk::k_batch_tiles<T, R, false, BATCH_KW_TILES, long>), dim3(grd), blk, 0, b->stream,
                               args, anchor, b->d_cum, b->d_dom, nd, b->planes, b->d_ctl);
            hipLaunchKernelGGL((fdtdk::k_batch_tiles<T, R, true, BATCH_KW_TILES, long>), dim3(grd), blk, 0, b->stream,
                               args, anchor, b->d_cum, b->d_dom, nd, b->planes, b->d_ctl);
        }
    } else if (o32) {
        hipLaunchKernelGGL((fdtdk::k_batch_step<T, R, false, BATCH_KW, unsigned>), dim3((unsigned)nwg), blk, 0,
                           b->stream, args, anchor, b->d_cum, b->d_dom, nd, b->chunk, b->planes, b->d_ctl);
        hipLaunchKernelGGL((fdtdk::k_batch_step<T, R, true, BATCH_KW, unsigned>), dim3((unsigned)nwg), blk, 0,
                           b->stream, args, anchor, b->d_cum, b->d_dom, nd, b->chunk, b->planes, b->d_ctl);
    } else {
        hipLaunchKernelGGL((fdtdk::k_batch_step<T, R, false, BATCH_KW, long>), dim3((unsigned)nwg), blk, 0,
                           b->stream, args, anchor, b->d_cum, b->d_dom, nd, b->chunk, b->planes, b->d_ctl);
        hipLaunchKernelGGL((fdtdk::k_batch_step<T, R, true, BATCH_KW, long>), dim3((unsigned)nwg), blk, 0,
                           b->stream, args, anchor, b->d_cum, b->d_dom, nd, b->chunk, b->planes, b->d_ctl);
    }
    if (!b->h_probes.empty())
        hipLaunchKernelGGL((fdtdk::k_batch_probes<T, R>), dim3((unsigned)((b->h_probes.size() + 63) / 64)), dim3(64), 0,
                           b->stream, args, anchor, b->d_dom, b->d_probes, (int)b->h_probes.size(), b->d_ctl);
    if (!b->h_dfts.empty()) {
        long mx = 1;
        for (const auto& m : b->h_dfts) mx = std::max(mx, m.d.ncell);
        hipLaunchKernelGGL((fdtdk::k_batch_dft<T, R>), dim3((unsigned)std::min<long>((mx + 255) / 256, 256),
                           (unsigned)b->h_dfts.size()), dim3(256), 0, b->stream, args, anchor, b->d_dom, b->d_dfts,
                           b->d_ctl);
    }
    // the monitor kernel itself skips members that are not on a monitor step
    // or have stopped, so the launch depends on the step phases only
    bool mon = false;
    for (int d = 0; d < nd && !mon; ++d) {
        const int iv = b->doms[d]->mon_interval;
        mon = iv > 0 && (n0[d] + s) % iv == 0;
    }
    if (mon && b->slots > 0)
        hipLaunchKernelGGL((fdtdk::k_batch_monitor<T, R>), dim3(64, (unsigned)nd), dim3(256), 0, b->stream, args,
                           anchor, b->d_dom, b->d_mon, b->slots, b->d_ctl);
    if (sizeof(T) == 2)
        for (int d = 0; d < nd; ++d)
            if (s < msteps[d] && (n0[d] + s) % b->doms[d]->fp16_interval == 0) fp16_maintain(b->doms[d], b->stream);
}

template <typename T, typename R>
int batch_step(fdtd_batch* b, int nsteps, const long* n0, const int* msteps, const double* const* srcv, double* ms) {
    const int nd = (int)b->doms.size();
    if (int rc = batch_tables<T, R>(b, nsteps, n0, msteps)) return rc;
    for (int d = 0; d < nd; ++d) {
        fdtd_sim* s = b->doms[d];
        if (s->nsrc > 0 && (!srcv || !srcv[d])) return fail(FDTD_EINVAL, "source values missing for a batch member");
    }
    // uploads for this call (outside any graph)
    b->h_srcv.resize(std::max<size_t>(1, (size_t)b->srcv_off.back() + (size_t)nsteps * b->doms.back()->nsrc));
    for (int d = 0; d < nd; ++d)
        if (b->doms[d]->nsrc > 0)
            std::memcpy(b->h_srcv.data() + b->srcv_off[d], srcv[d], (size_t)msteps[d] * b->doms[d]->nsrc * sizeof(double));
    HIPCHK(hipMemcpyAsync(b->d_srcv, b->h_srcv.data(), b->h_srcv.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_args, b->h_args.data(), b->h_args.size(), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_cum, b->h_cum.data(), b->h_cum.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_dom, b->h_dom.data(), b->h_dom.size() * sizeof(fdtdk::BatchDom), hipMemcpyHostToDevice,
                          b->stream));
    if (!b->h_probes.empty())
        HIPCHK(hipMemcpyAsync(b->d_probes, b->h_probes.data(), b->h_probes.size() * sizeof(fdtdk::BatchProbe),
                              hipMemcpyHostToDevice, b->stream));
    if (!b->h_dfts.empty())
        HIPCHK(hipMemcpyAsync(b->d_dfts, b->h_dfts.data(), b->h_dfts.size() * sizeof(fdtdk::BatchDft),
                              hipMemcpyHostToDevice, b->stream));
    if (b->slots > 0)
        HIPCHK(hipMemsetAsync(b->d_mon, 0, (size_t)nd * b->slots * sizeof(MonOut), b->stream));
    hipLaunchKernelGGL(fdtdk::k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->d_ctl);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, b->stream));
    }
    // The monitor and FP16 rescaling run on each member's own cadence: a
    // captured block is valid for the same step phases only.  Members that
    // stop inside the block idle on the device (their step count is in the
    // domain table), so the count enters the key only where the host decides
    // per member (FP16 rescaling).
    std::vector<int> phase;
    for (int d = 0; d < nd; ++d) {
        const fdtd_sim* m = b->doms[d];
        phase.push_back(m->mon_interval > 0 ? (int)(n0[d] % m->mon_interval) : -1);
        if (b->prec == FDTD_FP16) {
            phase.push_back(msteps[d]);
            phase.push_back((int)(n0[d] % m->fp16_interval));
        }
    }
    if (b->use_graph) {
        if (!b->gexec || b->g_nsteps != nsteps || b->g_version != b->version || b->g_phase != phase) {
            if (b->gexec) {
                (void)hipGraphExecDestroy(b->gexec);
                b->gexec = nullptr;
            }
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            for (int s = 0; s < nsteps; ++s) batch_launch_step<T, R>(b, s, n0, msteps);
            HIPCHK(hipStreamEndCapture(b->stream, &g));
            const hipError_t e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(FDTD_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            b->g_nsteps = nsteps;
            b->g_version = b->version;
            b->g_phase = phase;
        }
        HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    } else {
        for (int s = 0; s < nsteps; ++s) batch_launch_step<T, R>(b, s, n0, msteps);
    }
    HIPCHK(hipGetLastError());
    if (ms) {
        HIPCHK(hipEventRecord(e1, b->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    // per-member bookkeeping: probe counts, monitor records and policy
    std::vector<MonOut> mo((size_t)nd * std::max(b->slots, 1));
    if (b->slots > 0)
        HIPCHK(hipMemcpyAsync(mo.data(), b->d_mon, (size_t)nd * b->slots * sizeof(MonOut), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    bool ok = true;
    for (int d = 0; d < nd; ++d) {
        fdtd_sim* s = b->doms[d];
        const long n_end = n0[d] + msteps[d];
        if (msteps[d] == 0) continue;
        for (Probe& p : s->probes) p.count = std::min(p.cap, (n_end - 1) / p.save_every + 1);
        s->checks.clear();
        const fdtdk::BatchDom& m = b->h_dom[d];
        if (m.interval <= 0) continue;
        for (long n = m.mon_first * m.interval; n < n_end; n += m.interval) {
            const long slot = n / m.interval - m.mon_first;
            if (slot >= b->slots) break;
            const MonOut& q = mo[(size_t)d * b->slots + slot];
            fdtd_check_record r{};
            const unsigned long long inf = 0x7ff0000000000000ull;
            r.step = n;
            r.nan_E = q.maxE >= inf; r.nan_H = q.maxH >= inf;
            r.nan_P = q.maxP >= inf; r.nan_N = q.nanN >= inf;
            r.max_E = bits_to_double(q.maxE);
            r.max_H = bits_to_double(q.maxH);
            r.max_P = bits_to_double(q.maxP);
            r.pop_total = q.pop_total;
            r.pop_initial = q.pop_init;
            s->checks.push_back(r);
            if (!apply_policy(s, r)) ok = false;
        }
    }
    return ok ? FDTD_OK : fail(FDTD_EUNSTABLE, "a batch member is unstable");
}

}  // namespace

extern "C" {

int fdtd_batch_create(int precision, fdtd_batch_handle* out) {
    if (!out || precision < FDTD_FP64 || precision > FDTD_FP16) return fail(FDTD_EINVAL, "bad batch arguments");
    *out = nullptr;
    auto* b = new fdtd_batch();
    b->prec = precision;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(FDTD_EDEVICE, "no HIP device for the batch");
    }
    *out = b;
    return FDTD_OK;
}

int fdtd_batch_destroy(fdtd_batch_handle b) {
    if (!b) return FDTD_OK;
    for (fdtd_sim* s : b->doms) s->batch = nullptr;
    delete b;
    return FDTD_OK;
}

int fdtd_batch_add(fdtd_batch_handle b, fdtd_handle h) {
    if (!b || !h) return fail(FDTD_EINVAL, "null argument");
    if (h->batch) return fail(FDTD_EINVAL, "domain already belongs to a batch");
    if (h->prec != b->prec) return fail(FDTD_EINVAL, "domain precision differs from the batch's");
    // the batch steps in place in whichever buffer holds the current state
    if (int rc = step_finish(h)) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->batch = b;
    b->doms.push_back(h);
    ++b->version;
    return (int)b->doms.size() - 1;
}

int fdtd_batch_remove(fdtd_batch_handle b, fdtd_handle h) {
    if (!b || !h) return fail(FDTD_EINVAL, "null argument");
    auto it = std::find(b->doms.begin(), b->doms.end(), h);
    if (it == b->doms.end()) return fail(FDTD_EINVAL, "domain is not a member of this batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->doms.erase(it);
    h->batch = nullptr;
    ++b->version;
    return FDTD_OK;
}

int fdtd_batch_count(fdtd_batch_handle b) { return b ? (int)b->doms.size() : fail(FDTD_EINVAL, "null batch"); }

int fdtd_batch_set_tuning(fdtd_batch_handle b, const int* tuning) {
    if (!b || !tuning) return fail(FDTD_EINVAL, "null argument");
    const int rows = tuning[0], wg_per_cu = tuning[1], chunk = tuning[2], planes = tuning[3];
    if ((rows != 1 && rows != 2 && rows != 4) || wg_per_cu < 1 || wg_per_cu > 16 || chunk < 0 || chunk > 64 ||
        planes < 1 || planes > 16)
        return fail(FDTD_EINVAL, "tuning out of range");
    b->rows = rows;
    b->wg_per_cu = wg_per_cu;
    b->chunk = chunk;
    b->planes = planes;
    ++b->version;
    return FDTD_OK;
}

int fdtd_batch_get_tuning(fdtd_batch_handle b, int* tuning) {
    if (!b || !tuning) return fail(FDTD_EINVAL, "null argument");
    tuning[0] = b->rows;
    tuning[1] = b->wg_per_cu;
    tuning[2] = b->chunk;
    tuning[3] = b->planes;
    return FDTD_OK;
}

int fdtd_batch_set_graph(fdtd_batch_handle b, int on) {
    if (!b) return fail(FDTD_EINVAL, "null batch");
    b->use_graph = on != 0;
    return FDTD_OK;
}

int fdtd_batch_step(fdtd_batch_handle b, int nsteps, const long* n0, const int* member_steps,
                    const double* const* src_values, double* ms) {
    if (!b || nsteps < 0 || (!b->doms.empty() && !n0)) return fail(FDTD_EINVAL, "bad batch step arguments");
    if (nsteps == 0 || b->doms.empty()) {
        if (ms) *ms = 0;
        return FDTD_OK;
    }
    std::vector<int> ms_(b->doms.size(), nsteps);
    for (fdtd_sim* m : b->doms) ++m->gen;
    for (size_t d = 0; d < b->doms.size(); ++d) {
        if (n0[d] < 0) return fail(FDTD_EINVAL, "negative step index");
        if (member_steps) {
            if (member_steps[d] < 0 || member_steps[d] > nsteps) return fail(FDTD_EINVAL, "member step count out of range");
            ms_[d] = member_steps[d];
        }
    }
    switch (b->prec) {
    case FDTD_FP64: return batch_step<double, double>(b, nsteps, n0, ms_.data(), src_values, ms);
    case FDTD_FP32: return batch_step<float, float>(b, nsteps, n0, ms_.data(), src_values, ms);
    default: return batch_step<uint16_t, float>(b, nsteps, n0, ms_.data(), src_values, ms);
    }
}

}  // extern "C"

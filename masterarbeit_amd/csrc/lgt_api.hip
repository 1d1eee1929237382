// lgt_api.hip -- C ABI of the engine (include/lgt_su2.h): handle, device
// memory, stream, launch sequencing and host transfers.  No kernels here;
// they are instantiated per precision in lgt_f64.hip / lgt_f32.hip.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <new>
#include <algorithm>
#include <map>
#include <thread>
#include <vector>

#include "../../include/lgt_su2.h"
#include "launch.hpp"

using namespace lgt;

namespace {

// message of the last failed call on this host thread (a fixed buffer: the
// allocation-free entry point reports through it too)
thread_local char g_last_error[512];

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t err, const char* what)
{
    return fail(LGT_EDEVICE, "%s: %s", what, hipGetErrorString(err));
}

#define HIP_TRY(call, what)                              \
    do {                                                 \
        hipError_t _e = (call);                          \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

#define TRY(x)                     \
    do {                           \
        int _rc = (x);             \
        if (_rc != LGT_OK) return _rc; \
    } while (0)

struct ProfEvent {
    int cls;
    hipEvent_t a, b;
};

// device state of the APE adjoint: checkpoints, one segment of recomputed
// steps with their |M|, two cotangent buffers and the objective's gradient
struct VjpState {
    int mode = 0, nsteps = 0, seg = 0, nck = 0;
    double sp = 0.0;
    bool fwd = false, cot = false;
    bool top_saved = false;   // segment buffers hold the forward's last segment
    size_t fbytes = 0, nbytes = 0, gbytes = 0;   // one field / norm / cotangent buffer of the layout
    std::vector<void*> ck, ckM, segU, segM;
    std::vector<RawPlanes> ckP, segUP;
    std::vector<RawNorms> ckMN, segMN;   // ckMN[c]: |M| of the step before checkpoint c (c >= 1)
    void* g[2] = {nullptr, nullptr};
    RawPlanes gP[2];
    void* fixed = nullptr;   // temporal mode: the spatial links, which the smearing leaves alone
    RawPlanes fixedP;
    void* gobj = nullptr;
    RawPlanes gobjP;
    void* lines = nullptr;   // straight-line fields of the loop gradient
    size_t lines_bytes = 0;
};

}  // namespace

namespace lgt {
int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace lgt

struct lgt_engine {
    int dev = 0;
    int L[4] = {0, 0, 0, 0};
    int prec = LGT_FP64;
    uint64_t seed = 0;
    uint32_t sweep = 0;
    int64_t vol = 0;
    Geo geo;
    size_t link_bytes = 0;
    hipStream_t stream = nullptr;
    void* field = nullptr;
    void* field_alt = nullptr;
    void* wline = nullptr;      // Wilson loops: the second spatial-line field (directions 1..3)
    RawPlanes wl2{};
    RawPlanes cur, alt;
    void* partial = nullptr;
    size_t partial_bytes = 0;
    void* dres = nullptr;
    size_t dres_bytes = 0;
    double* hres = nullptr;
    size_t hres_bytes = 0;
    void* pin[2] = {nullptr, nullptr};           // pinned chunks of the host <-> device copies
    hipEvent_t pin_ev[2] = {nullptr, nullptr};   // last DMA that used each chunk
    void* staging = nullptr;
    size_t staging_bytes = 0;
    void* link_buf = nullptr;   // per-link calls: staple / new link (8 doubles), status (2 ints)
    size_t link_buf_bytes = 0;
    void* uni_buf = nullptr;    // per-link heatbath: the injected uniforms
    size_t uni_buf_bytes = 0;
    void* qline = nullptr;      // improved charge: lines of 2 and 3 links (two fields)
    size_t qline_bytes = 0;
    RawPlanes qL2, qL3;
    int profile = 0;
    int order = -1;   // tile order of every kernel; -1: per kernel (see kgeo)
    int group = 0;    // LGT_OPT_GROUP (0: per kernel, see kgeo)
    int alternate = 1;  // chain launches alternate between forward and reversed tile order
    uint32_t flip = 0;  // reverses the next chain launch (see ugeo)
    std::vector<ProfEvent> events;
    double prof_ms[LGT_K_COUNT] = {0};
    int64_t prof_n[LGT_K_COUNT] = {0};
    int64_t dev_bytes = 0;
    VjpState vjp;
};

namespace {

int set_dev(lgt_engine* e)
{
    HIP_TRY(hipSetDevice(e->dev), "hipSetDevice");
    return LGT_OK;
}

int ensure(lgt_engine* e, void** buf, size_t* have, size_t need, const char* what)
{
    if (*have >= need) return LGT_OK;
    if (*buf) {
        hipFree(*buf);
        e->dev_bytes -= (int64_t)*have;
        *buf = nullptr;
        *have = 0;
    }
    hipError_t err = hipMalloc(buf, need);
    if (err != hipSuccess) {
        *buf = nullptr;
        return fail(LGT_ENOMEM, "hipMalloc(%zu) for %s: %s", need, what, hipGetErrorString(err));
    }
    *have = need;
    e->dev_bytes += (int64_t)need;
    return LGT_OK;
}

int ensure_host(lgt_engine* e, size_t need)
{
    if (e->hres_bytes >= need) return LGT_OK;
    if (e->hres) hipHostFree(e->hres);
    e->hres = nullptr;
    e->hres_bytes = 0;
    hipError_t err = hipHostMalloc((void**)&e->hres, need, hipHostMallocDefault);
    if (err != hipSuccess) return fail(LGT_ENOMEM, "hipHostMalloc(%zu): %s", need, hipGetErrorString(err));
    e->hres_bytes = need;
    return LGT_OK;
}

// Host <-> device copies of whole host arrays (fields, gradients, densities).
// A copy from or to pageable memory ran at 10-29 GB/s; here the array moves in
// PIN_CHUNK pieces through two pinned buffers: host threads copy one piece
// while the DMA engine moves the other, so the rate is set by the host memory
// copy across up to PIN_THREADS cores and the link.
constexpr size_t PIN_CHUNK = (size_t)32 << 20;
constexpr unsigned PIN_THREADS = 8;

void par_copy(void* dst, const void* src, size_t n)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<size_t>({(size_t)PIN_THREADS, (size_t)hw, n / ((size_t)1 << 20) + 1});
    if (nt <= 1) {
        memcpy(dst, src, n);
        return;
    }
    const size_t part = (n / nt + 63) & ~(size_t)63;
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; t++) {
        // the last thread also takes what the rounded-down parts leave
        const size_t off = std::min(n, part * t), len = (t + 1 == nt) ? n - off : std::min(n - off, part);
        if (!len) continue;
        try {
            pool.emplace_back([=] { memcpy((char*)dst + off, (const char*)src + off, len); });
        } catch (...) {   // no thread to be had: this part in the calling thread
            memcpy((char*)dst + off, (const char*)src + off, len);
        }
    }
    memcpy(dst, src, std::min(n, part));
    for (auto& th : pool) th.join();
}

int ensure_pinned(lgt_engine* e)
{
    for (int b = 0; b < 2; b++) {
        if (!e->pin[b]) HIP_TRY(hipHostMalloc(&e->pin[b], PIN_CHUNK, hipHostMallocDefault), "hipHostMalloc");
        if (!e->pin_ev[b]) HIP_TRY(hipEventCreateWithFlags(&e->pin_ev[b], hipEventDisableTiming), "hipEventCreate");
    }
    return LGT_OK;
}

// enqueues host -> device on the engine's stream; returns when the host array
// has been read (the last DMA may still run on the stream)
int h2d(lgt_engine* e, void* dst, const void* src, size_t bytes)
{
    TRY(ensure_pinned(e));
    for (size_t off = 0, c = 0; off < bytes; off += PIN_CHUNK, c++) {
        const int b = (int)(c & 1);
        const size_t n = std::min(PIN_CHUNK, bytes - off);
        // the chunk's last DMA (of this or an earlier copy) has read it
        HIP_TRY(hipEventSynchronize(e->pin_ev[b]), "hipEventSynchronize");
        par_copy(e->pin[b], (const char*)src + off, n);
        HIP_TRY(hipMemcpyAsync((char*)dst + off, e->pin[b], n, hipMemcpyHostToDevice, e->stream), "hipMemcpyAsync H2D");
        HIP_TRY(hipEventRecord(e->pin_ev[b], e->stream), "hipEventRecord");
    }
    return LGT_OK;
}

// device -> host after the work already queued on the engine's stream;
// returns with the host array complete
int d2h(lgt_engine* e, void* dst, const void* src, size_t bytes)
{
    TRY(ensure_pinned(e));
    const size_t nch = (bytes + PIN_CHUNK - 1) / PIN_CHUNK;
    auto issue = [&](size_t c) -> int {
        const int b = (int)(c & 1);
        const size_t off = c * PIN_CHUNK, n = std::min(PIN_CHUNK, bytes - off);
        HIP_TRY(hipMemcpyAsync(e->pin[b], (const char*)src + off, n, hipMemcpyDeviceToHost, e->stream), "hipMemcpyAsync D2H");
        HIP_TRY(hipEventRecord(e->pin_ev[b], e->stream), "hipEventRecord");
        return LGT_OK;
    };
    if (nch) TRY(issue(0));
    for (size_t c = 0; c < nch; c++) {
        const int b = (int)(c & 1);
        const size_t off = c * PIN_CHUNK, n = std::min(PIN_CHUNK, bytes - off);
        if (c + 1 < nch) TRY(issue(c + 1));   // the other chunk was emptied by the last pass
        HIP_TRY(hipEventSynchronize(e->pin_ev[b]), "hipEventSynchronize");
        par_copy((char*)dst + off, e->pin[b], n);
    }
    return LGT_OK;
}

void set_planes(RawPlanes& P, void* base, uint32_t vh, size_t link_bytes)
{
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++) P.p[mu][p] = (char*)base + ((size_t)(mu * 2 + p) * vh) * link_bytes;
}

// runs one launcher, bracketed by events when profiling
template <class F>
int run(lgt_engine* e, int cls, const char* what, F&& f)
{
    hipError_t err;
    if (e->profile) {
        ProfEvent ev;
        ev.cls = cls;
        HIP_TRY(hipEventCreate(&ev.a), "hipEventCreate");
        HIP_TRY(hipEventCreate(&ev.b), "hipEventCreate");
        HIP_TRY(hipEventRecord(ev.a, e->stream), "hipEventRecord");
        err = f();
        HIP_TRY(hipEventRecord(ev.b, e->stream), "hipEventRecord");
        e->events.push_back(ev);
    } else {
        err = f();
    }
    if (err != hipSuccess) return hip_fail(err, what);
    return LGT_OK;
}

int flush_profile(lgt_engine* e)
{
    if (e->events.empty()) return LGT_OK;
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    for (auto& ev : e->events) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b), "hipEventElapsedTime");
        e->prof_ms[ev.cls] += ms;
        e->prof_n[ev.cls] += 1;
        hipEventDestroy(ev.a);
        hipEventDestroy(ev.b);
    }
    e->events.clear();
    return LGT_OK;
}

#define DISPATCH(e, call) ((e)->prec == LGT_FP64 ? Launch<double>::call : Launch<float>::call)

uint32_t nbx(const lgt_engine* e) { return grid_x(e->geo); }

// Group size of the grouped tile order.  Auto (0): the largest divisor of the
// tiles of one time slice that is at most an eighth of them, so that with the
// XCD-contiguous order each of the 8 XCDs walks its own spatial slab through
// all time slices (measured best on MI355X at 48^4, 27 tiles, and at 96^4, 216
// tiles).  Returns false for a size that does not divide the tiles of a time
// slice.
bool set_group(Geo& g, int gs)
{
    if ((g.vh / (uint32_t)g.L[0]) % 256u != 0) {
        g.gs = 0;
        return gs == 0;
    }
    uint32_t tiles = g.vh / (uint32_t)g.L[0] / 256u;
    if (gs == 0) {
        uint32_t cap = std::max(1u, tiles / 8u);
        for (uint32_t d = std::min(cap, tiles); d >= 1; d--)
            if (tiles % d == 0) {
                gs = (int)d;
                break;
            }
    } else if (gs < 0 || tiles % (uint32_t)gs != 0) {
        return false;
    }
    g.gs = gs;
    g.dgs = make_fastdiv((uint32_t)g.gs);
    g.dgt = make_fastdiv((uint32_t)g.gs * (uint32_t)g.L[0]);
    return true;
}

// the time-fastest and grouped orders need whole 256-site tiles per time slice
int tile_order_ok(const Geo& g, int order)
{
    if ((order & 4) && g.gs == 0) order &= ~4;
    if ((order & 2) && ((g.vh / (uint32_t)g.L[0]) % 256u) != 0) order &= ~2;
    return order;
}

// Geometry of one launch with its tile order.  Measured on MI355X at 48^4 and
// 96^4: the grouped, XCD-contiguous order (5) serves the staple gathers of
// every kernel best, except the loop-gradient kernel, whose long temporal
// walks prefer plain time-fastest tiles (3; -20..30 %); without whole tiles
// per time slice the XCD-contiguous linear order (1) is used.
Geo kgeo(const lgt_engine* e, int cls, int mu)
{
    (void)mu;
    Geo g = e->geo;
    int order = (e->order >= 0) ? e->order : (cls == LGT_K_WLOOP_GRAD ? 3 : (g.gs ? 5 : 1));
    // The FP64 loops read 16 time slices ahead of a tile in 32-byte links: a
    // third of the automatic group keeps that window in L2 (48^4, 160 loops:
    // 20.4 instead of 26.9 ms; FP32 is fastest with the automatic group).
    if (cls == LGT_K_WLOOP && e->prec == LGT_FP64 && e->group == 0 && g.gs > 0) {
        const uint32_t tiles = g.vh / (uint32_t)g.L[0] / 256u;
        for (uint32_t d = std::max(1u, (uint32_t)g.gs / 3u); d >= 1; d--)
            if (tiles % d == 0) {
                set_group(g, (int)d);
                break;
            }
    }
    g.order = tile_order_ok(g, order);
    return g;
}

// Geometry of a launch in a chain of field passes (heatbath, overrelaxation,
// Metropolis, APE steps and their adjoint).  A sweep reads the 48^4 FP32 field
// (340 MB) about twice per launch pair, more than the 256 MB memory-side cache
// holds; walking the tiles of every second launch backwards starts it on the
// links the previous launch touched last, which are still in that cache
// (measured at 48^4: sweep -6 %, 4D APE step -5 %).
Geo ugeo(lgt_engine* e, int cls, int mu)
{
    Geo g = kgeo(e, cls, mu);
    if (e->alternate && (e->flip ^= 1u)) g.order |= 8;
    return g;
}

// reduces `nrows` rows of `ncols` double-doubles to host (hi, lo) pairs
int reduce_to_host_dd(lgt_engine* e, int nrows, int ncols, double* out2)
{
    TRY(ensure(e, &e->dres, &e->dres_bytes, sizeof(double) * 2 * (size_t)nrows, "results"));
    TRY(ensure_host(e, sizeof(double) * 2 * (size_t)nrows));
    TRY(run(e, LGT_K_REDUCE, "reduce", [&] { return launch_reduce_rows(e->stream, e->partial, nrows, ncols, e->dres); }));
    HIP_TRY(hipMemcpyAsync(e->hres, e->dres, sizeof(double) * 2 * nrows, hipMemcpyDeviceToHost, e->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    for (int r = 0; r < 2 * nrows; r++) out2[r] = e->hres[r];
    return LGT_OK;
}

// same, rounded to one double per row
int reduce_to_host(lgt_engine* e, int nrows, int ncols, double* out)
{
    std::vector<double> dd(2 * (size_t)nrows);
    TRY(reduce_to_host_dd(e, nrows, ncols, dd.data()));
    for (int r = 0; r < nrows; r++) out[r] = dd[2 * r] + dd[2 * r + 1];
    return LGT_OK;
}

// Per-sweep acceptance of the update loops: the 8 nb block partials of sweep
// s are reduced on the stream into the double-double dres[s mod ACC_BATCH],
// and the host reads a batch of those at once instead of waiting for every
// sweep.
constexpr int ACC_BATCH = 1024;

int acc_prepare(lgt_engine* e, int nsweeps)
{
    const size_t batch = (size_t)std::min(std::max(nsweeps, 1), ACC_BATCH);
    TRY(ensure(e, &e->partial, &e->partial_bytes, sizeof(double) * 2 * 8 * (size_t)nbx(e), "partials"));
    TRY(ensure(e, &e->dres, &e->dres_bytes, sizeof(double) * 2 * batch, "results"));
    TRY(ensure_host(e, sizeof(double) * 2 * batch));
    return LGT_OK;
}

// enqueues the reduction of sweep s's partials (after its update launches)
int acc_reduce(lgt_engine* e, int s)
{
    return run(e, LGT_K_REDUCE, "reduce", [&] {
        return launch_reduce_rows(e->stream, e->partial, 1, (int)(8 * nbx(e)), (double*)e->dres + 2 * (s % ACC_BATCH));
    });
}

// at the end of sweep s: when the batch is full or s is the last sweep, reads
// it back and hands each of its sweeps' sums (hi + lo, as reduce_to_host
// rounds them) to add, in sweep order
template <class F>
int acc_collect(lgt_engine* e, int s, int nsweeps, F&& add)
{
    if ((s + 1) % ACC_BATCH != 0 && s + 1 != nsweeps) return LGT_OK;
    const int n = s % ACC_BATCH + 1;
    HIP_TRY(hipMemcpyAsync(e->hres, e->dres, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, e->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    for (int j = 0; j < n; j++) add(e->hres[2 * j] + e->hres[2 * j + 1]);
    return LGT_OK;
}

// The launches of one update sweep (update.c:102-176): the heatbath
// sub-sweeps (with acc, each block's sum of 1/trials goes to the partials),
// n_or overrelaxation sweeps, and after a heatbath-only sweep the projection
// of the whole field.  In FP64 that projection is a pass of its own, so that
// the staples of the sweep read the heatbath's unprojected links as the
// reference's do; in FP32 the heatbath stores its links projected instead,
// which saves that pass (one read and one write of the field).
int enqueue_sweep(lgt_engine* e, int n_or, double beta, bool acc)
{
    const uint32_t nb = nbx(e);
    const bool fold_projection = n_or == 0 && e->prec == LGT_FP32;
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++) {
            void* ap = acc ? (char*)e->partial + sizeof(double) * 2 * (size_t)nb * (mu * 2 + p) : nullptr;
            TRY(run(e, LGT_K_HEATBATH, "heatbath", [&] {
                return DISPATCH(e, heatbath(e->stream, ugeo(e, LGT_K_HEATBATH, mu), e->cur, mu, p, beta, e->seed,
                                            e->sweep, ap, fold_projection));
            }));
        }
    for (int k = 0; k < n_or; k++)
        for (int mu = 0; mu < 4; mu++)
            for (int p = 0; p < 2; p++)
                TRY(run(e, LGT_K_OVERRELAX, "overrelax",
                        [&] { return DISPATCH(e, overrelax(e->stream, ugeo(e, LGT_K_OVERRELAX, mu), e->cur, mu, p)); }));
    if (n_or == 0 && !fold_projection)
        TRY(run(e, LGT_K_PROJECT, "project", [&] { return DISPATCH(e, project(e->stream, e->geo, e->cur)); }));
    return LGT_OK;
}

int alloc_field(lgt_engine* e, void** base, RawPlanes* P)
{
    size_t bytes = (size_t)4 * e->vol * e->link_bytes, have = 0;
    *base = nullptr;
    TRY(ensure(e, base, &have, bytes, "adjoint field"));
    set_planes(*P, *base, e->geo.vh, e->link_bytes);
    return LGT_OK;
}

// Buffers of the smearing adjoint hold only the directions they need: the
// fields of the steps the links the mode smears (spatial 1..3, all 0..3,
// temporal 0; the spatial links of temporal smearing, which never change, are
// kept once in `fixed`), |M| those same links, the cotangents every link a
// staple touches (spatial 1..3: the temporal links are in no staple; else
// 0..3).  Planes [mu][p] of a buffer for mu in [lo, hi] lie at
// ((mu - lo) * 2 + p) * VOL/2 elements; the others are null (or `fixed`).
struct DirRange { int lo, hi; };
DirRange vjp_link_dirs(int mode) { return mode == 0 ? DirRange{1, 3} : (mode == 1 ? DirRange{0, 3} : DirRange{0, 0}); }
DirRange vjp_cot_dirs(int mode) { return mode == 0 ? DirRange{1, 3} : DirRange{0, 3}; }

size_t dir_bytes(const lgt_engine* e, DirRange r, size_t el) { return (size_t)(r.hi - r.lo + 1) * e->vol * el; }

template <class P>
int alloc_dirs(lgt_engine* e, DirRange r, size_t el, void** base, P* out, const char* what)
{
    size_t have = 0;
    *base = nullptr;
    TRY(ensure(e, base, &have, dir_bytes(e, r, el), what));
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++)
            out->p[mu][p] = (mu >= r.lo && mu <= r.hi)
                                ? (void*)((char*)*base + ((size_t)((mu - r.lo) * 2 + p) * e->geo.vh) * el)
                                : nullptr;
    return LGT_OK;
}

// a field buffer of the mode's smeared links; in temporal mode its spatial
// planes are those of `fixed`
int alloc_vjp_field(lgt_engine* e, int mode, void** base, RawPlanes* P)
{
    TRY(alloc_dirs(e, vjp_link_dirs(mode), e->link_bytes, base, P, "adjoint field"));
    if (mode == 2)
        for (int mu = 1; mu < 4; mu++)
            for (int p = 0; p < 2; p++) P->p[mu][p] = e->vjp.fixedP.p[mu][p];
    return LGT_OK;
}

void free_buf(lgt_engine* e, void*& b, size_t bytes)
{
    if (b) {
        (void)hipFree(b);
        e->dev_bytes -= (int64_t)bytes;
    }
    b = nullptr;
}

void free_vjp(lgt_engine* e, bool keep_cot)
{
    VjpState& v = e->vjp;
    for (auto& b : v.ck) free_buf(e, b, v.fbytes);
    for (auto& b : v.ckM) free_buf(e, b, v.nbytes);
    for (auto& b : v.segU) free_buf(e, b, v.fbytes);
    for (auto& b : v.segM) free_buf(e, b, v.nbytes);
    v.ck.clear(); v.ckM.clear(); v.segU.clear(); v.segM.clear();
    v.ckP.clear(); v.ckMN.clear(); v.segUP.clear(); v.segMN.clear();
    v.top_saved = false;
    free_buf(e, v.g[0], v.gbytes);
    free_buf(e, v.g[1], v.gbytes);
    free_buf(e, v.fixed, dir_bytes(e, DirRange{1, 3}, e->link_bytes));
    if (!keep_cot) {
        free_buf(e, v.gobj, (size_t)4 * e->vol * e->link_bytes);
        free_buf(e, v.lines, v.lines_bytes);
        v.lines_bytes = 0;
        v.cot = false;
    }
    v.fwd = false;
    v.nsteps = v.seg = v.nck = 0;
}

// copies the planes of directions r from src to dst
int copy_dirs(lgt_engine* e, const RawPlanes& dst, const RawPlanes& src, DirRange r)
{
    size_t bytes = (size_t)e->geo.vh * e->link_bytes;
    for (int mu = r.lo; mu <= r.hi; mu++)
        for (int p = 0; p < 2; p++)
            HIP_TRY(hipMemcpyAsync(dst.p[mu][p], src.p[mu][p], bytes, hipMemcpyDeviceToDevice, e->stream),
                    "hipMemcpyAsync D2D");
    return LGT_OK;
}

int check_loop(const lgt_engine* e, int R, int T)
{
    if (T < 1 || T > e->L[0] - 1 || T > WMAXT_HOST || R < 1 || R > 4096)
        return fail(LGT_EINVAL, "loop R=%d T=%d: need 1 <= R <= 4096 and 1 <= T <= %d", R, T, e->L[0] - 1);
    return LGT_OK;
}

}  // namespace

extern "C" {

const char* lgt_strerror(int status)
{
    switch (status) {
    case LGT_OK: return "ok";
    case LGT_EINVAL: return "invalid argument";
    case LGT_ENOMEM: return "out of memory";
    case LGT_EDEVICE: return "HIP device error";
    default: return "unknown status";
    }
}

const char* lgt_last_error(void) { return g_last_error; }

int lgt_create(const int extents[4], int precision, uint64_t seed, lgt_engine** out)
{
    if (!extents || !out) return fail(LGT_EINVAL, "lgt_create: null argument");
    *out = nullptr;
    if (precision != LGT_FP64 && precision != LGT_FP32) return fail(LGT_EINVAL, "lgt_create: precision %d", precision);
    int64_t vol = 1;
    for (int d = 0; d < 4; d++) {
        if (extents[d] < 2 || (extents[d] & 1) || extents[d] > 4096)
            return fail(LGT_EINVAL, "lgt_create: extent[%d] = %d must be even and in [2, 4096]", d, extents[d]);
        vol *= extents[d];
    }
    if (extents[3] < 4) return fail(LGT_EINVAL, "lgt_create: extent[3] = %d must be >= 4", extents[3]);
    // kernels address both parity planes of a direction with 32-bit byte offsets
    const int64_t lb = (precision == LGT_FP64) ? 32 : 16;
    if (vol * lb >= (int64_t(1) << 32))
        return fail(LGT_EINVAL, "lgt_create: volume %lld too large for %s links on one device", (long long)vol,
                    precision == LGT_FP64 ? "FP64" : "FP32");

    lgt_engine* e = new lgt_engine();
    hipError_t err = hipGetDevice(&e->dev);
    if (err != hipSuccess) {
        delete e;
        return hip_fail(err, "hipGetDevice");
    }
    for (int d = 0; d < 4; d++) e->L[d] = extents[d];
    e->prec = precision;
    e->seed = seed;
    e->vol = vol;
    Geo& g = e->geo;
    for (int d = 0; d < 4; d++) g.L[d] = extents[d];
    g.hz = extents[3] / 2;
    g.vh = (uint32_t)(vol / 2);
    g.stride[3] = 0;
    g.stride[2] = (uint32_t)g.hz;
    g.stride[1] = (uint32_t)extents[2] * g.hz;
    g.stride[0] = (uint32_t)extents[1] * extents[2] * g.hz;
    g.dz = make_fastdiv((uint32_t)g.hz);
    g.dy = make_fastdiv((uint32_t)extents[2]);
    g.dx = make_fastdiv((uint32_t)extents[1]);
    g.dt = make_fastdiv((uint32_t)extents[0]);
    g.order = 0;
    g.gs = 0;
    set_group(g, 0);
    e->link_bytes = (precision == LGT_FP64) ? 32 : 16;

    err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete e;
        return hip_fail(err, "hipStreamCreate");
    }
    size_t bytes = (size_t)4 * vol * e->link_bytes;
    err = hipMalloc(&e->field, bytes);
    if (err != hipSuccess) {
        hipStreamDestroy(e->stream);
        delete e;
        return fail(LGT_ENOMEM, "hipMalloc(%zu) for the link field: %s", bytes, hipGetErrorString(err));
    }
    e->dev_bytes = (int64_t)bytes;
    set_planes(e->cur, e->field, g.vh, e->link_bytes);
    int rc = lgt_init_field(e, 0);
    if (rc != LGT_OK) {
        lgt_destroy(e);
        return rc;
    }
    *out = e;
    return LGT_OK;
}

int lgt_destroy(lgt_engine* e)
{
    if (!e) return LGT_OK;
    hipSetDevice(e->dev);
    if (e->stream) hipStreamSynchronize(e->stream);
    for (auto& ev : e->events) {
        hipEventDestroy(ev.a);
        hipEventDestroy(ev.b);
    }
    free_vjp(e, false);
    if (e->field) hipFree(e->field);
    if (e->field_alt) hipFree(e->field_alt);
    if (e->wline) hipFree(e->wline);
    if (e->partial) hipFree(e->partial);
    if (e->dres) hipFree(e->dres);
    if (e->staging) hipFree(e->staging);
    if (e->link_buf) hipFree(e->link_buf);
    if (e->uni_buf) hipFree(e->uni_buf);
    if (e->qline) hipFree(e->qline);
    if (e->hres) hipHostFree(e->hres);
    for (int b = 0; b < 2; b++) {
        if (e->pin[b]) hipHostFree(e->pin[b]);
        if (e->pin_ev[b]) hipEventDestroy(e->pin_ev[b]);
    }
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return LGT_OK;
}

int lgt_set_option(lgt_engine* e, int option, int value)
{
    if (!e) return fail(LGT_EINVAL, "lgt_set_option: null engine");
    switch (option) {
    case LGT_OPT_TILE_ORDER:
        if (value < -1 || value > 15) return fail(LGT_EINVAL, "lgt_set_option: tile order %d", value);
        e->order = value;
        return LGT_OK;
    case LGT_OPT_GROUP:
        if (!set_group(e->geo, value)) return fail(LGT_EINVAL, "lgt_set_option: group size %d", value);
        e->group = value;
        return LGT_OK;
    case LGT_OPT_ALTERNATE:
        if (value != 0 && value != 1) return fail(LGT_EINVAL, "lgt_set_option: alternate %d", value);
        e->alternate = value;
        return LGT_OK;
    case LGT_OPT_PROFILE:
        if (!value) TRY(flush_profile(e));
        e->profile = value ? 1 : 0;
        return LGT_OK;
    default: return fail(LGT_EINVAL, "lgt_set_option: unknown option %d", option);
    }
}

int lgt_get_extents(const lgt_engine* e, int extents[4], int* precision)
{
    if (!e) return fail(LGT_EINVAL, "lgt_get_extents: null engine");
    if (extents)
        for (int d = 0; d < 4; d++) extents[d] = e->L[d];
    if (precision) *precision = e->prec;
    return LGT_OK;
}

int64_t lgt_device_bytes(const lgt_engine* e) { return e ? e->dev_bytes : 0; }

int lgt_init_field(lgt_engine* e, int start)
{
    if (!e || (start != 0 && start != 1)) return fail(LGT_EINVAL, "lgt_init_field: start must be 0 or 1");
    TRY(set_dev(e));
    if (start == 0)
        return run(e, LGT_K_FIELD, "cold start", [&] { return DISPATCH(e, cold(e->stream, e->geo, e->cur)); });
    return run(e, LGT_K_FIELD, "hot start", [&] { return DISPATCH(e, hot(e->stream, e->geo, e->cur, e->seed)); });
}

int lgt_upload(lgt_engine* e, const double* aos)
{
    if (!e || !aos) return fail(LGT_EINVAL, "lgt_upload: null argument");
    TRY(set_dev(e));
    size_t bytes = sizeof(double) * 16 * (size_t)e->vol;
    TRY(ensure(e, &e->staging, &e->staging_bytes, bytes, "staging"));
    TRY(h2d(e, e->staging, aos, bytes));
    TRY(run(e, LGT_K_FIELD, "from_aos",
            [&] { return DISPATCH(e, from_aos(e->stream, e->geo, e->cur, (const double*)e->staging)); }));
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_upload_plaquette(lgt_engine* e, const double* aos, double* plaq)
{
    if (!e || !aos || !plaq) return fail(LGT_EINVAL, "lgt_upload_plaquette: null argument");
    TRY(set_dev(e));
    size_t bytes = sizeof(double) * 16 * (size_t)e->vol;
    TRY(ensure(e, &e->staging, &e->staging_bytes, bytes, "staging"));
    TRY(h2d(e, e->staging, aos, bytes));
    const int nb = (int)((2 * (size_t)e->geo.vh + 255) / 256);   // one partial per 256 sites
    TRY(ensure(e, &e->partial, &e->partial_bytes, sizeof(double) * 2 * (size_t)nb, "partials"));
    TRY(run(e, LGT_K_PLAQUETTE, "plaquette of the uploaded field",
            [&] { return launch_plaquette_aos(e->stream, e->geo, (const double*)e->staging, e->partial); }));
    double sum = 0.0;
    TRY(reduce_to_host(e, 1, nb, &sum));
    *plaq = sum / (double)(2 * 6 * e->vol);
    TRY(run(e, LGT_K_FIELD, "from_aos",
            [&] { return DISPATCH(e, from_aos(e->stream, e->geo, e->cur, (const double*)e->staging)); }));
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_download(lgt_engine* e, double* aos)
{
    if (!e || !aos) return fail(LGT_EINVAL, "lgt_download: null argument");
    TRY(set_dev(e));
    size_t bytes = sizeof(double) * 16 * (size_t)e->vol;
    TRY(ensure(e, &e->staging, &e->staging_bytes, bytes, "staging"));
    TRY(run(e, LGT_K_FIELD, "to_aos", [&] { return DISPATCH(e, to_aos(e->stream, e->geo, e->cur, (double*)e->staging)); }));
    return d2h(e, aos, e->staging, bytes);
}

// ---- configuration files (config_IO.c:40-117,537-621) ----
static_assert(sizeof(int32_t) == 4 && sizeof(double) == 8, "config format");

namespace {
bool host_little_endian()
{
    const uint16_t one = 1;
    return *(const uint8_t*)&one == 1;
}

struct FileCloser {
    FILE* f;
    ~FileCloser() { if (f) fclose(f); }
};
}  // namespace

int lgt_write_config(lgt_engine* e, const char* path)
{
    if (!e || !path) return fail(LGT_EINVAL, "lgt_write_config: null argument");
    if (!host_little_endian()) return fail(LGT_EINVAL, "lgt_write_config: big-endian host");
    double plaq = 0.0;
    TRY(lgt_plaquette(e, &plaq));
    std::vector<double> aos;
    try {
        aos.resize(16 * (size_t)e->vol);
    } catch (const std::bad_alloc&) {
        return fail(LGT_ENOMEM, "lgt_write_config: host buffer");
    }
    TRY(lgt_download(e, aos.data()));
    FileCloser fc{fopen(path, "wb")};
    if (!fc.f) return fail(LGT_EINVAL, "lgt_write_config: cannot create %s", path);
    const int32_t info[2] = {4, 2};
    const int32_t ls[4] = {e->L[0], e->L[1], e->L[2], e->L[3]};
    bool ok = fwrite(info, sizeof(int32_t), 2, fc.f) == 2 && fwrite(ls, sizeof(int32_t), 4, fc.f) == 4 &&
              fwrite(&plaq, sizeof(double), 1, fc.f) == 1 && fwrite(aos.data(), sizeof(double), aos.size(), fc.f) == aos.size();
    if (!ok) return fail(LGT_EINVAL, "lgt_write_config: write error on %s", path);
    if (fclose(fc.f) != 0) {
        fc.f = nullptr;
        return fail(LGT_EINVAL, "lgt_write_config: write error on %s", path);
    }
    fc.f = nullptr;
    return LGT_OK;
}

namespace {
// header of a configuration file; checks DIM, SUN and the expected extents
int read_config_header(FILE* f, const char* path, const int expect[4], double* plaq0, const char* who)
{
    int32_t info[2], ls[4];
    if (fread(info, sizeof(int32_t), 2, f) != 2 || fread(ls, sizeof(int32_t), 4, f) != 4 ||
        fread(plaq0, sizeof(double), 1, f) != 1)
        return fail(LGT_EINVAL, "%s: read error on %s", who, path);
    if (info[0] != 4 || info[1] != 2) return fail(LGT_EINVAL, "%s: DIM %d SUN %d, expected 4 2", who, info[0], info[1]);
    for (int d = 0; d < 4; d++)
        if (ls[d] != expect[d])
            return fail(LGT_EINVAL, "%s: lattice %dx%dx%dx%d, expected %dx%dx%dx%d", who, ls[0], ls[1], ls[2], ls[3],
                        expect[0], expect[1], expect[2], expect[3]);
    return LGT_OK;
}

// uploads, recomputes the plaquette of the links as read (FP64, before an FP32
// engine rounds them) and applies readConfig's test (config_IO.c:609-614)
int load_and_check(lgt_engine* e, const double* aos, double plaq0, double* plaq_file, double* plaq_calc, const char* who)
{
    double plaq = 0.0;
    TRY(lgt_upload_plaquette(e, aos, &plaq));
    if (plaq_file) *plaq_file = plaq0;
    if (plaq_calc) *plaq_calc = plaq;
    const double eps = std::sqrt(6.0 * (double)e->vol) * DBL_EPSILON;
    if (!(std::fabs(plaq - plaq0) <= eps))
        return fail(LGT_EINVAL, "%s: plaquette test failed (file %.16g, field %.16g)", who, plaq0, plaq);
    return LGT_OK;
}
}  // namespace

int lgt_read_config(lgt_engine* e, const char* path, double* plaq_file, double* plaq_calc)
{
    if (!e || !path) return fail(LGT_EINVAL, "lgt_read_config: null argument");
    if (!host_little_endian()) return fail(LGT_EINVAL, "lgt_read_config: big-endian host");
    FileCloser fc{fopen(path, "rb")};
    if (!fc.f) return fail(LGT_EINVAL, "lgt_read_config: cannot open %s", path);
    double plaq0 = 0.0;
    TRY(read_config_header(fc.f, path, e->L, &plaq0, "lgt_read_config"));
    std::vector<double> aos;
    try {
        aos.resize(16 * (size_t)e->vol);
    } catch (const std::bad_alloc&) {
        return fail(LGT_ENOMEM, "lgt_read_config: host buffer");
    }
    if (fread(aos.data(), sizeof(double), aos.size(), fc.f) != aos.size())
        return fail(LGT_EINVAL, "lgt_read_config: %s is short", path);
    return load_and_check(e, aos.data(), plaq0, plaq_file, plaq_calc, "lgt_read_config");
}

int lgt_read_config_doubled(lgt_engine* e, const char* path, double* plaq_file, double* plaq_calc)
{
    if (!e || !path) return fail(LGT_EINVAL, "lgt_read_config_doubled: null argument");
    if (!host_little_endian()) return fail(LGT_EINVAL, "lgt_read_config_doubled: big-endian host");
    int half[4];
    for (int d = 0; d < 4; d++) {
        if (e->L[d] % 2) return fail(LGT_EINVAL, "lgt_read_config_doubled: odd extent %d", e->L[d]);
        half[d] = e->L[d] / 2;
    }
    FileCloser fc{fopen(path, "rb")};
    if (!fc.f) return fail(LGT_EINVAL, "lgt_read_config_doubled: cannot open %s", path);
    double plaq0 = 0.0;
    TRY(read_config_header(fc.f, path, half, &plaq0, "lgt_read_config_doubled"));
    const int64_t vs = e->vol / 16;
    std::vector<double> aos, small;
    try {
        aos.resize(16 * (size_t)e->vol);
        small.resize(16 * (size_t)vs);
    } catch (const std::bad_alloc&) {
        return fail(LGT_ENOMEM, "lgt_read_config_doubled: host buffer");
    }
    if (fread(small.data(), sizeof(double), small.size(), fc.f) != small.size())
        return fail(LGT_EINVAL, "lgt_read_config_doubled: %s is short", path);
    const int* L = e->L;
    for (int64_t n = 0; n < vs; n++) {
        int c[4];
        int64_t r = n;
        for (int d = 3; d >= 0; d--) {
            c[d] = (int)(r % half[d]);
            r /= half[d];
        }
        for (int mu = 0; mu < 4; mu++) {
            const double* u = &small[16 * n + 4 * mu];
            for (int k = 0; k < 16; k++) {   // sublattice k: bit j mirrors direction j
                int64_t idx = 0;
                for (int j = 0; j < 4; j++) {
                    int x = c[j];
                    if ((k >> j) & 1) x = (j == mu) ? L[j] - c[j] - 1 : (c[j] == 0 ? half[j] : L[j] - c[j]);
                    idx = idx * L[j] + x;
                }
                double* o = &aos[16 * idx + 4 * mu];
                const double sg = ((k >> mu) & 1) ? -1.0 : 1.0;   // mirrored along the link: dagger
                o[0] = u[0];
                o[1] = sg * u[1];
                o[2] = sg * u[2];
                o[3] = sg * u[3];
            }
        }
    }
    return load_and_check(e, aos.data(), plaq0, plaq_file, plaq_calc, "lgt_read_config_doubled");
}

int lgt_metropolis_update(lgt_engine* e, int nsweeps, double beta, double eps, double* acceptance)
{
    if (!e || nsweeps < 0) return fail(LGT_EINVAL, "lgt_metropolis_update: nsweeps must be >= 0");
    if (!(beta >= 0.0) || !std::isfinite(beta) || !std::isfinite(eps))
        return fail(LGT_EINVAL, "lgt_metropolis_update: beta = %g, eps = %g", beta, eps);
    TRY(set_dev(e));
    const uint32_t nb = nbx(e);
    double accepted = 0.0;
    if (acceptance) TRY(acc_prepare(e, nsweeps));
    for (int s = 0; s < nsweeps; s++) {
        for (int mu = 0; mu < 4; mu++)
            for (int p = 0; p < 2; p++) {
                void* ap = acceptance ? (char*)e->partial + sizeof(double) * 2 * (size_t)nb * (mu * 2 + p) : nullptr;
                TRY(run(e, LGT_K_METROPOLIS, "metropolis", [&] {
                    return DISPATCH(e, metropolis(e->stream, ugeo(e, LGT_K_METROPOLIS, mu), e->cur, mu, p, beta, eps,
                                                  e->seed, e->sweep, ap));
                }));
            }
        if (acceptance) TRY(acc_reduce(e, s));
        e->sweep++;
        if (acceptance) TRY(acc_collect(e, s, nsweeps, [&](double v) { accepted += v; }));
    }
    TRY(run(e, LGT_K_PROJECT, "project", [&] { return DISPATCH(e, project(e->stream, e->geo, e->cur)); }));
    if (acceptance) *acceptance = nsweeps > 0 ? accepted / ((double)(4 * e->vol) * nsweeps) : 0.0;
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_heatbath_subsweep(lgt_engine* e, int mu, int parity, double beta, double* acc_sum)
{
    if (!e || mu < 0 || mu > 3 || parity < 0 || parity > 1) return fail(LGT_EINVAL, "lgt_heatbath_subsweep: mu/parity");
    TRY(set_dev(e));
    void* ap = nullptr;
    if (acc_sum) {
        TRY(ensure(e, &e->partial, &e->partial_bytes, sizeof(double) * 2 * nbx(e), "partials"));
        ap = e->partial;
    }
    TRY(run(e, LGT_K_HEATBATH, "heatbath", [&] {
        return DISPATCH(e, heatbath(e->stream, ugeo(e, LGT_K_HEATBATH, mu), e->cur, mu, parity, beta, e->seed, e->sweep, ap));
    }));
    if (acc_sum) TRY(reduce_to_host(e, 1, (int)nbx(e), acc_sum));
    return LGT_OK;
}

int lgt_overrelax_subsweep(lgt_engine* e, int mu, int parity)
{
    if (!e || mu < 0 || mu > 3 || parity < 0 || parity > 1) return fail(LGT_EINVAL, "lgt_overrelax_subsweep: mu/parity");
    TRY(set_dev(e));
    return run(e, LGT_K_OVERRELAX, "overrelax",
               [&] { return DISPATCH(e, overrelax(e->stream, ugeo(e, LGT_K_OVERRELAX, mu), e->cur, mu, parity)); });
}

int lgt_project(lgt_engine* e)
{
    if (!e) return fail(LGT_EINVAL, "lgt_project: null engine");
    TRY(set_dev(e));
    return run(e, LGT_K_PROJECT, "project", [&] { return DISPATCH(e, project(e->stream, e->geo, e->cur)); });
}

int lgt_gauge_transform(lgt_engine* e, uint64_t seed, double* g_out)
{
    if (!e) return fail(LGT_EINVAL, "lgt_gauge_transform: null engine");
    TRY(set_dev(e));
    TRY(run(e, LGT_K_FIELD, "gauge transformation",
            [&] { return DISPATCH(e, gauge(e->stream, e->geo, e->cur, seed)); }));
    if (g_out) {
        const size_t bytes = sizeof(double) * 4 * (size_t)e->vol;
        TRY(ensure(e, &e->staging, &e->staging_bytes, sizeof(double) * 16 * (size_t)e->vol, "staging"));
        TRY(run(e, LGT_K_FIELD, "gauge elements",
                [&] { return DISPATCH(e, gauge_elements(e->stream, e->geo, seed, (double*)e->staging)); }));
        TRY(d2h(e, g_out, e->staging, bytes));
    }
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_set_sweep_index(lgt_engine* e, uint32_t sweep)
{
    if (!e) return fail(LGT_EINVAL, "lgt_set_sweep_index: null engine");
    e->sweep = sweep;
    return LGT_OK;
}

uint32_t lgt_get_sweep_index(const lgt_engine* e) { return e ? e->sweep : 0u; }

int lgt_update(lgt_engine* e, int nsweeps, int n_or, double beta, double* acceptance)
{
    if (!e || nsweeps < 0 || n_or < 0) return fail(LGT_EINVAL, "lgt_update: nsweeps/n_or must be >= 0");
    if (!(beta >= 0.0) || !std::isfinite(beta)) return fail(LGT_EINVAL, "lgt_update: beta = %g", beta);
    TRY(set_dev(e));
    double a = 0.0;
    if (acceptance) TRY(acc_prepare(e, nsweeps));
    for (int s = 0; s < nsweeps; s++) {
        TRY(enqueue_sweep(e, n_or, beta, acceptance != nullptr));
        if (acceptance) TRY(acc_reduce(e, s));
        e->sweep++;
        if (acceptance) TRY(acc_collect(e, s, nsweeps, [&](double v) { a += v / (double)(4 * e->vol); }));
    }
    if (acceptance) *acceptance = nsweeps > 0 ? a / nsweeps : 0.0;
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_plaquette(lgt_engine* e, double* plaq)
{
    if (!e || !plaq) return fail(LGT_EINVAL, "lgt_plaquette: null argument");
    TRY(set_dev(e));
    const uint32_t nb = nbx(e);
    TRY(ensure(e, &e->partial, &e->partial_bytes, sizeof(double) * 2 * 2 * (size_t)nb, "partials"));
    TRY(run(e, LGT_K_PLAQUETTE, "plaquette",
            [&] { return DISPATCH(e, plaquette(e->stream, kgeo(e, LGT_K_PLAQUETTE, 0), e->cur, e->partial)); }));
    double s = 0.0;
    TRY(reduce_to_host(e, 1, (int)(2 * nb), &s));
    *plaq = s / (double)(2 * 6 * e->vol);
    return LGT_OK;
}

int lgt_staple(lgt_engine* e, int64_t site, int mu, double A[4])
{
    if (!e || !A) return fail(LGT_EINVAL, "lgt_staple: null argument");
    if (site < 0 || site >= e->vol || mu < 0 || mu > 3)
        return fail(LGT_EINVAL, "lgt_staple: link (%lld, %d) outside the lattice", (long long)site, mu);
    TRY(set_dev(e));
    TRY(ensure(e, &e->link_buf, &e->link_buf_bytes, 8 * sizeof(double) + 2 * sizeof(int), "link buffer"));
    TRY(ensure_host(e, 8 * sizeof(double) + 2 * sizeof(int)));
    TRY(run(e, LGT_K_FIELD, "staple", [&] {
        return DISPATCH(e, staple_link(e->stream, e->geo, e->cur, (uint32_t)site, mu, (double*)e->link_buf));
    }));
    HIP_TRY(hipMemcpyAsync(e->hres, e->link_buf, 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream), "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    for (int k = 0; k < 4; k++) A[k] = e->hres[k];
    return LGT_OK;
}

int lgt_local_heatbath(lgt_engine* e, int64_t site, int mu, double beta, const double* uniforms, int n_uniforms,
                       double link[4], int* trials, int* needed)
{
    if (!e || !trials || !needed || (n_uniforms > 0 && !uniforms) || n_uniforms < 0)
        return fail(LGT_EINVAL, "lgt_local_heatbath: bad arguments");
    if (site < 0 || site >= e->vol || mu < 0 || mu > 3)
        return fail(LGT_EINVAL, "lgt_local_heatbath: link (%lld, %d) outside the lattice", (long long)site, mu);
    if (!(beta >= 0.0) || !std::isfinite(beta)) return fail(LGT_EINVAL, "lgt_local_heatbath: beta = %g", beta);
    TRY(set_dev(e));
    TRY(ensure(e, &e->link_buf, &e->link_buf_bytes, 8 * sizeof(double) + 2 * sizeof(int), "link buffer"));
    TRY(ensure(e, &e->uni_buf, &e->uni_buf_bytes, sizeof(double) * (size_t)std::max(n_uniforms, 1), "uniforms"));
    TRY(ensure_host(e, 8 * sizeof(double) + 2 * sizeof(int)));
    if (n_uniforms > 0)
        HIP_TRY(hipMemcpyAsync(e->uni_buf, uniforms, sizeof(double) * n_uniforms, hipMemcpyHostToDevice, e->stream),
                "hipMemcpyAsync H2D");
    double* dlink = (double*)e->link_buf + 4;
    int* dstat = (int*)((double*)e->link_buf + 8);
    TRY(run(e, LGT_K_HEATBATH, "local heatbath", [&] {
        return DISPATCH(e, local_heatbath(e->stream, e->geo, e->cur, (uint32_t)site, mu, beta,
                                          (const double*)e->uni_buf, n_uniforms, dlink, dstat));
    }));
    HIP_TRY(hipMemcpyAsync(e->hres, dlink, 4 * sizeof(double) + 2 * sizeof(int), hipMemcpyDeviceToHost, e->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    const int* st = (const int*)(e->hres + 4);
    *trials = st[0];
    *needed = st[1];
    if (link && st[0] >= 0)
        for (int k = 0; k < 4; k++) link[k] = e->hres[k];
    return LGT_OK;
}

int lgt_gauge_action(lgt_engine* e, double beta, double* action)
{
    if (!action) return fail(LGT_EINVAL, "lgt_gauge_action: null argument");
    double p = 0.0;
    TRY(lgt_plaquette(e, &p));
    *action = beta * 6 * (double)e->vol * (1. - p);
    return LGT_OK;
}

// the second field: APE output (swapped in after every step) and, for the
// Wilson loops, the spatial-line field
static int ensure_alt(lgt_engine* e)
{
    if (e->field_alt) return LGT_OK;
    size_t have = 0;
    TRY(ensure(e, &e->field_alt, &have, (size_t)4 * e->vol * e->link_bytes, "APE buffer"));
    set_planes(e->alt, e->field_alt, e->geo.vh, e->link_bytes);
    return LGT_OK;
}

int lgt_ape(lgt_engine* e, int mode, int nsteps, double sp)
{
    if (!e || mode < 0 || mode > 2 || nsteps < 0) return fail(LGT_EINVAL, "lgt_ape: mode %d nsteps %d", mode, nsteps);
    TRY(set_dev(e));
    if (nsteps == 0) return LGT_OK;
    TRY(ensure_alt(e));
    int mu_lo = (mode == LGT_APE_SPATIAL) ? 1 : 0, mu_hi = (mode == LGT_APE_TEMPORAL) ? 0 : 3;
    for (int s = 0; s < nsteps; s++) {
        TRY(run(e, LGT_K_APE, "ape", [&] { return DISPATCH(e, ape(e->stream, ugeo(e, LGT_K_APE, 0), e->cur, e->alt, mode, sp)); }));
        for (int mu = mu_lo; mu <= mu_hi; mu++)
            for (int p = 0; p < 2; p++) std::swap(e->cur.p[mu][p], e->alt.p[mu][p]);
    }
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

// the loops of `groups` (R ascending) into the partial rows.  The spatial
// lines S_R of the groups alternate between the APE buffer (planes 1..3) and a
// second line field: k_sline builds those of the first group, and every loop
// launch but the last writes the next group's.
static int run_wloops(lgt_engine* e, const std::vector<WGroupHost>& groups, const WGroupHost* dgroups,
                      const int* dslots, int ncols)
{
    if (groups.size() > 1 && !e->wline) {
        size_t have = 0;
        TRY(ensure(e, &e->wline, &have, (size_t)3 * e->vol * e->link_bytes, "second line field"));
        for (int d = 0; d < 4; d++)
            for (int p = 0; p < 2; p++)
                e->wl2.p[d][p] = d == 0 ? nullptr
                                        : (char*)e->wline + ((size_t)((d - 1) * 2 + p) * e->geo.vh) * e->link_bytes;
    }
    RawPlanes S = e->alt, Sn = e->wl2;
    TRY(run(e, LGT_K_WLOOP, "spatial lines", [&] {
        return DISPATCH(e, sline(e->stream, kgeo(e, LGT_K_WLOOP, 0), e->cur, S, 0, groups[0].R));
    }));
    for (size_t k = 0; k < groups.size(); k++) {
        const int Rn = (k + 1 < groups.size()) ? groups[k + 1].R : 0;
        TRY(run(e, LGT_K_WLOOP, "wilson loops", [&] {
            return DISPATCH(e, wloops(e->stream, kgeo(e, LGT_K_WLOOP, 0), e->cur, S, dgroups, (int)k, dslots,
                                      e->partial, ncols, Sn, Rn));
        }));
        std::swap(S, Sn);
    }
    return LGT_OK;
}

int lgt_wilson_loops(lgt_engine* e, int nloop, const int* R, const int* T, double* w)
{
    if (!e || nloop <= 0 || !R || !T || !w) return fail(LGT_EINVAL, "lgt_wilson_loops: empty loop list");
    for (int i = 0; i < nloop; i++) {
        if (T[i] < 1 || T[i] > e->L[0] - 1 || T[i] > WMAXT_HOST)
            return fail(LGT_EINVAL, "lgt_wilson_loops: T = %d outside [1, %d] (lattice too small)", T[i], e->L[0] - 1);
        if (R[i] < 1 || R[i] > 4096) return fail(LGT_EINVAL, "lgt_wilson_loops: R = %d outside [1, 4096]", R[i]);
    }
    TRY(set_dev(e));
    TRY(ensure_alt(e));
    // unique (R,T) pairs, grouped by R (ascending: the line field grows)
    std::map<std::pair<int, int>, int> uniq;
    std::vector<int> order_R;
    std::map<int, int> tmax;
    for (int i = 0; i < nloop; i++) {
        auto key = std::make_pair(R[i], T[i]);
        if (!uniq.count(key)) {
            int id = (int)uniq.size();
            uniq[key] = id;
        }
        if (!tmax.count(R[i])) {
            order_R.push_back(R[i]);
            tmax[R[i]] = 0;
        }
        tmax[R[i]] = std::max(tmax[R[i]], T[i]);
    }
    std::sort(order_R.begin(), order_R.end());
    std::vector<WGroupHost> groups;
    std::vector<int> slots;
    for (int r : order_R) {
        WGroupHost gh{r, tmax[r], (int)slots.size(), 0};
        for (int t = 1; t <= tmax[r]; t++) {
            auto it = uniq.find(std::make_pair(r, t));
            slots.push_back(it == uniq.end() ? -1 : it->second);
        }
        groups.push_back(gh);
    }
    const int nuniq = (int)uniq.size();
    const int ncols = 6 * (int)nbx(e);
    size_t tab_bytes = sizeof(WGroupHost) * groups.size() + sizeof(int) * slots.size();
    size_t part_bytes = sizeof(double) * 2 * (size_t)nuniq * ncols;
    TRY(ensure(e, &e->partial, &e->partial_bytes, part_bytes + tab_bytes + 256, "partials"));
    char* tab = (char*)e->partial + ((part_bytes + 255) & ~(size_t)255);
    WGroupHost* dgroups = (WGroupHost*)tab;
    int* dslots = (int*)(tab + sizeof(WGroupHost) * groups.size());
    HIP_TRY(hipMemcpyAsync(dgroups, groups.data(), sizeof(WGroupHost) * groups.size(), hipMemcpyHostToDevice, e->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(dslots, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice, e->stream),
            "hipMemcpyAsync");
    TRY(run_wloops(e, groups, dgroups, dslots, ncols));
    std::vector<double> sums(nuniq);
    TRY(reduce_to_host(e, nuniq, ncols, sums.data()));
    for (int i = 0; i < nloop; i++) w[i] = sums[uniq[std::make_pair(R[i], T[i])]] / (double)(3 * e->vol * 2);
    return LGT_OK;
}

int lgt_qtop(lgt_engine* e, int kind, double* q, double* abs_sum, double* density)
{
    if (!e || !q || (kind != LGT_Q_IMPROVED && kind != LGT_Q_CLOVER)) return fail(LGT_EINVAL, "lgt_qtop: kind %d", kind);
    TRY(set_dev(e));
    const uint32_t nb = nbx(e);
    TRY(ensure(e, &e->partial, &e->partial_bytes, sizeof(double) * 2 * 4 * (size_t)nb, "partials"));
    double* dq = nullptr;
    if (density) {
        TRY(ensure(e, &e->staging, &e->staging_bytes, sizeof(double) * (size_t)e->vol, "staging"));
        dq = (double*)e->staging;
    }
    if (kind == LGT_Q_IMPROVED) {
        const size_t fb = (size_t)4 * e->vol * e->link_bytes;
        if (!e->qline) {
            TRY(ensure(e, &e->qline, &e->qline_bytes, 2 * fb, "charge line fields"));
            set_planes(e->qL2, e->qline, e->geo.vh, e->link_bytes);
            set_planes(e->qL3, (char*)e->qline + fb, e->geo.vh, e->link_bytes);
        }
        TRY(run(e, LGT_K_QTOP, "charge lines",
                [&] { return DISPATCH(e, qlines(e->stream, e->geo, e->cur, e->qL2, e->qL3)); }));
    }
    TRY(run(e, LGT_K_QTOP, "qtop", [&] {
        return DISPATCH(e, qtop(e->stream, kgeo(e, LGT_K_QTOP, 0), e->cur, kind, e->qL2, e->qL3, e->partial, dq));
    }));
    double s[2];
    TRY(reduce_to_host(e, 2, (int)(2 * nb), s));
    const double PI = 3.14159265358979323846;
    double norm = (kind == LGT_Q_IMPROVED) ? (32. * PI * PI) : (4. * PI * PI);
    *q = s[0] / norm;
    if (abs_sum) *abs_sum = s[1] / norm;
    if (density) {
        TRY(d2h(e, density, dq, sizeof(double) * (size_t)e->vol));
    }
    return LGT_OK;
}

int lgt_ape_vjp_forward(lgt_engine* e, int mode, int nsteps, double sp, int segment)
{
    if (!e || nsteps < 1 || segment < 0 || mode < 0 || mode > 2)
        return fail(LGT_EINVAL, "lgt_ape_vjp_forward: mode %d nsteps %d segment %d", mode, nsteps, segment);
    TRY(set_dev(e));
    VjpState& v = e->vjp;
    int seg = segment > 0 ? segment : (int)std::ceil(std::sqrt((double)nsteps));
    if (seg > nsteps) seg = nsteps;
    const int nck = (nsteps + seg - 1) / seg;
    const DirRange lr = vjp_link_dirs(mode), cr = vjp_cot_dirs(mode);
    // checkpoints, segment fields and norms, cotangents: allocated here once
    // for a layout (mode, seg, nck) and kept for later calls with the same
    // layout, so neither pass allocates when the adjoint is repeated
    if (v.seg != seg || v.nck != nck || v.mode != mode) {
        free_vjp(e, true);
        v.mode = mode;
        v.fbytes = dir_bytes(e, lr, e->link_bytes);
        v.nbytes = dir_bytes(e, lr, e->link_bytes / 4);
        v.gbytes = dir_bytes(e, cr, e->link_bytes);
        int rc = LGT_OK;
        if (mode == 2) rc = alloc_dirs(e, DirRange{1, 3}, e->link_bytes, &v.fixed, &v.fixedP, "adjoint field");
        v.ck.assign(nck, nullptr);
        v.ckP.resize(nck);
        v.ckM.assign(nck, nullptr);
        v.ckMN.resize(nck);
        for (int c = 0; c < nck && rc == LGT_OK; c++) rc = alloc_vjp_field(e, mode, &v.ck[c], &v.ckP[c]);
        for (int c = 1; c < nck && rc == LGT_OK; c++)
            rc = alloc_dirs(e, lr, e->link_bytes / 4, &v.ckM[c], &v.ckMN[c], "adjoint norms");
        if (rc == LGT_OK) rc = alloc_dirs(e, cr, e->link_bytes, &v.g[0], &v.gP[0], "adjoint cotangent");
        if (rc == LGT_OK) rc = alloc_dirs(e, cr, e->link_bytes, &v.g[1], &v.gP[1], "adjoint cotangent");
        v.segU.assign(seg, nullptr);
        v.segUP.resize(seg);
        v.segM.assign(seg, nullptr);
        v.segMN.resize(seg);
        for (int j = 0; j < seg && rc == LGT_OK; j++) {
            rc = alloc_vjp_field(e, mode, &v.segU[j], &v.segUP[j]);
            if (rc == LGT_OK) rc = alloc_dirs(e, lr, e->link_bytes / 4, &v.segM[j], &v.segMN[j], "adjoint norms");
        }
        if (rc != LGT_OK) {
            free_vjp(e, true);
            return rc;
        }
        v.seg = seg;
        v.nck = nck;
    }
    v.nsteps = nsteps;
    v.sp = sp;
    v.fwd = false;
    v.cot = false;
    TRY(ensure_alt(e));
    if (mode == 2) TRY(copy_dirs(e, v.fixedP, e->cur, DirRange{1, 3}));
    // Steps before the last segment smear cur -> alt (and swap); the step just
    // before a checkpoint also keeps its |M| (the backward then emits the
    // tangent cotangent across the segment boundary itself).  The last
    // segment's steps keep their outputs and |M| in the segment buffers, so
    // the backward starts without recomputing it; its final step writes alt.
    const int n0_last = (nck - 1) * seg;
    v.top_saved = false;
    for (int n = 0; n < nsteps; n++) {
        if (n % seg == 0) TRY(copy_dirs(e, v.ckP[n / seg], e->cur, lr));
        if (n < n0_last) {
            if ((n + 1) % seg == 0) {
                TRY(run(e, LGT_K_APE, "ape", [&] {
                    return DISPATCH(e, ape_save(e->stream, ugeo(e, LGT_K_APE, 0), e->cur, e->alt, v.ckMN[(n + 1) / seg],
                                                mode, sp));
                }));
            } else {
                TRY(run(e, LGT_K_APE, "ape",
                        [&] { return DISPATCH(e, ape(e->stream, ugeo(e, LGT_K_APE, 0), e->cur, e->alt, mode, sp)); }));
            }
        } else {
            const RawPlanes& in = (n == n0_last) ? e->cur : v.segUP[n - n0_last - 1];
            const RawPlanes& out = (n == nsteps - 1) ? e->alt : v.segUP[n - n0_last];
            TRY(run(e, LGT_K_APE, "ape", [&] {
                return DISPATCH(e, ape_save(e->stream, ugeo(e, LGT_K_APE, 0), in, out, v.segMN[n - n0_last], mode, sp));
            }));
        }
        if (n < n0_last || n == nsteps - 1)
            for (int mu = lr.lo; mu <= lr.hi; mu++)
                for (int p = 0; p < 2; p++) std::swap(e->cur.p[mu][p], e->alt.p[mu][p]);
    }
    v.top_saved = true;
    v.fwd = true;
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_wloop_sum(lgt_engine* e, int R, int T, double sum[2])
{
    if (!e || !sum) return fail(LGT_EINVAL, "lgt_wloop_sum: null argument");
    TRY(check_loop(e, R, T));
    TRY(set_dev(e));
    const int ncols = 6 * (int)nbx(e);
    std::vector<int> slots(T, -1);
    slots[T - 1] = 0;
    WGroupHost gh{R, T, 0, 0};
    size_t part_bytes = sizeof(double) * 2 * (size_t)ncols;
    size_t off = (part_bytes + 255) & ~(size_t)255;
    TRY(ensure(e, &e->partial, &e->partial_bytes, off + sizeof(gh) + sizeof(int) * T, "partials"));
    WGroupHost* dg = (WGroupHost*)((char*)e->partial + off);
    int* ds = (int*)((char*)e->partial + off + sizeof(gh));
    HIP_TRY(hipMemcpyAsync(dg, &gh, sizeof(gh), hipMemcpyHostToDevice, e->stream), "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(ds, slots.data(), sizeof(int) * T, hipMemcpyHostToDevice, e->stream), "hipMemcpyAsync");
    TRY(ensure_alt(e));
    TRY(run_wloops(e, std::vector<WGroupHost>{gh}, dg, ds, ncols));
    return reduce_to_host_dd(e, 1, ncols, sum);
}

static int field_to_host(lgt_engine* e, const RawPlanes& P, double* aos)
{
    size_t bytes = sizeof(double) * 16 * (size_t)e->vol;
    TRY(ensure(e, &e->staging, &e->staging_bytes, bytes, "staging"));
    TRY(run(e, LGT_K_FIELD, "to_aos", [&] { return DISPATCH(e, to_aos(e->stream, e->geo, P, (double*)e->staging)); }));
    return d2h(e, aos, e->staging, bytes);
}

int lgt_wloop_grad(lgt_engine* e, int R, int T, double* grad)
{
    if (!e) return fail(LGT_EINVAL, "lgt_wloop_grad: null engine");
    TRY(check_loop(e, R, T));
    TRY(set_dev(e));
    VjpState& v = e->vjp;
    if (!v.gobj) TRY(alloc_field(e, &v.gobj, &v.gobjP));
    // T temporal and 3R spatial line fields of VOL links
    TRY(ensure(e, &v.lines, &v.lines_bytes, (size_t)(T + 3 * R) * e->vol * e->link_bytes, "loop lines"));
    TRY(run(e, LGT_K_WLOOP_GRAD, "loop lines",
            [&] { return DISPATCH(e, lines(e->stream, kgeo(e, LGT_K_WLOOP_GRAD, 0), e->cur, v.lines, R, T)); }));
    TRY(run(e, LGT_K_WLOOP_GRAD, "wilson-loop gradient",
            [&] { return DISPATCH(e, wloop_grad(e->stream, kgeo(e, LGT_K_WLOOP_GRAD, 0), v.lines, R, T, v.gobjP)); }));
    v.cot = true;
    if (grad) return field_to_host(e, v.gobjP, grad);
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_ape_vjp_backward(lgt_engine* e, const double* cotangent, double* grad)
{
    if (!e) return fail(LGT_EINVAL, "lgt_ape_vjp_backward: null engine");
    VjpState& v = e->vjp;
    if (!v.fwd) return fail(LGT_EINVAL, "lgt_ape_vjp_backward: no forward pass (call lgt_ape_vjp_forward)");
    TRY(set_dev(e));
    if (cotangent) {
        if (!v.gobj) TRY(alloc_field(e, &v.gobj, &v.gobjP));
        size_t bytes = sizeof(double) * 16 * (size_t)e->vol;
        TRY(ensure(e, &e->staging, &e->staging_bytes, bytes, "staging"));
        TRY(h2d(e, e->staging, cotangent, bytes));
        TRY(run(e, LGT_K_FIELD, "from_aos",
                [&] { return DISPATCH(e, from_aos(e->stream, e->geo, v.gobjP, (const double*)e->staging)); }));
        v.cot = true;
    }
    if (!v.cot) return fail(LGT_EINVAL, "lgt_ape_vjp_backward: no cotangent (pass one or call lgt_wloop_grad)");
    int h = 0;   // gP[h]: the running cotangent
    const double sp = v.sp;
    for (int s = v.nck - 1; s >= 0; s--) {
        const int n0 = s * v.seg, n1 = std::min(n0 + v.seg, v.nsteps);
        if (!(s == v.nck - 1 && v.top_saved)) {
            v.top_saved = false;
            // below the top segment the last step's output and |M| are not
            // read (the next segment starts from its checkpoint and the |M|
            // the forward kept), so that step is not recomputed
            const int n_end = (s == v.nck - 1) ? n1 : n1 - 1;
            for (int n = n0; n < n_end; n++) {
                const RawPlanes& in = (n == n0) ? v.ckP[s] : v.segUP[n - n0 - 1];
                TRY(run(e, LGT_K_APE_SAVE, "ape recompute", [&] {
                    return DISPATCH(e, ape_save(e->stream, ugeo(e, LGT_K_APE_SAVE, 0), in, v.segUP[n - n0], v.segMN[n - n0],
                                                v.mode, sp));
                }));
            }
        }
        if (s == v.nck - 1) {
            TRY(run(e, LGT_K_VJP, "adjoint tangent", [&] {
                return DISPATCH(e, vjp_tangent(e->stream, e->geo, e->cur, v.gobjP, v.segMN[n1 - 1 - n0], v.gP[h], v.mode));
            }));
        }
        for (int n = n1 - 1; n >= n0; n--) {
            const RawPlanes& Z = (n == n0) ? v.ckP[s] : v.segUP[n - n0 - 1];
            // past the bottom of a segment (s >= 1) the |M| kept by the forward
            const RawNorms* mprev = (n > n0) ? &v.segMN[n - 1 - n0] : (s > 0 ? &v.ckMN[s] : nullptr);
            TRY(run(e, LGT_K_VJP, "adjoint step", [&] {
                return DISPATCH(e, vjp_step(e->stream, ugeo(e, LGT_K_VJP, 0), Z, v.gP[h], mprev, v.gP[1 - h], v.mode, sp));
            }));
            h = 1 - h;
        }
    }
    // the carried links: the adjoint's result; the temporal links of spatial
    // smearing (in no staple): the cotangent itself
    RawPlanes out = v.gP[h];
    if (v.mode == 0) {
        out.p[0][0] = v.gobjP.p[0][0];
        out.p[0][1] = v.gobjP.p[0][1];
    }
    if (grad) return field_to_host(e, out, grad);
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

int lgt_heatbath_harness(int n, const double* staples, double beta, const double* uniforms, const int* offsets,
                         const int* lens, double* links, int* counts)
{
    if (n <= 0 || !staples || !uniforms || !offsets || !lens || !links || !counts)
        return fail(LGT_EINVAL, "lgt_heatbath_harness: bad arguments");
    int64_t nuni = 0;
    for (int i = 0; i < n; i++) {
        if (offsets[i] < 0 || lens[i] < 0) return fail(LGT_EINVAL, "lgt_heatbath_harness: negative offset");
        nuni = std::max<int64_t>(nuni, (int64_t)offsets[i] + lens[i]);
    }
    double *ds = nullptr, *du = nullptr, *dl = nullptr;
    int *doff = nullptr, *dlen = nullptr, *dc = nullptr;
    int rc = LGT_OK;
    hipError_t err = hipSuccess;
    if ((err = hipMalloc(&ds, sizeof(double) * 4 * n)) != hipSuccess || (err = hipMalloc(&du, sizeof(double) * (nuni + 1))) != hipSuccess ||
        (err = hipMalloc(&dl, sizeof(double) * 4 * n)) != hipSuccess || (err = hipMalloc(&doff, sizeof(int) * n)) != hipSuccess ||
        (err = hipMalloc(&dlen, sizeof(int) * n)) != hipSuccess || (err = hipMalloc(&dc, sizeof(int) * n)) != hipSuccess) {
        rc = fail(LGT_ENOMEM, "lgt_heatbath_harness: hipMalloc: %s", hipGetErrorString(err));
    }
    if (rc == LGT_OK) {
        hipMemcpy(ds, staples, sizeof(double) * 4 * n, hipMemcpyHostToDevice);
        hipMemcpy(du, uniforms, sizeof(double) * nuni, hipMemcpyHostToDevice);
        hipMemcpy(doff, offsets, sizeof(int) * n, hipMemcpyHostToDevice);
        hipMemcpy(dlen, lens, sizeof(int) * n, hipMemcpyHostToDevice);
        err = launch_hb_harness(nullptr, n, ds, beta, du, doff, dlen, dl, dc);
        if (err == hipSuccess) err = hipDeviceSynchronize();
        if (err == hipSuccess) err = hipMemcpy(links, dl, sizeof(double) * 4 * n, hipMemcpyDeviceToHost);
        if (err == hipSuccess) err = hipMemcpy(counts, dc, sizeof(int) * n, hipMemcpyDeviceToHost);
        if (err != hipSuccess) rc = hip_fail(err, "lgt_heatbath_harness");
    }
    hipFree(ds); hipFree(du); hipFree(dl); hipFree(doff); hipFree(dlen); hipFree(dc);
    return rc;
}

int lgt_overrelax_harness(int n, const double* staples, const double* links_in, double* links_out)
{
    if (n <= 0 || !staples || !links_in || !links_out) return fail(LGT_EINVAL, "lgt_overrelax_harness: bad arguments");
    double *ds = nullptr, *di = nullptr, *dl = nullptr;
    int rc = LGT_OK;
    hipError_t err = hipSuccess;
    if ((err = hipMalloc(&ds, sizeof(double) * 4 * n)) != hipSuccess || (err = hipMalloc(&di, sizeof(double) * 4 * n)) != hipSuccess ||
        (err = hipMalloc(&dl, sizeof(double) * 4 * n)) != hipSuccess)
        rc = fail(LGT_ENOMEM, "lgt_overrelax_harness: hipMalloc: %s", hipGetErrorString(err));
    if (rc == LGT_OK) {
        hipMemcpy(ds, staples, sizeof(double) * 4 * n, hipMemcpyHostToDevice);
        hipMemcpy(di, links_in, sizeof(double) * 4 * n, hipMemcpyHostToDevice);
        err = launch_or_harness(nullptr, n, ds, di, dl);
        if (err == hipSuccess) err = hipDeviceSynchronize();
        if (err == hipSuccess) err = hipMemcpy(links_out, dl, sizeof(double) * 4 * n, hipMemcpyDeviceToHost);
        if (err != hipSuccess) rc = hip_fail(err, "lgt_overrelax_harness");
    }
    hipFree(ds); hipFree(di); hipFree(dl);
    return rc;
}

int lgt_metropolis_harness(int n, const double* staples, const double* links_in, double beta, double eps,
                           const double* uniforms, double* links_out, int* accepted)
{
    if (n <= 0 || !staples || !links_in || !uniforms || !links_out || !accepted)
        return fail(LGT_EINVAL, "lgt_metropolis_harness: bad arguments");
    double *ds = nullptr, *di = nullptr, *du = nullptr, *dl = nullptr;
    int* da = nullptr;
    int rc = LGT_OK;
    hipError_t err = hipSuccess;
    if ((err = hipMalloc(&ds, sizeof(double) * 4 * n)) != hipSuccess || (err = hipMalloc(&di, sizeof(double) * 4 * n)) != hipSuccess ||
        (err = hipMalloc(&du, sizeof(double) * 4 * n)) != hipSuccess || (err = hipMalloc(&dl, sizeof(double) * 4 * n)) != hipSuccess ||
        (err = hipMalloc(&da, sizeof(int) * n)) != hipSuccess)
        rc = fail(LGT_ENOMEM, "lgt_metropolis_harness: hipMalloc: %s", hipGetErrorString(err));
    if (rc == LGT_OK) {
        err = hipMemcpy(ds, staples, sizeof(double) * 4 * n, hipMemcpyHostToDevice);
        if (err == hipSuccess) err = hipMemcpy(di, links_in, sizeof(double) * 4 * n, hipMemcpyHostToDevice);
        if (err == hipSuccess) err = hipMemcpy(du, uniforms, sizeof(double) * 4 * n, hipMemcpyHostToDevice);
        if (err == hipSuccess) err = launch_metro_harness(nullptr, n, ds, di, beta, eps, du, dl, da);
        if (err == hipSuccess) err = hipDeviceSynchronize();
        if (err == hipSuccess) err = hipMemcpy(links_out, dl, sizeof(double) * 4 * n, hipMemcpyDeviceToHost);
        if (err == hipSuccess) err = hipMemcpy(accepted, da, sizeof(int) * n, hipMemcpyDeviceToHost);
        if (err != hipSuccess) rc = hip_fail(err, "lgt_metropolis_harness");
    }
    hipFree(ds); hipFree(di); hipFree(du); hipFree(dl); hipFree(da);
    return rc;
}

int lgt_profile_read(lgt_engine* e, int kernel_class, double* ms_total, int64_t* launches)
{
    if (!e || kernel_class < 0 || kernel_class >= LGT_K_COUNT) return fail(LGT_EINVAL, "lgt_profile_read: class");
    TRY(set_dev(e));
    TRY(flush_profile(e));
    if (ms_total) *ms_total = e->prof_ms[kernel_class];
    if (launches) *launches = e->prof_n[kernel_class];
    return LGT_OK;
}

int lgt_profile_reset(lgt_engine* e)
{
    if (!e) return fail(LGT_EINVAL, "lgt_profile_reset: null engine");
    TRY(set_dev(e));
    TRY(flush_profile(e));
    for (int k = 0; k < LGT_K_COUNT; k++) {
        e->prof_ms[k] = 0.0;
        e->prof_n[k] = 0;
    }
    return LGT_OK;
}

int lgt_sync(lgt_engine* e)
{
    if (!e) return fail(LGT_EINVAL, "lgt_sync: null engine");
    TRY(set_dev(e));
    HIP_TRY(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return LGT_OK;
}

}  // extern "C"

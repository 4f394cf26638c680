// C ABI of libxtreme_hip.so: operator handle, operator apply, FGMRES with the inner-GMRES right
// preconditioner, and the side-by-side solve of several right-hand sides.
//
// The Krylov solver restates Extreme.Fgmres (Gmres.cs, FlexibleGmresSolver.cs,
// ClassicGmresSolver.cs, FlexibleGmresWithGmresPreconditioner.cs) with the basis vectors in
// HBM: the host keeps the (m+1)^2 Hessenberg matrix and the Givens rotations in FP64 and
// drives the kernels; every Gram-Schmidt pass costs one device->host copy of the j+1
// projections and the new norm, which is exactly what the iterated classical Gram-Schmidt
// test (Gmres.cs:297-300) needs to decide on a second pass.
#include "xtr_green.h"
#include "xtr_kernels.h"
#include "xtr_learned.h"
#include "xtr_common.h"
#include "../../include/xtreme_hip.h"

#include <array>
#include <map>
#include <tuple>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

using cd = std::complex<double>;

namespace {

thread_local std::string g_last_error;

constexpr double kMu0 = 4.0 * 3.14159265358979323846 * 1.0E-07;   // OmegaModelUtils.cs:12
const double kEps0 = 1.0 / (kMu0 * 299792458.0 * 299792458.0);      // OmegaModelUtils.cs:13

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return -1;
}

void require(bool ok, const std::string &msg)
{
    if (!ok) throw std::invalid_argument(msg);
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t count = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    void alloc(size_t n)
    {
        release();
        if (n) XTR_CHECK(hipMalloc((void **)&p, n * sizeof(T)));
        count = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        count = 0;
    }
    ~DevBuf() { release(); }
    void swap(DevBuf &o)
    {
        std::swap(p, o.p);
        std::swap(count, o.count);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Pinned host memory that kernels write directly (mapped into the device address space,
// fine-grained coherent): small results the host reads after the stream has passed the writer.
template <typename T>
struct MappedBuf {
    T *p = nullptr;     // host address
    T *dev = nullptr;   // device address of the same memory
    explicit MappedBuf(size_t n)
    {
        XTR_CHECK(hipHostMalloc((void **)&p, n * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
        XTR_CHECK(hipHostGetDevicePointer((void **)&dev, p, 0));
    }
    ~MappedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    MappedBuf(const MappedBuf &) = delete;
    MappedBuf &operator=(const MappedBuf &) = delete;
};

// FFT tile choices measured per process: (precision, nx, ny, nz) -> tile of each pass (the number
// of right-hand sides only repeats the same workgroups along the grid's y).
struct TileKey {
    int p, nx, ny, nz;
    bool operator<(const TileKey &o) const { return std::tie(p, nx, ny, nz) < std::tie(o.p, o.nx, o.ny, o.nz); }
};
std::map<TileKey, std::array<int, 4>> &tile_cache()
{
    static std::map<TileKey, std::array<int, 4>> m;
    return m;
}
std::mutex &tile_cache_mu()
{
    static std::mutex m;
    return m;
}

// Small-nz contraction kernel measured per process: (precision, nx, ny, nz, right-hand sides) ->
// CONTRACT_VECTOR or CONTRACT_MFMA (guarded by tile_cache_mu as well).
struct ContractKey {
    int p, nx, ny, nz, nrhs;
    bool operator<(const ContractKey &o) const
    {
        return std::tie(p, nx, ny, nz, nrhs) < std::tie(o.p, o.nx, o.ny, o.nz, o.nrhs);
    }
};
std::map<ContractKey, int> &contract_cache()
{
    static std::map<ContractKey, int> m;
    return m;
}

// Shortest decimal that reads back as the same double (what .NET's default double formatting
// prints in the reference's messages, up to the exponent's spelling).
std::string shortest(double v)
{
    char b[40];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(b, sizeof(b), "%.*g", p, v);
        if (std::strtod(b, nullptr) == v) break;
    }
    return b;
}

// Waits for everything submitted to `s` so far by polling an event: the host resumes as soon as
// the GPU passes the marker instead of being woken from a blocking wait.
inline void spin_wait(hipStream_t s, hipEvent_t ev)
{
    XTR_CHECK(hipEventRecord(ev, s));
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    XTR_CHECK(e);
}

}  // namespace

namespace {
struct Slot;   // Krylov storage, stream and reduction scratch of one right-hand side (below)
}

struct xtr_op {
    xtr::Geom g{};
    int precision = XTR_FP64;
    int active = XTR_FP64;        // arithmetic of the next applies when precision == XTR_MIXED
    int max_rhs = 1;
    int type = -1;
    double omega = 0;
    double dx = 0, dy = 0;
    std::vector<double> depth, thick;
    xtr::Section section;
    int number_of_hankels = 10;
    std::vector<double> hank0, hank1;

    int device = 0;               // the device current at xtr_op_create (the stream's)
    hipStream_t stream = nullptr;
    DevBuf<double2> G64;          // [nblk][gblk] FP64 Green tensor
    DevBuf<float2> G32;           // [nblk][gblk] complex64 copy for XTR_FP32 and XTR_MIXED
    bool green_ready = false;
    bool folded = false;          // G holds blocks kx <= Nx, ky <= Ny only (nblk = Lq, else L)
    int64_t nblk() const { return folded ? g.Lq() : g.L(); }
    DevBuf<double2> R;            // [nx*ny*nz] of model 0
    std::vector<std::unique_ptr<DevBuf<double2>>> Rm;   // R of models 1 .. nmodels-1 (same Green tensor)
    int nmodels = 1;

    const double2 *model_r(int m) const
    {
        require(m >= 0 && m < nmodels, "model index out of range (see xtr_op_prepare_models)");
        return m == 0 ? R.p : Rm[m - 1]->p;
    }
    DevBuf<double> f, b;          // [nz]
    DevBuf<uint16_t> packed_rc;   // [S][2]
    DevBuf<double2> tw64_x, tw64_y;
    DevBuf<float2> tw32_x, tw32_y;
    DevBuf<unsigned char> buf;    // max_rhs * M spectrum
    DevBuf<double2> stage_x, stage_y;   // host-pointer apply staging
    DevBuf<double2> Go[2];        // AtoO level tensors [L][ncomp][nz]: [0] magnetic, [1] electric
    bool go_ready[2] = {false, false};
    DevBuf<double2> lvl_spec, lvl_small;   // level-field scratch: L * 3nz and L * 3

    // Per-stage timing: one set of six events per apply while profiling is on, summed on demand.
    bool profiling = false;
    std::array<int, 4> last_tiles{{-1, -1, -1, -1}};   // FFT column tiles of the latest apply
    std::vector<std::array<hipEvent_t, 6>> ev_sets;
    std::vector<int> ev_nrhs;                        // right-hand sides of the apply each set timed
    size_t ev_used = 0;
    std::mutex apply_mu;

    // Per right-hand side: a stream, the Krylov bases of the outer FGMRES and the inner GMRES, the
    // refinement vectors and the reduction scratch, allocated on first use and kept for later solves
    // (grown when a solve needs more), so that a solve allocates nothing once the handle has run one
    // of its shape. Slot kMaxRhs holds the FP64 norms of the mixed-precision refinement.
    std::vector<std::unique_ptr<Slot>> slots;
    Slot &slot(int r);

    ~xtr_op();

    hipEvent_t *next_events(int nrhs)
    {
        if (ev_used == ev_sets.size()) {
            std::array<hipEvent_t, 6> set{};
            for (auto &e : set) XTR_CHECK(hipEventCreate(&e));
            ev_sets.push_back(set);
            ev_nrhs.push_back(0);
        }
        ev_nrhs[ev_used] = nrhs;
        return ev_sets[ev_used++].data();
    }

    size_t elem() const { return precision == XTR_FP32 ? sizeof(float2) : sizeof(double2); }
    int arith() const { return precision == XTR_MIXED ? active : precision; }

    // Column tiles of the four FFT passes for this apply's precision and grid: the
    // XTR_FFT_TILES override (four letters d/n/w/m/r/s: default, narrow, wide, narrow with 512
    // threads, register-resident stages, register-resident stages with the other column count),
    // else the choice measured once per process and shape. Measuring runs whole applies on this
    // apply's own buffers (they rewrite the spectrum and y, which the apply then recomputes from
    // x; x and y never overlap, which apply_raw checks).
    std::array<int, 4> fft_tiles(xtr::ApplyArgs &a, hipStream_t s)
    {
        const int P[4] = {2 * g.ny, 2 * g.nx, 2 * g.nx, 2 * g.ny};
        if (const char *e = std::getenv("XTR_FFT_TILES")) {
            std::array<int, 4> t{};
            for (int k = 0; k < 4; ++k) {
                const char c = e[0] && std::strlen(e) == 4 ? e[k] : 'd';
                t[k] = c == 'w'   ? xtr::FFT_TILE_WIDE
                       : c == 'n' ? xtr::FFT_TILE_NARROW
                       : c == 'm' ? xtr::FFT_TILE_NARROW_T512
                       : c == 'r' ? xtr::FFT_TILE_REG
                       : c == 's' ? xtr::FFT_TILE_REG_ALT
                                  : xtr::FFT_TILE_DEFAULT;
            }
            return t;
        }
        const TileKey key{a.precision, g.nx, g.ny, g.nz};
        {
            std::lock_guard<std::mutex> lk(tile_cache_mu());
            auto it = tile_cache().find(key);
            if (it != tile_cache().end()) return it->second;
        }
        using Pass = void (*)(const xtr::ApplyArgs &, hipStream_t);
        // the apply's five launches in order; the FFT passes are stages 0, 1, 3 and 4
        const Pass stage[5] = {xtr::launch_fwd_y, xtr::launch_fwd_x, xtr::launch_contract, xtr::launch_inv_x,
                               xtr::launch_inv_y};
        const int stage_of_pass[4] = {0, 1, 3, 4};
        hipEvent_t e0, e1;
        XTR_CHECK(hipEventCreate(&e0));
        XTR_CHECK(hipEventCreate(&e1));
        std::array<int, 4> best{};
        for (int k = 0; k < 4; ++k) a.fft_tile[k] = xtr::FFT_TILE_DEFAULT;
        try {
            // Each candidate is timed inside whole applies (one warm-up, three timed), so the
            // contraction's stream of the Green tensor between the passes leaves the caches as a
            // solve does; repeating one pass alone would time it on cache-resident data.
            for (int k = 0; k < 4; ++k) {
                int opts[8];
                const int nopt = xtr::fft_tile_options(P[k], a.precision, opts);
                float best_ms = 0;
                for (int o = 0; o < nopt; ++o) {
                    const int t = opts[o];
                    a.fft_tile[k] = t;
                    float total = 0;
                    for (int r = 0; r < 4; ++r) {
                        for (int st = 0; st < 5; ++st) {
                            if (st == stage_of_pass[k]) XTR_CHECK(hipEventRecord(e0, s));
                            stage[st](a, s);
                            if (st == stage_of_pass[k]) XTR_CHECK(hipEventRecord(e1, s));
                        }
                        XTR_CHECK(hipEventSynchronize(e1));
                        float ms = 0;
                        XTR_CHECK(hipEventElapsedTime(&ms, e0, e1));
                        if (r > 0) total += ms;
                    }
                    if (o == 0 || total < best_ms) {
                        best_ms = total;
                        best[k] = t;
                    }
                }
                a.fft_tile[k] = xtr::FFT_TILE_DEFAULT;
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::lock_guard<std::mutex> lk(tile_cache_mu());
        tile_cache()[key] = best;
        return best;
    }

    // Contraction kernel of a folded small-nz apply (contract_foldn_kernel or contract_foldm_kernel),
    // measured once per process and shape the way fft_tiles measures tiles: whole applies on this
    // apply's own buffers, one warm-up and three timed per kernel, the contraction stage timed.
    int contract_variant(xtr::ApplyArgs &a, hipStream_t s)
    {
        if (!xtr::contract_variants_apply(a)) return xtr::CONTRACT_RULE;
        const ContractKey key{a.precision, g.nx, g.ny, g.nz, a.nrhs};
        {
            std::lock_guard<std::mutex> lk(tile_cache_mu());
            auto it = contract_cache().find(key);
            if (it != contract_cache().end()) return it->second;
        }
        using Pass = void (*)(const xtr::ApplyArgs &, hipStream_t);
        const Pass stage[5] = {xtr::launch_fwd_y, xtr::launch_fwd_x, xtr::launch_contract, xtr::launch_inv_x,
                               xtr::launch_inv_y};
        hipEvent_t e0, e1;
        XTR_CHECK(hipEventCreate(&e0));
        XTR_CHECK(hipEventCreate(&e1));
        int best = xtr::CONTRACT_RULE;
        try {
            float best_ms = 0;
            for (int variant : {xtr::CONTRACT_VECTOR, xtr::CONTRACT_MFMA}) {
                a.contract_variant = variant;
                float total = 0;
                for (int r = 0; r < 4; ++r) {
                    for (int st = 0; st < 5; ++st) {
                        if (st == 2) XTR_CHECK(hipEventRecord(e0, s));
                        stage[st](a, s);
                        if (st == 2) XTR_CHECK(hipEventRecord(e1, s));
                    }
                    XTR_CHECK(hipEventSynchronize(e1));
                    float ms = 0;
                    XTR_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0) total += ms;
                }
                if (best == xtr::CONTRACT_RULE || total < best_ms) {
                    best_ms = total;
                    best = variant;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::lock_guard<std::mutex> lk(tile_cache_mu());
        contract_cache()[key] = best;
        return best;
    }

    // One operator apply for nrhs vectors on stream s (ConvolutionOperator.CalculateOperatorAorKr);
    // vec32: the vectors are complex64 (the FP32 apply of a mixed-precision correction solve).
    void apply_raw(const xtr::RhsPtrs &v, int nrhs, hipStream_t s, bool vec32 = false)
    {
        require(green_ready, "the Green tensor is not set");
        require(type == XTR_OP_A || type == XTR_OP_CHI0, "the operator is not prepared");
        require(nrhs >= 1 && nrhs <= max_rhs, "nrhs out of range");
        require(!vec32 || arith() == XTR_FP32, "complex64 vectors need the FP32 apply");
        // the last pass reads x while it writes y, element by element, and the tile measurement
        // repeats that pass: y must not overlap any x
        const size_t vb = (size_t)g.n() * (vec32 ? sizeof(float2) : sizeof(double2));
        for (int r = 0; r < nrhs; ++r)
            for (int q = 0; q < nrhs; ++q) {
                const char *xb = (const char *)v.x[q], *yb = (const char *)v.y[r];
                require(yb + vb <= xb || xb + vb <= yb, "the result vector overlaps an input vector");
            }
        std::lock_guard<std::mutex> lk(apply_mu);
        const int p = arith();
        xtr::ApplyArgs a{};
        a.g = g;
        a.precision = p;
        a.nrhs = nrhs;
        a.type = type;
        a.R = R.p;
        a.f = f.p;
        a.b = b.p;
        a.G = p == XTR_FP64 ? (const void *)G64.p : (const void *)G32.p;
        a.folded = folded ? 1 : 0;
        a.packed_rc = packed_rc.p;
        a.buf = buf.p;
        a.tw_x = p == XTR_FP64 ? (const void *)tw64_x.p : (const void *)tw32_x.p;
        a.tw_y = p == XTR_FP64 ? (const void *)tw64_y.p : (const void *)tw32_y.p;
        a.vec32 = vec32 ? 1 : 0;
        a.v = v;
        last_tiles = fft_tiles(a, s);
        for (int k = 0; k < 4; ++k) a.fft_tile[k] = last_tiles[k];
        a.contract_variant = contract_variant(a, s);
        hipEvent_t *ev = profiling ? next_events(nrhs) : nullptr;
        if (ev) XTR_CHECK(hipEventRecord(ev[0], s));
        xtr::launch_fwd_y(a, s);
        if (ev) XTR_CHECK(hipEventRecord(ev[1], s));
        xtr::launch_fwd_x(a, s);
        if (ev) XTR_CHECK(hipEventRecord(ev[2], s));
        xtr::launch_contract(a, s);
        if (ev) XTR_CHECK(hipEventRecord(ev[3], s));
        xtr::launch_inv_x(a, s);
        if (ev) XTR_CHECK(hipEventRecord(ev[4], s));
        xtr::launch_inv_y(a, s);
        if (ev) XTR_CHECK(hipEventRecord(ev[5], s));
    }
};

namespace {

std::vector<double2> twiddles64(int p)
{
    std::vector<double2> t(p);
    for (int k = 0; k < p; ++k) {
        const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)p;
        t[k] = make_double2(std::cos(a), -std::sin(a));
    }
    return t;
}

std::vector<float2> twiddles32(int p)
{
    auto d = twiddles64(p);
    std::vector<float2> t(p);
    for (int k = 0; k < p; ++k) t[k] = make_float2((float)d[k].x, (float)d[k].y);
    return t;
}

template <typename T>
void upload(DevBuf<T> &d, const std::vector<T> &h)
{
    d.alloc(h.size());
    XTR_CHECK(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// ModelUtils.FindCorrespondingBackgroundLayerIndex (Model/ModelUtils.cs:35-52), in exact micrometres
int background_layer(const xtr::Section &s, double start, double end)
{
    const int i = xtr::background_layer_index(s, start, end);
    if (i < 0) throw std::invalid_argument("an anomaly layer is not inside one background layer");
    return i;
}

cd zeta_of(double sigma, double omega) { return cd(sigma, -omega * kEps0); }   // OmegaModelUtils.cs:41-42

// ---------------------------------------------------------------------------------------------
// FGMRES / GMRES on device vectors
// ---------------------------------------------------------------------------------------------

// One operator apply of a solve's vector x into y: complex128 vectors, or complex64 ones in the
// correction solves of the mixed-precision refinement.
using ApplyFn = std::function<void(const void *x, void *y)>;

struct SolveCtx {
    hipStream_t stream;
    ApplyFn apply;
    int64_t matvecs = 0;
    int64_t dots = 0;
    xtr_iteration_fn on_iter = nullptr;
    void *user = nullptr;
    int rhs_index = 0;
    // reduction scratch shared by the outer and inner solvers of one solve
    DevBuf<double2> partial_c;   // [32][kRedBlocks]
    DevBuf<double> partial_r;    // [kRedBlocks]
    DevBuf<double2> h_dev;       // [32] projections
    DevBuf<double> norm_dev;     // [1]
    // pairs (v_j, z_j) of the flexible preconditioner, copied out while count < cap
    double2 *const *harvest_v = nullptr;
    double2 *const *harvest_z = nullptr;
    int harvest_cap = 0, harvest_count = 0;
    MappedBuf<double2> h_host{32};   // the projections, written by the reduction kernel
    MappedBuf<double> nn_host{1};    // the squared norm, likewise
    hipEvent_t ev = nullptr;
    SolveCtx() : partial_c(32 * xtr::kRedBlocks), partial_r(xtr::kRedBlocks), h_dev(32), norm_dev(1)
    {
        XTR_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    ~SolveCtx()
    {
        if (ev) (void)hipEventDestroy(ev);
    }
    SolveCtx(const SolveCtx &) = delete;
    SolveCtx &operator=(const SolveCtx &) = delete;
    void wait() { spin_wait(stream, ev); }
    // ready for a new solve on stream s: counters, callbacks and harvest cleared
    void reset(hipStream_t s)
    {
        stream = s;
        apply = nullptr;
        matvecs = dots = 0;
        on_iter = nullptr;
        user = nullptr;
        rhs_index = 0;
        harvest_v = harvest_z = nullptr;
        harvest_cap = harvest_count = 0;
    }
};

// The vectors of one Gmres instance: basis V (m columns), flexible basis Z (m columns, FGMRES
// only), the work vector w and the residual r0.
template <typename VT>
struct GmresBufs {
    DevBuf<VT> V, Z, w, r0;
};

template <typename VT>
inline void grow(DevBuf<VT> &b, size_t need)
{
    if (b.count < need) b.alloc(need);
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;   // the solve's vector is ready for the shared apply (ApplyGate)
    SolveCtx ctx;
    GmresBufs<double2> outer, inner;
    GmresBufs<float2> outer32, inner32;   // complex64 bases of the mixed-precision correction solves
    DevBuf<double2> r, d;         // refinement residual and correction (complex128)
    DevBuf<float2> r32, d32;      // the same in complex64, for correction solves on complex64 vectors
    template <typename VT> GmresBufs<VT> &outer_bufs();
    template <typename VT> GmresBufs<VT> &inner_bufs();
    Slot()
    {
        XTR_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        const hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamDefault);
        if (e != hipSuccess) {
            (void)hipEventDestroy(ready);
            XTR_CHECK(e);
        }
    }
    ~Slot()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (ready) (void)hipEventDestroy(ready);
    }
    Slot(const Slot &) = delete;
    Slot &operator=(const Slot &) = delete;
};

}  // namespace

template <> GmresBufs<double2> &Slot::outer_bufs<double2>() { return outer; }
template <> GmresBufs<double2> &Slot::inner_bufs<double2>() { return inner; }
template <> GmresBufs<float2> &Slot::outer_bufs<float2>() { return outer32; }
template <> GmresBufs<float2> &Slot::inner_bufs<float2>() { return inner32; }

Slot &xtr_op::slot(int r)
{
    require(r >= 0 && r <= xtr::kMaxRhs, "right-hand side slot out of range");
    if ((int)slots.size() <= r) slots.resize(r + 1);
    if (!slots[r]) slots[r] = std::make_unique<Slot>();
    return *slots[r];
}

xtr_op::~xtr_op()
{
    if (stream) (void)hipStreamSynchronize(stream);
    slots.clear();
    for (auto &set : ev_sets)
        for (auto &e : set)
            if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

// Learned right preconditioner (no counterpart in the reference): the network's device state.
struct xtr_net {
    xtr::NetGeom g{};
    hipStream_t stream = nullptr;
    DevBuf<unsigned short> x0, h, dyb, dpb, W1, W2, WT2;
    DevBuf<float> pre, y, tgt, dy, dh, dp, part;
    DevBuf<float> w1, w2, m1, m2, s1, s2, g1, g2;   // FP32 master weights, Adam moments, gradients
    DevBuf<double> partial_r, nn, lpart;
    DevBuf<float> feat;                               // [cells][2] feature channels, when set
    bool has_feat = false;
    float phi = 0.0f;                                 // the frequency feature (log10 f), 0 without features
    int step = 0;
    std::mutex mu;
    // spectral branch (K = 0: none): FP32 master weights [2][NM][Q][Q][2], Adam moments, gradient,
    // BF16 copies, twiddle tables and the transform buffers
    xtr::SpecGeom sg{};
    DevBuf<float> sw, sm, sv, sgrad;
    DevBuf<unsigned short> swb;
    DevBuf<float2> twx, twy, sT, sU, sY, sgY;
    xtr::SpecBuffers sbufs() { return xtr::SpecBuffers{twx.p, twy.p, sT.p, sU.p, sY.p, sgY.p, swb.p}; }

    ~xtr_net()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
    xtr::NetBuffers bufs()
    {
        return xtr::NetBuffers{x0.p,  h.p,    pre.p, y.p,  tgt.p,  dy.p,  dyb.p,
                               dh.p,  dp.p,   dpb.p, part.p, W1.p, W2.p, WT2.p, has_feat ? feat.p : nullptr};
    }
};

namespace {

// z = the network applied to v (complex n-vectors on the device), on stream s
void net_precondition(xtr_net *net, const double2 *v, double2 *z, hipStream_t s)
{
    const int64_t n = net->g.cells * 3;
    xtr::launch_norm2(v, n, net->partial_r.p, s);
    xtr::launch_reduce_r(net->partial_r.p, net->nn.p, s);
    const xtr::NetBuffers b = net->bufs();
    xtr::net_pack(net->g, v, nullptr, net->nn.p, b, s);
    xtr::net_forward(net->g, b, s);
    if (net->sg.K > 0) xtr::spec_forward(net->sg, b, net->sbufs(), net->phi, s);
    xtr::net_unpack(net->g, v, net->nn.p, b, z, s);
}

// Gradient of the mean relative loss over the samples into net->g1, net->g2; returns the mean loss.
double net_gradients(xtr_net *net, int ns, const double2 *const *v, const double2 *const *z)
{
    const int64_t n = net->g.cells * 3;
    const hipStream_t s = net->stream;
    const xtr::NetBuffers b = net->bufs();
    const int nb = xtr::net_loss_blocks(net->g);
    std::vector<double> lp(2 * (size_t)nb);
    double total = 0;
    for (int k = 0; k < ns; ++k) {
        xtr::launch_norm2(v[k], n, net->partial_r.p, s);
        xtr::launch_reduce_r(net->partial_r.p, net->nn.p, s);
        xtr::net_pack(net->g, v[k], z[k], net->nn.p, b, s);
        xtr::net_forward(net->g, b, s);
        if (net->sg.K > 0) xtr::spec_forward(net->sg, b, net->sbufs(), net->phi, s);
        xtr::net_loss(net->g, b, net->lpart.p, s);
        XTR_CHECK(hipMemcpyAsync(lp.data(), net->lpart.p, lp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        XTR_CHECK(hipStreamSynchronize(s));
        double se = 0, st = 0;
        for (int i = 0; i < nb; ++i) {
            se += lp[2 * i];
            st += lp[2 * i + 1];
        }
        const float scale = st > 0 ? (float)(1.0 / (st * ns)) : 0.0f;
        xtr::net_backward(net->g, b, net->g1.p, net->g2.p, scale, k > 0, s);
        if (net->sg.K > 0) xtr::spec_backward(net->sg, b, net->sbufs(), net->phi, net->sgrad.p, scale, k > 0, s);
        total += st > 0 ? se / st : 0.0;
    }
    XTR_CHECK(hipStreamSynchronize(s));
    return total / ns;
}

// GmresSolver (Gmres.cs) on vectors of element type VT: double2 (complex128, the reference's
// arithmetic) or float2 (complex64, the correction solves of the mixed-precision refinement). Dots,
// norms, the Hessenberg matrix and the Givens rotations are FP64 either way.
template <typename VT>
struct Gmres {
    int64_t n;
    int m;
    bool flexible;
    double tol = 1e-10;
    int max_repeat = 1;
    int backward_check = 0;
    SolveCtx *ctx;
    Gmres<VT> *inner = nullptr;
    xtr_net *net = nullptr;       // learned right preconditioner in place of the inner GMRES
    VT *V = nullptr, *Z = nullptr, *w = nullptr, *r0 = nullptr;   // in the slot's GmresBufs
    std::vector<cd> H;      // (m+1) x (m+1), column-major (FortranMatrix.cs:11-52)
    std::vector<double> gc;
    std::vector<cd> gs;
    int iter_counter = 0;

    Gmres(int64_t n_, int m_, bool flex, SolveCtx *c, GmresBufs<VT> &b)
        : n(n_), m(m_), flexible(flex), ctx(c), H((size_t)(m_ + 1) * (m_ + 1)), gc(m_), gs(m_)
    {
        require(m_ >= 1 && m_ <= 31, "buffer length must be within 1..31");
        grow(b.V, (size_t)n_ * m_);
        if (flex) grow(b.Z, (size_t)n_ * m_);
        grow(b.w, (size_t)n_);
        grow(b.r0, (size_t)n_);
        V = b.V.p;
        Z = flex ? b.Z.p : nullptr;
        w = b.w.p;
        r0 = b.r0.p;
    }

    cd &h(int i, int j) { return H[(size_t)j * (m + 1) + i]; }
    VT *col(VT *B, int j) { return B + (int64_t)j * n; }

    void matvec(const VT *x, VT *y)
    {
        ctx->apply(x, y);
        ctx->matvecs++;
    }

    double norm(const VT *v)
    {
        xtr::launch_norm2(v, n, ctx->partial_r.p, ctx->stream);
        xtr::launch_reduce_r(ctx->partial_r.p, ctx->norm_dev.p, ctx->stream, ctx->nn_host.dev);
        ctx->wait();
        ctx->dots++;
        return std::sqrt(ctx->nn_host.p[0]);
    }

    // ClassicalGramSchmidtIteration (Gmres.cs:302-313) fused with the norm of the new w.
    void cgs(int jH, double &dloo, double &dnormw)
    {
        const int cnt = jH + 1;
        xtr::launch_block_dots(V, cnt, w, n, ctx->partial_c.p, ctx->stream);
        xtr::launch_reduce_c(ctx->partial_c.p, cnt, ctx->h_dev.p, ctx->stream, ctx->h_host.dev);
        xtr::launch_cgs_update(V, cnt, ctx->h_dev.p, w, n, ctx->partial_r.p, ctx->stream);
        xtr::launch_reduce_r(ctx->partial_r.p, ctx->norm_dev.p, ctx->stream, ctx->nn_host.dev);
        ctx->wait();
        double s = 0;
        for (int k = 0; k < cnt; ++k) {
            const cd d(ctx->h_host.p[k].x, ctx->h_host.p[k].y);
            h(k, jH) += d;                              // Zaxpy(jH+1, One, dots, H[:,jH])
            s += d.real() * d.real() + d.imag() * d.imag();
        }
        dloo = s;                                       // Dznrm2^2
        dnormw = std::sqrt(ctx->nn_host.p[0]);
        ctx->dots += cnt + 1;
    }

    // zrot on one pair (green_tensor.c:21-28)
    static void zrot1(cd &x, cd &y, double c, cd s)
    {
        const cd t = c * x + s * y;
        y = c * y - std::conj(s) * x;
        x = t;
    }

    // reference BLAS zrotg (green_tensor.c:30-33)
    static void zrotg(cd a, cd b, double &c, cd &s)
    {
        const double aa = std::abs(a);
        if (aa == 0.0) {
            c = 0.0;
            s = 1.0;
            return;
        }
        const double scale = aa + std::abs(b);
        const double t1 = std::abs(a / scale), t2 = std::abs(b / scale);
        const double nrm = scale * std::sqrt(t1 * t1 + t2 * t2);
        const cd alpha = a / aa;
        c = aa / nrm;
        s = alpha * std::conj(b) / nrm;
    }

    // ApplyPreconditionerRight + MatrixVectorMult: w = A M^-1 v_jH (Gmres.cs:239-241)
    void precond_and_apply(int jH)
    {
        if (flexible) {
            if (net) {
                if constexpr (std::is_same<VT, double2>::value)
                    net_precondition(net, col(V, jH), col(Z, jH), ctx->stream);
                else
                    throw std::invalid_argument("the learned preconditioner runs on complex128 vectors");
            } else {
                double bea, be;
                int conv;
                inner->solve(col(V, jH), col(V, jH), col(Z, jH), bea, be, conv);   // FlexibleGmresWithGmresPreconditioner.cs:57
            }
            if (ctx->harvest_count < ctx->harvest_cap) {
                const int k = ctx->harvest_count++;
                XTR_CHECK(hipMemcpyAsync(ctx->harvest_v[k], col(V, jH), n * sizeof(VT), hipMemcpyDeviceToDevice,
                                         ctx->stream));
                XTR_CHECK(hipMemcpyAsync(ctx->harvest_z[k], col(Z, jH), n * sizeof(VT), hipMemcpyDeviceToDevice,
                                         ctx->stream));
            }
            matvec(col(Z, jH), w);
        } else {
            matvec(col(V, jH), w);
        }
    }

    // ConstructCurrentSolution (Gmres.cs:187-196): y = H^-1 g, x += B y
    void construct(int jH, VT *x)
    {
        std::vector<cd> y(jH);
        for (int i = 0; i < jH; ++i) y[i] = h(i, m);
        for (int i = jH - 1; i >= 0; --i) {          // SimplifiedZtrsv (green_tensor.c:34-37)
            cd t = y[i];
            for (int k = i + 1; k < jH; ++k) t -= h(i, k) * y[k];
            y[i] = t / h(i, i);
        }
        xtr::Coef32 c{};
        for (int i = 0; i < jH; ++i) c.c[i] = make_double2(y[i].real(), y[i].imag());
        if (jH > 0) xtr::launch_update_solution(flexible ? Z : V, jH, c, x, n, ctx->stream);
    }

    // CheckBackwardError (Gmres.cs:200-230)
    double check_backward_error(double bn, const VT *rhs, const VT *x)
    {
        matvec(x, w);
        xtr::launch_sub(rhs, w, w, n, ctx->stream);
        const double be = norm(w) / bn;
        if (be >= tol && backward_check == 2)
            throw std::runtime_error("be = " + shortest(be) + ", tolerance = " + shortest(tol));
        return be;
    }

    double final_actions(int jH, double bn, const VT *rhs, VT *x)
    {
        construct(jH, x);
        if (backward_check != 0) return check_backward_error(bn, rhs, x);
        return -1;
    }

    // KrylovSubspaceConstruction (Gmres.cs:233-295)
    double krylov(double bn, int &jH)
    {
        double bea = -1;
        do {
            precond_and_apply(jH);
            for (int j = 0; j <= jH; ++j) h(j, jH) = 0;
            int north = 0;
            double dloo = 0, dnormw = 0;
            do {
                ++north;
                cgs(jH, dloo, dnormw);
                if (!std::isfinite(dnormw) || !std::isfinite(dloo))
                    throw std::runtime_error("a Krylov vector holds NaN or Inf: the operator produced non-finite values");
            } while ((2.0 * dnormw <= std::sqrt(dloo)) && (north < 3));
            h(jH + 1, jH) = dnormw;
            if (jH < m - 1) xtr::launch_div(w, dnormw, col(V, jH + 1), n, ctx->stream);
            for (int j = 0; j < jH; ++j) zrot1(h(j, jH), h(j + 1, jH), gc[j], gs[j]);
            zrotg(h(jH, jH), h(jH + 1, jH), gc[jH], gs[jH]);
            zrot1(h(jH, m), h(jH + 1, m), gc[jH], gs[jH]);
            zrot1(h(jH, jH), h(jH + 1, jH), gc[jH], gs[jH]);
            h(jH + 1, jH) = 0;
            bea = std::abs(h(jH + 1, m)) / bn;
            ++iter_counter;
            if (flexible && ctx->on_iter) ctx->on_iter(ctx->user, ctx->rhs_index, iter_counter, bea);
            ++jH;
        } while (bea > tol && jH < m);
        return bea;
    }

    // PrepareRestart, ComputeTheTrue (Gmres.cs:145-162)
    double restart(const VT *rhs, VT *x, int jH)
    {
        construct(jH, x);
        matvec(x, r0);
        xtr::launch_sub(rhs, r0, w, n, ctx->stream);
        return norm(w);
    }

    // GmresSolver.Solve (Gmres.cs:77-143), initial guess supplied
    void solve(const VT *rhs, const VT *x0, VT *result, double &bea_out, double &be_out, int &conv)
    {
        iter_counter = 0;
        const double bn = norm(rhs);
        // NaN or Inf would end the iteration as if converged (bea > tol is false): refuse instead
        if (!std::isfinite(bn)) throw std::runtime_error("the right-hand side holds NaN or Inf");
        if (bn == 0) {
            bea_out = 0; be_out = 0; conv = 1;
            return;
        }
        matvec(x0, r0);
        xtr::launch_sub(rhs, r0, w, n, ctx->stream);
        double beta = norm(w);
        if (!std::isfinite(beta)) throw std::runtime_error("the initial residual holds NaN or Inf (operator or initial guess)");
        if (beta == 0) {   // CheckForExactSolution: ResultInfo.Converged(0, 0, 1) (Gmres.cs:104-107)
            if (result != x0) XTR_CHECK(hipMemcpyAsync(result, x0, n * sizeof(VT), hipMemcpyDeviceToDevice, ctx->stream));
            bea_out = 0; be_out = 0; conv = 1;
            iter_counter = 1;
            return;
        }
        double bea = -1, be = -1;
        int iter = 0;
        if (result != x0) XTR_CHECK(hipMemcpyAsync(result, x0, n * sizeof(VT), hipMemcpyDeviceToDevice, ctx->stream));
        for (;;) {
            xtr::launch_mul(w, 1 / beta, col(V, 0), n, ctx->stream);
            h(0, m) = beta;
            for (int i = 1; i <= m; ++i) h(i, m) = 0;
            int jH = 0;
            bea = krylov(bn, jH);
            if (bea < tol) {
                be = final_actions(jH, bn, rhs, result);
                bea_out = bea; be_out = be; conv = 1;
                return;
            }
            ++iter;
            if (iter >= max_repeat) break;
            beta = restart(rhs, result, jH);
        }
        be = final_actions(m, bn, rhs, result);
        bea_out = bea; be_out = be; conv = 0;
    }
};

// FlexibleGmresWithGmresPreconditioner (:29-58) on one right-hand side.
template <typename VT>
void run_fgmres(int64_t n, const xtr_gmres_params &p, Slot &sl, const VT *b, VT *x, xtr_result_info &info,
                xtr_net *net = nullptr)
{
    SolveCtx &ctx = sl.ctx;
    if (!std::is_same<VT, double2>::value)
        require(ctx.harvest_cap == 0, "pairs are harvested from complex128 solves only");
    require(p.outer_buffer_length >= 1 && p.outer_buffer_length <= 31, "OuterBufferLength must be 1..31");
    require(p.inner_buffer_length >= 1 && p.inner_buffer_length <= 31, "InnerBufferLength must be 1..31");
    require(p.max_repeat_number >= 1, "MaxRepeatsNumber must be >= 1");
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<Gmres<VT>> inner;
    if (!net) {
        inner = std::make_unique<Gmres<VT>>(n, p.inner_buffer_length, false, &ctx, sl.inner_bufs<VT>());
        inner->tol = p.tolerance;
        inner->max_repeat = 1;
        inner->backward_check = 0;
    }
    Gmres<VT> outer(n, p.outer_buffer_length, true, &ctx, sl.outer_bufs<VT>());
    outer.tol = p.tolerance;
    outer.max_repeat = p.max_repeat_number;
    outer.backward_check = p.backward_error_checking;
    outer.inner = inner.get();
    outer.net = net;
    double bea = -1, be = -1;
    int conv = 0;
    outer.solve(b, x, x, bea, be, conv);
    XTR_CHECK(hipStreamSynchronize(ctx.stream));
    auto t1 = std::chrono::steady_clock::now();
    info.converged = conv;
    info.arnoldi_backward_error = bea;
    info.backward_error = be;
    info.number_of_iterations = outer.iter_counter;
    info.matvecs = ctx.matvecs;
    info.dot_products = ctx.dots;
    info.seconds = std::chrono::duration<double>(t1 - t0).count();
}

// Pairs up the operator applies of solves that run side by side on their own streams: the
// apply fires when every still-running solve has submitted one vector, so the Green tensor
// is streamed once per round for all of them.
struct ApplyGate {
    xtr_op *op;
    bool vec32;                   // the solves' vectors are complex64
    std::mutex mu;
    std::condition_variable cv;
    int active;
    int pending = 0;
    uint64_t round = 0;
    struct Sub { const void *x; void *y; hipStream_t s; hipEvent_t ready; const double2 *R; };
    std::vector<Sub> subs;
    hipEvent_t done = nullptr;
    std::exception_ptr err;

    ApplyGate(xtr_op *o, int n, bool v32) : op(o), vec32(v32), active(n)
    {
        XTR_CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    }
    ~ApplyGate()
    {
        if (done) (void)hipEventDestroy(done);
    }

    void fire()
    {
        try {
            xtr::RhsPtrs v{};
            for (size_t i = 0; i < subs.size(); ++i) {
                XTR_CHECK(hipStreamWaitEvent(op->stream, subs[i].ready, 0));
                v.x[i] = (const double2 *)subs[i].x;
                v.y[i] = (double2 *)subs[i].y;
                v.R[i] = subs[i].R;
            }
            op->apply_raw(v, (int)subs.size(), op->stream, vec32);
            XTR_CHECK(hipEventRecord(done, op->stream));
            for (auto &s : subs) XTR_CHECK(hipStreamWaitEvent(s.s, done, 0));
        } catch (...) {
            err = std::current_exception();
        }
        subs.clear();
        pending = 0;
        ++round;
        cv.notify_all();
    }

    void submit(const void *x, void *y, hipStream_t s, hipEvent_t ready, const double2 *R)
    {
        XTR_CHECK(hipEventRecord(ready, s));
        std::unique_lock<std::mutex> lk(mu);
        if (err) std::rethrow_exception(err);
        subs.push_back({x, y, s, ready, R});
        ++pending;
        const uint64_t mine = round;
        if (pending == active) fire();
        else cv.wait(lk, [&] { return round != mine; });
        if (err) std::rethrow_exception(err);
    }

    void leave()
    {
        std::unique_lock<std::mutex> lk(mu);
        --active;
        if (pending > 0 && pending == active) fire();
    }
};

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================

// Side-by-side FGMRES solves of nrhs right-hand sides with one ApplyGate (shared applies). Right-hand
// side r runs on its own host thread with the handle's slot r (stream, bases, scratch).
template <typename VT = double2>
static void solve_batch(xtr_op *op, const xtr_gmres_params *params, int nrhs, const void *const *b, void *const *x,
                        xtr_result_info *info, xtr_iteration_fn on_iteration, void *user, const int *model = nullptr)
{
    ApplyGate gate(op, nrhs, std::is_same<VT, float2>::value);
    std::vector<Slot *> sl(nrhs, nullptr);
    std::vector<std::exception_ptr> errs(nrhs);
    std::mutex cb_mu;
    struct CbUser { xtr_iteration_fn fn; void *user; std::mutex *mu; };
    CbUser cbu{on_iteration, user, &cb_mu};
    auto cb = [](void *u, int r, int it, double bea) {
        auto *c = (CbUser *)u;
        std::lock_guard<std::mutex> lk(*c->mu);
        c->fn(c->user, r, it, bea);
    };
    try {
        for (int r = 0; r < nrhs; ++r) {
            sl[r] = &op->slot(r);
            SolveCtx &c = sl[r]->ctx;
            c.reset(sl[r]->stream);
            c.rhs_index = r;
            if (on_iteration) {
                c.on_iter = cb;
                c.user = &cbu;
            }
            hipStream_t s = sl[r]->stream;
            hipEvent_t ev = sl[r]->ready;
            ApplyGate *gp = &gate;
            const double2 *Rp = model ? op->model_r(model[r]) : nullptr;   // null: the handle's R
            c.apply = [gp, s, ev, Rp](const void *xx, void *yy) { gp->submit(xx, yy, s, ev, Rp); };
        }
        // inputs written on the caller's default stream must be visible to the solve streams
        XTR_CHECK(hipDeviceSynchronize());
        std::vector<std::thread> th;
        for (int r = 0; r < nrhs; ++r) {
            th.emplace_back([&, r] {
                try {
                    std::memset(&info[r], 0, sizeof(info[r]));
                    run_fgmres<VT>(op->g.n(), *params, *sl[r], (const VT *)b[r], (VT *)x[r], info[r]);
                } catch (...) {
                    errs[r] = std::current_exception();
                }
                gate.leave();
            });
        }
        for (auto &t : th) t.join();
        XTR_CHECK(hipStreamSynchronize(op->stream));
    } catch (...) {
        for (auto *p : sl)
            if (p) (void)hipStreamSynchronize(p->stream);
        throw;
    }
    for (auto *p : sl) XTR_CHECK(hipStreamSynchronize(p->stream));
    for (auto &e : errs)
        if (e) std::rethrow_exception(e);
}

extern "C" {

int xtr_last_error(char *buf, int len)
{
    const int n = (int)g_last_error.size();
    if (buf && len > 0) {
        const int c = n < len - 1 ? n : len - 1;
        std::memcpy(buf, g_last_error.data(), c);
        buf[c] = 0;
    }
    return n;
}

int xtr_op_create(const xtr_op_desc *d, xtr_op **out)
{
    return guarded([&] {
        require(d && out, "null argument");
        require(d->nx > 0 && d->ny > 0 && d->nz > 0, "grid sizes must be positive");
        require(3 * d->nz * 1 <= 1024, "Nz too large for the contraction kernel");
        require(xtr::fft_length_supported(2 * d->nx) && xtr::fft_length_supported(2 * d->ny),
                "2*Nx and 2*Ny must be one of 4,6,8,12,16,24,32,48,64,96,128,192,256,384,512");
        require(d->precision == XTR_FP64 || d->precision == XTR_FP32 || d->precision == XTR_MIXED,
                "precision must be XTR_FP64, XTR_FP32 or XTR_MIXED");
        require(d->max_rhs >= 1 && d->max_rhs <= xtr::kMaxRhs, "max_rhs must be 1..8");
        require(3 * d->nz * d->max_rhs <= 1024, "3*Nz*max_rhs must not exceed 1024");
        require(d->nlayers >= 3, "the section needs the air, at least one layer and the lower half-space");
        auto op = std::make_unique<xtr_op>();
        op->g = xtr::Geom{d->nx, d->ny, d->nz};
        op->precision = d->precision;
        op->max_rhs = d->max_rhs;
        op->omega = d->omega;
        op->dx = d->dx;
        op->dy = d->dy;
        op->depth.assign(d->depth, d->depth + d->nz);
        op->thick.assign(d->thick, d->thick + d->nz);
        op->section.zero_level = d->zero_level;
        op->section.thickness.assign(d->layer_thickness, d->layer_thickness + d->nlayers);
        op->section.sigma.assign(d->layer_sigma, d->layer_sigma + d->nlayers);
        op->number_of_hankels = d->number_of_hankels;
        if (d->hank0 && d->hank1 && d->ntaps > 0) {
            op->hank0.assign(d->hank0, d->hank0 + d->ntaps);
            op->hank1.assign(d->hank1, d->hank1 + d->ntaps);
        }
        for (int k = 0; k < d->nz; ++k) background_layer(op->section, op->depth[k], op->depth[k] + op->thick[k]);
        XTR_CHECK(hipGetDevice(&op->device));
        XTR_CHECK(hipStreamCreateWithFlags(&op->stream, hipStreamDefault));
        const xtr::Geom &g = op->g;
        // packed (rc, tr) table in CombineSymm order (TranceiverUtils.cs:65-74)
        std::vector<uint16_t> prc(2 * g.S());
        for (int tr = 0; tr < g.nz; ++tr)
            for (int rc = 0; rc <= tr; ++rc) {
                const int o = rc + tr * (tr + 1) / 2;
                prc[2 * o] = (uint16_t)rc;
                prc[2 * o + 1] = (uint16_t)tr;
            }
        upload(op->packed_rc, prc);
        upload(op->tw64_x, twiddles64(2 * g.nx));
        upload(op->tw64_y, twiddles64(2 * g.ny));
        upload(op->tw32_x, twiddles32(2 * g.nx));
        upload(op->tw32_y, twiddles32(2 * g.ny));
        op->buf.alloc((size_t)op->max_rhs * g.M() * op->elem());
        op->R.alloc((size_t)g.nx * g.ny * g.nz);
        op->f.alloc(g.nz);
        op->b.alloc(g.nz);
        *out = op.release();
    });
}

void xtr_op_destroy(xtr_op *op)
{
    if (op) {
        (void)hipStreamSynchronize(op->stream);
        delete op;
    }
}

static xtr::GreenInput green_input(const xtr_op *op)
{
    xtr::GreenInput in;
    in.g = op->g;
    in.dx = op->dx;
    in.dy = op->dy;
    in.depth = op->depth;
    in.thick = op->thick;
    in.section = op->section;
    in.omega = op->omega;
    in.number_of_hankels = op->number_of_hankels;
    in.hank0 = op->hank0;
    in.hank1 = op->hank1;
    return in;
}

static void finish_green(xtr_op *op)
{
    if (op->precision == XTR_FP32 || op->precision == XTR_MIXED) {
        op->G32.alloc(op->G64.count);
        xtr::launch_to_f32(op->G64.p, op->G32.p, (int64_t)op->G64.count, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
        if (op->precision == XTR_FP32) op->G64.release();
    }
    op->green_ready = true;
}

// XTR_GREEN_LAYOUT=full keeps every wavenumber's block (to compare the two layouts).
static bool green_fold_allowed()
{
    const char *e = std::getenv("XTR_GREEN_LAYOUT");
    return !(e && std::strcmp(e, "full") == 0);
}

// Takes a full FP64 tensor [L][gblk] in device layout and stores it folded (fold) or as it is.
static void store_green(xtr_op *op, DevBuf<double2> &full, bool fold)
{
    const xtr::Geom &g = op->g;
    op->G32.release();
    op->folded = fold;
    if (fold) {
        op->G64.alloc((size_t)(g.Lq() * g.gblk()));
        xtr::launch_green_fold(full.p, op->G64.p, g, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
        full.release();
    } else {
        op->G64.swap(full);
        full.release();
    }
    finish_green(op);
}

int xtr_op_set_green(xtr_op *op, const xtr_complex *xx, const xtr_complex *xy, const xtr_complex *xz,
                     const xtr_complex *yy, const xtr_complex *yz, const xtr_complex *zz)
{
    return guarded([&] {
        require(op && xx && xy && xz && yy && yz && zz, "null argument");
        const xtr::Geom &g = op->g;
        const int64_t S = g.S(), A = (int64_t)g.nz * g.nz, L = g.L();
        const xtr_complex *host[6] = {xx, xy, xz, yy, yz, zz};
        const int64_t sizes[6] = {S, S, A, S, A, S};
        DevBuf<double2> tmp[6];
        const double2 *dev[6];
        for (int c = 0; c < 6; ++c) {
            tmp[c].alloc((size_t)(L * sizes[c]));
            XTR_CHECK(hipMemcpy(tmp[c].p, host[c], L * sizes[c] * sizeof(double2), hipMemcpyHostToDevice));
            dev[c] = tmp[c].p;
        }
        op->green_ready = false;
        op->G64.release();
        op->G32.release();
        DevBuf<double2> full((size_t)(L * g.gblk()));
        xtr::launch_green_interleave(dev, full.p, g, op->stream);
        for (auto &t : tmp) t.release();
        // fold only a tensor with the mirror symmetry: every mirrored block within 1e-12 of its
        // source, relative to the component's largest element
        bool fold = green_fold_allowed();
        if (fold) {
            DevBuf<double> stat(12);
            xtr::launch_green_mirror_check(full.p, g, stat.p, op->stream);
            double h[12];
            XTR_CHECK(hipMemcpyAsync(h, stat.p, sizeof(h), hipMemcpyDeviceToHost, op->stream));
            XTR_CHECK(hipStreamSynchronize(op->stream));
            for (int c = 0; c < 6; ++c) fold = fold && h[2 * c] <= 1e-12 * h[2 * c + 1];
        } else {
            XTR_CHECK(hipStreamSynchronize(op->stream));
        }
        store_green(op, full, fold);
    });
}

int xtr_op_build_green(xtr_op *op)
{
    return guarded([&] {
        require(op, "null argument");
        require(op->hank0.size() == 431 && op->hank1.size() == 431, "the 431-tap Hankel filter is required");
        require(op->number_of_hankels > 0, "NumberOfHankels must be positive");
        const xtr::Geom &g = op->g;
        op->green_ready = false;
        op->G64.release();
        op->G32.release();
        DevBuf<double2> full((size_t)(g.L() * g.gblk()));
        xtr::build_green_atoa(green_input(op), full.p, op->tw64_x.p, op->tw64_y.p, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
        // the assembled tensor has the mirror symmetry by construction (signed mirroring)
        store_green(op, full, green_fold_allowed());
    });
}

int xtr_op_get_green(xtr_op *op, xtr_complex *xx, xtr_complex *xy, xtr_complex *xz, xtr_complex *yy,
                     xtr_complex *yz, xtr_complex *zz)
{
    return guarded([&] {
        require(op && xx && xy && xz && yy && yz && zz, "null argument");
        require(op->green_ready, "the Green tensor is not set");
        const xtr::Geom &g = op->g;
        const int64_t S = g.S(), A = (int64_t)g.nz * g.nz, L = g.L();
        xtr_complex *host[6] = {xx, xy, xz, yy, yz, zz};
        const int64_t sizes[6] = {S, S, A, S, A, S};
        DevBuf<double2> tmp[6];
        double2 *dev[6];
        for (int c = 0; c < 6; ++c) {
            tmp[c].alloc((size_t)(L * sizes[c]));
            dev[c] = tmp[c].p;
        }
        DevBuf<double2> wide, full;
        const double2 *src = op->G64.p;
        if (op->precision == XTR_FP32) {
            // widen the complex64 copy for export
            std::vector<float2> h32(op->G32.count);
            XTR_CHECK(hipMemcpy(h32.data(), op->G32.p, h32.size() * sizeof(float2), hipMemcpyDeviceToHost));
            std::vector<double2> h64(h32.size());
            for (size_t i = 0; i < h32.size(); ++i) h64[i] = xtr::to_d(h32[i]);
            wide.alloc(h64.size());
            XTR_CHECK(hipMemcpy(wide.p, h64.data(), h64.size() * sizeof(double2), hipMemcpyHostToDevice));
            src = wide.p;
        }
        if (op->folded) {
            full.alloc((size_t)(L * g.gblk()));
            xtr::launch_green_unfold(src, full.p, g, op->stream);
            src = full.p;
        }
        xtr::launch_green_deinterleave(src, dev, g, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
        for (int c = 0; c < 6; ++c)
            XTR_CHECK(hipMemcpy(host[c], dev[c], L * sizes[c] * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int xtr_op_prepare(xtr_op *op, int type, const xtr_complex *zeta)
{
    const xtr_complex *z[1] = {zeta};
    return xtr_op_prepare_models(op, type, 1, z);
}

int xtr_op_prepare_models(xtr_op *op, int type, int nmodels, const xtr_complex *const *zeta)
{
    return guarded([&] {
        require(op && zeta, "null argument");
        require(type == XTR_OP_A || type == XTR_OP_CHI0, "type must be XTR_OP_A or XTR_OP_CHI0");
        require(nmodels >= 1 && nmodels <= xtr::kMaxRhs, "nmodels must be 1..8");
        for (int m = 0; m < nmodels; ++m) require(zeta[m] != nullptr, "null zeta");
        const xtr::Geom &g = op->g;
        const int nz = g.nz;
        // GetBackgroundZeta (ConvolutionOperator.cs:80-88): the background, so shared by the models
        std::vector<cd> z0(nz);
        for (int k = 0; k < nz; ++k) {
            const int li = background_layer(op->section, op->depth[k], op->depth[k] + op->thick[k]);
            z0[k] = zeta_of(op->section.sigma[li], op->omega);
        }
        // PrepareForwardFactors / PrepareBackwardFactors (:90-111)
        std::vector<double> f(nz), b(nz);
        const double nf = 1 / ((double)g.nx * (double)g.ny * 4);
        for (int k = 0; k < nz; ++k) {
            f[k] = std::sqrt(z0[k].real());
            b[k] = 2 * std::sqrt(z0[k].real()) / op->thick[k] * nf;
        }
        XTR_CHECK(hipStreamSynchronize(op->stream));
        while ((int)op->Rm.size() < nmodels - 1) op->Rm.push_back(std::make_unique<DevBuf<double2>>((size_t)g.nx * g.ny * nz));
        // PrepareRFunction (:113-143) per model
        std::vector<double2> R((size_t)g.nx * g.ny * nz);
        for (int m = 0; m < nmodels; ++m) {
            for (int64_t idx = 0; idx < (int64_t)R.size(); ++idx) {
                const int k = (int)(idx % nz);
                const cd z(zeta[m][idx].re, zeta[m][idx].im);
                const cd cz0 = std::conj(z0[k]);
                const cd r = type == XTR_OP_CHI0 ? std::sqrt(cd(z0[k].real(), 0)) / (z + cz0) : (z - z0[k]) / (z + cz0);
                R[idx] = make_double2(r.real(), r.imag());
            }
            double2 *dst = m == 0 ? op->R.p : op->Rm[m - 1]->p;
            XTR_CHECK(hipMemcpy(dst, R.data(), R.size() * sizeof(double2), hipMemcpyHostToDevice));
        }
        XTR_CHECK(hipMemcpy(op->f.p, f.data(), nz * sizeof(double), hipMemcpyHostToDevice));
        XTR_CHECK(hipMemcpy(op->b.p, b.data(), nz * sizeof(double), hipMemcpyHostToDevice));
        op->type = type;
        op->nmodels = nmodels;
    });
}

int xtr_op_apply_models_device(xtr_op *op, int nrhs, const int *model, const void *const *x, void *const *y)
{
    return guarded([&] {
        require(op && model && x && y, "null argument");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs out of range");
        xtr::RhsPtrs v{};
        for (int r = 0; r < nrhs; ++r) {
            v.x[r] = (const double2 *)x[r];
            v.y[r] = (double2 *)y[r];
            v.R[r] = op->model_r(model[r]);
        }
        XTR_CHECK(hipDeviceSynchronize());   // inputs written on other streams
        op->apply_raw(v, nrhs, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

int xtr_op_apply_device(xtr_op *op, const void *x, void *y, int nrhs)
{
    return guarded([&] {
        require(op && x && y, "null argument");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs out of range");
        xtr::RhsPtrs v{};
        const int64_t n = op->g.n();
        for (int r = 0; r < nrhs; ++r) {
            v.x[r] = (const double2 *)x + r * n;
            v.y[r] = (double2 *)y + r * n;
        }
        op->apply_raw(v, nrhs, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

int xtr_op_apply_c64_device(xtr_op *op, const void *x, void *y, int nrhs)
{
    return guarded([&] {
        require(op && x && y, "null argument");
        require(op->precision == XTR_FP32 || op->precision == XTR_MIXED,
                "complex64 applies need a handle created with XTR_FP32 or XTR_MIXED");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs out of range");
        xtr::RhsPtrs v{};
        const int64_t n = op->g.n();
        for (int r = 0; r < nrhs; ++r) {
            v.x[r] = (const double2 *)((const float2 *)x + r * n);
            v.y[r] = (double2 *)((float2 *)y + r * n);
        }
        struct ActiveGuard {
            xtr_op *o;
            int prev;
            ~ActiveGuard() { o->active = prev; }
        } guard{op, op->active};
        op->active = XTR_FP32;
        op->apply_raw(v, nrhs, op->stream, true);
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

int xtr_op_apply(xtr_op *op, const xtr_complex *x, xtr_complex *y, int nrhs)
{
    return guarded([&] {
        require(op && x && y, "null argument");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs out of range");
        const int64_t n = op->g.n();
        if (op->stage_x.count < (size_t)(nrhs * n)) {
            op->stage_x.alloc((size_t)(op->max_rhs * n));
            op->stage_y.alloc((size_t)(op->max_rhs * n));
        }
        XTR_CHECK(hipMemcpy(op->stage_x.p, x, nrhs * n * sizeof(double2), hipMemcpyHostToDevice));
        xtr::RhsPtrs v{};
        for (int r = 0; r < nrhs; ++r) {
            v.x[r] = op->stage_x.p + r * n;
            v.y[r] = op->stage_y.p + r * n;
        }
        op->apply_raw(v, nrhs, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
        XTR_CHECK(hipMemcpy(y, op->stage_y.p, nrhs * n * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int xtr_op_set_profiling(xtr_op *op, int enable)
{
    return guarded([&] {
        require(op, "null argument");
        std::lock_guard<std::mutex> lk(op->apply_mu);
        op->profiling = enable != 0;
        op->ev_used = 0;
    });
}

int xtr_op_stage_times(xtr_op *op, double *out5, int64_t *applies)
{
    return xtr_op_stage_times_rhs(op, 0, out5, applies);
}

int xtr_op_stage_times_rhs(xtr_op *op, int nrhs, double *out5, int64_t *applies)
{
    return guarded([&] {
        require(op && out5, "null argument");
        require(nrhs >= 0, "nrhs must be >= 0");
        std::lock_guard<std::mutex> lk(op->apply_mu);
        for (int i = 0; i < 5; ++i) out5[i] = 0;
        int64_t counted = 0;
        for (size_t k = 0; k < op->ev_used; ++k) {
            if (nrhs > 0 && op->ev_nrhs[k] != nrhs) continue;
            ++counted;
            auto &ev = op->ev_sets[k];
            XTR_CHECK(hipEventSynchronize(ev[5]));
            for (int i = 0; i < 5; ++i) {
                float ms = 0;
                XTR_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                out5[i] += ms;
            }
        }
        if (applies) *applies = counted;
    });
}

int xtr_op_fft_tiles(xtr_op *op, int *tiles4)
{
    return guarded([&] {
        require(op && tiles4, "null argument");
        std::lock_guard<std::mutex> lk(op->apply_mu);
        for (int k = 0; k < 4; ++k) tiles4[k] = op->last_tiles[k];
    });
}

int xtr_op_green_layout(xtr_op *op, int *folded, int64_t *blocks)
{
    return guarded([&] {
        require(op, "null argument");
        require(op->green_ready, "the Green tensor is not set");
        if (folded) *folded = op->folded ? 1 : 0;
        if (blocks) *blocks = op->nblk();
    });
}

int xtr_op_green_mirror_check(xtr_op *op, double *out12)
{
    return guarded([&] {
        require(op && out12, "null argument");
        require(op->green_ready && op->G64.p, "the FP64 Green tensor is not set");
        require(!op->folded, "the Green tensor is folded: the check needs every wavenumber's block");
        DevBuf<double> stat(12);
        xtr::launch_green_mirror_check(op->G64.p, op->g, stat.p, op->stream);
        XTR_CHECK(hipMemcpyAsync(out12, stat.p, 12 * sizeof(double), hipMemcpyDeviceToHost, op->stream));
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

int xtr_op_sizes(xtr_op *op, int64_t *n, int64_t *green_bytes, int64_t *buffer_bytes)
{
    return guarded([&] {
        require(op, "null argument");
        if (n) *n = op->g.n();
        const int64_t per = op->precision == XTR_FP64 ? 16 : op->precision == XTR_FP32 ? 8 : 24;
        if (green_bytes) *green_bytes = (int64_t)(op->nblk() * op->g.gblk() * per);
        if (buffer_bytes) *buffer_bytes = (int64_t)(op->g.M() * op->elem());
    });
}

int xtr_probe_stream(const void *src, void *dst, int64_t blocks, int read_kb, int write_kb, void *stream)
{
    return guarded([&] {
        require(src && (dst || write_kb == 0), "null argument");
        xtr::launch_probe_stream(src, dst, blocks, read_kb, write_kb, (hipStream_t)stream);
    });
}

int xtr_fgmres_solve_device(xtr_op *op, const xtr_gmres_params *params, const void *b, void *x,
                            xtr_result_info *info, xtr_iteration_fn on_iteration, void *user)
{
    return guarded([&] {
        require(op && params && b && x && info, "null argument");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before solving");
        Slot &sl = op->slot(0);
        SolveCtx &ctx = sl.ctx;
        ctx.reset(op->stream);
        ctx.on_iter = on_iteration;
        ctx.user = user;
        xtr_op *o = op;
        ctx.apply = [o](const void *xx, void *yy) {
            xtr::RhsPtrs v{};
            v.x[0] = (const double2 *)xx;
            v.y[0] = (double2 *)yy;
            o->apply_raw(v, 1, o->stream);
        };
        std::memset(info, 0, sizeof(*info));
        run_fgmres(op->g.n(), *params, sl, (const double2 *)b, (double2 *)x, *info);
    });
}

int xtr_fgmres_solve(xtr_op *op, const xtr_gmres_params *params, const xtr_complex *b, xtr_complex *x,
                     xtr_result_info *info, xtr_iteration_fn on_iteration, void *user)
{
    return guarded([&] {
        require(op && params && b && x && info, "null argument");
        const int64_t n = op->g.n();
        DevBuf<double2> db(n), dx(n);
        XTR_CHECK(hipMemcpy(db.p, b, n * sizeof(double2), hipMemcpyHostToDevice));
        XTR_CHECK(hipMemcpy(dx.p, x, n * sizeof(double2), hipMemcpyHostToDevice));
        if (xtr_fgmres_solve_device(op, params, db.p, dx.p, info, on_iteration, user) != 0)
            throw std::runtime_error(g_last_error);
        XTR_CHECK(hipMemcpy(x, dx.p, n * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int xtr_fgmres_solve_batch_device(xtr_op *op, const xtr_gmres_params *params, int nrhs, const void *const *b,
                                  void *const *x, xtr_result_info *info, xtr_iteration_fn on_iteration, void *user)
{
    return guarded([&] {
        require(op && params && b && x && info, "null argument");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before solving");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs must be within 1..max_rhs");
        solve_batch(op, params, nrhs, b, x, info, on_iteration, user);
    });
}


int xtr_fgmres_solve_models_device(xtr_op *op, const xtr_gmres_params *params, int nrhs, const int *model,
                                   const void *const *b, void *const *x, xtr_result_info *info,
                                   xtr_iteration_fn on_iteration, void *user)
{
    return guarded([&] {
        require(op && params && model && b && x && info, "null argument");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before solving");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs must be within 1..max_rhs");
        for (int r = 0; r < nrhs; ++r) (void)op->model_r(model[r]);
        solve_batch(op, params, nrhs, b, x, info, on_iteration, user, model);
    });
}

// The refined solve of nrhs right-hand sides; model[r] (when model is not null) names right-hand side
// r's model among those of xtr_op_prepare_models, whose R both the FP64 residual applies and the
// complex64 correction solves use.
static void refined_solve(xtr_op *op, const xtr_refine_params *params, int nrhs, const int *model,
                          const void *const *b, void *const *x, xtr_refine_info *info)
{
    {
        require(op && params && b && x && info, "null argument");
        require(op->precision == XTR_MIXED, "the refined solve needs an operator created with XTR_MIXED");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before solving");
        require(nrhs >= 1 && nrhs <= op->max_rhs, "nrhs must be within 1..max_rhs");
        require(params->tolerance > 0 && params->max_refinements >= 1, "tolerance > 0 and max_refinements >= 1");
        if (model)
            for (int r = 0; r < nrhs; ++r) (void)op->model_r(model[r]);
        const int64_t n = op->g.n();
        const auto t0 = std::chrono::steady_clock::now();
        XTR_CHECK(hipDeviceSynchronize());   // inputs written on other streams
        SolveCtx &ctx = op->slot(xtr::kMaxRhs).ctx;   // reduction scratch for the FP64 norms
        ctx.reset(op->stream);
        const bool vec32 = params->vector_precision == XTR_FP32;
        require(params->vector_precision == XTR_FP64 || vec32, "vector_precision must be XTR_FP64 or XTR_FP32");
        std::vector<DevBuf<double2> *> r(nrhs), d(nrhs);
        std::vector<DevBuf<float2> *> r32(nrhs), d32(nrhs);
        for (int k = 0; k < nrhs; ++k) {
            Slot &sk = op->slot(k);
            grow(sk.r, (size_t)n);
            r[k] = &sk.r;
            if (vec32) {
                grow(sk.r32, (size_t)n);
                grow(sk.d32, (size_t)n);
                r32[k] = &sk.r32;
                d32[k] = &sk.d32;
            } else {
                grow(sk.d, (size_t)n);
                d[k] = &sk.d;
            }
        }
        auto norm = [&](const double2 *v) {
            xtr::launch_norm2(v, n, ctx.partial_r.p, ctx.stream);
            xtr::launch_reduce_r(ctx.partial_r.p, ctx.norm_dev.p, ctx.stream, ctx.nn_host.dev);
            ctx.wait();
            return std::sqrt(ctx.nn_host.p[0]);
        };
        std::vector<double> bn(nrhs), rn(nrhs, 0.0);
        std::vector<int> live;
        for (int k = 0; k < nrhs; ++k) {
            std::memset(&info[k], 0, sizeof(info[k]));
            bn[k] = norm((const double2 *)b[k]);
            live.push_back(k);
        }
        struct ActiveGuard {
            xtr_op *o;
            ~ActiveGuard() { o->active = XTR_FP64; }
        } guard{op};
        for (int step = 0;; ++step) {
            // FP64 true residuals r = b - A x of the unconverged right-hand sides, one batched apply
            op->active = XTR_FP64;
            xtr::RhsPtrs v{};
            for (size_t i = 0; i < live.size(); ++i) {
                v.x[i] = (const double2 *)x[live[i]];
                v.y[i] = r[live[i]]->p;
                if (model) v.R[i] = op->model_r(model[live[i]]);
            }
            op->apply_raw(v, (int)live.size(), op->stream);
            std::vector<int> next;
            for (int k : live) {
                info[k].matvecs64++;
                xtr::launch_sub((const double2 *)b[k], r[k]->p, r[k]->p, n, op->stream);
                rn[k] = bn[k] > 0 ? norm(r[k]->p) : 0.0;
                if (!std::isfinite(rn[k])) throw std::runtime_error("the FP64 residual holds NaN or Inf");
                info[k].residual = bn[k] > 0 ? rn[k] / bn[k] : 0.0;
                if (info[k].residual <= params->tolerance) info[k].converged = 1;
                else if (step < params->max_refinements) next.push_back(k);
            }
            live.swap(next);
            if (live.empty()) break;
            // FGMRES corrections A d = r with the complex64 operator, side by side, from d = 0; on
            // complex64 vectors the residual goes in scaled to unit norm, r / |r|, and the correction
            // comes back scaled by |r|
            op->active = XTR_FP32;
            std::vector<const void *> rb;
            std::vector<void *> dx;
            xtr_gmres_params ip = params->inner;
            double tol_in = 1.0;
            for (int k : live) {
                if (vec32) {
                    xtr::launch_mul(r[k]->p, 1.0 / rn[k], r32[k]->p, n, op->stream);
                    XTR_CHECK(hipMemsetAsync(d32[k]->p, 0, n * sizeof(float2), op->stream));
                    rb.push_back(r32[k]->p);
                    dx.push_back(d32[k]->p);
                } else {
                    XTR_CHECK(hipMemsetAsync(d[k]->p, 0, n * sizeof(double2), op->stream));
                    rb.push_back(r[k]->p);
                    dx.push_back(d[k]->p);
                }
                tol_in = std::min(tol_in, std::max(params->inner_floor, 0.5 * params->tolerance * bn[k] / rn[k]));
            }
            ip.tolerance = tol_in;
            XTR_CHECK(hipStreamSynchronize(op->stream));
            std::vector<xtr_result_info> ii(live.size());
            std::vector<int> ml;
            for (int k : live) ml.push_back(model ? model[k] : 0);
            const int *mp = model ? ml.data() : nullptr;
            if (vec32) solve_batch<float2>(op, &ip, (int)live.size(), rb.data(), dx.data(), ii.data(), nullptr, nullptr, mp);
            else solve_batch<double2>(op, &ip, (int)live.size(), rb.data(), dx.data(), ii.data(), nullptr, nullptr, mp);
            for (size_t i = 0; i < live.size(); ++i) {
                const int k = live[i];
                info[k].refinements++;
                info[k].matvecs32 += ii[i].matvecs;
                info[k].inner_iterations += ii[i].number_of_iterations;
                xtr::Coef32 c{};
                if (vec32) {
                    c.c[0] = make_double2(rn[k], 0.0);
                    xtr::launch_update_solution(d32[k]->p, 1, c, (double2 *)x[k], n, op->stream);   // x += |r| d
                } else {
                    c.c[0] = make_double2(1.0, 0.0);
                    xtr::launch_update_solution(d[k]->p, 1, c, (double2 *)x[k], n, op->stream);     // x += d
                }
            }
        }
        XTR_CHECK(hipStreamSynchronize(op->stream));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int k = 0; k < nrhs; ++k) info[k].seconds = secs;
    }
}

int xtr_refined_solve_batch_device(xtr_op *op, const xtr_refine_params *params, int nrhs, const void *const *b,
                                   void *const *x, xtr_refine_info *info)
{
    return guarded([&] { refined_solve(op, params, nrhs, nullptr, b, x, info); });
}

int xtr_refined_solve_models_device(xtr_op *op, const xtr_refine_params *params, int nrhs, const int *model,
                                    const void *const *b, void *const *x, xtr_refine_info *info)
{
    return guarded([&] {
        require(model, "null argument");
        refined_solve(op, params, nrhs, model, b, x, info);
    });
}


int xtr_op_build_green_atoo(xtr_op *op, double level_z, double x_shift, double y_shift, int electric)
{
    return guarded([&] {
        require(op, "null argument");
        require(op->hank0.size() == 431 && op->hank1.size() == 431, "the 431-tap Hankel filter is required");
        require(op->number_of_hankels > 0, "NumberOfHankels must be positive");
        const int e = electric ? 1 : 0;
        int slot[9];
        const int ncomp = xtr::atoo_components(e, slot);
        op->go_ready[e] = false;
        op->Go[e].alloc((size_t)(op->g.L() * ncomp * op->g.nz));
        require(std::isfinite(x_shift) && std::isfinite(y_shift), "level shift must be finite");
        xtr::build_green_atoo_level(green_input(op), level_z, x_shift, y_shift, e, op->Go[e].p, op->tw64_x.p,
                                    op->tw64_y.p, op->stream);
        op->go_ready[e] = true;
    });
}

int xtr_op_get_green_atoo(xtr_op *op, int electric, xtr_complex *const *comps)
{
    return guarded([&] {
        require(op && comps, "null argument");
        const int e = electric ? 1 : 0;
        require(op->go_ready[e], "the level tensor is not built");
        int slot[9];
        const int ncomp = xtr::atoo_components(e, slot);
        const int64_t L = op->g.L(), nz = op->g.nz;
        std::vector<double2> h((size_t)(L * ncomp * nz));
        XTR_CHECK(hipMemcpy(h.data(), op->Go[e].p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
        for (int t = 0; t < 9; ++t) {
            if (!comps[t] || slot[t] < 0) continue;
            auto *dst = reinterpret_cast<double2 *>(comps[t]);
            for (int64_t l = 0; l < L; ++l)
                std::memcpy(dst + l * nz, h.data() + (l * ncomp + slot[t]) * nz, nz * sizeof(double2));
        }
    });
}

int xtr_op_jq_model_device(xtr_op *op, int model, const void *chi, const void *js, void *jq)
{
    return guarded([&] {
        require(op && chi && js && jq, "null argument");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before computing jQ");
        const double2 *R = op->model_r(model);
        XTR_CHECK(hipDeviceSynchronize());   // chi and js may have been written on other streams
        xtr::launch_jq((const double2 *)chi, (const double2 *)js, R, op->f.p, op->g, (double2 *)jq, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

int xtr_op_jq_device(xtr_op *op, const void *chi, const void *js, void *jq)
{
    if (op && op->nmodels != 1) {
        g_last_error = "the operator holds several models: name one with xtr_op_jq_model_device";
        return -1;
    }
    return xtr_op_jq_model_device(op, 0, chi, js, jq);
}

}  // extern "C"

// ---- allocation-free level field ---------------------------------------------------------------
// Scratch of xtr_op_level_field_noalloc_device: the zero-padded spectrum of jQ (4 Nx Ny * 3 Nz) and
// the level spectrum (4 Nx Ny * 3), complex128, in one file-scope device array of 128 MiB: grids with
// 4 Nx Ny * 3 (Nz + 1) <= 8 Mi elements (128 x 128 x 32 needs 103.8 MB). Each GPU the library runs on
// holds its own copy until the process ends. In host code the array's name is not a device address:
// the host uses only the address hipGetSymbolAddress returns for the current device.
namespace {
constexpr size_t kScratchElems = (size_t)8 << 20;                     // complex128 elements
static __device__ __align__(256) double scratch[2 * kScratchElems];   // re, im interleaved
std::mutex g_scratch_mu;
std::map<int, double2 *> g_scratch_ptr;                               // device ordinal -> address

// The current device's copy of scratch, looked up once per device (with that device current).
double2 *scratch_on_device(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void *p = nullptr;
    XTR_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = (double2 *)p;
    return (double2 *)p;
}
}  // namespace

extern "C" {

int xtr_op_level_field_noalloc_device(xtr_op *op, int electric, const void *jq, void *field)
{
    return guarded([&] {
        require(op && jq && field, "null argument");
        const xtr::Geom &g = op->g;
        const size_t need = (size_t)g.M() + 3 * (size_t)g.L();
        if (need > kScratchElems)
            throw std::invalid_argument("the level field of this grid needs " + std::to_string(need * 16 >> 20) +
                                        " MiB of scratch, more than the 128 MiB of the allocation-free entry point "
                                        "(4 Nx Ny * 3 (Nz + 1) must not exceed 8388608); use xtr_op_level_field_device");
        const int e = electric ? 1 : 0;
        require(op->go_ready[e], "build the level tensor first (xtr_op_build_green_atoo)");
        int dev = 0;
        XTR_CHECK(hipGetDevice(&dev));
        require(dev == op->device, "the handle was created on another device than the current one");
        double2 *ptr = scratch_on_device(dev);
        XTR_CHECK(hipDeviceSynchronize());   // jQ may have been written on another stream
        xtr::launch_level_field(op->Go[e].p, e, (const double2 *)jq, g, ptr, ptr + g.M(), op->tw64_x.p, op->tw64_y.p,
                                (double2 *)field, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

int xtr_op_level_field_device(xtr_op *op, int electric, const void *jq, void *field)
{
    return guarded([&] {
        require(op && jq && field, "null argument");
        const int e = electric ? 1 : 0;
        require(op->go_ready[e], "build the level tensor first (xtr_op_build_green_atoo)");
        const xtr::Geom &g = op->g;
        if (op->lvl_spec.count < (size_t)g.M()) op->lvl_spec.alloc((size_t)g.M());
        if (op->lvl_small.count < (size_t)(3 * g.L())) op->lvl_small.alloc((size_t)(3 * g.L()));
        XTR_CHECK(hipDeviceSynchronize());   // jQ may have been written on another stream
        xtr::launch_level_field(op->Go[e].p, e, (const double2 *)jq, g, op->lvl_spec.p, op->lvl_small.p, op->tw64_x.p,
                                op->tw64_y.p, (double2 *)field, op->stream);
        XTR_CHECK(hipStreamSynchronize(op->stream));
    });
}

}  // extern "C"

// ---- learned right preconditioner ----------------------------------------------------------------

int xtr_net_create(int nx, int ny, int nz, uint64_t seed, xtr_net **out)
{
    return guarded([&] {
        require(out && nx >= 1 && ny >= 1 && nz >= 1, "invalid network shape");
        *out = nullptr;
        auto net = std::make_unique<xtr_net>();
        net->g = xtr::NetGeom{nx, ny, nz, (int64_t)nx * ny * nz, 0};
        const int64_t c = net->g.cells;
        XTR_CHECK(hipStreamCreateWithFlags(&net->stream, hipStreamDefault));
        net->x0.alloc(c * 8);
        net->h.alloc(c * 16);
        net->dyb.alloc(c * 8);
        net->dpb.alloc(c * 16);
        net->pre.alloc(c * 16);
        net->y.alloc(c * 16);
        net->tgt.alloc(c * 8);
        net->dy.alloc(c * 8);
        net->dh.alloc(c * 16);
        net->dp.alloc(c * 16);
        net->part.alloc((size_t)xtr::net_wgrad_chunks(net->g) * xtr::kNetW2);
        net->W1.alloc(xtr::kNetKpad8 * 16);
        net->W2.alloc(27 * 16 * 16);
        net->WT2.alloc(xtr::kNetKpad8 * 16);
        for (auto *d : {&net->W1, &net->WT2})
            XTR_CHECK(hipMemset(d->p, 0, xtr::kNetKpad8 * 16 * sizeof(unsigned short)));   // rows 216..223 stay 0
        for (auto *d : {&net->w1, &net->m1, &net->s1, &net->g1}) {
            d->alloc(xtr::kNetW1);
            XTR_CHECK(hipMemset(d->p, 0, xtr::kNetW1 * sizeof(float)));
        }
        for (auto *d : {&net->w2, &net->m2, &net->s2, &net->g2}) {
            d->alloc(xtr::kNetW2);
            XTR_CHECK(hipMemset(d->p, 0, xtr::kNetW2 * sizeof(float)));
        }
        net->partial_r.alloc(xtr::kRedBlocks);
        net->nn.alloc(1);
        net->lpart.alloc(2 * (size_t)xtr::net_loss_blocks(net->g));
        // layer 1: He-scaled normal values (6 live input channels x 27 taps) from SplitMix64 +
        // Box-Muller; layer 2: zero, so the untrained network is the identity preconditioner
        std::vector<float> w1(xtr::kNetW1, 0.0f);
        uint64_t st = seed;
        auto next = [&st]() {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
        };
        const double sd = std::sqrt(2.0 / (27.0 * 6.0));
        for (int e = 0; e < xtr::kNetW1; ++e) {
            const double u1 = next(), u2 = next();
            const int ci = (e / 16) % 8;
            w1[e] = ci < 6 ? (float)(sd * std::sqrt(-2.0 * std::log(u1 > 0 ? u1 : 1e-300)) * std::cos(2 * M_PI * u2)) : 0.0f;
        }
        XTR_CHECK(hipMemcpy(net->w1.p, w1.data(), w1.size() * sizeof(float), hipMemcpyHostToDevice));
        xtr::net_weights(net->w1.p, net->w2.p, net->bufs(), net->stream);
        XTR_CHECK(hipStreamSynchronize(net->stream));
        *out = net.release();
    });
}

void xtr_net_destroy(xtr_net *net) { delete net; }

int xtr_net_set_features(xtr_net *net, const float *feat)
{
    return guarded([&] {
        require(net, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        if (!feat) {
            net->has_feat = false;
            net->phi = 0.0f;
            return;
        }
        if (!net->feat.p) net->feat.alloc((size_t)net->g.cells * 2);
        XTR_CHECK(hipMemcpy(net->feat.p, feat, (size_t)net->g.cells * 2 * sizeof(float), hipMemcpyHostToDevice));
        net->has_feat = true;
        net->phi = feat[1];   // the frequency feature, the same in every cell, modulates the spectral branch
    });
}

int xtr_net_set_linear(xtr_net *net, int linear)
{
    return guarded([&] {
        require(net, "null argument");
        require(linear == 0 || linear == 1, "linear must be 0 or 1");
        std::lock_guard<std::mutex> lk(net->mu);
        net->g.linear = linear;
    });
}

int xtr_net_set_weights(xtr_net *net, const float *w1, const float *w2)
{
    return guarded([&] {
        require(net && w1 && w2, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        XTR_CHECK(hipMemcpy(net->w1.p, w1, xtr::kNetW1 * sizeof(float), hipMemcpyHostToDevice));
        XTR_CHECK(hipMemcpy(net->w2.p, w2, xtr::kNetW2 * sizeof(float), hipMemcpyHostToDevice));
        xtr::net_weights(net->w1.p, net->w2.p, net->bufs(), net->stream);
        XTR_CHECK(hipStreamSynchronize(net->stream));
    });
}

int xtr_net_get_weights(xtr_net *net, float *w1, float *w2)
{
    return guarded([&] {
        require(net && w1 && w2, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        XTR_CHECK(hipStreamSynchronize(net->stream));
        XTR_CHECK(hipMemcpy(w1, net->w1.p, xtr::kNetW1 * sizeof(float), hipMemcpyDeviceToHost));
        XTR_CHECK(hipMemcpy(w2, net->w2.p, xtr::kNetW2 * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int xtr_net_set_spectral(xtr_net *net, int modes)
{
    return guarded([&] {
        require(net, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        const xtr::NetGeom &g = net->g;
        require(modes >= 0 && modes <= 12, "spectral modes must be 0..12");
        require(modes == 0 || (g.nz % 8 == 0 && modes < g.nx && modes < g.ny),
                "the spectral branch needs Nz divisible by 8 and modes below Nx and Ny");
        net->sg = xtr::SpecGeom{g.nx, g.ny, g.nz, modes};
        if (modes == 0) return;
        const xtr::SpecGeom &sg = net->sg;
        const int64_t nw = 2 * sg.weights();   // floats
        for (auto *d : {&net->sw, &net->sm, &net->sv, &net->sgrad}) {
            d->alloc(nw);
            XTR_CHECK(hipMemset(d->p, 0, nw * sizeof(float)));   // zero: the branch starts as no change
        }
        net->swb.alloc(nw);
        XTR_CHECK(hipMemset(net->swb.p, 0, nw * sizeof(unsigned short)));
        auto table = [](int p) {   // e^{-2 pi i t / p}, t < p, from FP64
            std::vector<float2> t(p);
            for (int k = 0; k < p; ++k) {
                const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)p;
                t[k] = make_float2((float)std::cos(a), (float)-std::sin(a));
            }
            return t;
        };
        upload(net->twx, table(2 * g.nx));
        upload(net->twy, table(2 * g.ny));
        const int nb = 2 * modes + 1, Q = sg.q();
        net->sT.alloc((size_t)g.nx * nb * Q);
        net->sU.alloc((size_t)sg.nm() * Q);
        net->sY.alloc((size_t)sg.nm() * Q);
        net->sgY.alloc((size_t)sg.nm() * Q);
    });
}

int xtr_net_spectral_size(xtr_net *net, int *modes, int64_t *floats)
{
    return guarded([&] {
        require(net, "null argument");
        if (modes) *modes = net->sg.K;
        if (floats) *floats = net->sg.K > 0 ? 2 * net->sg.weights() : 0;
    });
}

int xtr_net_set_spectral_weights(xtr_net *net, const float *w)
{
    return guarded([&] {
        require(net && w, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        require(net->sg.K > 0, "the network has no spectral branch (xtr_net_set_spectral)");
        XTR_CHECK(hipMemcpy(net->sw.p, w, 2 * net->sg.weights() * sizeof(float), hipMemcpyHostToDevice));
        xtr::spec_weights(net->sg, net->sw.p, net->swb.p, net->stream);
        XTR_CHECK(hipStreamSynchronize(net->stream));
    });
}

int xtr_net_get_spectral_weights(xtr_net *net, float *w)
{
    return guarded([&] {
        require(net && w, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        require(net->sg.K > 0, "the network has no spectral branch (xtr_net_set_spectral)");
        XTR_CHECK(hipStreamSynchronize(net->stream));
        XTR_CHECK(hipMemcpy(w, net->sw.p, 2 * net->sg.weights() * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int xtr_net_spectral_gradient(xtr_net *net, float *g)
{
    return guarded([&] {
        require(net && g, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        require(net->sg.K > 0, "the network has no spectral branch (xtr_net_set_spectral)");
        XTR_CHECK(hipStreamSynchronize(net->stream));
        XTR_CHECK(hipMemcpy(g, net->sgrad.p, 2 * net->sg.weights() * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int xtr_net_apply_device(xtr_net *net, const void *v, void *z)
{
    return guarded([&] {
        require(net && v && z, "null argument");
        std::lock_guard<std::mutex> lk(net->mu);
        XTR_CHECK(hipDeviceSynchronize());   // inputs written on other streams
        net_precondition(net, (const double2 *)v, (double2 *)z, net->stream);
        XTR_CHECK(hipStreamSynchronize(net->stream));
    });
}

int xtr_net_gradients_device(xtr_net *net, int nsamples, const void *const *v, const void *const *z, float *g1,
                             float *g2, double *loss)
{
    return guarded([&] {
        require(net && v && z && g1 && g2 && nsamples >= 1, "invalid arguments");
        std::lock_guard<std::mutex> lk(net->mu);
        XTR_CHECK(hipDeviceSynchronize());
        const double l = net_gradients(net, nsamples, (const double2 *const *)v, (const double2 *const *)z);
        XTR_CHECK(hipMemcpy(g1, net->g1.p, xtr::kNetW1 * sizeof(float), hipMemcpyDeviceToHost));
        XTR_CHECK(hipMemcpy(g2, net->g2.p, xtr::kNetW2 * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) *loss = l;
    });
}

int xtr_net_train_device(xtr_net *net, int nsamples, const void *const *v, const void *const *z, float lr,
                         double *loss)
{
    return guarded([&] {
        require(net && v && z && nsamples >= 1 && lr > 0, "invalid arguments");
        std::lock_guard<std::mutex> lk(net->mu);
        XTR_CHECK(hipDeviceSynchronize());
        const double l = net_gradients(net, nsamples, (const double2 *const *)v, (const double2 *const *)z);
        ++net->step;
        xtr::net_adam(net->w1.p, net->m1.p, net->s1.p, net->g1.p, xtr::kNetW1, 1.0f, lr, net->step, net->stream);
        xtr::net_adam(net->w2.p, net->m2.p, net->s2.p, net->g2.p, xtr::kNetW2, 1.0f, lr, net->step, net->stream);
        xtr::net_weights(net->w1.p, net->w2.p, net->bufs(), net->stream);
        if (net->sg.K > 0) {
            xtr::net_adam(net->sw.p, net->sm.p, net->sv.p, net->sgrad.p, (int)(2 * net->sg.weights()), 1.0f, lr,
                          net->step, net->stream);
            xtr::spec_weights(net->sg, net->sw.p, net->swb.p, net->stream);
        }
        XTR_CHECK(hipStreamSynchronize(net->stream));
        if (loss) *loss = l;
    });
}

int xtr_fgmres_solve_learned_device(xtr_op *op, const xtr_gmres_params *params, xtr_net *net, const void *b, void *x,
                                    xtr_result_info *info)
{
    return guarded([&] {
        require(op && params && net && b && x && info, "null argument");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before solving");
        require(net->g.nx == op->g.nx && net->g.ny == op->g.ny && net->g.nz == op->g.nz,
                "the network was made for another grid");
        std::lock_guard<std::mutex> lk(net->mu);
        XTR_CHECK(hipDeviceSynchronize());
        Slot &sl = op->slot(0);
        SolveCtx &ctx = sl.ctx;
        ctx.reset(op->stream);
        xtr_op *o = op;
        ctx.apply = [o](const void *xx, void *yy) {
            xtr::RhsPtrs v{};
            v.x[0] = (const double2 *)xx;
            v.y[0] = (double2 *)yy;
            o->apply_raw(v, 1, o->stream);
        };
        std::memset(info, 0, sizeof(*info));
        run_fgmres(op->g.n(), *params, sl, (const double2 *)b, (double2 *)x, *info, net);
    });
}

int xtr_fgmres_solve_harvest_device(xtr_op *op, const xtr_gmres_params *params, const void *b, void *x,
                                    xtr_result_info *info, int capacity, void *const *v_out, void *const *z_out,
                                    int *count)
{
    return guarded([&] {
        require(op && params && b && x && info && count && capacity >= 0, "invalid arguments");
        require(capacity == 0 || (v_out && z_out), "null harvest buffers");
        require(op->type == XTR_OP_A, "prepare the operator as XTR_OP_A before solving");
        XTR_CHECK(hipDeviceSynchronize());
        Slot &sl = op->slot(0);
        SolveCtx &ctx = sl.ctx;
        ctx.reset(op->stream);
        ctx.harvest_v = (double2 *const *)v_out;
        ctx.harvest_z = (double2 *const *)z_out;
        ctx.harvest_cap = capacity;
        xtr_op *o = op;
        ctx.apply = [o](const void *xx, void *yy) {
            xtr::RhsPtrs v{};
            v.x[0] = (const double2 *)xx;
            v.y[0] = (double2 *)yy;
            o->apply_raw(v, 1, o->stream);
        };
        std::memset(info, 0, sizeof(*info));
        run_fgmres(op->g.n(), *params, sl, (const double2 *)b, (double2 *)x, *info);
        XTR_CHECK(hipStreamSynchronize(op->stream));
        *count = ctx.harvest_count;
    });
}

// C ABI of libsoftdft_hip: context creation, objective, Anderson solve,
// observables.  Host orchestration only; all grid work runs in the kernels of
// sdft_kernels.hip on one HIP stream per context.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "sdft_launch.h"

using namespace sdft;

struct sdft_ctx {
    Ctx c;
    Pipeline<float> p32;
    Pipeline<double> p64;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;       // the FP64 pipeline's stream while both precisions run
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int32_t concurrent = 1;              // run the two pipelines side by side
    int device = 0;
    std::vector<DevPoint> pts;
    std::vector<void*> allocs;
    std::string err;
    bool solved = false;
    bool have_eval = false;
    double last_solve_ms = 0.0;
    double last_iters_total = 0.0;
    int32_t last_iters32 = 0, last_iters64 = 0;  // batch iterations that ran each pipeline
    std::vector<int32_t> last_switch;             // per point: iterations done at the FP32 -> FP64 switch
    int32_t obj_band32 = 1, obj_band64 = 0;       // bond operator of sdft_objective per precision
    int nsets = 0;
    int S_max = 1;
    std::vector<int> nseg;
    std::vector<int> S;
    std::vector<int> seg_states_h;
    std::vector<int> state_off;
    std::vector<int> multistate;
    int hard_walls = 0, charged_walls = 0;
    std::vector<double> valences;
    double lB = 0.0;
    int fwd_terms = 0, back_terms = 0;   // convolution terms per point (algorithmic flop count)
    double* d_obs = nullptr;             // sdft_observables' work buffer (obs_layout), allocated on first use
    double* d_profile_fx = nullptr;      // sdft_kernel_profile's scratch F(x), allocated on first use
    double* d_scratch = nullptr;         // this device's copy of the static scratch array (sdft_create)
    // contact-value theorem table (built by sdft_create): per point the entries off[p]..off[p+1) of
    // (wall side, bead field offset j*NP + index or -1 for a constant term, factor); one gather kernel
    // sums them into d_cvt[P][2]
    int32_t* d_cvt_off = nullptr;
    int32_t* d_cvt_idx = nullptr;
    int32_t* d_cvt_side = nullptr;
    double* d_cvt_fac = nullptr;
    bool cvt_any = false;
    // convolution plan inputs, kept to rebuild the plans with other tile shapes (sdft_autotune)
    std::vector<ConvJob> jobs_h;
    std::vector<std::vector<int>> canon;
    std::vector<int32_t> set_L_h;
    std::vector<int64_t> set_kbase_h;
    std::vector<int> set_shape[3];       // per stage (fwd, pre, back) and kernel set: tile shape, -1 heuristic
    std::vector<sdft_tune_entry_t> tuning;   // decisions of the last sdft_autotune
    int32_t graph_capture = 1;           // sdft_solve runs its poll windows as captured hipGraphs
    int64_t graphs_built = 0;            // graphs instantiated by sdft_solve over the context's life
};

// Executable graphs of one sdft_solve, one per pair of active pipelines (FP32 | FP64 << 1);
// released when the solve returns, after the stream has drained.
struct WindowGraphs {
    hipStream_t stream;
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    explicit WindowGraphs(hipStream_t s) : stream(s) {}
    WindowGraphs(const WindowGraphs&) = delete;
    WindowGraphs& operator=(const WindowGraphs&) = delete;
    ~WindowGraphs() {
        bool any = false;
        for (hipGraphExec_t e : exec) any = any || e;
        if (!any) return;
        (void)hipStreamSynchronize(stream);
        for (hipGraphExec_t& e : exec)
            if (e) (void)hipGraphExecDestroy(e);
    }
};

// Two timing events destroyed on every exit path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    hipError_t create() {
        hipError_t e = hipEventCreate(&e0);
        return e != hipSuccess ? e : hipEventCreate(&e1);
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Saves the per-point evaluation flags and writes them back on every exit path (profiling and
// tuning force every point to one precision while they run).
struct EvalflagGuard {
    hipStream_t stream;
    int32_t* d_flags;
    std::vector<int32_t> saved;
    bool armed = false;
    EvalflagGuard(hipStream_t s, int32_t* d, int P) : stream(s), d_flags(d), saved(P) {}
    EvalflagGuard(const EvalflagGuard&) = delete;
    EvalflagGuard& operator=(const EvalflagGuard&) = delete;
    hipError_t save() {
        hipError_t e = hipMemcpyAsync(saved.data(), d_flags, saved.size() * sizeof(int32_t), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        armed = e == hipSuccess;
        return e;
    }
    hipError_t restore() {
        if (!armed) return hipSuccess;
        armed = false;
        hipError_t e = hipMemcpyAsync(d_flags, saved.data(), saved.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream);
        return e != hipSuccess ? e : hipStreamSynchronize(stream);
    }
    ~EvalflagGuard() { (void)restore(); }
};

// Device scratch of sdft_observables_noalloc: a file-scope array of SDFT_SCRATCH_DOUBLES doubles
// (8 MiB) that every GPU the library runs on holds until the process ends. Its name is a host-side
// stand-in: host code uses only the per-device address from hipGetSymbolAddress, looked up once per
// device by sdft_create with that device current. Calls on one GPU share the array, so calls that
// overlap on one GPU race; each call rewrites every entry it reads.
#define SDFT_SCRATCH_DOUBLES (1 << 20)
static __device__ __align__(256) double scratch[SDFT_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch_ptr;   // device ordinal -> address of its copy of `scratch`

static thread_local std::string g_last_error;

static void set_error(sdft_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    (void)code;
    g_last_error = buf;
    if (ctx) ctx->err = buf;
}

template <typename T>
static int dalloc(sdft_ctx* ctx, T** p, int64_t count) {
    if (count <= 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)count * sizeof(T));
    if (e != hipSuccess) {
        set_error(ctx, SDFT_E_HIP, "hipMalloc(%lld bytes) failed: %s", (long long)(count * sizeof(T)), hipGetErrorString(e));
        return SDFT_E_HIP;
    }
    e = hipMemsetAsync(q, 0, (size_t)count * sizeof(T), ctx->stream);
    if (e != hipSuccess) {
        set_error(ctx, SDFT_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
        return SDFT_E_HIP;
    }
    ctx->allocs.push_back(q);
    *p = (T*)q;
    return SDFT_OK;
}

// releases a buffer from dalloc before the context ends
static void dfree(sdft_ctx* ctx, void* p) {
    if (!p) return;
    for (size_t i = 0; i < ctx->allocs.size(); ++i)
        if (ctx->allocs[i] == p) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(p);
            ctx->allocs.erase(ctx->allocs.begin() + (std::ptrdiff_t)i);
            return;
        }
}

#define TRY(expr)                      \
    do {                               \
        int _r = (expr);               \
        if (_r != SDFT_OK) return _r;  \
    } while (0)

template <typename T>
static int upload(sdft_ctx* ctx, T** dst, const T* src, int64_t count) {
    TRY(dalloc(ctx, dst, count));
    if (count > 0) HIP_CHECK_RET(hipMemcpyAsync(*dst, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return SDFT_OK;
}

extern "C" int32_t sdft_abi_version(void) { return SDFT_ABI_VERSION; }

extern "C" const char* sdft_last_error(const sdft_ctx_t* ctx) {
    if (ctx && !ctx->err.empty()) return ctx->err.c_str();
    return g_last_error.c_str();
}

extern "C" void sdft_destroy(sdft_ctx_t* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    for (void* p : ctx->allocs) (void)hipFree(p);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// Convolution job table (functionals A14-A20): forward weighted densities; shared back-projection
// partials together with the lng and Psi projections; the bead back projections into F(x).
//  * n_V = w_V * rho_j enters only summed over beads (no |.|), so it is one job per diameter class on
//    the class sum of rho_j trapez_j.
//  * mu_ex_j = w2_j * ch0_j + w3_j * ch1_j - wV_j * chv_j. With FMT, ch1 (dPhi4) is the same field for
//    every bead and chv (dPhi5/(2 pi s) + dPhi8) the same within a diameter class, whose beads share
//    w3 and wV: w3 * ch1 - wV * chv is computed once per class (outkind 2) and added to each bead's
//    w2 * ch0_j (linearity; the reference sums the three per bead).
//  * the two ends of a bond with equal diameters share kernel and input: one job writes both lng fields.
//  * square well (square_well.jl:98-123): mu_i -= Re(sum_j rho_hat_j w_ij eps_ij) in k-space. With the
//    half-sample mirror the real part is the average of the convolution at n and n-1, so it is one
//    convolution with the kernel averaged over d and d-1, on rho_j trapez_j: a shared partial per bead
//    type, added to that bead's back projection. The plain convolution (the free energy's, :88-95) is a
//    forward weighted field read only by the grand potential.
static int build_jobs(const sdft_desc_t* d, const DevSystem& sys, std::vector<ConvJob>& fwd,
                      std::vector<ConvJob>& pre, std::vector<ConvJob>& back, int& Fws) {
    const int NB = d->NB, MB = d->MB;
    const int F = d->functionals;
    const bool fmt = F & (SDFT_F_MFMT | SDFT_F_TPT1MFMT);
    const bool msa = F & (SDFT_F_MSA | SDFT_F_TPT1MSA);
    const bool tpt1 = F & (SDFT_F_TPT1MFMT | SDFT_F_TPT1MSA);
    auto job = [](int outkind, int outfield, int absval, double scale) {
        ConvJob j;
        memset(&j, 0, sizeof(j));
        j.outkind = outkind; j.outfield = outfield; j.absval = absval; j.scale = scale;
        j.addfield = -1; j.outfield2 = -1; j.addfield2 = -1;
        return j;
    };
    auto add = [](ConvJob& j, int kern, int odd, int infam, int infield) {
        ConvTerm& t = j.t[j.nterms++];
        t.kern = kern; t.odd = odd; t.infam = infam; t.infield = infield;
    };
    for (int j = 0; j < NB; ++j) {
        if (fmt || msa) { ConvJob a = job(0, j, 1, 1.0); add(a, 3 * j + 0, 0, 0, j); fwd.push_back(a); }
        if (fmt) { ConvJob b = job(0, NB + j, 1, 1.0); add(b, 3 * j + 1, 0, 0, j); fwd.push_back(b); }
    }
    // vector weighted density (no |.|, linear): one per diameter class on the class density sum
    if (fmt) {
        for (int c = 0; c < sys.ncls; ++c) {
            int j0 = 0;
            while (sys.cls[j0] != c) ++j0;
            ConvJob v = job(0, 2 * NB + c, 0, 1.0);
            add(v, 3 * j0 + 2, 1, 0, NB + 2 * MB + c);
            fwd.push_back(v);
        }
    }
    if (tpt1) {
        for (int u = 0; u < MB; ++u)
            for (int e = 0; e < 2; ++e) {
                const int bt = d->bond_types[2 * u + e];
                const double s = d->diameters[bt];
                ConvJob a = job(0, 3 * NB + 2 * u + e, 1, 1.0 / (M_PI * s * s));
                add(a, 3 * bt + 0, 0, 0, NB + 2 * u + e);
                fwd.push_back(a);
            }
    }
    Fws = 0;
    if (fmt) {
        Fws = sys.ncls;
        for (int c = 0; c < sys.ncls; ++c) {
            int j0 = 0;
            while (sys.cls[j0] != c) ++j0;
            ConvJob w = job(2, c, 0, 1.0);
            add(w, 3 * j0 + 1, 0, 1, NB + j0);
            add(w, 3 * j0 + 2, 2, 1, 2 * NB + j0);
            pre.push_back(w);
        }
    }
    // square-well partners of each participating bead type: for every participating j the first pair
    // (min(i, j), max(i, j)) in list order (findfirst, square_well.jl:78-79)
    std::vector<int> sw_slot(NB, -1);
    const int sw_ws = Fws;
    for (int s = 0; s < sys.n_sw_beads; ++s) {
        const int i = sys.sw_bead[s];
        ConvJob f = job(0, sys.nw_sw_base + s, 0, 1.0);
        ConvJob m = job(2, sw_ws + s, 0, -1.0);
        for (int t = 0; t < sys.n_sw_beads; ++t) {
            const int j = sys.sw_bead[t];
            const int lo = std::min(i, j), hi = std::max(i, j);
            int u = 0;
            while (u < sys.n_sw && !(sys.sw_i[u] == lo && sys.sw_j[u] == hi)) ++u;
            if (u == sys.n_sw) continue;
            if (f.nterms == SDFT_SW_TERMS) return SDFT_E_UNSUPPORTED;
            add(f, sys.sw_kbase + 2 * u + 1, 0, 0, j);
            add(m, sys.sw_kbase + 2 * u, 0, 0, j);
        }
        fwd.push_back(f);
        pre.push_back(m);
        sw_slot[i] = sw_ws + s;
    }
    Fws += sys.n_sw_beads;
    for (int j = 0; j < NB; ++j) {
        ConvJob a = job(1, j, 0, 1.0);
        if (fmt || msa) add(a, 3 * j + 0, 0, 1, j);
        if (fmt) a.addfield = sys.cls[j];
        a.addfield2 = sw_slot[j];
        back.push_back(a);
    }
    // lng and Psi need no shared partial: they run in the same launch as the partials
    for (int u = 0; u < MB; ++u) {
        const bool same = d->diameters[d->bond_types[2 * u]] == d->diameters[d->bond_types[2 * u + 1]];
        for (int e = 0; e < 2; ++e) {
            if (e == 1 && same) continue;  // written by the e = 0 job
            const int bt = d->bond_types[2 * u + e];
            ConvJob a = job(1, NB + 2 * u + e, 0, 1.0);
            if (tpt1) add(a, 3 * bt + 0, 0, 1, 3 * NB + 2 * u + e);
            if (same) a.outfield2 = NB + 2 * u + 1;
            pre.push_back(a);
        }
    }
    ConvJob psi = job(1, NB + 2 * MB, 0, 1.0);
    add(psi, 3 * NB + MB, 0, 1, 3 * NB + 2 * MB);
    pre.push_back(psi);
    return SDFT_OK;
}

// ---------------------------------------------------------------------------
// MFMA operator-GEMM plans: group columns that share a kernel set and the same
// (canonical) operator terms, then cut each group into row x column tiles.
// ---------------------------------------------------------------------------
template <typename V>
static int upload_vec(sdft_ctx* ctx, void** dst, const std::vector<V>& v) {
    V* d = nullptr;
    TRY(upload(ctx, &d, v.data(), (int64_t)v.size()));
    *dst = d;
    return SDFT_OK;
}

// Workgroups are dealt to the 8 XCDs round-robin by ID, each XCD with its own L2. Reorder the
// tiles (built column block by column block, rows fastest) so that all row tiles of one column
// block land on the same XCD and share its L2 copy of the block's inputs: column block k goes to
// XCD k % 8 and the lists are interleaved, padded with empty tiles (ncols 0 exits at once).
static std::vector<OpTile> xcd_order(const std::vector<OpTile>& tiles) {
    constexpr int NX = 8;
    std::vector<std::vector<OpTile>> by(NX);
    int blk = -1;
    OpTile prev{-1, -1, -1, 0};
    for (const OpTile& t : tiles) {
        if (t.group != prev.group || t.col0 != prev.col0) ++blk;
        by[blk % NX].push_back(t);
        prev = t;
    }
    size_t mx = 0;
    for (auto& v : by) mx = std::max(mx, v.size());
    std::vector<OpTile> out;
    out.reserve(mx * NX);
    for (size_t k = 0; k < mx; ++k)
        for (int x = 0; x < NX; ++x) out.push_back(k < by[x].size() ? by[x][k] : OpTile{0, 0, 0, 0});
    return out;  // a multiple of 8 long: split launches (ID = tile + ntiles * part) keep parts on one XCD
}

// distinct input fields (infam, infield) of a job range, in first-use order
static std::vector<FftSlot> stage_slots(const std::vector<ConvJob>& jobs, int job_base, int njobs) {
    std::vector<FftSlot> v;
    for (int j = 0; j < njobs; ++j) {
        const ConvJob& J = jobs[job_base + j];
        for (int t = 0; t < J.nterms; ++t) {
            bool seen = false;
            for (const FftSlot& q : v) seen = seen || (q.infam == J.t[t].infam && q.infield == J.t[t].infield);
            if (!seen) v.push_back(FftSlot{J.t[t].infam, J.t[t].infield});
        }
    }
    return v;
}

// the FFT path takes a kernel set of a stage when its mirror length is a power of two in range
// and the stage's input spectra fit the middle pass's LDS
static bool fft_stage_ok(int L, int nslot) {
    if (!fft_length_ok(L) || nslot < 1) return false;
    int L1, L2;
    fft_split(L, &L1, &L2);
    return fft_mid_lds_bytes(nslot, L1, sizeof(double)) <= 160 * 1024;
}

// kernel spectra and twiddles of kernel set s (once per context)
static int fft_prepare_set(sdft_ctx* ctx, int s, int L) {
    Ctx& c = ctx->c;
    if ((int)c.fft_res.size() < ctx->nsets) c.fft_res.resize(ctx->nsets);
    FftRes& r = c.fft_res[s];
    if (r.L == L) return SDFT_OK;
    const int64_t H1 = fft_spectrum_len(L);
    double* W64 = nullptr;
    float* W32 = nullptr;
    double* ch64 = nullptr;
    float* ch32 = nullptr;
    double* tmp = nullptr;
    TRY(dalloc(ctx, &W64, 2 * (int64_t)L));
    TRY(dalloc(ctx, &W32, 2 * (int64_t)L));
    TRY(dalloc(ctx, &ch64, 2 * H1 * c.K));
    TRY(dalloc(ctx, &ch32, 2 * H1 * c.K));
    TRY(dalloc(ctx, &tmp, 2 * H1 * c.K));
    launch_fft_twiddles(ctx->stream, L, W64, W32);
    launch_fft_kernel_spectra(ctx->stream, L, c.K, ctx->p64.kern + ctx->set_kbase_h[s] * c.K, W64, tmp, ch64, ch32);
    HIP_CHECK_RET(hipGetLastError());
    dfree(ctx, tmp);  // synchronises the stream first
    r.L = L;
    r.W64 = W64;
    r.W32 = W32;
    r.chat64 = ch64;
    r.chat32 = ch32;
    return SDFT_OK;
}

// spectra work buffers of both precisions for n_in input and n_out output spectra of length H1
static int fft_reserve(sdft_ctx* ctx, int64_t n_in, int64_t n_out) {
    Ctx& c = ctx->c;
    if (n_in > c.fft_sa_n) {
        for (int w = 0; w < 2; ++w) {
            dfree(ctx, c.fft_sa[w]);
            c.fft_sa[w] = nullptr;
        }
        float* a32 = nullptr;
        double* a64 = nullptr;
        TRY(dalloc(ctx, &a32, 2 * n_in));
        TRY(dalloc(ctx, &a64, 2 * n_in));
        c.fft_sa[0] = a32;
        c.fft_sa[1] = a64;
        c.fft_sa_n = n_in;
    }
    if (n_out > c.fft_sb_n) {
        for (int w = 0; w < 2; ++w) {
            dfree(ctx, c.fft_sb[w]);
            c.fft_sb[w] = nullptr;
        }
        float* b32 = nullptr;
        double* b64 = nullptr;
        TRY(dalloc(ctx, &b32, 2 * n_out));
        TRY(dalloc(ctx, &b64, 2 * n_out));
        c.fft_sb[0] = b32;
        c.fft_sb[1] = b64;
        c.fft_sb_n = n_out;
    }
    return SDFT_OK;
}

// Convolution plan of one stage: columns (point, job) grouped by (kernel set, operator terms), each
// group tiled in the shape of its kernel set (set_shape[kset]; < 0: the fixed rule, the batched FFT
// where the mirror length allows it, else 64 x 32 tiles when the set's groups average >= 48
// columns, else 128 x 16). Points with point_mask[p] == 0 are
// left out (the autotuner's per-grid-length plans).
static int build_conv_plan(sdft_ctx* ctx, const std::vector<ConvJob>& jobs, int job_base, int njobs,
                           const std::vector<std::vector<int>>& canon, const std::vector<int>& set_shape,
                           const std::vector<char>* point_mask, const std::vector<int32_t>& set_L,
                           const std::vector<int64_t>& set_kbase, Ctx::ConvStage& stage) {
    Ctx& c = ctx->c;
    struct Key {
        int kset, nterms, kern[3], sign[3];
        bool operator<(const Key& o) const {
            if (kset != o.kset) return kset < o.kset;
            if (nterms != o.nterms) return nterms < o.nterms;
            for (int i = 0; i < 3; ++i) {
                if (kern[i] != o.kern[i]) return kern[i] < o.kern[i];
                if (sign[i] != o.sign[i]) return sign[i] < o.sign[i];
            }
            return false;
        }
    };
    std::map<Key, std::vector<OpCol>> groups;
    for (int p = 0; p < c.P; ++p) {
        if (point_mask && !(*point_mask)[p]) continue;
        const int s = ctx->pts[p].kset;
        for (int j = 0; j < njobs; ++j) {
            const ConvJob& J = jobs[job_base + j];
            Key k;
            memset(&k, 0, sizeof(k));
            k.kset = s;
            k.nterms = J.nterms;
            for (int t = 0; t < J.nterms; ++t) {
                k.kern[t] = canon[s][J.t[t].kern];
                k.sign[t] = J.t[t].odd == 2 ? -1 : 1;
            }
            groups[k].push_back(OpCol{p, job_base + j});
        }
    }
    // shape of each kernel set in this stage
    std::map<int, std::pair<size_t, size_t>> set_cols;  // kset -> (columns, groups)
    for (auto& kv : groups) {
        auto& e = set_cols[kv.first.kset];
        e.first += kv.second.size();
        e.second += 1;
    }
    const std::vector<FftSlot> slots = stage_slots(jobs, job_base, njobs);
    auto shape_of = [&](int kset) {
        const bool fft_ok = fft_stage_ok(set_L[kset], (int)slots.size());
        if (kset < (int)set_shape.size() && set_shape[kset] >= 0 && (set_shape[kset] != CONV_FFT || fft_ok))
            return set_shape[kset];
        // fixed rule: the FFT wherever it applies (measured faster than the best GEMM tile from NP 2048
        // up, FP32 + FP64 forward / back stage: 1.06x / 1.5x at NP 2048, 2.4x / 3.1x at 4096,
        // 4.9x / 6.4x at 8192, ~9x at 16384);
        // else 64 x 32 tiles where the set's groups average >= 48 columns, 128 x 16 otherwise
        if (fft_ok) return (int)CONV_FFT;
        const auto& e = set_cols[kset];
        return e.first >= 48 * e.second ? (int)CONV_MID : (int)CONV_TALL;
    };
    for (int sh = 0; sh < CONV_NSHAPES; ++sh) {
        OpPlan& plan = stage.sub[sh];
        plan = OpPlan();
        plan.shape = sh;
        plan.wide = sh != CONV_TALL;
        if (sh == CONV_FFT) {
            // per kernel set: its points, the stage's input slots and one output column per job
            std::map<int, std::pair<std::vector<int32_t>, std::vector<FftJob>>> sets;
            for (auto& kv : groups) {
                const Key& k = kv.first;
                if (shape_of(k.kset) != CONV_FFT) continue;
                auto& e = sets[k.kset];
                for (const OpCol& col : kv.second) {
                    if (std::find(e.first.begin(), e.first.end(), col.p) == e.first.end()) e.first.push_back(col.p);
                    bool seen = false;
                    for (const FftJob& fj : e.second) seen = seen || fj.job == col.job;
                    if (seen) continue;
                    FftJob fj;
                    memset(&fj, 0, sizeof(fj));
                    fj.job = col.job;
                    fj.nterms = k.nterms;
                    const ConvJob& J = jobs[col.job];
                    for (int t = 0; t < k.nterms; ++t) {
                        fj.kern[t] = k.kern[t];
                        fj.sign[t] = k.sign[t];
                        for (size_t q = 0; q < slots.size(); ++q)
                            if (slots[q].infam == J.t[t].infam && slots[q].infield == J.t[t].infield) fj.slot[t] = (int)q;
                    }
                    e.second.push_back(fj);
                }
            }
            for (auto& kv : sets) {
                FftSetPlan f;
                f.kset = kv.first;
                f.L = set_L[kv.first];
                fft_split(f.L, &f.L1, &f.L2);
                f.lg1 = 0;
                while ((1 << f.lg1) < f.L1) ++f.lg1;
                f.lg2 = 0;
                while ((1 << f.lg2) < f.L2) ++f.lg2;
                f.H = f.L2 / 2 + 1;
                f.R = 256 / f.L1;
                f.npts = (int)kv.second.first.size();
                f.nslot = (int)slots.size();
                f.njob = (int)kv.second.second.size();
                void* dp = nullptr;
                TRY(upload_vec(ctx, &dp, kv.second.first));
                f.d_pts = (int32_t*)dp;
                TRY(upload_vec(ctx, &dp, slots));
                f.d_slots = (FftSlot*)dp;
                TRY(upload_vec(ctx, &dp, kv.second.second));
                f.d_jobs = (FftJob*)dp;
                TRY(fft_prepare_set(ctx, f.kset, f.L));
                const int64_t H1 = fft_spectrum_len(f.L);
                TRY(fft_reserve(ctx, (int64_t)f.npts * f.nslot * H1, (int64_t)f.npts * f.njob * H1));
                plan.fft.push_back(f);
            }
            continue;
        }
        const int TM = conv_tile_rows(sh), TN = conv_tile_cols(sh);
        std::vector<OpGroup> gv;
        std::vector<OpCol> cols;
        std::vector<OpTile> tiles;
        for (auto& kv : groups) {
            const Key& k = kv.first;
            if (shape_of(k.kset) != sh) continue;
            OpGroup g;
            memset(&g, 0, sizeof(g));
            g.kset = k.kset;
            g.L = set_L[k.kset];
            g.NP = g.L / 2;
            g.nterms = k.nterms;
            for (int t = 0; t < 3; ++t) { g.kern[t] = k.kern[t]; g.sign[t] = k.sign[t]; }
            g.col_off = (int32_t)cols.size();
            g.ncols = (int32_t)kv.second.size();
            g.kbase = set_kbase[k.kset];
            g.band = -1;
            const int gi = (int)gv.size();
            gv.push_back(g);
            cols.insert(cols.end(), kv.second.begin(), kv.second.end());
            for (int c0 = 0; c0 < g.ncols; c0 += TN)
                for (int r = 0; r < g.NP; r += TM) tiles.push_back(OpTile{gi, r, c0, std::min(TN, g.ncols - c0)});
        }
        tiles = xcd_order(tiles);
        plan.ntiles = (int)tiles.size();
        if (plan.ntiles == 0) continue;
        TRY(upload_vec(ctx, &plan.d_tiles, tiles));
        TRY(upload_vec(ctx, &plan.d_groups, gv));
        TRY(upload_vec(ctx, &plan.d_cols, cols));
        plan.ngroups = (int)gv.size();
        for (int k = 0; k < 2; ++k) {
            OpCol* act = nullptr;
            TRY(dalloc(ctx, &act, (int64_t)cols.size()));
            plan.d_cols_act[k] = act;
            TRY(dalloc(ctx, &plan.d_count[k], (int64_t)gv.size()));
        }
    }
    return SDFT_OK;
}

// releases the device tables of a stage's plans (autotuner candidates, replaced plans)
static void free_stage(sdft_ctx* ctx, Ctx::ConvStage& stage) {
    for (OpPlan& pl : stage.sub) {
        std::vector<void*> ptrs = {pl.d_tiles, pl.d_groups, pl.d_cols, pl.d_cols_act[0], pl.d_cols_act[1],
                                   pl.d_count[0], pl.d_count[1]};
        for (const FftSetPlan& f : pl.fft) {
            ptrs.push_back(f.d_pts);
            ptrs.push_back(f.d_slots);
            ptrs.push_back(f.d_jobs);
        }
        for (void* q : ptrs) {
            if (!q) continue;
            auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), q);
            if (it != ctx->allocs.end()) ctx->allocs.erase(it);
            (void)hipFree(q);
        }
        pl = OpPlan();
    }
}

// job range of convolution stage k (0 forward, 1 shared partials, 2 back projections); with the
// merged back stage, stage 1 is empty and stage 2 takes the partials and the projections from the
// table's second copy
static void stage_range(const Ctx& c, int k, int* base, int* n) {
    if (k == 0) {
        *base = 0;
        *n = c.n_fwd_jobs;
    } else if (c.merge_back) {
        *base = c.n_jobs + c.n_fwd_jobs;
        *n = k == 1 ? 0 : c.n_pre_jobs + c.n_back_jobs;
    } else {
        *base = k == 1 ? c.n_fwd_jobs : c.n_fwd_jobs + c.n_pre_jobs;
        *n = k == 1 ? c.n_pre_jobs : c.n_back_jobs;
    }
}

static int rebuild_conv_plans(sdft_ctx* ctx) {
    Ctx& c = ctx->c;
    Ctx::ConvStage* st[3] = {&c.fwd_plan, &c.pre_plan, &c.back_plan};
    for (int k = 0; k < 3; ++k) {
        free_stage(ctx, *st[k]);
        int base = 0, n = 0;
        stage_range(c, k, &base, &n);
        TRY(build_conv_plan(ctx, ctx->jobs_h, base, n, ctx->canon, ctx->set_shape[k], nullptr, ctx->set_L_h,
                            ctx->set_kbase_h, *st[k]));
    }
    return SDFT_OK;
}

static int bond_of(const DevSystem& sys, int MB, int si, int sj) {
    const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
    for (int u = 0; u < MB; ++u)
        if (sys.bond_i[u] == lo && sys.bond_j[u] == hi) return u;
    return -1;
}

// Canonical bond operator of every (sweep, a, b) column of chain species u (ck[(sweep*S + a)*S + b],
// -1 where the column never occurs). False if a column's operator changes along the chain or a
// bond type is missing: such chains keep the direct kernel.
static bool column_kernels(sdft_ctx* ctx, const DevSystem& sys, int u, const std::vector<std::vector<int>>& canon,
                           std::vector<int>& ck) {
    Ctx& c = ctx->c;
    const SpeciesHost& sh = c.species[u];
    const int N = sh.nseg, S = sh.S, NB = c.NB, MB = c.MB;
    const int* st = ctx->seg_states_h.data() + sh.state_off;
    ck.assign(2 * S * S, -2);
    for (int sw = 0; sw < 2; ++sw)
        for (int s = 1; s < N; ++s) {
            const int outseg = sw == 0 ? N - 1 - s : s;
            const int inseg = sw == 0 ? outseg + 1 : outseg - 1;
            for (int a = 0; a < S; ++a)
                for (int b = 0; b < S; ++b) {
                    const int si = st[outseg * S + a], sj = st[inseg * S + b];
                    if (si < 0 || sj < 0) continue;
                    const int bond = bond_of(sys, MB, si, sj);
                    if (bond < 0) return false;
                    for (size_t ks = 0; ks < canon.size(); ++ks) {
                        const int k = canon[ks][3 * NB + bond];
                        int& slot = ck[(sw * S + a) * S + b];
                        if (slot == -2) slot = k;
                        else if (slot != k) return false;
                    }
                }
        }
    for (int& k : ck)
        if (k == -2) k = -1;
    return true;
}

// Persistent banded chain plan of species u: free (ungrafted) linear chains with fixed column
// operators whose band fits the mirror construction (2 Be + 1 <= NP) and whose working set fits LDS.
static int build_chain_plan(sdft_ctx* ctx, const DevSystem& sys, int u, const std::vector<std::vector<int>>& canon,
                            const std::vector<int32_t>& band, ChainPlanHost& ch) {
    Ctx& c = ctx->c;
    const SpeciesHost& sh = c.species[u];
    ch = ChainPlanHost();
    const int N = sh.nseg, S = sh.S;
    if (N < 2 || sh.fixed_seg >= 0 || S > SDFT_FAST_S || sh.model == SDFT_MODEL_MONOMERBEAD) return SDFT_OK;
    std::vector<int> ck;
    if (!column_kernels(ctx, sys, u, canon, ck)) return SDFT_OK;
    for (int sw = 0; sw < 2; ++sw)
        for (int k = 0; k < SDFT_FAST_S * SDFT_FAST_S; ++k) ch.ck[sw][k] = k < S * S ? ck[sw * S * S + k] : -1;
    const int* st = ctx->seg_states_h.data() + sh.state_off;
    ch.homog = true;
    for (int q = 1; q < N; ++q)
        for (int a = 0; a < S; ++a) ch.homog = ch.homog && st[q * S + a] == st[a];
    ch.S = S;
    for (int p = 0; p < c.P; ++p) {
        const int NP = ctx->pts[p].NP;
        const int Be = chain_half_width(band[ctx->pts[p].kset]);  // the chain kernel's half-width
        if (2 * Be + 1 > NP) return SDFT_OK;
        ch.npad_max = std::max(ch.npad_max, (NP + 255) / 256 * 256);
        ch.be_max = std::max(ch.be_max, Be);
        ch.xl_need = std::max(ch.xl_need, NP + 2 * Be);
    }
    ch.ok32 = chain_lds_bytes(ch, sizeof(float), false) <= 160 * 1024;
    ch.ok64 = chain_lds_bytes(ch, sizeof(double), false) <= 160 * 1024;
    // output-factor table in LDS for homogeneous chains where it fits without costing occupancy:
    // FP32 runs two workgroups per CU, so within 80 KB when the workgroup fitted there before; the
    // FP64 variants run one per CU (registers), so within 160 KB
    auto hv_fits = [&](size_t elem, bool ok) {
        if (!ok || !ch.homog) return false;
        const size_t b0 = chain_lds_bytes(ch, elem, false), b1 = chain_lds_bytes(ch, elem, true);
        const size_t lim = (elem == 4 && b0 <= 80 * 1024) ? 80 * 1024 : 160 * 1024;
        return b1 <= lim;
    };
    ch.hv32 = hv_fits(sizeof(float), ch.ok32);
    ch.hv64 = hv_fits(sizeof(double), ch.ok64);
    // homogeneous chains with at most one (a, tile) unit per wave run on k_chain_pp
    {
        const int units = S * (ch.npad_max / 256);
        const size_t full = 160 * 1024 - 256;
        // both sweeps of a point in one workgroup when more points are active than half the CUs
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 2)
            ncu = 256;
        ch.pp_min = ncu / 2 + 1;
        const bool pp_ok = ch.homog && units <= 8;
        // one bond operator for every valid column: k_chain_pp runs both input states of a unit as
        // two accumulation chains over one operator window
        ch.shared_a = S == 2;
        int k0 = -1;
        for (int sw = 0; sw < 2; ++sw)
            for (int col = 0; col < S * S; ++col) {
                const int k = ch.ck[sw][col];
                if (k < 0) continue;
                if (k0 < 0) k0 = k;
                ch.shared_a = ch.shared_a && k == k0;
            }
        ch.pp32 = pp_ok && ch.ok32 && chain_pp_lds_bytes(ch, sizeof(float)) <= full;
        ch.pp64 = pp_ok && ch.ok64 && chain_pp_lds_bytes(ch, sizeof(double)) <= full;
    }
    ch.ok = ch.ok32 || ch.ok64;
    ch.elig32 = ch.ok32;
    ch.elig64 = ch.ok64;
    if (!ch.ok) return SDFT_OK;
    // column table of every step of both sweeps (child_propagators.jl / parent_propagators.jl walk
    // the same linear chain from opposite ends): input bead and bond fields of each (a, b) pair
    const int C = S * S;
    std::vector<ChainCol> tab((size_t)2 * (N - 1) * C);
    for (int sw = 0; sw < 2; ++sw)
        for (int q = 1; q < N; ++q) {
            const int outseg = sw == 0 ? N - 1 - q : q;
            const int inseg = sw == 0 ? outseg + 1 : outseg - 1;
            for (int col = 0; col < C; ++col) {
                const int a = col / S, b = col - a * S;
                const int si = st[outseg * S + a], sj = st[inseg * S + b];
                ChainCol cc;
                cc.valid = (si >= 0 && sj >= 0 && ch.ck[sw][col] >= 0) ? 1 : 0;
                cc.sj = sj < 0 ? 0 : sj;
                int bond = 0;
                if (cc.valid) {
                    const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                    for (int u2 = 0; u2 < c.MB; ++u2)
                        if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                }
                cc.hl_in = 2 * bond + (si <= sj ? 1 : 0);
                cc.hl_out = 2 * bond + (si <= sj ? 0 : 1);
                tab[((size_t)sw * (N - 1) + (q - 1)) * C + col] = cc;
            }
        }
    TRY(upload(ctx, &ch.d_info, tab.data(), (int64_t)tab.size()));
    return SDFT_OK;
}

// Propagator plan of species u: valid only if every (sweep, a, b) column applies the same
// canonical bond operator at every step (homogeneous chains); otherwise the direct kernel runs.
static int build_prop_plan(sdft_ctx* ctx, const DevSystem& sys, int u, const std::vector<std::vector<int>>& canon,
                           const std::vector<int32_t>& set_L, const std::vector<int64_t>& set_kbase,
                           const std::vector<int32_t>& band, OpPlan& plan) {
    Ctx& c = ctx->c;
    const SpeciesHost& sh = c.species[u];
    const int S = sh.S;
    // more than two states: the direct step kernel (its general-state instantiation)
    if (sh.nseg < 2 || sh.model == SDFT_MODEL_MONOMERBEAD || S > SDFT_FAST_S) return SDFT_OK;
    std::vector<int> ck;
    if (!column_kernels(ctx, sys, u, canon, ck)) return SDFT_OK;  // not homogeneous: keep the direct kernel
    struct Key { int kset, kern; bool operator<(const Key& o) const { return kset != o.kset ? kset < o.kset : kern < o.kern; } };
    std::map<Key, std::vector<PropCol>> groups;
    for (int p = 0; p < c.P; ++p)
        for (int sw = 0; sw < 2; ++sw)
            for (int a = 0; a < S; ++a)
                for (int b = 0; b < S; ++b) {
                    const int k = ck[(sw * S + a) * S + b];
                    if (k < 0) continue;
                    groups[Key{ctx->pts[p].kset, k}].push_back(PropCol{p, sw, a, b});
                }
    size_t tot = 0;
    for (auto& g : groups) tot += g.second.size();
    plan.wide = groups.empty() ? true : (tot / groups.size() >= 48);
    // fused chaining needs every (sweep, a, b) column, b fastest, in one group per kernel set
    {
        bool full = S <= 2;
        for (int k = 0; k < 2 * S * S; ++k) full = full && ck[k] >= 0 && ck[k] == ck[0];
        plan.full_pairs = full;
    }
    const int TM = opgemm_tile_rows(plan.wide), TN = opgemm_tile_cols(plan.wide);
    std::vector<OpGroup> gv;
    std::vector<PropCol> cols;
    std::vector<OpTile> tiles;
    for (auto& kv : groups) {
        OpGroup g;
        memset(&g, 0, sizeof(g));
        g.kset = kv.first.kset;
        g.L = set_L[g.kset];
        g.NP = g.L / 2;
        g.nterms = 1;
        g.kern[0] = kv.first.kern;
        g.sign[0] = 1;
        g.col_off = (int32_t)cols.size();
        g.ncols = (int32_t)kv.second.size();
        g.kbase = set_kbase[g.kset];
        g.band = band[g.kset];
        const int gi = (int)gv.size();
        gv.push_back(g);
        cols.insert(cols.end(), kv.second.begin(), kv.second.end());
        for (int c0 = 0; c0 < g.ncols; c0 += TN)
            for (int r = 0; r < g.NP; r += TM) tiles.push_back(OpTile{gi, r, c0, std::min(TN, g.ncols - c0)});
    }
    tiles = xcd_order(tiles);
    plan.ntiles = (int)tiles.size();
    if (plan.ntiles == 0) return SDFT_OK;
    TRY(upload_vec(ctx, &plan.d_tiles, tiles));
    TRY(upload_vec(ctx, &plan.d_groups, gv));
    TRY(upload_vec(ctx, &plan.d_cols, cols));
    return SDFT_OK;
}

// Host-only checks of a descriptor, run before any device call: required tables present, grid sizes
// in range, and every contact index inside its point's grid.
static int validate_desc(const sdft_desc_t* d) {
    const int P = d->n_points, NB = d->NB, nsp = d->n_species;
    if (!d->diameters || !d->valences || (d->MB > 0 && !d->bond_types) || !d->species_nseg || !d->species_model ||
        !d->species_nstates || !d->seg_states || !d->NP || !d->dz || !d->delta_muH || !d->mu_species || !d->Ext ||
        !d->trapez || !d->surf) {
        set_error(nullptr, SDFT_E_ARG, "a required descriptor table is NULL");
        return SDFT_E_ARG;
    }
    for (int p = 0; p < P; ++p)
        if (d->NP[p] < 4 || d->NP[p] > SDFT_MAX_TILES * SDFT_ACC_POINTS || !(d->dz[p] > 0.0)) {
            set_error(nullptr, SDFT_E_ARG, "point %d: NP=%d dz=%g", p, d->NP[p], d->dz[p]);
            return SDFT_E_ARG;
        }
    if (d->hard_walls && (!d->wall_lo || !d->wall_hi)) {
        set_error(nullptr, SDFT_E_ARG, "hard_walls = 1 needs both contact-index tables wall_lo and wall_hi");
        return SDFT_E_ARG;
    }
    if (d->ndim > 1) {
        if (d->ndim > 3 || !d->grid || !d->bins) {
            set_error(nullptr, SDFT_E_ARG, "ndim = %d needs 2 or 3 dimensions and the grid and bins tables", d->ndim);
            return SDFT_E_ARG;
        }
        for (int p = 0; p < P; ++p) {
            int64_t prod = 1;
            double vol = 1.0;
            for (int k = 0; k < d->ndim; ++k) {
                const int n = d->grid[(size_t)p * 3 + k];
                const double h = d->bins[(size_t)p * 3 + k];
                if (n < 2 || n > 128 || !(h > 0.0)) {
                    set_error(nullptr, SDFT_E_ARG, "point %d dimension %d: %d points of width %g (2..128 points)", p, k, n, h);
                    return SDFT_E_ARG;
                }
                if (n != d->grid[k] || h != d->bins[k]) {
                    set_error(nullptr, SDFT_E_ARG, "point %d: every point of a multi-dimensional context needs the same grid", p);
                    return SDFT_E_ARG;
                }
                prod *= n;
                vol *= h;
            }
            if (prod != d->NP[p] || fabs(d->dz[p] - vol) > 1e-12 * vol) {
                set_error(nullptr, SDFT_E_ARG, "point %d: NP = %d and dz = %g must be the grid's %lld points and volume element %g",
                          p, d->NP[p], d->dz[p], (long long)prod, vol);
                return SDFT_E_ARG;
            }
        }
        if (d->hard_walls || d->wall_charge || d->sw_wall_idx || (d->functionals & SDFT_F_SW)) {
            set_error(nullptr, SDFT_E_UNSUPPORTED, "multi-dimensional grids: no contact-value tables, wall charges table or "
                                                   "square-well terms (the wall charge enters through surf)");
            return SDFT_E_UNSUPPORTED;
        }
        for (int u = 0; u < nsp; ++u)
            if (d->ndim != 3 && d->species_fixed_seg && d->species_fixed_seg[u] >= 0) {
                set_error(nullptr, SDFT_E_UNSUPPORTED, "two-dimensional grids: species %d has a grafted segment "
                                                       "(the reference places grafted chains in 3-D only)", u);
                return SDFT_E_UNSUPPORTED;
            }
    }
    if ((d->sw_wall_idx != nullptr) != (d->sw_wall_fac != nullptr)) {
        set_error(nullptr, SDFT_E_ARG, "square-well walls need both sw_wall_idx and sw_wall_fac");
        return SDFT_E_ARG;
    }
    for (int p = 0; p < P; ++p) {
        const int NPp = d->NP[p];
        for (int j = 0; j < NB && d->hard_walls; ++j) {
            const int32_t lo = d->wall_lo[(size_t)p * NB + j], hi = d->wall_hi[(size_t)p * NB + j];
            if (lo < 0 || lo >= NPp || hi < 0 || hi >= NPp) {
                set_error(nullptr, SDFT_E_ARG, "point %d bead %d: hard-wall contact indices (%d, %d) outside [0, %d)",
                          p, j, lo, hi, NPp);
                return SDFT_E_ARG;
            }
        }
        for (int k = 0; k < 2 * NB && d->sw_wall_idx; ++k) {
            const int32_t i = d->sw_wall_idx[(size_t)p * NB * 2 + k];
            if (i < -1 || i >= NPp) {
                set_error(nullptr, SDFT_E_ARG, "point %d: square-well wall index %d outside the grid", p, i);
                return SDFT_E_ARG;
            }
        }
    }
    (void)nsp;
    return SDFT_OK;
}

// delta_muH, mu_species and the chain normalisations that follow from them, for one point: normative
// (child_propagators.jl:41-50) is the product over segments of (sum_states e^{-dmuH})^(1/N); a free
// chain is normalised by normative^(N-1) (handle_fixed.jl:105-107), a free monomerbead species by 1
// (monomerbead/handle_fixed.jl:39-41). Needs the context's species tables (nseg, S, state_off,
// seg_states_h, model).
static void set_point_chemistry(const sdft_ctx_t* ctx, DevPoint& q, const double* dmu, const double* mus) {
    const int NB = ctx->c.NB, nsp = ctx->c.nsp;
    for (int j = 0; j < NB; ++j) q.delta_muH[j] = dmu[j];
    for (int u = 0; u < nsp; ++u) {
        q.mu_species[u] = mus[u];
        const int N = ctx->nseg[u], S = ctx->S[u];
        double normative = 1.0;
        for (int s2 = N - 1; s2 >= 0; --s2) {
            double t = 0.0;
            for (int a = 0; a < S; ++a) {
                const int b = ctx->seg_states_h[ctx->state_off[u] + s2 * S + a];
                if (b >= 0) t += exp(-q.delta_muH[b]);
            }
            normative *= pow(t, 1.0 / N);
        }
        q.normative[u] = normative;
        q.chain_norm[u] = ctx->c.species[u].model == SDFT_MODEL_MONOMERBEAD ? 1.0 : pow(normative, (double)(N - 1));
    }
}

extern "C" int32_t sdft_create(const sdft_desc_t* d, sdft_ctx_t** out) {
    if (!out || !d) {
        set_error(nullptr, SDFT_E_ARG, "null argument");
        return SDFT_E_ARG;
    }
    *out = nullptr;
    if (d->abi_version != SDFT_ABI_VERSION) {
        set_error(nullptr, SDFT_E_ARG, "abi_version %d != %d", d->abi_version, SDFT_ABI_VERSION);
        return SDFT_E_ARG;
    }
    const int P = d->n_points, NB = d->NB, MB = d->MB, nsp = d->n_species;
    if (P <= 0 || NB <= 0 || NB > SDFT_MAX_NB || MB < 0 || MB > SDFT_MAX_MB || nsp <= 0 || nsp > SDFT_MAX_SPECIES) {
        set_error(nullptr, SDFT_E_ARG, "bad sizes P=%d NB=%d MB=%d n_species=%d", P, NB, MB, nsp);
        return SDFT_E_ARG;
    }
    const int F = d->functionals;
    if (((F & SDFT_F_MFMT) && (F & SDFT_F_TPT1MFMT)) || ((F & SDFT_F_MSA) && (F & SDFT_F_TPT1MSA))) {
        set_error(nullptr, SDFT_E_UNSUPPORTED, "list at most one FMT and one MSA functional");
        return SDFT_E_UNSUPPORTED;
    }
    {
        const int vr = validate_desc(d);
        if (vr != SDFT_OK) return vr;
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) {
        set_error(nullptr, SDFT_E_HIP, "no HIP device available");
        return SDFT_E_HIP;
    }
    sdft_ctx* ctx = new sdft_ctx();
    (void)hipGetDevice(&ctx->device);
    {
        // this device's copy of the static scratch array, looked up once per device (device current)
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        auto it = g_scratch_ptr.find(ctx->device);
        if (it == g_scratch_ptr.end()) {
            void* ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                set_error(ctx, SDFT_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
                delete ctx;
                return SDFT_E_HIP;
            }
            it = g_scratch_ptr.emplace(ctx->device, static_cast<double*>(ptr)).first;
        }
        ctx->d_scratch = it->second;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(ctx, SDFT_E_HIP, "hipStreamCreate failed");
        delete ctx;
        return SDFT_E_HIP;
    }
    // second stream and fork / join events: iterations with points in both precisions run the FP32
    // and FP64 pipelines side by side (they touch disjoint points)
    if (hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
        set_error(ctx, SDFT_E_HIP, "hipStreamCreate / hipEventCreate failed");
        sdft_destroy(ctx);
        return SDFT_E_HIP;
    }
    int rc = SDFT_OK;
    auto fail = [&](int code) {
        sdft_destroy(ctx);
        return code;
    };
    Ctx& c = ctx->c;
    const int n_sw = (F & SDFT_F_SW) ? d->n_sw_pairs : 0;
    if ((F & SDFT_F_SW) && (n_sw <= 0 || n_sw > SDFT_MAX_SW || !d->sw_pairs || !d->sw_energy || !d->sw_lambda)) {
        set_error(ctx, SDFT_E_ARG, "square well: need 1..%d pairs with energies and lambdas (got %d)", SDFT_MAX_SW, n_sw);
        return fail(SDFT_E_ARG);
    }
    c.P = P; c.NB = NB; c.MB = MB; c.nsp = nsp; c.K = 3 * NB + MB + 1 + 2 * n_sw;
    if (d->ndim > 1) {
        c.nd = d->ndim;
        c.ndp.D = d->ndim;
        c.ndp.NPt = d->NP[0];
        for (int k = 0; k < d->ndim; ++k) {
            c.ndp.n[k] = d->grid[k];
            c.ndp.h[k] = d->bins[k];
        }
    }
    ctx->lB = d->bjerrum;
    ctx->valences.assign(d->valences, d->valences + NB);

    // ---- system
    DevSystem sys;
    memset(&sys, 0, sizeof(sys));
    sys.NB = NB; sys.MB = MB; sys.n_species = nsp; sys.K = c.K; sys.functionals = F; sys.lB = d->bjerrum;
    for (int j = 0; j < NB; ++j) {
        if (!(d->diameters[j] > 0.0)) {
            set_error(ctx, SDFT_E_ARG, "diameter of bead %d must be positive", j);
            return fail(SDFT_E_ARG);
        }
        sys.sigma[j] = d->diameters[j];
        sys.z[j] = d->valences[j];
        int k = 0;
        while (k < sys.ncls && sys.cls_sigma[k] != sys.sigma[j]) ++k;
        if (k == sys.ncls) sys.cls_sigma[sys.ncls++] = sys.sigma[j];
        sys.cls[j] = k;
    }
    for (int u = 0; u < MB; ++u) {
        sys.bond_i[u] = d->bond_types[2 * u];
        sys.bond_j[u] = d->bond_types[2 * u + 1];
        if (sys.bond_i[u] < 0 || sys.bond_j[u] >= NB || sys.bond_i[u] > sys.bond_j[u]) {
            set_error(ctx, SDFT_E_ARG, "bond %d must satisfy 0 <= i <= j < NB", u);
            return fail(SDFT_E_ARG);
        }
    }
    c.ncls = sys.ncls;
    // square-well pairs and the bead types they involve (ascending)
    sys.n_sw = n_sw;
    sys.sw_kbase = 3 * NB + MB + 1;
    for (int u = 0; u < n_sw; ++u) {
        sys.sw_i[u] = d->sw_pairs[2 * u];
        sys.sw_j[u] = d->sw_pairs[2 * u + 1];
        sys.sw_eps[u] = d->sw_energy[u];
        sys.sw_lam[u] = d->sw_lambda[u];
        if (sys.sw_i[u] < 0 || sys.sw_j[u] >= NB || sys.sw_i[u] > sys.sw_j[u]) {
            set_error(ctx, SDFT_E_ARG, "square-well pair %d must satisfy 0 <= i <= j < NB", u);
            return fail(SDFT_E_ARG);
        }
    }
    for (int j = 0; j < NB; ++j) {
        bool in = false;
        for (int u = 0; u < n_sw; ++u) in = in || sys.sw_i[u] == j || sys.sw_j[u] == j;
        if (in) sys.sw_bead[sys.n_sw_beads++] = j;
    }
    sys.nw_sw_base = 3 * NB + 2 * MB;
    sys.hard_walls = d->hard_walls;
    sys.charged_walls = d->wall_charge != nullptr;
    ctx->hard_walls = d->hard_walls;
    ctx->charged_walls = d->wall_charge != nullptr;
    // Contact-value theorem table (contact_value_theorem.jl:1-20): hard walls contribute
    // sum_j rho_j at the contact index of each wall (hard_walls.jl:49-79), charged walls the constant
    // -2 pi lB q^2 (charged_walls.jl:60-81), square-well walls rho_j at the well edge times
    // (exp(energy) - 1) (sw_walls.jl:61-103). validate_desc has checked every index against its grid.
    std::vector<int32_t> cvt_off(P + 1, 0), cvt_idx, cvt_side;
    std::vector<double> cvt_fac;
    for (int p = 0; p < P; ++p) {
        const int NPp = d->NP[p];
        if (d->hard_walls) {
            for (int j = 0; j < NB; ++j) {
                const int32_t lo = d->wall_lo[(size_t)p * NB + j], hi = d->wall_hi[(size_t)p * NB + j];
                cvt_side.push_back(0); cvt_idx.push_back(j * NPp + lo); cvt_fac.push_back(1.0);
                cvt_side.push_back(1); cvt_idx.push_back(j * NPp + hi); cvt_fac.push_back(1.0);
            }
        }
        if (d->wall_charge) {
            for (int w = 0; w < 2; ++w) {
                const double q = d->wall_charge[2 * p + w];
                cvt_side.push_back(w); cvt_idx.push_back(-1); cvt_fac.push_back(-2.0 * M_PI * d->bjerrum * q * q);
            }
        }
        if (d->sw_wall_idx) {
            for (int j = 0; j < NB; ++j)
                for (int w = 0; w < 2; ++w) {
                    const size_t k = ((size_t)p * NB + j) * 2 + w;
                    const int32_t i = d->sw_wall_idx[k];
                    if (i < 0) continue;
                    cvt_side.push_back(w); cvt_idx.push_back(j * NPp + i); cvt_fac.push_back(d->sw_wall_fac[k]);
                }
        }
        cvt_off[p + 1] = (int32_t)cvt_idx.size();
    }
    ctx->cvt_any = !cvt_idx.empty();
    if ((rc = upload(ctx, &ctx->d_cvt_off, cvt_off.data(), P + 1)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &ctx->d_cvt_idx, cvt_idx.data(), (int64_t)cvt_idx.size())) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &ctx->d_cvt_side, cvt_side.data(), (int64_t)cvt_side.size())) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &ctx->d_cvt_fac, cvt_fac.data(), (int64_t)cvt_fac.size())) != SDFT_OK) return fail(rc);

    // ---- species
    std::vector<DevSpecies> spec(nsp);
    std::vector<int64_t> goffs(nsp);
    int state_off = 0, maxd_off = 0;
    int64_t goff = 0;
    c.species.resize(nsp);
    for (int u = 0; u < nsp; ++u) {
        const int N = d->species_nseg[u], S = d->species_nstates[u];
        if (N <= 0 || S <= 0 || S > SDFT_MAX_S) {
            set_error(ctx, SDFT_E_UNSUPPORTED, "species %d: nseg=%d states=%d (at most %d states per segment)", u, N, S, SDFT_MAX_S);
            return fail(SDFT_E_UNSUPPORTED);
        }
        const bool mbead = d->species_model[u] == SDFT_MODEL_MONOMERBEAD;
        if (d->species_model[u] != SDFT_MODEL_FREELY_JOINTED && !mbead) {
            set_error(ctx, SDFT_E_ARG, "species %d: unknown species model %d", u, d->species_model[u]);
            return fail(SDFT_E_ARG);
        }
        const int fseg = d->species_fixed_seg ? d->species_fixed_seg[u] : -1;
        if (fseg > 0 || (fseg == 0 && N < 2 && !mbead)) {
            set_error(ctx, SDFT_E_UNSUPPORTED, "species %d: only a grafted first segment is supported (of a chain, or "
                      "of a monomerbead species)", u);
            return fail(SDFT_E_UNSUPPORTED);
        }
        DevSpecies& s = spec[u];
        memset(&s, 0, sizeof(s));
        s.nseg = N; s.S = S; s.model = d->species_model[u]; s.fixed_seg = fseg;
        s.state_off = state_off; s.maxd_off = maxd_off;
        int multi = 0;
        for (int q = 0; q < N; ++q) {
            int cnt = 0;
            for (int a = 0; a < S; ++a) {
                const int b = d->seg_states[state_off + q * S + a];
                if (b >= NB) {
                    set_error(ctx, SDFT_E_ARG, "species %d segment %d: bead type %d out of range", u, q, b);
                    return fail(SDFT_E_ARG);
                }
                if (b >= 0) ++cnt;
            }
            if (cnt == 0 || d->seg_states[state_off + q * S] < 0) {
                set_error(ctx, SDFT_E_ARG, "species %d segment %d has no leading state", u, q);
                return fail(SDFT_E_ARG);
            }
            if (cnt > 1) multi = 1;
        }
        s.multistate = multi;
        s.homog = 1;
        for (int q = 1; q < N; ++q)
            for (int a = 0; a < S; ++a)
                if (d->seg_states[state_off + q * S + a] != d->seg_states[state_off + a]) s.homog = 0;
        s.rescale = (N >= SDFT_RESCALE_MIN_SEG && s.model == SDFT_MODEL_FREELY_JOINTED) ? 1 : 0;
        goffs[u] = goff;
        SpeciesHost& sh = c.species[u];
        sh.nseg = N; sh.S = S; sh.model = s.model; sh.fixed_seg = fseg; sh.multistate = multi;
        sh.rescale = s.rescale;
        c.nmax_seg = std::max(c.nmax_seg, N);
        c.any_rescale = c.any_rescale || s.rescale;
        sh.state_off = state_off; sh.maxd_off = maxd_off; sh.goff = goff;
        ctx->nseg.push_back(N);
        ctx->S.push_back(S);
        ctx->state_off.push_back(state_off);
        ctx->multistate.push_back(multi);
        ctx->S_max = std::max(ctx->S_max, S);
        c.states_max = std::max(c.states_max, S);
        goff += (int64_t)N * S * S;
        state_off += N * S;
        maxd_off += N;
    }
    ctx->seg_states_h.assign(d->seg_states, d->seg_states + state_off);
    c.GF = (int)goff;
    c.Frt = NB + 2 * MB + sys.ncls;  // bead and bond-end densities x trapez, then one class sum per diameter
    c.Fnw = 3 * NB + 2 * MB + sys.n_sw_beads;  // ... then one square-well field per participating bead type
    c.Fdp = 3 * NB + 2 * MB + 1;
    if (c.nd > 1) {   // one vector block of NB fields per dimension
        c.Fnw = (2 + c.nd) * NB + 2 * MB;
        c.Fdp = (2 + c.nd) * NB + 2 * MB + 1;
    }

    // ---- points and kernel sets (one set per distinct (NP, dz))
    std::map<std::pair<int, uint64_t>, int> setmap;
    std::vector<int32_t> set_L;
    std::vector<int64_t> set_kbase;
    std::vector<double> set_dz;
    ctx->pts.resize(P);
    int64_t base = 0, xoff = 0, kb = 0, gsum = 0;
    for (int p = 0; p < P; ++p) {
        DevPoint& q = ctx->pts[p];
        memset(&q, 0, sizeof(q));
        const int NP = d->NP[p];
        if (NP < 4 || NP > SDFT_MAX_TILES * SDFT_ACC_POINTS || !(d->dz[p] > 0.0)) {
            set_error(ctx, SDFT_E_ARG, "point %d: NP=%d dz=%g", p, NP, d->dz[p]);
            return fail(SDFT_E_ARG);
        }
        q.NP = NP; q.L = 2 * NP; q.base = base; q.xoff = xoff;
        q.gstride = (NP + 31) / 32 * 32;
        q.gbase = gsum * c.GF;
        gsum += q.gstride;
        q.n = (int64_t)NP * (NB + 2 * MB + 1) + 1;
        q.dz = d->dz[p];   // multi-dimensional grids: the volume element prod(bins) (validated)
        q.nd = 1;
        q.gn[0] = NP;
        q.gh[0] = d->dz[p];
        if (d->ndim > 1) {
            q.nd = d->ndim;
            for (int k = 0; k < d->ndim; ++k) {
                q.gn[k] = d->grid[(size_t)p * 3 + k];
                q.gh[k] = d->bins[(size_t)p * 3 + k];
            }
        }
        q.total_charge = d->total_charge ? d->total_charge[p] : 0.0;
        uint64_t bits;
        memcpy(&bits, &d->dz[p], 8);
        auto key = std::make_pair(NP, bits);
        auto it = setmap.find(key);
        if (it == setmap.end()) {
            const int sidx = (int)set_L.size();
            setmap[key] = sidx;
            set_L.push_back(2 * NP);
            set_kbase.push_back(kb);
            set_dz.push_back(d->dz[p]);
            kb += 2 * NP;
            q.kset = sidx;
        } else {
            q.kset = it->second;
        }
        q.kbase = set_kbase[q.kset];
        for (int u = 0; u < nsp; ++u) {
            q.fixed_coord[u] = d->fixed_coord ? d->fixed_coord[(int64_t)p * nsp + u] : -1;
            if (spec[u].fixed_seg >= 0 && (q.fixed_coord[u] < 0 || q.fixed_coord[u] >= NP)) {
                set_error(ctx, SDFT_E_ARG, "point %d species %d: grafting coordinate %d outside [0, %d)", p, u, q.fixed_coord[u], NP);
                return fail(SDFT_E_ARG);
            }
        }
        set_point_chemistry(ctx, q, d->delta_muH + (int64_t)p * NB, d->mu_species + (int64_t)p * nsp);
        for (int u = 0; u < nsp; ++u) q.graft_density[u] = d->fixed_density ? d->fixed_density[u] : 0.0;
        for (int j = 0; j < NB; ++j) {
            q.wall_lo[j] = d->wall_lo ? d->wall_lo[(int64_t)p * NB + j] : 0;
            q.wall_hi[j] = d->wall_hi ? d->wall_hi[(int64_t)p * NB + j] : 0;
        }
        if (d->wall_charge) { q.wall_q[0] = d->wall_charge[2 * p]; q.wall_q[1] = d->wall_charge[2 * p + 1]; }
        base += NP;
        xoff += q.n;
        c.maxNP = std::max(c.maxNP, NP);
        c.max_n = std::max(c.max_n, q.n);
    }
    c.sumNP = base;
    c.sumGS = gsum;
    c.total_n = xoff;
    c.sumL_sets = kb;
    ctx->nsets = (int)set_L.size();

    // ---- jobs
    std::vector<ConvJob> fwd, pre, back;
    if (build_jobs(d, sys, fwd, pre, back, c.Fws) != SDFT_OK) {
        set_error(ctx, SDFT_E_UNSUPPORTED, "square well: at most %d partner bead types per bead type", SDFT_SW_TERMS);
        return fail(SDFT_E_UNSUPPORTED);
    }
    c.n_fwd_jobs = (int)fwd.size();
    c.n_pre_jobs = (int)pre.size();
    c.n_back_jobs = (int)back.size();
    for (const ConvJob& j : fwd) ctx->fwd_terms += j.nterms;
    for (const ConvJob& j : pre) ctx->back_terms += j.nterms;
    for (const ConvJob& j : back) ctx->back_terms += j.nterms;
    std::vector<ConvJob> jobs(fwd);
    jobs.insert(jobs.end(), pre.begin(), pre.end());
    jobs.insert(jobs.end(), back.begin(), back.end());
    // second copy for the merged back stage: the projections are stored without the shared
    // partials, which k_add_partials adds once the launch that computes both has finished
    c.n_jobs = (int)jobs.size();
    for (int j = 0; j < c.n_jobs; ++j) {
        ConvJob J = jobs[j];
        if (j >= c.n_fwd_jobs + c.n_pre_jobs) {
            J.addfield = -1;
            J.addfield2 = -1;
        }
        jobs.push_back(J);
    }

    // ---- uploads
    if ((rc = upload(ctx, &c.d_sys, &sys, 1)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &c.d_pts, ctx->pts.data(), P)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &c.d_spec, spec.data(), nsp)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &c.d_goffs, goffs.data(), nsp)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &c.d_seg_states, d->seg_states, state_off)) != SDFT_OK) return fail(rc);
    c.h_seg_states.assign(d->seg_states, d->seg_states + state_off);
    c.h_bond_i.assign(sys.bond_i, sys.bond_i + c.MB);
    c.h_bond_j.assign(sys.bond_j, sys.bond_j + c.MB);
    std::vector<double> maxd(maxd_off, 0.0);
    if (d->max_distance) maxd.assign(d->max_distance, d->max_distance + maxd_off);
    if ((rc = upload(ctx, &c.d_maxd, maxd.data(), maxd_off)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &c.d_jobs, jobs.data(), (int64_t)jobs.size())) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &c.d_Ext, d->Ext, c.sumNP * NB)) != SDFT_OK) return fail(rc);
    std::vector<double> surf(c.sumNP, 0.0);
    if (d->surf) surf.assign(d->surf, d->surf + c.sumNP);
    if ((rc = upload(ctx, &c.d_surf, surf.data(), c.sumNP)) != SDFT_OK) return fail(rc);
    double* d_trap64 = nullptr;
    if ((rc = upload(ctx, &d_trap64, d->trapez, c.sumNP * NB)) != SDFT_OK) return fail(rc);
    ctx->p64.trapez = d_trap64;
    if ((rc = dalloc(ctx, &ctx->p32.trapez, c.sumNP * NB)) != SDFT_OK) return fail(rc);
    launch_convert_f32(ctx->stream, d_trap64, ctx->p32.trapez, c.sumNP * NB);

    if (c.nd > 1) {
        std::string nerr;
        const int nrc = nd_setup(c, sys, fwd, pre, back, ctx->allocs, nerr);
        if (nrc != SDFT_OK) {
            set_error(ctx, nrc, "%s", nerr.c_str());
            return fail(nrc);
        }
        hipError_t e = prepare_kernel_attributes(c);
        if (e == hipSuccess) e = nd_prepare_attributes(c);
        if (e != hipSuccess) {
            set_error(ctx, SDFT_E_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
            return fail(SDFT_E_HIP);
        }
        c.use_gemm = false;
        c.merge_back = false;
        ctx->jobs_h = jobs;
        c.prop_plan.assign(nsp, OpPlan());
        c.chain_plan.assign(nsp, ChainPlanHost());
    } else {
    // kernel vectors
    int32_t* d_setL = nullptr;
    int64_t* d_setkb = nullptr;
    double* d_setdz = nullptr;
    double* d_wbuf = nullptr;
    if ((rc = upload(ctx, &d_setL, set_L.data(), ctx->nsets)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &d_setkb, set_kbase.data(), ctx->nsets)) != SDFT_OK) return fail(rc);
    if ((rc = upload(ctx, &d_setdz, set_dz.data(), ctx->nsets)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &d_wbuf, c.sumL_sets * c.K)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &ctx->p64.kern, c.sumL_sets * c.K)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &ctx->p32.kern, c.sumL_sets * c.K)) != SDFT_OK) return fail(rc);
    {
        hipError_t e = prepare_kernel_attributes(c);
        if (e == hipSuccess) e = fft_prepare_attributes();
        if (e != hipSuccess) {
            set_error(ctx, SDFT_E_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
            return fail(SDFT_E_HIP);
        }
    }
    int maxL = 0;
    for (int L : set_L) maxL = std::max(maxL, L);
    launch_build_kernels(ctx->stream, c.d_sys, c.K, ctx->nsets, maxL, d_setL, d_setkb, d_setdz, d_wbuf, ctx->p64.kern,
                         ctx->p32.kern);
    HIP_CHECK_RET(hipGetLastError());
    // banded half-width of the bond kernels (FP32 phase): beyond it every |c| <= 1e-9 max|c|
    {
        std::vector<double> hk((size_t)c.sumL_sets * c.K);
        HIP_CHECK_RET(hipMemcpyAsync(hk.data(), ctx->p64.kern, hk.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
        std::vector<int32_t> band(ctx->nsets, 0);
        for (int s = 0; s < ctx->nsets; ++s) {
            const int L = set_L[s];
            int B = 0;
            for (int u = 0; u < MB; ++u) {
                const double* cv = hk.data() + set_kbase[s] * c.K + (int64_t)(3 * NB + u) * L;
                double mx = 0.0;
                for (int i = 0; i < L; ++i) mx = std::max(mx, fabs(cv[i]));
                int b = 0;
                for (int i = 0; i <= L / 2; ++i)
                    if (fabs(cv[i]) > 1e-9 * mx || fabs(cv[(L - i) % L]) > 1e-9 * mx) b = i;
                B = std::max(B, b);
            }
            band[s] = B;
        }
        if ((rc = upload(ctx, &c.d_band, band.data(), ctx->nsets)) != SDFT_OK) return fail(rc);
        HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
        // canonical kernel index per set: identical vectors (equal diameters) share one operator
        std::vector<std::vector<int>> canon(ctx->nsets, std::vector<int>(c.K));
        for (int s = 0; s < ctx->nsets; ++s) {
            const int L = set_L[s];
            for (int k = 0; k < c.K; ++k) {
                canon[s][k] = k;
                const double* ck = hk.data() + set_kbase[s] * c.K + (int64_t)k * L;
                for (int k2 = 0; k2 < k; ++k2) {
                    const double* c2 = hk.data() + set_kbase[s] * c.K + (int64_t)k2 * L;
                    if (memcmp(ck, c2, (size_t)L * sizeof(double)) == 0) { canon[s][k] = canon[s][k2]; break; }
                }
            }
        }
        const char* kenv = getenv("SOFTDFT_KERNELS");
        c.use_gemm = !(kenv && strcmp(kenv, "direct") == 0);
        c.merge_back = c.use_gemm && c.n_pre_jobs > 0;
        ctx->jobs_h = jobs;
        ctx->canon = canon;
        ctx->set_L_h = set_L;
        ctx->set_kbase_h = set_kbase;
        for (auto& v : ctx->set_shape) v.assign(ctx->nsets, -1);
        if ((rc = rebuild_conv_plans(ctx)) != SDFT_OK) return fail(rc);
        // banded propagator steps on k_bandstep: every point's applied half-width (the band rounded up
        // to even) fits its grid, and the workgroup's operator windows and inputs fit 64 KB of LDS
        for (int u = 0; u < nsp; ++u) {
            SpeciesHost& sh = c.species[u];
            sh.band_step = 0;
            sh.band_be = 0;
            if (sh.nseg < 2 || sh.S > SDFT_FAST_S) continue;
            bool ok = true;
            for (int p = 0; p < P; ++p) {
                const int Be = chain_half_width(band[ctx->pts[p].kset]);
                ok = ok && 2 * Be + 1 <= ctx->pts[p].NP;
                sh.band_be = std::max(sh.band_be, Be);
            }
            sh.band_step = ok && bandstep_lds_bytes(sh.S, sh.band_be, sizeof(double)) <= 65536 ? 1 : 0;
        }
        c.prop_plan.assign(nsp, OpPlan());
        for (int u = 0; u < nsp; ++u)
            if ((rc = build_prop_plan(ctx, sys, u, canon, set_L, set_kbase, band, c.prop_plan[u])) != SDFT_OK) return fail(rc);
        c.chain_plan.assign(nsp, ChainPlanHost());
        const char* cenv = getenv("SOFTDFT_CHAIN");
        const bool chain_on = c.use_gemm && !(cenv && strcmp(cenv, "off") == 0);
        for (int u = 0; u < nsp && chain_on; ++u) {
            if ((rc = build_chain_plan(ctx, sys, u, canon, band, c.chain_plan[u])) != SDFT_OK) return fail(rc);
            if (!c.chain_plan[u].ok) continue;
            const hipError_t e = chain_prepare(c.chain_plan[u]);
            if (e != hipSuccess) {
                set_error(ctx, SDFT_E_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
                return fail(SDFT_E_HIP);
            }
        }
    }
    }   // one-dimensional kernel vectors and plans

    // ---- work buffers
    auto alloc_pipe = [&](auto& pp) -> int {
        TRY(dalloc(ctx, &pp.elam, c.sumNP * NB));
        TRY(dalloc(ctx, &pp.hl, c.sumNP * 2 * MB));
        TRY(dalloc(ctx, &pp.gC, c.sumGS * c.GF));
        TRY(dalloc(ctx, &pp.gP, c.sumGS * c.GF));
        TRY(dalloc(ctx, &pp.smax, (int64_t)P * c.nsp * 2 * c.nmax_seg));
        TRY(dalloc(ctx, &pp.lsc, (int64_t)P * c.nsp * (c.nmax_seg + 1)));
        TRY(dalloc(ctx, &pp.rho_beads, c.sumNP * NB));
        TRY(dalloc(ctx, &pp.rho_bonds, c.sumNP * 2 * MB));
        TRY(dalloc(ctx, &pp.rhoT, c.sumNP * c.Frt));
        TRY(dalloc(ctx, &pp.nw, c.sumNP * c.Fnw));
        TRY(dalloc(ctx, &pp.dphi, c.sumNP * c.Fdp));
        if (c.nd > 1) {
            TRY(dalloc(ctx, &pp.nd_a, (int64_t)P * c.ndp.maxF * c.ndp.NPt));
            TRY(dalloc(ctx, &pp.nd_b, (int64_t)P * c.ndp.maxF * c.ndp.NPt));
        }
        return SDFT_OK;
    };
    {
        int Smax = 1;
        for (const auto& sh : c.species) Smax = std::max(Smax, sh.S);
        if ((rc = dalloc(ctx, &ctx->p32.fbuf, 2 * c.sumNP * 2 * Smax * Smax)) != SDFT_OK) return fail(rc);
        if ((rc = dalloc(ctx, &ctx->p64.fbuf, 2 * c.sumNP * 2 * Smax * Smax)) != SDFT_OK) return fail(rc);
    }
    if ((rc = alloc_pipe(ctx->p32)) != SDFT_OK) return fail(rc);
    if ((rc = alloc_pipe(ctx->p64)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_evalflag, P)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_errflag, P)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_acc_species, (int64_t)P * 2 * SDFT_MAX_SPECIES)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_acc_part, (int64_t)P * SDFT_MAX_TILES * 2 * SDFT_MAX_SPECIES)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_fixed_lambda, (int64_t)P * SDFT_MAX_SPECIES)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_x, c.total_n)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_fx, c.total_n)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_gram, (int64_t)P * SDFT_GRAM_BLOCKS * SDFT_GRAM_LEN)) != SDFT_OK) return fail(rc);
    if (c.Fws > 0 && (rc = dalloc(ctx, &c.d_wsum, c.sumNP * c.Fws)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_errbits, P)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_state, P)) != SDFT_OK) return fail(rc);
    if ((rc = dalloc(ctx, &c.d_n_active, 3)) != SDFT_OK) return fail(rc);
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    *out = ctx;
    return SDFT_OK;
}

extern "C" int64_t sdft_solution_length(const sdft_ctx_t* ctx) { return ctx ? ctx->c.total_n : -1; }

extern "C" int32_t sdft_solution_offsets(const sdft_ctx_t* ctx, int64_t* offsets) {
    if (!ctx || !offsets) return SDFT_E_ARG;
    for (int p = 0; p < ctx->c.P; ++p) offsets[p] = ctx->pts[p].xoff;
    offsets[ctx->c.P] = ctx->c.total_n;
    return SDFT_OK;
}

static int check_errflags(sdft_ctx* ctx) {
    std::vector<int32_t> ef(ctx->c.P);
    HIP_CHECK_RET(hipMemcpyAsync(ef.data(), ctx->c.d_errflag, ef.size() * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < ctx->c.P; ++p)
        if (ef[p] != 0) {
            set_error(ctx, ef[p], "Normalization error: non-finite or non-positive chain normalisation at point %d", p);
            return ef[p];
        }
    return SDFT_OK;
}

// evaluates F(x) for every point whose evalflag is 32 or 64; a pipeline with no such point is skipped
// When both precisions have points, the FP64 pipeline forks onto the second stream and joins before
// the Anderson step: a point is evaluated at one precision per iteration, and every buffer the two
// pipelines share (x, F(x), densities' partial sums, shared partials, error flags) is written per
// point, so they touch disjoint memory. Under graph capture the fork and join become graph edges.
static void run_objective(sdft_ctx* ctx, bool run32, bool run64) {
    if (run32 && run64 && ctx->concurrent) {
        (void)hipEventRecord(ctx->ev_fork, ctx->stream);
        (void)hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0);
        ctx->p64.objective(ctx->stream2, ctx->c, 64, ctx->c.d_x, ctx->c.d_fx);
        ctx->p32.objective(ctx->stream, ctx->c, 32, ctx->c.d_x, ctx->c.d_fx);
        (void)hipEventRecord(ctx->ev_join, ctx->stream2);
        (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        return;
    }
    if (run32) ctx->p32.objective(ctx->stream, ctx->c, 32, ctx->c.d_x, ctx->c.d_fx);
    if (run64) ctx->p64.objective(ctx->stream, ctx->c, 64, ctx->c.d_x, ctx->c.d_fx);
}

extern "C" int32_t sdft_objective(sdft_ctx_t* ctx, const double* x, double* fx, int32_t precision) {
    if (!ctx || !x || !fx) return SDFT_E_ARG;
    if (precision != SDFT_PREC_FP32 && precision != SDFT_PREC_FP64) {
        set_error(ctx, SDFT_E_ARG, "precision must be 32 or 64");
        return SDFT_E_ARG;
    }
    Ctx& c = ctx->c;
    HIP_CHECK_RET(hipMemcpyAsync(c.d_x, x, c.total_n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    launch_zero_i32(ctx->stream, c.d_errflag, c.P);
    launch_set_evalflag(ctx->stream, c.d_evalflag, precision, c.P);
    // record the evaluation precision in the solve state so observables read the right pipeline
    std::vector<DevSolveState> st(c.P);
    for (auto& s : st) { memset(&s, 0, sizeof(s)); s.prec = precision; s.converged = 1; s.mode = -1; }
    HIP_CHECK_RET(hipMemcpyAsync(c.d_state, st.data(), st.size() * sizeof(DevSolveState), hipMemcpyHostToDevice, ctx->stream));
    c.band_fp32 = ctx->obj_band32;
    c.band_fp64 = ctx->obj_band64;
    if (c.use_gemm) launch_compact_cols(ctx->stream, c, precision);
    run_objective(ctx, precision == 32, precision == 64);
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(hipMemcpyAsync(fx, c.d_fx, c.total_n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    ctx->have_eval = true;
    ctx->solved = false;
    return check_errflags(ctx);
}

extern "C" int32_t sdft_solve(sdft_ctx_t* ctx, const sdft_solver_opts_t* o, double* x_inout, int32_t* iters, double* err) {
    if (!ctx || !o || !x_inout) return SDFT_E_ARG;
    Ctx& c = ctx->c;
    if (o->anderm < 2 || o->anderm > SDFT_MAX_M || o->max_iters < 1) {
        set_error(ctx, SDFT_E_ARG, "anderm must be in [2, %d] and max_iters >= 1", SDFT_MAX_M);
        return SDFT_E_ARG;
    }
    const int prec = o->precision;
    if (prec != SDFT_PREC_FP32 && prec != SDFT_PREC_FP64 && prec != SDFT_PREC_MIXED) {
        set_error(ctx, SDFT_E_ARG, "bad precision %d", prec);
        return SDFT_E_ARG;
    }
    // (re)size the Anderson history
    if (!c.d_Xh || c.m != o->anderm) {
        dfree(ctx, c.d_Xh);
        dfree(ctx, c.d_Gh);
        c.d_Xh = c.d_Gh = nullptr;
        TRY(dalloc(ctx, &c.d_Xh, c.total_n * o->anderm));
        TRY(dalloc(ctx, &c.d_Gh, c.total_n * o->anderm));
        c.m = o->anderm;
    }
    if (o->band_fp32 < 0 || o->band_fp32 > 2) {
        set_error(ctx, SDFT_E_ARG, "band_fp32 must be 0, 1 or 2");
        return SDFT_E_ARG;
    }
    c.band_fp32 = o->band_fp32 ? 1 : 0;
    // the plain FP64 path stays dense unless every precision is asked to run banded (2)
    c.band_fp64 = ((prec == SDFT_PREC_MIXED && o->band_fp32) || o->band_fp32 == 2) ? 1 : 0;
    SolveParams sp;
    sp.tole = o->tole; sp.damping = o->damping; sp.mixing_max = o->mixing_max; sp.switch_tole = o->switch_tole;
    sp.fp32_switch = o->fp32_switch; sp.m = o->anderm; sp.max_iters = o->max_iters;
    sp.mixed = prec == SDFT_PREC_MIXED;
    sp.start_prec = prec == SDFT_PREC_FP64 ? 64 : 32;
    std::vector<DevSolveState> st(c.P);
    for (auto& s : st) {
        memset(&s, 0, sizeof(s));
        s.use_picard = 1;
        s.prec = sp.start_prec;
        s.mode = -1;
    }
    HIP_CHECK_RET(hipMemcpyAsync(c.d_state, st.data(), st.size() * sizeof(DevSolveState), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK_RET(hipMemcpyAsync(c.d_x, x_inout, c.total_n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    launch_zero_i32(ctx->stream, c.d_errflag, c.P);
    launch_set_evalflag(ctx->stream, c.d_evalflag, sp.start_prec, c.P);
    launch_zero_f64(ctx->stream, (double*)c.d_errbits, c.P);
    EventPair ev;
    HIP_CHECK_RET(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    HIP_CHECK_RET(hipEventRecord(e0, ctx->stream));
    const int check = o->check_every > 0 ? o->check_every : 10;
    // the host polls the active counts every `check` iterations; mixed-precision switches are
    // applied at these polls, so between two polls the set of precisions in use is fixed and a
    // pipeline with no point at its precision is not launched
    int32_t counts[3] = {c.P, sp.start_prec == 32 ? c.P : 0, 0};
    bool run32 = sp.start_prec == 32, run64 = sp.start_prec == 64;
    auto compact = [&]() {
        if (!c.use_gemm) return;
        if (run32) launch_compact_cols(ctx->stream, c, 32);
        if (run64) launch_compact_cols(ctx->stream, c, 64);
    };
    compact();
    // optional progress lines on stderr (SDFT_PROGRESS = seconds between lines), written at polls
    const char* prog_env = getenv("SDFT_PROGRESS");
    const double prog_every = prog_env ? atof(prog_env) : 0.0;
    auto prog_t0 = std::chrono::steady_clock::now();
    auto prog_last = prog_t0;
    // The `check` iterations between two polls are a fixed launch sequence for a given pair of
    // active pipelines (every grid size comes from the context, every per-iteration decision is
    // read on the device), so each full window runs as one captured hipGraph, instantiated once per
    // pipeline pair and solve. Windows cut short by max_iters run eagerly.
    WindowGraphs graphs(ctx->stream);
    const bool use_graph = ctx->graph_capture != 0;
    int it = 0, n32 = 0, n64 = 0;
    while (it < o->max_iters) {
        const int w = std::min(check, o->max_iters - it);
        const int key = (run32 ? 1 : 0) | (run64 ? 2 : 0);
        if (use_graph && w == check) {
            if (!graphs.exec[key]) {
                hipGraph_t g = nullptr;
                HIP_CHECK_RET(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < w; ++k) {
                    run_objective(ctx, run32, run64);
                    launch_anderson(ctx->stream, c, sp);
                }
                const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
                if (ec != hipSuccess || !g) {
                    set_error(ctx, SDFT_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
                    return SDFT_E_HIP;
                }
                const hipError_t ei = hipGraphInstantiate(&graphs.exec[key], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    graphs.exec[key] = nullptr;
                    set_error(ctx, SDFT_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
                    return SDFT_E_HIP;
                }
                ++ctx->graphs_built;
            }
            HIP_CHECK_RET(hipGraphLaunch(graphs.exec[key], ctx->stream));
        } else {
            for (int k = 0; k < w; ++k) {
                run_objective(ctx, run32, run64);
                launch_anderson(ctx->stream, c, sp);
            }
        }
        n32 += run32 ? w : 0;
        n64 += run64 ? w : 0;
        it += w;
        {
            HIP_CHECK_RET(hipMemcpyAsync(counts, c.d_n_active, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
            if (prog_every > 0.0) {
                const auto now = std::chrono::steady_clock::now();
                if (std::chrono::duration<double>(now - prog_last).count() >= prog_every) {
                    prog_last = now;
                    HIP_CHECK_RET(hipMemcpyAsync(st.data(), c.d_state, st.size() * sizeof(DevSolveState),
                                                 hipMemcpyDeviceToHost, ctx->stream));
                    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
                    double emax = 0.0;
                    for (const DevSolveState& q : st)
                        if (!q.converged) emax = q.err > emax || q.err != q.err ? q.err : emax;
                    fprintf(stderr, "sdft_solve: iteration %d, %d of %d points active (%d in FP32), max residual %.3e, %.1f s\n",
                            it, counts[0], c.P, counts[1] - counts[2], emax,
                            std::chrono::duration<double>(now - prog_t0).count());
                    fflush(stderr);
                }
            }
            if (counts[0] == 0) break;
            if (counts[2] > 0) launch_apply_switch(ctx->stream, c);
            run32 = counts[1] - counts[2] > 0;
            run64 = counts[0] - counts[1] + counts[2] > 0;
            compact();
        }
    }
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(hipEventRecord(e1, ctx->stream));
    HIP_CHECK_RET(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ctx->last_solve_ms = ms;
    ctx->last_iters32 = n32;
    ctx->last_iters64 = n64;
    HIP_CHECK_RET(hipMemcpyAsync(x_inout, c.d_x, c.total_n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipMemcpyAsync(st.data(), c.d_state, st.size() * sizeof(DevSolveState), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    int any_fail = 0;
    double tot = 0.0;
    ctx->last_switch.assign(c.P, 0);
    for (int p = 0; p < c.P; ++p) {
        ctx->last_switch[p] = st[p].switch_iter;
        if (iters) iters[p] = st[p].iter;
        if (err) err[p] = st[p].err;
        tot += st[p].iter;
        if (st[p].failed || !st[p].converged) any_fail = 1;
    }
    ctx->last_iters_total = tot;
    ctx->solved = true;
    ctx->have_eval = true;
    int rc = check_errflags(ctx);
    if (rc != SDFT_OK) return rc;
    if (any_fail) {
        set_error(ctx, SDFT_E_NOCONV, "at least one state point did not converge (max_iters or non-finite residual)");
        return SDFT_E_NOCONV;
    }
    return SDFT_OK;
}

extern "C" int32_t sdft_last_timing(const sdft_ctx_t* ctx, double* solve_ms, double* iterations_total) {
    if (!ctx) return SDFT_E_ARG;
    if (solve_ms) *solve_ms = ctx->last_solve_ms;
    if (iterations_total) *iterations_total = ctx->last_iters_total;
    return SDFT_OK;
}

extern "C" int32_t sdft_set_operator_band(sdft_ctx_t* ctx, int32_t fp32_banded, int32_t fp64_banded) {
    if (!ctx) return SDFT_E_ARG;
    ctx->obj_band32 = fp32_banded ? 1 : 0;
    ctx->obj_band64 = fp64_banded ? 1 : 0;
    return SDFT_OK;
}

extern "C" int32_t sdft_set_chemistry(sdft_ctx_t* ctx, const double* delta_muH, const double* mu_species) {
    if (!ctx || !delta_muH || !mu_species) return SDFT_E_ARG;
    const int P = ctx->c.P, NB = ctx->c.NB, nsp = ctx->c.nsp;
    for (int64_t i = 0; i < (int64_t)P * NB; ++i)
        if (!std::isfinite(delta_muH[i])) {
            set_error(ctx, SDFT_E_ARG, "delta_muH[%lld] is not finite", (long long)i);
            return SDFT_E_ARG;
        }
    for (int64_t i = 0; i < (int64_t)P * nsp; ++i)
        if (!std::isfinite(mu_species[i])) {
            set_error(ctx, SDFT_E_ARG, "mu_species[%lld] is not finite", (long long)i);
            return SDFT_E_ARG;
        }
    for (int p = 0; p < P; ++p)
        set_point_chemistry(ctx, ctx->pts[p], delta_muH + (int64_t)p * NB, mu_species + (int64_t)p * nsp);
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    HIP_CHECK_RET(hipMemcpyAsync(ctx->c.d_pts, ctx->pts.data(), sizeof(DevPoint) * P, hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    ctx->solved = false;
    ctx->have_eval = false;
    return SDFT_OK;
}

extern "C" int32_t sdft_set_graft_density(sdft_ctx_t* ctx, const double* density) {
    if (!ctx || !density) return SDFT_E_ARG;
    const int P = ctx->c.P, nsp = ctx->c.nsp;
    for (int64_t i = 0; i < (int64_t)P * nsp; ++i)
        if (!(std::isfinite(density[i]) && density[i] >= 0.0)) {
            set_error(ctx, SDFT_E_ARG, "grafting density[%lld] = %g is not a finite non-negative value",
                      (long long)i, density[i]);
            return SDFT_E_ARG;
        }
    for (int p = 0; p < P; ++p)
        for (int u = 0; u < nsp; ++u) ctx->pts[p].graft_density[u] = density[(int64_t)p * nsp + u];
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    HIP_CHECK_RET(hipMemcpyAsync(ctx->c.d_pts, ctx->pts.data(), sizeof(DevPoint) * P, hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    ctx->solved = false;
    ctx->have_eval = false;
    return SDFT_OK;
}

extern "C" int32_t sdft_last_pipeline_iters(const sdft_ctx_t* ctx, int32_t* fp32_iters, int32_t* fp64_iters) {
    if (!ctx) return SDFT_E_ARG;
    if (fp32_iters) *fp32_iters = ctx->last_iters32;
    if (fp64_iters) *fp64_iters = ctx->last_iters64;
    return SDFT_OK;
}

extern "C" int32_t sdft_last_switch_iters(const sdft_ctx_t* ctx, int32_t* switch_iter) {
    if (!ctx || !switch_iter) return SDFT_E_ARG;
    for (int p = 0; p < ctx->c.P; ++p)
        switch_iter[p] = p < (int)ctx->last_switch.size() ? ctx->last_switch[p] : 0;
    return SDFT_OK;
}

extern "C" int32_t sdft_set_launch_mode(sdft_ctx_t* ctx, int32_t graph, int32_t concurrent, int64_t* graphs_built) {
    if (!ctx) return SDFT_E_ARG;
    if (graph >= 0) ctx->graph_capture = graph ? 1 : 0;
    if (concurrent >= 0) ctx->concurrent = concurrent ? 1 : 0;
    if (graphs_built) *graphs_built = ctx->graphs_built;
    return SDFT_OK;
}

template <typename T>
static int download_field(sdft_ctx* ctx, const T* dptr, int64_t count, double* host) {
    std::vector<T> tmp(count);
    HIP_CHECK_RET(hipMemcpyAsync(tmp.data(), dptr, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < count; ++i) host[i] = (double)tmp[i];
    return SDFT_OK;
}

static int point_prec(sdft_ctx* ctx, int p) {
    DevSolveState s;
    (void)hipMemcpyAsync(&s, ctx->c.d_state + p, sizeof(s), hipMemcpyDeviceToHost, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    return s.prec == 32 ? 32 : 64;
}

extern "C" int32_t sdft_get_field(sdft_ctx_t* ctx, int32_t field_id, int32_t point, double* out, int64_t capacity) {
    if (!ctx || !out || point < 0 || point >= ctx->c.P) return SDFT_E_ARG;
    const Ctx& c = ctx->c;
    const DevPoint& P = ctx->pts[point];
    int64_t need = 0;
    switch (field_id) {
        case SDFT_FIELD_RHO_BEADS: need = (int64_t)P.NP * c.NB; break;
        case SDFT_FIELD_RHO_BONDS: need = (int64_t)P.NP * 2 * c.MB; break;
        case SDFT_FIELD_X: case SDFT_FIELD_FX: need = P.n; break;
        default:
            if (field_id >= SDFT_FIELD_RHO_SEGMENTS && field_id < SDFT_FIELD_RHO_SEGMENTS + c.nsp) {
                const SpeciesHost& sh = c.species[field_id - SDFT_FIELD_RHO_SEGMENTS];
                need = (int64_t)P.NP * sh.S * sh.nseg;
                break;
            }
            set_error(ctx, SDFT_E_ARG, "unknown field id %d", field_id);
            return SDFT_E_ARG;
    }
    if (capacity < need) {
        set_error(ctx, SDFT_E_ARG, "capacity %lld < %lld", (long long)capacity, (long long)need);
        return SDFT_E_ARG;
    }
    if (!ctx->have_eval) {
        set_error(ctx, SDFT_E_STATE, "no evaluation has run yet");
        return SDFT_E_STATE;
    }
    const int prec = point_prec(ctx, point);
    if (field_id >= SDFT_FIELD_RHO_SEGMENTS) {
        double* d = nullptr;
        HIP_CHECK_RET(hipMalloc(&d, need * sizeof(double)));
        const int u = field_id - SDFT_FIELD_RHO_SEGMENTS;
        if (prec == 64) ctx->p64.segment_profile(ctx->stream, c, u, point, d);
        else ctx->p32.segment_profile(ctx->stream, c, u, point, d);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out, d, need * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(d);
        if (e != hipSuccess) {
            set_error(ctx, SDFT_E_HIP, "segment profile: %s", hipGetErrorString(e));
            return SDFT_E_HIP;
        }
        return SDFT_OK;
    }
    if (field_id == SDFT_FIELD_X) return download_field(ctx, c.d_x + P.xoff, need, out);
    if (field_id == SDFT_FIELD_FX) return download_field(ctx, c.d_fx + P.xoff, need, out);
    if (field_id == SDFT_FIELD_RHO_BEADS)
        return prec == 64 ? download_field(ctx, ctx->p64.rho_beads + P.base * c.NB, need, out)
                          : download_field(ctx, ctx->p32.rho_beads + P.base * c.NB, need, out);
    return prec == 64 ? download_field(ctx, ctx->p64.rho_bonds + P.base * 2 * c.MB, need, out)
                      : download_field(ctx, ctx->p32.rho_bonds + P.base * 2 * c.MB, need, out);
}

// Work buffer of the observables, in doubles, in this order: Omega block partials [P][8], the segment
// sums of every multistate species ([P][N_u][S_u] each), contact values [P][2], Omega [P], alpha
// [P][n_species][S_max]. Returns its length and fills the offsets.
struct ObsOffsets {
    int64_t om, ss, cvt, omega, alpha, total;
};
static ObsOffsets obs_layout(const sdft_ctx* ctx, ObsLayout* lay) {
    const Ctx& c = ctx->c;
    const int64_t P = c.P;
    ObsOffsets o;
    o.om = 0;
    o.ss = P * SDFT_OMEGA_BLOCKS;
    int64_t ss = 0;
    for (int u = 0; u < SDFT_MAX_SPECIES; ++u) lay->ss_base[u] = 0;
    for (int u = 0; u < c.nsp; ++u) {
        lay->ss_base[u] = ss;
        if (ctx->multistate[u]) ss += P * ctx->nseg[u] * ctx->S[u];
    }
    o.cvt = o.ss + ss;
    o.omega = o.cvt + 2 * P;
    o.alpha = o.omega + P;
    o.total = o.alpha + P * c.nsp * ctx->S_max;
    return o;
}

// grand_potential / write_total_state_fractions / contact_value_theorem for every point, with
// every partial in the device buffer `work` (obs_layout) and the results copied to the caller's
// arrays (any of them may be NULL)
static int observables_run(sdft_ctx* ctx, double* work, double* omega, double* alpha, double* p_cvt) {
    Ctx& c = ctx->c;
    const int P = c.P;
    ObsLayout lay;
    const ObsOffsets o = obs_layout(ctx, &lay);
    double* d_om = work + o.om;
    double* d_ss = work + o.ss;
    ctx->p32.omega(ctx->stream, c, c.d_x, d_om);
    ctx->p64.omega(ctx->stream, c, c.d_x, d_om);
    for (int u = 0; u < c.nsp; ++u) {
        if (!ctx->multistate[u]) continue;
        const int64_t stride = (int64_t)ctx->nseg[u] * ctx->S[u];
        ctx->p32.segment_sums(ctx->stream, c, u, d_ss + lay.ss_base[u], stride);
        ctx->p64.segment_sums(ctx->stream, c, u, d_ss + lay.ss_base[u], stride);
    }
    // contact values of every point in one gather, read from the pipeline the point was evaluated in
    if (ctx->cvt_any)
        launch_cvt_gather(ctx->stream, c, ctx->d_cvt_off, ctx->d_cvt_idx, ctx->d_cvt_side, ctx->d_cvt_fac,
                          ctx->p32.rho_beads, ctx->p64.rho_beads, work + o.cvt);
    else
        launch_zero_f64(ctx->stream, work + o.cvt, 2 * (int64_t)P);
    launch_obs_finalize(ctx->stream, c, lay, ctx->S_max, d_om, d_ss, work + o.omega, work + o.alpha);
    HIP_CHECK_RET(hipGetLastError());
    if (omega)
        HIP_CHECK_RET(hipMemcpyAsync(omega, work + o.omega, (size_t)P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (alpha)
        HIP_CHECK_RET(hipMemcpyAsync(alpha, work + o.alpha, (size_t)P * c.nsp * ctx->S_max * sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->stream));
    if (p_cvt)
        HIP_CHECK_RET(hipMemcpyAsync(p_cvt, work + o.cvt, 2 * (size_t)P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    return SDFT_OK;
}

extern "C" int32_t sdft_observables(sdft_ctx_t* ctx, double* omega, double* alpha, double* p_cvt) {
    if (!ctx) return SDFT_E_ARG;
    if (!ctx->have_eval) {
        set_error(ctx, SDFT_E_STATE, "observables need a prior sdft_solve or sdft_objective");
        return SDFT_E_STATE;
    }
    // the work buffer is allocated on the first call and reused by every later one
    if (!ctx->d_obs) {
        ObsLayout lay;
        TRY(dalloc(ctx, &ctx->d_obs, obs_layout(ctx, &lay).total));
    }
    return observables_run(ctx, ctx->d_obs, omega, alpha, p_cvt);
}

extern "C" int32_t sdft_observables_noalloc(sdft_ctx_t* ctx, double* omega, double* alpha, double* p_cvt) {
    if (!ctx) return SDFT_E_ARG;
    if (!ctx->have_eval) {
        set_error(ctx, SDFT_E_STATE, "observables need a prior sdft_solve or sdft_objective");
        return SDFT_E_STATE;
    }
    if (!ctx->d_scratch) {
        set_error(ctx, SDFT_E_STATE, "the static scratch buffer was not resolved for this device");
        return SDFT_E_STATE;
    }
    ObsLayout lay;
    const int64_t need = obs_layout(ctx, &lay).total;
    if (need > SDFT_SCRATCH_DOUBLES) {
        set_error(ctx, SDFT_E_ARG,
                  "sdft_observables_noalloc: this batch needs %lld doubles of scratch, the static buffer holds %d "
                  "(P * (11 + n_species * S_max) + P * sum over multistate species of N_u * S_u)",
                  (long long)need, SDFT_SCRATCH_DOUBLES);
        return SDFT_E_ARG;
    }
    return observables_run(ctx, ctx->d_scratch, omega, alpha, p_cvt);
}

// HIP-event time of fn() on the context's stream: one warm-up call, then reps calls; us per call
template <class F>
static int time_launches(sdft_ctx* ctx, hipEvent_t e0, hipEvent_t e1, int reps, F&& fn, double* us) {
    fn();
    HIP_CHECK_RET(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; ++r) fn();
    HIP_CHECK_RET(hipEventRecord(e1, ctx->stream));
    HIP_CHECK_RET(hipEventSynchronize(e1));
    HIP_CHECK_RET(hipGetLastError());
    float ms = 0.f;
    HIP_CHECK_RET(hipEventElapsedTime(&ms, e0, e1));
    *us = 1e3 * (double)ms / reps;
    return SDFT_OK;
}

static int autotune_run(sdft_ctx* ctx, int reps, hipEvent_t e0, hipEvent_t e1) {
    Ctx& c = ctx->c;
    hipStream_t st = ctx->stream;
    // ---- propagators: one operator-GEMM launch per step vs the persistent chain kernel with its
    //      input weights held in registers or re-read every step, on the banded operator all of
    //      them take in the mixed path
    for (int u = 0; u < c.nsp; ++u) {
        if (!c.use_gemm || c.species[u].nseg < 2 || !c.chain_plan[u].ok) continue;
        ChainPlanHost& ch = c.chain_plan[u];
        for (int prec : {32, 64}) {
            bool& ok = prec == 32 ? ch.ok32 : ch.ok64;
            bool& rr = prec == 32 ? ch.reread32 : ch.reread64;
            if (!ok) continue;
            launch_set_evalflag(st, c.d_evalflag, prec, c.P);
            auto run = [&](int mode) {
                ok = mode != 0;
                rr = mode == 2;
                if (prec == 32) ctx->p32.propagate(st, c, 32, u, true);
                else ctx->p64.propagate(st, c, 64, u, true);
            };
            double t[3];
            for (int mode = 0; mode < 3; ++mode) TRY(time_launches(ctx, e0, e1, reps, [&]() { run(mode); }, &t[mode]));
            int best = 0;
            for (int mode = 1; mode < 3; ++mode)
                if (t[mode] < t[best]) best = mode;
            ok = best != 0;
            rr = best == 2;
            sdft_tune_entry_t e;
            e.kind = SDFT_TUNE_PROPAGATOR; e.key = u; e.precision = prec; e.choice = best;
            for (int mode = 0; mode < 3; ++mode) e.us[mode] = t[mode];
            e.us[3] = e.us[4] = e.us[5] = -1.0;
            ctx->tuning.push_back(e);
        }
    }
    // ---- convolution tile shapes per grid-length bin and stage
    if (!c.use_gemm) return SDFT_OK;
    std::map<int, std::vector<char>> bins;  // NP bound -> points of the bin
    std::vector<int> set_bin(ctx->nsets, 0);
    for (int p = 0; p < c.P; ++p) {
        const int bound = (ctx->pts[p].NP + 255) / 256 * 256;
        auto& m = bins[bound];
        if (m.empty()) m.assign(c.P, 0);
        m[p] = 1;
        set_bin[ctx->pts[p].kset] = bound;
    }
    int base[3], njob[3];
    for (int k = 0; k < 3; ++k) stage_range(c, k, &base[k], &njob[k]);
    for (auto& bin : bins) {
        for (int k = 0; k < 3; ++k) {
            if (njob[k] == 0) continue;
            double t[CONV_NSHAPES];
            const int nslot = (int)stage_slots(ctx->jobs_h, base[k], njob[k]).size();
            bool fft_ok = true;  // FFT candidate: every kernel set of the bin takes the FFT path
            for (int s = 0; s < ctx->nsets; ++s)
                if (set_bin[s] == bin.first && !fft_stage_ok(ctx->set_L_h[s], nslot)) fft_ok = false;
            for (int sh = 0; sh < CONV_NSHAPES; ++sh) {
                if (sh == CONV_FFT && !fft_ok) {
                    t[sh] = -1.0;
                    continue;
                }
                const std::vector<int> forced(ctx->nsets, sh);
                Ctx::ConvStage tmp;
                TRY(build_conv_plan(ctx, ctx->jobs_h, base[k], njob[k], ctx->canon, forced, &bin.second,
                                    ctx->set_L_h, ctx->set_kbase_h, tmp));
                const OpPlan* pl[1] = {&tmp.sub[sh]};
                t[sh] = 0.0;
                for (int prec : {32, 64}) {
                    launch_set_evalflag(st, c.d_evalflag, prec, c.P);
                    launch_compact_plans(st, c, prec, pl, 1);
                    double us = 0.0;
                    TRY(time_launches(ctx, e0, e1, reps, [&]() {
                        if (prec == 32)
                            launch_opgemm_conv<float>(st, c, 32, tmp.sub[sh], ctx->p32.kern, ctx->p32.rhoT,
                                                      ctx->p32.dphi, ctx->p32.nw, c.d_fx);
                        else
                            launch_opgemm_conv<double>(st, c, 64, tmp.sub[sh], ctx->p64.kern, ctx->p64.rhoT,
                                                       ctx->p64.dphi, ctx->p64.nw, c.d_fx);
                    }, &us));
                    t[sh] += us;
                }
                HIP_CHECK_RET(hipStreamSynchronize(st));
                free_stage(ctx, tmp);
            }
            int best = 0;
            for (int sh = 1; sh < CONV_NSHAPES; ++sh)
                if (t[sh] >= 0.0 && t[sh] < t[best]) best = sh;
            for (int s = 0; s < ctx->nsets; ++s)
                if (set_bin[s] == bin.first) ctx->set_shape[k][s] = best;
            sdft_tune_entry_t e;
            e.kind = SDFT_TUNE_CONV_FWD + k; e.key = bin.first; e.precision = 0; e.choice = best;
            for (int sh = 0; sh < CONV_NSHAPES; ++sh) e.us[sh] = t[sh];
            ctx->tuning.push_back(e);
        }
    }
    return rebuild_conv_plans(ctx);
}

extern "C" int32_t sdft_autotune(sdft_ctx_t* ctx, int32_t reps, sdft_tune_entry_t* entries, int32_t capacity,
                                 int32_t* n_entries) {
    if (!ctx || reps < 1 || (capacity > 0 && !entries)) return SDFT_E_ARG;
    Ctx& c = ctx->c;
    if (ctx->c.nd > 1) {
        set_error(ctx, SDFT_E_UNSUPPORTED, "multi-dimensional grids have no tunable launch decisions or kernel classes");
        return SDFT_E_UNSUPPORTED;
    }
    ctx->tuning.clear();
    EvalflagGuard flags(ctx->stream, c.d_evalflag, c.P);
    HIP_CHECK_RET(flags.save());
    const int band32 = c.band_fp32, band64 = c.band_fp64;
    c.band_fp32 = 1;
    c.band_fp64 = 1;
    EventPair ev;
    int rc = ev.create() == hipSuccess ? SDFT_OK : SDFT_E_HIP;
    if (rc == SDFT_OK) rc = autotune_run(ctx, reps, ev.e0, ev.e1);
    else set_error(ctx, SDFT_E_HIP, "hipEventCreate failed");
    c.band_fp32 = band32;
    c.band_fp64 = band64;
    ctx->solved = false;
    if (rc != SDFT_OK) return rc;
    HIP_CHECK_RET(flags.restore());
    const int n = (int)ctx->tuning.size();
    for (int i = 0; i < n && i < capacity; ++i) entries[i] = ctx->tuning[i];
    if (n_entries) *n_entries = n;
    return SDFT_OK;
}

extern "C" int32_t sdft_set_tuning(sdft_ctx_t* ctx, const sdft_tune_entry_t* entries, int32_t n) {
    if (!ctx || n < 0 || (n > 0 && !entries)) return SDFT_E_ARG;
    Ctx& c = ctx->c;
    if (ctx->c.nd > 1) {
        set_error(ctx, SDFT_E_UNSUPPORTED, "multi-dimensional grids have no tunable launch decisions or kernel classes");
        return SDFT_E_UNSUPPORTED;
    }
    bool conv = false;
    for (int i = 0; i < n; ++i) {
        const sdft_tune_entry_t& e = entries[i];
        if (e.kind == SDFT_TUNE_PROPAGATOR) {
            if (e.key < 0 || e.key >= c.nsp || e.choice < 0 || e.choice > 2 ||
                (e.precision != 0 && e.precision != 32 && e.precision != 64)) {
                set_error(ctx, SDFT_E_ARG, "tuning entry %d: bad propagator key / choice / precision", i);
                return SDFT_E_ARG;
            }
            ChainPlanHost& ch = c.chain_plan[e.key];
            if (e.precision != 64) {
                ch.ok32 = ch.elig32 && e.choice != 0;
                ch.reread32 = e.choice == 2;
            }
            if (e.precision != 32) {
                ch.ok64 = ch.elig64 && e.choice != 0;
                ch.reread64 = e.choice == 2;
            }
        } else if (e.kind >= SDFT_TUNE_CONV_FWD && e.kind <= SDFT_TUNE_CONV_BACK) {
            if (e.choice < -1 || e.choice >= CONV_NSHAPES) {
                set_error(ctx, SDFT_E_ARG, "tuning entry %d: tile shape %d out of range", i, e.choice);
                return SDFT_E_ARG;
            }
            const int k = e.kind - SDFT_TUNE_CONV_FWD;
            for (int p = 0; p < c.P; ++p) {
                const int bound = (ctx->pts[p].NP + 255) / 256 * 256;
                if (e.key == 0 || e.key == bound) ctx->set_shape[k][ctx->pts[p].kset] = e.choice;
            }
            conv = true;
        } else {
            set_error(ctx, SDFT_E_ARG, "tuning entry %d: unknown kind %d", i, e.kind);
            return SDFT_E_ARG;
        }
    }
    if (conv && c.use_gemm) {
        HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
        TRY(rebuild_conv_plans(ctx));
    }
    return SDFT_OK;
}

extern "C" int32_t sdft_bulk_solve(const sdft_bulk_desc_t* d, double tol, int32_t max_newton, sdft_bulk_out_t* out) {
    if (!d || !out) {
        set_error(nullptr, SDFT_E_ARG, "null argument");
        return SDFT_E_ARG;
    }
    if (d->abi_version != SDFT_ABI_VERSION) {
        set_error(nullptr, SDFT_E_ARG, "abi_version %d != %d", d->abi_version, SDFT_ABI_VERSION);
        return SDFT_E_ARG;
    }
    const int P = d->n_points, NB = d->NB, MB = d->MB, nsp = d->n_species;
    if (P <= 0 || NB <= 0 || NB > SDFT_MAX_NB || MB < 0 || MB > SDFT_MAX_MB || nsp <= 0 || nsp > SDFT_MAX_SPECIES ||
        max_newton < 1 || !(tol > 0.0)) {
        set_error(nullptr, SDFT_E_ARG, "bad sizes P=%d NB=%d MB=%d n_species=%d max_newton=%d", P, NB, MB, nsp, max_newton);
        return SDFT_E_ARG;
    }
    if (!d->diameters || !d->valences || (MB > 0 && !d->bond_types) || !d->species_nseg || !d->species_model ||
        !d->species_nstates || !d->seg_states || !d->adjustable || !d->delta_muH || !d->base_density ||
        !d->log_activity || !d->species_charge || !out->mu_species || !out->rho_species || !out->mu_ex ||
        !out->rho_beads || !out->lng || !out->rho_bonds || !out->pressure || !out->iterations) {
        set_error(nullptr, SDFT_E_ARG, "a required table is NULL");
        return SDFT_E_ARG;
    }
    BulkSysDev s;
    memset(&s, 0, sizeof(s));
    s.NB = NB; s.MB = MB; s.nsp = nsp; s.F = d->functionals; s.lB = d->bjerrum;
    for (int j = 0; j < NB; ++j) { s.sig[j] = d->diameters[j]; s.z[j] = d->valences[j]; }
    for (int u = 0; u < MB; ++u) {
        s.bi[u] = d->bond_types[2 * u];
        s.bj[u] = d->bond_types[2 * u + 1];
        if (s.bi[u] < 0 || s.bj[u] >= NB || s.bi[u] > s.bj[u]) {
            set_error(nullptr, SDFT_E_ARG, "bond %d must satisfy 0 <= i <= j < NB", u);
            return SDFT_E_ARG;
        }
    }
    int soff = 0, ws_len = 1;
    for (int u = 0; u < nsp; ++u) {
        const int N = d->species_nseg[u], S = d->species_nstates[u];
        if (N <= 0 || S <= 0 || S > SDFT_MAX_S) {
            set_error(nullptr, SDFT_E_UNSUPPORTED, "species %d: nseg=%d states=%d (at most %d states per segment)", u, N, S,
                      SDFT_MAX_S);
            return SDFT_E_UNSUPPORTED;
        }
        s.nseg[u] = N; s.S[u] = S; s.model[u] = d->species_model[u]; s.soff[u] = soff;
        s.adjustable[u] = d->adjustable[u];
        for (int k = 0; k < N * S; ++k) {
            const int b = d->seg_states[soff + k];
            if (b >= NB || b < -1 || (k % S == 0 && b < 0)) {
                set_error(nullptr, SDFT_E_ARG, "species %d: bad state table entry %d", u, b);
                return SDFT_E_ARG;
            }
        }
        soff += N * S;
        ws_len = std::max(ws_len, (S + 1) * N);
    }
    const int n_sw = (d->functionals & SDFT_F_SW) ? d->n_sw_pairs : 0;
    if (n_sw < 0 || n_sw > SDFT_MAX_SW || (n_sw > 0 && (!d->sw_pairs || !d->sw_energy || !d->sw_volume))) {
        set_error(nullptr, SDFT_E_ARG, "square well: need 0..%d pairs with energies and volumes", SDFT_MAX_SW);
        return SDFT_E_ARG;
    }
    s.n_sw = n_sw;
    for (int k = 0; k < n_sw; ++k) {
        s.swi[k] = d->sw_pairs[2 * k]; s.swj[k] = d->sw_pairs[2 * k + 1];
        s.sweps[k] = d->sw_energy[k]; s.swV[k] = d->sw_volume[k];
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) {
        set_error(nullptr, SDFT_E_HIP, "no HIP device available");
        return SDFT_E_HIP;
    }
    // one temporary context holds the device buffers and the stream (released on every exit path)
    struct Tmp {
        sdft_ctx* ctx = new sdft_ctx();
        ~Tmp() { sdft_destroy(ctx); }
    } tmp;
    sdft_ctx* ctx = tmp.ctx;
    (void)hipGetDevice(&ctx->device);
    HIP_CHECK_RET(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    int32_t* d_states = nullptr;
    double *d_dmu = nullptr, *d_base = nullptr, *d_la = nullptr, *d_q = nullptr, *d_ws = nullptr, *d_out = nullptr;
    int32_t* d_it = nullptr;
    TRY(upload(ctx, &d_states, d->seg_states, soff));
    TRY(upload(ctx, &d_dmu, d->delta_muH, (int64_t)P * NB));
    TRY(upload(ctx, &d_base, d->base_density, (int64_t)P * nsp));
    TRY(upload(ctx, &d_la, d->log_activity, (int64_t)P * nsp));
    TRY(upload(ctx, &d_q, d->species_charge, (int64_t)P * nsp));
    TRY(dalloc(ctx, &d_ws, (int64_t)P * 64 * ws_len));
    const int64_t per = 2 * nsp + 2 * NB + 3 * MB + 1;   // doubles of output per point
    TRY(dalloc(ctx, &d_out, (int64_t)P * per));
    TRY(dalloc(ctx, &d_it, P));
    BulkOutDev o;
    o.mu_species = d_out;
    o.rho_species = o.mu_species + (int64_t)P * nsp;
    o.mu_ex = o.rho_species + (int64_t)P * nsp;
    o.rho_beads = o.mu_ex + (int64_t)P * NB;
    o.lng = o.rho_beads + (int64_t)P * NB;
    o.rho_bonds = o.lng + (int64_t)P * MB;
    o.pressure = o.rho_bonds + (int64_t)P * 2 * MB;
    o.iterations = d_it;
    launch_bulk_solve(ctx->stream, P, s, d_states, d_dmu, d_base, d_la, d_q, tol, max_newton, d_ws, ws_len, o);
    HIP_CHECK_RET(hipGetLastError());
    auto down = [&](double* host, const double* dev, int64_t n) -> hipError_t {
        return hipMemcpyAsync(host, dev, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    };
    HIP_CHECK_RET(down(out->mu_species, o.mu_species, (int64_t)P * nsp));
    HIP_CHECK_RET(down(out->rho_species, o.rho_species, (int64_t)P * nsp));
    HIP_CHECK_RET(down(out->mu_ex, o.mu_ex, (int64_t)P * NB));
    HIP_CHECK_RET(down(out->rho_beads, o.rho_beads, (int64_t)P * NB));
    HIP_CHECK_RET(down(out->lng, o.lng, (int64_t)P * MB));
    HIP_CHECK_RET(down(out->rho_bonds, o.rho_bonds, (int64_t)P * 2 * MB));
    HIP_CHECK_RET(down(out->pressure, o.pressure, P));
    HIP_CHECK_RET(hipMemcpyAsync(out->iterations, d_it, (size_t)P * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK_RET(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < P; ++p)
        if (out->iterations[p] < 0) {
            set_error(nullptr, SDFT_E_NOCONV, "bulk state of point %d: %s", p,
                      out->iterations[p] == -2 ? "no compensating species can absorb the net charge"
                                               : "Newton did not reach the tolerance");
            return SDFT_E_NOCONV;
        }
    return SDFT_OK;
}

extern "C" int32_t sdft_kernel_profile(sdft_ctx_t* ctx, int32_t kc, int32_t reps, double* avg_ms, double* flops) {
    if (!ctx || reps < 1) return SDFT_E_ARG;
    if (ctx->c.nd > 1) {
        set_error(ctx, SDFT_E_UNSUPPORTED, "multi-dimensional grids have no tunable launch decisions or kernel classes");
        return SDFT_E_UNSUPPORTED;
    }
    if (!ctx->have_eval) {
        set_error(ctx, SDFT_E_STATE, "kernel profiling needs a prior sdft_solve or sdft_objective");
        return SDFT_E_STATE;
    }
    Ctx& c = ctx->c;
    if (kc < SDFT_KP_PROP_FP32 || kc > SDFT_KP_ACCUM_FP64) {
        set_error(ctx, SDFT_E_ARG, "unknown kernel class %d", kc);
        return SDFT_E_ARG;
    }
    const bool chain = kc == SDFT_KP_CHAIN_FP32 || kc == SDFT_KP_CHAIN_FP64;
    const int prec = (kc == SDFT_KP_PROP_FP32 || kc == SDFT_KP_CONV_FWD_FP32 || kc == SDFT_KP_CONV_BACK_FP32 ||
                      kc == SDFT_KP_CHAIN_FP32 || kc == SDFT_KP_ACCUM_FP32) ? 32 : 64;
    // pick the longest chain species for the propagator classes
    int u = 0;
    for (int k = 0; k < c.nsp; ++k)
        if (c.species[k].nseg > c.species[u].nseg) u = k;
    if ((kc == SDFT_KP_PROP_FP32 || kc == SDFT_KP_PROP_FP64 || chain) && c.species[u].nseg < 2) {
        set_error(ctx, SDFT_E_ARG, "no chain species to profile");
        return SDFT_E_ARG;
    }
    if (chain && !(prec == 32 ? c.chain_plan[u].ok32 : c.chain_plan[u].ok64)) {
        set_error(ctx, SDFT_E_ARG, "species %d does not run on the banded chain kernel", u);
        return SDFT_E_ARG;
    }
    // scratch output so the solution buffers are untouched (allocated on the first call, kept)
    if (!ctx->d_profile_fx) TRY(dalloc(ctx, &ctx->d_profile_fx, c.total_n));
    double* d_fx = ctx->d_profile_fx;
    EvalflagGuard flags(ctx->stream, c.d_evalflag, c.P);
    HIP_CHECK_RET(flags.save());
    launch_set_evalflag(ctx->stream, c.d_evalflag, prec, c.P);
    if (c.use_gemm) launch_compact_cols(ctx->stream, c, prec);  // every point active in the profiled launches
    const int s_mid = c.species[u].nseg / 2 > 0 ? c.species[u].nseg / 2 : 1;
    auto once = [&]() {
        switch (kc) {
            case SDFT_KP_PROP_FP32: ctx->p32.launch_prop_step(ctx->stream, c, 32, u, s_mid, c.band_fp32 != 0); break;
            case SDFT_KP_PROP_FP64: ctx->p64.launch_prop_step(ctx->stream, c, 64, u, s_mid, c.band_fp64 != 0); break;
            case SDFT_KP_CONV_FWD_FP32: ctx->p32.launch_conv(ctx->stream, c, 32, false, d_fx); break;
            case SDFT_KP_CONV_FWD_FP64: ctx->p64.launch_conv(ctx->stream, c, 64, false, d_fx); break;
            case SDFT_KP_CONV_BACK_FP32: ctx->p32.launch_conv(ctx->stream, c, 32, true, d_fx); break;
            case SDFT_KP_CONV_BACK_FP64: ctx->p64.launch_conv(ctx->stream, c, 64, true, d_fx); break;
            case SDFT_KP_CHAIN_FP32:
                launch_chain<float>(ctx->stream, c, 32, u, ctx->p32.kern, ctx->p32.elam, ctx->p32.hl, ctx->p32.trapez,
                                    ctx->p32.gC, ctx->p32.gP, ctx->p32.smax);
                break;
            case SDFT_KP_CHAIN_FP64:
                launch_chain<double>(ctx->stream, c, 64, u, ctx->p64.kern, ctx->p64.elam, ctx->p64.hl,
                                     ctx->p64.trapez, ctx->p64.gC, ctx->p64.gP, ctx->p64.smax);
                break;
            case SDFT_KP_ACCUM_FP32: ctx->p32.launch_accumulate(ctx->stream, c, 32); break;
            case SDFT_KP_ACCUM_FP64: ctx->p64.launch_accumulate(ctx->stream, c, 64); break;
            default: ctx->p64.launch_pointwise(ctx->stream, c, 64); break;
        }
    };
    once();  // warm
    EventPair ev;
    HIP_CHECK_RET(ev.create());
    HIP_CHECK_RET(hipEventRecord(ev.e0, ctx->stream));
    for (int r = 0; r < reps; ++r) once();
    HIP_CHECK_RET(hipEventRecord(ev.e1, ctx->stream));
    HIP_CHECK_RET(hipEventSynchronize(ev.e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
    HIP_CHECK_RET(flags.restore());
    HIP_CHECK_RET(hipGetLastError());
    if (avg_ms) *avg_ms = ms / reps;
    if (flops) {
        // algorithmic flops: 2 per multiply-add of the operator applied
        std::vector<int32_t> band(ctx->nsets);
        HIP_CHECK_RET(hipMemcpy(band.data(), c.d_band, ctx->nsets * sizeof(int32_t), hipMemcpyDeviceToHost));
        double f = 0.0;
        const int S = c.species[u].S;
        int npair = 0;
        {
            // valid (a, b) pairs at the profiled step: states of the two segments involved
            const int N = c.species[u].nseg;
            const int so = c.species[u].state_off;
            auto cnt = [&](int seg) { int k = 0; for (int a = 0; a < S; ++a) if (ctx->seg_states_h[so + seg * S + a] >= 0) ++k; return k; };
            npair = cnt(N - 1 - s_mid) * cnt(N - s_mid);  // child sweep; parent sweep similar
            (void)N;
        }
        // chain kernel: valid (a, b) columns summed over all steps of both sweeps, 2 B + 1 taps of the
        // banded operator (B: the bond kernel's band; the kernel's own padding to Be is not counted)
        double chain_cols = 0.0;
        if (chain) {
            const int N = c.species[u].nseg, S = c.species[u].S, so = c.species[u].state_off;
            auto cnt = [&](int seg) { int k = 0; for (int a = 0; a < S; ++a) if (ctx->seg_states_h[so + seg * S + a] >= 0) ++k; return k; };
            for (int s = 1; s < N; ++s) chain_cols += cnt(N - 1 - s) * cnt(N - s) + cnt(s) * cnt(s - 1);
        }
        for (int p = 0; p < c.P; ++p) {
            const double NP = ctx->pts[p].NP;
            if (chain) {
                f += chain_cols * 2.0 * NP * (2.0 * band[ctx->pts[p].kset] + 1.0);
            } else if (kc == SDFT_KP_PROP_FP32 || kc == SDFT_KP_PROP_FP64) {
                const bool bd = kc == SDFT_KP_PROP_FP32 ? c.band_fp32 != 0 : c.band_fp64 != 0;
                const double taps = bd ? std::min(NP, 2.0 * band[ctx->pts[p].kset] + 1.0) : NP;
                f += 2.0 /*sweeps*/ * npair * 2.0 * NP * taps;
            } else if (kc == SDFT_KP_POINTWISE_FP64 || kc == SDFT_KP_ACCUM_FP32 || kc == SDFT_KP_ACCUM_FP64) {
                f += 0.0;  // not counted in flops (memory-bound classes)
            } else {
                const bool back = kc == SDFT_KP_CONV_BACK_FP32 || kc == SDFT_KP_CONV_BACK_FP64;
                f += (back ? ctx->back_terms : ctx->fwd_terms) * 2.0 * NP * NP;
            }
        }
        *flops = f;
    }
    return SDFT_OK;
}

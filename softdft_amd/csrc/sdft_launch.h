// Host-side context and per-precision pipeline declarations shared by the
// kernel and API translation units.
#pragma once
#include <vector>
#include <string>
#include <algorithm>

#include "../../include/softdft_hip.h"
#include "sdft_internal.h"

namespace sdft {

// tile / group / column tables of the MFMA operator-GEMM (sdft_opgemm.hip)
struct OpTile {
    int32_t group;
    int32_t row0;
    int32_t col0;
    int32_t ncols;
};
struct OpGroup {
    int32_t kset;
    int32_t NP, L;
    int32_t nterms;
    int32_t kern[3];
    int32_t sign[3];
    int32_t col_off;
    int32_t ncols;
    int64_t kbase;
    int32_t band;
    int32_t pad;
};
struct OpCol { int32_t p, job; };
struct PropCol { int32_t p, sweep, a, b; };

// tile shapes of the convolution GEMMs: rows x columns per plan tile
enum ConvShape : int {
    CONV_TALL = 0,     // 128 x 16 (k_opgemm): few columns per operator (ragged batches)
    CONV_WIDE64 = 1,   // 64 x 64 (k_convgemm)
    CONV_WIDE128 = 2,  // 128 x 64 (k_convgemm; FP64 runs it as two 64-row workgroups per tile)
    CONV_MID = 3,      // 64 x 32 (k_convgemm): twice the tiles of 64 x 64, fewer registers
    CONV_NARROW = 4,   // 64 x 16 (k_convgemm)
    CONV_FFT = 5,      // batched FFT convolution (sdft_fft.hip): power-of-two mirror lengths 4096..65536
    CONV_NSHAPES = 6
};

// FFT convolution of one kernel set in one stage (sdft_fft.hip). The mirror length L = 2 NP is
// split as L = L1 x L2 (L2 = 256); spectra are kept as H = L2/2 + 1 rows of L1 values (the other
// rows follow from the Hermitian symmetry of real signals).
struct FftSlot { int32_t infam, infield; };   // one distinct input field of the stage's jobs
struct FftJob {                               // one output column of every point of the set
    int32_t job;                              // index into the job table (ConvJob: output semantics)
    int32_t nterms;
    int32_t kern[3];                          // canonical kernel index within the set
    int32_t sign[3];                          // sign of the Toeplitz part of each term's operator
    int32_t slot[3];                          // input slot of each term
};
struct FftSetPlan {
    int kset = 0, L = 0, L1 = 0, L2 = 0, lg1 = 0, lg2 = 0, H = 0, R = 0;
    int npts = 0, nslot = 0, njob = 0;
    int32_t* d_pts = nullptr;                 // [npts] points of the set
    FftSlot* d_slots = nullptr;               // [nslot]
    FftJob* d_jobs = nullptr;                 // [njob]
};
// per kernel set: the FFT of every kernel vector ([K][H][L1] complex) and the twiddles
// W[m] = exp(-2 pi i m / L), both precisions (prepared when a plan first uses the set)
struct FftRes {
    int L = 0;
    void* chat64 = nullptr;
    void* chat32 = nullptr;
    void* W64 = nullptr;
    void* W32 = nullptr;
};

struct OpPlan {
    int ntiles = 0;
    bool wide = true;          // propagator plans: 64 x 64 tiles (else 128 x 16)
    int shape = CONV_WIDE64;   // convolution plans
    bool full_pairs = false;   // propagator plans: every (sweep, a, b) column present, b fastest
    void* d_tiles = nullptr;
    void* d_groups = nullptr;
    void* d_cols = nullptr;
    int ngroups = 0;
    // convolution plans, per precision (0: FP32, 1: FP64): the columns of each group whose point
    // evaluates at that precision, compacted in order, and their number per group
    void* d_cols_act[2] = {nullptr, nullptr};
    int32_t* d_count[2] = {nullptr, nullptr};
    std::vector<FftSetPlan> fft;   // shape CONV_FFT: one entry per kernel set (no tiles or groups)
};

// applied half-width of the banded bond operator in the chain and banded step kernels: the band B
// rounded up to even, so that K = 2 Be + 16 taps per 16 outputs are a whole number of 4-tap MFMA steps
__host__ __device__ __forceinline__ int chain_half_width(int B) { return B + (B & 1); }

// one (a, b) column of one step of the chain kernel (a: state of the output segment, b: state of the
// input segment): valid pair, input bead, field of exp(lng/2) at the input and at the output end
struct ChainCol {
    int32_t valid, sj, hl_in, hl_out;
};

// persistent banded chain kernel of one species (ok: eligible for the banded operator)
struct ChainPlanHost {
    bool ok = false;                               // eligible in at least one precision
    bool ok32 = false, ok64 = false;               // runs at that precision (eligible and chosen)
    bool elig32 = false, elig64 = false;           // the working set fits LDS at that precision
    bool reread32 = false, reread64 = false;       // input weights re-read every step instead of
                                                   // held in registers (fewer registers per lane)
    bool homog = false;                            // every segment carries the same state beads
    int S = 1;
    int ck[2][SDFT_FAST_S * SDFT_FAST_S];          // canonical bond kernel of (sweep, a * S + b), -1 absent
    int npad_max = 0, be_max = 0;                  // largest padded NP and even band half-width
    int xl_need = 0;                               // largest NP + 2 Be (extended positions read)
    ChainCol* d_info = nullptr;                    // [2 sweeps][N-1 steps][S*S] column table (device)
    bool hv32 = false, hv64 = false;               // homogeneous chain: the output factors
                                                   // exp(lng/2)/normative tabulated in LDS
    bool pp32 = false, pp64 = false;               // both sweeps in one 16-wave workgroup (k_chain_pp)
    int pp_min = 0;                                // active points from which k_chain_pp pairs the sweeps
    bool shared_a = false;                         // every valid column of both sweeps: one bond operator
};

// multi-dimensional grids (sdft_nd.hip): one convolution column of the spectral path
struct NdTerm { int32_t kern, infield; };   // spectral kernel index, input field of the stage's family
struct NdJob {
    int32_t nterms;
    NdTerm t[4];
    int32_t odd_axis;   // the output is odd along this axis (vector weighted density), -1: even
    ConvJob out;        // output semantics (destination, |.|, scale, shared partials)
};
// every point of a multi-dimensional context shares one grid NP_1 x .. x NP_D (first fastest)
struct NdPlan {
    int D = 1;
    int n[3] = {1, 1, 1};
    double h[3] = {1.0, 1.0, 1.0};
    int64_t NPt = 0;
    int K = 0;                     // spectral kernels: (2 + D) per bead (w2, w3, w_V per axis), bonds, Coulomb
    int kbond0 = 0;                // first bond kernel
    double* kspec64 = nullptr;     // [K][NPt] weights on the mode grid
    float* kspec32 = nullptr;
    double* mat64[3][4] = {};      // per axis: forward even / odd, inverse even / odd (transposed n x n)
    float* mat32[3][4] = {};
    NdJob* d_fwd = nullptr;        // forward jobs (vector jobs expanded per axis)
    NdJob* d_back = nullptr;       // shared partials, lng and Psi columns, then the projections
    int nfwd = 0, npre = 0, nback = 0;
    int8_t* d_par_fwdout = nullptr;  // per forward job: odd axis of its output
    int8_t* d_par_dphi = nullptr;    // per dphi field: odd axis (vector block of axis v), -1 even
    int8_t* d_par_even = nullptr;    // all -1
    int maxF = 0;                  // most fields transformed at once per point
};

struct SpeciesHost {
    int nseg, S, model, fixed_seg, multistate, rescale;
    int state_off, maxd_off;
    int64_t goff;   // field offset in the propagator families
    int band_step = 0;   // banded steps run k_bandstep (every point's band fits its grid)
    int band_be = 0;     // largest applied half-width over the batch's grids (k_bandstep LDS)
};

// Device pointers and sizes shared by both precisions.
struct Ctx {
    int P = 0, NB = 0, MB = 0, nsp = 0, K = 0;
    int ncls = 1;                   // distinct bead diameters
    int maxNP = 0;
    int64_t sumNP = 0, sumL_sets = 0;
    int64_t sumGS = 0;              // sum over points of the propagator column stride (NP rounded up to 32)
    int64_t total_n = 0, max_n = 0;
    int GF = 0;                     // propagator fields per point (sum over species of nseg*S*S)
    int Frt = 0, Fnw = 0, Fdp = 0;  // family widths
    int n_fwd_jobs = 0, n_pre_jobs = 0, n_back_jobs = 0;  // job table: forward ++ shared partials ++ back
    int n_jobs = 0;                 // length of that table; d_jobs holds it twice, the second copy with
                                    // the back projections' partial additions removed (merged stage)
    bool merge_back = false;        // partials and back projections in one GEMM launch, the partials
                                    // added to F(x) afterwards by k_add_partials
    int Fws = 0;                    // shared back-projection partials per point (diameter classes)
    int band_fp32 = 1;
    int m = 5;                      // Anderson depth the history buffers were sized for
    int band_fp64 = 0;              // FP64 propagators banded (mixed-precision polish only)
    bool use_gemm = true;           // MFMA operator-GEMM path (else the direct reference kernels)
    // convolution stages (forward, shared back-projection partials, back projections), each split
    // by tile shape: the groups of a kernel set (one grid length) all run in that set's shape
    struct ConvStage { OpPlan sub[CONV_NSHAPES]; };
    ConvStage fwd_plan, pre_plan, back_plan;
    std::vector<OpPlan> prop_plan;  // per species (ntiles == 0: direct kernel)
    std::vector<ChainPlanHost> chain_plan;  // per species
    std::vector<SpeciesHost> species;

    DevSystem* d_sys = nullptr;
    DevPoint* d_pts = nullptr;
    DevSpecies* d_spec = nullptr;
    int64_t* d_goffs = nullptr;
    int32_t* d_seg_states = nullptr;
    std::vector<int32_t> h_seg_states;            // host copies of the state and bond tables
    std::vector<int32_t> h_bond_i, h_bond_j;
    double* d_maxd = nullptr;
    ConvJob* d_jobs = nullptr;
    double* d_wsum = nullptr;       // (sumNP * Fws) shared back-projection partials
    int32_t* d_band = nullptr;      // banded half-width of the bond kernels per kernel set
    double* d_Ext = nullptr;        // (sumNP * NB)
    double* d_surf = nullptr;       // (sumNP)
    int32_t* d_evalflag = nullptr;  // per point: 0 / 32 / 64 = evaluate in that precision
    int32_t* d_errflag = nullptr;
    double* d_acc_species = nullptr;   // [P][2*MAX_SPECIES] totals (written by k_neutrality)
    double* d_acc_part = nullptr;      // [P][MAX_TILES][2*MAX_SPECIES] per-tile partial sums
    double* d_fixed_lambda = nullptr;  // [P][MAX_SPECIES]
    // solve
    double* d_x = nullptr;
    double* d_fx = nullptr;
    double* d_Xh = nullptr;
    double* d_Gh = nullptr;
    double* d_gram = nullptr;
    unsigned long long* d_errbits = nullptr;
    DevSolveState* d_state = nullptr;
    int32_t* d_n_active = nullptr;  // [active, active at FP32, waiting for the FP64 switch]
    int nmax_seg = 1;               // longest species (propagator scale tables)
    int states_max = 1;             // most states per segment over the species
    bool any_rescale = false;       // some species propagates with per-step rescaling
    // FFT convolutions: per kernel set resources, and the spectra work buffers per precision
    // (0: FP32, 1: FP64; complex elements: inputs and outputs of the largest FFT stage)
    int nd = 1;                     // grid dimensions (2 or 3: the spectral path of sdft_nd.hip)
    NdPlan ndp;
    std::vector<FftRes> fft_res;
    void* fft_sa[2] = {nullptr, nullptr};
    void* fft_sb[2] = {nullptr, nullptr};
    int64_t fft_sa_n = 0, fft_sb_n = 0;
};

// Per-step propagator scale factors (species with DevSpecies::rescale): smax holds, per
// (point, species, sweep, step), the largest output value of that step as the bits of a
// non-negative double (atomicMax on the bits orders them like the values); the next step divides its
// input by it. lsc holds per (point, species, segment q) E(q), the summed logarithms of the factors
// divided out of gC[q] and gP[q], and at q = nmax the value E(q*) of the grafted normalisation
// segment: the stored product Gp(q) Qc(q) times exp(E(q)) is the unscaled one.
__device__ __forceinline__ int64_t scale_slot(int p, int u, int sweep, int s, int nsp, int nmax) {
    return (((int64_t)p * nsp + u) * 2 + sweep) * nmax + s;
}
__device__ __forceinline__ double scale_read(const unsigned long long* smax, int64_t i) {
    const double m = __longlong_as_double((long long)smax[i]);
    return (m > 0.0 && m <= 1.7976931348623157e308) ? m : 1.0;
}
__device__ __forceinline__ void scale_note(unsigned long long* smax, int64_t i, double v) {
    if (v > 0.0) atomicMax(smax + i, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ const double* scale_row(const double* lsc, int p, int u, int nsp, int nmax) {
    return lsc + ((int64_t)p * nsp + u) * (nmax + 1);
}

// arguments of one propagator step launch (species, step, layout of its propagator fields)
struct PropArgs {
    int u;            // species
    int s;            // step 1..N-1
    int nseg, S;
    int fixed_seg;
    int state_off, maxd_off;
    int64_t goff;     // field offset of this species in the gC / gP families
    int GF;           // fields per point of the gC / gP families
    int band;         // >= 0: banded Toeplitz window (FP32 phase), -1: dense
    int rescale;      // per-step rescaling (scale_slot tables)
    int nsp, nmax;
    unsigned long long* smax;
};

// squared distance between grid points n and Kf (compute_distance_squared, child_propagators.jl:165-182):
// one-dimensional grids (n - Kf) dz; multi-dimensional grids from the flat indices (first dimension
// fastest) with the bin width of each dimension
__device__ __forceinline__ double graft_dist2(const DevPoint& P, int n, int Kf) {
    if (P.nd <= 1) {
        const double d = (double)(n - Kf) * P.dz;
        return d * d;
    }
    double s = 0.0;
    int a = n, b = Kf;
    for (int v = 0; v < P.nd; ++v) {
        const int ia = a % P.gn[v], ib = b % P.gn[v];
        a /= P.gn[v];
        b /= P.gn[v];
        const double d = (double)(ia - ib) * P.gh[v];
        s += d * d;
    }
    return s;
}

// propagator input at an end of the recursion: 1 at a free chain end; for a grafted chain 1 within
// reach (max_distance) of the grafting point and 0 beyond, and at the grafted segment itself only at
// the grafting point (handle_fixed.jl)
template <typename T>
__device__ __forceinline__ T base_value(int seg, int n, int fixed_seg, int Kf, const DevPoint& P, const double* maxd) {
    if (fixed_seg < 0) return T(1);
    if (seg == fixed_seg) return n == Kf ? T(1) : T(0);
    return (graft_dist2(P, n, Kf) > maxd[seg] * maxd[seg]) ? T(0) : T(1);
}

// stores one convolution column's value y at grid point n of point P (ConvJob semantics)
template <typename T>
__device__ __forceinline__ void conv_store(const ConvJob& J, const DevPoint& P, int n, double y, T* nw, int Fnw,
                                           double* fx, double* wsum, int Fws) {
    y *= J.scale;
    if (J.absval) y = fabs(y);
    if (J.outkind == 0) {
        nw[P.base * (int64_t)Fnw + (int64_t)J.outfield * P.NP + n] = (T)y;
    } else if (J.outkind == 2) {
        wsum[P.base * (int64_t)Fws + (int64_t)J.outfield * P.NP + n] = y;
    } else {
        if (J.addfield >= 0) y += wsum[P.base * (int64_t)Fws + (int64_t)J.addfield * P.NP + n];
        if (J.addfield2 >= 0) y += wsum[P.base * (int64_t)Fws + (int64_t)J.addfield2 * P.NP + n];
        fx[P.xoff + (int64_t)J.outfield * P.NP + n] = y;
        if (J.outfield2 >= 0) fx[P.xoff + (int64_t)J.outfield2 * P.NP + n] = y;
    }
}

template <typename T>
struct Pipeline {
    T* kern = nullptr;      // convolution kernel vectors
    T* trapez = nullptr;    // (sumNP * NB)
    T* elam = nullptr;
    T* hl = nullptr;
    T* gC = nullptr;
    T* gP = nullptr;
    T* rho_beads = nullptr;
    T* rho_bonds = nullptr;
    T* rhoT = nullptr;
    T* nw = nullptr;
    T* dphi = nullptr;
    T* fbuf = nullptr;      // chained propagator inputs, two step buffers of (2 sweeps x S x S) fields
    unsigned long long* smax = nullptr;  // per-step scale factors (see scale_slot)
    double* lsc = nullptr;               // summed log scales per segment (see scale_row)
    T* nd_a = nullptr;                   // multi-dimensional grids: transform work buffers,
    T* nd_b = nullptr;                   // [P][maxF][NPt] each
    void objective(hipStream_t st, const Ctx& c, int prec, const double* x, double* fx);
    void omega(hipStream_t st, const Ctx& c, const double* x, double* omega_acc);
    void segment_sums(hipStream_t st, const Ctx& c, int u, double* out, int64_t stride);
    void segment_profile(hipStream_t st, const Ctx& c, int u, int p, double* out);
    // single-launch entry points used by sdft_kernel_profile
    void launch_prop_step(hipStream_t st, const Ctx& c, int prec, int u, int s, bool banded);
    void propagate(hipStream_t st, const Ctx& c, int prec, int u, bool allow_chain);
    void launch_conv(hipStream_t st, const Ctx& c, int prec, bool back, double* fx);
    void launch_pointwise(hipStream_t st, const Ctx& c, int prec);
    void launch_accumulate(hipStream_t st, const Ctx& c, int prec);
    // multi-dimensional grids (sdft_nd.hip)
    void propagate_nd(hipStream_t st, const Ctx& c, int prec, int u);
    void conv_nd(hipStream_t st, const Ctx& c, int prec, bool back, double* fx);
};

// multi-dimensional setup (sdft_nd.hip): spectral weights, axis matrices, job tables from the
// one-dimensional ones; device buffers are recorded in allocs (freed with the context)
int nd_setup(Ctx& c, const DevSystem& sys, const std::vector<ConvJob>& fwd1, const std::vector<ConvJob>& pre1,
             const std::vector<ConvJob>& back1, std::vector<void*>& allocs, std::string& err);
hipError_t nd_prepare_attributes(const Ctx& c);
size_t nd_axis_lds_bytes(int n, size_t elem);
void launch_scale_prefix(hipStream_t st, const Ctx& c, int prec, int u, unsigned long long* smax, double* lsc);
// grafted chain normalisation lambda (handle_fixed.jl:41-119) of species A.u into d_fixed_lambda
template <typename T>
void launch_fixed_lambda(hipStream_t st, const Ctx& c, int prec, const PropArgs& A, const T* elam, const T* trapez,
                         const T* gC, const T* gP);

void launch_compact_cols(hipStream_t st, const Ctx& c, int prec);
// FFT convolutions (sdft_fft.hip)
bool fft_length_ok(int L);
size_t fft_mid_lds_bytes(int nslot, int L1, size_t elem);
hipError_t fft_prepare_attributes();
void fft_split(int L, int* L1, int* L2);
int64_t fft_spectrum_len(int L);   // complex elements of one stored spectrum (H x L1)
void launch_fft_twiddles(hipStream_t st, int L, void* W64, void* W32);
void launch_fft_kernel_spectra(hipStream_t st, int L, int K, const double* c64, const void* W64, void* tmp,
                               void* chat64, void* chat32);
template <typename T>
void launch_fft_conv(hipStream_t st, const Ctx& c, int prec, const OpPlan& plan, const T* rhoT, const T* dphi,
                     T* nw, double* fx);
template <typename T>
void launch_opgemm_conv(hipStream_t st, const Ctx& c, int prec, const OpPlan& plan, const T* kern, const T* rhoT,
                        const T* dphi, T* nw, double* fx);
size_t chain_lds_bytes(const ChainPlanHost& ch, size_t elem, bool hv_table);
size_t chain_pp_lds_bytes(const ChainPlanHost& ch, size_t elem);
hipError_t chain_prepare(const ChainPlanHost& ch);
template <typename T>
void launch_chain(hipStream_t st, const Ctx& c, int prec, int u, const T* kern, const T* elam, const T* hl,
                  const T* trapez, T* gC, T* gP, unsigned long long* smax);
template <typename T>
void launch_bandstep(hipStream_t st, const Ctx& c, int prec, int u, int s, int skip_sweep, const T* kern,
                     const T* elam, const T* hl, const T* trapez, T* gC, T* gP, unsigned long long* smax);
size_t bandstep_lds_bytes(int S, int be, size_t elem);
template <typename T>
void launch_opgemm_prop(hipStream_t st, const Ctx& c, int prec, const OpPlan& plan, int u, int s, int skip_sweep,
                        bool banded, const T* kern, const T* elam, const T* hl, const T* trapez, T* gC, T* gP,
                        T* fbuf, unsigned long long* smax);
int opgemm_tile_rows(bool wide);
int opgemm_tile_cols(bool wide);
int conv_tile_rows(int shape);   // rows / columns per tile of the convolution GEMMs
int conv_tile_cols(int shape);
// compacts the active columns of the given convolution plans for one precision
void launch_compact_plans(hipStream_t st, const Ctx& c, int prec, const OpPlan* const* plans, int nplans);

void launch_build_kernels(hipStream_t st, const DevSystem* sys, int K, int nsets, int maxL, const int32_t* set_L,
                          const int64_t* set_kbase, const double* set_dz, double* wbuf, double* c64, float* c32);
void launch_anderson(hipStream_t st, const Ctx& c, const SolveParams& sp);
hipError_t prepare_kernel_attributes(const Ctx& c);
void launch_apply_switch(hipStream_t st, const Ctx& c);
// bulk state solve (sdft_bulk.hip): system layout shared by the batch, passed by value
struct BulkSysDev {
    int32_t NB, MB, nsp, F, n_sw;
    double lB;
    double sig[SDFT_MAX_NB], z[SDFT_MAX_NB];
    int32_t bi[SDFT_MAX_MB], bj[SDFT_MAX_MB];
    int32_t nseg[SDFT_MAX_SPECIES], S[SDFT_MAX_SPECIES], model[SDFT_MAX_SPECIES], soff[SDFT_MAX_SPECIES];
    int32_t adjustable[SDFT_MAX_SPECIES];
    int32_t swi[SDFT_MAX_SW], swj[SDFT_MAX_SW];
    double sweps[SDFT_MAX_SW], swV[SDFT_MAX_SW];
};
struct BulkOutDev {
    double *mu_species, *rho_species, *mu_ex, *rho_beads, *lng, *rho_bonds, *pressure;
    int32_t* iterations;
};
void launch_bulk_solve(hipStream_t st, int P, const BulkSysDev& s, const int32_t* seg_states, const double* dmu,
                       const double* base, const double* logact, const double* qch, double tol, int max_newton,
                       double* ws, int64_t ws_len, const BulkOutDev& out);

// per-species offsets (in doubles) of the segment-sum blocks inside an observables work buffer
struct ObsLayout {
    int64_t ss_base[SDFT_MAX_SPECIES];
};
void launch_obs_finalize(hipStream_t st, const Ctx& c, const ObsLayout& lay, int Smax, const double* om_blocks,
                         const double* segsums, double* omega, double* alpha);
void launch_cvt_gather(hipStream_t st, const Ctx& c, const int32_t* off, const int32_t* idx, const int32_t* side,
                       const double* fac, const float* rb32, const double* rb64, double* out);
void launch_set_evalflag(hipStream_t st, int32_t* evalflag, int value, int P);
void launch_zero_i32(hipStream_t st, int32_t* a, int64_t n);
void launch_zero_f64(hipStream_t st, double* a, int64_t n);
void launch_convert_f32(hipStream_t st, const double* a, float* b, int64_t n);

}  // namespace sdft

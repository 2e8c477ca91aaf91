// Internal device/host structures of libsoftdft_hip (gfx950).
//
// Memory layout in HBM (all per-point data contiguous per point):
//   grid family with F fields:   fam[base_p * F + f * NP_p + n]
//     base_p = sum_{q<p} NP_q,  n in [0, NP_p)
//   propagator families gC, gP (GF fields per point): g[gbase_p + f * gstride_p + n]
//     gstride_p = NP_p rounded up to 32 (every column starts on a 128-byte line in FP32, 256 in FP64),
//     gbase_p = GF * sum_{q<p} gstride_q
//   convolution kernels:          kern[kbase_s * K + k * L_s + d]
//     one set per distinct (NP, dz); L_s = 2 NP_s; K = 3 NB + MB + 1 + 2 n_sw
//   solution vectors (FP64):      x[xoff_p + i], i in [0, n_p), reference flatten order
//   Anderson history (FP64):      hist[hoff_p + col * n_p + i]  (hoff_p = m * xoff_p)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SDFT_MAX_NB 16
#define SDFT_MAX_MB 16
#define SDFT_MAX_S 4            // states per segment (state families, state_families.jl)
#define SDFT_FAST_S 2           // the MFMA chain, banded-step and operator-GEMM propagators and the
                                // two-state density accumulation cover species up to this many states
#define SDFT_MAX_SPECIES 16
#define SDFT_MAX_SW 16          // square-well pairs
#define SDFT_SW_TERMS 3         // square-well partners per bead type (terms of one convolution job)
#define SDFT_MAX_M 8
#define SDFT_ACC_POINTS 64          // grid points per density-accumulation block
#define SDFT_MAX_TILES 2048         // accumulation blocks per grid (NP <= 2048 * 64 = 131072)
#define SDFT_GRAM_BLOCKS 64         // most blocks per point for the Anderson reductions
#define SDFT_GRAM_CHUNK 2048        // solution entries per reduction block
#define SDFT_GRAM_LEN 44            // SDFT_MAX_M*(SDFT_MAX_M+1)/2 + SDFT_MAX_M
#define SDFT_OMEGA_BLOCKS 8

// kernel-vector index inside a kernel set
//   3*j + 0 : w2 (delta shell, R = sigma_j / 2)        even
//   3*j + 1 : w3 (step, R = sigma_j / 2)               even
//   3*j + 2 : wV (vector delta, i * w applied)          odd
//   3*NB + u: bond u, delta(sigma_ij)/(4 pi sigma_ij^2) x spectral filter   even
//   3*NB+MB : Coulomb 4 pi lB / lambda_FD(k)            even
//   3*NB+MB+1+2u   : square-well pair u, eps_u (step(lambda_u sigma_ij) - step(sigma_ij)),
//                    averaged over d and d-1 (the real part of the k-space product)
//   3*NB+MB+1+2u+1 : the same kernel unshifted (free energy)
struct DevPoint {
    int32_t NP;
    int32_t L;
    int32_t kset;        // index of the kernel set
    int32_t gstride;     // column stride of the propagator families: NP rounded up to 32 (aligned rows)
    int64_t base;        // sum of NP over previous points
    int64_t gbase;       // start of this point's propagator fields: GF * sum of gstride over previous points
    int64_t kbase;       // sum of L over previous kernel sets (of this point's set)
    int64_t xoff;        // offset of this point's solution vector
    int64_t n;           // solution length
    double dz;
    double total_charge;
    int32_t fixed_coord[SDFT_MAX_SPECIES];
    double mu_species[SDFT_MAX_SPECIES];
    double delta_muH[SDFT_MAX_NB];
    int32_t wall_lo[SDFT_MAX_NB];
    int32_t wall_hi[SDFT_MAX_NB];
    double wall_q[2];
    // per species: normative = prod_seg (sum_states exp(-delta_muH))^(1/N) (child_propagators.jl:41-50)
    // and the free-chain normalisation normative^(N-1) (handle_fixed.jl:105-107)
    double normative[SDFT_MAX_SPECIES];
    double chain_norm[SDFT_MAX_SPECIES];
    // multi-dimensional grids: dimensions, points and bin width per dimension (first fastest); nd <= 1
    // is the one-dimensional grid of NP points and width dz
    int32_t nd;
    int32_t gn[3];
    double gh[3];
    // grafting density of each grafted species (handle_fixed.jl:31: norm = density / lambda); per point
    // so that a batch may hold several grafting densities (a continuation in the grafting density)
    double graft_density[SDFT_MAX_SPECIES];
};

struct DevSpecies {
    int32_t nseg;
    int32_t S;            // max states per segment
    int32_t model;        // 0 freely jointed, 1 monomer bead
    int32_t fixed_seg;    // -1 or grafted segment (only 0 supported)
    int32_t multistate;   // any segment with > 1 state
    int32_t state_off;    // offset into the seg_states table ([nseg][S])
    int32_t segsum_off;   // offset into the per-point segment-sum table ([nseg][S][2])
    int32_t maxd_off;     // offset into the max-distance table ([nseg])
    int32_t rescale;      // propagators carry per-step scale factors (long chains)
    int32_t homog;        // every segment carries the same state beads (same bonds on every link)
};

// Chains of at least this many segments propagate with per-step rescaling: each step's input is
// divided by the largest value of the previous step's output (per point and sweep), and the
// logarithms of those factors are summed back in where densities and normalisations form products
// of the two sweeps. The reference's only rescaling is the constant normative per bond, which
// leaves 500-mer propagators outside the double range once a brush builds a potential.
#define SDFT_RESCALE_MIN_SEG 50

// System-wide constants shared by every point of a context
struct DevSystem {
    int32_t NB, MB, n_species, K;
    int32_t functionals;
    int32_t segsum_len;   // per point
    double lB;
    double sigma[SDFT_MAX_NB];
    double z[SDFT_MAX_NB];
    int32_t bond_i[SDFT_MAX_MB];
    int32_t bond_j[SDFT_MAX_MB];
    int32_t hard_walls;
    int32_t charged_walls;
    int32_t ncls;                     // distinct bead diameters (the MSA sums run over these classes)
    int32_t cls[SDFT_MAX_NB];         // class of each bead
    double cls_sigma[SDFT_MAX_NB];    // diameter of each class
    // square well: pair u couples bead types sw_i[u] <= sw_j[u]; its two kernels (depth times the
    // shell weight step(lambda sigma_ij) - step(sigma_ij)) sit at sw_kbase + 2u (the half-step
    // averaged form used for mu) and sw_kbase + 2u + 1 (plain, for the free energy)
    int32_t n_sw, sw_kbase;
    int32_t sw_i[SDFT_MAX_SW], sw_j[SDFT_MAX_SW];
    double sw_eps[SDFT_MAX_SW], sw_lam[SDFT_MAX_SW];
    int32_t n_sw_beads, nw_sw_base;   // bead types with a pair; their weighted fields start at nw_sw_base
    int32_t sw_bead[SDFT_MAX_NB];
};

// one term of a convolution job: out += sign-structured T(+-)H(kern) * in
struct ConvTerm {
    int32_t kern;     // kernel index within the set
    int32_t odd;      // 0: A = c[n-n'] + c[n+n'+1];  1: forward odd  A = c[n-n'] + c[n+n'+1] (c odd)
                      // 2: transpose of odd          A = -c[n-n'] + c[n+n'+1]
    int32_t infam;    // 0: rhoT family, 1: dphi family
    int32_t infield;
};

struct ConvJob {
    int32_t nterms;
    ConvTerm t[3];
    int32_t outkind;   // 0: weighted-density family (T), 1: F(x) (double) at field offset,
                       // 2: shared back-projection partial (double, one field per diameter class)
    int32_t outfield;  // field index in the family, or field offset in units of NP within the solution vector
    int32_t absval;    // take |.| of the result (scalar weighted densities)
    int32_t addfield;  // outkind 1: shared partial added to the result (-1: none)
    int32_t outfield2; // outkind 1: second F(x) field receiving the same values (-1: none)
    int32_t addfield2; // outkind 1: second shared partial added to the result (-1: none)
    double scale;
};

// Anderson / solve state per point
struct DevSolveState {
    int32_t iter;         // iterations done
    int32_t hist_base;    // iterations done when the current history started (0, or the precision switch)
    int32_t restarted;    // history was restarted at the precision switch
    int32_t use_picard;
    int32_t converged;
    int32_t prec;         // 32 or 64: precision of the next objective
    int32_t switch_iter;  // iteration of the FP32 -> FP64 switch (0 if none)
    int32_t mode;         // 0 Picard, 1 Anderson, for the pending update
    int32_t ncols;        // number of difference columns used by the pending update
    int32_t failed;       // 1: max_iters reached, 2: non-finite residual
    int32_t want64;       // mixed mode: residual below the switch threshold, switch at the next poll
    double err;
    double amix;
    double lam[SDFT_MAX_M];
};

struct SolveParams {
    double tole, damping, mixing_max, switch_tole, fp32_switch;
    int32_t m, max_iters, mixed, start_prec;
};

#define HIP_CHECK_RET(expr)                                                     \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            set_error(ctx, SDFT_E_HIP, "%s failed: %s (%s:%d)", #expr,          \
                      hipGetErrorString(_e), __FILE__, __LINE__);               \
            return SDFT_E_HIP;                                                  \
        }                                                                       \
    } while (0)

/*
 * libsoftdft_hip -- MI355X (gfx950) implementation of the SoftDFT iDFT
 * self-consistent-field (SCF) evaluation behind a plain C ABI.
 *
 * A context holds a batch of independent state points that share one
 * molecular system layout (bead types, bond types, species/segment topology,
 * functional list) but may differ in grid length, delta_muH, mu_species,
 * external fields and wall charges (pH / salt / slit-width sweeps).
 *
 * Conventions
 *  - every entry point returns 0 on success or a negative SDFT_E* code; the
 *    message of the last failure is available from sdft_last_error();
 *  - the caller owns every host buffer; the library owns device memory;
 *  - a context is bound to one HIP stream and is not re-entrant;
 *  - the solution vector of one state point uses the reference's flatten
 *    order (Julia column-major, first index fastest):
 *        mu_ex_K (NP, NB) ++ lng_K (NP, 2, MB) ++ Psi (NP) ++ PsiC (1)
 *    so n_p = NP*(NB + 2*MB + 1) + 1; batched buffers concatenate points.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *  - sdft_objective  <- objective_dft(problem_data) -> (mu_ex_K, lng_K, Psi, PsiC)
 *                       src/iDFT/evaluation/objective/objective_dft.jl:17-30
 *  - sdft_solve      <- solver_dft(dft_system) + solver_function(::AndersonMixing, ...)
 *                       src/iDFT/evaluation/objective/solver_dft.jl:5-25,
 *                       src/NonlinearSolvers/fixed_point/AndersonMixing.jl:56-169
 *  - sdft_observables <- grand_potential (src/iDFT/evaluation/analysis/grand_potential.jl:4-28),
 *                       write_total_state_fractions (degree_of_dissociation.jl:1-78),
 *                       contact_value_theorem (contact_value_theorem.jl:1-20)
 *  - sdft_get_field  <- reads of fields.rho_K / fields.excess after a solve
 *                       (src/iDFT/structures/{DensityFields,ExcessFields}.jl)
 *  - sdft_create     <- setup_iDFT (src/iDFT/setup/setup_iDFT.jl:9-30): takes the
 *                       host-built arrays (Ext, trapez, wall charge, grafting)
 *                       and precomputes the k-space weights of
 *                       src/iDFT/coordinate_system/Cartesian/fourier_transforms/ (delta, step, vector, coulomb weights)
 */
#ifndef SOFTDFT_HIP_H
#define SOFTDFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDFT_ABI_VERSION 5

/* error codes */
#define SDFT_OK 0
#define SDFT_E_ARG (-1)          /* bad descriptor / argument */
#define SDFT_E_HIP (-2)          /* HIP runtime error */
#define SDFT_E_NORMALIZATION (-3)/* non-finite or non-positive chain normalisation (handle_fixed.jl:110) */
#define SDFT_E_NOCONV (-4)       /* max_iters reached for at least one state point */
#define SDFT_E_UNSUPPORTED (-5)  /* topology / option outside this build */
#define SDFT_E_STATE (-6)        /* call order violated (e.g. observables before solve) */

/* functional flags (src/iDFT/functionals/Abstract_Types/functional_registry.jl:54-71) */
#define SDFT_F_MFMT 1
#define SDFT_F_TPT1MFMT 2
#define SDFT_F_MSA 4
#define SDFT_F_TPT1MSA 8
#define SDFT_F_SW 16      /* mean-field square well (square_well.jl:19-123); needs the sw_* arrays below */

/* species models */
#define SDFT_MODEL_FREELY_JOINTED 0
#define SDFT_MODEL_MONOMERBEAD 1

/* precision of the objective evaluation */
#define SDFT_PREC_FP64 64
#define SDFT_PREC_FP32 32
#define SDFT_PREC_MIXED 3264   /* FP32 until ||F(x)-x||_inf < switch, then FP64 */

/* field ids for sdft_get_field */
#define SDFT_FIELD_RHO_BEADS 1   /* (NP, NB) */
#define SDFT_FIELD_RHO_BONDS 2   /* (NP, 2, MB) */
#define SDFT_FIELD_X 3           /* current solution vector (n_p) */
#define SDFT_FIELD_FX 4          /* last F(x) (n_p) */
#define SDFT_FIELD_SEGMENT_SUMS 5 /* per species/segment/state: sum_K rho_seg (raw), then sum_K rho_seg*trapez*dz */
/* species u's segment density profiles rho_K.segments[u] (accumulate_densities.jl:95-184), layout
 * (NP, S_u, N_u) with S_u the species' states per segment (0 for absent states): field id
 * SDFT_FIELD_RHO_SEGMENTS + u */
#define SDFT_FIELD_RHO_SEGMENTS 16

typedef struct sdft_ctx sdft_ctx_t;

/*
 * Batch descriptor.  Per-point arrays are concatenated over points in point
 * order; grid arrays use the point's own NP.  All indices are 0-based.
 */
typedef struct sdft_desc {
    int32_t abi_version;       /* must equal SDFT_ABI_VERSION */
    int32_t n_points;
    int32_t NB;                /* bead (monomer state) types */
    int32_t MB;                /* bond types */
    int32_t n_species;
    int32_t functionals;       /* OR of SDFT_F_* */
    double bjerrum;            /* l_B in units of l_sc */
    const double* diameters;   /* [NB] */
    const double* valences;    /* [NB] */
    const int32_t* bond_types; /* [MB][2], i <= j */

    /* species: linear chains, segment 0 is the root (parse_topology.jl) */
    const int32_t* species_nseg;    /* [n_species] */
    const int32_t* species_model;   /* [n_species] SDFT_MODEL_* */
    const int32_t* species_nstates; /* [n_species] max states per segment S_u */
    const int32_t* seg_states;      /* concat over species of [nseg_u][S_u] bead types, -1 = unused */
    const int32_t* species_fixed_seg; /* [n_species] grafted segment (0) or -1 */
    const double* fixed_density;    /* [n_species] grafting density (used when fixed) */
    const double* max_distance;     /* concat over species of [nseg_u] max distance to the grafted segment */

    /* per point scalars */
    const int32_t* NP;             /* [P] */
    const double* dz;              /* [P] */
    const double* delta_muH;       /* [P][NB] */
    const double* mu_species;      /* [P][n_species] (grafted species already 0) */
    const double* total_charge;    /* [P] wall charge sum */
    const int32_t* fixed_coord;    /* [P][n_species] grid index of the grafted segment, -1 if none */

    /* per point grids */
    const double* Ext;             /* concat [P] of (NP, NB) column-major */
    const double* trapez;          /* concat [P] of (NP, NB) */
    const double* surf;            /* concat [P] of (NP): wall charge / dz on the primary half */

    /* walls (for the contact-value theorem); per point */
    int32_t hard_walls;            /* 0/1 */
    const int32_t* wall_lo;        /* [P][NB] contact index of the left wall */
    const int32_t* wall_hi;        /* [P][NB] */
    const double* wall_charge;     /* [P][2] or NULL */

    /* square-well pairs (properties.pairs[:square_well], process_pairs.jl:42-95), used with SDFT_F_SW */
    int32_t n_sw_pairs;
    const int32_t* sw_pairs;       /* [n_sw_pairs][2] bead types, i <= j */
    const double* sw_energy;       /* [n_sw_pairs] well depth eps_ij */
    const double* sw_lambda;       /* [n_sw_pairs] well width multiplier lambda_ij */

    /* square-well walls (sw_walls.jl:61-103), for the contact-value theorem only (their potential is
     * part of Ext): per point and bead type, the grid index of the well edge at the left and right
     * wall (-1: none) and exp(energy) - 1; both NULL without square-well walls */
    const int32_t* sw_wall_idx;    /* [P][NB][2] */
    const double* sw_wall_fac;     /* [P][NB][2] */

    /* multi-dimensional Cartesian grids (create_cartesian.jl:25-68; every dimension mirrored,
     * process_cartesian.jl:14-45): ndim 0 or 1 is the one-dimensional path above. With ndim 2 or 3
     * every point has grid[p][d] points and bin width bins[p][d] per dimension (first dimension
     * fastest in every grid array), the same for all points of the context; NP[p] is their product
     * (2 <= grid <= 128 per dimension) and dz[p] their volume element prod(bins). Grid arrays (Ext,
     * trapez, surf) are flattened over the grid. Not supported there: grafted segments on 2-D grids, the square-well
     * functional, and the contact-value tables (hard_walls must be 0: the host forms the per-dimension
     * contact values from the bead densities). grid, bins: [P][3] (entries past ndim ignored). */
    int32_t ndim;
    const int32_t* grid;
    const double* bins;
} sdft_desc_t;

typedef struct sdft_solver_opts {
    double tole;          /* ||F(x)-x||_inf convergence tolerance */
    double damping;
    double mixing_max;
    double switch_tole;   /* Picard -> Anderson switch */
    int32_t anderm;       /* Anderson depth (<= 8) */
    int32_t max_iters;
    int32_t precision;    /* SDFT_PREC_* */
    double fp32_switch;   /* mixed: switch to FP64 once ||F(x)-x||_inf < fp32_switch */
    int32_t band_fp32;    /* 1: the mixed mode's propagators (FP32 phase and FP64 polish) use the banded
                             bond kernel, the plain FP64 mode the dense one; 2: every mode banded (long
                             grafted chains, whose propagators span more decades than the dense or FFT
                             operator's round-off floor resolves); 0: dense everywhere */
    int32_t check_every;  /* host polls convergence every this many iterations */
} sdft_solver_opts_t;

int32_t sdft_abi_version(void);
const char* sdft_last_error(const sdft_ctx_t* ctx);   /* ctx may be NULL: global last error */

int32_t sdft_create(const sdft_desc_t* desc, sdft_ctx_t** out);
void sdft_destroy(sdft_ctx_t* ctx);

/* total length of the batched solution vector and per-point offsets (n_points+1 entries) */
int64_t sdft_solution_length(const sdft_ctx_t* ctx);
int32_t sdft_solution_offsets(const sdft_ctx_t* ctx, int64_t* offsets);

/* one evaluation of F(x) for every point; x and fx are host buffers of sdft_solution_length */
int32_t sdft_objective(sdft_ctx_t* ctx, const double* x, double* fx, int32_t precision);

/* Anderson solve of every point from x_inout (host, batched); iters/err/prec_switch_iter per point */
int32_t sdft_solve(sdft_ctx_t* ctx, const sdft_solver_opts_t* opts, double* x_inout,
                   int32_t* iters, double* err);

/* after sdft_solve: Omega[P], alpha[P][n_species*S_max] (NaN where undefined), p_cvt[P][2] */
int32_t sdft_observables(sdft_ctx_t* ctx, double* omega, double* alpha, double* p_cvt);

/* sdft_observables without any device allocation. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Same results as sdft_observables (the same kernels, in the same order), with every
 * partial sum in a static device array of 2^20 doubles (8 MiB per GPU, held until the process ends)
 * instead of the context's lazily allocated work buffer. Limit: the batch must need at most 2^20
 * doubles, P * (11 + n_species * S_max) + P * sum over multistate species of N_u * S_u; larger
 * batches return SDFT_E_ARG. Calls that overlap on one GPU (any contexts, any host threads) share
 * that array and race. */
int32_t sdft_observables_noalloc(sdft_ctx_t* ctx, double* omega, double* alpha, double* p_cvt);

/* copy one field of one point to host (layouts above) */
int32_t sdft_get_field(sdft_ctx_t* ctx, int32_t field_id, int32_t point, double* host_out, int64_t capacity);

/* replace every point's chemistry: delta_muH [P][NB] and mu_species [P][n_species] (grafted species
 * 0), with the chain normalisations that follow from them (normative, child_propagators.jl:41-50;
 * normative^(N-1), handle_fixed.jl:105-107). This is the state a pseudo-arclength step changes when
 * its continuation variable is the pSpecies of H+: objective_pac (src/NonlinearSolvers/continuation/
 * objective_pac.jl:230-260) rebuilds the system at the new pH (update_cont_var + establish_dft_system)
 * before each objective_dft. Everything else of the context (grids, walls, topology) is kept. */
int32_t sdft_set_chemistry(sdft_ctx_t* ctx, const double* delta_muH, const double* mu_species);

/* replace every point's grafting densities [P][n_species] (entries of species without a grafted
 * segment are stored and unused): the chain normalisation of a grafted species is density / lambda
 * (handle_fixed.jl:31-43). This is the state objective_pac_graft changes each evaluation
 * (update_cont_var_graft, src/NonlinearSolvers/continuation/objective_pac_graft.jl:286-288: the first
 * species' fixed density = exp(continuation variable)). */
int32_t sdft_set_graft_density(sdft_ctx_t* ctx, const double* density);

/* timing of the last sdft_solve: device milliseconds of the whole solve and total iterations */
int32_t sdft_last_timing(const sdft_ctx_t* ctx, double* solve_ms, double* iterations_total);

/* bond operator applied by sdft_objective at each precision: 1 = banded (the filtered bond kernel
 * truncated where |c| <= 1e-9 max|c|, half-width rounded up to a multiple of 8), 0 = dense.
 * Defaults: FP32 banded, FP64 dense.  sdft_solve takes its own choice from its options.
 * (precision control; no counterpart in the reference) */
int32_t sdft_set_operator_band(sdft_ctx_t* ctx, int32_t fp32_banded, int32_t fp64_banded);

/* batch iterations of the last sdft_solve during which the FP32 and the FP64 pipeline ran
 * (instrumentation; no counterpart in the reference) */
int32_t sdft_last_pipeline_iters(const sdft_ctx_t* ctx, int32_t* fp32_iters, int32_t* fp64_iters);

/* per point, the iterations done when the last mixed-precision sdft_solve moved it from FP32 to FP64
 * (0: no switch, e.g. a plain FP32 or FP64 solve); with the point's total iterations this splits its
 * work between the two precisions (instrumentation; no counterpart in the reference) */
int32_t sdft_last_switch_iters(const sdft_ctx_t* ctx, int32_t* switch_iter);

/* launch mode of sdft_solve: graph = 1 (default) runs each window of check_every iterations
 * between two host polls as one captured hipGraph (instantiated once per pair of active
 * pipelines and solve), 0 launches every kernel on the stream; concurrent = 1 (default) runs the
 * FP32 and FP64 pipelines on two streams in iterations where both have points (they evaluate
 * disjoint points), 0 one after the other; < 0 leaves a setting unchanged.  graphs_built (may be
 * NULL) receives the number of graphs instantiated so far.  Results are identical in every mode.
 * (launch control; no counterpart in the reference) */
int32_t sdft_set_launch_mode(sdft_ctx_t* ctx, int32_t graph, int32_t concurrent, int64_t* graphs_built);

/* Autotuner (no counterpart in the reference; the reference has one CPU code path).  Measures,
 * with HIP events on the context's stream, the candidate launch configurations of this batch and
 * keeps the fastest:
 *   SDFT_TUNE_PROPAGATOR  per chain species and precision: the block-per-problem mapping of the
 *                         propagator recursion.  choice 0 = one operator-GEMM launch per step
 *                         (workgroups tile the grid of every point), 1 = persistent chain kernel (one
 *                         workgroup per state point and sweep runs every step) with the input weights
 *                         held in registers, 2 = the same kernel re-reading them every step.
 *                         key = species index.
 *   SDFT_TUNE_CONV_FWD / _PRE / _BACK
 *                         per grid-length bin (kernel sets whose NP rounds up to the same multiple of
 *                         256; key = that bound) and convolution stage: the tile shape, choice
 *                         0 = 128 x 16, 1 = 64 x 64, 2 = 128 x 64, 3 = 64 x 32, 4 = 64 x 16 rows x
 *                         columns of the operator-GEMM, or 5 = batched FFT convolution (mirror
 *                         lengths 2 NP that are powers of two from 4096 to 65536; elsewhere not a
 *                         candidate, and a forced 5 falls back to the fixed rule).  Timed at FP32
 *                         and FP64 and chosen on the sum.
 * us[k] is the measured microseconds per launch sequence of candidate k (< 0: not a candidate).
 * Without a call the library uses fixed rules (chain kernel wherever eligible; the FFT convolution
 * where the mirror length allows it, else 64 x 32 tiles where a grid length's operators average >= 48
 * columns, else 128 x 16).  The choice changes summation
 * orders (FP64 results move in the last bits) and the banded operator's padding, never the algorithm.
 * Call it before sdft_solve / sdft_objective: it overwrites the work buffers. */
#define SDFT_TUNE_PROPAGATOR 1
#define SDFT_TUNE_CONV_FWD 2
#define SDFT_TUNE_CONV_PRE 3
#define SDFT_TUNE_CONV_BACK 4
typedef struct {
    int32_t kind;        /* SDFT_TUNE_* */
    int32_t key;         /* species index, or the bin's NP bound */
    int32_t precision;   /* 32 / 64 (propagators), 0 (convolutions: FP32 + FP64) */
    int32_t choice;
    double us[6];
} sdft_tune_entry_t;
int32_t sdft_autotune(sdft_ctx_t* ctx, int32_t reps, sdft_tune_entry_t* entries, int32_t capacity,
                      int32_t* n_entries);
/* applies decisions without measuring (a saved sdft_autotune result for the same batch layout, or
 * a forced configuration): PROPAGATOR entries with precision 0 apply to both precisions, and
 * choices 1 and 2 take effect only where the chain kernel is eligible; CONV entries with key 0 apply to
 * every grid length, choice -1 restores the fixed rule.  us[] is ignored. */
int32_t sdft_set_tuning(sdft_ctx_t* ctx, const sdft_tune_entry_t* entries, int32_t n);

/*
 * Bulk (homogeneous) state of every state point of a batch, solved on the device: the inputs
 * sdft_create takes from the bulk (mu_species; x0 = bulk mu_ex and lng) and the bulk pressure.
 * Replaces the iLST bulk solve, src/iLST/evaluation/objective/solver_bulkstate.jl:31-103 (with
 * residual_bulkstate.jl:26-42, functions/balance_charge.jl:21-79, the bulk chain propagators
 * src/iLST/species_model/) and eval_bulk_pressure (src/iLST/evaluation/analysis/eval_pressure.jl:10-25).
 * Linear chains (freely jointed or monomerbead); one wavefront per point runs Newton on the
 * self-consistency of (mu_ex, lng) with a forward-difference Jacobian to max |residual| < tol.
 */
typedef struct sdft_bulk_desc {
    int32_t abi_version;
    int32_t n_points, NB, MB, n_species;
    int32_t functionals;             /* OR of SDFT_F_* (the ideal term is always present) */
    double bjerrum;
    const double* diameters;         /* [NB] */
    const double* valences;          /* [NB] */
    const int32_t* bond_types;       /* [MB][2], i <= j */
    const int32_t* species_nseg;     /* [n_species] */
    const int32_t* species_model;    /* [n_species] SDFT_MODEL_* */
    const int32_t* species_nstates;  /* [n_species] */
    const int32_t* seg_states;       /* concat over species of [nseg_u][S_u] bead types, -1 = unused */
    const int32_t* adjustable;       /* [n_species] 1: charge compensator (balance_charge.jl) */
    const double* delta_muH;         /* [P][NB] */
    const double* base_density;      /* [P][n_species] chains per volume before the charge balance */
    const double* log_activity;      /* [P][n_species] ln(c 10^-pSpecies) for pSpecies species (their
                                        density is exp(log_activity + ln Xi)), NaN otherwise */
    const double* species_charge;    /* [P][n_species] charge per chain used by the balance */
    int32_t n_sw_pairs;              /* square-well pairs (SDFT_F_SW) */
    const int32_t* sw_pairs;         /* [n_sw_pairs][2] */
    const double* sw_energy;         /* [n_sw_pairs] */
    const double* sw_volume;         /* [n_sw_pairs] 4 pi / 3 sigma_ij^3 (lambda^3 - 1) */
} sdft_bulk_desc_t;

typedef struct sdft_bulk_out {       /* host arrays written by sdft_bulk_solve */
    double* mu_species;              /* [P][n_species] */
    double* rho_species;             /* [P][n_species] */
    double* mu_ex;                   /* [P][NB] */
    double* rho_beads;               /* [P][NB] */
    double* lng;                     /* [P][MB] */
    double* rho_bonds;               /* [P][2][MB] */
    double* pressure;                /* [P] */
    int32_t* iterations;             /* [P] Newton iterations; -1: not converged, -2: no compensator */
} sdft_bulk_out_t;

/* returns SDFT_E_NOCONV if any point did not converge (its iterations entry says why) */
int32_t sdft_bulk_solve(const sdft_bulk_desc_t* desc, double tol, int32_t max_newton, sdft_bulk_out_t* out);

/* kernel classes for sdft_kernel_profile */
#define SDFT_KP_PROP_FP32 0      /* one propagator step, FP32, banded bond kernel */
#define SDFT_KP_PROP_FP64 1      /* one propagator step, FP64, bond operator as in the last solve */
#define SDFT_KP_CONV_FWD_FP32 2  /* weighted-density convolutions, FP32 */
#define SDFT_KP_CONV_FWD_FP64 3
#define SDFT_KP_CONV_BACK_FP32 4 /* back-projection convolutions, FP32 */
#define SDFT_KP_CONV_BACK_FP64 5
#define SDFT_KP_POINTWISE_FP64 6 /* FMT/MSA derivatives (FP64 arithmetic) on FP64 fields */
#define SDFT_KP_CHAIN_FP32 7     /* persistent banded chain propagation (all steps, both sweeps), FP32 */
#define SDFT_KP_CHAIN_FP64 8     /* the same in FP64 */
#define SDFT_KP_ACCUM_FP32 9     /* segment, bead and bond densities from the stored propagators, FP32 */
#define SDFT_KP_ACCUM_FP64 10    /* the same in FP64 */

/* Re-launch one kernel class `reps` times for every point on the context's
 * stream, bracketed by HIP events, after a solve or objective has filled the
 * fields.  Returns the average duration of one launch and its algorithmic
 * flop count (2 flops per multiply-add of the operator actually applied).
 * Profiling aid; no counterpart in the reference. */
int32_t sdft_kernel_profile(sdft_ctx_t* ctx, int32_t kernel_class, int32_t reps, double* avg_ms, double* flops);

#ifdef __cplusplus
}
#endif
#endif

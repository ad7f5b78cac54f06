/*
 * oakg.h -- C ABI of the MI355X (gfx950) local ensemble square-root analysis.
 *
 * The entry points replace, for OAK's local scheme (schemetype = 1), the
 * analysis interfaces below; every array is column-major (Fortran layout) with
 * an explicit leading dimension, indices are 0-based, and state vectors are in
 * OAK's zone-permuted order (each zone contiguous, assimilation.F90:869-880).
 *
 *   oakg_local_analysis      <- locAnalysis(zoneSize, selectObservations, xf, Hxf,
 *                               yo, Sf, HSf, R, xa, Sa, amplitudes, localise_obs)
 *                               rrsqrt.F90:433-466 (zone loop :258-426, per-zone
 *                               solve analysisIncrement :100-190), with the
 *                               selection of assimilation.F90:3683-3771 done on
 *                               the GPU and R = DiagCovar(rmse^2)
 *                               (covariance.F90:70-79, assimilation.F90:2120-2127)
 *   oakg_local_analysis_csr  <- the same call with an arbitrary selectObservations
 *                               callback, evaluated by the caller per zone
 *                               (e.g. test/test_rrsqrt.F90:254-271)
 *   oakg_ensemble_analysis   <- oak_assim(ntime, n, r, Ef, Ea)
 *                               assimilation.F90:4919-4927, i.e. Assim's ensemble
 *                               path :3094-3135, :3299-3364, :3558-3563 (H·E,
 *                               anomalies, local analysis, inflation,
 *                               maxCorrection, ensemble rebuild)
 *   oakg_ensemble_analysis_f32  mixed-precision variant of the above (FP32
 *                               ensembles, FP64 eigensolve); no OAK counterpart
 *                               beyond oak_assim built with PRECISION=single
 *   oakg_report              <- report(unit, prefix, mjd, ingrid, R, HS, yo_Hx,
 *                               exclude_obs) assimilation.F90:3584-3628 (the
 *                               statistics; the caller writes the log lines)
 *   oakg_ensemble_report     <- the forecast/analysis report calls of Assim's
 *                               ensemble path, assimilation.F90:3389-3391,
 *                               3424-3429
 *   oakg_coo_to_csr          <- SparseMatrix (matoper.F90:30-39) as used by
 *                               obsoper (assimilation.F90:2836-2898)
 *   oakg_partition           <- initPartition (assimilation.F90:578-641) after
 *                               the gap relabelling of :408-412
 *   oakg_create/oakg_destroy <- oak_init/oak_done (assimilation.F90:4911-4933)
 *                               for the device context only (no init file)
 *
 * Conventions: every function returns OAKG_OK (0) or an error code; the text
 * of the last error of a context is returned by oakg_last_error.  With host
 * pointers, outputs are written only on success.  A context is used by one
 * host thread at a time; different contexts are independent.  Calls block
 * until the results are in the output arrays.
 */
#ifndef OAKG_H
#define OAKG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OAKG_OK 0
#define OAKG_EARG 1   /* invalid argument (sizes, NULL arrays, rmse <= 0)           */
#define OAKG_ENAN 2   /* NaN in the amplitudes: rrsqrt.F90:145-149 (ERROR_STOP)     */
#define OAKG_EHIP 3   /* HIP runtime error                                          */
#define OAKG_ENOMEM 4 /* device allocation failed                                   */
#define OAKG_EUNSUP 5 /* configuration not supported (e.g. N above the maximum)     */
#define OAKG_ECONV 6  /* eigensolver did not finish a zone (N > 128: no fallback)   */

#define OAKG_FP64 0  /* plain FP64 path                                             */
#define OAKG_MIXED 1 /* FP32/BF16 contractions, FP64 eigensolve                     */

#define OAKG_DEVICE_PTRS 1u /* flag: every array argument is a device pointer        */

#define OAKG_CARTESIAN 0        /* metrictype, assimilation.F90:108-112 */
#define OAKG_SPHERICAL 1
#define OAKG_SPHERICAL_APPROX 2

#define OAKG_MAX_N 256 /* largest ensemble size of this build (the likelihood: 128) */

typedef struct oakg_ctx oakg_ctx;

/* Zone layout and localisation geometry (assimilation.F90:210-229, 400-451). */
typedef struct {
  int64_t nzones;
  const int64_t* zone_ptr; /* [nzones+1] start of each zone in the permuted state   */
  const double* zpos;      /* [nzones*4] x,y,z,t of each zone's FIRST element,       */
                           /*   one row of 4 per zone (assimilation.F90:3713-3740)  */
  const double* corr_len;  /* [nzones] hCorrLengthToObs of that element  (L)        */
  const double* max_len;   /* [nzones] hMaxCorrLengthToObs of that element (cutoff) */
  int32_t metric;          /* OAKG_CARTESIAN / OAKG_SPHERICAL / ..._APPROX          */
  int32_t loctype;         /* 1 horizontal distance, 2 |dz|, 3 |dt|                 */
} oakg_zones;

/* Observations (obsGridX/Y/Z/T, yo and the diagonal of R). */
typedef struct {
  int64_t m;
  const double* x; /* [m] obsGridX (longitude or x)                  */
  const double* y; /* [m] obsGridY (latitude or y)                   */
  const double* z; /* [m] obsGridZ, may be NULL unless loctype == 2  */
  const double* t; /* [m] obsGridT, may be NULL unless loctype == 3  */
  const double* yo;       /* [m] observations                                     */
  const double* rmse2;    /* [m] R = diag(rmse^2), every entry > 0                */
  const uint8_t* excluded; /* [m] 1 = excluded (assimilation.F90:3086-3092), or NULL */
} oakg_obs;

/* Observation operator H (m x n), compressed by observation row. */
typedef struct {
  int64_t m, n, nnz;
  const int64_t* row_ptr; /* [m+1]                                   */
  const int32_t* col;     /* [nnz] column = permuted state index     */
  const double* val;      /* [nnz]                                   */
} oakg_csr;

/* Per-call statistics (filled by every analysis call). */
typedef struct {
  int64_t zones;           /* zones in the call                                  */
  int64_t zones_analysed;  /* zones with at least one relevant observation       */
  int64_t obs_selected;    /* sum over zones of the local observation count p_z  */
  int64_t pmax;            /* largest p_z                                        */
  int32_t sweeps_max;      /* most Jacobi sweeps used by one zone                */
  int32_t eig_fallbacks;   /* zones the QL eigensolver left to Jacobi            */
  double ms_total;         /* device time of the whole call                      */
  /* device time per phase, recorded only with oakg_set_timing(ctx, 1):
     prep (H·E, anomalies), select, gram (S^T R^-1 S), eig (eigensolve),
     tmat (transform T), update (Xa) */
  double ms_prep, ms_select, ms_gram, ms_eig, ms_tmat, ms_update;
} oakg_stats;

int oakg_version(void);
/* The context keeps a device workspace for the zone chunks of a call, sized
   to a quarter of the device memory between 8 and 64 GiB, or to
   OAKG_WORKSPACE_BYTES from the environment (at least 1 MiB) when that is set
   (allocated on first use and kept until oakg_destroy). */
int oakg_create(oakg_ctx** ctx, int device, int precision);
int oakg_destroy(oakg_ctx* ctx);
const char* oakg_last_error(const oakg_ctx* ctx);
/* Launch on this HIP stream (hipStream_t) instead of the default stream.
   With 4096 zones or more (and phase timing off) the zone chunks also run on
   internal non-blocking streams of the context (default priority; two with
   the default OAKG_TUNE_SLOTS),
   ordered after the work already on `stream` and joined back into it by
   events before the call returns; the call synchronises `stream`. */
int oakg_set_stream(oakg_ctx* ctx, void* stream);
/* 1 = record per-phase device times into oakg_stats (adds events; the zone
   chunks then run in order on the context's stream). */
int oakg_set_timing(oakg_ctx* ctx, int on);
/* Eigensolver of the N x N ensemble-space matrix (OAK calls LAPACK dsyev,
   matoper_inc.F90:986-1005):
     OAKG_EIG_QL      Householder tridiagonalisation + implicit QL (default;
                      zones where QL fails fall back to Jacobi)
     OAKG_EIG_JACOBI  one-sided Jacobi for every zone                        */
#define OAKG_EIG_JACOBI 0
#define OAKG_EIG_QL 1
int oakg_set_eigensolver(oakg_ctx* ctx, int method);
int oakg_get_stats(oakg_ctx* ctx, oakg_stats* out);
/* Tuning knobs of the zone loop (each zone is computed alone, so every
 * setting but OAKG_TUNE_DUAL and OAKG_TUNE_TRIDIAG_SPLIT = 2, 3, 4 gives the same
 * bits; those two give the same analysis to rounding):
 *   OAKG_TUNE_SLOTS       zone chunks in flight on as many streams (1..4,
 *                         default 3; from 4096 zones, phase timing off)
 *   OAKG_TUNE_QL_ZONES    zones per wave of the QL kernel (4, 8, 16, 32; 16)
 *   OAKG_TUNE_GRAPH       1: the zone loop (Gram, eigensolve, transform,
 *                         update of every chunk, on all its streams) is
 *                         captured in a HIP graph on first use and replayed
 *                         while its kernel arguments and buffers stay the
 *                         same (a caller that repeats an analysis on the same
 *                         device buffers); 0: launched eagerly (default)
 *   OAKG_TUNE_ORDER       1: zones run in order of decreasing observation
 *                         count (quarter-octave buckets; every zone chunk
 *                         gets the same mix), the heaviest first; 0: in
 *                         zone order (default)
 *   OAKG_TUNE_TRIDIAG_SPLIT  1: at N > 64 the Householder reduction runs as
 *                         two kernels (the first half of the steps, then the
 *                         rest with fewer registers and more zones per CU);
 *                         2: three kernels, the middle quarter of the steps
 *                         on 6 x 6 register tiles (sums grouped differently);
 *                         3: two kernels, the first with the lower triangle
 *                         only (half the registers per zone, four zones per
 *                         CU; sums grouped differently); 4: as 3, the second
 *                         on the lower triangle as well (one wave per zone);
 *                         0: one kernel (default)
 *   OAKG_TUNE_DUAL        at N > 64, zones with 1 <= p < value local
 *                         observations (value 16, 32 or 64; 0 = off, the
 *                         default) are solved in the p x p observation space:
 *                         G = S~ S~^T instead of M = I + S~^T S~, which have
 *                         the same nonzero spectrum.  The analysis is the
 *                         same to rounding (not bit for bit).
 *   OAKG_TUNE_PERSIST     at N > 64, the Gram kernel as a persistent grid
 *                         (two workgroups per CU) taking zones from work
 *                         counters in the zone loop's order: 2 a zone per
 *                         workgroup, 1 each wave the next zone of its
 *                         quarter of the tiles; 0: a workgroup per zone
 *                         (default)
 * oakg_get_tuning also reads OAKG_TUNE_GRAPH_REPLAYS / _CAPTURES (counts). */
#define OAKG_TUNE_SLOTS 1
#define OAKG_TUNE_QL_ZONES 2
#define OAKG_TUNE_GRAPH 3
#define OAKG_TUNE_GRAPH_REPLAYS 4
#define OAKG_TUNE_GRAPH_CAPTURES 5
#define OAKG_TUNE_ORDER 6
#define OAKG_TUNE_TRIDIAG_SPLIT 7
#define OAKG_TUNE_DUAL 8
#define OAKG_TUNE_PERSIST 9
int oakg_set_tuning(oakg_ctx* ctx, int key, int value);
int oakg_get_tuning(oakg_ctx* ctx, int key, int64_t* value);
/* Per-zone local observation counts p_z of the last call (host array). */
int oakg_get_counts(oakg_ctx* ctx, int64_t* counts, int64_t nzones);

/* locAnalysis: xa = xf + Sf·a_z per zone, Sa = Sf·T_z (rrsqrt.F90:100-190,
 * 258-466).  xf, xa: [n]; Sf, Sa: n x N; Hxf: [m]; HSf: m x N.  amplitudes
 * (N x nzones, may be NULL) is filled only when localise_obs == 0, as in
 * rrsqrt.F90:378-383.  localise_obs = 1 is OAK's default. */
int oakg_local_analysis(oakg_ctx* ctx, const oakg_zones* zones, const oakg_obs* obs, int32_t N,
                        const double* xf, const double* Hxf, const double* Sf, int64_t ldSf,
                        const double* HSf, int64_t ldHSf, double* xa, double* Sa, int64_t ldSa,
                        double* amplitudes, int32_t localise_obs, uint32_t flags);

/* Same, with the selection supplied per zone as CSR: zone z uses observations
 * sel_idx[sel_ptr[z] .. sel_ptr[z+1]) with taper weights sel_w (the weight
 * array that selectObservations returns).  With localise_obs == 0 the list of
 * a zone must hold all m observations (or none). */
int oakg_local_analysis_csr(oakg_ctx* ctx, int64_t nzones, const int64_t* zone_ptr,
                            const int64_t* sel_ptr, const int32_t* sel_idx, const double* sel_w,
                            const oakg_obs* obs, int32_t N, const double* xf, const double* Hxf,
                            const double* Sf, int64_t ldSf, const double* HSf, int64_t ldHSf,
                            double* xa, double* Sa, int64_t ldSa, double* amplitudes,
                            int32_t localise_obs, uint32_t flags);

/* oakg_local_analysis_csr without any device allocation.  No counterpart in
 * OAK: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as oakg_local_analysis_csr, with every array a
 * device pointer (flags must hold OAKG_DEVICE_PTRS).  Every internal buffer
 * (the observation-space copy HS, the per-zone matrices of the eigensolve,
 * the error word) comes from one fixed device array of
 * OAKG_NOALLOC_SCRATCH_BYTES per GPU, reserved when the library is loaded
 * and held until the process ends; calls whose m, N and zone count need
 * more than that (about m (NP + 2) 8 bytes + 4 bytes per zone + 64 zones of
 * eigensolver workspace, 52 MiB at N = 128) return OAKG_EARG before any
 * device work.  The zones run in chunks that fit, in order on the context's
 * stream; no events, streams or phase timers are created (ms_total and the
 * phase times read 0).  Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (other host threads, other contexts or
 * streams) race and must be serialised by the caller. */
#define OAKG_NOALLOC_SCRATCH_BYTES (256ull << 20)
int oakg_local_analysis_csr_noalloc(oakg_ctx* ctx, int64_t nzones, const int64_t* zone_ptr,
                                    const int64_t* sel_ptr, const int32_t* sel_idx,
                                    const double* sel_w, const oakg_obs* obs, int32_t N,
                                    const double* xf, const double* Hxf, const double* Sf,
                                    int64_t ldSf, const double* HSf, int64_t ldHSf, double* xa,
                                    double* Sa, int64_t ldSa, double* amplitudes,
                                    int32_t localise_obs, uint32_t flags);

/* The selection alone (assimilation.F90:3683-3771) for every zone: writes
 * sel_ptr[nzones+1] and, if sel_idx/sel_w are not NULL, up to cap entries of
 * observation indices and taper weights (zone by zone, in an order of the
 * implementation's choosing).  *total receives the number of entries. */
int oakg_select_observations(oakg_ctx* ctx, const oakg_zones* zones, const oakg_obs* obs,
                             int32_t localise_obs, int64_t* sel_ptr, int32_t* sel_idx,
                             double* sel_w, int64_t cap, int64_t* total, uint32_t flags);

/* oak_assim: Ef (n x N ensemble, zone-permuted) -> Ea.  H·Ef + Hshift gives the
 * observed ensemble; anomalies are scaled by 1/sqrt(N-1) (ppdef.h:52);
 * Sa is multiplied by inflation (inflation.mult) and xa - xf is clamped to
 * +-maxCorrection (assimilation.F90:3301-3312).  OAK's maxCorrection is a
 * per-element array, filled per variable from Model.maxCorrection(v) and
 * permuted like the state (assimilation.F90:462-476): max_correction_elem
 * gives it ([n], zone-permuted, every entry >= 0), or NULL to clamp every
 * element to the scalar max_correction (HUGE_VAL for none, OAK's default). */
int oakg_ensemble_analysis(oakg_ctx* ctx, const oakg_zones* zones, const oakg_obs* obs,
                           const oakg_csr* H, const double* Hshift, int32_t N,
                           const double* Ef, int64_t ldEf, double* Ea, int64_t ldEa,
                           double inflation, double max_correction,
                           const double* max_correction_elem, uint32_t flags);

/* Mixed-precision ensemble analysis: FP32 ensembles, the contractions
 * S = H·A', S^T R^-1 S and A'·W in FP32/BF16 with FP32 accumulation, the
 * N x N eigensolve, (1+lambda)^-1/2 and the R^-1 taper scaling in FP64.
 * Requires a context created with OAKG_MIXED. */
int oakg_ensemble_analysis_f32(oakg_ctx* ctx, const oakg_zones* zones, const oakg_obs* obs,
                               const oakg_csr* H, const float* Hshift, int32_t N,
                               const float* Ef, int64_t ldEf, float* Ea, int64_t ldEa,
                               double inflation, double max_correction,
                               const double* max_correction_elem, uint32_t flags);

/* Innovation statistics of report (assimilation.F90:3584-3628), which Assim
 * writes to the log for the forecast and the analysis (:3389-3391) and for
 * each observed variable (:3410-3430).  For each observation range r,
 * [range_ptr[r], range_ptr[r+1]) (range_ptr NULL and nranges = 1: all m),
 * with d = yo_Hx, D the exclusion mask (0 = excluded; excluded may be NULL),
 * R^-1 d = D (D d) / rmse2 (DCDCovar, covariance.F90:612-619) and
 * ingrid = count(D != 0) in the range:
 *   stats[5r+0]  ingrid
 *   stats[5r+1]  rms_yo-Hx                  sqrt(sum(d^2, D != 0) / ingrid)
 *   stats[5r+2]  bias_yo-Hx                 sum(d, D != 0) / ingrid
 *   stats[5r+3]  rms_invsqrtR_yo-Hx         sqrt(sum(d R^-1 d) / ingrid)
 *   stats[5r+4]  projection_yo-Hx_into_HSf  sqrt(|HS^T R^-1 d|^2 / ingrid)
 * (NaN where ingrid = 0, as in OAK).  HS is m x N (ldHS), N <= 256; every
 * rmse2 entry > 0.  The sums run in a fixed order: repeated calls give the
 * same bits. */
int oakg_report(oakg_ctx* ctx, int64_t m, int32_t N, const double* yo_Hx, const double* HS,
                int64_t ldHS, const double* rmse2, const uint8_t* excluded, int32_t nranges,
                const int64_t* range_ptr, double* stats, uint32_t flags);

/* The same statistics for an ensemble E (n x N, n = H->n) as Assim's
 * ensemble path forms them: HE = H·E + Hshift, Hx = mean(HE, 2),
 * HS = (HE - Hx)/sqrt(N-1), d = yo - Hx (assimilation.F90:3112-3134 for the
 * forecast Ef, :3325-3357 for the analysis Ea).  Only obs->yo, rmse2 and
 * excluded are used.  The _f32 variant takes FP32 ensembles (the mixed
 * path's state), with HS rounded to FP32 and the sums in FP64. */
int oakg_ensemble_report(oakg_ctx* ctx, const oakg_obs* obs, const oakg_csr* H,
                         const double* Hshift, int32_t N, const double* E, int64_t ldE,
                         int32_t nranges, const int64_t* range_ptr, double* stats,
                         uint32_t flags);
int oakg_ensemble_report_f32(oakg_ctx* ctx, const oakg_obs* obs, const oakg_csr* H,
                             const float* Hshift, int32_t N, const float* E, int64_t ldE,
                             int32_t nranges, const int64_t* range_ptr, double* stats,
                             uint32_t flags);

/* Innovation negative log-likelihood of the local analysis and its gradient
   (a hand-written adjoint; no counterpart in OAK, which has no
   differentiation: it serves the tuning of the analysis hyperparameters).
   With the selection and taper of oakg_local_analysis, per zone z:
     J_z = 1/2 [ d^T D^-1 d + log det D ],  D = a^2 S S^T + diag(sigma^2 / w^2)
   over the zone's observations (d = yo - Hxf, S = HSf rows, a = inflation),
   evaluated through the ensemble-space eigendecomposition.  Outputs:
     nll             sum_z J_z (1 value)
     grad_log_sigma  dJ/dlog sigma_l, sigma_l^2 = obs->rmse2[l]   [m]
     grad_corr_len   dJ/dL_z for the zone lengths zones->corr_len [nzones]
                     (the cutoffs zones->max_len are held fixed)
     grad_inflation  dJ/da (1 value)
     nll_zone        J_z [nzones] (may be NULL)
   grad_log_sigma is accumulated with atomic adds (order-dependent to the
   last bits); the other outputs are deterministic. */
int oakg_local_nll(oakg_ctx* ctx, const oakg_zones* zones, const oakg_obs* obs, int32_t N,
                   const double* Hxf, const double* HSf, int64_t ldHSf, double inflation,
                   double* nll, double* grad_log_sigma, double* grad_corr_len,
                   double* grad_inflation, double* nll_zone, uint32_t flags);

/* A model variable's grid for the observation operator (ModelGrid(v) of
 * assimilation.F90:335-361, with explicit coordinates as OAK's grid type
 * keeps them, and the memory layout of the variable in the state vector). */
typedef struct {
  int32_t ndim;               /* 1..4 (ModML%ndim(v))                                */
  int64_t shape[4];           /* nodes per dimension, at least 2 each (varshape)     */
  const double* coord[4];     /* [prod(shape)] node coordinate along each dimension, */
                              /*   first index fastest (the files of Model.gridX/Y/Z/T) */
  const uint8_t* masked;      /* [prod(shape)] 1 = land / invalid node, or NULL      */
  const int64_t* state_index; /* [prod(shape)] 0-based column of H for each node: the */
                              /*   state index after SeaIndex and the zone permutation */
                              /*   (assimilation.F90:1781-1814, 2404-2409), -1 = not in */
                              /*   the state; NULL = the node index itself              */
} oakg_grid;

/* genObservationOper (assimilation.F90:2471-2656) + packSparseMatrix
 * (:4132-4208) + the zone permutation of H's columns (:2404-2409) on the
 * GPU: observation l of model variable obs_var[l] (0-based; -1 or >= nvar:
 * unknown) at (ox, oy, oz, ot)[l] (only the first ndim are read) is located
 * in its grid by OAK's box search (ndgrid_inc.F90:921-990) and interpolated
 * by cinterp's hyper-tetrahedral weights (ndgrid.F90:357-665, 1183-1257):
 * row l of H holds the 2^ndim corners of its cell (cinterp's order), or
 * nothing when it is outside the grid, next to a masked node or of an
 * unknown variable.  Entries of nodes outside the state are dropped
 * (packSparseMatrix).  Outputs: row_ptr[m+1] (CSR by observation), up to cap
 * entries col/val, valid[m] (0: OAK excludes the observation, validobs of
 * loadObservationOper :2382-2387) and *nnz (the entries H holds; with
 * *nnz > cap nothing past cap is written and OAKG_EARG is returned).
 * Cells whose simplices are degenerate (|det| <= 1e-8, e.g. a dimension of
 * one node; OAK's SVD branch, ndgrid.F90:514-627) return OAKG_EUNSUP. */
int oakg_obs_operator(oakg_ctx* ctx, int32_t nvar, const oakg_grid* grids, int64_t m,
                      const int32_t* obs_var, const double* ox, const double* oy,
                      const double* oz, const double* ot, int64_t* row_ptr, int32_t* col,
                      double* val, int64_t cap, uint8_t* valid, int64_t* nnz, uint32_t flags);

/* The zone-permuted state and the model's field files (loadVector /
 * loadEnsemble / saveVector with ModML%permute, assimilation.F90:869-880,
 * 901, 1268-1359): with idx[l] = invindex(zoneIndex(l)) (0-based; the full
 * field index, land included, of permuted state row l),
 *   oakg_gather_rows:  out[l + k ldo] = in[idx[l] + k ldi]
 *   oakg_scatter_rows: out[idx[l] + k ldo] = in[l + k ldi]
 * for l < n and members k < N; elem_size 4 (FP32) or 8 (FP64); full_len is
 * the number of rows of the full field array (every idx must lie in
 * [0, full_len), else OAKG_EARG); rows of the scatter's out that no idx
 * names are left as they are (land values). */
int oakg_gather_rows(oakg_ctx* ctx, int64_t n, int32_t N, const int64_t* idx, int64_t full_len,
                     const void* in, int64_t ldi, void* out, int64_t ldo, int32_t elem_size,
                     uint32_t flags);
int oakg_scatter_rows(oakg_ctx* ctx, int64_t n, int32_t N, const int64_t* idx, int64_t full_len,
                      const void* in, int64_t ldi, void* out, int64_t ldo, int32_t elem_size,
                      uint32_t flags);

/* Host helpers ------------------------------------------------------------ */
/* Stable COO -> CSR by row (entries of a row keep their COO order). */
int oakg_coo_to_csr(int64_t m, int64_t nnz, const int32_t* row, const int32_t* col,
                    const double* val, int64_t* row_ptr, int32_t* col_out, double* val_out);
/* Zone layout from a partition vector: returns nzones (or -1), fills
 * zone_size[nzones], zone_index[n] (permuted position -> original index,
 * stable) and inv_zone_index[n]. */
int64_t oakg_partition(int64_t n, const int64_t* partition, int64_t* zone_size,
                       int64_t* zone_index, int64_t* inv_zone_index);

#ifdef __cplusplus
}
#endif
#endif /* OAKG_H */

/* qmcx.h — C ABI of the MI355X-native VMC particle-move engine (libqmcx.so).
 *
 * Plain C types only: opaque handles, host pointers and sizes.  Every call
 * returns 0 on success or a negative qmcx_status; qmcx_last_error() gives the
 * message of the last failure on the calling thread.  Host buffers are
 * borrowed for the duration of the call.  A batch is driven by one HIP stream;
 * calls on one batch must not overlap.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the miniqmc source tree).
 */
#ifndef QMCX_H
#define QMCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qmcx_system_s* qmcx_system_t;
typedef struct qmcx_batch_s* qmcx_batch_t;
typedef struct qmcx_group_s* qmcx_group_t;

enum qmcx_status {
  QMCX_OK = 0,
  QMCX_ERR_ARG = -1,     /* invalid argument or size */
  QMCX_ERR_HIP = -2,     /* HIP runtime error */
  QMCX_ERR_STATE = -3,   /* call out of order (e.g. no spline table yet) */
  QMCX_ERR_NOMEM = -4
};

enum qmcx_precision {
  QMCX_FP64 = 0,   /* all-FP64: spline table, distances, Jastrows, determinant */
  QMCX_FP32 = 1    /* FP32 spline/distance/Jastrow/Woodbury, FP64 LU recompute, log-det and acceptance */
};

enum qmcx_accept_mode {
  QMCX_ACCEPT_REFERENCE = 0,  /* accept when u < 0.5 (src/Drivers/miniqmc.cpp:445) */
  QMCX_ACCEPT_METROPOLIS = 1  /* accept when u < |Psi'/Psi|^2 */
};

const char* qmcx_last_error(void);
int qmcx_device_count(int* count);

/* ---- system: lattice, ions, Jastrow functors, SPO table --------------------
 * replaces build_ions/tile_cell (src/Particle/ParticleSet_builder.cpp:23,
 * src/Input/nio.hpp:23), buildJ1/buildJ2 (src/Input/nio.hpp:82,127) and
 * build_SPOSet + einspline_spo::set (src/QMCWaveFunctions/SPOSet_builder.cpp:19,
 * src/QMCWaveFunctions/einspline_spo.hpp:128-160).
 * lattice9: rows are the cell vectors (bohr).  ion_pos: [nions][3] Cartesian,
 * grouped by species.  nup == ndn is required. */
int qmcx_system_create(qmcx_system_t* out, const double* lattice9, int nions, const double* ion_pos,
                       const int* ion_species, int nspecies, int nup, int ndn);
int qmcx_system_destroy(qmcx_system_t sys);

/* BsplineFunctor::setupParameters (src/QMCWaveFunctions/Jastrow/BsplineFunctor.h:120-130)
 * for the J1 functor of one ion species (OneBodyJastrow::addFunc, OneBodyJastrow.h:101-106). */
int qmcx_system_set_j1(qmcx_system_t sys, int species, int nparams, double rcut, double cusp, const double* params);
/* J2 functor of a spin pair: 0 = same spin (uu/dd), 1 = opposite spin (ud/du)
 * (TwoBodyJastrow::addFunc, TwoBodyJastrow.h:204-240). */
int qmcx_system_set_j2(qmcx_system_t sys, int pair, int nparams, double rcut, double cusp, const double* params);

/* Three-body electron-electron-ion Jastrow (ThreeBodyJastrow::addFunc, src/QMCWaveFunctions/Jastrow/
 * ThreeBodyJastrow.h:118-152, with PolynomialFunctor3D::resize/reset_gamma, PolynomialFunctor3D.h:65-260;
 * miniqmc's -j option, src/Input/nio.hpp:170-244): the functor of ion species `species` and spin pair
 * `pair` (0 = same, 1 = opposite), polynomial degrees neI, nee in [1, 3], cutoff rcut (electrons count
 * within rcut/2 of the ion), params[nparams] the free parameters left by the no-cusp constraints
 * (26 for (3, 3); QMCX_ERR_ARG names the count when it differs).  All functors of one species take one
 * rcut; at most 256 ions.  Set it before qmcx_batch_create: a batch made earlier returns QMCX_ERR_STATE.
 * With a three-body Jastrow, qmcx_flex_evaluate_ratios and qmcx_nlpp_evaluate include its factor
 * (ThreeBodyJastrow::evaluateRatios, ThreeBodyJastrow.h:262-273), and an electron position may have at
 * most 6 ions within the cutoff (else QMCX_ERR_STATE). */
int qmcx_system_set_j3(qmcx_system_t sys, int species, int pair, int neI, int nee, double rcut, const double* params,
                       int nparams);

/* Spline table upload: coefs is [ng0+3][ng1+3][ng2+3][npad] FP32, orbitals
 * innermost (BsplineAllocator::allocateMultiBspline layout,
 * src/Numerics/Spline2/BsplineAllocator.hpp:84-136).  norb must equal nup.
 * spo_lattice9 is the lattice the SPO maps positions with (einspline_spo::Lattice).
 * precision_mask: bit0 keeps an FP32 device table, bit1 an FP64 copy. */
int qmcx_system_set_spline(qmcx_system_t sys, const int* ng, int norb, int npad, const float* coefs,
                           const double* spo_lattice9, int precision_mask);

/* ---- stand-alone kernels (parity surface) ----------------------------------
 * einspline_spo::evaluate_vgh -> MultiBsplineEval::evaluate_vgh
 * (src/QMCWaveFunctions/einspline_spo.hpp:194-202, src/Numerics/Spline2/MultiBspline.hpp:163-280):
 * npts Cartesian points -> out[npts][10][norb] = v, gx, gy, gz, hxx, hxy, hxz, hyy, hyz, hzz. */
int qmcx_spline_vgh(qmcx_system_t sys, int precision, int npts, const double* r, double* out);

/* DTD_BConds::computeDistances (src/Particle/Lattice/ParticleBConds.h:154-217): rows from
 * npts sources to the nt targets, flip index per source; out r[npts][nt], dr[npts][nt][3]. */
int qmcx_distance_rows(qmcx_system_t sys, int precision, int npts, const double* src, const int* flip_ind,
                       int nt, const double* targets, double* r, double* dr);

/* DiracMatrix::invert_transpose (src/QMCWaveFunctions/DiracMatrix.h:93-121), batched FP64 LU with
 * partial pivoting: a[nmat][n][n] (rows = electrons) -> psiM[nmat][n][n] = (a^-1)^T,
 * logdet[nmat] = log|det a|, phase[nmat] in {0, pi}. */
int qmcx_invert_transpose(int n, int nmat, const double* a, double* psiM, double* logdet, double* phase);

/* qmcx_invert_transpose on caller-owned DEVICE buffers that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernel and results as qmcx_invert_transpose.
 * d_a[nmat][n][n] -> d_psiM[nmat][n][n], d_logdet[nmat], d_phase[nmat]; asynchronous on `stream` (a
 * hipStream_t; NULL = the null stream).  The elimination workspace is a 32 MiB device array owned by
 * the library, one per GPU, held until the process ends; it takes floor(2^21 / n^2) matrices per launch
 * and larger batches run as several launches in order.  n <= 1024, else QMCX_ERR_ARG.  Calls that
 * overlap on one GPU (other threads or streams) share the array and race. */
int qmcx_invert_transpose_noalloc(int n, int nmat, const double* d_a, double* d_psiM, double* d_logdet,
                                  double* d_phase, void* stream);

/* ---- walker batch -----------------------------------------------------------
 * Device-resident walker state for nwalkers walkers (ParticleSet + WaveFunction
 * per Mover, src/Drivers/Mover.hpp:46-64; DiracDeterminant<DU_TYPE>,
 * src/QMCWaveFunctions/DiracDeterminant.h:35-165, with the DelayedUpdate engine
 * src/QMCWaveFunctions/DelayedUpdate.h:29-232 as its DU_TYPE).
 * delay_rank k >= 1 (k = 1 is Sherman-Morrison). */
int qmcx_batch_create(qmcx_batch_t* out, qmcx_system_t sys, int nwalkers, int delay_rank, int precision);
int qmcx_batch_destroy(qmcx_batch_t b);
int qmcx_batch_set_positions(qmcx_batch_t b, const double* R /* [nw][N][3] */);
int qmcx_batch_get_positions(qmcx_batch_t b, double* R);
/* global walker index of walker 0 of this batch (selects the random streams) */
int qmcx_batch_set_walker_offset(qmcx_batch_t b, int walker0);
/* spline stencil slices per workgroup (autotuner knob; 0 = default) */
int qmcx_batch_set_tile(qmcx_batch_t b, int spline_slices);
/* Delayed-update flush engine (DelayedUpdate::updateInvMat, src/QMCWaveFunctions/DelayedUpdate.h:176-193):
 * 0 = FP32 matrix cores (v_mfma_f32_16x16x4f32) when the batch is FP32 and n % 16 == 0, 1 = vector FMA,
 * 2 = matrix cores with the new Slater rows read from HBM rather than staged in LDS. */
int qmcx_batch_set_flush(qmcx_batch_t b, int kind);
/* Form of the delayed updates (DelayedUpdate.h:96-231; no reference counterpart for the choice):
 * 0 = automatic: a chunk's stale inverse rows, the capacitance inverse and the products U_d . Ainv[row]
 * stay in LDS (the in-kernel flush when it fits), or the streamed form when the delay rank's rows do not
 * fit the LDS (e.g. k = 64 and 128 at 384 orbitals); 1 = streamed: the stale rows are read from the
 * inverse in HBM, p = U . Ainv[row] is formed at each move as getInvRow does, the capacitance inverse and
 * the row list live in the chunk's delayed-update state, and a flush launch applies each chunk.  The flex_*
 * calls always use the streamed form with k > 1 and keep each walker's accepted rows pending across calls
 * until it has k of them or qmcx_flex_complete_updates.  get reports the form the batch's sweeps use. */
int qmcx_batch_set_delay_form(qmcx_batch_t b, int form);
/* FP64 recompute cadence (no reference counterpart; the FP32 form's "FP64 LU inverse recomputation every
 * rows accepted rows per walker"): with rows > 0 each walker's inverse of a determinant is rebuilt in FP64
 * from its Slater rows (DiracMatrix::invert_transpose, DiracMatrix.h:93-121) after every `rows` accepted
 * moves of that determinant, and its log|det| and phase are replaced by the FP64 values; the walker's
 * pending delayed rows are dropped there rather than flushed.  The count runs across sweeps and restarts at
 * every FP64 inversion (evaluateLog, the per-sweep refresh of qmcx_run_sweeps).  0 = off.  Needs
 * Sherman-Morrison moves or the in-kernel flush (FP32, n a multiple of 16 up to 384, k <= 32). */
int qmcx_batch_set_recompute_rows(qmcx_batch_t b, int rows);
int qmcx_batch_get_delay_form(qmcx_batch_t b, int* form);

/* WaveFunction::evaluateLog (src/QMCWaveFunctions/WaveFunction.cpp:178-195): SPO for all
 * electrons, FP64 LU, J1/J2 from scratch.  log_out[nw] (may be NULL). */
int qmcx_evaluate_log(qmcx_batch_t b, double* log_out);

/* The VMC sweep (src/Drivers/miniqmc.cpp:422-467 and the flex_* batched form,
 * src/Drivers/miniqmc_sync_move.cpp:459-513): for every electron setActive,
 * evalGrad, makeMove(sqrt(tau) * N(0,1)), ratioGrad, accept/reject, then
 * completeUpdates and evaluateGL.  Random numbers: Philox4x64-10 keyed by seed,
 * counter (electron, sweep, walker).  recompute_every > 0 refreshes every
 * inverse with FP64 LU after every that many sweeps.
 * log_out [nsweeps][nw] LogValue after each sweep (may be NULL). */
int qmcx_run_sweeps(qmcx_batch_t b, int nsweeps, uint64_t seed, double tau, int accept_mode, int recompute_every,
                    double* log_out);

/* Graph replay of qmcx_run_sweeps (no reference counterpart; the reference issues every kernel from the
 * host): with enable != 0 a run is captured into a HIP graph on its first call (the three streams' launches
 * and event dependencies of all its sweeps) and replayed by later calls with the same nsweeps, seed, tau,
 * acceptance rule and refresh cadence; other parameters capture a new graph.  The sweep numbers of the
 * random streams and of the log history come from device counters, so a replay continues the chains
 * exactly as the launches would.  Changing a batch setting (trace, flush kind, tile, band, walker offset)
 * drops the graph.  Not with kernel timing on. */
int qmcx_batch_set_graph(qmcx_batch_t b, int enable);

/* ---- flex_* stage calls (the batched boundary of src/Drivers/miniqmc_sync_move.cpp:459-513) ----
 * One electron of every walker per call.
 * qmcx_flex_ratio_grad: ParticleSet::flex_setActive + WaveFunction::flex_evalGrad + ParticleSet::flex_makeMove
 *   + WaveFunction::flex_ratioGrad (src/Particle/ParticleSet.cpp:296-346, src/QMCWaveFunctions/WaveFunction.cpp:343-417):
 *   displ[nw][3] -> ratio[nw], grad_now[nw][3], grad_new[nw][3] (outputs may be NULL); no state changes.
 * qmcx_flex_accept_restore: WaveFunction::flex_acceptrestoreMove + ParticleSet::acceptMove/rejectMove
 *   (WaveFunction.cpp:418-485, miniqmc_sync_move.cpp:496-505) for the same electron, accepted[nw] != 0.
 * qmcx_flex_complete_updates: WaveFunction::flex_completeUpdates (WaveFunction.cpp:486-499): the delayed rows
 *   that flex moves left pending reach the inverse (DiracDeterminant.cpp:294-321).  With delay rank k > 1 an
 *   accepted flex move is a pending delayed row of its walker until the walker has k of them, as in
 *   DelayedUpdate::acceptRow (DelayedUpdate.h:151-152); calls that read the inverse (evaluate_gl, ratios,
 *   NonLocalPP, run_sweeps, get_inverse) apply pending rows first.
 * qmcx_flex_evaluate_gl: WaveFunction::flex_evaluateGL; log_out[nw] (G, L via qmcx_get_gl). */
int qmcx_flex_ratio_grad(qmcx_batch_t b, int iel, const double* displ, double* ratio, double* grad_now,
                         double* grad_new);
int qmcx_flex_accept_restore(qmcx_batch_t b, int iel, const int* accepted);
int qmcx_flex_complete_updates(qmcx_batch_t b);
int qmcx_flex_evaluate_gl(qmcx_batch_t b, double* log_out);
/* WaveFunction::evaluateRatios (src/QMCWaveFunctions/WaveFunction.cpp:281-297) batched over walkers: the
 * virtual-move ratios Psi(r_iel -> vpos[w][k]) / Psi for electron iel of every walker, as the non-local
 * pseudopotential's quadrature asks for them (src/Drivers/NonLocalPP.hpp:115-140).  Value-only SPO row dotted
 * with the inverse row (DiracDeterminant.cpp:216-224, SPOSet.h:72-83) times the J1 and J2 ratios
 * (OneBodyJastrow.h:136-140, TwoBodyJastrow.h:103-107).  vpos[nw][nvp][3] -> ratios[nw][nvp]; no state
 * changes; 1 <= nvp <= 65536. */
int qmcx_flex_evaluate_ratios(qmcx_batch_t b, int iel, int nvp, const double* vpos, double* ratios);
/* NonLocalPP::evaluate (src/Drivers/NonLocalPP.hpp:115-140; driver src/Drivers/miniqmc.cpp:470-495) for every
 * electron of every walker: for each ion within rmax of the electron (minimum image, ions in index order) the twelve
 * ratios at dist * spheres[w][k] + displ + r_e, with spheres[nw][12][3] each walker's rotated quadrature knots of
 * this evaluation (the reference rotates per mover and evaluation, NonLocalPP.hpp:100-113, miniqmc.cpp:470).
 * ratios[nw][N][max_pairs][12] and npairs[nw][N] (either may be NULL; unused slots are left as they were);
 * 1 <= max_pairs <= 64; QMCX_ERR_ARG when an electron has more than max_pairs ions inside rmax. */
int qmcx_nlpp_evaluate(qmcx_batch_t b, double rmax, const double* spheres, int max_pairs, double* ratios,
                       int* npairs);
/* Tuning knob of qmcx_nlpp_evaluate: knots of one electron-ion pair per work unit (1, 2, 4 or 6; 0 = chosen by
 * orbital count and precision).  QMCX_ERR_ARG at evaluation when the unit's value rows do not fit a workgroup. */
int qmcx_batch_set_nlpp_unit(qmcx_batch_t b, int knots_per_unit);

/* Per-move record of the most recent sweep, [nw][N][9]:
 * ratio, det ratio, accepted, grad_now[3], grad_new[3].  enable must be set before the sweep. */
int qmcx_set_trace(qmcx_batch_t b, int enable);
int qmcx_get_trace(qmcx_batch_t b, double* trace);

/* Identical accept/reject decisions in the FP32 form under Metropolis acceptance (no reference
 * counterpart: the reference is FP64 throughout).  With band > 0 a move whose uniform u satisfies
 * |u - |R|^2| <= band |R|^2 for the FP32 ratio R is decided on its FP64 ratio: FP64 Slater rows with
 * the determinant ratio by iterative refinement from the FP32 inverse, J1/J2 (and J3) from scratch.  With a
 * three-body Jastrow the re-evaluation needs LDS beside its pair slots (QMCX_ERR_ARG names the need
 * when the delay rank leaves too little).  While the
 * band is on, the batch also keeps the FP64 Slater rows (the FP64 refresh then inverts those instead of
 * the FP32 rows); they are filled by the next qmcx_evaluate_log, which must come before the next
 * qmcx_run_sweeps (else QMCX_ERR_STATE).  band in [0, 1]; 0 (the default) switches it off.  FP64
 * batches ignore it.  qmcx_get_arbitrations: per-walker count of such decisions since the batch was
 * created. */
int qmcx_batch_set_accept_band(qmcx_batch_t b, double band);
int qmcx_get_arbitrations(qmcx_batch_t b, long long* narb);

/* FP64 Slater rows in the FP32 form without the band (no reference counterpart): the proposals also
 * sum the FP64 spline values, accepted moves store them, and the FP64 refresh inverts those, so
 * log|Psi| after a refresh carries no FP32 rounding of the matrix, however ill-conditioned.  As with
 * the band, the next qmcx_evaluate_log fills them before the next qmcx_run_sweeps.  enable 0 (the
 * default) or 1; FP64 batches ignore it. */
int qmcx_batch_set_fp64_rows(qmcx_batch_t b, int enable);

/* Outputs of the last evaluateGL: G[nw][N][3], L[nw][N], log[nw]; per-walker accept counts. */
int qmcx_get_gl(qmcx_batch_t b, double* G, double* L, double* log_out);
int qmcx_get_accepts(qmcx_batch_t b, long long* naccept);

/* Inverse matrix psiM of one determinant of one walker as FP64, [n][n] (DiracDeterminant::psiM). */
int qmcx_get_inverse(qmcx_batch_t b, int walker, int det, double* psiM);

/* Per-stage device time: with timing enabled, HIP events on the batch's streams bracket every
 * launch of qmcx_run_sweeps.  Index 0: move kernel; 1: delayed-update flush; 2: FP64 inverse
 * recompute; 3: evaluateGL; 4: SPO evaluation at the proposals.  ms5[i] accumulates
 * milliseconds and launches5[i] the number of bracketed launches.  Stages 2 and 4 run on their
 * own streams, beside the moves.  Enabling resets the totals. */
int qmcx_set_kernel_timing(qmcx_batch_t b, int enable);
int qmcx_get_kernel_timing(qmcx_batch_t b, double* ms5, long long* launches5);

/* ---- delayed-update engine of one determinant (DU_TYPE) --------------------
 * The plug point DiracDeterminant<DU_TYPE> (src/QMCWaveFunctions/DiracDeterminant.h:35,127,
 * instantiated for a GPU engine at DiracDeterminant.cpp:339-341) with the DelayedUpdate interface
 * (src/QMCWaveFunctions/DelayedUpdate.h:60-231).  The inverse psiM lives on the device; host
 * matrices are row-major n x n doubles.
 *   create        DelayedUpdate::resize(norb, delay)                         DelayedUpdate.h:65-74
 *   invert_transpose  DelayedUpdate::invert_transpose -> DiracMatrix      DelayedUpdate.h:75-82, DiracMatrix.h:93-121
 *                 logdetT[i][j] = phi_j(r_i); Ainv receives (logdetT^-1)^T (may be NULL)
 *   set_inverse   DelayedUpdate::initializeInv                               DelayedUpdate.h:87-91
 *   get_inv_row   DelayedUpdate::getInvRow                                   DelayedUpdate.h:96-114
 *   accept_row    DelayedUpdate::acceptRow (flushes when delay is reached)   DelayedUpdate.h:124-153
 *   update_inv_mat DelayedUpdate::updateInvMat                               DelayedUpdate.h:158-231
 * As in the reference, acceptRow expects getInvRow of the same row just before it. */
typedef struct qmcx_du_s* qmcx_du_t;
int qmcx_du_create(qmcx_du_t* out, int norb, int delay, int precision);
int qmcx_du_destroy(qmcx_du_t h);
int qmcx_du_invert_transpose(qmcx_du_t h, const double* logdetT, double* Ainv, double* log_value, double* phase);
int qmcx_du_set_inverse(qmcx_du_t h, const double* Ainv);
int qmcx_du_get_inverse(qmcx_du_t h, double* Ainv);
int qmcx_du_get_inv_row(qmcx_du_t h, int row, double* inv_row);
int qmcx_du_accept_row(qmcx_du_t h, int row, const double* psiV);
int qmcx_du_update_inv_mat(qmcx_du_t h);
int qmcx_du_delay_count(qmcx_du_t h, int* count);

/* Device time of the last qmcx_run_sweeps (HIP events on the batch stream), milliseconds. */
int qmcx_last_sweep_ms(qmcx_batch_t b, double* ms);

/* ---- mixed-size groups (no reference counterpart: the reference advances one system's walkers per
 * process; miniqmc_sync_move.cpp:459-513 is the batched step this generalizes) ----------------------
 * Walkers of up to 4 batches, each of its own system (its own spline table, Jastrows and size) and delay
 * rank, advanced together: every move stage and every proposals stage is one launch over all the
 * batches' walkers (the workgroup size and LDS are the largest the batches need); the FP64 refresh runs
 * one launch per batch (matrix size), evaluateGL per batch.  Batches: FP32, delay rank >= 2 with the
 * delayed updates applied inside the move kernel (n a multiple of 16, rank <= 32), no three-body Jastrow,
 * no near-tie band, all with FP64 Slater rows (qmcx_batch_set_fp64_rows) or none, one device.  The
 * batches stay usable on their own; each keeps its own state,
 * random streams (its walker offset) and log history.
 * qmcx_group_run_sweeps: as qmcx_run_sweeps for every batch at once; log_out [nsweeps][sum of nw] (the
 * batches' walkers in group order, may be NULL).  qmcx_group_set_graph: as qmcx_batch_set_graph. */
int qmcx_group_create(qmcx_group_t* out, const qmcx_batch_t* batches, int nbatches);
int qmcx_group_destroy(qmcx_group_t g);
int qmcx_group_set_graph(qmcx_group_t g, int enable);
int qmcx_group_run_sweeps(qmcx_group_t g, int nsweeps, uint64_t seed, double tau, int accept_mode,
                          int recompute_every, double* log_out);
int qmcx_group_last_ms(qmcx_group_t g, double* ms);

#ifdef __cplusplus
}
#endif
#endif

/*
 * goofit_hip.h -- C ABI of the MI355X (gfx950) likelihood engine.
 *
 * This is the drop-in boundary that sits under the GooFit C++ API
 * (GooPdf / PdfBase / Amp3Body / FitManager).  Every entry point takes plain
 * pointers and sizes; no C++ exception and no torch type crosses it.  Each
 * declaration names the reference interface it replaces (paths relative to
 * the GooFit 2.4.0 source tree).
 *
 * Data model, identical to the reference's flattened PDF tree
 * (src/PDFs/PdfBase.cu:268-311, include/goofit/PDFs/ParameterContainer.h:9-47):
 *   params[], consts[], obs[], norms[]  -- per function a block [count, v0, v1, ...]
 *                                          (norms blocks are [1, 1/integral])
 *   funcs[]                             -- one gf_func code per function, in the
 *                                          depth-first order of the tree walk
 *                                          (replaces host_function_table's device
 *                                          function pointers, src/PDFs/detail/Globals.cu:13)
 * Events are uploaded as structure-of-arrays columns, column k holding the
 * observable whose GooFit index is k (src/goofit/PdfBase.cpp:109-117).
 *
 * Errors: every call returns gf_status; gf_last_error() gives the message of the
 * last failing call on the calling thread.  Threading: one gf_ctx per host
 * thread; calls on one context are not re-entrant (the reference is not either,
 * src/PDFs/detail/Globals.cu:9-34).
 */
#ifndef GOOFIT_HIP_H
#define GOOFIT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gf_status;
#define GF_OK 0
#define GF_ERR_INVALID 1  /* bad argument / program */
#define GF_ERR_HIP 2      /* HIP runtime error */
#define GF_ERR_NOMEM 3    /* device allocation failed */
#define GF_ERR_STATE 4    /* call order violated (e.g. no events uploaded) */
#define GF_ERR_NODEVICE 5 /* no GPU visible */

/* Function codes.  Each replaces one registered device function pointer. */
enum gf_func {
    GF_FN_NONE              = 0,
    GF_FN_GAUSSIAN          = 1,  /* device_Gaussian        src/PDFs/basic/GaussianPdf.cu:7-17 */
    GF_FN_EXP               = 2,  /* device_Exp             src/PDFs/basic/ExpPdf.cu:7-17 */
    GF_FN_EXP_OFFSET        = 3,  /* device_ExpOffset       src/PDFs/basic/ExpPdf.cu:19-30 */
    GF_FN_EXP_POLY          = 4,  /* device_ExpPoly         src/PDFs/basic/ExpPdf.cu:32-48 */
    GF_FN_EXP_POLY_OFFSET   = 5,  /* device_ExpPolyOffset   src/PDFs/basic/ExpPdf.cu:50-67 */
    GF_FN_EXPGAUS           = 6,  /* device_ExpGaus         src/PDFs/basic/ExpGausPdf.cu:6-24 */
    GF_FN_POLYNOMIAL        = 7,  /* device_Polynomial      src/PDFs/basic/PolynomialPdf.cu:8-25 */
    GF_FN_OFFSET_POLYNOMIAL = 8,  /* device_OffsetPolynomial src/PDFs/basic/PolynomialPdf.cu:28-45 */
    GF_FN_MULTI_POLYNOMIAL  = 9,  /* device_MultiPolynomial src/PDFs/basic/PolynomialPdf.cu:47-118 */
    GF_FN_ADD               = 10, /* device_AddPdfs         src/PDFs/combine/AddPdf.cu:16-51 */
    GF_FN_ADD_EXT           = 11, /* device_AddPdfsExt      src/PDFs/combine/AddPdf.cu:53-79 */
    GF_FN_PROD              = 12, /* device_ProdPdfs        src/PDFs/combine/ProdPdf.cu:9-26 */
    GF_FN_DALITZ            = 13, /* device_DalitzPlot      src/PDFs/physics/Amp3Body.cu:60-122 */
    GF_FN_RBW               = 14, /* plainBW<1>             src/PDFs/physics/resonances/RBW.cu:9-86 */
    GF_FN_RBW_SYM           = 15, /* plainBW<2>             src/PDFs/physics/resonances/RBW.cu:89 */
    GF_FN_GS                = 16, /* gouSak                 src/PDFs/physics/resonances/GS.cu:9-75 */
    GF_FN_NONRES            = 17, /* nonres                 src/PDFs/physics/resonances/NonRes.cu:8-11 */
    GF_FN_TDDP              = 18, /* device_Tddp            src/PDFs/physics/Amp3Body_TD.cu:41-257 */
    GF_FN_TRUTH_RES         = 19, /* device_truth_resolution src/PDFs/physics/TruthResolution.cu:6-27 */
    GF_FN_THREEGAUSS_RES    = 20, /* device_threegauss_resolution src/PDFs/physics/ThreeGaussResolution.cu:63-127 */
    GF_FN_COUNT             = 21
};

/* Metric codes: same numbering as the reference's EvalFunc
 * (include/goofit/FitControl.h:10); functions in src/PDFs/MetricPointer.cu:8-58. */
enum gf_metric {
    GF_METRIC_EVAL         = 0,
    GF_METRIC_NLL          = 1,
    GF_METRIC_PROB         = 2,
    GF_METRIC_BINAVG       = 3,
    GF_METRIC_BINWITHERROR = 4,
    GF_METRIC_CHISQ        = 5
};

typedef struct gf_ctx gf_ctx;

/* ---- context ------------------------------------------------------------ */

/* Number of visible GPUs (0 when none).  No reference counterpart. */
int gf_device_count(void);

/* Create a context on `device`: one HIP stream plus the device buffers of the
 * flattened tree.  Replaces the process-global SmartVector mirrors
 * (include/goofit/detail/SmartVectorGPU.h, src/PDFs/detail/Globals.cu:9-13). */
gf_status gf_ctx_create(gf_ctx **out, int device);

/* Free every device buffer of the context.  Replaces GooFit::cleanup()
 * (src/PDFs/detail/Globals.cu:37-43) and gooFree (src/PDFs/PdfBase.cu:573-580). */
void gf_ctx_destroy(gf_ctx *ctx);

/* Message of the last failing call on this thread ("" if none). */
const char *gf_last_error(void);

/* Launch-shape knobs of a context.  No reference counterpart: the reference
 * leaves the grid to thrust::transform_reduce (src/PDFs/GooPdf.cu:65-70,100-105);
 * an autotuner, or a caller sharing the device with other work, sets them. */
enum gf_tune {
    GF_TUNE_INLINE_ARGS        = 0, /* 1: parameter and normalisation values travel in the kernel
                                       arguments (no upload copy); 0: device buffers (default 1) */
    GF_TUNE_FAST_DALITZ        = 1, /* 1: dedicated coherent-sum kernel for a lone Amp3Body (default 1) */
    GF_TUNE_BLOCKS_PER_CU      = 2, /* interpreter reduction grid cap, workgroups per CU (1..64, default 4) */
    GF_TUNE_FAST_BLOCKS_PER_CU = 3, /* coherent-sum kernel grid cap, workgroups per CU (1..64, default 3) */
    GF_TUNE_SPIN_US            = 4, /* a reduction's caller polls the result for up to this many
                                       microseconds before blocking on the stream (0..1e6, default 2000) */
    GF_TUNE_ZERO_OUTSIDE       = 5, /* 1: with a constant efficiency and a wave cache filled from the
                                       current events, the coherent-sum kernel takes |sum c_i W_i|^2 as 0
                                       outside the Dalitz region from the cache (which holds 0 there)
                                       instead of reading m12 and m13 (default 1) */
    GF_TUNE_FLAT               = 6, /* a tree that is one leaf, or one product or sum of 1-D leaves
                                       (Gaussian, Exp, ExpGaus) over at most 4 columns, runs a dedicated
                                       kernel that loads every column of a tile before evaluating it, with
                                       2 or 4 events per lane (the value), or 1: 4 from 2^20 events on,
                                       2 below; 0: the general interpreter (default 1) */
    GF_TUNE_FUSED_FINISH       = 7, /* 1: the coherent-sum and flat reductions add their workgroups'
                                       partials inside the sweep (the last workgroup to finish does it,
                                       in k_finish's order: the same bits) instead of in a second launch;
                                       0: a second launch (default 1) */
    GF_TUNE_COH_MFMA           = 8, /* 1: the multi-point NLL of a lone Amp3Body with a constant
                                       efficiency (gf_reduce_metric_multi) forms its coherent sums on the
                                       FP64 matrix cores -- the same bits as on the vector ALUs; 0: on the
                                       vector ALUs (default 1) */
    GF_TUNE_BATCH_MX_PTS       = 9, /* toy batches of a lone Amp3Body's NLL with a constant efficiency:
                                       a toy's group of at least this many points (of 16) forms its
                                       coherent sums on the FP64 matrix cores, a smaller one on the vector
                                       ALUs (2..17, 17: never); 0: groups of 4 or more when the model has
                                       more than 8 resonances, else never (default 0) */
    GF_TUNE_COUNT              = 10
};

/* Set / read one knob (gf_tune); values out of range give GF_ERR_INVALID. */
gf_status gf_ctx_set_tuning(gf_ctx *ctx, int key, int value);
gf_status gf_ctx_get_tuning(gf_ctx *ctx, int key, int *value);

/* ---- events ------------------------------------------------------------- */

/* Upload n events as nobs host columns (cols[k] = observable with index k).
 * Replaces the AoS pack + gooMalloc + MEMCPY of PdfBase::setData
 * (src/PDFs/PdfBase.cu:329-437) and sets c_totalEvents (PdfBase.cu:438). */
gf_status gf_events_upload(gf_ctx *ctx, int nobs, int64_t n, const double *const *cols);

/* Same, from the reference's AoS host array host_array[i*nobs + k]
 * (src/PDFs/PdfBase.cu:398-403). */
gf_status gf_events_upload_aos(gf_ctx *ctx, int nobs, int64_t n, const double *aos);

/* Allocate nobs device columns of n events without host data (for events
 * generated on the GPU); column pointers via gf_events_column.
 * No reference counterpart (the reference generates on the host). */
gf_status gf_events_alloc(gf_ctx *ctx, int nobs, int64_t n);

/* Device address of column k (valid until the next upload/alloc). */
gf_status gf_events_column(gf_ctx *ctx, int k, double **dev_ptr);

/* Declare that column k holds the event number of each row and that it equals
 * the row index (0..n-1), as UnbinnedDataSet + EventNumber produce it.  The
 * Dalitz coherent sum then indexes its wave cache by row instead of reading the
 * column (src/PDFs/physics/Amp3Body.cu:82-84).  Checked on the host data when
 * the events came through gf_events_upload*. */
gf_status gf_events_set_identity_column(gf_ctx *ctx, int k);

/* ---- flattened PDF tree ------------------------------------------------- */

/* Install a flattened tree.  Replaces PdfBase::pre_run (src/PDFs/PdfBase.cu:24-31),
 * i.e. the sync of host_function_table/parameters/constants/observables/
 * normalizations to the device.  The tree is validated (counts walk exactly
 * as ParameterContainer::incrementIndex, src/PDFs/ParameterContainer.cu:48-60)
 * and compiled into a linear device program. */
gf_status gf_program_set(gf_ctx *ctx,
                         const int32_t *funcs, int nfunc,
                         const double *params, int np,
                         const double *consts, int nc,
                         const double *obs, int no,
                         const double *norms, int nn);

/* Replace the parameter values (same layout and size as at gf_program_set).
 * Replaces PdfBase::pre_call -> host_parameters.smart_sync
 * (src/PDFs/PdfBase.cu:33-36). */
gf_status gf_params_update(gf_ctx *ctx, const double *params, int np);

/* Replace the normalisation values.  Replaces host_normalizations.sync
 * (src/PDFs/GooPdf.cu:264). */
gf_status gf_norms_update(gf_ctx *ctx, const double *norms, int nn);

/* Three-body decay constants {motherMass, daug1Mass, daug2Mass, daug3Mass,
 * meson_radius, mother_meson_radius}.  Replaces the MEMCPY_TO_SYMBOL of
 * c_motherMass .. c_mother_meson_radius in the Amp3Body constructor
 * (src/PDFs/physics/Amp3Body.cu:146-151). */
gf_status gf_decay3_set(gf_ctx *ctx, const double decay[6]);

/* ---- reductions over events and grids ----------------------------------- */

/* Sum over all events of metric(pdf(evt), norm of the top node).  Replaces
 * GooPdf::reduce_with_metric (src/PDFs/GooPdf.cu:45-78) with the MetricTaker
 * unbinned operator (src/PDFs/MetricTaker.cu:17-44).  The sum is accumulated
 * with compensated FP64 arithmetic; *sum is the rounded total, *comp (may be
 * NULL) the residual correction term that was folded in. */
gf_status gf_reduce_metric(gf_ctx *ctx, int metric, double *sum, double *comp);

/* Result block of gf_reduce_metric_noalloc, written to device memory by the
 * reduction's last kernel: the compensated total, the residual correction that
 * was folded in, the device error word of the sweep (bit 0: an Amp3Body event
 * number outside its wave cache; bit 1: a node kind the kernel was not built
 * with; 0 = no error) and the call's sequence number on its context. */
typedef struct gf_result {
    double sum;
    double residual;
    int32_t error;
    uint32_t seq;
    int32_t pad[2];
} gf_result;

/* Largest number of per-workgroup partials the no-allocation reduction holds
 * (its static scratch is GF_NOALLOC_PARTIALS x 16 bytes per GPU). */
#define GF_NOALLOC_PARTIALS 4096

/* gf_reduce_metric without any allocation, for callers that cannot allocate
 * device or pinned memory while they run: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  No counterpart in the
 * reference.  The same kernels as gf_reduce_metric (identical sum, bit for bit)
 * are enqueued on `stream` (a hipStream_t; NULL: the context's stream) without
 * waiting; the result block goes to the caller's device memory *dev_out.  The
 * per-workgroup partials live in a fixed device array of the library
 * (GF_NOALLOC_PARTIALS pairs = 64 KiB on each GPU, set up by the first
 * gf_ctx_create on that GPU and held until the process ends), not in the
 * context.  Refused with GF_ERR_STATE: parameter or normalisation arrays larger
 * than the kernel arguments hold (256 and 128 values) or GF_TUNE_INLINE_ARGS
 * off, and a grid beyond the scratch.  Work the context queued on its own
 * stream (uploads, wave caches) must be complete, or `stream` must be that
 * stream.  All calls on one GPU share the scratch array: calls whose launches
 * overlap on one GPU (two streams, two contexts on the same device) race and
 * must be ordered by the caller. */
gf_status gf_reduce_metric_noalloc(gf_ctx *ctx, int metric, gf_result *dev_out, void *stream);

/* The metric sums for npts parameter sets over the same events: params holds
 * npts consecutive copies of the flattened parameter array (np values each, np
 * as in gf_params_update), norms npts copies of the normalisation array;
 * out[k] = the sum gf_reduce_metric returns with set k installed.  For a lone
 * Amp3Body (the coherent-sum path) one sweep serves 16 sets (32 when the NLL's
 * coherent sums run on the matrix cores, GF_TUNE_COH_MFMA), reading each event's
 * waves once, with each set's result bit-identical to its own gf_reduce_metric;
 * otherwise the sets are reduced one after another.  The
 * context's own arrays are left as they were.  No reference counterpart: it
 * serves the independent points of a numerical derivative
 * (Minuit2::FCNBase::EvaluateMany of include/Minuit2/Minuit2.h) in one request. */
gf_status gf_reduce_metric_multi(gf_ctx *ctx, int metric, int npts, const double *params, int np,
                                 const double *norms, int nn, double *out);

/* Per-event metric values into host array out[n] (n = number of events).
 * Replaces GooPdf::evaluate_with_metric (src/PDFs/GooPdf.cu:112-173). */
gf_status gf_eval_metric(gf_ctx *ctx, int metric, double *out, int64_t n);

/* Same, into a device buffer (no host copy). */
gf_status gf_eval_metric_device(gf_ctx *ctx, int metric, double *dev_out, int64_t n);

/* Sum of the raw value of function func_idx over the midpoint grid given by
 * ranges[3*i+0..2] = {lower, upper, nbins} of the function's k observables
 * (first observable fastest).  Replaces GooPdf::reduce_with_bins
 * (src/PDFs/GooPdf.cu:83-108) with the MetricTaker binned operator
 * (src/PDFs/MetricTaker.cu:48-96). */
gf_status gf_reduce_bins(gf_ctx *ctx, int func_idx, const double *ranges, int k, double *sum);

/* The per-bin raw values that gf_reduce_bins adds, into host array out[nbins]
 * (nbins = the grid's bin count; bin b in the first-observable-fastest order).
 * Replaces the transform step of GooPdf::reduce_with_bins over the MetricTaker
 * binned operator (src/PDFs/GooPdf.cu:83-108, src/PDFs/MetricTaker.cu:48-96). */
gf_status gf_eval_bins(gf_ctx *ctx, int func_idx, const double *ranges, int k, double *out, int64_t nbins);

/* The cell centres every grid kernel uses, into host array out[d * nbins + b]
 * (coordinate d of bin b; ranges and order as in gf_reduce_bins); each centre is
 * ((upper - lower) / n) * (local bin + 0.5) + lower, one rounding per operation.
 * Replaces the bin-centre arithmetic of MetricTaker (src/PDFs/MetricTaker.cu:73-86)
 * and GooPdf::makeGrid (src/PDFs/GooPdf.cpp:92-99). */
gf_status gf_grid_cells(gf_ctx *ctx, const double *ranges, int k, double *out, int64_t nbins);

/* ---- three-body Dalitz amplitude ---------------------------------------- */

/* Allocate the wave cache of Dalitz node dalitz_func for n entries, read from
 * event rows offset..offset+n-1.  Replaces Amp3Body::setDataSize
 * (src/PDFs/physics/Amp3Body.cu:204-236). */
gf_status gf_dalitz_set_data_size(gf_ctx *ctx, int dalitz_func, int64_t n, int64_t offset);

/* Recompute the cached lineshape W_i[e] of every resonance i with bit i set in
 * res_mask (0 outside the Dalitz region).  Replaces the per-resonance
 * thrust::transform with SpecialResonanceCalculator
 * (src/PDFs/physics/Amp3Body.cu:292-310, detail/SpecialResonanceCalculator.cu:12-46). */
gf_status gf_dalitz_cache(gf_ctx *ctx, int dalitz_func, uint64_t res_mask);

/* Normalisation integrals I_ij = sum_cells eff * B_i * conj(B_j) over the
 * m12 x m13 midpoint grid grid6 = {l12, u12, n12, l13, u13, n13}, as one
 * fused lineshape + FP64-MFMA Gram kernel; out[2*(i*n+j)+{0,1}] = {Re, Im}.
 * eff_func is the function index of the efficiency.  Replaces the n*n
 * transform_reduce calls with SpecialResonanceIntegrator
 * (src/PDFs/physics/Amp3Body.cu:312-330, detail/SpecialResonanceIntegrator.cu:7-121). */
gf_status gf_dalitz_gram(gf_ctx *ctx, int dalitz_func, const double grid6[6], int eff_func,
                         double *out_re_im);

/* The integration cells of gf_dalitz_gram (time_dependent = 0) or gf_tddp_gram
 * (time_dependent != 0) on grid6, as those kernels compute them: cell p has m12 bin
 * p % n12 and m13 bin p / n12 (p < ncells = n12 * n13), centres into host arrays
 * m12[p], m13[p].  The time-integrated grid takes the m13 bin width from n12 as
 * SpecialResonanceIntegrator.cu:69-76 does; the time-dependent one its own n13
 * (SpecialDalitzIntegrator.cu:75-82). */
gf_status gf_dalitz_grid_cells(gf_ctx *ctx, const double grid6[6], int time_dependent, double *m12, double *m13,
                               int64_t ncells);

/* Same integrals over a list of np points (m12[p], m13[p]) with weights w[p]
 * (w may be NULL for 1): I_ij = sum_p w_p eff_p B_i(p) conj(B_j(p)).  Host
 * arrays.  Normalisation by phase-space Monte Carlo; no reference counterpart. */
gf_status gf_dalitz_gram_points(gf_ctx *ctx, int dalitz_func, const double *m12, const double *m13,
                                const double *w, int64_t np, int eff_func, double *out_re_im);

/* Same, with the point arrays already in device memory. */
gf_status gf_dalitz_gram_points_device(gf_ctx *ctx, int dalitz_func, const double *dev_m12,
                                       const double *dev_m13, const double *dev_w, int64_t np,
                                       int eff_func, double *out_re_im);

/* Copy cached wave `res` (n complex values, re/im interleaved) to the host.
 * Replaces Amp3Body::getCachedWave (src/PDFs/physics/Amp3Body.cu:370-375).  For an
 * Amp3Body_TD node res in [nres, 2 nres) selects B wave res - nres
 * (Amp3Body_TD::getCachedWaveNoCopy, include/goofit/PDFs/physics/Amp3Body_TD.h:136-138). */
gf_status gf_dalitz_get_cache(gf_ctx *ctx, int dalitz_func, int res, double *out_re_im, int64_t n);

/* Data-event Gram G_ij = sum_e W_i[e] conj(W_j[e]) over the wave cache, used for
 * fit fractions.  Replaces the n*n transform_reduce calls of
 * Amp3Body::fit_fractions (src/PDFs/physics/Amp3Body.cu:376-428). */
gf_status gf_dalitz_cache_gram(gf_ctx *ctx, int dalitz_func, double *out_re_im);

/* ---- time-dependent three-body Dalitz (Amp3Body_TD) -------------------------
 * A GF_FN_TDDP node takes the Amp3Body entry points above for its wave cache:
 * gf_dalitz_set_data_size allocates the A waves W_i(m12, m13) and the B waves
 * W_i(m13, m12) (WaveHolder_s, Amp3Body_TD.cu:531-580) and gf_dalitz_cache fills
 * both (SpecialWaveCalculator.cu:13-63).  Its normalisation integrals: */

/* ThreeComplex integrals of every resonance pair over the m12 x m13 midpoint grid
 * grid6 = {l12, u12, n12, l13, u13, n13}: out_aa = sum eff a_i conj(a_j), out_ab =
 * sum eff a_i conj(b_j), out_bb = sum eff b_i conj(b_j), each nres x nres complex
 * (re, im interleaved, row-major), with the efficiency of the node evaluated on
 * the fake event {2, .., m12, m13} at each cell centre.  One fused lineshape +
 * FP64-MFMA Gram of the stacked waves (a, b); above 8 resonances, pairs of
 * 8-wave groups.  Replaces the nres^2 transform_reduce calls with
 * SpecialDalitzIntegrator (src/PDFs/physics/Amp3Body_TD.cu:664-682,
 * detail/SpecialDalitzIntegrator.cu:9-126). */
gf_status gf_tddp_gram(gf_ctx *ctx, int tddp_func, const double grid6[6], double *out_aa, double *out_ab,
                       double *out_bb);

/* ---- normalisation by Monte Carlo: uniform and learned proposals ------------
 * No reference counterpart (the reference integrates Amp3Body on a fixed grid,
 * src/PDFs/physics/detail/SpecialResonanceIntegrator.cu:52-121).  Point sets go
 * to gf_dalitz_gram_points_device: with weights w_p = 1 / (N q(x_p)) the Gram
 * sum is an unbiased estimate of the integrals over the box. */

/* n uniform points of the box {l12, u12, l13, u13} into device arrays, w = area / n.
 * Counter-based generator: the same seed gives the same points. */
gf_status gf_sample_uniform(gf_ctx *ctx, const double box[4], int64_t n, uint64_t seed, double *dev_m12,
                            double *dev_m13, double *dev_w);

typedef struct gf_mlp gf_mlp;

/* Number of parameters of the proposal network (2 -> 128 -> 128 -> 1, SiLU):
 * W1[128][2], b1[128], W2[128][128] (row = output neuron), b2[128], w3[128], b3. */
int gf_mlp_param_count(void);

/* A network on ctx's device with seeded initial weights; batch_cap bounds the
 * samples of one forward/backward pass (rounded up to a multiple of 64).
 * Fails with GF_ERR_HIP if the device's BF16 MFMA result layout is not the one
 * the kernels were written for. */
gf_status gf_mlp_create(gf_ctx *ctx, uint64_t seed, int batch_cap, gf_mlp **out);
void gf_mlp_destroy(gf_mlp *m);
gf_status gf_mlp_get_params(gf_mlp *m, float *out, int n);
gf_status gf_mlp_set_params(gf_mlp *m, const float *in, int n);

/* f(x, y) for n host points xy[2i], xy[2i+1] in [-1, 1]^2 (u = (x+1)/2 of the box). */
gf_status gf_mlp_forward(gf_mlp *m, const float *xy, int64_t n, float *out);

/* grad = sum_i coef[i] * d f(x_i) / d theta over n host points (n <= batch_cap). */
gf_status gf_mlp_gradient(gf_mlp *m, const float *xy, const float *coef, int64_t n, float *grad);

/* Maximum-likelihood training on ndata events (device arrays m12, m13 inside box):
 * `steps` Adam steps (cosine-decayed learning rate from lr), each on batch/2 data
 * events drawn with replacement and batch/2 uniform points of the box
 * (loss = -mean_data f + log mean_uniform exp f).  loss_first/loss_last may be NULL. */
gf_status gf_mlp_train(gf_mlp *m, const double *dev_m12, const double *dev_m13, int64_t ndata, const double box[4],
                       int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last);

/* on = 0: launch each training step's kernels from the host instead of
 * replaying one captured step (the results are identical). */
gf_status gf_mlp_use_graph(gf_mlp *m, int on);

/* Same, with the events in host arrays. */
gf_status gf_mlp_train_host(gf_mlp *m, const double *m12, const double *m13, int64_t ndata, const double box[4],
                            int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last);

/* gf_mlp_train on nevents events of Amp3Body dalitz_func's density at its current
 * parameters, generated on the device (gf_dalitz_generate with gen_seed and the
 * estimated envelope) and never copied to the host. */
gf_status gf_mlp_train_generated(gf_mlp *m, int dalitz_func, int64_t nevents, uint64_t gen_seed, const double box[4],
                                 int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last);

/* Tabulate the proposal on a grid x grid cell-centre table over box:
 * p_cell = (1 - alpha) exp(f) / sum exp(f) + alpha / grid^2 (0 < alpha <= 1). */
gf_status gf_mlp_build_proposal(gf_mlp *m, const double box[4], int grid, double alpha);

/* n points of the tabulated proposal and their weights 1 / (n q(x)). */
gf_status gf_mlp_sample(gf_mlp *m, int64_t n, uint64_t seed, double *dev_m12, double *dev_m13, double *dev_w);

/* Draw and keep n normalisation points for Amp3Body dalitz_func: from the
 * tabulated proposal when proposal is not NULL, else uniform over box. */
gf_status gf_dalitz_set_norm_points(gf_ctx *ctx, int dalitz_func, gf_mlp *proposal, int64_t n, uint64_t seed,
                                    const double box[4]);

/* Keep the n points and weights of a deterministic quadrature rule (host arrays)
 * as the normalisation points of Amp3Body dalitz_func, in place of drawn ones:
 * the integrals become sum_k w_k eff B_i conj(B_j) over them.  No reference
 * counterpart (the reference integrates on the midpoint grid,
 * src/PDFs/physics/detail/SpecialResonanceIntegrator.cu:52-121). */
gf_status gf_dalitz_set_norm_rule(gf_ctx *ctx, int dalitz_func, const double *m12, const double *m13,
                                  const double *w, int64_t n);

/* Integrals I_ij = sum_p w_p eff_p B_i(p) conj(B_j(p)) over the kept points
 * [offset, offset + count) (weights as drawn for all n points). */
gf_status gf_dalitz_gram_norm_points(gf_ctx *ctx, int dalitz_func, int eff_func, int64_t offset, int64_t count,
                                     double *out_re_im);

/* n events of Amp3Body dalitz_func's density |sum c_i B_i|^2 * eff (current
 * parameters) by accept-reject on uniform candidates of box: candidate k of the
 * seeded counter-based stream is accepted when u_k * fmax < value_k; accepted
 * events keep the stream's order, so (seed, fmax) fix the events.  fmax <= 0:
 * 1.25 x the largest value over the first 2^22 candidates of the stream (which
 * are then skipped).  Fails if a candidate exceeds fmax.  host_out != 0:
 * m12/m13 are host arrays, else device arrays.  tried / fmax_used may be NULL.
 * Takes the place of the reference's MCBooster generation
 * (Amp3Body::GenerateSig, src/PDFs/physics/Amp3Body.cu:430-547) and of its
 * grid-based DalitzPlotter::fillDataSetMC (include/goofit/PDFs/physics/DalitzPlotter.h:64-95). */
gf_status gf_dalitz_generate(gf_ctx *ctx, int dalitz_func, int64_t n, uint64_t seed, const double box[4],
                             double fmax, int host_out, double *m12, double *m13, int64_t *tried,
                             double *fmax_used);

/* ---- toy batches ----------------------------------------------------------
 * Many fits of one model at once (no reference counterpart; the reference fits
 * one data set per FitManager, src/goofit/FitManagerMinuit2.cpp:31-119).  Toy t
 * owns rows [first[t], first[t] + count[t]) of the context's events and its own
 * copies of the parameter and normalisation arrays (layout of gf_program_set).
 * One gf_batch_reduce is one captured hipGraph: the staged arrays are copied in,
 * one sweep covers every active toy's rows, and one workgroup per toy adds its
 * partials in a fixed order. */
typedef struct gf_batch gf_batch;

gf_status gf_batch_create(gf_ctx *ctx, int ntoys, const int64_t *first, const int64_t *count, gf_batch **out);
void gf_batch_destroy(gf_batch *b);
/* Stage toy's parameter / normalisation arrays (either may be NULL: unchanged). */
gf_status gf_batch_set_toy(gf_batch *b, int toy, const double *params, const double *norms);
/* Metric sums of the toys with active[t] != 0 into out[t] (others untouched). */
gf_status gf_batch_reduce(gf_batch *b, const uint8_t *active, int metric, double *out);
/* Events per workgroup (a multiple of 256); default 8192. */
gf_status gf_batch_set_chunk(gf_batch *b, int chunk);
gf_status gf_batch_get_chunk(gf_batch *b, int *chunk);
/* on = 0: plain stream launches instead of the captured graph. */
gf_status gf_batch_use_graph(gf_batch *b, int on);
/* Time an active-all reduction for chunk in {2048, 4096, 8192, 16384} (one
 * warm-up and two timed launches each), keep the fastest. */
gf_status gf_batch_autotune(gf_batch *b, int metric, int *chunk_out, double *us_out);

/* Several parameter points per toy in one launch (no reference counterpart: the
 * points of one derivative round of every running fit, FCNBase::EvaluateMany).
 * gf_batch_add_point appends a point (params: np values, norms: nn values, the
 * layout of gf_program_set) to toy `toy`; a toy's points are added one after the
 * other, at most GF_BATCH_MAX_POINTS per toy.  gf_batch_reduce_points evaluates
 * every added point in one launch sequence (plain stream launches, the staged
 * points copied in with one copy), writes the metric sum of toy t's k-th point to
 * out[t * GF_BATCH_MAX_POINTS + k] (other entries untouched) and clears the
 * points; gf_batch_clear_points drops them without a launch.  A lone Amp3Body
 * reads each event's waves once for up to 16 points.  Each sum is added in the
 * order of gf_batch_reduce at the same chunk, so a lone point gives the same
 * value as gf_batch_reduce for the same arrays. */
#define GF_BATCH_MAX_POINTS 64
gf_status gf_batch_add_point(gf_batch *b, int toy, const double *params, const double *norms);
gf_status gf_batch_reduce_points(gf_batch *b, int metric, double *out);
gf_status gf_batch_clear_points(gf_batch *b);
/* The two halves of gf_batch_reduce_points: submit launches the added points and
 * returns at once with a ticket; points added after it go to a second buffer, so
 * the next request can be staged while the launch runs.  At most two submissions
 * are in flight (adding or submitting while both are fails until one is waited
 * for); wait blocks until the ticket's launch has finished and writes its sums as
 * gf_batch_reduce_points does. */
gf_status gf_batch_submit_points(gf_batch *b, int metric, uint32_t *ticket);
gf_status gf_batch_wait_points(gf_batch *b, uint32_t ticket, double *out);

/* Normalisation of the points on the device.  The reference normalises on the
 * host at every FCN call: GooPdf::normalize (src/PDFs/GooPdf.cu:340-396) with the
 * analytic integrals of GaussianPdf (src/PDFs/basic/GaussianPdf.cu:28-35) and
 * ExpPdf (src/PDFs/basic/ExpPdf.cu:121-132), and Amp3Body::normalize's quadratic
 * form of the couplings over the cached integrals (src/PDFs/physics/
 * Amp3Body.cu:334-358).  With ops installed, a point added with norms == NULL gets
 * its normalisation array inside the launch of gf_batch_submit_points: the
 * template array installed here, then norms[op.out] = 1 / I for each op, with I
 * from that point's own parameters P:
 *   GF_NORM_SCALE   I = P[p] * c0                       (GaussianPdf: c0 = sqrt 2 sqrt pi)
 *   GF_NORM_EXP     I = (exp(a c1) - exp(a c0)) / a, a = P[p]; c1 - c0 if a == 0  (ExpPdf on [c0, c1])
 *   GF_NORM_DALITZ  I = c0 Re sum_ij c_i conj(c_j) T_ij, c_i = P[p + 2i] + i P[p + 2i + 1],
 *                   T_ij = table[g + 2 (i n + j)] + i table[g + 2 (i n + j) + 1], i, j < n
 * each in the host's operation order.  nops = 0 removes the ops (points then need
 * their norms again).  The table (nt doubles) and the template (nn values) are
 * copied. */
enum gf_norm_kind { GF_NORM_SCALE = 1, GF_NORM_EXP = 2, GF_NORM_DALITZ = 3 };
typedef struct gf_norm_op {
    int32_t kind;  /* gf_norm_kind */
    int32_t out;   /* index in the normalisation array */
    int32_t p;     /* index in the parameter array */
    int32_t n;     /* GF_NORM_DALITZ: resonance count (<= 16) */
    int32_t g;     /* GF_NORM_DALITZ: first entry in the table */
    int32_t pad;
    double c0, c1;
} gf_norm_op;
gf_status gf_batch_set_norm_ops(gf_batch *b, const gf_norm_op *ops, int nops, const double *norms_template,
                                const double *table, int nt);

/* ---- timing ------------------------------------------------------------- */

/* Run gf_reduce_metric `reps` times back to back on the context stream and
 * return the elapsed device time between HIP events recorded on that stream
 * before the first and after the last launch.  No reference counterpart. */
gf_status gf_time_reduce_metric(gf_ctx *ctx, int metric, int reps, double *ms_total);

/* Same launches (reps <= 4096), with events around each event sweep: ms_sweep is
 * the summed device time of the sweep kernels alone (without the kernel that
 * adds the per-workgroup partials, or including that work where the sweep does it
 * itself, GF_TUNE_FUSED_FINISH), ms_total (optional) the whole span.
 * No reference counterpart. */
gf_status gf_time_reduce_sweep(gf_ctx *ctx, int metric, int reps, double *ms_sweep, double *ms_total);

/* The event sweep alone, reps <= 4096 launches back to back between two events
 * (no finishing kernel in between and no result read; a sweep that adds its
 * partials itself still does): ms_per_sweep = elapsed / reps, the sweep kernel's
 * duration plus the gap between back-to-back dispatches.  The next reduction of
 * the context is unaffected.  No reference counterpart. */
gf_status gf_time_sweep_back_to_back(gf_ctx *ctx, int metric, int reps, double *ms_per_sweep);

/* Same for gf_dalitz_cache (all resonances) and gf_dalitz_gram. */
gf_status gf_time_dalitz_cache(gf_ctx *ctx, int dalitz_func, int reps, double *ms_total);
gf_status gf_time_dalitz_gram(gf_ctx *ctx, int dalitz_func, const double grid6[6], int eff_func, int reps,
                              double *ms_total);

#ifdef __cplusplus
}
#endif

#endif /* GOOFIT_HIP_H */

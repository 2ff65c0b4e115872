// gf_internal.h -- structures shared by the C-ABI host code and the gfx950 kernels.
//
// A flattened PDF tree (the reference's four arrays + function table,
// src/PDFs/PdfBase.cu:268-311) is compiled on the host into a linear program
// of GfIns instructions.  The kernels interpret it with wave-uniform control
// flow: every lane of a wavefront runs the same instruction on its own events,
// so the instruction fetch and dispatch live in scalar registers and only the
// arithmetic is per lane.  Tree recursion becomes a small stack of per-level
// accumulators (acc[level]); a leaf leaves its value in T.
#pragma once

#include <stdint.h>

#define GF_MAX_OBS 8      // observable columns per event
#define GF_MAX_LVL 4      // nesting depth of combining nodes
#define GF_MAX_RES 16     // resonances per Dalitz node (Amp3Body.h:88 caps it at 16 as well)
#define GF_MAX_DALITZ 4   // Dalitz nodes per program
#define GF_BLOCK 256      // threads per block of the event kernels
#define GF_MAX_PARTIALS 4096
// An empty slot of the in-kernel finish: a signalling NaN.  Arithmetic never
// returns one (IEEE operations return quiet NaNs), so no partial equals it.
#define GF_SLOT_EMPTY 0x7FF4A5A5A5A5A5A5ull
// Device error word bit: a partial of the in-kernel finish did not become visible
// within the finishing workgroup's bounded wait.
#define GF_DEVERR_FINISH 4

// Instruction codes of the compiled program.  Leaf codes compute T; the
// combining steps act on the accumulator of one nesting level and are encoded
// as GI_LVL + 4*kind + level, so each (kind, level) pair is its own case with
// the accumulator fixed at compile time (no register array indexed at run time).
enum GfOp : int32_t {
    GI_END = 0,
    GI_GAUSS,          // T = exp(-0.5*(x-m)*(x-m)/(s*s))
    GI_EXP,            // T = exp(a*x)
    GI_EXP_OFF,        // T = exp(a*(x-x0))
    GI_EXP_POLY,       // T = exp(sum_i x^i p_i)
    GI_EXP_POLY_OFF,   // T = exp(sum_{i>=1} (x-x0)^i p_i)
    GI_EXPGAUS,        // exponentially modified Gaussian
    GI_POLY,           // sum_i p_i x^(d+i)
    GI_OFF_POLY,       // sum_{i>=2} p_i x^(d+i-2)
    GI_MULTI_POLY,     // multivariate polynomial
    GI_LVL = 16
};
enum GfLvlKind : int32_t {
    GK_PUSH0 = 0,      // acc = 0
    GK_PUSH1,          // acc = 1
    GK_ADD_ACC,        // acc += w * T * norm
    GK_ADD_LAST,       // acc += (1 - sum w) * T * norm
    GK_ADDEXT_END,     // T = acc / sum w
    GK_PROD_ACC,       // acc *= T * norm
    GK_POP,            // T = acc
    GK_DALITZ,         // acc = inside ? |sum_i c_i W_i|^2 : -1
    GK_DALITZ_END,     // T = (acc == -1) ? 0 : acc * T
    GK_TDDP,           // acc = time-dependent rate of A = sum c_i W_i, B = sum c_i W'_i
    GK_TDDP_END,       // T = inside ? acc * T : 0
    GK_COUNT
};
#define GF_LVL_OP(kind, lvl) (GI_LVL + 4 * (kind) + (lvl))

struct GfIns {
    int32_t op;   // GfOp
    int32_t lvl;  // accumulator level
    int32_t a;    // observable column of a leaf; resonance count of GI_DALITZ
    int32_t p;    // index in params[] of the first parameter value
    int32_t c;    // index in consts[] of the first constant value
    int32_t n;    // index in norms[] of the normalisation value used by a combine step
    int32_t k;    // count (number of parameters / weights / observables)
    int32_t x;    // GI_MULTI_POLY: index in obsid[] of its first column id; GI_DALITZ: node slot
};

// Lineshape kinds of the resonances of a Dalitz node.
enum GfRes : int32_t { GR_RBW = 0, GR_RBW_SYM = 1, GR_GS = 2, GR_NONRES = 3 };

struct GfResDesc {
    int32_t kind;  // GfRes
    int32_t p;     // index in params[] of mass (width at p+1)
    int32_t c;     // index in consts[] of spin (cyclic index at c+1, norm flag at c+2)
    int32_t swap;  // 1: evaluated at (m13, m12) -- the B wave of a time-dependent node
};

// Time-resolution models of a time-dependent Dalitz node.
enum GfTimeRes : int32_t { GT_TRUTH = 0, GT_THREEGAUSS = 1 };

struct GfDecay {
    double motherMass, daug1Mass, daug2Mass, daug3Mass, meson_radius, mother_meson_radius;
};

// Per-Dalitz-node data the coherent sum needs.  A time-dependent node also has
// the B waves W'_i(m12, m13) = W_i(m13, m12) in wave_b (same layout), its time,
// time-error, mistag and charm-tag columns (-1: absent) and its resolution model.
struct GfDalitzSlot {
    const double2 *wave;  // nres x ncache complex, resonance-major (W[i*ncache + e])
    int64_t ncache;
    int32_t col_m12, col_m13, col_num;
    int32_t nres;
    const double2 *wave_b;
    int32_t col_time, col_sigma, col_mis, col_tag;
    int32_t res_kind;     // GfTimeRes
    int32_t res_p;        // index in params[] of the resolution's first parameter
};

// Everything an event-evaluating kernel needs, passed by value as the kernel
// argument.  Arrays indexed by a program operand (column pointers, Dalitz
// slots) live in device memory so that a wave-uniform index becomes a scalar
// load instead of a dynamically indexed copy of the argument block.
struct GfProgArgs {
    const GfIns *ins;            // program (device)
    const double *params;        // device copies of the flattened arrays
    const double *consts;
    const double *norms;
    const int32_t *obsid;        // observable ids of the obs[] array as ints (device)
    const double *const *cols;   // [GF_MAX_OBS] device column pointers
    const GfDalitzSlot *dz;      // [GF_MAX_DALITZ] Dalitz node data
    int64_t nevents;
    int32_t identity_num;        // 1: event-number column equals the row index
    int32_t nobs;
    double total_events;         // c_totalEvents
    GfDecay decay;
    int32_t td;                  // 1: the program holds a time-dependent Dalitz node
    int32_t has_dz;              // 1: the program holds a Dalitz node (either kind)
    int32_t *err;                // device error word (bit 0: cache index out of range, bit 1: node kind not compiled in)
    uint32_t seq;                // sequence number k_finish stamps on its GfResult
    // In-kernel finish (lone-Amp3Body and flat reductions; null: k_finish adds the
    // partials).  fuse_slots: 2 x GF_MAX_PARTIALS words (sum, residual per workgroup),
    // GF_SLOT_EMPTY when unused and between launches; fuse_result: the block k_finish
    // would write.
    unsigned long long *fuse_slots;
    struct GfResult *fuse_result;
    void *mark;                  // host only: hipEvent_t recorded between a sweep and its k_finish
    int32_t sweep_only;          // host only: launch the sweep without its k_finish (timing)
    int32_t coh_mx_nres;         // host only: > 0 runs a lone Amp3Body's multi-point NLL (constant efficiency,
                                 // waves 0 outside) on the matrix cores, k_coh_mx, for this many resonances
    int32_t coh_group;           // 1: k_dalitz_fast groups the NLL's events as k_coh_mx does (set when the
                                 // multi-point sums of the context run there)
};

#define GF_METRIC_BINAVG_CODE 3  // metrics from here on read (bin entry, bin volume/error) columns

// Parameter and normalisation values passed inside the kernel arguments of the
// per-call kernels (no host-to-device copy per likelihood evaluation) when the
// flattened arrays fit; larger trees read the device copies instead.
#define GF_PIN 256
#define GF_NIN 128
struct GfInline {
    double p[GF_PIN];
    double n[GF_NIN];
};

// Result block of a reduction, written by k_finish to host-mapped memory:
// the values first, then (after a system-scope fence) the launch's sequence
// number, so a host that sees seq reads complete values.
struct GfResult {
    double sum;
    double comp;
    int32_t err;
    uint32_t seq;
    int32_t pad[2];
};

// One workgroup's share of a toy batch: rows [start, end) of toy `toy`
// (total = the toy's event count, for the binned metrics).
struct GfBatchWork {
    int32_t toy;
    int32_t pad;
    int64_t start, end;
    double total;
};

// A normalisation computed on the device for a batched point (gf_norm_op of the C ABI).
enum GfNormKind : int32_t { GN_SCALE = 1, GN_EXP = 2, GN_DALITZ = 3 };
struct GfNormOp {
    int32_t kind, out, p, n, g, pad;
    double c0, c1;
};

// Host-side launchers (gf_kernels.hip).
namespace gfk {
struct Partial {
    double s, c;
};
// inl: parameter/normalisation values in the kernel arguments (null: device arrays).
// Reductions: one partial per workgroup, then one workgroup adds them into *result.
int launch_reduce_metric(const GfProgArgs &a, const GfInline *inl, int metric, Partial *partials, GfResult *result,
                         int nblocks, void *stream);
int launch_eval_metric(const GfProgArgs &a, const GfInline *inl, int metric, double *out, int nblocks, void *stream);
// Flat program (one leaf, or one AddPdf / ProdPdf level of Gaussian / Exp / Exp-offset /
// ExpGaus leaves over columns 0..ncols-1, ncols <= GF_FLAT_COLS), unbinned metrics:
// the same sum as launch_reduce_metric; ept = 2 or 4 events per lane.
#define GF_FLAT_COLS 4
int launch_reduce_flat(const GfProgArgs &a, const GfInline *inl, int metric, int ncols, int ept, Partial *partials,
                       GfResult *result, int nblocks, void *stream);
// Coherent-sum fast path: program [DALITZ, MULTI_POLY efficiency, DALITZ_END, END],
// event number = row index, every row inside the wave cache.
// const_eff: the efficiency polynomial has degree 0 (a constant coefficient).
// The coherent-sum metric of npts <= GF_MULTI_PTS parameter sets (params: npts x np,
// norms: npts x nn, device) over the same events in one sweep; out[2p], out[2p+1] =
// sum and residual of set p (device).  partials: npts x nblocks.
#define GF_MULTI_PTS 16
#define GF_MX_SETS 32  // sets per sweep of the matrix-core NLL path (two groups of GF_MULTI_PTS)
int launch_dalitz_multi(const GfProgArgs &a, int nres, const double *params, int np, const double *norms, int nn,
                        int npts, bool const_eff, bool zero_outside, int metric, Partial *partials, double *out,
                        int nblocks, void *stream);
// zero_outside (with const_eff): the wave cache was filled from these events, so m12/m13
// are not read (the waves are 0 outside the Dalitz region).
int launch_dalitz_fast(const GfProgArgs &a, const GfInline *inl, bool const_eff, bool zero_outside, int metric,
                       Partial *partials, GfResult *result, double *out, int nblocks, void *stream);
int launch_reduce_bins(const GfProgArgs &a, int pc0, const double *ranges, const int *slots, int k, int64_t total,
                       Partial *partials, GfResult *result, int nblocks, void *stream);
// Per-bin values of the same grid into out[total] (device).
int launch_eval_bins(const GfProgArgs &a, int pc0, const double *ranges, const int *slots, int k, int64_t total,
                     double *out, int nblocks, void *stream);
// Cell centres of a k-dimensional grid: out[d * total + bin] (device).
int launch_grid_cells(const double *ranges, int k, int64_t total, double *out, void *stream);
// Mark every slot of the in-kernel finish empty.
int launch_fuse_init(unsigned long long *slots, void *stream);
// Cell centres of a Dalitz integration grid grid6 (host array), cells 0..ncells-1 (device outputs).
int launch_dalitz_cells(const double *grid6, int td_grid, int64_t ncells, double *m12, double *m13, void *stream);
// Toy batch: per-workgroup sweep with per-toy parameters, per-toy finishing
// reduction into result[2t..2t+1] (host-mapped), then the completion stamp.
int launch_batch(const GfProgArgs &a, const GfBatchWork *work, int nblocks, const double *params, int np,
                 const double *norms, int nn, const uint8_t *active, int metric, Partial *partials, const int *first,
                 int ntoys, double *result, uint32_t *counter, uint32_t *stamp, void *stream);
// Toy batch with several points per toy: tbl[2t], tbl[2t + 1] = first stage slot, points;
// each stage slot holds np parameters then nn normalisation values.  The grid is
// nblocks x ceil(max points / GF_MULTI_PTS); partials[b * GF_BATCH_PTS + p];
// result[2 (t GF_BATCH_PTS + p) ..] (host-mapped), then the stamp.
// fast: the coherent-sum program of launch_dalitz_fast.
#define GF_BATCH_PTS 64
// Normalisation arrays of nslots staged points (slot k: np parameters then nn norms at
// stage + k (np + nn)): the template, then each op's 1 / integral from the slot's parameters.
int launch_batch_norms(double *stage, int nslots, int np, int nn, const double *tmpl, const GfNormOp *ops, int nops,
                       const double *table, void *stream);
int launch_batch_points(const GfProgArgs &a, int nres, const GfBatchWork *work, int nblocks, int groups,
                        const double *stage, int np, int nn, const int *tbl, bool fast, bool const_eff,
                        bool zero_outside, int metric, int mx_min, Partial *partials, const int *first, int ntoys,
                        double *result, uint32_t *counter, uint32_t *stamp, void *stream);
// Accept-reject generation: candidates (uniform in box) with accept flags, and
// the ordered compaction of the accepted ones.
// eff_pc >= 0: efficiency subprogram; -1: efficiency 1; -2: the constant eff_const.
int launch_dalitz_candidates(const GfProgArgs &a, const GfResDesc *res, int nres, int p_coup, int eff_pc,
                             double eff_const, int id12, int id13, const double box[4], uint64_t key, int64_t base,
                             int64_t nc, double fmax, double *c12, double *c13, uint8_t *flag, int *bcount,
                             double *bmax, void *stream);
int launch_compact(const double *c12, const double *c13, const uint8_t *flag, int64_t nc, const int64_t *boff,
                   int64_t limit, double *o12, double *o13, void *stream);
int launch_dalitz_cache(const GfProgArgs &a, int slot, int64_t offset, int64_t n, uint32_t mask,
                        const GfResDesc *res, int nres, int nblocks, void *stream);
// mode 0: grid cells (grid6 host array); 1: points (m12, m13, w device arrays); 2: data cache of `slot`.
// Writes 768 doubles (tiles RR, II, RI in MFMA register layout) to tiles_dev.
// td_grid: the grid of a time-dependent node (m13 bin count grid6[5]); res[i].swap
// evaluates entry i at (m13, m12).
int launch_gram(const GfProgArgs &a, int mode, const double *grid6, const double *m12, const double *m13,
                const double *w, int64_t npoints, int slot, int id_m12, int id_m13, int eff_pc, double eff_const,
                int td_grid, const GfResDesc *res, int nres, double *partials, double *tiles_dev, int nblocks,
                void *stream);
int launch_identity_check(const double *col, int64_t n, int32_t *flag, void *stream);
int mfma_probe(const double *dev_a, const double *dev_b, double *dev_out, void *stream);
}  // namespace gfk

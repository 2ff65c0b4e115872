// gf_capi.hip -- the C ABI of include/goofit_hip.h: contexts, event columns,
// flattened-tree compilation, and the launches of gf_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <complex>
#include <map>
#include <string>
#include <vector>

#include "../../include/goofit_hip.h"
#include "gf_internal.h"
#include "gf_device.h"
#include "gf_mlp.h"

#include <mutex>
#include <random>

// Per-workgroup partials of gf_reduce_metric_noalloc: a fixed device array that
// every GPU holds once, for the life of the process.  In host code `scratch` only
// names the symbol; its address on a device comes from hipGetSymbolAddress with
// that device current (noalloc_partials), and only that address is passed to
// kernels.
static __device__ __align__(256) gfk::Partial scratch[GF_NOALLOC_PARTIALS];
static_assert(GF_NOALLOC_PARTIALS >= GF_MAX_PARTIALS, "the no-allocation scratch holds every reduction grid");
static_assert(sizeof(gf_result) == sizeof(GfResult), "gf_result is the device result block");

namespace {

thread_local std::string g_last_error;

gf_status fail(gf_status code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define GF_HIP(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if(e_ != hipSuccess)                                                                                           \
            return fail(GF_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));                                    \
    } while(0)

// Wave-cache storage of one Dalitz node, keyed by its cacheToUse constant as the
// reference keys cResonances[16 * cacheToUse + i] (Amp3Body.cu:54, 222-233).
struct WaveCache {
    double2 *wave = nullptr;    // GF_MAX_RES x n complex
    double2 *wave_b = nullptr;  // time-dependent nodes: the B waves, same layout
    int64_t n = 0, offset = 0;
    // resonances whose waves were computed from the current event columns (events_gen,
    // m12/m13 columns): with all of them filled, the cache is 0 exactly outside the plot
    uint64_t filled = 0, gen = 0;
    int col12 = -1, col13 = -1;
};

// A Dalitz node of the installed program.
struct DalitzNode {
    int func = -1;        // function index
    int cache_key = 0;    // cacheToUse constant
    int nres = 0;
    std::vector<int> res_funcs;
    std::vector<GfResDesc> res;
    int eff_func = -1;    // efficiencyFunction (last function of the subtree, Amp3Body.cu:198-202)
    int eff_root = -1;    // first function of the efficiency subtree
    int col_m12 = 0, col_m13 = 1, col_num = 2;
    int p_coup = 0;       // params index of (re, im) of the first coupling
    // time-dependent node (Amp3Body_TD): columns, mixing parameters, resolution
    bool td = false;
    int col_time = -1, col_sigma = -1, col_mis = -1, col_tag = -1;
    int p_mix = 0;        // params index of tau (x, y, dx, dy follow)
    int res_func = -1;    // resolution function index
    int res_kind = GT_TRUTH;
    int res_p = 0;        // params index of the resolution's first parameter
};

}  // namespace

struct gf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;

    // events
    int ncols        = 0;
    int64_t n        = 0;
    double *col_base = nullptr;  // ncols * n doubles, column k at col_base + k*n
    double **d_cols  = nullptr;  // device array [GF_MAX_OBS]
    int identity_col = -1;
    bool identity_ok = false;

    // program (host copies)
    std::vector<int32_t> funcs;
    std::vector<double> params, consts, obs, norms;
    std::vector<GfIns> ins;           // root program followed by compiled subprograms
    std::map<int, int> sub_pc;        // function index -> first instruction
    int max_col = -1;                 // largest observable column any instruction reads
    std::vector<DalitzNode> dalitz;   // Dalitz nodes of the program, slot order
    bool has_td = false;              // one of them is time-dependent

    // device copies
    double *d_params = nullptr, *d_consts = nullptr, *d_norms = nullptr;
    int32_t *d_obsid = nullptr;
    GfIns *d_ins = nullptr;
    size_t cap_params = 0, cap_consts = 0, cap_norms = 0, cap_obsid = 0, cap_ins = 0;
    GfDalitzSlot *d_dz = nullptr;
    GfResDesc *d_res = nullptr;  // GF_MAX_DALITZ * GF_MAX_RES
    GfDecay decay{};

    std::map<int, WaveCache> caches;

    // Monte-Carlo normalisation points per Amp3Body (keyed by cacheToUse):
    // buf = m12[n] | m13[n] | w[n]
    struct NormPoints {
        double *buf = nullptr;
        int64_t n = 0, cap = 0;
    };
    std::map<int, NormPoints> norm_points;

    // reduction scratch
    gfk::Partial *d_partials = nullptr;
    unsigned *d_counter      = nullptr;
    double *d_result         = nullptr;  // scratch of the grid/probe paths
    GfResult *h_res          = nullptr;  // host-mapped result block written by the last block
    GfResult *d_res_mapped   = nullptr;  // its device address
    double *h_result         = nullptr;  // pinned staging for small copies
    int32_t *d_err           = nullptr;  // device error word (bit 0: cache index out of range, bit 1: node kind not compiled in)
    unsigned long long *d_fuse_slots = nullptr;  // in-kernel finish: 2 x GF_MAX_PARTIALS words
    bool fast_dalitz         = false;    // program is [DALITZ, MULTI_POLY, DALITZ_END, END]
    bool flat                = false;    // program is one level of 1-D leaves (flat_program)
    double *d_gram_partials  = nullptr;
    size_t cap_gram_partials = 0;
    double *d_gram_tiles     = nullptr;  // 768
    double *h_gram_tiles     = nullptr;  // pinned 768
    int mfma_layout          = -1;       // -1 unknown, else a layout_ij() code

    // pending uploads of params/norms through pinned staging
    double *h_stage       = nullptr;
    size_t cap_stage      = 0;
    bool params_dirty     = false;
    bool norms_dirty      = false;
    bool stage_in_flight  = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> ev_sweep;  // gf_time_reduce_sweep

    uint32_t res_seq = 0;  // sequence number of the last reduction launched
    uint32_t noalloc_seq = 0;  // of the last gf_reduce_metric_noalloc

    // launch-shape knobs (gf_tune)
    int tune[GF_TUNE_COUNT] = {1, 1, 4, 3, 2000, 1, 1, 1, 1, 0};

    uint64_t events_gen = 1;  // bumped whenever the event columns may change

    // gf_reduce_metric_multi: parameter/norm sets, partials and results
    // d_multi_out: the device address of the host-mapped h_multi_out (the multi-point results)
    double *d_multi = nullptr, *h_multi = nullptr, *d_multi_out = nullptr, *h_multi_out = nullptr;
    size_t cap_multi = 0;
    gfk::Partial *d_multi_partials = nullptr;
};

namespace {

bool valid_ctx(gf_ctx *ctx) {
    if(!ctx) {
        fail(GF_ERR_INVALID, "null context");
        return false;
    }
    return hipSetDevice(ctx->device) == hipSuccess;
}

// Device address of `scratch` on each GPU, looked up once per device (by the
// first gf_ctx_create on it, with that device current).
std::mutex g_scratch_mu;
std::map<int, gfk::Partial *> g_scratch_addr;

gf_status noalloc_partials_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(g_scratch_addr.count(device))
        return GF_OK;
    void *ptr    = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if(e != hipSuccess || !ptr)
        return fail(GF_ERR_HIP, "hipGetSymbolAddress of the no-allocation scratch on device %d failed: %s", device,
                    hipGetErrorString(e));
    g_scratch_addr[device] = (gfk::Partial *)ptr;
    return GF_OK;
}

gfk::Partial *noalloc_partials(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    return it == g_scratch_addr.end() ? nullptr : it->second;
}

template <class T>
gf_status ensure_dev(T *&ptr, size_t &cap, size_t count, hipStream_t s) {
    if(count <= cap && ptr)
        return GF_OK;
    if(ptr) {
        GF_HIP(hipStreamSynchronize(s));
        GF_HIP(hipFree(ptr));
        ptr = nullptr;
    }
    size_t c = count < 16 ? 16 : count;
    if(hipMalloc((void **)&ptr, c * sizeof(T)) != hipSuccess) {
        ptr = nullptr;
        cap = 0;
        return fail(GF_ERR_NOMEM, "device allocation of %zu bytes failed", c * sizeof(T));
    }
    cap = c;
    return GF_OK;
}

// ------------------------------------------------------- tree compilation

// ParameterContainer cursor (include/goofit/PDFs/ParameterContainer.h:17-21)
struct Cur {
    int p = 0, c = 0, o = 0, n = 0, f = 0;
};

struct Compiler {
    gf_ctx *ctx;
    std::vector<GfIns> out;
    std::string err;
    int max_col = -1;

    bool in_range(const Cur &q) const {
        return q.p >= 0 && q.p < (int)ctx->params.size() && q.c >= 0 && q.c < (int)ctx->consts.size() && q.o >= 0
               && q.o < (int)ctx->obs.size() && q.n >= 0 && q.n < (int)ctx->norms.size() && q.f >= 0
               && q.f < (int)ctx->funcs.size();
    }
    int np(const Cur &q) const { return (int)ctx->params[q.p]; }
    int nc(const Cur &q) const { return (int)ctx->consts[q.c]; }
    int no(const Cur &q) const { return (int)ctx->obs[q.o]; }
    int nn(const Cur &q) const { return (int)ctx->norms[q.n]; }

    // slow incrementIndex(), src/PDFs/ParameterContainer.cu:48-60
    bool inc(Cur &q) {
        if(!in_range(q)) {
            err = "flattened arrays end before function " + std::to_string(q.f);
            return false;
        }
        int a = np(q), b = nc(q), c = no(q), d = nn(q);
        if(a < 0 || b < 0 || c < 0 || d < 0) {
            err = "negative count in the block of function " + std::to_string(q.f);
            return false;
        }
        q.f++;
        q.p += a + 1;
        q.c += b + 1;
        q.o += c + 1;
        q.n += d + 1;
        return true;
    }

    bool expect(const Cur &q, int p, int c, int o, const char *name) {
        if(!in_range(q)) {
            err = std::string(name) + ": arrays end at function " + std::to_string(q.f);
            return false;
        }
        if((p >= 0 && np(q) != p) || (c >= 0 && nc(q) != c) || (o >= 0 && no(q) != o) || nn(q) != 1) {
            char b[256];
            snprintf(b, sizeof(b), "%s at function %d: block counts (%d,%d,%d,%d) differ from (%d,%d,%d,1)", name,
                     q.f, np(q), nc(q), no(q), nn(q), p, c, o);
            err = b;
            return false;
        }
        // every value of the blocks must exist
        if(np(q) > 0 && q.p + np(q) > (int)ctx->params.size() - 1) {
            err = std::string(name) + ": parameter block runs past the array";
            return false;
        }
        if(nc(q) > 0 && q.c + nc(q) > (int)ctx->consts.size() - 1) {
            err = std::string(name) + ": constant block runs past the array";
            return false;
        }
        if(no(q) > 0 && q.o + no(q) > (int)ctx->obs.size() - 1) {
            err = std::string(name) + ": observable block runs past the array";
            return false;
        }
        if(q.n + 1 > (int)ctx->norms.size() - 1) {
            err = std::string(name) + ": normalisation block runs past the array";
            return false;
        }
        return true;
    }

    bool col_of(const Cur &q, int i, int &col) {
        double v = ctx->obs[q.o + 1 + i];
        col      = (int)v;
        if(col < 0 || col >= GF_MAX_OBS || (double)col != v) {
            err = "observable index " + std::to_string(v) + " not in [0," + std::to_string(GF_MAX_OBS) + ")";
            return false;
        }
        if(col > max_col)
            max_col = col;
        return true;
    }

    void push(int op, int lvl, int a, int p, int c, int n, int k, int x) {
        GfIns g;
        g.op  = op;
        g.lvl = lvl;
        g.a   = a;
        g.p   = p;
        g.c   = c;
        g.n   = n;
        g.k   = k;
        g.x   = x;
        out.push_back(g);
    }

    // Emit the subtree rooted at q.f; on return q points past it.  lvl is the first
    // free accumulator level.
    bool emit(Cur &q, int lvl) {
        if(!in_range(q)) {
            err = "function index " + std::to_string(q.f) + " outside the flattened arrays";
            return false;
        }
        const int f = ctx->funcs[q.f];
        int col     = 0;
        switch(f) {
        case GF_FN_GAUSSIAN:
            if(!expect(q, 2, 0, 1, "GaussianPdf") || !col_of(q, 0, col))
                return false;
            push(GI_GAUSS, lvl, col, q.p + 1, q.c + 1, q.n + 1, 2, 0);
            return inc(q);
        case GF_FN_EXP:
            if(!expect(q, 1, 0, 1, "ExpPdf") || !col_of(q, 0, col))
                return false;
            push(GI_EXP, lvl, col, q.p + 1, q.c + 1, q.n + 1, 1, 0);
            return inc(q);
        case GF_FN_EXP_OFFSET:
            if(!expect(q, 2, 0, 1, "ExpPdf(offset)") || !col_of(q, 0, col))
                return false;
            push(GI_EXP_OFF, lvl, col, q.p + 1, q.c + 1, q.n + 1, 2, 0);
            return inc(q);
        case GF_FN_EXP_POLY:
        case GF_FN_EXP_POLY_OFFSET:
            if(!expect(q, -1, 0, 1, "ExpPdf(poly)") || !col_of(q, 0, col))
                return false;
            push(f == GF_FN_EXP_POLY ? GI_EXP_POLY : GI_EXP_POLY_OFF, lvl, col, q.p + 1, q.c + 1, q.n + 1, np(q), 0);
            return inc(q);
        case GF_FN_EXPGAUS:
            if(!expect(q, 3, 0, 1, "ExpGausPdf") || !col_of(q, 0, col))
                return false;
            push(GI_EXPGAUS, lvl, col, q.p + 1, q.c + 1, q.n + 1, 3, 0);
            return inc(q);
        case GF_FN_POLYNOMIAL:
        case GF_FN_OFFSET_POLYNOMIAL:
            if(!expect(q, -1, 1, 1, "PolynomialPdf") || !col_of(q, 0, col))
                return false;
            push(f == GF_FN_POLYNOMIAL ? GI_POLY : GI_OFF_POLY, lvl, col, q.p + 1, q.c + 1, q.n + 1, np(q), 0);
            return inc(q);
        case GF_FN_MULTI_POLYNOMIAL: {
            if(!expect(q, -1, -1, -1, "PolynomialPdf(multi)"))
                return false;
            const int nobs = no(q);
            if(nobs < 1 || nobs > GF_MAX_OBS || nc(q) < 1) {
                err = "PolynomialPdf(multi): bad observable/constant counts";
                return false;
            }
            for(int i = 0; i < nobs; i++)
                if(!col_of(q, i, col))
                    return false;
            // coefficient count the device loop will read (PolynomialPdf.cu:61-111)
            int maxDegree = (int)ctx->consts[q.c + 1] + 1, boxes = 1, used = 1;
            for(int i = 0; i < nobs; i++)
                boxes *= maxDegree;
            for(int b = 1; b < boxes; b++) {
                int ci = b, sum = 0;
                for(int o = 0; o < nobs; o++) {
                    sum += ci % maxDegree;
                    ci /= maxDegree;
                }
                if(sum < maxDegree)
                    used++;
            }
            if(np(q) < nobs + used) {
                err = "PolynomialPdf(multi): fewer coefficients than terms";
                return false;
            }
            push(GI_MULTI_POLY, lvl, nobs, q.p + 1, q.c + 1, q.n + 1, np(q), q.o + 1);
            return inc(q);
        }
        case GF_FN_ADD:
        case GF_FN_ADD_EXT: {
            if(!expect(q, -1, -1, -1, "AddPdf"))
                return false;
            if(lvl >= GF_MAX_LVL) {
                err = "PDF tree nests deeper than " + std::to_string(GF_MAX_LVL) + " combining levels";
                return false;
            }
            const int nw   = np(q);
            const int wbase = q.p + 1;
            const bool ext = f == GF_FN_ADD_EXT;
            if(!inc(q))
                return false;
            push(GF_LVL_OP(GK_PUSH0, lvl), lvl, 0, 0, 0, 0, 0, 0);
            for(int i = 0; i < nw; i++) {
                const int normslot = q.n + 1;
                if(!emit(q, lvl + 1))
                    return false;
                push(GF_LVL_OP(GK_ADD_ACC, lvl), lvl, 0, wbase + i, 0, normslot, 0, 0);
            }
            if(ext) {
                push(GF_LVL_OP(GK_ADDEXT_END, lvl), lvl, 0, wbase, 0, 0, nw, 0);
            } else {
                const int normslot = q.n + 1;
                if(!emit(q, lvl + 1))
                    return false;
                push(GF_LVL_OP(GK_ADD_LAST, lvl), lvl, 0, wbase, 0, normslot, nw, 0);
                push(GF_LVL_OP(GK_POP, lvl), lvl, 0, 0, 0, 0, 0, 0);
            }
            return true;
        }
        case GF_FN_PROD: {
            // device_ProdPdfs walks incrementIndex(1, 0, numCons, numObs, 1) (ProdPdf.cu:15)
            if(!expect(q, 0, -1, -1, "ProdPdf"))
                return false;
            if(nc(q) < 1) {
                err = "ProdPdf without component count";
                return false;
            }
            if(lvl >= GF_MAX_LVL) {
                err = "PDF tree nests deeper than " + std::to_string(GF_MAX_LVL) + " combining levels";
                return false;
            }
            const int ncomp = (int)ctx->consts[q.c + 1];
            if(!inc(q))
                return false;
            push(GF_LVL_OP(GK_PUSH1, lvl), lvl, 0, 0, 0, 0, 0, 0);
            for(int i = 0; i < ncomp; i++) {
                const int normslot = q.n + 1;
                if(!emit(q, lvl + 1))
                    return false;
                push(GF_LVL_OP(GK_PROD_ACC, lvl), lvl, 0, 0, 0, normslot, 0, 0);
            }
            push(GF_LVL_OP(GK_POP, lvl), lvl, 0, 0, 0, 0, 0, 0);
            return true;
        }
        case GF_FN_DALITZ: {
            if(!expect(q, -1, 2, -1, "Amp3Body"))
                return false;
            const int nres = (int)ctx->consts[q.c + 1];
            if(nres < 0 || nres > GF_MAX_RES || np(q) != 2 * nres || no(q) < 3) {
                err = "Amp3Body: bad resonance/parameter/observable counts";
                return false;
            }
            if(lvl >= GF_MAX_LVL) {
                err = "PDF tree nests deeper than " + std::to_string(GF_MAX_LVL) + " combining levels";
                return false;
            }
            int slot = -1;
            for(size_t s = 0; s < ctx->dalitz.size(); s++)
                if(ctx->dalitz[s].func == q.f)
                    slot = (int)s;
            if(slot < 0) {
                err = "Amp3Body node not registered";
                return false;
            }
            const int pbase = q.p + 1;
            if(!inc(q))
                return false;
            for(int i = 0; i < nres; i++)
                if(!inc(q))
                    return false;
            push(GF_LVL_OP(GK_DALITZ, lvl), lvl, nres, pbase, 0, 0, 0, slot);
            if(!emit(q, lvl + 1))
                return false;
            push(GF_LVL_OP(GK_DALITZ_END, lvl), lvl, 0, 0, 0, 0, 0, 0);
            return true;
        }
        case GF_FN_TDDP: {
            // device_Tddp walks incrementIndex(1, nP, nC, nO, 1), the resonances and the
            // resolution, then evaluates the efficiency (Amp3Body_TD.cu:208-253)
            if(!expect(q, -1, 4, -1, "Amp3Body_TD"))
                return false;
            const int nres = (int)ctx->consts[q.c + 1];
            if(nres < 0 || nres > GF_MAX_RES || np(q) != 5 + 2 * nres || no(q) < 5 || no(q) > 7) {
                err = "Amp3Body_TD: bad resonance/parameter/observable counts";
                return false;
            }
            if(lvl >= GF_MAX_LVL) {
                err = "PDF tree nests deeper than " + std::to_string(GF_MAX_LVL) + " combining levels";
                return false;
            }
            int slot = -1;
            for(size_t s = 0; s < ctx->dalitz.size(); s++)
                if(ctx->dalitz[s].func == q.f)
                    slot = (int)s;
            if(slot < 0) {
                err = "Amp3Body_TD node not registered";
                return false;
            }
            const int pbase = q.p + 1;
            if(!inc(q))
                return false;
            for(int i = 0; i < nres + 1; i++)  // the resonances, then the resolution
                if(!inc(q))
                    return false;
            push(GF_LVL_OP(GK_TDDP, lvl), lvl, nres, pbase, 0, 0, 0, slot);
            if(!emit(q, lvl + 1))
                return false;
            push(GF_LVL_OP(GK_TDDP_END, lvl), lvl, 0, 0, 0, 0, 0, slot);
            return true;
        }
        case GF_FN_RBW:
        case GF_FN_RBW_SYM:
        case GF_FN_GS:
        case GF_FN_NONRES:
            err = "a resonance is only valid as a component of Amp3Body or Amp3Body_TD";
            return false;
        case GF_FN_TRUTH_RES:
        case GF_FN_THREEGAUSS_RES:
            err = "a time resolution is only valid as a component of Amp3Body_TD";
            return false;
        default: err = "unknown function code " + std::to_string(f); return false;
        }
    }
};

// Skip the subtree at q (no code emitted).
bool skip_subtree(Compiler &C, Cur &q) {
    if(!C.in_range(q)) {
        C.err = "subtree runs past the arrays";
        return false;
    }
    const int f  = C.ctx->funcs[q.f];
    int nchild   = 0;
    if(f == GF_FN_ADD)
        nchild = C.np(q) + 1;
    else if(f == GF_FN_ADD_EXT)
        nchild = C.np(q);
    else if(f == GF_FN_PROD)
        nchild = C.nc(q) >= 1 ? (int)C.ctx->consts[q.c + 1] : 0;
    else if(f == GF_FN_DALITZ)
        nchild = C.nc(q) >= 1 ? (int)C.ctx->consts[q.c + 1] + 1 : 0;
    else if(f == GF_FN_TDDP)  // resonances, resolution, efficiency
        nchild = C.nc(q) >= 1 ? (int)C.ctx->consts[q.c + 1] + 2 : 0;
    if(!C.inc(q))
        return false;
    for(int i = 0; i < nchild; i++)
        if(!skip_subtree(C, q))
            return false;
    return true;
}

bool cursor_at(Compiler &C, int func, Cur &q) {
    q = Cur();
    while(q.f < func)
        if(!C.inc(q))
            return false;
    return true;
}

// Discover the Dalitz nodes (resonance layout, efficiency index, observable columns).
gf_status scan_dalitz(gf_ctx *ctx) {
    ctx->dalitz.clear();
    ctx->has_td = false;
    Compiler C{ctx, {}, {}, -1};
    for(int f = 0; f < (int)ctx->funcs.size(); f++) {
        if(ctx->funcs[f] != GF_FN_DALITZ && ctx->funcs[f] != GF_FN_TDDP)
            continue;
        const bool td = ctx->funcs[f] == GF_FN_TDDP;
        if((int)ctx->dalitz.size() >= GF_MAX_DALITZ)
            return fail(GF_ERR_INVALID, "more than %d Amp3Body nodes in one tree", GF_MAX_DALITZ);
        Cur q;
        if(!cursor_at(C, f, q))
            return fail(GF_ERR_INVALID, "%s", C.err.c_str());
        DalitzNode d;
        d.func = f;
        d.td   = td;
        if(!td && (C.nc(q) != 2 || C.no(q) < 3))
            return fail(GF_ERR_INVALID, "Amp3Body at %d: expected 2 constants and >= 3 observables", f);
        if(td && (C.nc(q) != 4 || C.no(q) < 5 || C.no(q) > 7))
            return fail(GF_ERR_INVALID, "Amp3Body_TD at %d: expected 4 constants and 5 to 7 observables", f);
        d.nres = (int)ctx->consts[q.c + 1];
        // cacheToUse counts Amp3Body and Amp3Body_TD objects separately (Amp3Body.cu:157,
        // Amp3Body_TD.cu:313), so the two kinds key their caches apart
        d.cache_key = (int)ctx->consts[q.c + 2] + (td ? (1 << 20) : 0);
        if(d.nres < 0 || d.nres > GF_MAX_RES)
            return fail(GF_ERR_INVALID, "Amp3Body at %d: %d resonances (at most %d)", f, d.nres, GF_MAX_RES);
        if(td) {
            // observables dtime, sigma_t, m12, m13, event number[, mistag[, charm tag]]
            // (Amp3Body_TD.cu:262-295)
            int ct, cs, c12, c13, cnum;
            if(!C.col_of(q, 0, ct) || !C.col_of(q, 1, cs) || !C.col_of(q, 2, c12) || !C.col_of(q, 3, c13)
               || !C.col_of(q, 4, cnum))
                return fail(GF_ERR_INVALID, "%s", C.err.c_str());
            d.col_time  = ct;
            d.col_sigma = cs;
            d.col_m12   = c12;
            d.col_m13   = c13;
            d.col_num   = cnum;
            if(C.no(q) > 5 && !C.col_of(q, 5, d.col_mis))
                return fail(GF_ERR_INVALID, "%s", C.err.c_str());
            if(C.no(q) > 6 && !C.col_of(q, 6, d.col_tag))
                return fail(GF_ERR_INVALID, "%s", C.err.c_str());
            if(C.np(q) != 5 + 2 * d.nres)
                return fail(GF_ERR_INVALID, "Amp3Body_TD at %d: %d parameters for %d couplings", f, C.np(q), d.nres);
            d.p_mix  = q.p + 1;
            d.p_coup = q.p + 1 + 5;
            ctx->has_td = true;
        } else {
            int c12, c13, cnum;
            if(!C.col_of(q, 0, c12) || !C.col_of(q, 1, c13) || !C.col_of(q, 2, cnum))
                return fail(GF_ERR_INVALID, "%s", C.err.c_str());
            d.col_m12 = c12;
            d.col_m13 = c13;
            d.col_num = cnum;
            d.p_coup  = q.p + 1;
            if(C.np(q) < 2 * d.nres)
                return fail(GF_ERR_INVALID, "Amp3Body at %d: %d parameters for %d couplings", f, C.np(q), d.nres);
        }
        if(!C.inc(q))
            return fail(GF_ERR_INVALID, "%s", C.err.c_str());
        for(int i = 0; i < d.nres; i++) {
            if(!C.in_range(q))
                return fail(GF_ERR_INVALID, "Amp3Body at %d: resonance %d outside the arrays", f, i);
            int rf = ctx->funcs[q.f];
            GfResDesc r{};
            if(rf == GF_FN_RBW || rf == GF_FN_RBW_SYM || rf == GF_FN_GS) {
                if(C.np(q) != 2 || C.nc(q) != 3 || C.no(q) != 0)
                    return fail(GF_ERR_INVALID, "resonance at %d: expected 2 parameters and 3 constants", q.f);
                r.kind = rf == GF_FN_RBW ? GR_RBW : (rf == GF_FN_RBW_SYM ? GR_RBW_SYM : GR_GS);
            } else if(rf == GF_FN_NONRES) {
                r.kind = GR_NONRES;
            } else {
                return fail(GF_ERR_INVALID, "Amp3Body at %d: component %d is not a resonance", f, i);
            }
            r.p = q.p + 1;
            r.c = q.c + 1;
            d.res_funcs.push_back(q.f);
            d.res.push_back(r);
            if(!C.inc(q))
                return fail(GF_ERR_INVALID, "%s", C.err.c_str());
        }
        if(td) {
            // the resolution follows the resonances (Amp3Body_TD.cu:514-515)
            if(!C.in_range(q))
                return fail(GF_ERR_INVALID, "Amp3Body_TD at %d: no resolution function", f);
            const int rf = ctx->funcs[q.f];
            if(rf == GF_FN_TRUTH_RES) {
                if(C.np(q) != 0 || C.nc(q) != 0 || C.no(q) != 0)
                    return fail(GF_ERR_INVALID, "TruthResolution at %d: expected no parameters", q.f);
                d.res_kind = GT_TRUTH;
            } else if(rf == GF_FN_THREEGAUSS_RES) {
                if(C.np(q) != 13 || C.nc(q) != 0 || C.no(q) != 0)
                    return fail(GF_ERR_INVALID, "ThreeGaussResolution at %d: expected 13 parameters", q.f);
                d.res_kind = GT_THREEGAUSS;
            } else {
                return fail(GF_ERR_INVALID, "Amp3Body_TD at %d: component %d is not a time resolution", f, q.f);
            }
            d.res_func = q.f;
            d.res_p    = q.p + 1;
            if(!C.inc(q))
                return fail(GF_ERR_INVALID, "%s", C.err.c_str());
        }
        d.eff_root = q.f;
        if(!skip_subtree(C, q))
            return fail(GF_ERR_INVALID, "%s", C.err.c_str());
        // Amp3Body: the last function of the subtree (Amp3Body.cu:198-202);
        // Amp3Body_TD: the root of the efficiency subtree (Amp3Body_TD.cu:518)
        d.eff_func = td ? d.eff_root : q.f - 1;
        ctx->dalitz.push_back(d);
    }
    return GF_OK;
}

gf_status upload_dalitz_table(gf_ctx *ctx) {
    GfDalitzSlot slots[GF_MAX_DALITZ];
    GfResDesc res[GF_MAX_DALITZ * GF_MAX_RES];
    memset(slots, 0, sizeof(slots));
    memset(res, 0, sizeof(res));
    for(size_t s = 0; s < ctx->dalitz.size(); s++) {
        const DalitzNode &d = ctx->dalitz[s];
        auto it             = ctx->caches.find(d.cache_key);
        slots[s].wave       = it != ctx->caches.end() ? it->second.wave : nullptr;
        slots[s].ncache     = it != ctx->caches.end() ? it->second.n : 0;
        slots[s].col_m12    = d.col_m12;
        slots[s].col_m13    = d.col_m13;
        slots[s].col_num    = d.col_num;
        slots[s].nres       = d.nres;
        slots[s].wave_b     = it != ctx->caches.end() ? it->second.wave_b : nullptr;
        slots[s].col_time   = d.col_time;
        slots[s].col_sigma  = d.col_sigma;
        slots[s].col_mis    = d.col_mis;
        slots[s].col_tag    = d.col_tag;
        slots[s].res_kind   = d.res_kind;
        slots[s].res_p      = d.res_p;
        for(int i = 0; i < d.nres; i++)
            res[s * GF_MAX_RES + i] = d.res[i];
    }
    GF_HIP(hipMemcpyAsync(ctx->d_dz, slots, sizeof(slots), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(ctx->d_res, res, sizeof(res), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

gf_status upload_ins(gf_ctx *ctx) {
    gf_status st = ensure_dev(ctx->d_ins, ctx->cap_ins, ctx->ins.size(), ctx->stream);
    if(st)
        return st;
    GF_HIP(hipMemcpyAsync(ctx->d_ins, ctx->ins.data(), ctx->ins.size() * sizeof(GfIns), hipMemcpyHostToDevice,
                          ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

// Instruction index of the program that evaluates the subtree at `func` (compiled on first use).
gf_status subprogram(gf_ctx *ctx, int func, int &pc) {
    auto it = ctx->sub_pc.find(func);
    if(it != ctx->sub_pc.end()) {
        pc = it->second;
        return GF_OK;
    }
    Compiler C{ctx, {}, {}, -1};
    Cur q;
    if(func < 0 || func >= (int)ctx->funcs.size() || !cursor_at(C, func, q))
        return fail(GF_ERR_INVALID, "function index %d outside the tree", func);
    if(!C.emit(q, 0))
        return fail(GF_ERR_INVALID, "%s", C.err.c_str());
    C.push(GI_END, 0, 0, 0, 0, 0, 0, 0);
    pc = (int)ctx->ins.size();
    ctx->ins.insert(ctx->ins.end(), C.out.begin(), C.out.end());
    if(C.max_col > ctx->max_col)
        ctx->max_col = C.max_col;
    ctx->sub_pc[func] = pc;
    return upload_ins(ctx);
}

// The root program is flat (k_reduce_flat): [leaf, END], or one level-0 AddPdf /
// ProdPdf whose children are all leaves -- PUSH, (leaf, combine)..., [POP], END --
// with Gaussian / Exp / Exp-offset / ExpGaus leaves over columns < GF_FLAT_COLS.
bool flat_program(const gf_ctx *ctx) {
    if(!ctx->dalitz.empty() || ctx->max_col < 0 || ctx->max_col >= GF_FLAT_COLS)
        return false;
    auto leaf = [](int op) { return op == GI_GAUSS || op == GI_EXP || op == GI_EXP_OFF || op == GI_EXPGAUS; };
    const std::vector<GfIns> &I = ctx->ins;
    size_t pc = 0;
    for(; pc < I.size() && I[pc].op != GI_END; ++pc) {
        const int op = I[pc].op;
        if(op < GI_LVL) {
            if(!leaf(op))
                return false;
            continue;
        }
        const int kind = (op - GI_LVL) >> 2, lvl = (op - GI_LVL) & 3;
        if(lvl != 0 || kind == GK_DALITZ || kind == GK_DALITZ_END || kind == GK_TDDP || kind == GK_TDDP_END)
            return false;
    }
    return pc < I.size() && pc > 0;
}

// Flush pending parameter/normalisation updates through the pinned staging buffer.
gf_status flush_uploads(gf_ctx *ctx) {
    if(!ctx->params_dirty && !ctx->norms_dirty)
        return GF_OK;
    if(ctx->stage_in_flight) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stage_in_flight = false;
    }
    const size_t np = ctx->params.size(), nn = ctx->norms.size();
    if(np + nn > ctx->cap_stage) {
        if(ctx->h_stage)
            GF_HIP(hipHostFree(ctx->h_stage));
        ctx->h_stage   = nullptr;
        ctx->cap_stage = 0;
        size_t c       = (np + nn) * 2 + 64;
        GF_HIP(hipHostMalloc((void **)&ctx->h_stage, c * sizeof(double), hipHostMallocDefault));
        ctx->cap_stage = c;
    }
    if(ctx->params_dirty && np) {
        memcpy(ctx->h_stage, ctx->params.data(), np * sizeof(double));
        GF_HIP(hipMemcpyAsync(ctx->d_params, ctx->h_stage, np * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    if(ctx->norms_dirty && nn) {
        memcpy(ctx->h_stage + np, ctx->norms.data(), nn * sizeof(double));
        GF_HIP(hipMemcpyAsync(ctx->d_norms, ctx->h_stage + np, nn * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    ctx->params_dirty = ctx->norms_dirty = false;
    ctx->stage_in_flight                 = true;
    return GF_OK;
}

GfProgArgs prog_args(gf_ctx *ctx) {
    GfProgArgs a;
    a.ins          = ctx->d_ins;
    a.params       = ctx->d_params;
    a.consts       = ctx->d_consts;
    a.norms        = ctx->d_norms;
    a.obsid        = ctx->d_obsid;
    a.cols         = ctx->d_cols;
    a.dz           = ctx->d_dz;
    a.nevents      = ctx->n;
    a.identity_num = ctx->identity_ok ? 1 : 0;
    a.nobs         = ctx->ncols;
    a.total_events = (double)ctx->n;
    a.decay        = ctx->decay;
    a.td           = ctx->has_td ? 1 : 0;
    a.has_dz       = ctx->dalitz.empty() ? 0 : 1;
    a.err          = ctx->d_err;
    a.seq          = ctx->res_seq;
    a.mark         = nullptr;
    a.sweep_only   = 0;
    a.fuse_slots   = nullptr;
    a.fuse_result  = nullptr;
    a.coh_mx_nres  = 0;
    a.coh_group    = 0;
    return a;
}

int event_blocks(gf_ctx *ctx, int64_t n) {
    int64_t per = (int64_t)GF_BLOCK * 2;
    int64_t nb  = (n + per - 1) / per;
    int64_t cap = (int64_t)ctx->cu_count * ctx->tune[GF_TUNE_BLOCKS_PER_CU];
    if(cap > GF_MAX_PARTIALS)
        cap = GF_MAX_PARTIALS;
    if(nb > cap)
        nb = cap;
    if(nb < 1)
        nb = 1;
    return (int)nb;
}

gf_status check_events_for_program(gf_ctx *ctx, int metric) {
    if(ctx->ins.empty())
        return fail(GF_ERR_STATE, "no program installed (gf_program_set)");
    if(!ctx->col_base)  // an upload of zero events is a valid empty set: every sum is 0
        return fail(GF_ERR_STATE, "no events uploaded");
    if(ctx->max_col >= ctx->ncols)
        return fail(GF_ERR_STATE, "program reads observable column %d but only %d columns are uploaded", ctx->max_col,
                    ctx->ncols);
    if(metric < 0 || metric > GF_METRIC_CHISQ)
        return fail(GF_ERR_INVALID, "unknown metric %d", metric);
    if(metric >= GF_METRIC_BINAVG && ctx->ncols < 2)
        return fail(GF_ERR_STATE, "binned metric needs (bin entry, bin volume) columns");
    for(const DalitzNode &d : ctx->dalitz) {
        auto it = ctx->caches.find(d.cache_key);
        if(it == ctx->caches.end() || !it->second.wave)
            return fail(GF_ERR_STATE, "Amp3Body at %d has no wave cache (gf_dalitz_set_data_size)", d.func);
    }
    return GF_OK;
}

// Wait for the reduction launched last and read the block k_finish wrote to
// host-mapped memory ({sum, residual, error word}; the kernel clears the error).
// The host polls the block's sequence number for up to GF_TUNE_SPIN_US
// microseconds (a finished reduction is seen within ~1 us instead of after a
// stream synchronisation); after that, or when polling is off, it blocks on
// the stream, which also reports a failed kernel.
gf_status read_result(gf_ctx *ctx, double *sum, double *comp) {
    const volatile GfResult *h = ctx->h_res;  // written by the device
    const uint32_t want        = ctx->res_seq;
    bool seen                  = false;
    if(ctx->tune[GF_TUNE_SPIN_US] > 0) {
        const auto t0    = std::chrono::steady_clock::now();
        const auto limit = std::chrono::microseconds(ctx->tune[GF_TUNE_SPIN_US]);
        for(unsigned it = 0;; ++it) {
            if(__atomic_load_n(&ctx->h_res->seq, __ATOMIC_ACQUIRE) == want) {
                seen = true;
                break;
            }
            if((it & 255u) == 255u && std::chrono::steady_clock::now() - t0 > limit)
                break;
            __builtin_ia32_pause();
        }
    }
    if(!seen) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        if(__atomic_load_n(&ctx->h_res->seq, __ATOMIC_ACQUIRE) != want)
            return fail(GF_ERR_HIP, "reduction %u finished without writing its result", want);
    }
    // every command before k_finish on the stream is complete
    ctx->stage_in_flight = false;
    GfResult r;
    r.sum  = h->sum;
    r.comp = h->comp;
    r.err  = h->err;
    if(r.err & GF_DEVERR_FINISH)
        return fail(GF_ERR_HIP, "in-kernel finish: a workgroup's partial was not visible to the finishing workgroup "
                                "(error word %d)",
                    r.err);
    if(r.err & 2)
        return fail(GF_ERR_STATE, "internal: a Dalitz node reached a kernel compiled without it (error word %d)",
                    r.err);
    if(r.err)
        return fail(GF_ERR_INVALID, "Dalitz event number outside the wave cache (error word %d)", r.err);
    if(sum)
        *sum = r.sum;
    if(comp)
        *comp = r.comp;
    return GF_OK;
}

// Parameter and normalisation values for the kernel arguments (false if they do not fit).
bool make_inline(const gf_ctx *ctx, GfInline &inl) {
    if(!ctx->tune[GF_TUNE_INLINE_ARGS] || ctx->params.size() > GF_PIN || ctx->norms.size() > GF_NIN)
        return false;
    std::copy(ctx->params.begin(), ctx->params.end(), inl.p);
    std::copy(ctx->norms.begin(), ctx->norms.end(), inl.n);
    return true;
}

// The coherent-sum fast path applies to this launch.
bool use_fast_dalitz(const gf_ctx *ctx, int metric) {
    if(!ctx->tune[GF_TUNE_FAST_DALITZ] || !ctx->fast_dalitz || !ctx->identity_ok || metric >= GF_METRIC_BINAVG || ctx->dalitz.empty())
        return false;
    auto it = ctx->caches.find(ctx->dalitz[0].cache_key);
    return it != ctx->caches.end() && it->second.wave && it->second.n >= ctx->n;
}

// The flat kernel applies to this launch.
bool use_flat(const gf_ctx *ctx, int metric) {
    return ctx->tune[GF_TUNE_FLAT] != 0 && ctx->flat && metric < GF_METRIC_BINAVG;
}

// Events per lane of the flat kernel: 4 hides more latency on large sets (1e7 events of
// the 3-D ProdPdf: 123 us against 140 us with 2), 2 keeps small sets on more workgroups.
int flat_ept(const gf_ctx *ctx) {
    const int t = ctx->tune[GF_TUNE_FLAT];
    return t == 1 ? (ctx->n >= ((int64_t)1 << 20) ? 4 : 2) : t;
}

// The fast path's efficiency polynomial has maximum degree 0.
bool const_efficiency(const gf_ctx *ctx) { return (int)ctx->consts[ctx->ins[1].c] == 0; }

// The fast path may skip m12/m13: constant finite efficiency, and every resonance's wave
// computed from the current event columns starting at row 0, so the cache is exactly 0
// outside the Dalitz region (k_dalitz_cache writes 0 there) and |A|^2 eff is 0 there too.
bool zero_outside(const gf_ctx *ctx) {
    if(!ctx->tune[GF_TUNE_ZERO_OUTSIDE] || !const_efficiency(ctx) || ctx->dalitz.empty())
        return false;
    const DalitzNode &d = ctx->dalitz[0];
    auto it             = ctx->caches.find(d.cache_key);
    if(it == ctx->caches.end())
        return false;
    const WaveCache &w  = it->second;
    const uint64_t all = d.nres >= 64 ? ~0ull : ((1ull << d.nres) - 1ull);
    const double eff   = ctx->params[ctx->ins[1].p + ctx->ins[1].a];
    return w.offset == 0 && w.gen == ctx->events_gen && w.col12 == d.col_m12 && w.col13 == d.col_m13
           && (w.filled & all) == all && std::isfinite(eff);
}

// Efficiency handling of the point kernels (Gram, generator): a MultiPolynomial
// of maximum degree 0 is its single coefficient (device_MultiPolynomial,
// PolynomialPdf.cu:47-118, with numBoxes = 1), passed as a constant (mode -2);
// any other efficiency runs as a compiled subprogram (pc >= 0); none: -1.
gf_status efficiency_mode(gf_ctx *ctx, int eff_func, int &pc, double &value) {
    pc    = -1;
    value = 1.0;
    if(eff_func < 0)
        return GF_OK;
    if(ctx->funcs.at(eff_func) == GF_FN_MULTI_POLYNOMIAL) {
        Compiler C{ctx, {}, {}, -1};
        Cur q;
        if(cursor_at(C, eff_func, q) && C.nc(q) >= 1 && (int)ctx->consts[q.c + 1] == 0) {
            const int nobs = C.no(q);
            if(C.np(q) >= nobs + 1) {
                pc    = -2;
                value = ctx->params[q.p + 1 + nobs];
                return GF_OK;
            }
        }
    }
    return subprogram(ctx, eff_func, pc);
}

// The multi-point NLL of a lone Amp3Body on the matrix cores (k_coh_mx): the knob on, the
// result layout of v_mfma_f64_16x16x4f64 the one the kernel assumes (probed at context
// creation), and the NLL of a constant efficiency with the waves settling inDalitz.
// Returns the resonance count for GfProgArgs::coh_mx_nres, or 0.
int coh_mx_nres(const gf_ctx *ctx, int metric) {
    if(!ctx->tune[GF_TUNE_COH_MFMA] || ctx->mfma_layout != 2 || metric != GF_METRIC_NLL || ctx->dalitz.empty()
       || !const_efficiency(ctx) || !zero_outside(ctx))
        return 0;
    return ctx->dalitz[0].nres;
}

int fast_blocks(const gf_ctx *ctx, int64_t n) {
    int64_t nb  = (n + 255) / 256;
    int64_t cap = (int64_t)ctx->cu_count * ctx->tune[GF_TUNE_FAST_BLOCKS_PER_CU];
    if(cap > GF_MAX_PARTIALS)
        cap = GF_MAX_PARTIALS;
    return (int)std::max<int64_t>(1, std::min(nb, cap));
}

// Launch the likelihood reduction (fast path or interpreter) without waiting;
// mark (optional) is recorded between the event sweep and k_finish.
gf_status launch_reduce(gf_ctx *ctx, int metric, hipEvent_t mark = nullptr, bool sweep_only = false) {
    GfInline inl;
    const bool have_inl = make_inline(ctx, inl);
    gf_status st        = have_inl ? GF_OK : flush_uploads(ctx);
    if(st)
        return st;
    if(!sweep_only)
        ++ctx->res_seq;
    GfProgArgs a = prog_args(ctx);
    a.mark       = mark;
    a.sweep_only = sweep_only ? 1 : 0;
    const bool fast = use_fast_dalitz(ctx, metric);
    if(fast)
        a.coh_group = coh_mx_nres(ctx, metric) > 0;
    if(ctx->tune[GF_TUNE_FUSED_FINISH] && (fast || use_flat(ctx, metric))) {
        a.fuse_slots  = ctx->d_fuse_slots;
        a.fuse_result = ctx->d_res_mapped;
    }
    int rc;
    if(fast)
        rc = gfk::launch_dalitz_fast(a, have_inl ? &inl : nullptr, const_efficiency(ctx), zero_outside(ctx), metric, ctx->d_partials,
                                     ctx->d_res_mapped, nullptr, fast_blocks(ctx, ctx->n),
                                     ctx->stream);
    else if(use_flat(ctx, metric))
        rc = gfk::launch_reduce_flat(a, have_inl ? &inl : nullptr, metric, ctx->max_col + 1, flat_ept(ctx),
                                     ctx->d_partials, ctx->d_res_mapped, event_blocks(ctx, ctx->n), ctx->stream);
    else
        rc = gfk::launch_reduce_metric(a, have_inl ? &inl : nullptr, metric, ctx->d_partials, ctx->d_res_mapped,
                                       event_blocks(ctx, ctx->n), ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "reduce launch failed (%d)", rc);
    return GF_OK;
}

DalitzNode *find_dalitz(gf_ctx *ctx, int func, int *slot) {
    for(size_t s = 0; s < ctx->dalitz.size(); s++)
        if(ctx->dalitz[s].func == func) {
            if(slot)
                *slot = (int)s;
            return &ctx->dalitz[s];
        }
    return nullptr;
}

// Candidate result layouts of v_mfma_f64_16x16x4f64: which D[i][j] register r of
// lane l holds.  The probe in ensure_layout picks the one the hardware uses
// (on gfx950: layout 2, D[4r + l/16][l%16]).
void layout_ij(int layout, int l, int r, int &i, int &j) {
    switch(layout) {
    case 0: i = 4 * (l / 16) + r; j = l % 16; break;
    case 1: j = 4 * (l / 16) + r; i = l % 16; break;
    case 2: i = 4 * r + l / 16; j = l % 16; break;
    default: j = 4 * r + l / 16; i = l % 16; break;
    }
}

// Run one MFMA on host operands a[64], b[64]; out[256] = D registers by lane.
gf_status mfma_run(gf_ctx *ctx, const double *a, const double *b, double *out) {
    double *d = ctx->d_gram_tiles;  // 768 doubles: a at 0, b at 64, result at 256
    GF_HIP(hipMemcpyAsync(d, a, 64 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(d + 64, b, 64 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = gfk::mfma_probe(d, d + 64, d + 256, ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "MFMA probe launch failed (%d)", rc);
    GF_HIP(hipMemcpyAsync(out, d + 256, 256 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

gf_status ensure_layout(gf_ctx *ctx) {
    if(ctx->mfma_layout >= 0)
        return GF_OK;
    // A[i][0] = i+1, B[0][j] = 1, A[i][1] = 1, B[1][j] = 100*(j+1) under the operand
    // layout k_gram uses (A[i][k] in lane i + 16k, B[k][j] in lane j + 16k), so
    // D[i][j] = (i+1) + 100*(j+1); accept either result register layout.
    double a[64], b[64], out[256];
    for(int l = 0; l < 64; l++) {
        const int k = l >> 4, i = l & 15;
        a[l]        = (k == 0) ? (double)(i + 1) : (k == 1 ? 1.0 : 0.0);
        b[l]        = (k == 0) ? 1.0 : (k == 1 ? 100.0 * (double)(i + 1) : 0.0);
    }
    gf_status st = mfma_run(ctx, a, b, out);
    if(st)
        return st;
    bool ok[4] = {true, true, true, true};
    for(int l = 0; l < 64; l++)
        for(int r = 0; r < 4; r++)
            for(int c = 0; c < 4; c++) {
                int i, j;
                layout_ij(c, l, r, i, j);
                if(out[4 * l + r] != (double)(i + 1) + 100.0 * (double)(j + 1))
                    ok[c] = false;
            }
    for(int c = 0; c < 4; c++)
        if(ok[c]) {
            ctx->mfma_layout = c;
            return GF_OK;
        }
    return fail(GF_ERR_HIP, "unexpected v_mfma_f64_16x16x4f64 layout: lane0 %g %g %g %g lane1 %g lane16 %g lane17 %g",
                out[0], out[1], out[2], out[3], out[4], out[64], out[68]);
}

// Map the three MFMA tiles (RR, II, RI) into I_ij = sum e B_i conj(B_j), nres x nres.
void tiles_to_matrix(gf_ctx *ctx, const double *tiles, int nres, double *out) {
    double RR[16][16], II[16][16], RI[16][16];
    for(int t = 0; t < 3; t++)
        for(int r = 0; r < 4; r++)
            for(int l = 0; l < 64; l++) {
                int i, j;
                layout_ij(ctx->mfma_layout, l, r, i, j);
                double v = tiles[(t * 4 + r) * 64 + l];
                (t == 0 ? RR : (t == 1 ? II : RI))[i][j] = v;
            }
    for(int i = 0; i < nres; i++)
        for(int j = 0; j < nres; j++) {
            out[2 * (i * nres + j)]     = RR[i][j] + II[i][j];
            out[2 * (i * nres + j) + 1] = RI[j][i] - RI[i][j];
        }
}

// Gram of the node's lineshapes over grid cells (mode 0), points (1) or its data
// cache (2) into out (nres x nres complex).  stack/nstack (device array of at most
// GF_MAX_RES descriptors) replaces the node's resonance list when given;
// td_grid selects the time-dependent grid's m13 bin count.
gf_status run_gram(gf_ctx *ctx, const DalitzNode &d, int slot, int mode, const double *grid6, const double *m12,
                   const double *m13, const double *w, int64_t np, int eff_func, double *out,
                   const GfResDesc *stack = nullptr, int nstack = 0, int td_grid = 0) {
    gf_status st = ensure_layout(ctx);
    if(st)
        return st;
    int eff_pc      = -1;
    double eff_cval = 1.0;
    if(mode != 2 && eff_func >= 0) {
        st = efficiency_mode(ctx, eff_func, eff_pc, eff_cval);
        if(st)
            return st;
    }
    st = flush_uploads(ctx);
    if(st)
        return st;
    int64_t nb = (np + 127) / 128;
    int64_t cap = (int64_t)ctx->cu_count * 4;
    if(nb > cap)
        nb = cap;
    if(nb < 1)
        nb = 1;
    st = ensure_dev(ctx->d_gram_partials, ctx->cap_gram_partials, (size_t)nb * 768, ctx->stream);
    if(st)
        return st;
    GfProgArgs a         = prog_args(ctx);
    const GfResDesc *res = stack ? stack : ctx->d_res + slot * GF_MAX_RES;
    const int nres       = stack ? nstack : d.nres;
    int rc = gfk::launch_gram(a, mode, grid6, m12, m13, w, np, slot, d.col_m12, d.col_m13, eff_pc, eff_cval, td_grid,
                              res, nres, ctx->d_gram_partials, ctx->d_gram_tiles, (int)nb, ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "Gram launch failed (%d)", rc);
    GF_HIP(hipMemcpyAsync(ctx->h_gram_tiles, ctx->d_gram_tiles, 768 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stage_in_flight = false;
    tiles_to_matrix(ctx, ctx->h_gram_tiles, nres, out);
    return GF_OK;
}

}  // namespace

// ============================================================== C ABI

extern "C" {

int gf_device_count(void) {
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

const char *gf_last_error(void) { return g_last_error.c_str(); }

gf_status gf_ctx_set_tuning(gf_ctx *ctx, int key, int value) {
    if(!ctx)
        return fail(GF_ERR_INVALID, "null context");
    switch(key) {
    case GF_TUNE_FLAT:
        if(value != 0 && value != 1 && value != 2 && value != 4)
            return fail(GF_ERR_INVALID, "tuning key %d takes 0, 1, 2 or 4 (got %d)", key, value);
        break;
    case GF_TUNE_INLINE_ARGS:
    case GF_TUNE_FAST_DALITZ:
    case GF_TUNE_ZERO_OUTSIDE:
    case GF_TUNE_FUSED_FINISH:
    case GF_TUNE_COH_MFMA:
        if(value != 0 && value != 1)
            return fail(GF_ERR_INVALID, "tuning key %d takes 0 or 1 (got %d)", key, value);
        break;
    case GF_TUNE_BLOCKS_PER_CU:
    case GF_TUNE_FAST_BLOCKS_PER_CU:
        if(value < 1 || value > 64)
            return fail(GF_ERR_INVALID, "tuning key %d takes 1..64 workgroups per CU (got %d)", key, value);
        break;
    case GF_TUNE_BATCH_MX_PTS:
        if(value != 0 && (value < 2 || value > GF_MULTI_PTS + 1))
            return fail(GF_ERR_INVALID, "tuning key %d takes 0 or 2..%d points (got %d)", key, GF_MULTI_PTS + 1,
                        value);
        break;
    case GF_TUNE_SPIN_US:
        if(value < 0 || value > 1000000)
            return fail(GF_ERR_INVALID, "tuning key %d takes 0..1000000 microseconds (got %d)", key, value);
        break;
    default: return fail(GF_ERR_INVALID, "unknown tuning key %d", key);
    }
    if(key == GF_TUNE_INLINE_ARGS && value == 0 && ctx->tune[key] == 1) {
        // the device buffers may lag the host copies while values travel inline
        ctx->params_dirty = true;
        ctx->norms_dirty  = true;
    }
    ctx->tune[key] = value;
    return GF_OK;
}

gf_status gf_ctx_get_tuning(gf_ctx *ctx, int key, int *value) {
    if(!ctx || !value)
        return fail(GF_ERR_INVALID, "null argument");
    if(key < 0 || key >= GF_TUNE_COUNT)
        return fail(GF_ERR_INVALID, "unknown tuning key %d", key);
    *value = ctx->tune[key];
    return GF_OK;
}

gf_status gf_ctx_create(gf_ctx **out, int device) {
    if(!out)
        return fail(GF_ERR_INVALID, "null output pointer");
    *out  = nullptr;
    int n = gf_device_count();
    if(n <= 0)
        return fail(GF_ERR_NODEVICE, "no GPU visible to HIP");
    if(device < 0 || device >= n)
        return fail(GF_ERR_INVALID, "device %d out of range (%d visible)", device, n);
    GF_HIP(hipSetDevice(device));
    gf_status sst = noalloc_partials_init(device);
    if(sst)
        return sst;
    gf_ctx *c = new gf_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->cu_count = prop.multiProcessorCount;
    gf_status st = GF_OK;
#define TRY(x)                                                                                                         \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if(e_ != hipSuccess) {                                                                                         \
            st = fail(GF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));                                         \
            gf_ctx_destroy(c);                                                                                         \
            return st;                                                                                                 \
        }                                                                                                              \
    } while(0)
    TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    TRY(hipMalloc((void **)&c->d_cols, GF_MAX_OBS * sizeof(double *)));
    TRY(hipMalloc((void **)&c->d_dz, GF_MAX_DALITZ * sizeof(GfDalitzSlot)));
    TRY(hipMalloc((void **)&c->d_res, GF_MAX_DALITZ * GF_MAX_RES * sizeof(GfResDesc)));
    TRY(hipMalloc((void **)&c->d_partials, GF_MAX_PARTIALS * sizeof(gfk::Partial)));
    TRY(hipMalloc((void **)&c->d_counter, 64));
    TRY(hipMalloc((void **)&c->d_result, 64));
    TRY(hipMalloc((void **)&c->d_err, 64));
    TRY(hipMalloc((void **)&c->d_fuse_slots, 2 * GF_MAX_PARTIALS * sizeof(unsigned long long)));
    TRY(hipHostMalloc((void **)&c->h_res, sizeof(GfResult), hipHostMallocMapped | hipHostMallocCoherent));
    TRY(hipHostGetDevicePointer((void **)&c->d_res_mapped, c->h_res, 0));
    // Pinned blocks are recycled between contexts: a block freed by an earlier context
    // still holds that context's last {sum, seq}, and this context's first reduction
    // (seq 1) would read it as its own.  Start from seq 0.
    memset((void *)c->h_res, 0, sizeof(GfResult));
    TRY(hipMalloc((void **)&c->d_gram_tiles, 768 * sizeof(double)));
    TRY(hipHostMalloc((void **)&c->h_result, 64, hipHostMallocDefault));
    TRY(hipHostMalloc((void **)&c->h_gram_tiles, 768 * sizeof(double), hipHostMallocDefault));
    // device memory is written on the context's own stream: the stream is non-blocking,
    // so work on the null stream (plain hipMemset / hipMemcpy) is not ordered with it
    TRY(hipMemsetAsync(c->d_counter, 0, 64, c->stream));
    TRY(hipMemsetAsync(c->d_result, 0, 64, c->stream));
    TRY(hipMemsetAsync(c->d_err, 0, 64, c->stream));
    TRY(hipMemsetAsync(c->d_cols, 0, GF_MAX_OBS * sizeof(double *), c->stream));
    TRY(hipMemsetAsync(c->d_dz, 0, GF_MAX_DALITZ * sizeof(GfDalitzSlot), c->stream));
    if(gfk::launch_fuse_init(c->d_fuse_slots, c->stream)) {
        st = fail(GF_ERR_HIP, "in-kernel finish slots: init launch failed");
        gf_ctx_destroy(c);
        return st;
    }
    TRY(hipStreamSynchronize(c->stream));
    TRY(hipEventCreate(&c->ev0));
    TRY(hipEventCreate(&c->ev1));
#undef TRY
    // the result layout of v_mfma_f64_16x16x4f64, probed once here so that the choice of
    // the likelihood kernels (k_coh_mx needs layout 2) does not depend on call history
    if((st = ensure_layout(c))) {
        gf_ctx_destroy(c);
        return st;
    }
    *out = c;
    return GF_OK;
}

void gf_ctx_destroy(gf_ctx *c) {
    if(!c)
        return;
    (void)hipSetDevice(c->device);
    if(c->stream)
        (void)hipStreamSynchronize(c->stream);
    for(auto &kv : c->caches) {
        if(kv.second.wave)
            (void)hipFree(kv.second.wave);
        if(kv.second.wave_b)
            (void)hipFree(kv.second.wave_b);
    }
    for(auto &kv : c->norm_points)
        if(kv.second.buf)
            (void)hipFree(kv.second.buf);
    if(c->d_multi)
        (void)hipFree(c->d_multi);
    if(c->d_multi_partials)
        (void)hipFree(c->d_multi_partials);
    if(c->h_multi)
        (void)hipHostFree(c->h_multi);
    if(c->h_multi_out)
        (void)hipHostFree(c->h_multi_out);
    void *dev[] = {c->col_base,  c->d_cols,   c->d_params,   c->d_consts,        c->d_norms,
                   c->d_obsid,   c->d_ins,    c->d_dz,       c->d_res,           c->d_partials,
                   c->d_counter, c->d_result, c->d_err, c->d_gram_partials, c->d_gram_tiles,
                   c->d_fuse_slots};
    for(void *p : dev)
        if(p)
            (void)hipFree(p);
    if(c->h_result)
        (void)hipHostFree(c->h_result);
    if(c->h_res)
        (void)hipHostFree(c->h_res);
    if(c->h_gram_tiles)
        (void)hipHostFree(c->h_gram_tiles);
    if(c->h_stage)
        (void)hipHostFree(c->h_stage);
    if(c->ev0)
        (void)hipEventDestroy(c->ev0);
    if(c->ev1)
        (void)hipEventDestroy(c->ev1);
    for(hipEvent_t e : c->ev_sweep)
        (void)hipEventDestroy(e);
    if(c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

// ---------------------------------------------------------------- events

static gf_status alloc_columns(gf_ctx *ctx, int ncols, int64_t n) {
    if(ncols < 1 || ncols > GF_MAX_OBS)
        return fail(GF_ERR_INVALID, "column count %d not in [1,%d]", ncols, GF_MAX_OBS);
    if(n < 0)
        return fail(GF_ERR_INVALID, "negative event count");
    GF_HIP(hipStreamSynchronize(ctx->stream));
    if(ctx->col_base) {
        GF_HIP(hipFree(ctx->col_base));
        ctx->col_base = nullptr;
    }
    ctx->ncols       = ncols;
    ctx->n           = n;
    ctx->events_gen++;
    ctx->identity_ok = false;
    ctx->identity_col = -1;
    size_t bytes     = (size_t)ncols * (size_t)(n > 0 ? n : 1) * sizeof(double);
    if(hipMalloc((void **)&ctx->col_base, bytes) != hipSuccess) {
        ctx->col_base = nullptr;
        ctx->n        = 0;
        return fail(GF_ERR_NOMEM, "cannot allocate %zu bytes of event columns", bytes);
    }
    double *ptrs[GF_MAX_OBS];
    for(int k = 0; k < GF_MAX_OBS; k++)
        ptrs[k] = k < ncols ? ctx->col_base + (size_t)k * (size_t)n : ctx->col_base;
    GF_HIP(hipMemcpyAsync(ctx->d_cols, ptrs, sizeof(ptrs), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

gf_status gf_events_alloc(gf_ctx *ctx, int nobs, int64_t n) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    return alloc_columns(ctx, nobs, n);
}

gf_status gf_events_upload(gf_ctx *ctx, int nobs, int64_t n, const double *const *cols) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!cols)
        return fail(GF_ERR_INVALID, "null column array");
    for(int k = 0; k < nobs; k++)
        if(!cols[k] && n > 0)
            return fail(GF_ERR_INVALID, "column %d is null", k);
    gf_status st = alloc_columns(ctx, nobs, n);
    if(st)
        return st;
    for(int k = 0; k < nobs && n > 0; k++)
        GF_HIP(hipMemcpyAsync(ctx->col_base + (size_t)k * n, cols[k], (size_t)n * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

gf_status gf_events_upload_aos(gf_ctx *ctx, int nobs, int64_t n, const double *aos) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!aos && n > 0)
        return fail(GF_ERR_INVALID, "null event array");
    if(nobs < 1 || nobs > GF_MAX_OBS)
        return fail(GF_ERR_INVALID, "column count %d not in [1,%d]", nobs, GF_MAX_OBS);
    std::vector<std::vector<double>> cols(nobs, std::vector<double>((size_t)n));
    for(int64_t i = 0; i < n; i++)
        for(int k = 0; k < nobs; k++)
            cols[k][i] = aos[i * nobs + k];
    std::vector<const double *> ptrs(nobs);
    for(int k = 0; k < nobs; k++)
        ptrs[k] = cols[k].data();
    return gf_events_upload(ctx, nobs, n, ptrs.data());
}

gf_status gf_events_column(gf_ctx *ctx, int k, double **dev_ptr) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!dev_ptr || k < 0 || k >= ctx->ncols || !ctx->col_base)
        return fail(GF_ERR_INVALID, "column %d not allocated", k);
    *dev_ptr = ctx->col_base + (size_t)k * ctx->n;
    ctx->events_gen++;  // the caller may write the column
    return GF_OK;
}

gf_status gf_events_set_identity_column(gf_ctx *ctx, int k) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(k < 0 || k >= ctx->ncols || !ctx->col_base)
        return fail(GF_ERR_INVALID, "column %d not allocated", k);
    int32_t one    = 1;
    int32_t *dflag = reinterpret_cast<int32_t *>(ctx->d_counter) + 4;  // word 4 of the counter block
    GF_HIP(hipMemcpyAsync(dflag, &one, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    int rc = gfk::launch_identity_check(ctx->col_base + (size_t)k * ctx->n, ctx->n, dflag, ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "identity check launch failed (%d)", rc);
    int32_t flag = 0;
    GF_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    if(!flag) {
        ctx->identity_ok  = false;
        ctx->identity_col = -1;
        return fail(GF_ERR_INVALID, "column %d is not the row index 0..n-1", k);
    }
    ctx->identity_ok  = true;
    ctx->identity_col = k;
    return GF_OK;
}

// ---------------------------------------------------------------- program

gf_status gf_program_set(gf_ctx *ctx, const int32_t *funcs, int nfunc, const double *params, int np,
                         const double *consts, int nc, const double *obs, int no, const double *norms, int nn) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(nfunc < 1 || !funcs || np < 1 || !params || nc < 1 || !consts || no < 1 || !obs || nn < 2 || !norms)
        return fail(GF_ERR_INVALID, "empty flattened arrays");
    ctx->funcs.assign(funcs, funcs + nfunc);
    ctx->params.assign(params, params + np);
    ctx->consts.assign(consts, consts + nc);
    ctx->obs.assign(obs, obs + no);
    ctx->norms.assign(norms, norms + nn);
    ctx->ins.clear();
    ctx->sub_pc.clear();
    ctx->max_col = -1;
    gf_status st = scan_dalitz(ctx);
    if(st) {
        ctx->funcs.clear();
        return st;
    }
    Compiler C{ctx, {}, {}, -1};
    Cur q;
    if(!C.emit(q, 0)) {
        ctx->funcs.clear();
        return fail(GF_ERR_INVALID, "%s", C.err.c_str());
    }
    if(q.f != nfunc) {
        std::string e = "the tree at function 0 covers " + std::to_string(q.f) + " of " + std::to_string(nfunc)
                        + " functions";
        ctx->funcs.clear();
        return fail(GF_ERR_INVALID, "%s", e.c_str());
    }
    C.push(GI_END, 0, 0, 0, 0, 0, 0, 0);
    ctx->ins     = C.out;
    ctx->max_col = C.max_col;
    ctx->sub_pc[0] = 0;
    ctx->fast_dalitz = ctx->ins.size() >= 4 && ctx->ins[0].op == GF_LVL_OP(GK_DALITZ, 0)
                       && ctx->ins[1].op == GI_MULTI_POLY && ctx->ins[2].op == GF_LVL_OP(GK_DALITZ_END, 0)
                       && ctx->ins[3].op == GI_END;
    ctx->flat = flat_program(ctx);
    for(const DalitzNode &d : ctx->dalitz) {
        int cols[7] = {d.col_m12, d.col_m13, d.col_num, d.col_time, d.col_sigma, d.col_mis, d.col_tag};
        for(int c : cols)
            if(c > ctx->max_col)
                ctx->max_col = c;
    }
    std::vector<int32_t> obsid(no);
    for(int i = 0; i < no; i++)
        obsid[i] = (int32_t)obs[i];
    if((st = ensure_dev(ctx->d_params, ctx->cap_params, (size_t)np, ctx->stream))
       || (st = ensure_dev(ctx->d_consts, ctx->cap_consts, (size_t)nc, ctx->stream))
       || (st = ensure_dev(ctx->d_norms, ctx->cap_norms, (size_t)nn, ctx->stream))
       || (st = ensure_dev(ctx->d_obsid, ctx->cap_obsid, (size_t)no, ctx->stream)))
        return st;
    GF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stage_in_flight = false;
    GF_HIP(hipMemcpyAsync(ctx->d_params, params, np * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(ctx->d_consts, consts, nc * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(ctx->d_norms, norms, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(ctx->d_obsid, obsid.data(), no * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->params_dirty = ctx->norms_dirty = false;
    if((st = upload_ins(ctx)))
        return st;
    return upload_dalitz_table(ctx);
}

gf_status gf_params_update(gf_ctx *ctx, const double *params, int np) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(ctx->funcs.empty())
        return fail(GF_ERR_STATE, "no program installed");
    if(np != (int)ctx->params.size() || !params)
        return fail(GF_ERR_INVALID, "parameter array has %d entries, the program %zu", np, ctx->params.size());
    for(int i = 0; i < np; i++)
        if(ctx->params[i] != params[i]) {
            // counts are structure: a changed count would change the tree walk
            ctx->params.assign(params, params + np);
            ctx->params_dirty = true;
            break;
        }
    return GF_OK;
}

gf_status gf_norms_update(gf_ctx *ctx, const double *norms, int nn) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(ctx->funcs.empty())
        return fail(GF_ERR_STATE, "no program installed");
    if(nn != (int)ctx->norms.size() || !norms)
        return fail(GF_ERR_INVALID, "normalisation array has %d entries, the program %zu", nn, ctx->norms.size());
    for(int i = 0; i < nn; i++)
        if(ctx->norms[i] != norms[i]) {
            ctx->norms.assign(norms, norms + nn);
            ctx->norms_dirty = true;
            break;
        }
    return GF_OK;
}

gf_status gf_decay3_set(gf_ctx *ctx, const double decay[6]) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!decay)
        return fail(GF_ERR_INVALID, "null decay array");
    ctx->decay.motherMass          = decay[0];
    ctx->decay.daug1Mass           = decay[1];
    ctx->decay.daug2Mass           = decay[2];
    ctx->decay.daug3Mass           = decay[3];
    ctx->decay.meson_radius        = decay[4];
    ctx->decay.mother_meson_radius = decay[5];
    return GF_OK;
}

// ------------------------------------------------------------- reductions

gf_status gf_reduce_metric(gf_ctx *ctx, int metric, double *sum, double *comp) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if((st = launch_reduce(ctx, metric)))
        return st;
    return read_result(ctx, sum, comp);
}

gf_status gf_reduce_metric_noalloc(gf_ctx *ctx, int metric, gf_result *dev_out, void *stream) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!dev_out)
        return fail(GF_ERR_INVALID, "null result block");
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    GfInline inl;
    if(!make_inline(ctx, inl))
        return fail(GF_ERR_STATE,
                    "no-allocation reduction: %zu parameters / %zu normalisation values do not fit the kernel "
                    "arguments (%d / %d), or GF_TUNE_INLINE_ARGS is off",
                    ctx->params.size(), ctx->norms.size(), GF_PIN, GF_NIN);
    gfk::Partial *partials = noalloc_partials(ctx->device);
    if(!partials)
        return fail(GF_ERR_STATE, "no-allocation scratch of device %d not set up", ctx->device);
    const bool fast = use_fast_dalitz(ctx, metric);
    const bool flat = !fast && use_flat(ctx, metric);
    const int nblocks = fast ? fast_blocks(ctx, ctx->n) : event_blocks(ctx, ctx->n);
    if(nblocks > GF_NOALLOC_PARTIALS)
        return fail(GF_ERR_STATE, "no-allocation reduction: %d workgroups exceed the scratch of %d partials", nblocks,
                    GF_NOALLOC_PARTIALS);
    GfProgArgs a = prog_args(ctx);
    a.seq        = ++ctx->noalloc_seq;
    if(fast)
        a.coh_group = coh_mx_nres(ctx, metric) > 0;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    GfResult *res = reinterpret_cast<GfResult *>(dev_out);
    int rc;
    if(fast)
        rc = gfk::launch_dalitz_fast(a, &inl, const_efficiency(ctx), zero_outside(ctx), metric, partials, res,
                                     nullptr, nblocks, s);
    else if(flat)
        rc = gfk::launch_reduce_flat(a, &inl, metric, ctx->max_col + 1, flat_ept(ctx), partials, res,
                                     nblocks, s);
    else
        rc = gfk::launch_reduce_metric(a, &inl, metric, partials, res, nblocks, s);
    if(rc)
        return fail(GF_ERR_HIP, "no-allocation reduce launch failed (%d)", rc);
    return GF_OK;
}

gf_status gf_reduce_metric_multi(gf_ctx *ctx, int metric, int npts, const double *params, int np,
                                 const double *norms, int nn, double *out) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if(npts < 0 || (npts > 0 && (!params || !norms || !out)))
        return fail(GF_ERR_INVALID, "null arrays or negative point count");
    if(np != (int)ctx->params.size() || nn != (int)ctx->norms.size())
        return fail(GF_ERR_INVALID, "arrays of %d / %d entries, the program has %zu / %zu", np, nn,
                    ctx->params.size(), ctx->norms.size());
    if(npts == 0)
        return GF_OK;
    if(!use_fast_dalitz(ctx, metric)) {
        // one reduction per set, then the context's own arrays back
        const std::vector<double> p0 = ctx->params, n0 = ctx->norms;
        for(int k = 0; k < npts && !st; k++) {
            if(!(st = gf_params_update(ctx, params + (size_t)k * np, np))
               && !(st = gf_norms_update(ctx, norms + (size_t)k * nn, nn)) && !(st = launch_reduce(ctx, metric)))
                st = read_result(ctx, out + k, nullptr);
        }
        gf_params_update(ctx, p0.data(), np);
        gf_norms_update(ctx, n0.data(), nn);
        return st;
    }
    // coherent-sum path: GF_MULTI_PTS sets per sweep (GF_MX_SETS on the matrix cores), the
    // waves read once per sweep
    const size_t vals = (size_t)npts * (size_t)(np + nn);
    if(vals > ctx->cap_multi) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        if(ctx->d_multi)
            GF_HIP(hipFree(ctx->d_multi));
        if(ctx->h_multi)
            GF_HIP(hipHostFree(ctx->h_multi));
        if(ctx->h_multi_out)
            GF_HIP(hipHostFree(ctx->h_multi_out));
        ctx->d_multi = ctx->h_multi = ctx->d_multi_out = ctx->h_multi_out = nullptr;
        ctx->cap_multi = 0;
        const size_t c = vals + vals / 2 + 64;
        const size_t o = 2 * (c / 2 + 64);  // result pairs: at least 2 per set
        if(hipMalloc((void **)&ctx->d_multi, c * sizeof(double)) != hipSuccess
           || hipHostMalloc((void **)&ctx->h_multi, c * sizeof(double), hipHostMallocDefault) != hipSuccess
           || hipHostMalloc((void **)&ctx->h_multi_out, o * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent)
                  != hipSuccess
           || hipHostGetDevicePointer((void **)&ctx->d_multi_out, ctx->h_multi_out, 0) != hipSuccess)
            return fail(GF_ERR_NOMEM, "cannot allocate %zu parameter values for %d sets", c, npts);
        ctx->cap_multi = c;
    }
    if(!ctx->d_multi_partials
       && hipMalloc((void **)&ctx->d_multi_partials, (size_t)GF_MX_SETS * GF_MAX_PARTIALS * sizeof(gfk::Partial))
              != hipSuccess) {
        ctx->d_multi_partials = nullptr;
        return fail(GF_ERR_NOMEM, "cannot allocate the multi-set partials");
    }
    if(ctx->stage_in_flight) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stage_in_flight = false;
    }
    double *hp = ctx->h_multi, *hn = ctx->h_multi + (size_t)npts * np;
    memcpy(hp, params, (size_t)npts * np * sizeof(double));
    memcpy(hn, norms, (size_t)npts * nn * sizeof(double));
    GF_HIP(hipMemcpyAsync(ctx->d_multi, ctx->h_multi, vals * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // the zero-outside sweep needs every set's efficiency constant finite
    bool zo = zero_outside(ctx);
    for(int k = 0; k < npts && zo; k++)
        zo = std::isfinite(params[(size_t)k * np + ctx->ins[1].p + ctx->ins[1].a]);
    const bool ce = const_efficiency(ctx);
    GfProgArgs a  = prog_args(ctx);
    const int nb  = fast_blocks(ctx, ctx->n);
    a.coh_mx_nres = coh_mx_nres(ctx, metric);
    const int chunk = (a.coh_mx_nres > 0 && ce && zo) ? GF_MX_SETS : GF_MULTI_PTS;
    // the result slots' empty marker (no launch of this context that writes them is in
    // flight: every earlier request was read back before it returned)
    for(int k = 0; k < npts; k++) {
        const unsigned long long empty = GF_SLOT_EMPTY;
        memcpy(ctx->h_multi_out + 2 * (size_t)k, &empty, sizeof empty);
    }
    for(int k0 = 0; k0 < npts; k0 += chunk) {
        const int cnt = std::min(chunk, npts - k0);
        int rc = gfk::launch_dalitz_multi(a, ctx->dalitz[0].nres, ctx->d_multi + (size_t)k0 * np, np,
                                          ctx->d_multi + (size_t)npts * np + (size_t)k0 * nn, nn, cnt, ce, zo, metric,
                                          ctx->d_multi_partials, ctx->d_multi_out + 2 * (size_t)k0, nb, ctx->stream);
        if(rc)
            return fail(GF_ERR_HIP, "multi-set launch failed (%d)", rc);
    }
    // k_finish_multi stores each sum over the empty marker left below, after its residual
    // and a system-scope fence; the host polls the sums as read_result polls its block
    // (up to GF_TUNE_SPIN_US), then blocks on the stream, which also reports a failed kernel.
    volatile unsigned long long *res = (volatile unsigned long long *)ctx->h_multi_out;
    auto ready = [&]() {
        for(int k = 0; k < npts; k++)
            if(__atomic_load_n((const unsigned long long *)&res[2 * k], __ATOMIC_ACQUIRE) == GF_SLOT_EMPTY)
                return false;
        return true;
    };
    bool seen = false;
    if(ctx->tune[GF_TUNE_SPIN_US] > 0) {
        const auto t0    = std::chrono::steady_clock::now();
        const auto limit = std::chrono::microseconds(ctx->tune[GF_TUNE_SPIN_US]);
        for(unsigned it = 0;; ++it) {
            if(ready()) {
                seen = true;
                break;
            }
            if((it & 255u) == 255u && std::chrono::steady_clock::now() - t0 > limit)
                break;
            __builtin_ia32_pause();
        }
    }
    if(!seen) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        if(!ready())
            return fail(GF_ERR_HIP, "multi-set reduction finished without writing its results");
    }
    ctx->stage_in_flight = false;  // every command before the last k_finish_multi is complete
    for(int k = 0; k < npts; k++)
        out[k] = ctx->h_multi_out[2 * k];
    return GF_OK;
}

gf_status gf_eval_metric_device(gf_ctx *ctx, int metric, double *dev_out, int64_t n) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if(!dev_out || n != ctx->n)
        return fail(GF_ERR_INVALID, "output has %lld entries, %lld events uploaded", (long long)n, (long long)ctx->n);
    GfInline inl;
    const bool have_inl = make_inline(ctx, inl);
    if(!have_inl && (st = flush_uploads(ctx)))
        return st;
    GfProgArgs a = prog_args(ctx);
    int rc;
    if(use_fast_dalitz(ctx, metric))
        rc = gfk::launch_dalitz_fast(a, have_inl ? &inl : nullptr, const_efficiency(ctx), zero_outside(ctx), metric, ctx->d_partials,
                                     ctx->d_res_mapped, dev_out, fast_blocks(ctx, ctx->n),
                                     ctx->stream);
    else
        rc = gfk::launch_eval_metric(a, have_inl ? &inl : nullptr, metric, dev_out, event_blocks(ctx, ctx->n),
                                     ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "eval launch failed (%d)", rc);
    return GF_OK;
}

gf_status gf_eval_metric(gf_ctx *ctx, int metric, double *out, int64_t n) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!out || n != ctx->n)
        return fail(GF_ERR_INVALID, "output has %lld entries, %lld events uploaded", (long long)n, (long long)ctx->n);
    double *d_out = nullptr;
    if(hipMalloc((void **)&d_out, (size_t)(n > 0 ? n : 1) * sizeof(double)) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate the output of %lld events", (long long)n);
    gf_status st = gf_eval_metric_device(ctx, metric, d_out, n);
    if(!st) {
        int32_t err = 0;
        if(hipMemcpyAsync(out, d_out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess
           || hipMemcpyAsync(&err, ctx->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess
           || hipStreamSynchronize(ctx->stream) != hipSuccess)
            st = fail(GF_ERR_HIP, "copy of the evaluated events failed");
        else if(err) {
            (void)hipMemsetAsync(ctx->d_err, 0, sizeof(int32_t), ctx->stream);
            (void)hipStreamSynchronize(ctx->stream);
            st = (err & 2) ? fail(GF_ERR_STATE, "internal: a Dalitz node reached a kernel compiled without it "
                                                "(error word %d)", err)
                           : fail(GF_ERR_INVALID, "Dalitz event number outside the wave cache (error word %d)", err);
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    return st;
}

// Total bin count of a grid ranges[3 i + {0, 1, 2}] = {lower, upper, nbins}, i < k.
static gf_status grid_total(const double *ranges, int k, int64_t &total) {
    if(!ranges || k < 1 || k > GF_MAX_OBS)
        return fail(GF_ERR_INVALID, "grid needs 1..%d observables", GF_MAX_OBS);
    total = 1;
    for(int i = 0; i < k; i++) {
        double nb = floor(ranges[3 * i + 2] + 0.5);
        if(!(nb >= 1) || nb > 2147483647.0)
            return fail(GF_ERR_INVALID, "bin count %g of grid dimension %d", ranges[3 * i + 2], i);
        total *= (int64_t)nb;
        if(total > ((int64_t)1 << 40))
            return fail(GF_ERR_INVALID, "grid of more than 2^40 bins");
    }
    return GF_OK;
}

// Checks of a grid evaluation of function func_idx (gf_reduce_bins, gf_eval_bins): the
// subprogram's first instruction, the observable slot of each grid dimension and the
// bin count; fails for a subtree holding an Amp3Body without wave cache or an Amp3Body_TD.
static gf_status prepare_bins(gf_ctx *ctx, int func_idx, const double *ranges, int k, int &pc, int *slots,
                              int64_t &total) {
    if(ctx->funcs.empty())
        return fail(GF_ERR_STATE, "no program installed");
    gf_status st = grid_total(ranges, k, total);
    if(st)
        return st;
    pc = 0;
    if((st = subprogram(ctx, func_idx, pc)))
        return st;
    // observable slot of each grid dimension: the function's own observables in order
    Compiler C{ctx, {}, {}, -1};
    Cur q;
    if(!cursor_at(C, func_idx, q))
        return fail(GF_ERR_INVALID, "%s", C.err.c_str());
    int nobs_f = C.no(q);
    for(int i = 0; i < k; i++) {
        int idx = (k == 1) ? 0 : i;
        if(idx >= nobs_f)
            return fail(GF_ERR_INVALID, "function %d has %d observables, grid has %d", func_idx, nobs_f, k);
        if(!C.col_of(q, idx, slots[i]))
            return fail(GF_ERR_INVALID, "%s", C.err.c_str());
    }
    for(const DalitzNode &d : ctx->dalitz)
        if(d.func >= func_idx) {
            if(d.td) {
                // the bin kernels run without the time-dependent code: not for a subtree holding one
                Cur r;
                Compiler C2{ctx, {}, {}, -1};
                if(cursor_at(C2, func_idx, r) && skip_subtree(C2, r) && d.func < r.f)
                    return fail(GF_ERR_INVALID, "binned integral over a subtree holding Amp3Body_TD at %d", d.func);
            }
            auto it = ctx->caches.find(d.cache_key);
            if(it == ctx->caches.end() || !it->second.wave) {
                Cur r;
                // only an error if the Dalitz node sits inside this subtree
                Compiler C2{ctx, {}, {}, -1};
                if(cursor_at(C2, func_idx, r) && skip_subtree(C2, r) && d.func < r.f)
                    return fail(GF_ERR_STATE, "Amp3Body at %d has no wave cache", d.func);
            }
        }
    return flush_uploads(ctx);
}

gf_status gf_reduce_bins(gf_ctx *ctx, int func_idx, const double *ranges, int k, double *sum) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!sum)
        return fail(GF_ERR_INVALID, "null output");
    int pc = 0, slots[GF_MAX_OBS] = {0};
    int64_t total = 0;
    gf_status st  = prepare_bins(ctx, func_idx, ranges, k, pc, slots, total);
    if(st)
        return st;
    ++ctx->res_seq;
    GfProgArgs a = prog_args(ctx);
    int rc = gfk::launch_reduce_bins(a, pc, ranges, slots, k, total, ctx->d_partials,
                                     ctx->d_res_mapped, event_blocks(ctx, total), ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "grid reduce launch failed (%d)", rc);
    return read_result(ctx, sum, nullptr);
}

// Copy n doubles from a device buffer the call allocated to the host, then free it.
static gf_status finish_device_output(gf_ctx *ctx, double *dev, double *host, int64_t n, const char *what) {
    gf_status st = GF_OK;
    if(hipMemcpyAsync(host, dev, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess
       || hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = fail(GF_ERR_HIP, "copy of the %s failed", what);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dev);
    return st;
}

gf_status gf_eval_bins(gf_ctx *ctx, int func_idx, const double *ranges, int k, double *out, int64_t nbins) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!out)
        return fail(GF_ERR_INVALID, "null output");
    int pc = 0, slots[GF_MAX_OBS] = {0};
    int64_t total = 0;
    gf_status st  = prepare_bins(ctx, func_idx, ranges, k, pc, slots, total);
    if(st)
        return st;
    if(nbins != total)
        return fail(GF_ERR_INVALID, "output holds %lld values, the grid has %lld bins", (long long)nbins,
                    (long long)total);
    double *d_out = nullptr;
    if(hipMalloc((void **)&d_out, (size_t)total * sizeof(double)) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate %lld bin values", (long long)total);
    GfProgArgs a = prog_args(ctx);
    int rc = gfk::launch_eval_bins(a, pc, ranges, slots, k, total, d_out, event_blocks(ctx, total), ctx->stream);
    if(rc) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(d_out);
        return fail(GF_ERR_HIP, "grid evaluation launch failed (%d)", rc);
    }
    return finish_device_output(ctx, d_out, out, total, "bin values");
}

gf_status gf_grid_cells(gf_ctx *ctx, const double *ranges, int k, double *out, int64_t nbins) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!out)
        return fail(GF_ERR_INVALID, "null output");
    int64_t total = 0;
    gf_status st  = grid_total(ranges, k, total);
    if(st)
        return st;
    if(nbins != total)
        return fail(GF_ERR_INVALID, "output holds %lld cells, the grid has %lld bins", (long long)nbins,
                    (long long)total);
    double *d_out = nullptr;
    if(hipMalloc((void **)&d_out, (size_t)total * k * sizeof(double)) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate %lld cell centres", (long long)(total * k));
    int rc = gfk::launch_grid_cells(ranges, k, total, d_out, ctx->stream);
    if(rc) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(d_out);
        return fail(GF_ERR_HIP, "grid cell launch failed (%d)", rc);
    }
    return finish_device_output(ctx, d_out, out, total * k, "cell centres");
}

gf_status gf_dalitz_grid_cells(gf_ctx *ctx, const double grid6[6], int time_dependent, double *m12, double *m13,
                               int64_t ncells) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!grid6 || !m12 || !m13)
        return fail(GF_ERR_INVALID, "null grid or output");
    double n12 = floor(grid6[2] + 0.5), n13 = floor(grid6[5] + 0.5);
    if(!(n12 >= 1) || !(n13 >= 1) || n12 * n13 > 2147483647.0)
        return fail(GF_ERR_INVALID, "bad grid bin counts %g x %g", grid6[2], grid6[5]);
    if(ncells != (int64_t)n12 * (int64_t)n13)
        return fail(GF_ERR_INVALID, "output holds %lld cells, the grid has %lld", (long long)ncells,
                    (long long)((int64_t)n12 * (int64_t)n13));
    double *d_out = nullptr;
    if(hipMalloc((void **)&d_out, (size_t)ncells * 2 * sizeof(double)) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate %lld cell centres", (long long)ncells);
    int rc = gfk::launch_dalitz_cells(grid6, time_dependent ? 1 : 0, ncells, d_out, d_out + ncells, ctx->stream);
    if(rc) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(d_out);
        return fail(GF_ERR_HIP, "Dalitz cell launch failed (%d)", rc);
    }
    gf_status st = GF_OK;
    if(hipMemcpyAsync(m12, d_out, (size_t)ncells * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess
       || hipMemcpyAsync(m13, d_out + ncells, (size_t)ncells * sizeof(double), hipMemcpyDeviceToHost, ctx->stream)
              != hipSuccess
       || hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = fail(GF_ERR_HIP, "copy of the Dalitz cell centres failed");
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    return st;
}

// ---------------------------------------------------------------- Dalitz

gf_status gf_dalitz_set_data_size(gf_ctx *ctx, int dalitz_func, int64_t n, int64_t offset) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, nullptr);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(n < 1 || offset < 0)
        return fail(GF_ERR_INVALID, "bad cache size %lld / offset %lld", (long long)n, (long long)offset);
    WaveCache &wc = ctx->caches[d->cache_key];
    GF_HIP(hipStreamSynchronize(ctx->stream));
    if(wc.wave) {
        GF_HIP(hipFree(wc.wave));
        wc.wave = nullptr;
    }
    if(wc.wave_b) {
        GF_HIP(hipFree(wc.wave_b));
        wc.wave_b = nullptr;
    }
    wc.n = 0;
    // GF_MAX_RES waves per event (Amp3Body.cu:222-233); a time-dependent node keeps
    // its B waves beside them (WaveHolder_s, Amp3Body_TD.cu:531-580)
    size_t bytes = (size_t)GF_MAX_RES * (size_t)n * sizeof(double2);
    if(hipMalloc((void **)&wc.wave, bytes) != hipSuccess) {
        wc.wave = nullptr;
        return fail(GF_ERR_NOMEM, "cannot allocate a wave cache of %zu bytes", bytes);
    }
    // zeroed on the context stream, ordered before the kernels that fill it (a memset on
    // the null stream can still be running when they write)
    GF_HIP(hipMemsetAsync(wc.wave, 0, bytes, ctx->stream));
    if(d->td) {
        if(hipMalloc((void **)&wc.wave_b, bytes) != hipSuccess) {
            wc.wave_b = nullptr;
            (void)hipFree(wc.wave);
            wc.wave = nullptr;
            return fail(GF_ERR_NOMEM, "cannot allocate a wave cache of %zu bytes", bytes);
        }
        GF_HIP(hipMemsetAsync(wc.wave_b, 0, bytes, ctx->stream));
    }
    GF_HIP(hipStreamSynchronize(ctx->stream));
    wc.n      = n;
    wc.offset = offset;
    wc.filled = 0;
    return upload_dalitz_table(ctx);
}

gf_status gf_dalitz_cache(gf_ctx *ctx, int dalitz_func, uint64_t res_mask) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, &slot);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    auto it = ctx->caches.find(d->cache_key);
    if(it == ctx->caches.end() || !it->second.wave)
        return fail(GF_ERR_STATE, "no wave cache (gf_dalitz_set_data_size)");
    const WaveCache &wc = it->second;
    if(!ctx->col_base || wc.offset + wc.n > ctx->n)
        return fail(GF_ERR_STATE, "wave cache rows %lld..%lld exceed the %lld uploaded events", (long long)wc.offset,
                    (long long)(wc.offset + wc.n), (long long)ctx->n);
    if(d->col_m12 >= ctx->ncols || d->col_m13 >= ctx->ncols)
        return fail(GF_ERR_STATE, "m12/m13 columns not uploaded");
    uint32_t mask = (uint32_t)(res_mask & ((d->nres >= 32) ? 0xffffffffu : ((1u << d->nres) - 1u)));
    if(!mask)
        return GF_OK;
    gf_status st = flush_uploads(ctx);
    if(st)
        return st;
    int64_t nb = (wc.n + 255) / 256;
    if(nb > (int64_t)ctx->cu_count * 8)
        nb = (int64_t)ctx->cu_count * 8;
    GfProgArgs a = prog_args(ctx);
    int rc = gfk::launch_dalitz_cache(a, slot, wc.offset, wc.n, mask, ctx->d_res + slot * GF_MAX_RES, d->nres,
                                      (int)(nb < 1 ? 1 : nb), ctx->stream);
    if(rc)
        return fail(GF_ERR_HIP, "cache launch failed (%d)", rc);
    GF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stage_in_flight = false;
    WaveCache &w = it->second;
    if(w.gen != ctx->events_gen || w.col12 != d->col_m12 || w.col13 != d->col_m13) {
        w.filled = 0;
        w.gen    = ctx->events_gen;
        w.col12  = d->col_m12;
        w.col13  = d->col_m13;
    }
    w.filled |= mask;
    return GF_OK;
}

gf_status gf_dalitz_gram(gf_ctx *ctx, int dalitz_func, const double grid6[6], int eff_func, double *out_re_im) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, &slot);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(d->td)
        return fail(GF_ERR_INVALID, "function %d is an Amp3Body_TD node (gf_tddp_gram)", dalitz_func);
    if(!grid6 || !out_re_im)
        return fail(GF_ERR_INVALID, "null grid or output");
    double n12 = floor(grid6[2] + 0.5), n13 = floor(grid6[5] + 0.5);
    // the cell index is an int (SpecialResonanceIntegrator.cu:59)
    if(!(n12 >= 1) || !(n13 >= 1) || n12 * n13 > 2147483647.0)
        return fail(GF_ERR_INVALID, "bad grid bin counts %g x %g", grid6[2], grid6[5]);
    if(eff_func >= (int)ctx->funcs.size())
        return fail(GF_ERR_INVALID, "efficiency function %d outside the tree", eff_func);
    return run_gram(ctx, *d, slot, 0, grid6, nullptr, nullptr, nullptr, (int64_t)n12 * (int64_t)n13, eff_func,
                    out_re_im);
}

gf_status gf_dalitz_gram_points_device(gf_ctx *ctx, int dalitz_func, const double *dev_m12, const double *dev_m13,
                                       const double *dev_w, int64_t np, int eff_func, double *out_re_im) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, &slot);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(d->td)
        return fail(GF_ERR_INVALID, "point normalisation is not available for Amp3Body_TD");
    if(!dev_m12 || !dev_m13 || !out_re_im || np < 0)
        return fail(GF_ERR_INVALID, "null point arrays or output");
    if(eff_func >= (int)ctx->funcs.size())
        return fail(GF_ERR_INVALID, "efficiency function %d outside the tree", eff_func);
    return run_gram(ctx, *d, slot, 1, nullptr, dev_m12, dev_m13, dev_w, np, eff_func, out_re_im);
}

gf_status gf_dalitz_gram_points(gf_ctx *ctx, int dalitz_func, const double *m12, const double *m13, const double *w,
                                int64_t np, int eff_func, double *out_re_im) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!m12 || !m13 || np < 1)
        return fail(GF_ERR_INVALID, "null or empty point arrays");
    double *buf  = nullptr;
    size_t bytes = (size_t)np * sizeof(double) * (w ? 3 : 2);
    if(hipMalloc((void **)&buf, bytes) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate %zu bytes of points", bytes);
    gf_status st = GF_OK;
    if(hipMemcpyAsync(buf, m12, np * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess
       || hipMemcpyAsync(buf + np, m13, np * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess
       || (w && hipMemcpyAsync(buf + 2 * np, w, np * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
       || hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = fail(GF_ERR_HIP, "copy of the points failed");
    if(!st)
        st = gf_dalitz_gram_points_device(ctx, dalitz_func, buf, buf + np, w ? buf + 2 * np : nullptr, np, eff_func,
                                          out_re_im);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    return st;
}

gf_status gf_dalitz_get_cache(gf_ctx *ctx, int dalitz_func, int res, double *out_re_im, int64_t n) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, nullptr);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    auto it = ctx->caches.find(d->cache_key);
    if(it == ctx->caches.end() || !it->second.wave)
        return fail(GF_ERR_STATE, "no wave cache");
    const int nwaves = d->td ? 2 * d->nres : d->nres;  // time-dependent: A waves, then B waves
    if(res < 0 || res >= nwaves || !out_re_im || n != it->second.n)
        return fail(GF_ERR_INVALID, "wave %d / size %lld do not match the cache (%d, %lld)", res, (long long)n,
                    nwaves, (long long)it->second.n);
    const double2 *src = res < d->nres ? it->second.wave + (size_t)res * it->second.n
                                       : it->second.wave_b + (size_t)(res - d->nres) * it->second.n;
    GF_HIP(hipMemcpyAsync(out_re_im, src, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

gf_status gf_dalitz_cache_gram(gf_ctx *ctx, int dalitz_func, double *out_re_im) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, &slot);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(d->td)
        return fail(GF_ERR_INVALID, "function %d is an Amp3Body_TD node", dalitz_func);
    auto it = ctx->caches.find(d->cache_key);
    if(it == ctx->caches.end() || !it->second.wave)
        return fail(GF_ERR_STATE, "no wave cache");
    if(!out_re_im)
        return fail(GF_ERR_INVALID, "null output");
    return run_gram(ctx, *d, slot, 2, nullptr, nullptr, nullptr, nullptr, it->second.n, -1, out_re_im);
}

gf_status gf_tddp_gram(gf_ctx *ctx, int tddp_func, const double grid6[6], double *out_aa, double *out_ab,
                       double *out_bb) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, tddp_func, &slot);
    if(!d || !d->td)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body_TD node", tddp_func);
    if(!grid6 || !out_aa || !out_ab || !out_bb)
        return fail(GF_ERR_INVALID, "null grid or output");
    double n12 = floor(grid6[2] + 0.5), n13 = floor(grid6[5] + 0.5);
    // the cell index is an int (SpecialDalitzIntegrator.cu:65)
    if(!(n12 >= 1) || !(n13 >= 1) || n12 * n13 > 2147483647.0)
        return fail(GF_ERR_INVALID, "bad grid bin counts %g x %g", grid6[2], grid6[5]);
    const int n = d->nres;
    if(n == 0)
        return GF_OK;
    // Stacked waves v = (a_0..a_{n-1}, b_0..b_{n-1}); the Hermitian Gram
    // G_kl = sum_cells eff v_k conj(v_l) holds ai aj*, ai bj* and bi bj* as its blocks.
    // One launch covers 16 entries; above that, groups of 8 are taken in pairs.
    const int m = 2 * n;
    std::vector<GfResDesc> all(m);
    for(int k = 0; k < n; k++) {
        all[k]          = d->res[k];
        all[k].swap     = 0;
        all[n + k]      = d->res[k];
        all[n + k].swap = 1;
    }
    std::vector<int> groups;  // start of each group; the last group ends at m
    const int gsize = m <= GF_MAX_RES ? m : GF_MAX_RES / 2;
    for(int g = 0; g < m; g += gsize)
        groups.push_back(g);
    const int ng = (int)groups.size();
    std::vector<std::complex<double>> G((size_t)m * m);
    std::vector<char> have((size_t)m * m, 0);
    GfResDesc *dstack = nullptr;
    if(hipMalloc((void **)&dstack, GF_MAX_RES * sizeof(GfResDesc)) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate the Gram stack");
    gf_status st = GF_OK;
    for(int u = 0; u < ng && !st; u++) {
        for(int v = (ng == 1 ? 0 : u + 1); v < ng && !st; v++) {
            std::vector<int> idx;  // stacked entries of this launch
            for(int g : {u, v}) {
                if(g == v && u == v && !idx.empty())
                    break;
                const int lo = groups[g], hi = std::min(m, lo + gsize);
                for(int k = lo; k < hi; k++)
                    idx.push_back(k);
            }
            std::vector<GfResDesc> st_res(idx.size());
            for(size_t k = 0; k < idx.size(); k++)
                st_res[k] = all[idx[k]];
            if(hipMemcpyAsync(dstack, st_res.data(), st_res.size() * sizeof(GfResDesc), hipMemcpyHostToDevice,
                              ctx->stream) != hipSuccess
               || hipStreamSynchronize(ctx->stream) != hipSuccess) {
                st = fail(GF_ERR_HIP, "copy of the Gram stack failed");
                break;
            }
            const int ns = (int)idx.size();
            std::vector<double> out(2 * (size_t)ns * ns);
            st = run_gram(ctx, *d, slot, 0, grid6, nullptr, nullptr, nullptr, (int64_t)n12 * (int64_t)n13,
                          d->eff_func, out.data(), dstack, ns, 1);
            for(int i = 0; i < ns && !st; i++)
                for(int j = 0; j < ns; j++) {
                    const size_t at = (size_t)idx[i] * m + idx[j];
                    if(!have[at]) {
                        G[at]    = {out[2 * (i * ns + j)], out[2 * (i * ns + j) + 1]};
                        have[at] = 1;
                    }
                }
        }
    }
    (void)hipFree(dstack);
    if(st)
        return st;
    for(int i = 0; i < n; i++)
        for(int j = 0; j < n; j++) {
            const size_t o = 2 * ((size_t)i * n + j);
            const std::complex<double> aa = G[(size_t)i * m + j], ab = G[(size_t)i * m + n + j],
                                       bb = G[(size_t)(n + i) * m + n + j];
            out_aa[o] = aa.real(), out_aa[o + 1] = aa.imag();
            out_ab[o] = ab.real(), out_ab[o + 1] = ab.imag();
            out_bb[o] = bb.real(), out_bb[o + 1] = bb.imag();
        }
    return GF_OK;
}

// Internal hook (not in the public header): one FP64 MFMA on host operands.
gf_status gf_internal_mfma_f64_probe(gf_ctx *ctx, const double *a64, const double *b64, double *out256) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!a64 || !b64 || !out256)
        return fail(GF_ERR_INVALID, "null operand");
    return mfma_run(ctx, a64, b64, out256);
}

// ---------------------------------------------------------------- timing

gf_status gf_time_reduce_metric(gf_ctx *ctx, int metric, int reps, double *ms_total) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if(reps < 1 || !ms_total)
        return fail(GF_ERR_INVALID, "reps must be >= 1");
    GF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for(int r = 0; r < reps; r++)
        if((st = launch_reduce(ctx, metric)))
            return st;
    GF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    GF_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    GF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_total = ms;
    return read_result(ctx, nullptr, nullptr);
}

gf_status gf_time_reduce_sweep(gf_ctx *ctx, int metric, int reps, double *ms_sweep, double *ms_total) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if(reps < 1 || reps > 4096 || !ms_sweep)
        return fail(GF_ERR_INVALID, "reps must be in 1..4096");
    while((int)ctx->ev_sweep.size() < 2 * reps + 1) {
        hipEvent_t e;
        GF_HIP(hipEventCreate(&e));
        ctx->ev_sweep.push_back(e);
    }
    std::vector<hipEvent_t> &ev = ctx->ev_sweep;
    GF_HIP(hipEventRecord(ev[0], ctx->stream));
    for(int r = 0; r < reps; r++) {
        if(r)
            GF_HIP(hipEventRecord(ev[2 * r], ctx->stream));
        if((st = launch_reduce(ctx, metric, ev[2 * r + 1])))
            return st;
    }
    GF_HIP(hipEventRecord(ev[2 * reps], ctx->stream));
    GF_HIP(hipEventSynchronize(ev[2 * reps]));
    double sweep = 0.0;
    for(int r = 0; r < reps; r++) {
        float ms = 0;
        GF_HIP(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
        sweep += ms;
    }
    *ms_sweep = sweep;
    if(ms_total) {
        float ms = 0;
        GF_HIP(hipEventElapsedTime(&ms, ev[0], ev[2 * reps]));
        *ms_total = ms;
    }
    return read_result(ctx, nullptr, nullptr);
}

gf_status gf_time_sweep_back_to_back(gf_ctx *ctx, int metric, int reps, double *ms_per_sweep) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if(reps < 1 || reps > 4096 || !ms_per_sweep)
        return fail(GF_ERR_INVALID, "reps must be in 1..4096");
    if(ctx->stage_in_flight) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stage_in_flight = false;
    }
    GF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for(int r = 0; r < reps; r++)
        if((st = launch_reduce(ctx, metric, nullptr, true)))
            return st;
    GF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    GF_HIP(hipEventSynchronize(ctx->ev1));
    ctx->stage_in_flight = false;  // every staged upload has been consumed
    float ms = 0;
    GF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_per_sweep = ms / reps;
    return GF_OK;
}

gf_status gf_time_dalitz_cache(gf_ctx *ctx, int dalitz_func, int reps, double *ms_total) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(reps < 1 || !ms_total)
        return fail(GF_ERR_INVALID, "reps must be >= 1");
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, &slot);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    auto it = ctx->caches.find(d->cache_key);
    if(it == ctx->caches.end() || !it->second.wave || it->second.offset + it->second.n > ctx->n)
        return fail(GF_ERR_STATE, "no wave cache over the uploaded events");
    gf_status st = flush_uploads(ctx);
    if(st)
        return st;
    const WaveCache &wc = it->second;
    int64_t nb          = (wc.n + 255) / 256;
    if(nb > (int64_t)ctx->cu_count * 8)
        nb = (int64_t)ctx->cu_count * 8;
    GfProgArgs a  = prog_args(ctx);
    uint32_t mask = d->nres >= 32 ? 0xffffffffu : ((1u << d->nres) - 1u);
    GF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for(int r = 0; r < reps; r++) {
        int rc = gfk::launch_dalitz_cache(a, slot, wc.offset, wc.n, mask, ctx->d_res + slot * GF_MAX_RES, d->nres,
                                          (int)(nb < 1 ? 1 : nb), ctx->stream);
        if(rc)
            return fail(GF_ERR_HIP, "cache launch failed (%d)", rc);
    }
    GF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    GF_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    GF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_total            = ms;
    ctx->stage_in_flight = false;
    return GF_OK;
}

gf_status gf_time_dalitz_gram(gf_ctx *ctx, int dalitz_func, const double grid6[6], int eff_func, int reps,
                              double *ms_total) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(reps < 1 || !ms_total || !grid6)
        return fail(GF_ERR_INVALID, "reps must be >= 1");
    std::vector<double> tmp(2 * GF_MAX_RES * GF_MAX_RES);
    // first call compiles the efficiency program and probes the MFMA layout
    gf_status st = gf_dalitz_gram(ctx, dalitz_func, grid6, eff_func, tmp.data());
    if(st)
        return st;
    GF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for(int r = 0; r < reps; r++) {
        st = gf_dalitz_gram(ctx, dalitz_func, grid6, eff_func, tmp.data());
        if(st)
            return st;
    }
    GF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    GF_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    GF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_total = ms;
    return GF_OK;
}


// ------------------------------------------------------------ learned proposal

}  // extern "C"

struct gf_mlp {
    gf_ctx *ctx = nullptr;
    GfMlpDev D{};
    int adam_t     = 0;
    bool use_graph = true;  // replay one captured training step
    // tabulated proposal
    double *d_cdf = nullptr, *d_pcell = nullptr;
    int G         = 0;
    GfBox box{};
};

namespace {

void mlp_free(gf_mlp *m) {
    void *p[] = {m->D.P,     m->D.m,       m->D.v,       m->D.grad,     m->D.w2b,     m->D.w2t,   m->D.xy,
                 m->D.z1,    m->D.z2,      m->D.h1,      m->D.delta2,   m->D.out,     m->D.coef,  m->D.loss,
                 m->D.part_in, m->D.part_w2, m->D.part_out, m->d_cdf, m->d_pcell, m->D.state};
    for(void *q : p)
        if(q)
            (void)hipFree(q);
}

// The kernels assume lane l of v_mfma_f32_16x16x32_bf16 takes A[l%16][8(l/16)+t],
// B[8(l/16)+t][l%16] and returns D[GF_BF16_D_ROW(l, r)][l%16]: verify on the device.
gf_status mlp_layout_check(gf_ctx *ctx) {
    std::vector<uint16_t> a8(64 * 8), b8(64 * 8);
    std::vector<float> A(16 * 32), B(32 * 16), d(256);
    std::mt19937 g(7);
    auto bf = [](float x) {
        uint32_t u;
        memcpy(&u, &x, 4);
        return (uint16_t)(u >> 16);  // small integers: exact
    };
    for(auto &x : A)
        x = (float)((int)(g() % 7) - 3);
    for(auto &x : B)
        x = (float)((int)(g() % 7) - 3);
    for(int l = 0; l < 64; l++)
        for(int t = 0; t < 8; t++) {
            a8[8 * l + t] = bf(A[(l % 16) * 32 + 8 * (l / 16) + t]);
            b8[8 * l + t] = bf(B[(8 * (l / 16) + t) * 16 + l % 16]);
        }
    uint16_t *da = nullptr;
    float *dd    = nullptr;
    GF_HIP(hipMalloc((void **)&da, 2 * 64 * 8 * sizeof(uint16_t)));
    if(hipMalloc((void **)&dd, 256 * sizeof(float)) != hipSuccess) {
        (void)hipFree(da);
        return fail(GF_ERR_NOMEM, "probe buffer");
    }
    gf_status st = GF_OK;
    if(hipMemcpyAsync(da, a8.data(), 512 * 2, hipMemcpyHostToDevice, ctx->stream) != hipSuccess
       || hipMemcpyAsync(da + 512, b8.data(), 512 * 2, hipMemcpyHostToDevice, ctx->stream) != hipSuccess
       || gfm::launch_bf16_probe(da, da + 512, dd, ctx->stream)
       || hipMemcpyAsync(d.data(), dd, 256 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess
       || hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = fail(GF_ERR_HIP, "BF16 MFMA probe failed");
    (void)hipFree(da);
    (void)hipFree(dd);
    if(st)
        return st;
    for(int l = 0; l < 64; l++)
        for(int r = 0; r < 4; r++) {
            const int i = GF_BF16_D_ROW(l, r), j = l % 16;
            float want  = 0.f;
            for(int k = 0; k < 32; k++)
                want += A[i * 32 + k] * B[k * 16 + j];
            if(d[4 * l + r] != want)
                return fail(GF_ERR_HIP, "BF16 MFMA result layout differs from the compiled one (lane %d reg %d)", l, r);
        }
    return GF_OK;
}

inline int64_t round64(int64_t n) { return (n + 63) / 64 * 64; }

gf_status mlp_copy_xy(gf_mlp *m, const float *xy, int64_t n) {
    GF_HIP(hipMemcpyAsync(m->D.xy, xy, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, m->ctx->stream));
    const int64_t np = round64(n);
    if(np > n)
        GF_HIP(hipMemsetAsync(m->D.xy + n, 0, (size_t)(np - n) * 2 * sizeof(float), m->ctx->stream));
    return GF_OK;
}

}  // namespace

extern "C" {

int gf_mlp_param_count(void) { return GF_MLP_NPARAM; }

gf_status gf_mlp_create(gf_ctx *ctx, uint64_t seed, int batch_cap, gf_mlp **out) {
    if(!out)
        return fail(GF_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(batch_cap < 64 || batch_cap > (1 << 22))
        return fail(GF_ERR_INVALID, "batch_cap %d outside 64..4194304", batch_cap);
    gf_status st = mlp_layout_check(ctx);
    if(st)
        return st;
    gf_mlp *m  = new gf_mlp();
    m->ctx     = ctx;
    GfMlpDev &D = m->D;
    D.cap       = round64(batch_cap);
    const size_t cap = (size_t)D.cap, H = GF_MLP_H;
    const size_t nb = cap / 64;
    bool ok = hipMalloc((void **)&D.P, GF_MLP_NPARAM * 4) == hipSuccess
              && hipMalloc((void **)&D.m, GF_MLP_NPARAM * 4) == hipSuccess
              && hipMalloc((void **)&D.v, GF_MLP_NPARAM * 4) == hipSuccess
              && hipMalloc((void **)&D.grad, GF_MLP_NPARAM * 4) == hipSuccess
              && hipMalloc((void **)&D.w2b, H * H * 2) == hipSuccess
              && hipMalloc((void **)&D.w2t, H * H * 2) == hipSuccess
              && hipMalloc((void **)&D.xy, cap * 8) == hipSuccess
              && hipMalloc((void **)&D.z1, cap * H * 4) == hipSuccess
              && hipMalloc((void **)&D.z2, cap * H * 4) == hipSuccess
              && hipMalloc((void **)&D.h1, cap * H * 2) == hipSuccess
              && hipMalloc((void **)&D.delta2, cap * H * 2) == hipSuccess
              && hipMalloc((void **)&D.out, cap * 4) == hipSuccess
              && hipMalloc((void **)&D.coef, cap * 4) == hipSuccess
              && hipMalloc((void **)&D.loss, 64) == hipSuccess
              && hipMalloc((void **)&D.part_in, nb * GF_MLP_PART_IN * 4) == hipSuccess
              && hipMalloc((void **)&D.part_w2, (size_t)GF_MLP_W2_BLOCKS * H * H * 4) == hipSuccess
              && hipMalloc((void **)&D.part_out, nb * GF_MLP_PART_OUT * 4) == hipSuccess
              && hipMalloc((void **)&D.state, sizeof(GfTrainState)) == hipSuccess;
    if(!ok) {
        mlp_free(m);
        delete m;
        return fail(GF_ERR_NOMEM, "cannot allocate the network buffers (batch_cap %d)", batch_cap);
    }
    // seeded initial weights: first-layer ridges spread over the square, Kaiming-uniform
    // second layer, small output layer
    std::vector<float> P(GF_MLP_NPARAM, 0.f);
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<float> u1(-3.f, 3.f);
    const float b2 = std::sqrt(6.f / GF_MLP_H);
    std::uniform_real_distribution<float> u2(-b2, b2), u3(-0.05f, 0.05f);
    for(int i = GF_MLP_W1; i < GF_MLP_W2; i++)
        P[i] = u1(g);
    for(int i = GF_MLP_W2; i < GF_MLP_B2; i++)
        P[i] = u2(g);
    for(int i = GF_MLP_W3; i < GF_MLP_B3; i++)
        P[i] = u3(g);
    st = gf_mlp_set_params(m, P.data(), GF_MLP_NPARAM);
    if(st) {
        gf_mlp_destroy(m);
        return st;
    }
    *out = m;
    return GF_OK;
}

void gf_mlp_destroy(gf_mlp *m) {
    if(!m)
        return;
    if(m->ctx) {
        (void)hipSetDevice(m->ctx->device);
        (void)hipStreamSynchronize(m->ctx->stream);
    }
    mlp_free(m);
    delete m;
}

gf_status gf_mlp_set_params(gf_mlp *m, const float *in, int n) {
    if(!m || !in || n != GF_MLP_NPARAM)
        return fail(GF_ERR_INVALID, "parameter vector must have %d entries", GF_MLP_NPARAM);
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    hipStream_t s = m->ctx->stream;
    GF_HIP(hipMemcpyAsync(m->D.P, in, GF_MLP_NPARAM * 4, hipMemcpyHostToDevice, s));
    GF_HIP(hipMemsetAsync(m->D.m, 0, GF_MLP_NPARAM * 4, s));
    GF_HIP(hipMemsetAsync(m->D.v, 0, GF_MLP_NPARAM * 4, s));
    m->adam_t = 0;
    if(gfm::launch_w2_bf16(m->D, s))
        return fail(GF_ERR_HIP, "bf16 conversion launch failed");
    GF_HIP(hipStreamSynchronize(s));
    return GF_OK;
}

gf_status gf_mlp_get_params(gf_mlp *m, float *out, int n) {
    if(!m || !out || n != GF_MLP_NPARAM)
        return fail(GF_ERR_INVALID, "parameter vector must have %d entries", GF_MLP_NPARAM);
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    GF_HIP(hipMemcpyAsync(out, m->D.P, GF_MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    GF_HIP(hipStreamSynchronize(m->ctx->stream));
    return GF_OK;
}

gf_status gf_mlp_forward(gf_mlp *m, const float *xy, int64_t n, float *out) {
    if(!m || !xy || !out || n < 1)
        return fail(GF_ERR_INVALID, "null or empty arrays");
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    hipStream_t s = m->ctx->stream;
    for(int64_t off = 0; off < n; off += m->D.cap) {
        const int64_t c = std::min<int64_t>(m->D.cap, n - off);
        gf_status st    = mlp_copy_xy(m, xy + 2 * off, c);
        if(st)
            return st;
        if(gfm::launch_forward(m->D, m->D.xy, (int)round64(c), m->D.out, s))
            return fail(GF_ERR_HIP, "forward launch failed");
        GF_HIP(hipMemcpyAsync(out + off, m->D.out, (size_t)c * 4, hipMemcpyDeviceToHost, s));
    }
    GF_HIP(hipStreamSynchronize(s));
    return GF_OK;
}

gf_status gf_mlp_gradient(gf_mlp *m, const float *xy, const float *coef, int64_t n, float *grad) {
    if(!m || !xy || !coef || !grad || n < 1)
        return fail(GF_ERR_INVALID, "null or empty arrays");
    if(n > m->D.cap)
        return fail(GF_ERR_INVALID, "%lld points exceed batch_cap %lld", (long long)n, (long long)m->D.cap);
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    hipStream_t s    = m->ctx->stream;
    const int64_t np = round64(n);
    gf_status st     = mlp_copy_xy(m, xy, n);
    if(st)
        return st;
    GF_HIP(hipMemcpyAsync(m->D.coef, coef, (size_t)n * 4, hipMemcpyHostToDevice, s));
    if(np > n)
        GF_HIP(hipMemsetAsync(m->D.coef + n, 0, (size_t)(np - n) * 4, s));
    if(gfm::launch_forward(m->D, m->D.xy, (int)np, m->D.out, s)
       || gfm::launch_backward(m->D, m->D.xy, m->D.coef, (int)np, m->D.grad, s))
        return fail(GF_ERR_HIP, "forward/backward launch failed");
    GF_HIP(hipMemcpyAsync(grad, m->D.grad, GF_MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    return GF_OK;
}

gf_status gf_mlp_train(gf_mlp *m, const double *dev_m12, const double *dev_m13, int64_t ndata, const double box[4],
                       int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last) {
    if(!m || !dev_m12 || !dev_m13 || !box || ndata < 1 || steps < 1)
        return fail(GF_ERR_INVALID, "bad training arguments");
    if(batch < 128 || batch % 128 || batch > m->D.cap)
        return fail(GF_ERR_INVALID, "batch %d must be a multiple of 128 and <= batch_cap %lld", batch,
                    (long long)m->D.cap);
    if(!(box[1] > box[0]) || !(box[3] > box[2]))
        return fail(GF_ERR_INVALID, "empty box");
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    hipStream_t s = m->ctx->stream;
    const GfBox B{box[0], box[1], box[2], box[3]};
    GfTrainState st{};
    st.seed   = seed;
    st.k      = 0;
    st.steps  = steps;
    st.adam_t = m->adam_t;
    st.lr0    = lr;
    st.b1     = 0.9f;
    st.b2     = 0.999f;
    st.eps    = 1e-8f;
    GF_HIP(hipMemcpyAsync(m->D.state, &st, sizeof(st), hipMemcpyHostToDevice, s));
    // step 0 eagerly (its loss is the first loss)
    if(gfm::launch_train_step(m->D, dev_m12, dev_m13, ndata, B, batch, s))
        return fail(GF_ERR_HIP, "training step launch failed");
    if(loss_first) {
        GF_HIP(hipMemcpyAsync(loss_first, m->D.loss, 4, hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
    }
    if(steps > 1 && m->use_graph) {
        // one captured step replayed: all per-step values come from the device state
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        GF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc   = gfm::launch_train_step(m->D, dev_m12, dev_m13, ndata, B, batch, s);
        hipError_t e   = hipStreamEndCapture(s, &g);
        if(rc || e != hipSuccess) {
            if(g)
                (void)hipGraphDestroy(g);
            return fail(GF_ERR_HIP, "training step capture failed");
        }
        if(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(GF_ERR_HIP, "training graph instantiation failed");
        }
        gf_status gst = GF_OK;
        for(int k = 1; k < steps && !gst; k++)
            if(hipGraphLaunch(ge, s) != hipSuccess)
                gst = fail(GF_ERR_HIP, "training graph launch failed at step %d", k);
        (void)hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if(gst)
            return gst;
    } else {
        for(int k = 1; k < steps; k++)
            if(gfm::launch_train_step(m->D, dev_m12, dev_m13, ndata, B, batch, s))
                return fail(GF_ERR_HIP, "training step %d launch failed", k);
    }
    m->adam_t += steps;
    if(loss_last)
        GF_HIP(hipMemcpyAsync(loss_last, m->D.loss, 4, hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    return GF_OK;
}

gf_status gf_mlp_use_graph(gf_mlp *m, int on) {
    if(!m)
        return fail(GF_ERR_INVALID, "null network");
    m->use_graph = on != 0;
    return GF_OK;
}

gf_status gf_mlp_build_proposal(gf_mlp *m, const double box[4], int grid, double alpha) {
    if(!m || !box)
        return fail(GF_ERR_INVALID, "null argument");
    if(grid < 8 || grid > 4096 || grid % 8)
        return fail(GF_ERR_INVALID, "grid %d must be a multiple of 8 in 8..4096", grid);
    if(!(alpha > 0.0 && alpha <= 1.0))
        return fail(GF_ERR_INVALID, "alpha %g outside (0, 1]", alpha);
    if(!(box[1] > box[0]) || !(box[3] > box[2]))
        return fail(GF_ERR_INVALID, "empty box");
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    hipStream_t s      = m->ctx->stream;
    const int64_t nc   = (int64_t)grid * grid;
    float2 *gxy        = nullptr;
    float *gout        = nullptr;
    GF_HIP(hipMalloc((void **)&gxy, (size_t)nc * 8));
    if(hipMalloc((void **)&gout, (size_t)nc * 4) != hipSuccess) {
        (void)hipFree(gxy);
        return fail(GF_ERR_NOMEM, "proposal table");
    }
    std::vector<float> f(nc);
    gf_status st = GF_OK;
    if(gfm::launch_batch_grid(grid, gxy, s))
        st = fail(GF_ERR_HIP, "grid launch failed");
    for(int64_t off = 0; !st && off < nc; off += m->D.cap) {
        const int64_t c = std::min<int64_t>(m->D.cap, nc - off);  // multiple of 64: grid % 8 == 0
        if(gfm::launch_forward(m->D, gxy + off, (int)c, gout + off, s))
            st = fail(GF_ERR_HIP, "forward launch failed");
    }
    if(!st && (hipMemcpyAsync(f.data(), gout, (size_t)nc * 4, hipMemcpyDeviceToHost, s) != hipSuccess
               || hipStreamSynchronize(s) != hipSuccess))
        st = fail(GF_ERR_HIP, "copy of the table failed");
    (void)hipFree(gxy);
    (void)hipFree(gout);
    if(st)
        return st;
    double fmax = -INFINITY;
    for(float v : f)
        if(std::isfinite(v))
            fmax = std::max(fmax, (double)v);
    if(!std::isfinite(fmax))
        return fail(GF_ERR_INVALID, "the network output is not finite anywhere");
    std::vector<double> p(nc), cdf(nc);
    double z = 0.0;
    for(int64_t c = 0; c < nc; c++) {
        p[c] = std::isfinite(f[c]) ? std::exp((double)f[c] - fmax) : 0.0;
        z += p[c];
    }
    double acc = 0.0;
    for(int64_t c = 0; c < nc; c++) {
        p[c] = (1.0 - alpha) * p[c] / z + alpha / (double)nc;
        acc += p[c];
        cdf[c] = acc;
    }
    for(int64_t c = 0; c < nc; c++)  // exact 1 at the end; p matches the normalised table
        cdf[c] /= acc;
    for(int64_t c = 0; c < nc; c++)
        p[c] /= acc;
    cdf[nc - 1] = 1.0;
    if(m->G != grid) {
        if(m->d_cdf)
            (void)hipFree(m->d_cdf);
        if(m->d_pcell)
            (void)hipFree(m->d_pcell);
        m->d_cdf = m->d_pcell = nullptr;
        m->G                  = 0;
        GF_HIP(hipMalloc((void **)&m->d_cdf, (size_t)nc * 8));
        GF_HIP(hipMalloc((void **)&m->d_pcell, (size_t)nc * 8));
    }
    GF_HIP(hipMemcpyAsync(m->d_cdf, cdf.data(), (size_t)nc * 8, hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(m->d_pcell, p.data(), (size_t)nc * 8, hipMemcpyHostToDevice, s));
    GF_HIP(hipStreamSynchronize(s));
    m->G   = grid;
    m->box = GfBox{box[0], box[1], box[2], box[3]};
    return GF_OK;
}

gf_status gf_mlp_sample(gf_mlp *m, int64_t n, uint64_t seed, double *dev_m12, double *dev_m13, double *dev_w) {
    if(!m || !dev_m12 || !dev_m13 || !dev_w || n < 1)
        return fail(GF_ERR_INVALID, "null or empty arrays");
    if(!m->G)
        return fail(GF_ERR_STATE, "no proposal table (gf_mlp_build_proposal)");
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    hipStream_t s = m->ctx->stream;
    if(gfm::launch_sample(m->d_cdf, m->d_pcell, m->G, m->box, n, seed, dev_m12, dev_m13, dev_w, s))
        return fail(GF_ERR_HIP, "sampling launch failed");
    GF_HIP(hipStreamSynchronize(s));
    return GF_OK;
}

gf_status gf_sample_uniform(gf_ctx *ctx, const double box[4], int64_t n, uint64_t seed, double *dev_m12,
                            double *dev_m13, double *dev_w) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(!box || !dev_m12 || !dev_m13 || !dev_w || n < 1)
        return fail(GF_ERR_INVALID, "null or empty arrays");
    if(!(box[1] > box[0]) || !(box[3] > box[2]))
        return fail(GF_ERR_INVALID, "empty box");
    if(gfm::launch_sample_uniform(GfBox{box[0], box[1], box[2], box[3]}, n, seed, dev_m12, dev_m13, dev_w,
                                  ctx->stream))
        return fail(GF_ERR_HIP, "sampling launch failed");
    GF_HIP(hipStreamSynchronize(ctx->stream));
    return GF_OK;
}

gf_status gf_mlp_train_generated(gf_mlp *m, int dalitz_func, int64_t nevents, uint64_t gen_seed, const double box[4],
                                 int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last) {
    if(!m || nevents < 1 || !box)
        return fail(GF_ERR_INVALID, "need a network, a box and at least one event");
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    double *buf = nullptr;
    GF_HIP(hipMalloc((void **)&buf, (size_t)nevents * 2 * sizeof(double)));
    gf_status st = gf_dalitz_generate(m->ctx, dalitz_func, nevents, gen_seed, box, 0.0, 0, buf, buf + nevents,
                                      nullptr, nullptr);
    if(!st)
        st = gf_mlp_train(m, buf, buf + nevents, nevents, box, steps, batch, lr, seed, loss_first, loss_last);
    (void)hipStreamSynchronize(m->ctx->stream);
    (void)hipFree(buf);
    return st;
}

gf_status gf_mlp_train_host(gf_mlp *m, const double *m12, const double *m13, int64_t ndata, const double box[4],
                            int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last) {
    if(!m || !m12 || !m13 || ndata < 1)
        return fail(GF_ERR_INVALID, "null or empty event arrays");
    if(!valid_ctx(m->ctx))
        return GF_ERR_INVALID;
    double *buf = nullptr;
    GF_HIP(hipMalloc((void **)&buf, (size_t)ndata * 2 * sizeof(double)));
    gf_status st = GF_OK;
    if(hipMemcpyAsync(buf, m12, (size_t)ndata * 8, hipMemcpyHostToDevice, m->ctx->stream) != hipSuccess
       || hipMemcpyAsync(buf + ndata, m13, (size_t)ndata * 8, hipMemcpyHostToDevice, m->ctx->stream) != hipSuccess
       || hipStreamSynchronize(m->ctx->stream) != hipSuccess)
        st = fail(GF_ERR_HIP, "copy of the training events failed");
    if(!st)
        st = gf_mlp_train(m, buf, buf + ndata, ndata, box, steps, batch, lr, seed, loss_first, loss_last);
    (void)hipStreamSynchronize(m->ctx->stream);
    (void)hipFree(buf);
    return st;
}

gf_status gf_dalitz_set_norm_points(gf_ctx *ctx, int dalitz_func, gf_mlp *proposal, int64_t n, uint64_t seed,
                                    const double box[4]) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, nullptr);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(d->td)
        return fail(GF_ERR_INVALID, "point normalisation is not available for Amp3Body_TD");
    if(n < 1)
        return fail(GF_ERR_INVALID, "need at least one point");
    if(proposal && proposal->ctx != ctx)
        return fail(GF_ERR_INVALID, "the proposal belongs to another context");
    if(!proposal && !box)
        return fail(GF_ERR_INVALID, "uniform points need the box");
    auto &np = ctx->norm_points[d->cache_key];
    if(np.cap < n) {
        if(np.buf) {
            GF_HIP(hipStreamSynchronize(ctx->stream));
            GF_HIP(hipFree(np.buf));
        }
        np.buf = nullptr;
        np.cap = np.n = 0;
        if(hipMalloc((void **)&np.buf, (size_t)n * 3 * sizeof(double)) != hipSuccess)
            return fail(GF_ERR_NOMEM, "cannot allocate %lld normalisation points", (long long)n);
        np.cap = n;
    }
    np.n         = 0;
    gf_status st = proposal ? gf_mlp_sample(proposal, n, seed, np.buf, np.buf + n, np.buf + 2 * n)
                            : gf_sample_uniform(ctx, box, n, seed, np.buf, np.buf + n, np.buf + 2 * n);
    if(st)
        return st;
    np.n = n;
    return GF_OK;
}

gf_status gf_dalitz_set_norm_rule(gf_ctx *ctx, int dalitz_func, const double *m12, const double *m13,
                                  const double *w, int64_t n) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, nullptr);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(d->td)
        return fail(GF_ERR_INVALID, "point normalisation is not available for Amp3Body_TD");
    if(n < 1 || !m12 || !m13 || !w)
        return fail(GF_ERR_INVALID, "need at least one point with its weight");
    auto &np = ctx->norm_points[d->cache_key];
    if(np.cap < n) {
        if(np.buf) {
            GF_HIP(hipStreamSynchronize(ctx->stream));
            GF_HIP(hipFree(np.buf));
        }
        np.buf = nullptr;
        np.cap = np.n = 0;
        if(hipMalloc((void **)&np.buf, (size_t)n * 3 * sizeof(double)) != hipSuccess)
            return fail(GF_ERR_NOMEM, "cannot allocate %lld normalisation points", (long long)n);
        np.cap = n;
    }
    np.n = 0;
    GF_HIP(hipStreamSynchronize(ctx->stream));
    GF_HIP(hipMemcpyAsync(np.buf, m12, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(np.buf + n, m13, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipMemcpyAsync(np.buf + 2 * n, w, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(hipStreamSynchronize(ctx->stream));
    np.n = n;
    return GF_OK;
}

gf_status gf_dalitz_gram_norm_points(gf_ctx *ctx, int dalitz_func, int eff_func, int64_t offset, int64_t count,
                                     double *out_re_im) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, nullptr);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    auto it = ctx->norm_points.find(d->cache_key);
    if(it == ctx->norm_points.end() || it->second.n == 0)
        return fail(GF_ERR_STATE, "no normalisation points (gf_dalitz_set_norm_points)");
    const auto &np = it->second;
    if(offset < 0 || count < 1 || offset + count > np.n)
        return fail(GF_ERR_INVALID, "points [%lld, %lld) outside the %lld stored", (long long)offset,
                    (long long)(offset + count), (long long)np.n);
    return gf_dalitz_gram_points_device(ctx, dalitz_func, np.buf + offset, np.buf + np.n + offset,
                                        np.buf + 2 * np.n + offset, count, eff_func, out_re_im);
}

gf_status gf_dalitz_generate(gf_ctx *ctx, int dalitz_func, int64_t n, uint64_t seed, const double box[4],
                             double fmax, int host_out, double *m12, double *m13, int64_t *tried,
                             double *fmax_used) {
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    int slot      = 0;
    DalitzNode *d = find_dalitz(ctx, dalitz_func, &slot);
    if(!d)
        return fail(GF_ERR_INVALID, "function %d is not an Amp3Body node", dalitz_func);
    if(d->td)
        return fail(GF_ERR_INVALID, "event generation is not available for Amp3Body_TD");
    if(n < 1 || !m12 || !m13 || !box)
        return fail(GF_ERR_INVALID, "null output or n < 1");
    if(!(box[1] > box[0]) || !(box[3] > box[2]))
        return fail(GF_ERR_INVALID, "empty box");
    int eff_pc      = -1;
    double eff_cval = 1.0;
    gf_status st    = efficiency_mode(ctx, d->eff_func, eff_pc, eff_cval);
    if(st)
        return st;
    if((st = flush_uploads(ctx)))
        return st;
    const int64_t B  = 1 << 22;  // candidates per pass
    const int64_t nb = B / 256;
    double *c = nullptr, *o = nullptr, *bmax = nullptr;
    uint8_t *flag = nullptr;
    int *bcount = nullptr;
    int64_t *boff = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(ctx->stream);
        void *p[] = {c, o, bmax, flag, bcount, boff};
        for(void *q : p)
            if(q)
                (void)hipFree(q);
    };
    if(hipMalloc((void **)&c, (size_t)B * 2 * 8) != hipSuccess || hipMalloc((void **)&flag, (size_t)B) != hipSuccess
       || hipMalloc((void **)&bcount, (size_t)nb * 4) != hipSuccess
       || hipMalloc((void **)&bmax, (size_t)nb * 8) != hipSuccess
       || hipMalloc((void **)&boff, (size_t)nb * 8) != hipSuccess
       || (host_out && hipMalloc((void **)&o, (size_t)n * 2 * 8) != hipSuccess)) {
        release();
        return fail(GF_ERR_NOMEM, "cannot allocate the generator buffers");
    }
    double *o12 = host_out ? o : m12, *o13 = host_out ? o + n : m13;
    GfProgArgs a = prog_args(ctx);
    std::vector<int> hc(nb);
    std::vector<double> hm(nb);
    std::vector<int64_t> off(nb);
    const uint64_t key = gfd::mix64(seed ^ 0x6A09E667F3BCC909ull);
    int64_t have = 0, base = 0;
    auto pass = [&](double env) -> gf_status {
        if(gfk::launch_dalitz_candidates(a, ctx->d_res + slot * GF_MAX_RES, d->nres, d->p_coup, eff_pc, eff_cval,
                                         d->col_m12,
                                         d->col_m13, box, key, base, B, env, c, c + B, flag, bcount, bmax,
                                         ctx->stream))
            return fail(GF_ERR_HIP, "candidate launch failed");
        GF_HIP(hipMemcpyAsync(hc.data(), bcount, (size_t)nb * 4, hipMemcpyDeviceToHost, ctx->stream));
        GF_HIP(hipMemcpyAsync(hm.data(), bmax, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->stream));
        GF_HIP(hipStreamSynchronize(ctx->stream));
        return GF_OK;
    };
    if(!(fmax > 0.0)) {
        // envelope: 1.25 x the largest value over one pass of candidates
        st = pass(INFINITY);
        if(st) {
            release();
            return st;
        }
        double mx = 0.0;
        for(double v : hm)
            mx = std::max(mx, v);
        if(!(mx > 0.0)) {
            release();
            return fail(GF_ERR_INVALID, "|A|^2 * eff is zero on %lld candidates", (long long)B);
        }
        fmax = 1.25 * mx;
        base += B;
    }
    while(have < n) {
        if((st = pass(fmax))) {
            release();
            return st;
        }
        double mx    = 0.0;
        int64_t acc  = 0;
        for(int64_t b = 0; b < nb; b++) {
            off[b] = have + acc;
            acc += hc[b];
            mx = std::max(mx, hm[b]);
        }
        if(mx > fmax) {
            release();
            return fail(GF_ERR_INVALID, "envelope %g below the value %g found: raise fmax", fmax, mx);
        }
        if(acc > 0) {
            if(hipMemcpyAsync(boff, off.data(), (size_t)nb * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess
               || gfk::launch_compact(c, c + B, flag, B, boff, n, o12, o13, ctx->stream)) {
                release();
                return fail(GF_ERR_HIP, "compaction failed");
            }
        }
        have += acc;
        base += B;
        if(base > (int64_t)1 << 40) {
            release();
            return fail(GF_ERR_INVALID, "acceptance too low: %lld of %lld events after %lld candidates",
                        (long long)have, (long long)n, (long long)base);
        }
    }
    if(host_out) {
        if(hipMemcpyAsync(m12, o12, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess
           || hipMemcpyAsync(m13, o13, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) {
            release();
            return fail(GF_ERR_HIP, "copy of the events failed");
        }
    }
    release();
    if(tried)
        *tried = base;
    if(fmax_used)
        *fmax_used = fmax;
    return GF_OK;
}

// ------------------------------------------------------------ toy batches

}  // extern "C"

struct gf_batch {
    gf_ctx *ctx = nullptr;
    int ntoys = 0, np = 0, nn = 0;
    std::vector<int64_t> first, count;
    // pinned staging: params [T][np] | norms [T][nn] | active [T] (bytes, padded to 8)
    double *h_stage = nullptr, *d_stage = nullptr;
    size_t stage_bytes = 0;
    // work table and partials for the current chunk
    int chunk = 0, nblocks = 0;
    GfBatchWork *d_work = nullptr;
    int *d_first = nullptr;
    gfk::Partial *d_partials = nullptr;
    // results: mapped [2T] doubles + stamp
    double *h_res = nullptr, *d_res = nullptr;
    uint32_t *h_stamp = nullptr, *d_stamp = nullptr, *d_counter = nullptr;
    uint32_t expect = 0;
    // graph of (stage copy, sweep, per-toy finish, stamp) for one metric
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_metric = -1, graph_chunk = 0;
    bool use_graph = true;
    // several points per toy, two buffers (one filled while the other's launch runs):
    // pinned stage = table (int pairs {first slot, points} per toy, ntoys doubles)
    // then slots of np + nn doubles; device copy alike
    struct Points {
        double *h = nullptr, *d = nullptr;
        size_t cap = 0;  // slots
        int slots = 0, last = -1, max = 0;
        gfk::Partial *d_partials = nullptr;  // [nblocks][GF_BATCH_PTS]
        int partials_blocks = 0;
        double *h_res = nullptr, *d_res = nullptr;  // mapped [ntoys][GF_BATCH_PTS][2]
        bool in_flight = false;
        uint32_t ticket = 0;  // stamp value of its launch
    } pts[2];
    int pts_cur = 0;  // the buffer gf_batch_add_point fills
    // normalisation of the points on the device (gf_batch_set_norm_ops)
    int norm_nops = 0;
    double *d_norm_tmpl = nullptr;    // nn
    GfNormOp *d_norm_ops = nullptr;   // norm_nops
    double *d_norm_table = nullptr;
};
static_assert(sizeof(gf_norm_op) == sizeof(GfNormOp), "gf_norm_op is GfNormOp");
static_assert(GF_BATCH_PTS == GF_BATCH_MAX_POINTS, "point limits of the C ABI and the kernels");

namespace {

void batch_drop_graph(gf_batch *b) {
    if(b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if(b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec  = nullptr;
    b->graph = nullptr;
    b->graph_metric = -1;
}

double *batch_params(gf_batch *b) { return b->h_stage; }
double *batch_norms(gf_batch *b) { return b->h_stage + (size_t)b->ntoys * b->np; }
uint8_t *batch_active(gf_batch *b) { return (uint8_t *)(b->h_stage + (size_t)b->ntoys * (b->np + b->nn)); }

// Work table for `chunk` events per workgroup: each toy's rows split into
// ceil(count / chunk) pieces, toys in order.
gf_status batch_layout(gf_batch *b, int chunk) {
    std::vector<GfBatchWork> w;
    std::vector<int> first(b->ntoys + 1);
    for(int t = 0; t < b->ntoys; t++) {
        first[t] = (int)w.size();
        const int64_t n = b->count[t];
        for(int64_t s = 0; s < n || (s == 0 && n == 0); s += chunk) {
            GfBatchWork e;
            e.toy   = t;
            e.start = b->first[t] + s;
            e.end   = b->first[t] + std::min<int64_t>(n, s + chunk);
            e.total = (double)n;
            w.push_back(e);
            if(n == 0)
                break;
        }
    }
    first[b->ntoys] = (int)w.size();
    if(w.size() > (size_t)INT32_MAX / 2)
        return fail(GF_ERR_INVALID, "too many workgroups");
    batch_drop_graph(b);
    if(b->d_work)
        GF_HIP(hipFree(b->d_work));
    if(b->d_partials)
        GF_HIP(hipFree(b->d_partials));
    b->d_work     = nullptr;
    b->d_partials = nullptr;
    GF_HIP(hipMalloc((void **)&b->d_work, w.size() * sizeof(GfBatchWork)));
    GF_HIP(hipMalloc((void **)&b->d_partials, w.size() * sizeof(gfk::Partial)));
    GF_HIP(hipMemcpyAsync(b->d_work, w.data(), w.size() * sizeof(GfBatchWork), hipMemcpyHostToDevice, b->ctx->stream));
    GF_HIP(hipMemcpyAsync(b->d_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice,
                          b->ctx->stream));
    GF_HIP(hipStreamSynchronize(b->ctx->stream));
    b->chunk   = chunk;
    b->nblocks = (int)w.size();
    return GF_OK;
}

// The launch sequence of one batched reduction on the context stream.
gf_status batch_enqueue(gf_batch *b, int metric) {
    gf_ctx *ctx  = b->ctx;
    hipStream_t s = ctx->stream;
    GF_HIP(hipMemcpyAsync(b->d_stage, b->h_stage, b->stage_bytes, hipMemcpyHostToDevice, s));
    GfProgArgs a        = prog_args(ctx);
    const double *dpar  = b->d_stage;
    const double *dnorm = b->d_stage + (size_t)b->ntoys * b->np;
    const uint8_t *dact = (const uint8_t *)(b->d_stage + (size_t)b->ntoys * (b->np + b->nn));
    if(gfk::launch_batch(a, b->d_work, b->nblocks, dpar, b->np, dnorm, b->nn, dact, metric, b->d_partials, b->d_first,
                         b->ntoys, b->d_res, b->d_counter, b->d_stamp, s))
        return fail(GF_ERR_HIP, "batch launch failed");
    return GF_OK;
}

// Wait until the launch stamped `target` (or a later one on the stream) has finished.
gf_status batch_wait(gf_batch *b, uint32_t target) {
    const volatile uint32_t *stamp = b->h_stamp;
    auto reached                   = [&] { return (int32_t)(*stamp - target) >= 0; };
    const auto t0                  = std::chrono::steady_clock::now();
    for(unsigned it = 0;; ++it) {
        if(reached())
            return GF_OK;
        if((it & 255u) == 255u
           && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(b->ctx->tune[GF_TUNE_SPIN_US]))
            break;
        __builtin_ia32_pause();
    }
    GF_HIP(hipStreamSynchronize(b->ctx->stream));
    if(!reached())
        return fail(GF_ERR_HIP, "batched reduction finished without its stamp");
    return GF_OK;
}
gf_status batch_wait(gf_batch *b) { return batch_wait(b, b->expect); }

gf_status batch_run(gf_batch *b, int metric) {
    if(!b->use_graph) {
        gf_status st = batch_enqueue(b, metric);
        if(st)
            return st;
    } else {
        if(!b->exec || b->graph_metric != metric || b->graph_chunk != b->chunk) {
            batch_drop_graph(b);
            GF_HIP(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
            gf_status st = batch_enqueue(b, metric);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
            if(st) {
                if(g)
                    (void)hipGraphDestroy(g);
                return st;
            }
            if(e != hipSuccess)
                return fail(GF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            b->graph = g;
            GF_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            b->graph_metric = metric;
            b->graph_chunk  = b->chunk;
        }
        GF_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    }
    b->expect++;
    return batch_wait(b);
}

}  // namespace

extern "C" {

gf_status gf_batch_create(gf_ctx *ctx, int ntoys, const int64_t *first, const int64_t *count, gf_batch **out) {
    if(!out)
        return fail(GF_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if(!valid_ctx(ctx))
        return GF_ERR_INVALID;
    if(ntoys < 1 || !first || !count)
        return fail(GF_ERR_INVALID, "need at least one toy");
    if(ctx->ins.empty())
        return fail(GF_ERR_STATE, "no program installed (gf_program_set)");
    for(int t = 0; t < ntoys; t++)
        if(first[t] < 0 || count[t] < 0 || first[t] + count[t] > ctx->n)
            return fail(GF_ERR_INVALID, "toy %d rows [%lld, %lld) outside the %lld events", t, (long long)first[t],
                        (long long)(first[t] + count[t]), (long long)ctx->n);
    gf_batch *b = new gf_batch();
    b->ctx      = ctx;
    b->ntoys    = ntoys;
    b->np       = (int)ctx->params.size();
    b->nn       = (int)ctx->norms.size();
    b->first.assign(first, first + ntoys);
    b->count.assign(count, count + ntoys);
    const size_t nd = (size_t)ntoys * (b->np + b->nn) + ((size_t)ntoys + 7) / 8;
    b->stage_bytes  = nd * sizeof(double);
    gf_status st    = GF_OK;
    if(hipHostMalloc((void **)&b->h_stage, b->stage_bytes, hipHostMallocDefault) != hipSuccess
       || hipMalloc((void **)&b->d_stage, b->stage_bytes) != hipSuccess
       || hipMalloc((void **)&b->d_first, (size_t)(ntoys + 1) * sizeof(int)) != hipSuccess
       || hipHostMalloc((void **)&b->h_res, (size_t)2 * ntoys * sizeof(double),
                        hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess
       || hipHostGetDevicePointer((void **)&b->d_res, b->h_res, 0) != hipSuccess
       || hipHostMalloc((void **)&b->h_stamp, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess
       || hipHostGetDevicePointer((void **)&b->d_stamp, b->h_stamp, 0) != hipSuccess
       || hipMalloc((void **)&b->d_counter, 64) != hipSuccess
       || hipMemsetAsync(b->d_counter, 0, 64, ctx->stream) != hipSuccess
       || hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = fail(GF_ERR_NOMEM, "cannot allocate the batch buffers");
    if(!st) {
        *b->h_stamp = 0;
        memset(b->h_stage, 0, b->stage_bytes);
        for(int t = 0; t < ntoys; t++) {  // start from the context's arrays
            memcpy(batch_params(b) + (size_t)t * b->np, ctx->params.data(), b->np * sizeof(double));
            memcpy(batch_norms(b) + (size_t)t * b->nn, ctx->norms.data(), b->nn * sizeof(double));
        }
        st = batch_layout(b, 8192);
    }
    if(st) {
        gf_batch_destroy(b);
        return st;
    }
    *out = b;
    return GF_OK;
}

void gf_batch_destroy(gf_batch *b) {
    if(!b)
        return;
    if(b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    batch_drop_graph(b);
    void *dev[] = {b->d_stage,    b->d_first,         b->d_work,     b->d_partials,
                   b->d_counter,  b->pts[0].d,        b->pts[1].d,   b->pts[0].d_partials,
                   b->pts[1].d_partials, b->d_norm_tmpl, b->d_norm_ops, b->d_norm_table};
    for(void *p : dev)
        if(p)
            (void)hipFree(p);
    void *host[] = {b->h_stage, b->h_res, b->h_stamp, b->pts[0].h, b->pts[0].h_res, b->pts[1].h, b->pts[1].h_res};
    for(void *p : host)
        if(p)
            (void)hipHostFree(p);
    delete b;
}

gf_status gf_batch_set_toy(gf_batch *b, int toy, const double *params, const double *norms) {
    if(!b || toy < 0 || toy >= b->ntoys)
        return fail(GF_ERR_INVALID, "toy index out of range");
    if(params)
        memcpy(batch_params(b) + (size_t)toy * b->np, params, b->np * sizeof(double));
    if(norms)
        memcpy(batch_norms(b) + (size_t)toy * b->nn, norms, b->nn * sizeof(double));
    return GF_OK;
}

gf_status gf_batch_reduce(gf_batch *b, const uint8_t *active, int metric, double *out) {
    if(!b || !active || !out)
        return fail(GF_ERR_INVALID, "null argument");
    if(!valid_ctx(b->ctx))
        return GF_ERR_INVALID;
    gf_status st = check_events_for_program(b->ctx, metric);
    if(st)
        return st;
    if(b->np != (int)b->ctx->params.size() || b->nn != (int)b->ctx->norms.size())
        return fail(GF_ERR_STATE, "the context's program changed since the batch was created");
    if((st = flush_uploads(b->ctx)))  // device copies used by the wave cache and constants
        return st;
    memcpy(batch_active(b), active, b->ntoys);
    if((st = batch_run(b, metric)))
        return st;
    for(int t = 0; t < b->ntoys; t++)
        if(active[t])
            out[t] = ((const volatile double *)b->h_res)[2 * t];
    return GF_OK;
}

namespace {

using Points = gf_batch::Points;

int *batch_table(Points &q) { return (int *)q.h; }

// Room for `slots` points in the pinned stage of q and its device copy (contents kept).
gf_status batch_points_reserve(gf_batch *b, Points &q, size_t slots) {
    if(slots <= q.cap && q.h)
        return GF_OK;
    size_t cap = std::max<size_t>(q.cap ? q.cap : (size_t)std::max(b->ntoys, 256), 1);
    while(cap < slots)
        cap *= 2;
    const size_t W     = (size_t)(b->np + b->nn);
    const size_t bytes = ((size_t)b->ntoys + cap * W) * sizeof(double);
    double *h = nullptr, *d = nullptr;
    if(hipHostMalloc((void **)&h, bytes, hipHostMallocDefault) != hipSuccess)
        return fail(GF_ERR_NOMEM, "cannot allocate a stage of %zu points", cap);
    if(hipMalloc((void **)&d, bytes) != hipSuccess) {
        (void)hipHostFree(h);
        return fail(GF_ERR_NOMEM, "cannot allocate a device stage of %zu points", cap);
    }
    if(q.h) {
        memcpy(h, q.h, ((size_t)b->ntoys + (size_t)q.slots * W) * sizeof(double));
        GF_HIP(hipStreamSynchronize(b->ctx->stream));  // no copy from the old stage in flight
        (void)hipHostFree(q.h);
        (void)hipFree(q.d);
    } else {
        memset(h, 0, (size_t)b->ntoys * sizeof(double));
    }
    q.h   = h;
    q.d   = d;
    q.cap = cap;
    return GF_OK;
}

void batch_points_clear(gf_batch *b, Points &q) {
    if(q.h)
        memset(q.h, 0, (size_t)b->ntoys * sizeof(double));
    q.slots = 0;
    q.last  = -1;
    q.max   = 0;
}

}  // namespace

gf_status gf_batch_add_point(gf_batch *b, int toy, const double *params, const double *norms) {
    if(!b || toy < 0 || toy >= b->ntoys)
        return fail(GF_ERR_INVALID, "toy index out of range");
    if(!params || (!norms && b->norm_nops == 0))
        return fail(GF_ERR_INVALID, "null parameter or normalisation array");
    Points &q = b->pts[b->pts_cur];
    if(q.in_flight)
        return fail(GF_ERR_STATE, "two submissions in flight: wait for the earlier one first");
    gf_status st = batch_points_reserve(b, q, (size_t)q.slots + 1);
    if(st)
        return st;
    int *tbl = batch_table(q);
    if(tbl[2 * toy + 1] > 0 && q.last != toy)
        return fail(GF_ERR_INVALID, "the points of toy %d must be added one after the other", toy);
    if(tbl[2 * toy + 1] >= GF_BATCH_PTS)
        return fail(GF_ERR_INVALID, "toy %d already has %d points", toy, GF_BATCH_PTS);
    if(tbl[2 * toy + 1] == 0)
        tbl[2 * toy] = q.slots;
    const size_t W = (size_t)(b->np + b->nn);
    double *slot   = q.h + b->ntoys + (size_t)q.slots * W;
    memcpy(slot, params, (size_t)b->np * sizeof(double));
    if(b->norm_nops == 0)  // otherwise the launch computes them
        memcpy(slot + b->np, norms, (size_t)b->nn * sizeof(double));
    tbl[2 * toy + 1]++;
    q.slots++;
    q.last = toy;
    q.max  = std::max(q.max, tbl[2 * toy + 1]);
    return GF_OK;
}

gf_status gf_batch_clear_points(gf_batch *b) {
    if(!b)
        return fail(GF_ERR_INVALID, "null batch");
    if(!b->pts[b->pts_cur].in_flight)  // a buffer in flight holds no unsubmitted points
        batch_points_clear(b, b->pts[b->pts_cur]);
    return GF_OK;
}

gf_status gf_batch_submit_points(gf_batch *b, int metric, uint32_t *ticket) {
    if(!b || !ticket)
        return fail(GF_ERR_INVALID, "null argument");
    if(!valid_ctx(b->ctx))
        return GF_ERR_INVALID;
    gf_ctx *ctx  = b->ctx;
    gf_status st = check_events_for_program(ctx, metric);
    if(st)
        return st;
    if(b->np != (int)ctx->params.size() || b->nn != (int)ctx->norms.size())
        return fail(GF_ERR_STATE, "the context's program changed since the batch was created");
    Points &q = b->pts[b->pts_cur];
    if(q.in_flight)
        return fail(GF_ERR_STATE, "two submissions in flight: wait for the earlier one first");
    if(q.slots == 0)
        return fail(GF_ERR_STATE, "no points added since the last submission");
    if((st = flush_uploads(ctx)))
        return st;
    if(q.partials_blocks != b->nblocks) {
        GF_HIP(hipStreamSynchronize(ctx->stream));
        if(q.d_partials)
            GF_HIP(hipFree(q.d_partials));
        q.d_partials      = nullptr;
        q.partials_blocks = 0;
        if(hipMalloc((void **)&q.d_partials, (size_t)b->nblocks * GF_BATCH_PTS * sizeof(gfk::Partial)) != hipSuccess) {
            q.d_partials = nullptr;
            return fail(GF_ERR_NOMEM, "cannot allocate the partials of %d workgroups", b->nblocks);
        }
        q.partials_blocks = b->nblocks;
    }
    if(!q.h_res) {
        if(hipHostMalloc((void **)&q.h_res, (size_t)2 * b->ntoys * GF_BATCH_PTS * sizeof(double),
                         hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess
           || hipHostGetDevicePointer((void **)&q.d_res, q.h_res, 0) != hipSuccess) {
            if(q.h_res)
                (void)hipHostFree(q.h_res);
            q.h_res = q.d_res = nullptr;
            return fail(GF_ERR_NOMEM, "cannot allocate the batch results");
        }
    }
    const size_t W  = (size_t)(b->np + b->nn);
    const bool fast = use_fast_dalitz(ctx, metric);
    const bool ce   = fast && const_efficiency(ctx);
    bool zo         = ce && zero_outside(ctx);
    for(int k = 0; k < q.slots && zo; k++)  // every point's efficiency constant finite
        zo = std::isfinite(q.h[b->ntoys + (size_t)k * W + ctx->ins[1].p + ctx->ins[1].a]);
    GF_HIP(hipMemcpyAsync(q.d, q.h, ((size_t)b->ntoys + (size_t)q.slots * W) * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    if(b->norm_nops > 0
       && gfk::launch_batch_norms(q.d + b->ntoys, q.slots, b->np, b->nn, b->d_norm_tmpl, b->d_norm_ops, b->norm_nops,
                                  b->d_norm_table, ctx->stream))
        return fail(GF_ERR_HIP, "normalisation launch failed");
    const int groups = (q.max + GF_MULTI_PTS - 1) / GF_MULTI_PTS;
    GfProgArgs a     = prog_args(ctx);
    const int nres   = fast ? ctx->dalitz[0].nres : 0;
    // groups of at least mx_min points take the coherent sums on the matrix cores (0: none).
    // Automatic: measured on 1024-2048 Dalitz toys, groups of 4 or more gain at 12 and 16
    // resonances (toy batch 9 % and 13 % faster) and lose at 8, where the vector kernel
    // does half the work per point and keeps four waves per SIMD.
    int mx_min = 0;
    if(fast && ce && zo && metric == GF_METRIC_NLL && ctx->mfma_layout == 2) {
        const int knob = ctx->tune[GF_TUNE_BATCH_MX_PTS];
        mx_min         = knob ? knob : (nres > 8 ? 4 : 0);
    }
    if(gfk::launch_batch_points(a, nres, b->d_work, b->nblocks, groups, q.d + b->ntoys, b->np, b->nn,
                                (const int *)q.d, fast, ce, zo, metric, mx_min, q.d_partials, b->d_first, b->ntoys,
                                q.d_res, b->d_counter, b->d_stamp, ctx->stream))
        return fail(GF_ERR_HIP, "batch launch failed");
    q.ticket    = ++b->expect;
    q.in_flight = true;
    *ticket     = q.ticket;
    b->pts_cur ^= 1;
    return GF_OK;
}

gf_status gf_batch_wait_points(gf_batch *b, uint32_t ticket, double *out) {
    if(!b || !out)
        return fail(GF_ERR_INVALID, "null argument");
    if(!valid_ctx(b->ctx))
        return GF_ERR_INVALID;
    Points *q = nullptr;
    for(Points &c : b->pts)
        if(c.in_flight && c.ticket == ticket)
            q = &c;
    if(!q)
        return fail(GF_ERR_INVALID, "no submission with ticket %u in flight", ticket);
    gf_status st = batch_wait(b, ticket);
    if(st)
        return st;
    // the stage and its copy are no longer in use either: the launch has finished
    const int *tbl             = batch_table(*q);
    const volatile double *res = q->h_res;
    for(int t = 0; t < b->ntoys; t++)
        for(int k = 0; k < tbl[2 * t + 1]; k++)
            out[(size_t)t * GF_BATCH_PTS + k] = res[2 * ((size_t)t * GF_BATCH_PTS + k)];
    q->in_flight = false;
    batch_points_clear(b, *q);
    return GF_OK;
}

gf_status gf_batch_reduce_points(gf_batch *b, int metric, double *out) {
    if(!b || !out)
        return fail(GF_ERR_INVALID, "null argument");
    if(b->pts[b->pts_cur].slots == 0)
        return GF_OK;
    uint32_t ticket = 0;
    gf_status st    = gf_batch_submit_points(b, metric, &ticket);
    if(st)
        return st;
    return gf_batch_wait_points(b, ticket, out);
}

gf_status gf_batch_set_norm_ops(gf_batch *b, const gf_norm_op *ops, int nops, const double *norms_template,
                                const double *table, int nt) {
    if(!b || nops < 0 || (nops > 0 && (!ops || !norms_template)) || nt < 0 || (nt > 0 && !table))
        return fail(GF_ERR_INVALID, "bad normalisation ops");
    if(!valid_ctx(b->ctx))
        return GF_ERR_INVALID;
    for(const Points &q : b->pts)
        if(q.in_flight || q.slots > 0)
            return fail(GF_ERR_STATE, "normalisation ops change while points are staged or in flight");
    for(int o = 0; o < nops; o++) {
        const gf_norm_op &op = ops[o];
        bool ok = op.out >= 0 && op.out < b->nn && op.p >= 0 && op.p < b->np;
        if(op.kind == GF_NORM_DALITZ)
            ok = ok && op.n >= 1 && op.n <= GF_MAX_RES && op.p + 2 * op.n <= b->np && op.g >= 0
                 && (int64_t)op.g + 2 * (int64_t)op.n * op.n <= nt;
        else if(op.kind != GF_NORM_SCALE && op.kind != GF_NORM_EXP)
            ok = false;
        if(!ok)
            return fail(GF_ERR_INVALID, "normalisation op %d (kind %d, out %d, p %d, n %d, g %d) outside the arrays", o,
                        op.kind, op.out, op.p, op.n, op.g);
    }
    GF_HIP(hipStreamSynchronize(b->ctx->stream));
    void *old[] = {b->d_norm_tmpl, b->d_norm_ops, b->d_norm_table};
    for(void *p : old)
        if(p)
            GF_HIP(hipFree(p));
    b->d_norm_tmpl  = nullptr;
    b->d_norm_ops   = nullptr;
    b->d_norm_table = nullptr;
    b->norm_nops    = 0;
    if(nops == 0)
        return GF_OK;
    GF_HIP(hipMalloc((void **)&b->d_norm_tmpl, (size_t)b->nn * sizeof(double)));
    GF_HIP(hipMalloc((void **)&b->d_norm_ops, (size_t)nops * sizeof(GfNormOp)));
    GF_HIP(hipMalloc((void **)&b->d_norm_table, (size_t)std::max(nt, 1) * sizeof(double)));
    hipStream_t s = b->ctx->stream;
    GF_HIP(hipMemcpyAsync(b->d_norm_tmpl, norms_template, (size_t)b->nn * sizeof(double), hipMemcpyHostToDevice, s));
    GF_HIP(hipMemcpyAsync(b->d_norm_ops, ops, (size_t)nops * sizeof(GfNormOp), hipMemcpyHostToDevice, s));
    if(nt > 0)
        GF_HIP(hipMemcpyAsync(b->d_norm_table, table, (size_t)nt * sizeof(double), hipMemcpyHostToDevice, s));
    GF_HIP(hipStreamSynchronize(s));
    b->norm_nops = nops;
    return GF_OK;
}

gf_status gf_batch_set_chunk(gf_batch *b, int chunk) {
    if(!b || chunk < 256 || chunk % 256)
        return fail(GF_ERR_INVALID, "chunk must be a positive multiple of 256");
    if(!valid_ctx(b->ctx))
        return GF_ERR_INVALID;
    GF_HIP(hipStreamSynchronize(b->ctx->stream));
    return batch_layout(b, chunk);
}

gf_status gf_batch_get_chunk(gf_batch *b, int *chunk) {
    if(!b || !chunk)
        return fail(GF_ERR_INVALID, "null argument");
    *chunk = b->chunk;
    return GF_OK;
}

gf_status gf_batch_use_graph(gf_batch *b, int on) {
    if(!b)
        return fail(GF_ERR_INVALID, "null batch");
    b->use_graph = on != 0;
    return GF_OK;
}

gf_status gf_batch_autotune(gf_batch *b, int metric, int *chunk_out, double *us_out) {
    if(!b)
        return fail(GF_ERR_INVALID, "null batch");
    if(!valid_ctx(b->ctx))
        return GF_ERR_INVALID;
    std::vector<uint8_t> all(b->ntoys, 1);
    std::vector<double> tmp(b->ntoys);
    const int cand[] = {2048, 4096, 8192, 16384};
    int best = 0;
    double best_us = INFINITY;
    for(int c : cand) {
        gf_status st = gf_batch_set_chunk(b, c);
        if(st)
            return st;
        if((st = gf_batch_reduce(b, all.data(), metric, tmp.data())))  // warm-up (and graph build)
            return st;
        const auto t0 = std::chrono::steady_clock::now();
        for(int r = 0; r < 2; r++)
            if((st = gf_batch_reduce(b, all.data(), metric, tmp.data())))
                return st;
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / 2;
        if(us < best_us) {
            best_us = us;
            best    = c;
        }
    }
    gf_status st = gf_batch_set_chunk(b, best);
    if(st)
        return st;
    if(chunk_out)
        *chunk_out = best;
    if(us_out)
        *us_out = best_us;
    return GF_OK;
}

}  // extern "C"

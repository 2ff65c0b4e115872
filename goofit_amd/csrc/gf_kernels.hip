// gf_kernels.hip -- gfx950 kernels of the likelihood path.
//
//  * k_reduce_metric / k_eval_metric: the fused per-event evaluation of a
//    compiled PDF tree over SoA event columns, the metric (NLL, Prob, ...),
//    and a compensated FP64 sum: Neumaier per thread, wave shuffles, LDS
//    across waves, one partial per block and a last-block pass that adds the
//    partials in a fixed order (deterministic for a given grid).  Replaces the
//    thrust::transform_reduce / transform over MetricTaker
//    (src/PDFs/GooPdf.cu:45-78, 112-173; src/PDFs/MetricTaker.cu:17-44).
//  * k_reduce_bins: the midpoint-grid integral of one function
//    (src/PDFs/GooPdf.cu:83-108, src/PDFs/MetricTaker.cu:48-96).
//  * k_dalitz_cache: the resonance lineshape cache W_i[e]
//    (src/PDFs/physics/Amp3Body.cu:292-310, detail/SpecialResonanceCalculator.cu:12-46).
//  * k_gram: the per-amplitude normalisation integrals as one FP64-MFMA Gram
//    A^H diag(eff*w) A over grid cells, phase-space points or the data cache
//    (src/PDFs/physics/Amp3Body.cu:312-330 and 376-428,
//    detail/SpecialResonanceIntegrator.cu:7-121).
#include <hip/hip_runtime.h>

#include "gf_device.h"
#include "gf_internal.h"

namespace {

using gfk::Partial;

// ------------------------------------------------------------ event sources

// Events read from the SoA columns in HBM.
template <int EPT>
struct ColSrc {
    const double *const *cols;
    int64_t idx[EPT];
    __device__ __forceinline__ double get(int a, int j) const { return gfd::as_const(cols)[a][idx[j]]; }
};

// Synthetic events held in registers (grid points); slot chosen by a
// wave-uniform switch so no register array is indexed dynamically.
template <int EPT>
struct RegSrc {
    double v[EPT][GF_MAX_OBS];
    __device__ __forceinline__ double get(int a, int j) const {
        switch(a) {
        case 0: return v[j][0];
        case 1: return v[j][1];
        case 2: return v[j][2];
        case 3: return v[j][3];
        case 4: return v[j][4];
        case 5: return v[j][5];
        case 6: return v[j][6];
        default: return v[j][7];
        }
    }
    __device__ __forceinline__ void set(int a, int j, double x) {
        switch(a) {
        case 0: v[j][0] = x; break;
        case 1: v[j][1] = x; break;
        case 2: v[j][2] = x; break;
        case 3: v[j][3] = x; break;
        case 4: v[j][4] = x; break;
        case 5: v[j][5] = x; break;
        case 6: v[j][6] = x; break;
        default: v[j][7] = x; break;
        }
    }
};

// Per-level accumulators of the combining nodes.
template <int EPT>
struct Acc {
    double a0[EPT], a1[EPT], a2[EPT], a3[EPT];
};
template <int L, int EPT>
__device__ __forceinline__ double (&level(Acc<EPT> &a))[EPT] {
    if constexpr(L == 0)
        return a.a0;
    else if constexpr(L == 1)
        return a.a1;
    else if constexpr(L == 2)
        return a.a2;
    else
        return a.a3;
}

// ------------------------------------------------------------ interpreter

// One combining step on the accumulator of level L.  FEAT: the node kinds compiled
// in -- 0 none of the Dalitz nodes, 1 the time-integrated Dalitz node, 2 also the
// time-dependent one -- so that each kernel carries only the register budget of
// the programs it runs (a Dalitz node met by a kernel without it sets bit 1 of
// the error word instead of being skipped silently).
template <int L, int EPT, int FEAT, class Src>
__device__ __forceinline__ void step(int kind, const GfIns &in, const GfProgArgs &A, const GF_CONST double *__restrict__ P,
                                     const GF_CONST double *__restrict__ Nm, const Src &src, const int64_t (&row)[EPT],
                                     const bool (&valid)[EPT], Acc<EPT> &acc, double (&T)[EPT]) {
    double(&a)[EPT] = level<L>(acc);
    switch(kind) {
    case GK_PUSH0:
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            a[j] = 0.0;
        break;
    case GK_PUSH1:
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            a[j] = 1.0;
        break;
    case GK_ADD_ACC: {
        // device_AddPdfs / device_AddPdfsExt: ret += weight * curr * norm
        const double w = P[in.p], nm = Nm[in.n];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            a[j] += w * T[j] * nm;
    } break;
    case GK_ADD_LAST: {
        // AddPdf.cu:43-48: ret += (1 - totalWeight) * last * normFactor
        double tw = 0;
        for(int i = 0; i < in.k; ++i)
            tw += P[in.p + i];
        const double nm = Nm[in.n];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            a[j] += (1 - tw) * T[j] * nm;
    } break;
    case GK_ADDEXT_END: {
        // AddPdf.cu:76: ret /= totalWeight
        double tw = 0;
        for(int i = 0; i < in.k; ++i)
            tw += P[in.p + i];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = a[j] / tw;
    } break;
    case GK_PROD_ACC: {
        // ProdPdf.cu:19-23: curr *= norm; ret *= curr
        const double nm = Nm[in.n];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            double cur = T[j] * nm;
            a[j] *= cur;
        }
    } break;
    case GK_POP:
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = a[j];
        break;
    case GK_DALITZ: {
        if constexpr(FEAT < 1) {
            atomicOr(A.err, 2);
            break;
        }
        // device_DalitzPlot, src/PDFs/physics/Amp3Body.cu:60-122
        const GfDalitzSlot dz = gfd::as_const(A.dz)[in.x];
        const GfDecay &D      = A.decay;
        int64_t en[EPT];
        bool inside[EPT];
        double re[EPT], im[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            const double m12 = src.get(dz.col_m12, j), m13 = src.get(dz.col_m13, j);
            inside[j] = gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
            int64_t e = A.identity_num ? row[j] : (int64_t)floor(0.5 + src.get(dz.col_num, j));
            if(e < 0 || e >= dz.ncache) {
                if(inside[j] && valid[j])
                    atomicOr(A.err, 1);
                e = 0;
            }
            en[j] = e;
            re[j] = 0.0;
            im[j] = 0.0;
        }
        // Loads of GF_DZ_CHUNK resonances are issued back to back before any of them is
        // consumed, so each wave keeps 8 * EPT 16-byte loads in flight; the terms are
        // added in gfd::coh_sum's order (groups of four resonances), as every other
        // coherent-sum kernel adds them.
        constexpr int GF_DZ_CHUNK = 8;
        for(int i0 = 0; i0 < in.a; i0 += GF_DZ_CHUNK) {
            double2 w[GF_DZ_CHUNK][EPT];
#pragma unroll
            for(int u = 0; u < GF_DZ_CHUNK; ++u) {
                const bool use              = i0 + u < in.a;
                const double2 *__restrict__ b = dz.wave + (int64_t)(use ? i0 + u : 0) * dz.ncache;
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    w[u][j] = use ? b[en[j]] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                double2 wj[GF_DZ_CHUNK];
#pragma unroll
                for(int u = 0; u < GF_DZ_CHUNK; ++u)
                    wj[u] = w[u][j];
                gfd::coh_sum<GF_DZ_CHUNK>(
                    re[j], im[j],
                    [&](int u) { return make_double2(P[in.p + 2 * (i0 + u)], P[in.p + 2 * (i0 + u) + 1]); }, wj,
                    in.a - i0);
            }
        }
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            a[j] = inside[j] ? re[j] * re[j] + im[j] * im[j] : -1.0;
    } break;
    case GK_DALITZ_END:
        // Amp3Body.cu:118-121: ret *= eff (0 for events outside the Dalitz region)
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = (a[j] == -1.0) ? 0.0 : a[j] * T[j];
        break;
    case GK_TDDP:
        if constexpr(FEAT >= 2) {
            // device_Tddp, src/PDFs/physics/Amp3Body_TD.cu:41-257.  in.p: tau (then x, y,
            // dx, dy and the couplings); in.a: resonance count; in.x: node slot.
            const GfDalitzSlot dz = gfd::as_const(A.dz)[in.x];
            const GfDecay &D      = A.decay;
            int64_t en[EPT];
            double sar[EPT], sai[EPT], sbr[EPT], sbi[EPT];
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                const double m12 = src.get(dz.col_m12, j), m13 = src.get(dz.col_m13, j);
                const bool inside = gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
                int64_t e = A.identity_num ? row[j] : (int64_t)floor(0.5 + src.get(dz.col_num, j));
                if(e < 0 || e >= dz.ncache) {
                    if(inside && valid[j])
                        atomicOr(A.err, 1);
                    e = 0;
                }
                en[j]  = e;
                sar[j] = sai[j] = sbr[j] = sbi[j] = 0.0;
            }
            // A += a_i * c_i, B += b_i * c_i (Amp3Body_TD.cu:91-110).  The A and B waves of
            // GF_TD_CHUNK resonances are loaded before any is consumed (2 x 4 x EPT 16-byte
            // loads in flight per lane); the sums still run over i = 0, 1, ... in order.
            constexpr int GF_TD_CHUNK = 4;
            for(int i0 = 0; i0 < in.a; i0 += GF_TD_CHUNK) {
                double2 wa[GF_TD_CHUNK][EPT], wb[GF_TD_CHUNK][EPT];
#pragma unroll
                for(int u = 0; u < GF_TD_CHUNK; ++u) {
                    const bool use = i0 + u < in.a;
                    const int64_t off = (int64_t)(use ? i0 + u : 0) * dz.ncache;
#pragma unroll
                    for(int j = 0; j < EPT; ++j) {
                        wa[u][j] = use ? dz.wave[off + en[j]] : make_double2(0.0, 0.0);
                        wb[u][j] = use ? dz.wave_b[off + en[j]] : make_double2(0.0, 0.0);
                    }
                }
#pragma unroll
                for(int u = 0; u < GF_TD_CHUNK; ++u) {
                    if(i0 + u < in.a) {
                        const double cr = P[in.p + 5 + 2 * (i0 + u)], ci = P[in.p + 6 + 2 * (i0 + u)];
#pragma unroll
                        for(int j = 0; j < EPT; ++j) {
                            gfd::cmac(sar[j], sai[j], cr, ci, wa[u][j]);
                            gfd::cmac(sbr[j], sbi[j], cr, ci, wb[u][j]);
                        }
                    }
                }
            }
            const double tau = P[in.p], x0 = P[in.p + 1], y0 = P[in.p + 2], dx = P[in.p + 3], dy = P[in.p + 4];
            const GF_CONST double *__restrict__ R = P + dz.res_p;
            // mistag terms: A <-> B with fixed lifetime and mixing (Amp3Body_TD.cu:226-248)
            const double xfix = 0.0039, yfix = 0.0065, taufix = 0.4101;
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                const int tag = dz.col_tag >= 0 ? (int)src.get(dz.col_tag, j) : 0;
                const double x = tag == 1 ? x0 + dx : (tag == -1 ? x0 - dx : x0);
                const double y = tag == 1 ? y0 + dy : (tag == -1 ? y0 - dy : y0);
                const double t = src.get(dz.col_time, j), sg = src.get(dz.col_sigma, j);
                const double nA = sar[j] * sar[j] + sai[j] * sai[j];
                const double nB = sbr[j] * sbr[j] + sbi[j] * sbi[j];
                const double t1 = nA + nB, t2 = nA - nB;
                const double abr = sar[j] * sbr[j] + sai[j] * sbi[j];  // A conj(B)
                const double abi = sai[j] * sbr[j] - sar[j] * sbi[j];
                double ret = gfd::time_resolution(dz.res_kind, R, t1, t2, abr, abi, tau, t, x, y, sg);
                const double mis = dz.col_mis >= 0 ? src.get(dz.col_mis, j) : 0.0;
                if(mis > 0) {
                    ret *= (1 - 2 * mis);
                    if(dz.res_kind == GT_THREEGAUSS) {
                        // both mistag calls take the same (taufix, t, xfix, yfix, sigma): their
                        // Gaussian terms are evaluated once and combined twice (same values)
                        const gfd::TimeTerms Pf = gfd::threegauss_terms(R, taufix, t, xfix, yfix, sg);
                        ret += mis * gfd::threegauss_rate(Pf, t1, -t2, abr, -abi);
                        ret += mis * gfd::threegauss_rate(Pf, t1, t2, abr, abi);
                    } else {
                        ret += mis * gfd::truth_resolution(t1, -t2, abr, -abi, taufix, t, xfix, yfix);
                        ret += mis * gfd::truth_resolution(t1, t2, abr, abi, taufix, t, xfix, yfix);
                    }
                }
                a[j] = ret;
            }
        } else {
            atomicOr(A.err, 2);
        }
        break;
    case GK_TDDP_END:
        if constexpr(FEAT >= 2) {
            // Amp3Body_TD.cu:61-76, 253-255: ret *= eff inside, 0 outside
            const GfDalitzSlot dz = gfd::as_const(A.dz)[in.x];
            const GfDecay &D      = A.decay;
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                const bool inside = gfd::in_dalitz(src.get(dz.col_m12, j), src.get(dz.col_m13, j), D.motherMass,
                                                   D.daug1Mass, D.daug2Mass, D.daug3Mass);
                T[j] = inside ? a[j] * T[j] : 0.0;
            }
        } else {
            atomicOr(A.err, 2);
        }
        break;
    default: break;
    }
}


// Run the compiled program from instruction pc0 for EPT events per lane; the
// value of the tree ends up in T.  row[] are the event rows (used as wave-cache
// index when the event-number column is the identity), valid[] masks tail lanes.
template <int EPT, int FEAT = 1, class Src>
__device__ __forceinline__ void run(const GfProgArgs &A, const double *__restrict__ P_, const double *__restrict__ Nm_,
                                    int pc0, const Src &src, const int64_t (&row)[EPT], const bool (&valid)[EPT],
                                    double (&T)[EPT]) {
    const GF_CONST double *__restrict__ P  = gfd::as_const(P_);
    const GF_CONST double *__restrict__ Nm = gfd::as_const(Nm_);
    const GF_CONST double *__restrict__ Cn = gfd::as_const(A.consts);
    const GF_CONST GfIns *__restrict__ prog = gfd::as_const(A.ins);
    Acc<EPT> acc;
#pragma unroll
    for(int j = 0; j < EPT; ++j) {
        acc.a0[j] = acc.a1[j] = acc.a2[j] = acc.a3[j] = 0.0;
        T[j]                                          = 0.0;
    }
    for(int pc = pc0;; ++pc) {
        const GfIns in = prog[pc];
        switch(in.op) {
        case GI_END: return;
        case GI_GAUSS: {
            const double m = P[in.p], s = P[in.p + 1];
#pragma unroll
            for(int j = 0; j < EPT; ++j)
                T[j] = gfd::gaussian(src.get(in.a, j), m, s);
        } break;
        case GI_EXP: {
            const double al = P[in.p];
#pragma unroll
            for(int j = 0; j < EPT; ++j)
                T[j] = exp(al * src.get(in.a, j));
        } break;
        case GI_EXP_OFF: {
            const double x0 = P[in.p], al = P[in.p + 1];
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                double x = src.get(in.a, j);
                x -= x0;
                T[j] = exp(al * x);
            }
        } break;
        case GI_EXP_POLY: {
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                double x = src.get(in.a, j), e = 0;
                for(int i = 0; i < in.k; ++i)
                    e += pow(x, (double)i) * P[in.p + i];
                T[j] = exp(e);
            }
        } break;
        case GI_EXP_POLY_OFF: {
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                double x = src.get(in.a, j), e = 0;
                x -= P[in.p];
                for(int i = 1; i < in.k; ++i)
                    e += pow(x, (double)i) * P[in.p + i];
                T[j] = exp(e);
            }
        } break;
        case GI_EXPGAUS: {
            const double m = P[in.p], s = P[in.p + 1], al = P[in.p + 2];
#pragma unroll
            for(int j = 0; j < EPT; ++j)
                T[j] = gfd::expgaus(src.get(in.a, j), m, s, al);
        } break;
        case GI_POLY: {
            const int lowest = (int)Cn[in.c];
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                double x = src.get(in.a, j), r = 0;
                for(int i = 0; i < in.k; ++i)
                    r += P[in.p + i] * pow(x, (double)(lowest + i));
                T[j] = r;
            }
        } break;
        case GI_OFF_POLY: {
            const int lowest = (int)Cn[in.c];
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                double x = src.get(in.a, j), r = 0;
                for(int i = 2; i < in.k; ++i)
                    r += P[in.p + i] * pow(x, (double)(lowest + i - 2));
                T[j] = r;
            }
        } break;
        case GI_MULTI_POLY: {
            // device_MultiPolynomial, src/PDFs/basic/PolynomialPdf.cu:47-118
            const int nobs      = in.a;
            const int maxDegree = (int)Cn[in.c] + 1;
            int numBoxes        = 1;
            for(int i = 0; i < nobs; ++i)
                numBoxes *= maxDegree;
#pragma unroll
            for(int j = 0; j < EPT; ++j) {
                int coeff = nobs;
                double r  = P[in.p + coeff];
                coeff++;
                for(int i = 1; i < numBoxes; ++i) {
                    double term = 1;
                    int ci = i, sum = 0;
                    for(int o = 0; o < nobs; ++o) {
                        double x = src.get(gfd::as_const(A.obsid)[in.x + o], j);
                        x -= P[in.p + o];
                        int pw = ci % maxDegree;
                        ci /= maxDegree;
                        term *= pow(x, (double)pw);
                        sum += pw;
                    }
                    if(sum >= maxDegree)
                        continue;
                    term *= P[in.p + coeff];
                    coeff++;
                    r += term;
                }
                T[j] = r;
            }
        } break;
        default:
            if(in.op >= GI_LVL) {
                const int kind = (in.op - GI_LVL) >> 2;
                switch((in.op - GI_LVL) & 3) {
                case 0: step<0, EPT, FEAT>(kind, in, A, P, Nm, src, row, valid, acc, T); break;
                case 1: step<1, EPT, FEAT>(kind, in, A, P, Nm, src, row, valid, acc, T); break;
                case 2: step<2, EPT, FEAT>(kind, in, A, P, Nm, src, row, valid, acc, T); break;
                default: step<3, EPT, FEAT>(kind, in, A, P, Nm, src, row, valid, acc, T); break;
                }
                break;
            }
            return;
        }
    }
}

// the four result registers of a lane of v_mfma_f64_16x16x4f64
typedef double gf_double4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- reductions

// Reduce (s, c) over the block; the result is valid in thread 0.
__device__ __forceinline__ void block_reduce(double &s, double &c) {
    __shared__ double ls[GF_BLOCK / 64], lc[GF_BLOCK / 64];
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) {
        double s2 = __shfl_xor(s, off, 64);
        double c2 = __shfl_xor(c, off, 64);
        gfd::kb_merge(s, c, s2, c2);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if(lane == 0) {
        ls[wave] = s;
        lc[wave] = c;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        s = ls[0];
        c = lc[0];
        for(int w = 1; w < (int)(blockDim.x >> 6); ++w)
            gfd::kb_merge(s, c, ls[w], lc[w]);
    }
    __syncthreads();
}

// Reduce (s, c) over the 16 lanes of a lane's group (lanes 16 g .. 16 g + 15 of its
// wave): the result is valid in every lane of the group.
__device__ __forceinline__ void group_reduce(double &s, double &c) {
#pragma unroll
    for(int off = 8; off > 0; off >>= 1) {
        double s2 = __shfl_xor(s, off, 64);
        double c2 = __shfl_xor(c, off, 64);
        gfd::kb_merge(s, c, s2, c2);
    }
}

// Reduce the block's (sum, residual) pair and store it as partial blockIdx.x.
// No fence and no atomic: on a multi-XCD device an agent-scope release per
// workgroup writes back the XCD's L2 each time, which costs more than the
// whole sweep over the events once there are thousands of workgroups.  The
// partials are added by k_finish, whose launch orders it after every store.
__device__ __forceinline__ void store_partial(double s, double c, Partial *partials) {
    block_reduce(s, c);
    if(threadIdx.x == 0) {
        partials[blockIdx.x].s = s;
        partials[blockIdx.x].c = c;
    }
}

// One workgroup: add the nparts partials in a fixed order (thread t takes
// t, t+256, ...; then the wave butterfly and the waves in order), write
// {sum, residual correction, error word} to *result (host-mapped memory),
// clear the device error word, then stamp seq.
__global__ __launch_bounds__(GF_BLOCK) void k_finish(const Partial *partials, int nparts, GfResult *result,
                                                     int32_t *err, uint32_t seq) {
    double ts = 0.0, tc = 0.0;
    for(int i = threadIdx.x; i < nparts; i += GF_BLOCK)
        gfd::kb_merge(ts, tc, partials[i].s, partials[i].c);
    block_reduce(ts, tc);
    if(threadIdx.x == 0) {
        const int32_t e = *err;
        result->sum     = gfd::kb_total(ts, tc);
        result->comp    = tc;
        result->err     = e;
        if(e)
            *err = 0;
        __threadfence_system();
        __hip_atomic_store(&result->seq, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The in-kernel finish: k_finish's work done inside the sweep by its last
// workgroup (the highest index), with no second launch.  Every other workgroup
// stores its (sum, residual) pair into its slot with agent-scope atomic stores
// (coherent across the XCDs' L2 caches) and exits.  Workgroups are dispatched in
// index order on each XCD, so the last one starts after every workgroup it waits
// for has been placed and holds no resource they need.  It reads the slots with
// agent-scope atomic loads -- a thread's slots t, t+256, ... four at a time, all
// eight loads in flight together -- until none is GF_SLOT_EMPTY, adds them exactly
// as k_finish does (same order, same bits; its own pair taken from LDS), empties
// them for the next launch and writes the result block.  No workgroup runs a
// release fence: on this device an agent-scope release writes back the XCD's L2,
// once per workgroup.  The wait is bounded; a slot still empty after it sets
// GF_DEVERR_FINISH in the result's error word.
// finish_reduced: the same for a workgroup whose pair is reduced already (valid in
// thread 0); every thread of the workgroup calls it.
__device__ __forceinline__ void finish_reduced(double s, double c, const GfProgArgs &A) {
    constexpr uint32_t spin_limit = 1u << 18;
    constexpr int group           = 4;
    __shared__ double own_s, own_c;
    unsigned long long *slots = A.fuse_slots;
    const int nparts          = (int)gridDim.x;
    const int mine            = nparts - 1;
    if((int)blockIdx.x != mine) {
        if(threadIdx.x == 0) {
            __hip_atomic_store(slots + 2 * blockIdx.x, (unsigned long long)__double_as_longlong(s),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(slots + 2 * blockIdx.x + 1, (unsigned long long)__double_as_longlong(c),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    if(threadIdx.x == 0) {
        own_s = s;
        own_c = c;
    }
    __syncthreads();
    double ts = 0.0, tc = 0.0;
    int late  = 0;
    for(int i0 = threadIdx.x; i0 < nparts; i0 += group * GF_BLOCK) {
        unsigned long long vs[group], vc[group];
#pragma unroll
        for(int g = 0; g < group; ++g)
            vs[g] = vc[g] = GF_SLOT_EMPTY;
        for(uint32_t spin = 0;; ++spin) {
#pragma unroll
            for(int g = 0; g < group; ++g) {
                const int i = i0 + g * GF_BLOCK;
                if(i < mine && (vs[g] == GF_SLOT_EMPTY || vc[g] == GF_SLOT_EMPTY)) {
                    vs[g] = __hip_atomic_load(slots + 2 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    vc[g] = __hip_atomic_load(slots + 2 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            bool done = true;
#pragma unroll
            for(int g = 0; g < group; ++g)
                if(i0 + g * GF_BLOCK < mine && (vs[g] == GF_SLOT_EMPTY || vc[g] == GF_SLOT_EMPTY))
                    done = false;
            if(done)
                break;
            if(spin == spin_limit) {
                late = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
#pragma unroll
        for(int g = 0; g < group; ++g) {
            const int i = i0 + g * GF_BLOCK;
            if(i < mine) {
                gfd::kb_merge(ts, tc, __longlong_as_double((long long)vs[g]), __longlong_as_double((long long)vc[g]));
                __hip_atomic_store(slots + 2 * i, GF_SLOT_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(slots + 2 * i + 1, GF_SLOT_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if(i == mine) {
                gfd::kb_merge(ts, tc, own_s, own_c);
            }
        }
    }
    late = __syncthreads_or(late);
    block_reduce(ts, tc);
    if(threadIdx.x == 0) {
        GfResult *result = A.fuse_result;
        const int32_t e  = *A.err;  // written by earlier launches only
        result->sum      = gfd::kb_total(ts, tc);
        result->comp     = tc;
        result->err      = e | (late ? GF_DEVERR_FINISH : 0);
        if(e)
            *A.err = 0;
        __threadfence_system();
        __hip_atomic_store(&result->seq, A.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__device__ __forceinline__ void finish_in_kernel(double s, double c, const GfProgArgs &A) {
    block_reduce(s, c);
    finish_reduced(s, c, A);
}

__global__ __launch_bounds__(256) void k_fuse_init(unsigned long long *slots) {
    for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * GF_MAX_PARTIALS; i += gridDim.x * blockDim.x)
        slots[i] = GF_SLOT_EMPTY;
}

// ------------------------------------------------------------- event kernels

template <int EPT, bool INL, int FEAT>
__global__ __launch_bounds__(GF_BLOCK) void k_reduce_metric(GfProgArgs A, GfInline I, int metric, Partial *partials) {
    const double *__restrict__ P  = INL ? I.p : A.params;
    const double *__restrict__ Nm = INL ? I.n : A.norms;
    const int64_t n    = A.nevents;
    const int64_t tile = (int64_t)GF_BLOCK * EPT;
    const double norm  = Nm[1];
    const bool binned  = metric >= GF_METRIC_BINAVG_CODE;
    const bool nll     = metric == gfd::metric_nll;
    double s = nll ? 1.0 : 0.0, c = 0.0;  // the NLL as a product: (mantissa, exponent)
    for(int64_t base = (int64_t)blockIdx.x * tile; base < n; base += (int64_t)gridDim.x * tile) {
        ColSrc<EPT> src;
        src.cols = A.cols;
        int64_t row[EPT];
        bool valid[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            int64_t e  = base + (int64_t)j * GF_BLOCK + threadIdx.x;
            valid[j]   = e < n;
            row[j]     = valid[j] ? e : n - 1;
            src.idx[j] = row[j];
        }
        double T[EPT];
        run<EPT, FEAT>(A, P, Nm, 0, src, row, valid, T);
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            if(nll) {
                if(valid[j])
                    gfd::nll_add(s, c, T[j], norm);
                continue;
            }
            double e0 = 0.0, e1 = 0.0;
            if(binned) {
                e0 = A.cols[A.nobs - 2][row[j]];
                e1 = A.cols[A.nobs - 1][row[j]];
            }
            double v = gfd::metric(metric, T[j], e0, e1, norm, A.total_events);
            if(valid[j])
                gfd::kb_add(s, c, v);
        }
    }
    if(nll)
        s = gfd::nll_total(s, c), c = 0.0;
    store_partial(s, c, partials);
}

// ------------------------------------------------------------ flat trees

// The event columns of a flat program (at most GF_FLAT_COLS), loaded for the whole
// tile before any of them is used: a lane has every load of the tile in flight at
// once instead of waiting for one column per leaf.
template <int EPT>
struct FlatSrc {
    double v[EPT][GF_FLAT_COLS];
    __device__ __forceinline__ double get(int a, int j) const {
        switch(a) {
        case 0: return v[j][0];
        case 1: return v[j][1];
        case 2: return v[j][2];
        default: return v[j][3];
        }
    }
};

// A leaf of a flat program into T: the arithmetic of run()'s leaf cases.
template <int EPT>
__device__ __forceinline__ void flat_leaf(const GfIns &in, const GF_CONST double *__restrict__ P,
                                          const FlatSrc<EPT> &src, double (&T)[EPT]) {
    switch(in.op) {
    case GI_GAUSS: {
        const double m = P[in.p], sg = P[in.p + 1];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = gfd::gaussian(src.get(in.a, j), m, sg);
    } break;
    case GI_EXP: {
        const double al = P[in.p];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = exp(al * src.get(in.a, j));
    } break;
    case GI_EXP_OFF: {
        const double x0 = P[in.p], al = P[in.p + 1];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            double x = src.get(in.a, j);
            x -= x0;
            T[j] = exp(al * x);
        }
    } break;
    default: {  // GI_EXPGAUS
        const double m = P[in.p], sg = P[in.p + 1], al = P[in.p + 2];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = gfd::expgaus(src.get(in.a, j), m, sg, al);
    } break;
    }
}

// The reduction of a flat program -- one leaf, or one AddPdf / ProdPdf level over
// Gaussian, Exp, Exp-with-offset and ExpGaus leaves, reading columns 0 .. ncols-1
// (ncols <= GF_FLAT_COLS) -- with run()'s arithmetic in the same order (the same
// bits as k_reduce_metric) but one accumulator instead of a stack of levels, no
// Dalitz or polynomial cases, and the tile's columns loaded up front.
template <int EPT, bool INL>
__global__ __launch_bounds__(GF_BLOCK) void k_reduce_flat(GfProgArgs A, GfInline I, int metric, int ncols,
                                                          Partial *partials) {
    const GF_CONST double *__restrict__ P    = gfd::as_const(INL ? I.p : A.params);
    const GF_CONST double *__restrict__ Nm   = gfd::as_const(INL ? I.n : A.norms);
    const GF_CONST GfIns *__restrict__ prog  = gfd::as_const(A.ins);
    const auto cols                          = gfd::as_const(A.cols);
    const int64_t n    = A.nevents;
    const int64_t tile = (int64_t)GF_BLOCK * EPT;
    const double norm  = Nm[1];
    const bool nll     = metric == gfd::metric_nll;
    double s = nll ? 1.0 : 0.0, c = 0.0;  // the NLL as a product: (mantissa, exponent)
    FlatSrc<EPT> src;
#pragma unroll
    for(int j = 0; j < EPT; ++j)
#pragma unroll
        for(int k = 0; k < GF_FLAT_COLS; ++k)
            src.v[j][k] = 0.0;
    for(int64_t base = (int64_t)blockIdx.x * tile; base < n; base += (int64_t)gridDim.x * tile) {
        int64_t row[EPT];
        bool valid[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            const int64_t e = base + (int64_t)j * GF_BLOCK + threadIdx.x;
            valid[j]        = e < n;
            row[j]          = valid[j] ? e : n - 1;
        }
#pragma unroll
        for(int k = 0; k < GF_FLAT_COLS; ++k)
            if(k < ncols) {
                const double *__restrict__ col = cols[k];
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    src.v[j][k] = col[row[j]];
            }
        double T[EPT], acc[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            T[j] = acc[j] = 0.0;
        for(int pc = 0;; ++pc) {
            const GfIns in = prog[pc];
            if(in.op == GI_END)
                break;
            if(in.op < GI_LVL) {
                flat_leaf<EPT>(in, P, src, T);
                continue;
            }
            switch((in.op - GI_LVL) >> 2) {
            case GK_PUSH0:
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    acc[j] = 0.0;
                break;
            case GK_PUSH1:
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    acc[j] = 1.0;
                break;
            case GK_ADD_ACC: {
                const double w = P[in.p], nm = Nm[in.n];
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    acc[j] += w * T[j] * nm;
            } break;
            case GK_ADD_LAST: {
                double tw = 0;
                for(int i = 0; i < in.k; ++i)
                    tw += P[in.p + i];
                const double nm = Nm[in.n];
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    acc[j] += (1 - tw) * T[j] * nm;
            } break;
            case GK_ADDEXT_END: {
                double tw = 0;
                for(int i = 0; i < in.k; ++i)
                    tw += P[in.p + i];
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    T[j] = acc[j] / tw;
            } break;
            case GK_PROD_ACC: {
                const double nm = Nm[in.n];
#pragma unroll
                for(int j = 0; j < EPT; ++j) {
                    double cur = T[j] * nm;
                    acc[j] *= cur;
                }
            } break;
            default:  // GK_POP
#pragma unroll
                for(int j = 0; j < EPT; ++j)
                    T[j] = acc[j];
                break;
            }
        }
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            if(nll) {
                if(valid[j])
                    gfd::nll_add(s, c, T[j], norm);
                continue;
            }
            const double v = gfd::metric(metric, T[j], 0.0, 0.0, norm, A.total_events);
            if(valid[j])
                gfd::kb_add(s, c, v);
        }
    }
    if(nll)
        s = gfd::nll_total(s, c), c = 0.0;
    if(A.fuse_slots)
        finish_in_kernel(s, c, A);
    else
        store_partial(s, c, partials);
}

template <int EPT, bool INL, int FEAT>
__global__ __launch_bounds__(GF_BLOCK) void k_eval_metric(GfProgArgs A, GfInline I, int metric, double *out) {
    const double *__restrict__ P  = INL ? I.p : A.params;
    const double *__restrict__ Nm = INL ? I.n : A.norms;
    const int64_t n    = A.nevents;
    const int64_t tile = (int64_t)GF_BLOCK * EPT;
    const double norm  = Nm[1];
    const bool binned  = metric >= GF_METRIC_BINAVG_CODE;
    for(int64_t base = (int64_t)blockIdx.x * tile; base < n; base += (int64_t)gridDim.x * tile) {
        ColSrc<EPT> src;
        src.cols = A.cols;
        int64_t row[EPT];
        bool valid[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            int64_t e  = base + (int64_t)j * GF_BLOCK + threadIdx.x;
            valid[j]   = e < n;
            row[j]     = valid[j] ? e : n - 1;
            src.idx[j] = row[j];
        }
        double T[EPT];
        run<EPT, FEAT>(A, P, Nm, 0, src, row, valid, T);
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            double e0 = 0.0, e1 = 0.0;
            if(binned) {
                e0 = A.cols[A.nobs - 2][row[j]];
                e1 = A.cols[A.nobs - 1][row[j]];
            }
            double v = gfd::metric(metric, T[j], e0, e1, norm, A.total_events);
            if(valid[j])
                out[row[j]] = v;
        }
    }
}

// ------------------------------------------------------ coherent-sum fast path

// Multivariate polynomial of one event (device_MultiPolynomial,
// src/PDFs/basic/PolynomialPdf.cu:47-118), columns read at row e.
__device__ __forceinline__ double multipoly_at(const GfIns &in, const GfProgArgs &A,
                                               const GF_CONST double *__restrict__ P, int64_t e) {
    const int nobs      = in.a;
    const int maxDegree = (int)gfd::as_const(A.consts)[in.c] + 1;
    int numBoxes        = 1;
    for(int i = 0; i < nobs; ++i)
        numBoxes *= maxDegree;
    int coeff = nobs;
    double r  = P[in.p + coeff];
    coeff++;
    for(int i = 1; i < numBoxes; ++i) {
        double term = 1;
        int ci = i, sum = 0;
        for(int o = 0; o < nobs; ++o) {
            double x = gfd::as_const(A.cols)[gfd::as_const(A.obsid)[in.x + o]][e];
            x -= P[in.p + o];
            int pw = ci % maxDegree;
            ci /= maxDegree;
            term *= pow(x, (double)pw);
            sum += pw;
        }
        if(sum >= maxDegree)
            continue;
        term *= P[in.p + coeff];
        coeff++;
        r += term;
    }
    return r;
}

// |sum_i c_i W_i[e]|^2 * eff for a Dalitz node at the root of the tree.  All
// wave loads of an event are issued before any arithmetic (16 x 16 bytes per lane
// in flight); the rest is device_DalitzPlot's arithmetic, the coherent sum in
// gfd::coh_sum's grouped order (the matrix-core kernel's).
// ZERO_OUT (constant efficiency, cache filled from these events): the waves of an event
// outside the Dalitz region are all 0, so |A|^2 eff is the reference's 0 there without
// reading m12 and m13 (256 instead of 272 bytes per event).
template <bool WRITE, bool INL, bool CONST_EFF, bool ZERO_OUT = false>
__global__ __launch_bounds__(256) void k_dalitz_fast(GfProgArgs A, GfInline I, int metric, Partial *partials,
                                                     double *out) {
    const double *__restrict__ P  = INL ? I.p : A.params;
    const double *__restrict__ Nm = INL ? I.n : A.norms;
    const GfIns dzi       = A.ins[0];
    const GfIns effi      = A.ins[1];
    const GfDalitzSlot dz = A.dz[dzi.x];
    const GfDecay &D      = A.decay;
    const int nres        = dzi.a;
    const int64_t n       = A.nevents;
    const double norm     = Nm[1];
    const double *__restrict__ c12 = A.cols[dz.col_m12];
    const double *__restrict__ c13 = A.cols[dz.col_m13];
    const bool nll        = !WRITE && metric == gfd::metric_nll;
    // the events of the NLL grouped as k_coh_mx groups a set's (the multi-point sums of
    // these parameters run there): lanes 0-15 of a wave hold the products, below
    const bool grp        = nll && CONST_EFF && ZERO_OUT && A.coh_group;
    double s = nll ? 1.0 : 0.0, c = 0.0;  // the NLL as a product: (mantissa, exponent)
    for(int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e0 = base + threadIdx.x;
        const bool valid = e0 < n;
        const int64_t e  = valid ? e0 : n - 1;
        double2 w[GF_MAX_RES];
#pragma unroll
        for(int i = 0; i < GF_MAX_RES; ++i)
            w[i] = (i < nres) ? dz.wave[(int64_t)i * dz.ncache + e] : make_double2(0.0, 0.0);
        const double m12 = ZERO_OUT ? 0.0 : c12[e];
        const double m13 = ZERO_OUT ? 0.0 : c13[e];
        double re = 0.0, im = 0.0;
        gfd::coh_sum<GF_MAX_RES>(
            re, im, [&](int i) { return make_double2(P[dzi.p + 2 * i], P[dzi.p + 2 * i + 1]); }, w, nres);
        const bool inside =
            ZERO_OUT || gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
        // a degree-0 polynomial is its constant coefficient (numBoxes == 1)
        const double eff = CONST_EFF ? P[effi.p + effi.a] : multipoly_at(effi, A, gfd::as_const(P), e);
        const double T   = inside ? (re * re + im * im) * eff : 0.0;
        if(WRITE) {
            if(valid)
                out[e] = gfd::metric(metric, T, 0.0, 0.0, norm, A.total_events);
        } else if(grp) {
            // k_coh_mx's grouping: lane (0, j) takes events 16 t + j of the wave, in tile order
            const int64_t w0 = base + (threadIdx.x & ~63);
#pragma unroll
            for(int t = 0; t < 4; ++t) {
                const double v = t == 0 ? T : __shfl(T, 16 * t + (threadIdx.x & 15), 64);
                if((threadIdx.x & 63) < 16 && w0 + 16 * t + (threadIdx.x & 15) < n)
                    gfd::nll_add(s, c, v, norm);
            }
        } else if(nll) {
            if(valid)
                gfd::nll_add(s, c, T, norm);
        } else if(valid) {
            gfd::kb_add(s, c, gfd::metric(metric, T, 0.0, 0.0, norm, A.total_events));
        }
    }
    if(!WRITE && grp) {
        // as k_coh_mx reduces a set: the 16 lanes of the group, then the four waves in order
        __shared__ double gls[4], glc[4];
        double sv = (threadIdx.x & 63) < 16 ? gfd::nll_total(s, c) : 0.0, cv = 0.0;
        group_reduce(sv, cv);
        if((threadIdx.x & 63) == 0) {
            gls[threadIdx.x >> 6] = sv;
            glc[threadIdx.x >> 6] = cv;
        }
        __syncthreads();
        if(threadIdx.x == 0) {
            sv = gls[0];
            cv = glc[0];
            for(int wv = 1; wv < 4; ++wv)
                gfd::kb_merge(sv, cv, gls[wv], glc[wv]);
        }
        if(A.fuse_slots) {
            finish_reduced(sv, cv, A);
        } else if(threadIdx.x == 0) {
            partials[blockIdx.x].s = sv;
            partials[blockIdx.x].c = cv;
        }
    } else if(!WRITE) {
        if(nll)
            s = gfd::nll_total(s, c), c = 0.0;
        if(A.fuse_slots)
            finish_in_kernel(s, c, A);
        else
            store_partial(s, c, partials);
    }
}

// The coherent-sum metric sum of up to GF_MULTI_PTS parameter sets over the same
// events in one sweep: an event's waves are loaded once and every set's term is
// taken with the arithmetic, the event-to-lane assignment and the reduction
// order of k_dalitz_fast on the same grid, so each set's partials equal that
// kernel's for the same parameters bit for bit.  params / norms: npts
// consecutive copies of the flattened arrays (np and nn doubles each).
// partials: [npts][gridDim.x].
// The parameter sets of one multi-set sweep as the event loop reads them: the
// couplings of set p at c[p][i] (zero for resonances i >= nres, so that a resonance
// loop unrolled to NR >= nres adds exact zeros past nres), each set's constant
// efficiency and normalisation.  Staged once per workgroup in LDS (4.3 KB): the
// loop then reads them with broadcast LDS loads instead of one scalar load and
// wait per coupling.
struct MultiSets {
    // one row per set, padded by one entry: 16 lanes reading resonance i of sets 0..15
    // (the matrix-core kernels' A operands) then hit 16 different bank groups
    double2 c[GF_MULTI_PTS][GF_MAX_RES + 1];
    double eff[GF_MULTI_PTS], norm[GF_MULTI_PTS];
};
static_assert(GF_MULTI_PTS * GF_MAX_RES == 256, "one staging thread per (set, resonance)");

// Fill S from cnt sets: set p's parameter array at P0 + p * pstride, its
// normalisation array at N0 + p * nstride; couplings from index pc, the constant
// efficiency (CONST_EFF) at index pe.  All 256 threads of the workgroup take part.
template <bool CONST_EFF>
__device__ __forceinline__ void stage_sets(MultiSets &S, const double *__restrict__ P0, int64_t pstride,
                                           const double *__restrict__ N0, int64_t nstride, int cnt, int pc, int pe,
                                           int nres) {
    const int t = threadIdx.x, p = t >> 4, i = t & 15;
    double2 v   = make_double2(0.0, 0.0);
    if(p < cnt && i < nres) {
        const GF_CONST double *__restrict__ P = gfd::as_const(P0 + p * pstride);
        v                                     = make_double2(P[pc + 2 * i], P[pc + 2 * i + 1]);
    }
    S.c[p][i] = v;
    if(t < GF_MULTI_PTS) {
        S.eff[t]  = (CONST_EFF && t < cnt) ? gfd::as_const(P0 + t * pstride)[pe] : 0.0;
        S.norm[t] = t < cnt ? gfd::as_const(N0 + t * nstride)[1] : 0.0;
    }
    __syncthreads();
}

// The coherent-sum metric of up to GF_MULTI_PTS parameter sets over the same
// events in one sweep: an event's waves are loaded once and every set's term is
// taken with the arithmetic, the event-to-lane assignment and the reduction
// order of k_dalitz_fast on the same grid, so each set's partials equal that
// kernel's for the same parameters bit for bit (the resonance loop runs to
// NR >= nres with zero couplings and waves past nres, whose fused terms add
// exact zeros).  params / norms: npts consecutive copies of the flattened arrays
// (np and nn doubles each).  partials: [npts][gridDim.x].
template <bool CONST_EFF, bool ZERO_OUT, int NR, bool NLL>
__global__ __launch_bounds__(256) void k_dalitz_multi(GfProgArgs A, const double *__restrict__ params, int np,
                                                      const double *__restrict__ norms, int nn, int npts, int metric,
                                                      Partial *partials) {
    __shared__ MultiSets S;
    const GfIns dzi       = A.ins[0];
    const GfIns effi      = A.ins[1];
    const GfDalitzSlot dz = gfd::as_const(A.dz)[dzi.x];
    const GfDecay &D      = A.decay;
    const int nres        = dzi.a;
    const int64_t n       = A.nevents;
    const double *__restrict__ c12 = A.cols[dz.col_m12];
    const double *__restrict__ c13 = A.cols[dz.col_m13];
    stage_sets<CONST_EFF>(S, params, np, norms, nn, npts, dzi.p, effi.p + effi.a, nres);
    constexpr bool nll = NLL;  // the launcher sets it for metric == NLL
    double s[GF_MULTI_PTS], c[GF_MULTI_PTS];  // per set: (sum, residual), or the NLL's (mantissa, exponent)
#pragma unroll
    for(int p = 0; p < GF_MULTI_PTS; ++p) {
        s[p] = nll ? 1.0 : 0.0;
        c[p] = 0.0;
    }
    for(int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        // keep the LDS reads of the sets inside the loop: hoisted, the 16 x NR
        // loop-invariant couplings would not fit in registers
        asm volatile("" ::: "memory");
        const int64_t e0 = base + threadIdx.x;
        const bool valid = e0 < n;
        const int64_t e  = valid ? e0 : n - 1;
        double2 w[NR];
#pragma unroll
        for(int i = 0; i < NR; ++i)
            w[i] = (i < nres) ? dz.wave[(int64_t)i * dz.ncache + e] : make_double2(0.0, 0.0);
        const double m12  = ZERO_OUT ? 0.0 : c12[e];
        const double m13  = ZERO_OUT ? 0.0 : c13[e];
        const bool inside =
            ZERO_OUT || gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
#pragma unroll
        for(int p = 0; p < GF_MULTI_PTS; ++p) {
            if(p < npts) {
                double re = 0.0, im = 0.0;
                gfd::coh_sum<NR>(re, im, [&](int i) { return S.c[p][i]; }, w, nres);
                const double eff =
                    CONST_EFF ? S.eff[p] : multipoly_at(effi, A, gfd::as_const(params + (int64_t)p * np), e);
                const double T = inside ? (re * re + im * im) * eff : 0.0;
                if(nll) {
                    if(valid)
                        gfd::nll_add(s[p], c[p], T, S.norm[p]);
                } else if(valid) {
                    gfd::kb_add(s[p], c[p], gfd::metric(metric, T, 0.0, 0.0, S.norm[p], A.total_events));
                }
            }
        }
    }
#pragma unroll
    for(int p = 0; p < GF_MULTI_PTS; ++p) {
        if(p < npts) {
            double ss = s[p], cc = c[p];
            if(nll)
                ss = gfd::nll_total(ss, cc), cc = 0.0;
            block_reduce(ss, cc);
            if(threadIdx.x == 0) {
                partials[(int64_t)p * gridDim.x + blockIdx.x].s = ss;
                partials[(int64_t)p * gridDim.x + blockIdx.x].c = cc;
            }
        }
    }
}

// ------------------------------------------------------- coherent sums on MFMA
//
// The NLL of up to 16 coupling sets over the same events, constant efficiency, waves 0
// outside the Dalitz region (the k_dalitz_multi<true, true, NR, true> case), with the
// coherent sums of a wave's 64 events formed on the FP64 matrix cores.  Per tile of 16
// events, Re_p and Im_p are rows of real matrix products M W over K = 2 NR terms, on
// v_mfma_f64_16x16x4f64 (NR / 2 steps of K = 4).  The instruction adds its K terms to
// the accumulator one after the other, each as a fused multiply-add.  K-step 2 m takes
// the real parts of the waves of resonances 4 m .. 4 m + 3 (Re: cr wr; Im: ci wr),
// K-step 2 m + 1 their imaginary parts (Re: -ci wi; Im: cr wi): the order of
// gfd::coh_sum, so every element equals the vector sweeps' sum bit for bit.
// Operand layouts (gfx950; the host uses this kernel only after its probe of the
// instruction found the result layout assumed here): lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] of a K-step and holds D[4 r + l / 16][l % 16] in result register
// r.  The couplings are loop-invariant A operands (16 doubles a lane, read from LDS once).
// Lane (g = l / 16, j = l % 16) loads the wave of resonance 4 m + g of event 16 t + j
// (one 16-byte load per group of four resonances: a 256-byte run of 16 events per
// resonance) and gives its real part to K-step 2 m and its imaginary part to 2 m + 1,
// for both products.  Lane (g, j) ends tile t holding |A|^2 eff of sets 4 r + g of event
// 16 t + j, and multiplies it into those sets' NLL products: per set, lane (g, j) takes
// events 16 t + j of each of the wave's tiles in order.  The sets are then reduced over
// the 16 lanes of their group (xor 8, 4, 2, 1) and the four waves in order, partials
// [p][b].  k_dalitz_fast groups the events of a set the same way whenever this kernel
// serves the multi-point sums (GfProgArgs::coh_group), so a multi-point sum equals the
// single sweep of the same parameters.  The register budget is that of three waves per
// SIMD, the occupancy the grid of three workgroups per CU asks for.
//
// coh_mx_rows runs the tiles of this workgroup's events (rows first + k stride +
// 0..255 below end, wave w taking rows 64 w .. 64 w + 63 of each) for the NSG x 16 sets
// staged in S[0 .. NSG - 1] (npts of them in all) and leaves set p's (sum, residual) of
// the workgroup in ls[p], lc[p] (every thread returns after the workgroup barrier).
// With two set groups the wave loads of a tile serve both (the HBM bytes of a sweep are
// shared by 32 sets); the first group's couplings stay in registers, the second's are
// read from LDS per tile, so the kernel keeps three waves per SIMD.  Every set is
// reduced over the same lanes and events in the same order whichever group it is in.
// All 256 threads call it.
template <int NR, int NSG>
__device__ __forceinline__ void coh_mx_rows(const MultiSets *S, const GfDalitzSlot &dz, int nres, int64_t first,
                                            int64_t stride, int64_t end, int npts, double *ls, double *lc) {
    static_assert(NR % 4 == 0 && NR <= GF_MAX_RES, "NR: 4, 8, 12 or 16");
    static_assert(NSG == 1 || NSG == 2, "one or two groups of 16 sets");
    constexpr int NG = NR / 4;  // groups of four resonances: K-steps 2 m (real parts), 2 m + 1 (imaginary)
    __shared__ double wls[4][GF_MX_SETS], wlc[4][GF_MX_SETS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
    // A[j][k]: set j's factor of K term k = 4 kk + g (resonance 4 (kk / 2) + g): K-step 2 m
    // takes (cr | ci) against wr, 2 m + 1 (-ci | cr) against wi
    auto a_ops = [&](const double2 c, double &ar0, double &ai0, double &ar1, double &ai1) {
        ar0 = c.x;   // Re: cr wr
        ai0 = c.y;   // Im: ci wr
        ar1 = -c.y;  // Re: -ci wi
        ai1 = c.x;   // Im: cr wi
    };
    double are[2 * NG], aim[2 * NG];
#pragma unroll
    for(int m = 0; m < NG; ++m)
        a_ops(S[0].c[j][4 * m + g], are[2 * m], aim[2 * m], are[2 * m + 1], aim[2 * m + 1]);
    double m_[4 * NSG];  // per set: the NLL product's mantissa and (integer) exponent
    int ex[4 * NSG];
#pragma unroll
    for(int r = 0; r < 4 * NSG; ++r) {
        m_[r] = 1.0;
        ex[r] = 0;
    }
    // The wave's tiles in order: tile q = 4 p + t covers rows first + p stride + 64 wave +
    // 16 t + (0..15) (pass p of the strided loop, quarter t of the wave's 64 rows).
    // Tile q + 2's waves (q + 1's with two set groups) are requested before tile q's
    // products start, so loads stay in flight behind the matrix cores; a tile past the end
    // reads row end - 1 (one line per resonance) and is never used.
    const double2 *__restrict__ wv = dz.wave;
    const int64_t ntile            = first < end ? 4 * ((end - 1 - first) / stride + 1) : 0;
    auto tile_event = [&](int64_t q) { return first + (q >> 2) * stride + 64 * wave + 16 * (q & 3) + j; };
    // Only the last group can reach past nres (NR - 4 < nres <= NR): there the load reads
    // resonance nres - 1 and the value is replaced by zero, without a branch around it.
    auto load_tile = [&](int64_t q, double2(&w)[NG]) {
        const int64_t ev  = tile_event(q);
        const int64_t row = ev < end ? ev : end - 1;  // loads stay in bounds; the row is masked below
#pragma unroll
        for(int m = 0; m < NG; ++m) {
            const int i  = 4 * m + g;
            const int ir = (m == NG - 1 && i >= nres) ? nres - 1 : i;
            w[m]         = wv[(int64_t)ir * dz.ncache + row];
        }
    };
    auto run_tile = [&](int64_t q, const double2(&w)[NG]) {
        const bool live     = tile_event(q) < end;
        const bool all_live = tile_event(q) - j + 15 < end;  // uniform over the wave
#pragma unroll
        for(int sg = 0; sg < NSG; ++sg) {
            gf_double4 re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
#pragma unroll
            for(int m = 0; m < NG; ++m) {
                const bool pad   = m == NG - 1 && 4 * m + g >= nres;
                const double2 wk = pad ? make_double2(0.0, 0.0) : w[m];
                double ar0, ai0, ar1, ai1;
                if(sg == 0) {
                    ar0 = are[2 * m], ai0 = aim[2 * m], ar1 = are[2 * m + 1], ai1 = aim[2 * m + 1];
                } else {
                    a_ops(S[sg].c[j][4 * m + g], ar0, ai0, ar1, ai1);
                }
                re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar0, wk.x, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai0, wk.x, im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar1, wk.y, re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai1, wk.y, im, 0, 0, 0);
            }
            // The sets' efficiencies and normalisations are read from LDS per tile (held in
            // registers they would cost the occupancy of three waves per SIMD).  A set past
            // npts has efficiency and normalisation 0 (stage_sets), so its factor is 0, which
            // the product takes as exactly 1 (mantissa 1/2, exponent +1, renormalised back:
            // the pair is unchanged); a lane past the last row is given the factor 1 outright,
            // in the one tile per wave that reaches past it.
            asm volatile("" ::: "memory");
            double ef[4], nm[4];
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                ef[r] = S[sg].eff[4 * r + g];
                nm[r] = S[sg].norm[4 * r + g];
            }
            if(all_live) {
#pragma unroll
                for(int r = 0; r < 4; ++r)
                    gfmath::nll_add(m_[4 * sg + r], ex[4 * sg + r], ((re[r] * re[r] + im[r] * im[r]) * ef[r]) * nm[r]);
            } else {
#pragma unroll
                for(int r = 0; r < 4; ++r) {
                    const double v = ((re[r] * re[r] + im[r] * im[r]) * ef[r]) * nm[r];
                    gfmath::nll_add(m_[4 * sg + r], ex[4 * sg + r], live ? v : 1.0);
                }
            }
        }
    };
    if(ntile > 0 && NSG == 1) {
        double2 wa[NG], wb[NG], wc[NG];
        load_tile(0, wa);
        load_tile(1, wb);
        for(int64_t q = 0;; q += 3) {
            load_tile(q + 2, wc);
            run_tile(q, wa);
            if(q + 1 >= ntile)
                break;
            load_tile(q + 3, wa);
            run_tile(q + 1, wb);
            if(q + 2 >= ntile)
                break;
            load_tile(q + 4, wb);
            run_tile(q + 2, wc);
            if(q + 3 >= ntile)
                break;
        }
    } else if(ntile > 0) {
        // two groups: a tile's products take twice as long, so one tile of loads ahead
        // covers the same time (and the registers of a third buffer are not there)
        double2 wa[NG], wb[NG];
        load_tile(0, wa);
        for(int64_t q = 0;; q += 2) {
            load_tile(q + 1, wb);
            run_tile(q, wa);
            if(q + 1 >= ntile)
                break;
            load_tile(q + 2, wa);
            run_tile(q + 1, wb);
            if(q + 2 >= ntile)
                break;
        }
    }
    // set p: the 16 lanes of its group (xor 8, 4, 2, 1), then the four waves in order
#pragma unroll
    for(int r = 0; r < 4 * NSG; ++r) {
        double sv = gfd::nll_total(m_[r], (double)ex[r]), cv = 0.0;
        group_reduce(sv, cv);
        if(j == 0) {
            const int p    = GF_MULTI_PTS * (r >> 2) + 4 * (r & 3) + g;
            wls[wave][p]   = sv;
            wlc[wave][p]   = cv;
        }
    }
    __syncthreads();
    if((int)threadIdx.x < npts) {
        const int p = threadIdx.x;
        double sv   = wls[0][p], cv = wlc[0][p];
        for(int w = 1; w < 4; ++w)
            gfd::kb_merge(sv, cv, wls[w][p], wlc[w][p]);
        ls[p] = sv;
        lc[p] = cv;
    }
    __syncthreads();
}

// Up to NSG x 16 sets per sweep (partials [p][b]); NSG = 2 shares each wave load
// between two groups of 16 sets.
template <int NR, int NSG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_coh_mx(
    GfProgArgs A, const double *__restrict__ params, int np, const double *__restrict__ norms, int nn, int npts,
    Partial *partials) {
    __shared__ MultiSets S[NSG];
    __shared__ double ls[GF_MX_SETS], lc[GF_MX_SETS];
    const GfIns dzi       = A.ins[0];
    const GfIns effi      = A.ins[1];
    const GfDalitzSlot dz = gfd::as_const(A.dz)[dzi.x];
    const int nres        = dzi.a;
#pragma unroll
    for(int sg = 0; sg < NSG; ++sg)
        stage_sets<true>(S[sg], params + (int64_t)GF_MULTI_PTS * sg * np, np, norms + (int64_t)GF_MULTI_PTS * sg * nn,
                         nn, min(max(npts - GF_MULTI_PTS * sg, 0), GF_MULTI_PTS), dzi.p, effi.p + effi.a, nres);
    coh_mx_rows<NR, NSG>(S, dz, nres, (int64_t)blockIdx.x * blockDim.x, (int64_t)gridDim.x * blockDim.x, A.nevents,
                         npts, ls, lc);
    if((int)threadIdx.x < npts) {
        const int p                                    = threadIdx.x;
        partials[(int64_t)p * gridDim.x + blockIdx.x].s = ls[p];
        partials[(int64_t)p * gridDim.x + blockIdx.x].c = lc[p];
    }
}

// k_finish for several sums: workgroup p adds the nparts partials of set p in the
// order k_finish uses and writes {sum, residual} to out[2p] (host-mapped memory that
// the host polls): the residual, a system-scope fence, then the sum over the empty
// marker the host left there.
__global__ __launch_bounds__(GF_BLOCK) void k_finish_multi(const Partial *partials, int nparts, double *out) {
    const Partial *q = partials + (int64_t)blockIdx.x * nparts;
    double ts = 0.0, tc = 0.0;
    for(int i = threadIdx.x; i < nparts; i += GF_BLOCK)
        gfd::kb_merge(ts, tc, q[i].s, q[i].c);
    block_reduce(ts, tc);
    if(threadIdx.x == 0) {
        out[2 * blockIdx.x + 1] = tc;
        __threadfence_system();
        __hip_atomic_store((unsigned long long *)(out + 2 * blockIdx.x),
                           (unsigned long long)__double_as_longlong(gfd::kb_total(ts, tc)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace

// Grid description of k_reduce_bins (by value; indexed only by unrolled constants).
struct GfGrid {
    double lo[GF_MAX_OBS], hi[GF_MAX_OBS];
    int nb[GF_MAX_OBS];
    int slot[GF_MAX_OBS];
    int k;
};

namespace {

// The centres of global bin `bin` of grid G in every dimension (MetricTaker.cu:76-86:
// the first observable varies fastest, local bin = bin % n, then bin /= n).
__device__ __forceinline__ void grid_point(const GfGrid &G, int64_t bin, double (&x)[GF_MAX_OBS]) {
#pragma unroll
    for(int d = 0; d < GF_MAX_OBS; ++d) {
        x[d] = 0.0;
        if(d < G.k) {
            const int64_t localBin = bin % G.nb[d];
            x[d]                   = gfd::bin_centre(G.lo[d], G.hi[d], G.nb[d], localBin);
            bin /= G.nb[d];
        }
    }
}

// The synthetic event of grid bin `bin` in slot j of src: every observable slot 0,
// then each grid dimension's centre at its observable slot.
template <int EPT>
__device__ __forceinline__ void grid_event(const GfGrid &G, int64_t bin, RegSrc<EPT> &src, int j) {
    double x[GF_MAX_OBS];
    grid_point(G, bin, x);
#pragma unroll
    for(int o = 0; o < GF_MAX_OBS; ++o)
        src.v[j][o] = 0.0;
#pragma unroll
    for(int d = 0; d < GF_MAX_OBS; ++d)
        if(d < G.k)
            src.set(G.slot[d], j, x[d]);
}

template <int EPT>
__global__ __launch_bounds__(GF_BLOCK) void k_reduce_bins(GfProgArgs A, int pc0, GfGrid G, int64_t total,
                                                          Partial *partials) {
    const int64_t tile = (int64_t)GF_BLOCK * EPT;
    double s = 0.0, c = 0.0;
    for(int64_t base = (int64_t)blockIdx.x * tile; base < total; base += (int64_t)gridDim.x * tile) {
        RegSrc<EPT> src;
        int64_t row[EPT];
        bool valid[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            int64_t b = base + (int64_t)j * GF_BLOCK + threadIdx.x;
            valid[j]  = b < total;
            row[j]    = valid[j] ? b : 0;
            grid_event(G, row[j], src, j);
        }
        double T[EPT];
        run<EPT>(A, A.params, A.norms, pc0, src, row, valid, T);
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            if(valid[j])
                gfd::kb_add(s, c, T[j]);
    }
    store_partial(s, c, partials);
}

// Per-bin raw values of the subprogram at pc0 over grid G: the transform that
// k_reduce_bins sums (same cells, same evaluation), written to out[bin].
template <int EPT>
__global__ __launch_bounds__(GF_BLOCK) void k_eval_bins(GfProgArgs A, int pc0, GfGrid G, int64_t total, double *out) {
    const int64_t tile = (int64_t)GF_BLOCK * EPT;
    for(int64_t base = (int64_t)blockIdx.x * tile; base < total; base += (int64_t)gridDim.x * tile) {
        RegSrc<EPT> src;
        int64_t row[EPT];
        bool valid[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            int64_t b = base + (int64_t)j * GF_BLOCK + threadIdx.x;
            valid[j]  = b < total;
            row[j]    = valid[j] ? b : 0;
            grid_event(G, row[j], src, j);
        }
        double T[EPT];
        run<EPT>(A, A.params, A.norms, pc0, src, row, valid, T);
#pragma unroll
        for(int j = 0; j < EPT; ++j)
            if(valid[j])
                out[row[j]] = T[j];
    }
}

// The cell centres of grid G: out[d * total + bin] = coordinate d of bin.
__global__ __launch_bounds__(256) void k_grid_cells(GfGrid G, int64_t total, double *out) {
    for(int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < total; b += (int64_t)gridDim.x * blockDim.x) {
        double x[GF_MAX_OBS];
        grid_point(G, b, x);
#pragma unroll
        for(int d = 0; d < GF_MAX_OBS; ++d)
            if(d < G.k)
                out[(int64_t)d * total + b] = x[d];
    }
}

// ------------------------------------------------------------- Dalitz cache

__global__ __launch_bounds__(256) void k_dalitz_cache(GfProgArgs A, int slot, int64_t offset, int64_t n,
                                                      uint32_t mask, const GfResDesc *__restrict__ res, int nres) {
    const GfDalitzSlot dz = A.dz[slot];
    const GfDecay D       = A.decay;
    double2 *wave         = const_cast<double2 *>(dz.wave);
    double2 *wave_b       = const_cast<double2 *>(dz.wave_b);  // time-dependent nodes only
    const double *c12     = A.cols[dz.col_m12];
    const double *c13     = A.cols[dz.col_m13];
    for(int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const double m12  = c12[offset + e];
        const double m13  = c13[offset + e];
        const bool inside = gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
        // SpecialResonanceCalculator.cu:36-37
        const double m23 = D.motherMass * D.motherMass + D.daug1Mass * D.daug1Mass + D.daug2Mass * D.daug2Mass
                           + D.daug3Mass * D.daug3Mass - m12 - m13;
        for(int i = 0; i < nres; ++i) {
            if(!((mask >> i) & 1u))
                continue;
            double2 w = make_double2(0.0, 0.0);
            if(inside)
                w = gfd::lineshape(res[i], A.params, A.consts, D, m12, m13, m23);
            wave[(int64_t)i * dz.ncache + e] = w;
            if(wave_b) {
                // SpecialWaveCalculator.cu:55-58: the B wave at the reversed point, same m23
                double2 v = make_double2(0.0, 0.0);
                if(inside)
                    v = gfd::lineshape(res[i], A.params, A.consts, D, m13, m12, m23);
                wave_b[(int64_t)i * dz.ncache + e] = v;
            }
        }
    }
}

// ------------------------------------------------------------- toy batches

// One entry of a batch's work table: rows [start, end) of toy `toy`.
// (GfBatchWork is declared in gf_internal.h)

// The metric sum over rows [w.start, w.end) of one toy with parameter array P and
// normalisation array Nm, accumulated per thread in the tile order of
// k_batch_metric (EPT rows per thread per tile).
template <int EPT, int FEAT>
__device__ __forceinline__ void batch_rows(const GfProgArgs &A, const GfBatchWork &w, const double *__restrict__ P,
                                           const double *__restrict__ Nm, int metric, double &s, double &c) {
    const double norm  = Nm[1];
    const bool binned  = metric >= GF_METRIC_BINAVG_CODE;
    const bool nll     = metric == gfd::metric_nll;
    const int64_t tile = (int64_t)GF_BLOCK * EPT;
    if(nll)  // the NLL as a product: (mantissa, exponent)
        s = 1.0, c = 0.0;
    for(int64_t base = w.start; base < w.end; base += tile) {
        ColSrc<EPT> src;
        src.cols = A.cols;
        int64_t row[EPT];
        bool valid[EPT];
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            const int64_t r = base + (int64_t)j * GF_BLOCK + threadIdx.x;
            valid[j]        = r < w.end;
            row[j]          = valid[j] ? r : w.start;
            src.idx[j]      = row[j];
        }
        double T[EPT];
        run<EPT, FEAT>(A, P, Nm, 0, src, row, valid, T);
#pragma unroll
        for(int j = 0; j < EPT; ++j) {
            if(!valid[j])
                continue;
            if(nll) {
                gfd::nll_add(s, c, T[j], norm);
                continue;
            }
            double e0 = 0.0, e1 = 0.0;
            if(binned) {
                e0 = A.cols[A.nobs - 2][row[j]];
                e1 = A.cols[A.nobs - 1][row[j]];
            }
            gfd::kb_add(s, c, gfd::metric(metric, T[j], e0, e1, norm, w.total));
        }
    }
    if(nll)
        s = gfd::nll_total(s, c), c = 0.0;
}

// Metric sums of many toys in one launch: workgroup b evaluates the rows of
// work[b] with its toy's parameter and normalisation arrays (P = params +
// toy * np, Nm = norms + toy * nn) and stores one partial; workgroups of
// inactive toys store zero.  FEAT as for k_reduce_metric.
template <int EPT, int FEAT>
__global__ __launch_bounds__(GF_BLOCK) void k_batch_metric(GfProgArgs A, const GfBatchWork *__restrict__ work,
                                                           const double *__restrict__ params, int np,
                                                           const double *__restrict__ norms, int nn,
                                                           const uint8_t *__restrict__ active, int metric,
                                                           Partial *partials) {
    const GfBatchWork w = work[blockIdx.x];
    double s = 0.0, c = 0.0;
    if(active[w.toy])
        batch_rows<EPT, FEAT>(A, w, params + (int64_t)w.toy * np, norms + (int64_t)w.toy * nn, metric, s, c);
    store_partial(s, c, partials);
}

// Several points per toy, any program: workgroup (b, g) evaluates the rows of
// work[b] once per point p in [16 g, min(16 g + 16, npts)) of its toy and stores
// partial (b, p).  Each point's sum takes the path of k_batch_metric.
template <int EPT, int FEAT>
__global__ __launch_bounds__(GF_BLOCK) void k_batch_points(GfProgArgs A, const GfBatchWork *__restrict__ work,
                                                           const double *__restrict__ stage, int np, int nn,
                                                           const int *__restrict__ tbl, int metric,
                                                           Partial *partials) {
    const GfBatchWork w = work[blockIdx.x];
    const int slot0     = tbl[2 * w.toy];
    const int npts      = tbl[2 * w.toy + 1];
    const int p0        = (int)blockIdx.y * GF_MULTI_PTS;
    const int p1        = min(npts, p0 + GF_MULTI_PTS);
    for(int p = p0; p < p1; ++p) {
        const double *__restrict__ P = stage + (int64_t)(slot0 + p) * (np + nn);
        double s = 0.0, c = 0.0;
        batch_rows<EPT, FEAT>(A, w, P, P + np, metric, s, c);
        block_reduce(s, c);
        if(threadIdx.x == 0) {
            partials[(int64_t)blockIdx.x * GF_BATCH_PTS + p].s = s;
            partials[(int64_t)blockIdx.x * GF_BATCH_PTS + p].c = c;
        }
    }
}

// Several points per toy, coherent-sum program (as k_dalitz_multi): workgroup
// (b, g) reads the waves of each row of work[b] once and adds the row's metric
// for every point p in [16 g, min(16 g + 16, npts)) of its toy.  A group of
// mx_min points or more is left to k_batch_coh_mx (launched over the same grid).
template <bool CONST_EFF, bool ZERO_OUT, int NR, bool NLL>
__global__ __launch_bounds__(GF_BLOCK) void k_batch_dalitz_points(GfProgArgs A, const GfBatchWork *__restrict__ work,
                                                                  const double *__restrict__ stage, int np, int nn,
                                                                  const int *__restrict__ tbl, int metric,
                                                                  int mx_min, Partial *partials) {
    __shared__ MultiSets S;
    const GfBatchWork w = work[blockIdx.x];
    const int slot0     = tbl[2 * w.toy];
    const int npts      = tbl[2 * w.toy + 1];
    const int p0        = (int)blockIdx.y * GF_MULTI_PTS;
    const int cnt       = min(npts - p0, GF_MULTI_PTS);
    if(cnt <= 0 || cnt >= mx_min)
        return;  // uniform over the workgroup
    const GfIns dzi       = A.ins[0];
    const GfIns effi      = A.ins[1];
    const GfDalitzSlot dz = gfd::as_const(A.dz)[dzi.x];
    const GfDecay &D      = A.decay;
    const int nres        = dzi.a;
    const int64_t W       = np + nn;
    const double *__restrict__ S0  = stage + (int64_t)(slot0 + p0) * W;
    const double *__restrict__ c12 = A.cols[dz.col_m12];
    const double *__restrict__ c13 = A.cols[dz.col_m13];
    stage_sets<CONST_EFF>(S, S0, W, S0 + np, W, cnt, dzi.p, effi.p + effi.a, nres);
    constexpr bool nll = NLL;  // the launcher sets it for metric == NLL
    double s[GF_MULTI_PTS], c[GF_MULTI_PTS];  // per point: (sum, residual), or the NLL's (mantissa, exponent)
#pragma unroll
    for(int p = 0; p < GF_MULTI_PTS; ++p) {
        s[p] = nll ? 1.0 : 0.0;
        c[p] = 0.0;
    }
    for(int64_t base = w.start; base < w.end; base += GF_BLOCK) {
        asm volatile("" ::: "memory");  // the set reads stay in the loop (as k_dalitz_multi)
        const int64_t e0 = base + threadIdx.x;
        const bool valid = e0 < w.end;
        const int64_t e  = valid ? e0 : w.start;
        double2 wv[NR];
#pragma unroll
        for(int i = 0; i < NR; ++i)
            wv[i] = (i < nres) ? dz.wave[(int64_t)i * dz.ncache + e] : make_double2(0.0, 0.0);
        const double m12  = ZERO_OUT ? 0.0 : c12[e];
        const double m13  = ZERO_OUT ? 0.0 : c13[e];
        const bool inside =
            ZERO_OUT || gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
#pragma unroll
        for(int p = 0; p < GF_MULTI_PTS; ++p) {
            if(p < cnt) {
                // a toy's points are compared with one another, never with the single
                // sweep, so the sum keeps the per-resonance order (fewer live registers)
                double re = 0.0, im = 0.0;
#pragma unroll
                for(int i = 0; i < NR; ++i) {
                    const double2 cp = S.c[p][i];
                    gfd::cmac(re, im, cp.x, cp.y, wv[i]);
                }
                const double eff = CONST_EFF ? S.eff[p] : multipoly_at(effi, A, gfd::as_const(S0 + p * W), e);
                const double T   = inside ? (re * re + im * im) * eff : 0.0;
                if(nll) {
                    if(valid)
                        gfd::nll_add(s[p], c[p], T, S.norm[p]);
                } else if(valid) {
                    gfd::kb_add(s[p], c[p], gfd::metric(metric, T, 0.0, 0.0, S.norm[p], w.total));
                }
            }
        }
    }
#pragma unroll
    for(int p = 0; p < GF_MULTI_PTS; ++p) {
        if(p < cnt) {
            double ss = s[p], cc = c[p];
            if(nll)
                ss = gfd::nll_total(ss, cc), cc = 0.0;
            block_reduce(ss, cc);
            if(threadIdx.x == 0) {
                partials[(int64_t)blockIdx.x * GF_BATCH_PTS + p0 + p].s = ss;
                partials[(int64_t)blockIdx.x * GF_BATCH_PTS + p0 + p].c = cc;
            }
        }
    }
}

// The groups of k_batch_dalitz_points with mx_min points or more, for the NLL of a
// lone Amp3Body with constant efficiency and waves 0 outside the Dalitz region: the
// coherent sums on the FP64 matrix cores (coh_mx_rows over the rows of the
// workgroup's chunk, as k_coh_mx runs them over the whole sample).  A smaller group
// (a line-search point, say) exits here and is taken by the vector kernel.  A
// point's value depends on the size of the group it came in, deterministically; the
// points of a toy meet only in its minimiser, which sees differences of an ulp.
template <int NR>
__global__ __launch_bounds__(GF_BLOCK) void k_batch_coh_mx(GfProgArgs A, const GfBatchWork *__restrict__ work,
                                                           const double *__restrict__ stage, int np, int nn,
                                                           const int *__restrict__ tbl, int mx_min,
                                                           Partial *partials) {
    const GfBatchWork w = work[blockIdx.x];
    const int npts      = tbl[2 * w.toy + 1];
    const int p0        = (int)blockIdx.y * GF_MULTI_PTS;
    const int cnt       = min(npts - p0, GF_MULTI_PTS);
    if(cnt < mx_min)
        return;  // uniform over the workgroup
    __shared__ MultiSets S[1];
    __shared__ double ls[GF_MULTI_PTS], lc[GF_MULTI_PTS];
    const GfIns dzi                = A.ins[0];
    const GfIns effi               = A.ins[1];
    const GfDalitzSlot dz          = gfd::as_const(A.dz)[dzi.x];
    const int nres                 = dzi.a;
    const int64_t W                = np + nn;
    const double *__restrict__ S0  = stage + (int64_t)(tbl[2 * w.toy] + p0) * W;
    stage_sets<true>(S[0], S0, W, S0 + np, W, cnt, dzi.p, effi.p + effi.a, nres);
    coh_mx_rows<NR, 1>(S, dz, nres, w.start, GF_BLOCK, w.end, cnt, ls, lc);
    if((int)threadIdx.x < cnt) {
        const int p                                             = threadIdx.x;
        partials[(int64_t)blockIdx.x * GF_BATCH_PTS + p0 + p].s = ls[p];
        partials[(int64_t)blockIdx.x * GF_BATCH_PTS + p0 + p].c = lc[p];
    }
}

// The normalisation arrays of staged batch points, one lane per point: the template,
// then norms[op.out] = 1 / I for each op with I from the point's parameters in the
// host's operation order (GooPdf::normalize: ret = 1 * integral, norm = 1 / ret;
// GaussianPdf.cu:28-35, ExpPdf.cu:121-132, Amp3Body.cu:334-358).
__global__ __launch_bounds__(256) void k_batch_norms(double *__restrict__ stage, int nslots, int np, int nn,
                                                     const double *__restrict__ tmpl,
                                                     const GfNormOp *__restrict__ ops, int nops,
                                                     const double *__restrict__ table) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= nslots)
        return;
    const double *__restrict__ P = stage + (int64_t)k * (np + nn);
    double *__restrict__ Nm      = stage + (int64_t)k * (np + nn) + np;
    for(int i = 0; i < nn; ++i)
        Nm[i] = tmpl[i];
    for(int o = 0; o < nops; ++o) {
        const GfNormOp op = ops[o];
        double integral;
        if(op.kind == GN_SCALE) {
            integral = P[op.p];
            integral *= op.c0;
        } else if(op.kind == GN_EXP) {
            const double a = P[op.p];
            if(0 == a) {
                integral = op.c1 - op.c0;
            } else {
                integral = exp(a * op.c1) - exp(a * op.c0);
                integral /= a;
            }
        } else {
            // sum_ij c_i conj(c_j) T_ij, each complex product written out
            const double *__restrict__ cp = P + op.p;
            const double *__restrict__ T  = table + op.g;
            double sum_re = 0;
            for(int i = 0; i < op.n; ++i) {
                const double ar = cp[2 * i], ai = cp[2 * i + 1];
                for(int j = 0; j < op.n; ++j) {
                    const double br = cp[2 * j], bi = -cp[2 * j + 1];
                    const double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
                    const double ir = T[2 * (i * op.n + j)], ii = T[2 * (i * op.n + j) + 1];
                    sum_re += pr * ir - pi * ii;
                }
            }
            integral = sum_re;
            integral *= op.c0;
        }
        double ret = 1;
        ret *= integral;
        Nm[op.out] = 1.0 / ret;
    }
}

// One workgroup per toy: for each of its points, the partials of the toy's work
// entries [first[t], first[t + 1]) in the order of k_finish_toys into
// result[2 (t GF_BATCH_PTS + p) ..] (host-mapped).
__global__ __launch_bounds__(GF_BLOCK) void k_finish_toy_points(const Partial *__restrict__ partials,
                                                                const int *__restrict__ first,
                                                                const int *__restrict__ tbl,
                                                                double *__restrict__ result) {
    const int t    = blockIdx.x;
    const int npts = tbl[2 * t + 1];
    const int f0 = first[t], f1 = first[t + 1];
    if(f1 - f0 <= 64) {  // uniform over the workgroup
        // Every partial of the toy lies in wave 0 of block_reduce's layout (thread i holds
        // entry i), and waves 1-3 hold exact (+0, +0) pairs.  So one wave per point gives
        // block_reduce's bits -- the same butterfly, then three zero pairs merged in
        // order -- and the four waves take the points in turn, without a barrier.
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for(int p = wave; p < npts; p += GF_BLOCK / 64) {
            double s = 0.0, c = 0.0;
            if(f0 + lane < f1)
                gfd::kb_merge(s, c, partials[(int64_t)(f0 + lane) * GF_BATCH_PTS + p].s,
                              partials[(int64_t)(f0 + lane) * GF_BATCH_PTS + p].c);
#pragma unroll
            for(int off = 32; off > 0; off >>= 1) {
                double s2 = __shfl_xor(s, off, 64);
                double c2 = __shfl_xor(c, off, 64);
                gfd::kb_merge(s, c, s2, c2);
            }
            if(lane == 0) {
                for(int w = 1; w < GF_BLOCK / 64; ++w)
                    gfd::kb_merge(s, c, 0.0, 0.0);
                result[2 * ((int64_t)t * GF_BATCH_PTS + p)]     = gfd::kb_total(s, c);
                result[2 * ((int64_t)t * GF_BATCH_PTS + p) + 1] = c;
            }
        }
        return;
    }
    for(int p = 0; p < npts; ++p) {
        double s = 0.0, c = 0.0;
        for(int i = first[t] + threadIdx.x; i < first[t + 1]; i += GF_BLOCK)
            gfd::kb_merge(s, c, partials[(int64_t)i * GF_BATCH_PTS + p].s, partials[(int64_t)i * GF_BATCH_PTS + p].c);
        block_reduce(s, c);
        if(threadIdx.x == 0) {
            result[2 * ((int64_t)t * GF_BATCH_PTS + p)]     = gfd::kb_total(s, c);
            result[2 * ((int64_t)t * GF_BATCH_PTS + p) + 1] = c;
        }
    }
}

// One workgroup per toy: the toy's partials [first[t], first[t + 1]) in order
// into result[2t] = sum, result[2t + 1] = residual (host-mapped).
__global__ __launch_bounds__(GF_BLOCK) void k_finish_toys(const Partial *__restrict__ partials,
                                                          const int *__restrict__ first,
                                                          const uint8_t *__restrict__ active,
                                                          double *__restrict__ result) {
    const int t = blockIdx.x;
    if(!active[t])
        return;
    double s = 0.0, c = 0.0;
    for(int i = first[t] + threadIdx.x; i < first[t + 1]; i += GF_BLOCK)
        gfd::kb_merge(s, c, partials[i].s, partials[i].c);
    block_reduce(s, c);
    if(threadIdx.x == 0) {
        result[2 * t]     = gfd::kb_total(s, c);
        result[2 * t + 1] = c;
    }
}

// Completion stamp after the toys' results: counter + 1 into the mapped word.
__global__ void k_stamp(uint32_t *__restrict__ counter, uint32_t *__restrict__ mapped) {
    const uint32_t v = *counter + 1u;
    *counter         = v;
    __threadfence_system();
    __hip_atomic_store(mapped, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------------------- event generation

// Candidates for accept-reject generation from an Amp3Body: point i of the
// counter-based stream is uniform in the box; its value is
// |sum_r c_r B_r(m12, m13)|^2 * eff (0 outside the Dalitz plot, as
// device_DalitzPlot, Amp3Body.cu:60-122, with lineshapes computed directly);
// accepted when u_i * fmax < value.  Per block: accepted count and the largest
// value (to check the envelope).
template <bool GEN_EFF>
__global__ __launch_bounds__(256) void k_dalitz_candidates(GfProgArgs A, const GfResDesc *__restrict__ res, int nres,
                                                           int p_coup, int eff_pc, double eff_const, int id12,
                                                           int id13, double l12,
                                                           double u12, double l13, double u13, uint64_t key,
                                                           int64_t base, int64_t nc, double fmax,
                                                           double *__restrict__ c12, double *__restrict__ c13,
                                                           uint8_t *__restrict__ flag, int *__restrict__ bcount,
                                                           double *__restrict__ bmax) {
    __shared__ int scount[4];
    __shared__ double smax[4];
    const int64_t i   = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid  = i < nc;
    const GfDecay D   = A.decay;
    const uint64_t ct = (uint64_t)(base + (valid ? i : 0));
    const double m12  = l12 + gfd::unif01(key, 2 * ct) * (u12 - l12);
    const double m13  = l13 + gfd::unif01(key, 2 * ct + 1) * (u13 - l13);
    double value      = 0.0;
    if(valid && gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass)) {
        const double m23 = D.motherMass * D.motherMass + D.daug1Mass * D.daug1Mass + D.daug2Mass * D.daug2Mass
                           + D.daug3Mass * D.daug3Mass - m12 - m13;
        double re = 0.0, im = 0.0;
#pragma unroll 1
        for(int r = 0; r < nres; ++r) {
            const double2 b = gfd::lineshape(res[r], A.params, A.consts, D, m12, m13, m23);
            const double ar = A.params[p_coup + 2 * r], ai = A.params[p_coup + 2 * r + 1];
            re += ar * b.x - ai * b.y;
            im += ar * b.y + ai * b.x;
        }
        double eff = eff_pc == -2 ? eff_const : 1.0;
        if constexpr(GEN_EFF) {
            RegSrc<1> ev;
#pragma unroll
            for(int o = 0; o < GF_MAX_OBS; ++o)
                ev.v[0][o] = 0.0;
            ev.set(id12, 0, m12);
            ev.set(id13, 0, m13);
            const int64_t r0[1] = {0};
            const bool v0[1]    = {true};
            double T[1];
            run<1>(A, A.params, A.norms, eff_pc, ev, r0, v0, T);
            eff = T[0];
        }
        value = (re * re + im * im) * eff;
    }
    const bool acc = valid && gfd::unif01(key ^ 0xD1B54A32D192ED03ull, ct) * fmax < value;
    if(valid) {
        c12[i]  = m12;
        c13[i]  = m13;
        flag[i] = acc ? 1 : 0;
    }
    // block totals: count of accepted, largest value
    const unsigned long long ball = __ballot(acc);
    double vm = value;
    for(int off = 32; off > 0; off >>= 1)
        vm = ::fmax(vm, __shfl_xor(vm, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if(lane == 0) {
        scount[wave] = __popcll(ball);
        smax[wave]   = vm;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        int c    = 0;
        double m = 0.0;
        for(int w = 0; w < 4; ++w) {
            c += scount[w];
            m = ::fmax(m, smax[w]);
        }
        bcount[blockIdx.x] = c;
        bmax[blockIdx.x]   = m;
    }
}

// Write the accepted candidates in order: block b's go to boff[b] + their rank
// in the block (waves in order, lanes in order); rows >= limit are dropped.
__global__ __launch_bounds__(256) void k_compact(const double *__restrict__ c12, const double *__restrict__ c13,
                                                 const uint8_t *__restrict__ flag, int64_t nc,
                                                 const int64_t *__restrict__ boff, int64_t limit,
                                                 double *__restrict__ o12, double *__restrict__ o13) {
    __shared__ int wcount[4];
    const int64_t i  = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool acc   = i < nc && flag[i];
    const unsigned long long ball = __ballot(acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if(lane == 0)
        wcount[wave] = __popcll(ball);
    __syncthreads();
    int before = 0;
    for(int w = 0; w < wave; ++w)
        before += wcount[w];
    const unsigned long long lower = lane ? (ball & ((1ull << lane) - 1ull)) : 0ull;
    const int64_t dst = boff[blockIdx.x] + before + __popcll(lower);
    if(acc && dst < limit) {
        o12[dst] = c12[i];
        o13[dst] = c13[i];
    }
}

// ------------------------------------------------------------- FP64-MFMA Gram


#define GG_BLOCK 128  // two waves
#define GG_LD 65      // padded LDS row (doubles): 16 lanes reading one column hit distinct banks

struct GramSrc {
    int mode;  // 0 grid, 1 points, 2 data cache
    double l12, u12, l13, u13;
    int n12, n13;
    const double *m12, *m13, *w;
    int64_t npoints;
    int slot;          // mode 2
    int id_m12, id_m13;
    int eff_pc;        // >= 0: efficiency subprogram; -1: efficiency 1; -2: constant eff_const
    double eff_const;
    int td;            // grid of a time-dependent node: m13 bin count n13 (SpecialDalitzIntegrator.cu:75-82)
};

// Centre (m12, m13) of integration cell p of a Dalitz grid (SpecialResonanceIntegrator.cu:61-76:
// m12 fastest; the m13 bin count is read from the m12 entry there, numBinsM13 = numBinsM12,
// reproduced here; a time-dependent node's integrator reads its own m13 count,
// SpecialDalitzIntegrator.cu:75-82).  The cell index is an int there, as here.
__device__ __forceinline__ void dalitz_cell(const GramSrc &S, int64_t p, double &m12, double &m13) {
    int globalBinNumber  = (int)p;
    const int numBinsM12 = S.n12;
    const int binM12     = globalBinNumber % numBinsM12;
    m12                  = gfd::bin_centre(S.l12, S.u12, numBinsM12, binM12);
    globalBinNumber /= numBinsM12;
    const int numBinsM13 = S.td ? S.n13 : S.n12;
    m13                  = gfd::bin_centre(S.l13, S.u13, numBinsM13, globalBinNumber);
}

// The centres of the first npoints cells of the grid in S (mode 0).
__global__ __launch_bounds__(256) void k_dalitz_cells(GramSrc S, double *m12, double *m13) {
    for(int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < S.npoints; p += (int64_t)gridDim.x * blockDim.x)
        dalitz_cell(S, p, m12[p], m13[p]);
}

// One block: 128 points per pass; each lane computes the 16 complex amplitudes of
// its point (wave-uniform control flow), stages them in LDS component-major,
// and each wave multiplies its 64 points as 16 MFMA k-steps for each of three
// 16x16 output tiles: RR = R^T E R, II = I^T E I, RI = R^T E I, with
// E = diag(eff * weight).  Partials: 3 tiles x 64 lanes x 4 values per block.
template <bool GEN_EFF>
__global__ __launch_bounds__(GG_BLOCK) void k_gram(GfProgArgs A, GramSrc S, const GfResDesc *__restrict__ res,
                                                   int nres, double *partials) {
    __shared__ double Xr[2][16][GG_LD];
    __shared__ double Xi[2][16][GG_LD];
    __shared__ double Ew[2][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GfDecay D = A.decay;
    gf_double4 RR = {0, 0, 0, 0}, II = {0, 0, 0, 0}, RI = {0, 0, 0, 0};
    const int64_t npts = S.npoints;
    for(int64_t base = (int64_t)blockIdx.x * GG_BLOCK; base < npts; base += (int64_t)gridDim.x * GG_BLOCK) {
        const int64_t p  = base + tid;
        const bool valid = p < npts;
        double m12 = 0.0, m13 = 0.0, wt = 1.0;
        if(S.mode == 0) {
            dalitz_cell(S, valid ? p : 0, m12, m13);
        } else if(S.mode == 1) {
            m12 = valid ? S.m12[p] : 0.0;
            m13 = valid ? S.m13[p] : 0.0;
            wt  = (valid && S.w) ? S.w[p] : 1.0;
        }
        double ew = 0.0;
        if(S.mode == 2) {
            const GfDalitzSlot dz = A.dz[S.slot];
            const int64_t q       = valid ? p : 0;
            for(int i = 0; i < GF_MAX_RES; ++i) {
                double2 b = (i < nres) ? dz.wave[(int64_t)i * dz.ncache + q] : make_double2(0.0, 0.0);
                Xr[wave][i][lane] = b.x;
                Xi[wave][i][lane] = b.y;
            }
            ew = valid ? 1.0 : 0.0;
        } else {
            const bool inside = valid && gfd::in_dalitz(m12, m13, D.motherMass, D.daug1Mass, D.daug2Mass, D.daug3Mass);
            const double m23  = D.motherMass * D.motherMass + D.daug1Mass * D.daug1Mass + D.daug2Mass * D.daug2Mass
                               + D.daug3Mass * D.daug3Mass - m12 - m13;
            // one lineshape body in the loop (not 16 inlined copies): registers for occupancy
#pragma unroll 1
            for(int i = 0; i < GF_MAX_RES; ++i) {
                double2 b = make_double2(0.0, 0.0);
                if(i < nres && inside) {
                    const bool sw = res[i].swap != 0;  // B wave of a time-dependent node
                    b = gfd::lineshape(res[i], A.params, A.consts, D, sw ? m13 : m12, sw ? m12 : m13, m23);
                }
                Xr[wave][i][lane] = b.x;
                Xi[wave][i][lane] = b.y;
            }
            double eff = S.eff_pc == -2 ? S.eff_const : 1.0;
            if constexpr(GEN_EFF) {
                // The fake event of SpecialResonanceIntegrator.cu:97-101: slot 0 = 2,
                // then m12 and m13 at their observable slots; other slots 0.
                RegSrc<1> ev;
#pragma unroll
                for(int o = 0; o < GF_MAX_OBS; ++o)
                    ev.v[0][o] = 0.0;
                ev.v[0][0] = 2.0;
                ev.set(S.id_m12, 0, m12);
                ev.set(S.id_m13, 0, m13);
                const int64_t r0[1] = {0};
                const bool v0[1]    = {true};
                double T[1];
                run<1>(A, A.params, A.norms, S.eff_pc, ev, r0, v0, T);
                eff = T[0];
            }
            ew = valid ? eff * wt : 0.0;
        }
        Ew[wave][lane] = ew;
        __syncthreads();
        // A operand lane l: A[i = l%16][k = l/16]; B operand lane l: B[k = l/16][j = l%16].
        const int comp = lane & 15, sub = lane >> 4;
#pragma unroll
        for(int kk = 0; kk < 16; ++kk) {
            const int pt    = 4 * kk + sub;
            const double ar = Xr[wave][comp][pt];
            const double ai = Xi[wave][comp][pt];
            const double e  = Ew[wave][pt];
            const double br = e * ar, bi = e * ai;
            RR = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, RR, 0, 0, 0);
            II = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, II, 0, 0, 0);
            RI = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, RI, 0, 0, 0);
        }
        __syncthreads();
    }
    // add wave 1 into wave 0 through LDS, then write this block's partial
    double *stage = &Xr[0][0][0];  // 2*16*65 doubles >= 64*12
    if(wave == 1) {
#pragma unroll
        for(int r = 0; r < 4; ++r) {
            stage[(0 * 4 + r) * 64 + lane] = RR[r];
            stage[(1 * 4 + r) * 64 + lane] = II[r];
            stage[(2 * 4 + r) * 64 + lane] = RI[r];
        }
    }
    __syncthreads();
    if(wave == 0) {
        double *out = partials + (int64_t)blockIdx.x * 768;
#pragma unroll
        for(int r = 0; r < 4; ++r) {
            out[(0 * 4 + r) * 64 + lane] = RR[r] + stage[(0 * 4 + r) * 64 + lane];
            out[(1 * 4 + r) * 64 + lane] = II[r] + stage[(1 * 4 + r) * 64 + lane];
            out[(2 * 4 + r) * 64 + lane] = RI[r] + stage[(2 * 4 + r) * 64 + lane];
        }
    }
}

// Sum the per-block Gram partials: one block per output element, each thread a
// fixed stride of blocks, then a fixed-order tree (deterministic for a given grid).
__global__ __launch_bounds__(256) void k_gram_final(const double *__restrict__ partials, int nblocks, double *out) {
    const int e = blockIdx.x;  // 0..767
    double s = 0.0, c = 0.0;
    for(int b = threadIdx.x; b < nblocks; b += blockDim.x)
        gfd::kb_add(s, c, partials[(int64_t)b * 768 + e]);
    block_reduce(s, c);
    if(threadIdx.x == 0)
        out[e] = gfd::kb_total(s, c);
}

// One v_mfma_f64_16x16x4f64 on given per-lane operands: out[4*lane + r] is
// D-register r of the lane.  Used to establish the operand/result lane layout.
__global__ __launch_bounds__(64) void k_mfma_probe(const double *__restrict__ a, const double *__restrict__ b,
                                                   double *out) {
    const int l    = threadIdx.x;
    gf_double4 acc = {0, 0, 0, 0};
    acc            = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
    out[4 * l + 0] = acc[0];
    out[4 * l + 1] = acc[1];
    out[4 * l + 2] = acc[2];
    out[4 * l + 3] = acc[3];
}

// flag stays 1 only if floor(0.5 + col[i]) == i for every row (Amp3Body.cu:82-84).
__global__ __launch_bounds__(256) void k_identity_check(const double *__restrict__ col, int64_t n, int32_t *flag) {
    bool ok = true;
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        ok = ok && ((int64_t)floor(0.5 + col[i]) == i);
    if(!ok)
        atomicAnd(flag, 0);
}

inline int check(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

}  // namespace

// ------------------------------------------------------------------ launchers

namespace gfk {

// Add the nblocks partials of the launch before it on stream s into *result.
// A non-null a.mark event is recorded between the two launches (timing).
static int launch_finish(const GfProgArgs &a, Partial *partials, int nblocks, GfResult *result, hipStream_t s) {
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    if(a.mark && (rc = check(hipEventRecord((hipEvent_t)a.mark, s))))
        return rc;
    if(a.sweep_only || a.fuse_slots)  // a fused sweep has finished itself
        return 0;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(GF_BLOCK), 0, s, partials, nblocks, result, a.err, a.seq);
    return check(hipGetLastError());
}

// Node kinds a program needs compiled in (FEAT of the interpreter kernels).
static int feature_level(const GfProgArgs &a) { return a.td ? 2 : (a.has_dz ? 1 : 0); }

int launch_reduce_metric(const GfProgArgs &a, const GfInline *inl, int metric, Partial *partials, GfResult *result,
                         int nblocks, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if(a.fuse_slots)  // the interpreter has no in-kernel finish
        return -1;
    // time-dependent nodes: the kernels compiled with the TD ops (two events per lane
    // measured 1.5 % faster than one on the 16-resonance three-Gaussian workload);
    // programs without Dalitz nodes: the kernels compiled without them
    const dim3 g(nblocks), b(GF_BLOCK);
    const GfInline none{};
    switch(feature_level(a) * 2 + (inl ? 1 : 0)) {
    case 0: hipLaunchKernelGGL((k_reduce_metric<2, false, 0>), g, b, 0, s, a, none, metric, partials); break;
    case 1: hipLaunchKernelGGL((k_reduce_metric<2, true, 0>), g, b, 0, s, a, *inl, metric, partials); break;
    case 2: hipLaunchKernelGGL((k_reduce_metric<2, false, 1>), g, b, 0, s, a, none, metric, partials); break;
    case 3: hipLaunchKernelGGL((k_reduce_metric<2, true, 1>), g, b, 0, s, a, *inl, metric, partials); break;
    case 4: hipLaunchKernelGGL((k_reduce_metric<2, false, 2>), g, b, 0, s, a, none, metric, partials); break;
    default: hipLaunchKernelGGL((k_reduce_metric<2, true, 2>), g, b, 0, s, a, *inl, metric, partials); break;
    }
    return launch_finish(a, partials, nblocks, result, s);
}
int launch_reduce_flat(const GfProgArgs &a, const GfInline *inl, int metric, int ncols, int ept, Partial *partials,
                       GfResult *result, int nblocks, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if(ncols < 1 || ncols > GF_FLAT_COLS || metric >= GF_METRIC_BINAVG_CODE)
        return -1;
    const dim3 g(nblocks), b(GF_BLOCK);
    const GfInline none{};
    switch((ept == 4 ? 2 : 0) + (inl ? 1 : 0)) {
    case 0: hipLaunchKernelGGL((k_reduce_flat<2, false>), g, b, 0, s, a, none, metric, ncols, partials); break;
    case 1: hipLaunchKernelGGL((k_reduce_flat<2, true>), g, b, 0, s, a, *inl, metric, ncols, partials); break;
    case 2: hipLaunchKernelGGL((k_reduce_flat<4, false>), g, b, 0, s, a, none, metric, ncols, partials); break;
    default: hipLaunchKernelGGL((k_reduce_flat<4, true>), g, b, 0, s, a, *inl, metric, ncols, partials); break;
    }
    return launch_finish(a, partials, nblocks, result, s);
}

int launch_eval_metric(const GfProgArgs &a, const GfInline *inl, int metric, double *out, int nblocks, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const dim3 g(nblocks), b(GF_BLOCK);
    const GfInline none{};
    switch(feature_level(a) * 2 + (inl ? 1 : 0)) {
    case 0: hipLaunchKernelGGL((k_eval_metric<2, false, 0>), g, b, 0, s, a, none, metric, out); break;
    case 1: hipLaunchKernelGGL((k_eval_metric<2, true, 0>), g, b, 0, s, a, *inl, metric, out); break;
    case 2: hipLaunchKernelGGL((k_eval_metric<2, false, 1>), g, b, 0, s, a, none, metric, out); break;
    case 3: hipLaunchKernelGGL((k_eval_metric<2, true, 1>), g, b, 0, s, a, *inl, metric, out); break;
    case 4: hipLaunchKernelGGL((k_eval_metric<2, false, 2>), g, b, 0, s, a, none, metric, out); break;
    default: hipLaunchKernelGGL((k_eval_metric<2, true, 2>), g, b, 0, s, a, *inl, metric, out); break;
    }
    return check(hipGetLastError());
}

template <bool WRITE, bool INL, bool CE, bool ZO = false>
static void launch_fast_t(const GfProgArgs &a, const GfInline *inl, int metric, Partial *partials, double *out,
                          int nblocks, hipStream_t s) {
    hipLaunchKernelGGL((k_dalitz_fast<WRITE, INL, CE, ZO>), dim3(nblocks), dim3(256), 0, s, a,
                       inl ? *inl : GfInline{}, metric, partials, out);
}

// The multi-set kernels for nres resonances: the resonance loop unrolled to the
// next multiple of 4 (NR), the variant by efficiency and zero-outside use.
#define GF_MULTI_CASES(KERNEL, NLLB, ...)                                                                           \
    switch(v_ * 4 + nr_ / 4 - 1) {                                                                                   \
    case 0: hipLaunchKernelGGL((KERNEL<false, false, 4, NLLB>), __VA_ARGS__); break;                                \
    case 1: hipLaunchKernelGGL((KERNEL<false, false, 8, NLLB>), __VA_ARGS__); break;                                \
    case 2: hipLaunchKernelGGL((KERNEL<false, false, 12, NLLB>), __VA_ARGS__); break;                               \
    case 3: hipLaunchKernelGGL((KERNEL<false, false, 16, NLLB>), __VA_ARGS__); break;                               \
    case 4: hipLaunchKernelGGL((KERNEL<true, false, 4, NLLB>), __VA_ARGS__); break;                                 \
    case 5: hipLaunchKernelGGL((KERNEL<true, false, 8, NLLB>), __VA_ARGS__); break;                                 \
    case 6: hipLaunchKernelGGL((KERNEL<true, false, 12, NLLB>), __VA_ARGS__); break;                                \
    case 7: hipLaunchKernelGGL((KERNEL<true, false, 16, NLLB>), __VA_ARGS__); break;                                \
    case 8: hipLaunchKernelGGL((KERNEL<true, true, 4, NLLB>), __VA_ARGS__); break;                                  \
    case 9: hipLaunchKernelGGL((KERNEL<true, true, 8, NLLB>), __VA_ARGS__); break;                                  \
    case 10: hipLaunchKernelGGL((KERNEL<true, true, 12, NLLB>), __VA_ARGS__); break;                                \
    default: hipLaunchKernelGGL((KERNEL<true, true, 16, NLLB>), __VA_ARGS__); break;                                \
    }
// The variant for nres resonances (loop unrolled to the next multiple of 4), the
// efficiency / zero-outside use, and whether the metric is the NLL (its own
// accumulation, compiled apart so that each variant holds only its own path).
#define GF_MULTI_SWITCH(KERNEL, NRES, CE, ZO, NLL, ...)                                                             \
    do {                                                                                                             \
        const int nr_ = (NRES) <= 4 ? 4 : (NRES) <= 8 ? 8 : (NRES) <= 12 ? 12 : 16;                                  \
        const int v_  = (CE) ? ((ZO) ? 2 : 1) : 0;                                                                   \
        if(NLL) {                                                                                                    \
            GF_MULTI_CASES(KERNEL, true, __VA_ARGS__)                                                                \
        } else {                                                                                                     \
            GF_MULTI_CASES(KERNEL, false, __VA_ARGS__)                                                               \
        }                                                                                                            \
    } while(0)

// k_batch_coh_mx for nres resonances (NR = nres rounded up to 4).
#define GF_COH_MX(KERNEL, NRES, ...)                                                                                 \
    do {                                                                                                             \
        switch((NRES) <= 4 ? 4 : (NRES) <= 8 ? 8 : (NRES) <= 12 ? 12 : 16) {                                         \
        case 4: hipLaunchKernelGGL((KERNEL<4>), __VA_ARGS__); break;                                                \
        case 8: hipLaunchKernelGGL((KERNEL<8>), __VA_ARGS__); break;                                                \
        case 12: hipLaunchKernelGGL((KERNEL<12>), __VA_ARGS__); break;                                              \
        default: hipLaunchKernelGGL((KERNEL<16>), __VA_ARGS__); break;                                              \
        }                                                                                                            \
    } while(0)

// k_coh_mx<NR, NSG> for nres resonances (NR = nres rounded up to 4).
template <int NSG>
static void launch_coh_mx(int nres, int nblocks, hipStream_t s, const GfProgArgs &a, const double *params, int np,
                          const double *norms, int nn, int npts, Partial *partials) {
    const dim3 grid(nblocks), block(256);
    switch(nres <= 4 ? 4 : nres <= 8 ? 8 : nres <= 12 ? 12 : 16) {
    case 4: hipLaunchKernelGGL((k_coh_mx<4, NSG>), grid, block, 0, s, a, params, np, norms, nn, npts, partials); break;
    case 8: hipLaunchKernelGGL((k_coh_mx<8, NSG>), grid, block, 0, s, a, params, np, norms, nn, npts, partials); break;
    case 12:
        hipLaunchKernelGGL((k_coh_mx<12, NSG>), grid, block, 0, s, a, params, np, norms, nn, npts, partials);
        break;
    default:
        hipLaunchKernelGGL((k_coh_mx<16, NSG>), grid, block, 0, s, a, params, np, norms, nn, npts, partials);
        break;
    }
}

int launch_dalitz_multi(const GfProgArgs &a, int nres, const double *params, int np, const double *norms, int nn,
                        int npts, bool const_eff, bool zero_outside, int metric, Partial *partials, double *out,
                        int nblocks, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const bool mx = a.coh_mx_nres > 0 && const_eff && zero_outside && metric == gfd::metric_nll;
    if(npts < 1 || npts > (mx ? GF_MX_SETS : GF_MULTI_PTS) || nres < 1 || nres > GF_MAX_RES)
        return -1;
    if(mx) {
        if(npts > GF_MULTI_PTS)
            launch_coh_mx<2>(nres, nblocks, s, a, params, np, norms, nn, npts, partials);
        else
            launch_coh_mx<1>(nres, nblocks, s, a, params, np, norms, nn, npts, partials);
        int rc = check(hipGetLastError());
        if(rc)
            return rc;
        hipLaunchKernelGGL(k_finish_multi, dim3(npts), dim3(GF_BLOCK), 0, s, partials, nblocks, out);
        return check(hipGetLastError());
    }
    GF_MULTI_SWITCH(k_dalitz_multi, nres, const_eff, zero_outside && const_eff, metric == gfd::metric_nll, dim3(nblocks), dim3(256), 0, s, a,
                    params, np, norms, nn, npts, metric, partials);
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    hipLaunchKernelGGL(k_finish_multi, dim3(npts), dim3(GF_BLOCK), 0, s, partials, nblocks, out);
    return check(hipGetLastError());
}

int launch_dalitz_fast(const GfProgArgs &a, const GfInline *inl, bool const_eff, bool zero_outside, int metric,
                       Partial *partials, GfResult *result, double *out, int nblocks, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const int v   = (out ? 4 : 0) | (inl ? 2 : 0) | (const_eff ? 1 : 0);
    if(zero_outside && const_eff) {
        switch(v) {
        case 1: launch_fast_t<false, false, true, true>(a, inl, metric, partials, out, nblocks, s); break;
        case 3: launch_fast_t<false, true, true, true>(a, inl, metric, partials, out, nblocks, s); break;
        case 5: launch_fast_t<true, false, true, true>(a, inl, metric, partials, out, nblocks, s); break;
        default: launch_fast_t<true, true, true, true>(a, inl, metric, partials, out, nblocks, s); break;
        }
        if(out)
            return check(hipGetLastError());
        return launch_finish(a, partials, nblocks, result, s);
    }
    switch(v) {
    case 0: launch_fast_t<false, false, false>(a, inl, metric, partials, out, nblocks, s); break;
    case 1: launch_fast_t<false, false, true>(a, inl, metric, partials, out, nblocks, s); break;
    case 2: launch_fast_t<false, true, false>(a, inl, metric, partials, out, nblocks, s); break;
    case 3: launch_fast_t<false, true, true>(a, inl, metric, partials, out, nblocks, s); break;
    case 4: launch_fast_t<true, false, false>(a, inl, metric, partials, out, nblocks, s); break;
    case 5: launch_fast_t<true, false, true>(a, inl, metric, partials, out, nblocks, s); break;
    case 6: launch_fast_t<true, true, false>(a, inl, metric, partials, out, nblocks, s); break;
    default: launch_fast_t<true, true, true>(a, inl, metric, partials, out, nblocks, s); break;
    }
    if(out)
        return check(hipGetLastError());
    return launch_finish(a, partials, nblocks, result, s);
}

static GfGrid make_grid(const double *ranges, const int *slots, int k) {
    GfGrid G;
    for(int d = 0; d < GF_MAX_OBS; ++d) {
        G.lo[d]   = d < k ? ranges[3 * d + 0] : 0.0;
        G.hi[d]   = d < k ? ranges[3 * d + 1] : 1.0;
        G.nb[d]   = d < k ? (int)floor(ranges[3 * d + 2] + 0.5) : 1;
        G.slot[d] = (d < k && slots) ? slots[d] : 0;
    }
    G.k = k;
    return G;
}

int launch_reduce_bins(const GfProgArgs &a, int pc0, const double *ranges, const int *slots, int k, int64_t total,
                       Partial *partials, GfResult *result, int nblocks, void *stream) {
    const GfGrid G = make_grid(ranges, slots, k);
    hipStream_t s  = (hipStream_t)stream;
    if(a.fuse_slots)  // no in-kernel finish here
        return -1;
    hipLaunchKernelGGL(k_reduce_bins<2>, dim3(nblocks), dim3(GF_BLOCK), 0, s, a, pc0, G, total, partials);
    return launch_finish(a, partials, nblocks, result, s);
}

int launch_eval_bins(const GfProgArgs &a, int pc0, const double *ranges, const int *slots, int k, int64_t total,
                     double *out, int nblocks, void *stream) {
    const GfGrid G = make_grid(ranges, slots, k);
    hipLaunchKernelGGL(k_eval_bins<2>, dim3(nblocks), dim3(GF_BLOCK), 0, (hipStream_t)stream, a, pc0, G, total, out);
    return check(hipGetLastError());
}

static unsigned cell_blocks(int64_t total) {
    int64_t nb = (total + 255) / 256;
    return (unsigned)(nb > 8192 ? 8192 : (nb < 1 ? 1 : nb));
}

int launch_grid_cells(const double *ranges, int k, int64_t total, double *out, void *stream) {
    const GfGrid G = make_grid(ranges, nullptr, k);
    hipLaunchKernelGGL(k_grid_cells, dim3(cell_blocks(total)), dim3(256), 0, (hipStream_t)stream, G, total, out);
    return check(hipGetLastError());
}

int launch_fuse_init(unsigned long long *slots, void *stream) {
    hipLaunchKernelGGL(k_fuse_init, dim3(2 * GF_MAX_PARTIALS / 256), dim3(256), 0, (hipStream_t)stream, slots);
    return check(hipGetLastError());
}

int launch_batch(const GfProgArgs &a, const GfBatchWork *work, int nblocks, const double *params, int np,
                 const double *norms, int nn, const uint8_t *active, int metric, Partial *partials, const int *first,
                 int ntoys, double *result, uint32_t *counter, uint32_t *stamp, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    switch(feature_level(a)) {
    case 0:
        hipLaunchKernelGGL((k_batch_metric<2, 0>), dim3(nblocks), dim3(GF_BLOCK), 0, s, a, work, params, np, norms, nn,
                           active, metric, partials);
        break;
    case 1:
        hipLaunchKernelGGL((k_batch_metric<2, 1>), dim3(nblocks), dim3(GF_BLOCK), 0, s, a, work, params, np, norms, nn,
                           active, metric, partials);
        break;
    default:
        hipLaunchKernelGGL((k_batch_metric<2, 2>), dim3(nblocks), dim3(GF_BLOCK), 0, s, a, work, params, np, norms, nn,
                           active, metric, partials);
        break;
    }
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    hipLaunchKernelGGL(k_finish_toys, dim3(ntoys), dim3(GF_BLOCK), 0, s, partials, first, active, result);
    if((rc = check(hipGetLastError())))
        return rc;
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, counter, stamp);
    return check(hipGetLastError());
}

int launch_batch_norms(double *stage, int nslots, int np, int nn, const double *tmpl, const GfNormOp *ops, int nops,
                       const double *table, void *stream) {
    if(nslots <= 0)
        return 0;
    hipLaunchKernelGGL(k_batch_norms, dim3((nslots + 255) / 256), dim3(256), 0, (hipStream_t)stream, stage, nslots, np,
                       nn, tmpl, ops, nops, table);
    return check(hipGetLastError());
}

int launch_batch_points(const GfProgArgs &a, int nres, const GfBatchWork *work, int nblocks, int groups,
                        const double *stage, int np, int nn, const int *tbl, bool fast, bool const_eff,
                        bool zero_outside, int metric, int mx_min, Partial *partials, const int *first, int ntoys,
                        double *result, uint32_t *counter, uint32_t *stamp, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if(groups < 1 || groups * GF_MULTI_PTS > GF_BATCH_PTS || (fast && (nres < 1 || nres > GF_MAX_RES)))
        return -1;
    const dim3 grid(nblocks, groups);
    if(!fast && feature_level(a) == 2)
        hipLaunchKernelGGL((k_batch_points<2, 2>), grid, dim3(GF_BLOCK), 0, s, a, work, stage, np, nn, tbl, metric,
                           partials);
    else if(!fast && feature_level(a) == 1)
        hipLaunchKernelGGL((k_batch_points<2, 1>), grid, dim3(GF_BLOCK), 0, s, a, work, stage, np, nn, tbl, metric,
                           partials);
    else if(!fast)
        hipLaunchKernelGGL((k_batch_points<2, 0>), grid, dim3(GF_BLOCK), 0, s, a, work, stage, np, nn, tbl, metric,
                           partials);
    else {
        // groups of mx_min points or more on the matrix cores, the others on the vector ALUs
        const bool mx = mx_min > 0 && const_eff && zero_outside && metric == gfd::metric_nll;
        if(mx) {
            GF_COH_MX(k_batch_coh_mx, nres, grid, dim3(GF_BLOCK), 0, s, a, work, stage, np, nn, tbl, mx_min,
                      partials);
            int rc = check(hipGetLastError());
            if(rc)
                return rc;
        }
        GF_MULTI_SWITCH(k_batch_dalitz_points, nres, const_eff, zero_outside && const_eff, metric == gfd::metric_nll, grid, dim3(GF_BLOCK), 0, s,
                        a, work, stage, np, nn, tbl, metric, mx ? mx_min : GF_MULTI_PTS + 1, partials);
    }
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    hipLaunchKernelGGL(k_finish_toy_points, dim3(ntoys), dim3(GF_BLOCK), 0, s, partials, first, tbl, result);
    if((rc = check(hipGetLastError())))
        return rc;
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, counter, stamp);
    return check(hipGetLastError());
}

int launch_dalitz_candidates(const GfProgArgs &a, const GfResDesc *res, int nres, int p_coup, int eff_pc,
                             double eff_const, int id12, int id13, const double box[4], uint64_t key, int64_t base,
                             int64_t nc, double fmax, double *c12, double *c13, uint8_t *flag, int *bcount,
                             double *bmax, void *stream) {
    const int nb = (int)((nc + 255) / 256);
    if(eff_pc >= 0)
        hipLaunchKernelGGL(k_dalitz_candidates<true>, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, res, nres,
                           p_coup, eff_pc, eff_const, id12, id13, box[0], box[1], box[2], box[3], key, base, nc, fmax,
                           c12, c13, flag, bcount, bmax);
    else
        hipLaunchKernelGGL(k_dalitz_candidates<false>, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, res, nres,
                           p_coup, eff_pc, eff_const, id12, id13, box[0], box[1], box[2], box[3], key, base, nc, fmax,
                           c12, c13, flag, bcount, bmax);
    return check(hipGetLastError());
}

int launch_compact(const double *c12, const double *c13, const uint8_t *flag, int64_t nc, const int64_t *boff,
                   int64_t limit, double *o12, double *o13, void *stream) {
    const int nb = (int)((nc + 255) / 256);
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(256), 0, (hipStream_t)stream, c12, c13, flag, nc, boff, limit, o12,
                       o13);
    return check(hipGetLastError());
}

int launch_dalitz_cache(const GfProgArgs &a, int slot, int64_t offset, int64_t n, uint32_t mask, const GfResDesc *res,
                        int nres, int nblocks, void *stream) {
    hipLaunchKernelGGL(k_dalitz_cache, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, a, slot, offset, n, mask, res,
                       nres);
    return check(hipGetLastError());
}

int launch_gram(const GfProgArgs &a, int mode, const double *grid6, const double *m12, const double *m13,
                const double *w, int64_t npoints, int slot, int id_m12, int id_m13, int eff_pc, double eff_const,
                int td_grid, const GfResDesc *res, int nres, double *partials, double *tiles_dev, int nblocks,
                void *stream) {
    GramSrc S;
    S.mode = mode;
    S.td   = td_grid;
    S.l12 = S.u12 = S.l13 = S.u13 = 0.0;
    S.n12 = S.n13 = 1;
    if(grid6) {
        S.l12 = grid6[0];
        S.u12 = grid6[1];
        S.n12 = (int)floor(grid6[2] + 0.5);
        S.l13 = grid6[3];
        S.u13 = grid6[4];
        S.n13 = (int)floor(grid6[5] + 0.5);
    }
    S.m12     = m12;
    S.m13     = m13;
    S.w       = w;
    S.npoints = npoints;
    S.slot    = slot;
    S.id_m12  = id_m12;
    S.id_m13  = id_m13;
    S.eff_pc    = eff_pc;
    S.eff_const = eff_const;
    hipStream_t s = (hipStream_t)stream;
    if(eff_pc >= 0)
        hipLaunchKernelGGL(k_gram<true>, dim3(nblocks), dim3(GG_BLOCK), 0, s, a, S, res, nres, partials);
    else
        hipLaunchKernelGGL(k_gram<false>, dim3(nblocks), dim3(GG_BLOCK), 0, s, a, S, res, nres, partials);
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    hipLaunchKernelGGL(k_gram_final, dim3(768), dim3(256), 0, s, partials, nblocks, tiles_dev);
    return check(hipGetLastError());
}

int launch_dalitz_cells(const double *grid6, int td_grid, int64_t ncells, double *m12, double *m13, void *stream) {
    GramSrc S{};
    S.mode    = 0;
    S.td      = td_grid;
    S.l12     = grid6[0];
    S.u12     = grid6[1];
    S.n12     = (int)floor(grid6[2] + 0.5);
    S.l13     = grid6[3];
    S.u13     = grid6[4];
    S.n13     = (int)floor(grid6[5] + 0.5);
    S.npoints = ncells;
    hipLaunchKernelGGL(k_dalitz_cells, dim3(cell_blocks(ncells)), dim3(256), 0, (hipStream_t)stream, S, m12, m13);
    return check(hipGetLastError());
}

int launch_identity_check(const double *col, int64_t n, int32_t *flag, void *stream) {
    int64_t nb = (n + 255) / 256;
    if(nb > 2048)
        nb = 2048;
    if(nb < 1)
        nb = 1;
    hipLaunchKernelGGL(k_identity_check, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, col, n, flag);
    return check(hipGetLastError());
}

int mfma_probe(const double *dev_a, const double *dev_b, double *dev_out, void *stream) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, dev_a, dev_b, dev_out);
    return check(hipGetLastError());
}

}  // namespace gfk

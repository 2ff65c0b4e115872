// BatchFitter: many MIGRAD fits whose FCN calls are coalesced into batched GPU
// reductions (gf_batch_* of the C ABI).
#include <goofit/Engine.h>
#include <goofit/Error.h>
#include <goofit/PDFs/GooPdf.h>
#include <goofit/PDFs/physics/resonances/Resonance.h>
#include <goofit/fitting/BatchFitter.h>
#include <goofit/fitting/Params.h>

#include <Minuit2/FCNBase.h>
#include <Minuit2/FunctionMinimum.h>
#include <Minuit2/MnMigrad.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

namespace GooFit {

namespace {

// Shared state of one fit(): pending FCN requests and their answers.
struct Batcher {
    std::mutex mu;
    std::condition_variable cv_req;
    std::unique_ptr<std::condition_variable[]> cv_toy;  // one per toy: a delivery wakes only its toys
    std::vector<const std::vector<std::vector<double>> *> x;  // per toy: the requested Minuit vectors
    std::vector<size_t> x0, xn;                               // of those, the points of this request
    std::vector<uint8_t> pending, done;
    std::vector<std::vector<double>> result;
    int running    = 0;  // toys whose MIGRAD is in progress
    int waiting    = 0;  // of those, blocked in an FCN call
    int unprepared = 0;  // of those, whose request the batcher has not taken yet
    bool stop   = false;
};

// A toy's FCN: each call, or each group of independent points (a derivative
// round, FCNBase::EvaluateMany), becomes one request to the batcher; with
// `multi` false a group is served one point per request.
class ToyFCN : public Minuit2::FCNBase {
  public:
    ToyFCN(Batcher &b, int toy, bool multi)
        : b_(b)
        , toy_(toy)
        , multi_(multi) {}
    double operator()(const std::vector<double> &x) const override {
        std::vector<std::vector<double>> one{x};
        return request(one, 0, 1)[0];
    }
    std::vector<double> EvaluateMany(const std::vector<std::vector<double>> &xs) const override {
        std::vector<double> f;
        f.reserve(xs.size());
        const size_t step = multi_ ? (size_t)GF_BATCH_MAX_POINTS : 1;
        for(size_t k = 0; k < xs.size(); k += step) {
            const std::vector<double> r = request(xs, k, std::min(step, xs.size() - k));
            f.insert(f.end(), r.begin(), r.end());
        }
        return f;
    }
    double Up() const override { return 0.5; }

  private:
    std::vector<double> request(const std::vector<std::vector<double>> &xs, size_t k0, size_t n) const {
        std::unique_lock<std::mutex> lk(b_.mu);
        b_.x[toy_]       = &xs;
        b_.x0[toy_]      = k0;
        b_.xn[toy_]      = n;
        b_.done[toy_]    = 0;
        b_.pending[toy_] = 1;
        b_.waiting++;
        b_.unprepared++;
        b_.cv_req.notify_one();
        b_.cv_toy[toy_].wait(lk, [&] { return b_.done[toy_] != 0; });
        return b_.result[toy_];
    }
    Batcher &b_;
    int toy_;
    bool multi_;
};

}  // namespace

// A resonance's mass and width set the per-event lineshape cache that the rows of
// every toy share: a toy moving one would rebuild the cache under the other toys'
// sweeps, so a batched fit refuses them floating.
static void refuse_floating_lineshapes(PdfBase *p) {
    if(dynamic_cast<ResonancePdf *>(p) != nullptr)
        for(const Variable &v : p->getParameters())
            if(!v.IsFixed())
                throw GeneralError("BatchFitter: resonance parameter " + v.getName()
                                   + " floats; its lineshape cache is shared by all toys, so it must be fixed");
    for(PdfBase *c : p->getComponents())
        refuse_floating_lineshapes(c);
}

BatchFitter::BatchFitter(GooPdf *pdf, const std::vector<int64_t> &first, const std::vector<int64_t> &count)
    : pdf_(pdf)
    , first_(first)
    , count_(count) {
    if(!pdf || first.size() != count.size() || first.empty())
        throw GeneralError("BatchFitter: need a PDF and one (first, count) per toy");
    engine::flush();
    engine::check(gf_batch_create(engine::ctx(), (int)first.size(), first.data(), count.data(), &batch_),
                  "gf_batch_create");
}

BatchFitter::~BatchFitter() { gf_batch_destroy(batch_); }

void BatchFitter::useGraph(bool on) { engine::check(gf_batch_use_graph(batch_, on ? 1 : 0), "gf_batch_use_graph"); }

std::vector<double> BatchFitter::evaluate() {
    pdf_->normalize();
    engine::flush();
    std::vector<uint8_t> all(size(), 1);
    std::vector<double> out(size());
    for(size_t t = 0; t < size(); t++)
        engine::check(gf_batch_set_toy(batch_, (int)t, host_parameters.data(), host_normalizations.data()),
                      "gf_batch_set_toy");
    engine::check(gf_batch_reduce(batch_, all.data(), GF_METRIC_NLL, out.data()), "gf_batch_reduce");
    // calculateNLL's extended term (AddPdf.cu:187-204) with each toy's own event count
    for(size_t t = 0; t < size(); t++)
        out[t] += pdf_->nllExtraTermFor((fptype)count_[t]);
    return out;
}

void BatchFitter::fit(int threads, unsigned int strategy, double tolerance, bool autotune) {
    refuse_floating_lineshapes(pdf_);
    const int T = (int)size();
    Params params(*pdf_);
    std::vector<Variable> vars = params.GetGooFitParams();
    const std::vector<double> start = params.make_minuit_vector();
    // updateParameters() as a flat copy: the (Variable, slot) pairs of the tree
    std::vector<std::pair<Variable, unsigned int>> slots;
    pdf_->parameterSlots(slots);
    const auto t0 = std::chrono::steady_clock::now();
    stats_        = BatchFitStats{};
    pdf_->normalize();
    engine::flush();
    if(autotune) {
        for(int t = 0; t < T; t++)
            engine::check(gf_batch_set_toy(batch_, t, host_parameters.data(), host_normalizations.data()),
                          "gf_batch_set_toy");
        engine::check(gf_batch_autotune(batch_, GF_METRIC_NLL, &stats_.chunk, nullptr), "gf_batch_autotune");
    }
    // the points' normalisations inside the launches, when the model has device ops for
    // every normalisation the floating parameters move (the template: this normalize())
    std::vector<gf_norm_op> nops;
    std::vector<double> ntable;
    const bool dev_norm = multi_ && device_norms_ && pdf_->batchNormOps(nops, ntable);
    engine::check(gf_batch_set_norm_ops(batch_, dev_norm ? nops.data() : nullptr, dev_norm ? (int)nops.size() : 0,
                                        host_normalizations.data(), ntable.data(), (int)ntable.size()),
                  "gf_batch_set_norm_ops");
    stats_.device_norms = dev_norm;
    // with device norms a point is its flat parameter array: each Variable's value
    // (Params::from_minuit_vector: Minuit value minus blinding) at each of its slots
    struct ParamSlot {
        int fitter;
        unsigned int slot;
        fptype blind;
    };
    std::vector<ParamSlot> pmap;
    for(auto &sl : slots)
        if(sl.first.getFitterIndex() >= 0)
            pmap.push_back({sl.first.getFitterIndex(), sl.second, sl.first.getBlind(Variable::Key())});
    std::vector<fptype> pbuf(host_parameters.begin(), host_parameters.end());
    const std::vector<unsigned int> yields = pdf_->extendedYieldSlots();
    // the template norms serve points whose normalisation no floating parameter moves
    const fptype *point_norms = nops.empty() ? host_normalizations.data() : nullptr;
    Batcher B;
    B.cv_toy.reset(new std::condition_variable[T]);
    B.x.assign(T, nullptr);
    B.x0.assign(T, 0);
    B.xn.assign(T, 0);
    B.pending.assign(T, 0);
    B.done.assign(T, 0);
    B.result.assign(T, {});
    values_.assign(T, {});
    errors_.assign(T, {});
    valid_.assign(T, 0);
    nfcn_.assign(T, 0);
    fval_.assign(T, 0.0);
    int next    = 0;  // next toy to start (under B.mu)
    const int W = std::max(1, std::min(threads, T));
    B.running   = 0;
    std::exception_ptr worker_failure;  // the first exception of a toy's fit (under B.mu)
    auto worker = [&]() {
        for(;;) {
            int t;
            {
                std::lock_guard<std::mutex> lk(B.mu);
                if(next >= T)
                    break;
                t = next++;
                B.running++;
            }
            try {
                ToyFCN fcn(B, t, multi_);
                Minuit2::MnUserParameters up = static_cast<const Minuit2::MnUserParameters &>(params);
                Minuit2::MnMigrad migrad(fcn, up, Minuit2::MnStrategy(strategy));
                Minuit2::FunctionMinimum m = migrad(0, tolerance);
                values_[t] = m.UserParameters().Params();
                errors_[t] = m.UserParameters().Errors();
                valid_[t]  = m.IsValid() ? 1 : 0;
                nfcn_[t]   = (int)m.NFcn();
                fval_[t]   = m.Fval();
            } catch(...) {
                // the toy ends as an invalid fit; the exception is rethrown after the join
                valid_[t] = 0;
                std::lock_guard<std::mutex> lk(B.mu);
                if(!worker_failure)
                    worker_failure = std::current_exception();
            }
            std::lock_guard<std::mutex> lk(B.mu);
            B.running--;
            B.cv_req.notify_one();
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(W);
    for(int i = 0; i < W; i++)
        pool.emplace_back(worker);

    // batcher: prepare each request as it arrives (the Minuit work of the other toys
    // goes on meanwhile), launch once every running toy waits -- or when no new
    // request came for 200 us -- and answer the toys of that launch
    std::vector<uint8_t> act(T, 0), prepared(T, 0);
    // calculateNLL's extended term (AddPdf.cu:187-204) of each prepared point, with the
    // toy's own event count; zero for every other top-level PDF
    std::vector<std::vector<double>> extra(T);
    std::vector<double> out(multi_ ? (size_t)T * GF_BATCH_MAX_POINTS : (size_t)T);
    std::vector<int> taken, serve;
    std::exception_ptr failure;
    bool round_failed = false;
    int64_t points    = 0;
    auto prepare = [&](int t) {
        extra[t].assign(B.xn[t], 0.0);
        if(dev_norm) {
            for(size_t k = 0; k < B.xn[t]; k++) {
                const std::vector<double> &x = (*B.x[t])[B.x0[t] + k];
                for(const ParamSlot &m : pmap)
                    pbuf[m.slot] = x.at(m.fitter) - m.blind;
                if(!yields.empty()) {  // AddPdf::nllExtraTermFor with the toy's event count
                    fptype expEvents = 0;
                    for(unsigned int y : yields)
                        expEvents += pbuf[y];
                    extra[t][k] = expEvents - (fptype)count_[t] * std::log(expEvents);
                }
                engine::check(gf_batch_add_point(batch_, t, pbuf.data(), point_norms), "gf_batch_add_point");
            }
            return;
        }
        for(size_t k = 0; k < B.xn[t]; k++) {
            // push the point into the shared Variables, then the host arrays as
            // Params::from_minuit_vector: flag the values that move, so that an
            // Amp3Body recomputes its integrals only when a mass or width changes
            const std::vector<double> &x = (*B.x[t])[B.x0[t] + k];
            for(Variable &v : vars) {
                const double nv = x.at(v.getFitterIndex()) - v.getBlind(Variable::Key());
                v.setChanged(v.getValue() != nv);
                v.setValue(nv);
            }
            for(const auto &sl : slots)
                host_parameters[sl.second] = sl.first.getValue();
            pdf_->normalize();
            extra[t][k] = pdf_->nllExtraTermFor((fptype)count_[t]);
            if(multi_)
                engine::check(gf_batch_add_point(batch_, t, host_parameters.data(), host_normalizations.data()),
                              "gf_batch_add_point");
            else
                engine::check(gf_batch_set_toy(batch_, t, host_parameters.data(), host_normalizations.data()),
                              "gf_batch_set_toy");
        }
    };
    // With several points per toy the launches are pipelined: a stage goes to the GPU
    // once half of the running toys are prepared (or the round is complete), and the
    // next stage is prepared while it runs; at most one launch is in flight besides
    // the buffer being filled, and the batcher never blocks on new requests while a
    // launch is in flight.  When preparing is cheap against the GPU work (few points,
    // simple trees) a round stays one launch.
    struct Flight {
        uint32_t ticket = 0;
        std::vector<int> serve;
        int64_t points = 0;
        bool failed    = false;
    };
    Flight flight;
    bool in_flight     = false;
    bool split         = false;
    bool split_decided = false;
    auto deliver = [&](const std::vector<int> &who, bool failed) {
        {
            std::lock_guard<std::mutex> lk(B.mu);
            for(int t : who) {
                if(failed)
                    B.result[t].assign(B.xn[t], 1e300);
                else if(multi_)
                    B.result[t].assign(out.begin() + (size_t)t * GF_BATCH_MAX_POINTS,
                                       out.begin() + (size_t)t * GF_BATCH_MAX_POINTS + B.xn[t]);
                else
                    B.result[t].assign(1, out[t]);
                if(!failed)
                    for(size_t k = 0; k < B.result[t].size() && k < extra[t].size(); k++)
                        B.result[t][k] += extra[t][k];
                B.pending[t] = 0;
                B.done[t]    = 1;
                B.waiting--;
                prepared[t] = 0;
                act[t]      = 0;
            }
        }
        for(int t : who)
            B.cv_toy[t].notify_one();
    };
    auto land = [&]() {  // wait for the launch in flight and answer its toys
        const auto g0 = std::chrono::steady_clock::now();
        bool failed   = flight.failed;
        if(!failed) {
            try {
                engine::check(gf_batch_wait_points(batch_, flight.ticket, out.data()), "gf_batch_wait_points");
            } catch(...) {
                failure      = std::current_exception();
                failed       = true;
                round_failed = true;  // later requests are answered without launches
            }
        }
        stats_.gpu_wait_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - g0).count();
        deliver(flight.serve, failed);
        in_flight = false;
    };
    for(;;) {
        taken.clear();
        bool launch = false;
        {
            std::unique_lock<std::mutex> lk(B.mu);
            const auto w0 = std::chrono::steady_clock::now();
            // the requests of the toys blocked in an FCN call are all in hand only when
            // every running toy waits; new ones may still come until then
            bool arrived = B.unprepared > 0;
            if(!in_flight)
                arrived = B.cv_req.wait_for(
                    lk, serve.empty() ? std::chrono::microseconds(50000) : std::chrono::microseconds(200), [&] {
                        return B.unprepared > 0 || (!serve.empty() && B.waiting == B.running)
                               || (B.running == 0 && next >= T);
                    });
            stats_.gather_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
            if(B.running == 0 && next >= T && serve.empty() && !in_flight)
                break;
            for(int t = 0; t < T; t++)
                if(B.pending[t] && !prepared[t]) {
                    prepared[t] = 1;
                    taken.push_back(t);
                }
            B.unprepared -= (int)taken.size();
            // split rounds in two stages when preparing costs more than the GPU work:
            // then the GPU work of one stage hides behind the preparation of the other,
            // which outweighs the second launch's fixed cost (its grid still spans every
            // toy's rows).  Decided once, from the first twelve rounds, run unsplit (the
            // first ones, single points, would overstate the preparation)
            if(multi_ && !split_decided && (stages_ != 0 || stats_.launches >= 12)) {
                // a second stage costs one more launch per round (tens of microseconds) and
                // hides the preparation and the Minuit work of half the toys behind the GPU
                split         = stages_ == 2 || (stages_ == 0 && stats_.host_prep_s > 0.1 * stats_.gpu_wait_s);
                split_decided = true;
                stats_.two_stages = split;
            }
            const size_t stage = split ? (size_t)std::max(1, B.running / 2) : (size_t)T + 1;
            const size_t ready = serve.size() + taken.size();
            launch = ready > 0 && (B.waiting == B.running || (!in_flight && !arrived) || ready >= stage);
        }
        // the taken toys' requests stay untouched while those toys wait: prepare them
        // without holding the lock
        const auto p0 = std::chrono::steady_clock::now();
        for(int t : taken) {
            if(!round_failed) {
                try {
                    prepare(t);
                } catch(...) {
                    failure      = std::current_exception();
                    round_failed = true;
                }
            }
            points += (int64_t)B.xn[t];
            act[t] = 1;
            serve.push_back(t);
        }
        stats_.host_prep_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - p0).count();
        if(!launch) {
            if(in_flight && taken.empty())
                land();
            continue;
        }
        if(in_flight)
            land();  // frees the other point buffer
        stats_.launches++;
        stats_.fcn_calls += points;
        if(multi_) {
            flight.serve.swap(serve);
            flight.points = points;
            flight.failed = round_failed;
            if(!round_failed) {
                try {
                    engine::check(gf_batch_submit_points(batch_, GF_METRIC_NLL, &flight.ticket),
                                  "gf_batch_submit_points");
                } catch(...) {
                    failure       = std::current_exception();
                    round_failed  = true;
                    flight.failed = true;
                }
            }
            if(flight.failed)
                (void)gf_batch_clear_points(batch_);
            in_flight = true;
        } else {
            const auto g0 = std::chrono::steady_clock::now();
            if(!round_failed) {
                try {
                    engine::check(gf_batch_reduce(batch_, act.data(), GF_METRIC_NLL, out.data()), "gf_batch_reduce");
                } catch(...) {
                    failure      = std::current_exception();
                    round_failed = true;
                }
            }
            stats_.gpu_wait_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - g0).count();
            deliver(serve, round_failed);
        }
        serve.clear();
        points = 0;
    }
    for(auto &th : pool)
        th.join();
    // restore the starting values in the tree
    for(Variable &v : vars)
        v.setValue(start.at(v.getFitterIndex()));
    pdf_->updateParameters();
    stats_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if(failure)
        std::rethrow_exception(failure);
    if(worker_failure)
        std::rethrow_exception(worker_failure);
}

}  // namespace GooFit

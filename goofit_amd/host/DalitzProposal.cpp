// DalitzProposal: host side of the learned normalisation proposal (gf_mlp_* of
// the C ABI).
#include <goofit/Engine.h>
#include <goofit/Error.h>
#include <goofit/PDFs/physics/Amp3Body.h>
#include <goofit/PDFs/physics/DalitzProposal.h>

namespace GooFit {

DalitzProposal::DalitzProposal(Observable m12, Observable m13, unsigned long long seed, int batch_cap)
    : m12_(m12)
    , m13_(m13) {
    box_[0] = m12.getLowerLimit();
    box_[1] = m12.getUpperLimit();
    box_[2] = m13.getLowerLimit();
    box_[3] = m13.getUpperLimit();
    if(!(box_[1] > box_[0]) || !(box_[3] > box_[2]))
        throw GeneralError("DalitzProposal: empty (m12, m13) box");
    engine::check(gf_mlp_create(engine::ctx(), seed, batch_cap, &mlp_), "gf_mlp_create");
}

DalitzProposal::~DalitzProposal() { gf_mlp_destroy(mlp_); }

std::pair<float, float> DalitzProposal::train(const UnbinnedDataSet &events, int steps, int batch, float lr,
                                              unsigned long long seed) {
    const std::vector<fptype> a = events.getColumn(m12_), b = events.getColumn(m13_);
    if(a.empty())
        throw GeneralError("DalitzProposal::train: no events");
    float l0 = 0.f, l1 = 0.f;
    engine::check(gf_mlp_train_host(mlp_, a.data(), b.data(), (int64_t)a.size(), box_, steps, batch, lr, seed, &l0,
                                    &l1),
                  "gf_mlp_train_host");
    built_ = false;
    return {l0, l1};
}

std::pair<float, float> DalitzProposal::trainOnModel(Amp3Body &model, int64_t nevents, unsigned long long gen_seed,
                                                     int steps, int batch, float lr, unsigned long long seed) {
    if(nevents < 1)
        throw GeneralError("DalitzProposal::trainOnModel: no events");
    model.copyParams();
    engine::flush();
    float l0 = 0.f, l1 = 0.f;
    engine::check(gf_mlp_train_generated(mlp_, (int)model.getFunctionIndex(), nevents, gen_seed, box_, steps, batch,
                                         lr, seed, &l0, &l1),
                  "gf_mlp_train_generated");
    built_ = false;
    return {l0, l1};
}

void DalitzProposal::build(int grid, double alpha) {
    engine::check(gf_mlp_build_proposal(mlp_, box_, grid, alpha), "gf_mlp_build_proposal");
    built_ = true;
}

std::vector<float> DalitzProposal::evaluate(const std::vector<float> &xy) const {
    if(xy.size() % 2)
        throw GeneralError("DalitzProposal::evaluate: xy holds (x, y) pairs");
    std::vector<float> out(xy.size() / 2);
    if(!out.empty())
        engine::check(gf_mlp_forward(mlp_, xy.data(), (int64_t)out.size(), out.data()), "gf_mlp_forward");
    return out;
}

std::vector<float> DalitzProposal::getParameters() const {
    std::vector<float> p(gf_mlp_param_count());
    engine::check(gf_mlp_get_params(mlp_, p.data(), (int)p.size()), "gf_mlp_get_params");
    return p;
}

void DalitzProposal::setParameters(const std::vector<float> &p) {
    engine::check(gf_mlp_set_params(mlp_, p.data(), (int)p.size()), "gf_mlp_set_params");
    built_ = false;
}

void DalitzProposal::setRetraining(const Retraining &r) {
    if(r.events < 1 || r.steps < 1 || r.batch < 1 || !(r.lr > 0) || r.grid < 2 || !(r.alpha >= 0 && r.alpha <= 1)
       || !(r.factor > 1))
        throw GeneralError("DalitzProposal::setRetraining: need events, steps, batch >= 1, lr > 0, grid >= 2, "
                           "alpha in [0, 1] and factor > 1");
    retrain_    = r;
    retrain_on_ = true;
}

void DalitzProposal::retrain(Amp3Body &model, unsigned long long seed) {
    trainOnModel(model, retrain_.events, seed ^ 0x5eedf00dULL, retrain_.steps, retrain_.batch, retrain_.lr, seed + 1);
    build(retrain_.grid, retrain_.alpha);
    retrains_++;
}

}  // namespace GooFit

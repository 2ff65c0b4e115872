// Engine.cpp -- per-thread flattened tree and GPU context behind the GooFit API.
#include <goofit/Engine.h>
#include <goofit/Error.h>

#include <cstring>

namespace GooFit {

thread_local std::vector<fptype> host_parameters;
thread_local std::vector<fptype> host_constants;
thread_local std::vector<fptype> host_observables;
thread_local std::vector<fptype> host_normalizations;
thread_local std::vector<int> host_function_table;
thread_local std::vector<std::string> host_function_name;
thread_local std::vector<std::string> host_parameter_name;
thread_local unsigned int host_callnumber = 0;

namespace engine {

namespace {
int g_device = 0;

struct State {
    gf_ctx *ctx        = nullptr;
    bool program_dirty = true;
    bool decay_dirty   = false;
    double decay[6]    = {0, 0, 0, 0, 0, 0};
    ~State() {
        if(ctx)
            gf_ctx_destroy(ctx);
    }
};
thread_local State t_state;
}  // namespace

void set_device(int device) { g_device = device; }
int get_device() { return g_device; }

gf_ctx *ctx() {
    if(!t_state.ctx) {
        gf_ctx *c    = nullptr;
        gf_status st = gf_ctx_create(&c, g_device);
        if(st != GF_OK)
            throw GeneralError("GooFit engine: cannot create a GPU context on device {}: {}", g_device,
                               gf_last_error());
        t_state.ctx           = c;
        t_state.program_dirty = true;
        t_state.decay_dirty   = true;
    }
    return t_state.ctx;
}

void release() {
    if(t_state.ctx) {
        gf_ctx_destroy(t_state.ctx);
        t_state.ctx = nullptr;
    }
    t_state.program_dirty = true;
}

void mark_program_dirty() { t_state.program_dirty = true; }

namespace {
thread_local unsigned long long t_cache_epoch = 0;
}
unsigned long long cache_epoch() { return t_cache_epoch; }
void bump_cache_epoch() { ++t_cache_epoch; }
namespace {
thread_local unsigned long long t_layout_epoch = 0;
}
unsigned long long layout_epoch() { return t_layout_epoch; }
void bump_layout_epoch() { ++t_layout_epoch; }

void check(gf_status status, const char *what) {
    if(status != GF_OK)
        throw GeneralError("{}: {}", what, gf_last_error());
}

void set_decay(const double decay6[6]) {
    std::memcpy(t_state.decay, decay6, sizeof(t_state.decay));
    t_state.decay_dirty = true;
}

void flush() {
    gf_ctx *c = ctx();
    if(t_state.decay_dirty) {
        check(gf_decay3_set(c, t_state.decay), "gf_decay3_set");
        t_state.decay_dirty = false;
    }
    if(host_function_table.empty())
        throw GeneralError("GooFit engine: no PDF has been set up (setData / setIndices)");
    if(t_state.program_dirty) {
        check(gf_program_set(c, host_function_table.data(), (int)host_function_table.size(), host_parameters.data(),
                             (int)host_parameters.size(), host_constants.data(), (int)host_constants.size(),
                             host_observables.data(), (int)host_observables.size(), host_normalizations.data(),
                             (int)host_normalizations.size()),
              "gf_program_set");
        t_state.program_dirty = false;
        return;
    }
    check(gf_params_update(c, host_parameters.data(), (int)host_parameters.size()), "gf_params_update");
    check(gf_norms_update(c, host_normalizations.data(), (int)host_normalizations.size()), "gf_norms_update");
}

}  // namespace engine
}  // namespace GooFit

// Engine state behind the GooFit API: the flattened PDF tree of the last
// setIndices() (the reference keeps it in process-global SmartVectors,
// src/PDFs/detail/Globals.cu:9-13, include/goofit/PDFs/detail/Globals.h) and the
// gf_ctx (include/goofit_hip.h) holding its device copy, the events and the
// wave caches.  Both are per host thread, so independent fits can run on
// different threads, each on its own HIP stream.
#pragma once

#include <string>
#include <vector>

#include <goofit/GlobalCudaDefines.h>
#include <goofit_hip.h>

namespace GooFit {

extern thread_local std::vector<fptype> host_parameters;
extern thread_local std::vector<fptype> host_constants;
extern thread_local std::vector<fptype> host_observables;
extern thread_local std::vector<fptype> host_normalizations;
extern thread_local std::vector<int> host_function_table;  // gf_func codes
extern thread_local std::vector<std::string> host_function_name;
extern thread_local std::vector<std::string> host_parameter_name;
extern thread_local unsigned int host_callnumber;

namespace engine {
/// Device used by contexts created from now on (default 0).
void set_device(int device);
int get_device();
/// This thread's context, created on first use; throws GeneralError without a GPU.
gf_ctx *ctx();
/// Destroy this thread's context (device memory is released).
void release();
/// The flattened arrays changed: upload them before the next device call.
void mark_program_dirty();
/// Upload whatever changed since the last call (tree, parameters, normalisations,
/// decay constants).  Every device entry point of the API calls this first.
void flush();
/// Throw GeneralError("<what>: <gf_last_error()>") unless status == GF_OK.
void check(gf_status status, const char *what);
/// Three-body decay constants (c_motherMass .. c_mother_meson_radius).
void set_decay(const double decay6[6]);
/// Counter of wave-cache recomputations on this thread: derivative points that
/// see it unchanged share the caches and may be reduced in one request.
unsigned long long cache_epoch();
void bump_cache_epoch();
/// Counter of rebuilds of the flattened arrays' layout on this thread
/// (PdfBase::setIndices): plans of slots into host_parameters stay valid while
/// it is unchanged.
unsigned long long layout_epoch();
void bump_layout_epoch();
}  // namespace engine

}  // namespace GooFit

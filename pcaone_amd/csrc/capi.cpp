// extern "C" entry points: argument checks, exception -> status translation.
#include <string>

#include "../../include/pcaone_amd.h"
#include "common.hpp"
#include "synth_host.hpp"

namespace pcaone {

namespace {
thread_local std::string g_error;
}

void set_global_error(const std::string& msg) { g_error = msg; }
const char* global_error() { return g_error.c_str(); }

}  // namespace pcaone

using namespace pcaone;

namespace {

template <class F>
int guard_global(F&& f) {
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_global_error("host out of memory");
    return kErrInternal;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return kErrInternal;
  }
}

synth::Spec to_spec(const pcaone_synth_spec* s) {
  require(s != nullptr, kErrArgument, "null synth spec");
  return synth::Spec{s->seed,   s->nsamples, s->npops,     s->nsnps,          s->freq_lo,       s->freq_hi,
                     s->fst_lo, s->fst_hi,   s->miss_rate, s->high_miss_frac, s->high_miss_rate};
}

}  // namespace

extern "C" {

const char* pcaone_last_error(const pcaone_ctx* ctx) {
  (void)ctx;
  return global_error();
}

int pcaone_synth_packed(const pcaone_synth_spec* spec, uint8_t* packed_out, uint32_t* attempts_out) {
  return guard_global([&] {
    require(packed_out != nullptr, kErrArgument, "null output buffer");
    synth::generate(to_spec(spec), packed_out, attempts_out);
  });
}

int pcaone_synth_write_plink(const pcaone_synth_spec* spec, const char* prefix) {
  return guard_global([&] {
    require(prefix != nullptr, kErrArgument, "null prefix");
    synth::write_plink(to_spec(spec), prefix);
  });
}

}  // extern "C"

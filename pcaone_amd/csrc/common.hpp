// Shared error type and helpers for the host side of the library.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

namespace pcaone {

// Status codes returned by every C-ABI entry point.
enum Status : int {
  kOk = 0,
  kErrArgument = 1,  // invalid argument / shape
  kErrDevice = 2,    // HIP runtime failure
  kErrState = 3,     // call out of order (e.g. no data uploaded)
  kErrIO = 4,        // file open / read / write / format failure
  kErrNumeric = 5,   // numerical breakdown (e.g. non-positive-definite Gram)
  kErrNoDevice = 6,  // no usable GPU
  kErrInternal = 7,
};

class Error : public std::runtime_error {
 public:
  Error(Status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
  Status status;
};

[[noreturn]] inline void fail(Status s, const std::string& msg) { throw Error(s, msg); }

inline void require(bool cond, Status s, const std::string& msg) {
  if (!cond) fail(s, msg);
}

// Thread-local message of the last failed call made without a context.
void set_global_error(const std::string& msg);
const char* global_error();

inline uint64_t bytes_per_snp(uint64_t nsamples) { return (nsamples + 3) >> 2; }
inline uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

}  // namespace pcaone

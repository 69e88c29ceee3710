// The packed product without device allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops): every
// intermediate lives in `scratch`, a fixed array of this library in device
// memory, one copy per GPU, held until the process ends.
//
// Calls on one GPU share that GPU's copy, which keeps the last call's
// contents; calls that overlap on one GPU race.  Everything a call reads from
// the scratch it first writes in the same call (row padding, operand pieces,
// the int8 column maxima), with kernels only: no memset nodes.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "context.hpp"
#include "device.hpp"
#include "gemm.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"

namespace pcaone {

namespace {

constexpr uint64_t kScratchBytes = kNoAllocScratchBytes;

#define GRID_STRIDE(t, n) \
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (n); t += (uint64_t)gridDim.x * blockDim.x)

inline unsigned grid_for(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65535u * 4 ? 65535u * 4 : g));
}
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per GPU (the array's name in host code is only a
// stand-in for the symbol, never a device address).
unsigned char* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

unsigned char* scratch_of(int device) {
  require(device >= 0 && device < kMaxDevices, kErrArgument, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    // called with `device` current (the caller's hipSetDevice)
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) fail(kErrDevice, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch[device];
}

// Y (nrows x L, column-major FP64) from the product rows (rows x lp) in the
// form's accumulator: FP64, FP32, int32 times the int8 column factor, or FP32
// times the fp8 column factor.
__global__ void k_gemm_out(const void* __restrict__ acc, int kind, uint64_t nrows, int L, int lp,
                           const double* __restrict__ mul, double* __restrict__ Y) {
  GRID_STRIDE(t, nrows * (uint64_t)L) {
    const uint64_t l = t / nrows, r = t - l * nrows;
    const uint64_t idx = r * lp + l;
    double v;
    if (kind == 0)
      v = static_cast<const double*>(acc)[idx];
    else if (kind == 2)
      v = (double)static_cast<const int32_t*>(acc)[idx] * mul[l];
    else if (kind == 3)
      v = (double)static_cast<const float*>(acc)[idx] * mul[l];
    else
      v = static_cast<const float*>(acc)[idx];
    Y[t] = v;
  }
}

uint64_t carve(uint64_t& off, uint64_t bytes) {
  const uint64_t at = off;
  off = round_up(off + bytes, 256);
  return at;
}

}  // namespace

void packed_gemm_noalloc(Context& x, Form f, const uint8_t* rows, uint64_t nrows, uint32_t ncodes,
                         const double* W, uint32_t L, double* Y) {
  const int lp = (int)round_up(L, 16);
  const GemmShape shape = gemm_default_shape(f, lp);
  const uint64_t rows_pad = round_up(nrows, (uint64_t)gemm_row_tile(shape));
  const uint64_t K = round_up(ncodes, kKStep), pitch = K / 16, bps = bytes_per_snp(ncodes);
  const int P = form_pieces(f);
  uint64_t off = 0;
  const uint64_t oX = carve(off, rows_pad * pitch * 4);
  const uint64_t oW = carve(off, K * lp * 8);
  const uint64_t oWt = carve(off, (uint64_t)P * lp * K * form_wbytes(f));
  const uint64_t oY = carve(off, rows_pad * lp * form_accbytes(f));
  const uint64_t oS = carve(off, 2 * 128 * 8);
  const uint64_t oM = carve(off, 128 * 8);
  require(off <= kScratchBytes, kErrArgument,
          "packed_gemm_noalloc: the workspace of " + std::to_string(off) + " bytes exceeds the " +
              std::to_string(kScratchBytes) + "-byte scratch");
  unsigned char* base = scratch_of(x.device);
  uint32_t* X = reinterpret_cast<uint32_t*>(base + oX);
  double* dw = reinterpret_cast<double*>(base + oW);
  unsigned char* wt = base + oWt;
  unsigned char* acc = base + oY;
  double* i8s = reinterpret_cast<double*>(base + oS);
  unsigned long long* i8max = reinterpret_cast<unsigned long long*>(base + oM);
  hipStream_t s = x.stream;
  launch_fill_rows(X, pitch, 0, rows_pad, s);
  launch_interleave(rows, nrows, bps, ncodes, X, pitch, s);
  launch_colmajor_to_rows(W, ncodes, (int)L, lp, K, dw, s);
  if (form_colscaled(f)) launch_col_scales(f, dw, ncodes, lp, nullptr, i8max, i8s, i8s + 128, s);
  launch_make_pieces(f, dw, ncodes, K, lp, nullptr, nullptr, wt, K, (uint64_t)lp * K, nullptr, nullptr, s, i8s);
  GemmArgs a{};
  a.X = X;
  a.xpitch = pitch;
  a.rows = rows_pad;
  a.Wt = wt;
  a.wpitch = K;
  a.wpiece = (uint64_t)lp * K;
  a.Y = acc;
  a.ychunk = rows_pad * lp;
  a.k_begin = 0;
  a.k_end = K;
  a.kchunk = K;
  a.lp = lp;
  launch_packed_gemm(f, shape, a, s);
  const int kind = f == Form::FP64 ? 0 : (f == Form::I8 ? 2 : (f == Form::F8 ? 3 : 1));
  hipLaunchKernelGGL(k_gemm_out, dim3(grid_for(nrows * (uint64_t)L)), dim3(256), 0, s, acc, kind, nrows, (int)L, lp,
                     i8s + 128, Y);
  check_launch("gemm_out");
  PCAONE_HIP(hipStreamSynchronize(s));
}

}  // namespace pcaone

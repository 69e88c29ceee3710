/*
 * pcaone_amd: MI355X-native randomized-SVD hot path of PCAone (--svd 1 / --svd 2 / --emu)
 *
 * C ABI.  Every entry point returns 0 on success or a non-zero status
 * (PCAONE_ERR_*) and leaves a message readable with pcaone_last_error().
 * Matrices are column-major FP64 host arrays with leading dimension = rows,
 * the layout of the reference's Eigen::MatrixXd (src/Common.hpp:29).
 * Host pointers are borrowed for the duration of the call only; the context
 * owns every device buffer.  A context is single-threaded and stream-ordered.
 */
#ifndef PCAONE_AMD_H_
#define PCAONE_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCAONE_OK 0
#define PCAONE_ERR_ARGUMENT 1
#define PCAONE_ERR_DEVICE 2
#define PCAONE_ERR_STATE 3
#define PCAONE_ERR_IO 4
#define PCAONE_ERR_NUMERIC 5
#define PCAONE_ERR_NO_DEVICE 6
#define PCAONE_ERR_INTERNAL 7

/* Arithmetic form of the two genotype products G = A*Omega and H = A^T*G. */
#define PCAONE_PREC_FP64 0   /* FP64 MFMA (oracle-grade form)                          */
#define PCAONE_PREC_FP32 1   /* FP32 MFMA, plain form                                   */
#define PCAONE_PREC_BF16 2   /* BF16 MFMA, exact genotypes, hi/lo split of Omega and G  */
#define PCAONE_PREC_BF16X1 3 /* BF16 MFMA, exact genotypes, single-piece Omega and G    */

typedef struct pcaone_ctx pcaone_ctx;

/* Synthetic Balding-Nichols PLINK data (test and benchmark input; no
 * counterpart in the reference). */
typedef struct pcaone_synth_spec {
  uint64_t seed;
  uint32_t nsamples;
  uint32_t npops;
  uint64_t nsnps;
  double freq_lo, freq_hi; /* ancestral A1 frequency ~ U(freq_lo, freq_hi) */
  double fst_lo, fst_hi;   /* per-population Fst ~ U(fst_lo, fst_hi)      */
  double miss_rate;        /* MCAR missingness of every genotype          */
  double high_miss_frac;   /* fraction of samples with extra missingness  */
  double high_miss_rate;   /* extra missing rate of those samples         */
} pcaone_synth_spec;

/* ---- errors ------------------------------------------------------------ */
/* Message of the last failure on `ctx`, or of the calling thread's last
 * context-free failure when ctx is NULL.  Replaces cao.error()'s throw
 * (src/Logger.hpp:85-94). */
const char* pcaone_last_error(const pcaone_ctx* ctx);

/* ---- synthetic data ---------------------------------------------------- */
/* packed_out: nsnps x ceil(nsamples/4) bytes, the .bed payload after its
 * 3-byte magic.  attempts_out (nullable): per-SNP redraw counter. */
int pcaone_synth_packed(const pcaone_synth_spec* spec, uint8_t* packed_out, uint32_t* attempts_out);
/* Writes <prefix>.bed/.bim/.fam. */
int pcaone_synth_write_plink(const pcaone_synth_spec* spec, const char* prefix);

#ifdef __cplusplus
}
#endif

#endif /* PCAONE_AMD_H_ */

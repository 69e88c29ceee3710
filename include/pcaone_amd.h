/*
 * pcaone_amd: MI355X-native randomized-SVD hot path of PCAone (v0.7.2):
 * --svd 1 (single-pass RSVD), --svd 2 (window-based RSVD) and --emu.
 *
 * C ABI.  Every entry point returns 0 on success or a non-zero status
 * (PCAONE_ERR_*); the message is readable with pcaone_last_error().  This
 * replaces the reference's cao.error() throw (src/Logger.hpp:85-94) and
 * Param's exit(EXIT_FAILURE) (src/Cmd.cpp:240-264).
 *
 * Matrices are column-major FP64 host arrays with leading dimension = rows,
 * the layout of the reference's Mat2D = Eigen::MatrixXd (src/Common.hpp:29).
 * Host pointers are borrowed for the duration of a call only; the context
 * owns every device buffer.  A context is single-threaded and stream-ordered.
 * Reference paths below are relative to the PCAone source tree.
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

/* Arithmetic form of the genotype products G = A*Omega and H = A^T*G.
 * Genotypes enter every form exactly; Omega and G are rounded to the form's
 * operand type (split into BF16 hi/lo pieces for PCAONE_PREC_BF16, into two
 * int8 fixed-point pieces for PCAONE_PREC_INT8). */
#define PCAONE_PREC_FP64 0   /* FP64 MFMA, FP64 accumulation                 */
#define PCAONE_PREC_FP32 1   /* FP32 MFMA, FP32 accumulation (plain form)    */
#define PCAONE_PREC_BF16 2   /* BF16 MFMA, hi/lo split of Omega and G        */
#define PCAONE_PREC_BF16X1 3 /* BF16 MFMA, single BF16 piece of Omega and G  */
#define PCAONE_PREC_BF16MIX 4 /* BF16 MFMA, one piece of Omega, hi/lo of G   */
#define PCAONE_PREC_FP16 5   /* FP16 MFMA, single FP16 piece of Omega and G  */
#define PCAONE_PREC_FP16MIX 6 /* one FP16 piece of Omega, BF16 hi/lo of G    */
/* int8 MFMA: genotypes as v/2 in {+1, 0, -1}; each column of Omega and of
 * diag(sc) G as a 16-bit fixed-point value (column scale = max |value| /
 * 32639) split into two int8 pieces; exact int32 accumulation.  Single-matrix
 * path only (not the batched form). */
#define PCAONE_PREC_INT8 7
/* FP8 MFMA (f8f6f4, E4M3 on both sides): genotypes exact in E4M3; each column
 * of Omega and of diag(sc) G scaled by a power of two and split into three
 * E4M3 pieces (about 12 significant bits per value); FP32 accumulation.
 * Single-matrix path only (not the batched form). */
#define PCAONE_PREC_FP8 8

typedef struct pcaone_ctx pcaone_ctx;

/* Run settings: the fields of the reference's Param used on this path
 * (src/Cmd.hpp:16-98; defaults from pcaone_params_default match it). */
typedef struct pcaone_params {
  int svd;              /* -d/--svd: 1 sSVD (Halko), 2 winSVD (default 2)          */
  uint32_t k;           /* -k/--pc (10)                                             */
  uint32_t oversamples; /* --oversamples (10); raised to k as src/Cmd.cpp:216 does  */
  uint32_t maxp;        /* --maxp (20)                                              */
  uint32_t bands;       /* -w/--batches (64)                                        */
  uint32_t maxiter;     /* --maxiter (100), EM iterations with --emu                */
  double tol;           /* --tol-rsvd (1e-4)                                        */
  double tolem;         /* --tol-em (1e-5)                                          */
  double maf;           /* --maf (0: no filter)                                     */
  int seed;             /* --seed (112)                                             */
  int rand;             /* --rand: 1 Gaussian, 0 uniform (1)                        */
  int ploidy;           /* 2 (1 with --haploid)                                     */
  int noshuffle;        /* -S/--no-shuffle                                          */
  int emu;              /* --emu                                                    */
  int printv;           /* -V/--printv: also return / write loadings                */
  int verbose;          /* -v (0 here: silent)                                      */
  int precision;        /* PCAONE_PREC_* (PCAONE_PREC_FP32)                         */
  int wm_g, wm_h;       /* GEMM wave tiles (0 = tuned default)                      */
  int waves_g, waves_h; /* GEMM waves per workgroup, 4 or 8 (0 = tuned default)     */
  uint64_t kchunk_g;    /* split-K chunk of the G product in codes (0 = auto)       */
  uint64_t kchunk_h;    /* split-K chunk of the H product in codes (0 = auto)       */
  double memory;        /* -m/--memory in GB: > 0 runs the out-of-core schedule     */
                        /* (block plan, no flipOmg for --svd 1, blocks as winSVD    */
                        /* windows, permute_plink order); the genotypes stay in HBM */
  uint32_t buffer;      /* --buffer (2), GB of the permute_plink read buffer        */
  int ld;               /* -D/--ld: the PCA runs unscaled and pcaone_run_bfile also */
                        /* writes <out>.residuals and <out>.mbim                    */
  int ld_stats;         /* --ld-stats: 0 residuals of the PCs (default), 1 the      */
                        /* centered genotypes                                       */
  uint32_t imaxiter;    /* --imaxiter (1000): restarts of the IRAM solver, --svd 0  */
  double itol;          /* --itol (1e-6): IRAM convergence tolerance                */
  int stream;           /* pcaone_run_bfile with memory > 0: stream the .bed blocks */
                        /* of the -m plan from the mapped file every pass (data    */
                        /* beyond HBM) instead of loading the whole matrix         */
  int hardcall;         /* --hardcall: read the hardcall track of a .pgen that also */
                        /* stores dosages (without it such a file is rejected)      */
  uint32_t threads;     /* -n/--threads (12): with -m, a .pgen's shuffle depends on */
                        /* min(threads, host cores) (src/Main.cpp:53-54,169-172)    */
  int miss_gather;      /* 1: missing-entry corrections gather their source rows    */
                        /* per entry from L2 (the CSR form) even where the tiled    */
                        /* form applies (reduced precisions with L <= 32)           */
} pcaone_params;

typedef struct pcaone_run_info {
  int epochs;        /* epochs of the last computeUSV                */
  int total_epochs;  /* epochs over all computeUSV calls (EMU)      */
  int em_iters;      /* EM iterations run                           */
  int n_trace;       /* entries available from pcaone_run_trace     */
  int n_sv;          /* singular values returned (0 with --maxp 0)  */
  double seconds;    /* wall time of the run (inputs resident)      */
  double g_ms, h_ms; /* summed device time of the two products      */
  uint64_t g_launches, h_launches;
  double omega_ms;   /* host time of the Omega draws (initOmg), included in seconds */
  uint64_t householder_fallbacks; /* thin-Q of H by Householder QR (H near rank deficient) */
  uint64_t gram_second_passes;    /* QR of G with a second pass over G (ill conditioned)   */
  uint64_t epoch_rollbacks;       /* winSVD epochs rerun because an Omega update's device
                                     thin-Q met a near-singular Gram of H                   */
  uint64_t stream_loads;          /* streamed genotypes: blocks moved host -> device        */
  double stream_load_ms;          /* host wall time of those loads (gather, DMA, interleave) */
} pcaone_run_info;

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

/* ---- errors and context ------------------------------------------------ */
/* Message of the last failure on ctx, or of the calling thread's last
 * context-free failure when ctx is NULL. */
const char* pcaone_last_error(const pcaone_ctx* ctx);
int pcaone_device_count(int* n);
int pcaone_open(int device, pcaone_ctx** out);
int pcaone_close(pcaone_ctx* ctx);
void pcaone_params_default(pcaone_params* p);

/* ---- Data boundary (FileBed / Data) ------------------------------------ */
/* FileBed::FileBed (src/FilePlink.hpp:10-30) without a device: N = lines of
 * <prefix>.fam, M = lines of <prefix>.bim, the .bed opened and its 3-byte
 * magic 6c 1b 01 checked.  PCAONE_ERR_IO with the reference's messages
 * ("Cannot open bed file.", "Incorrect magic number in plink bed file."). */
int pcaone_bed_dims(const char* prefix, uint32_t* nsamples, uint64_t* nsnps);
/* Replaces FileBed::FileBed + read_all (src/FilePlink.hpp:10-30,
 * src/FilePlink.cpp:26-120) and the --maf filter (src/Data.cpp:88-105): reads
 * <prefix>.fam/.bim/.bed, checks the magic, moves the payload to HBM
 * undecoded and counts allele frequencies on the device. */
int pcaone_bed_read(pcaone_ctx* ctx, const char* prefix, double maf, uint32_t* nsamples, uint64_t* nsnps);
/* PLINK 2 input (the hardcall path of src/FilePgen.cpp:27-287).
 * pcaone_pgen_dims replaces FilePgen::FilePgen (src/FilePgen.hpp:14-35)
 * without a device: N = data lines of <prefix>.psam (not starting with '#'),
 * M = the variant count of <prefix>.pgen, whose header must give the same N.
 * pcaone_pgen_read replaces FilePgen::read_all (src/FilePgen.cpp:27-89) and
 * the --maf filter: the records (every storage mode and record type of the
 * format; the index may sit in <prefix>.pgen.pgi) are decoded on the host
 * into .bed rows (ALT count 2 / 1 / 0 / missing -> code 00 / 10 / 11 / 01,
 * the codes whose values are the reference's pgen2dosage) and uploaded as
 * pcaone_bed_read uploads them.  A file that stores dosages is read only with
 * hardcall != 0 (the reference's --hardcall); multi-allelic files are
 * rejected, as the reference's reader rejects them (PCAONE_ERR_IO).
 * pcaone_pgen_decode writes the rows [snp_begin, snp_end) as .bed rows
 * (ceil(N/4) bytes each) into host memory, without a device.
 * pcaone_stream_pgen opens the file for streamed runs (records decoded at
 * every block load, src/FilePgen.cpp:153-287). */
int pcaone_pgen_dims(const char* prefix, uint32_t* nsamples, uint64_t* nsnps);
int pcaone_pgen_read(pcaone_ctx* ctx, const char* prefix, double maf, int hardcall, uint32_t* nsamples,
                     uint64_t* nsnps);
int pcaone_pgen_decode(const char* prefix, int hardcall, uint64_t snp_begin, uint64_t snp_end, uint8_t* rows);
int pcaone_stream_pgen(pcaone_ctx* ctx, const char* prefix, int hardcall, uint64_t block_snps, uint32_t* nsamples,
                       uint64_t* nsnps);
/* The same from memory: reserve, then upload raw .bed rows (ceil(N/4) bytes
 * per SNP, the payload after the 3-byte magic), then finalize. */
int pcaone_bed_reserve(pcaone_ctx* ctx, uint32_t nsamples, uint64_t nsnps);
int pcaone_bed_upload(pcaone_ctx* ctx, const uint8_t* rows, uint64_t snp_begin, uint64_t nsnps);
int pcaone_bed_finalize(pcaone_ctx* ctx);
/* Synthetic data generated on the device, byte-identical to
 * pcaone_synth_packed (reserve + upload + finalize). */
int pcaone_bed_generate(pcaone_ctx* ctx, const pcaone_synth_spec* spec);
/* Streamed genotypes, for matrices beyond HBM (the reference's out-of-core
 * reads, src/FilePlink.cpp:122-298, src/Halko.cpp:129-152,225-268): the
 * packed rows stay on the host and each product loads the block of at most
 * block_snps SNPs it covers.  pcaone_stream_bed maps <prefix>.bed (magic
 * checked as src/FilePlink.hpp:24-27); pcaone_stream_host reads the caller's
 * rows (nsnps x ceil(N/4) bytes, kept alive and unchanged while the context
 * uses them).  Opening streams every block once to count the codes.  Runs
 * on streamed genotypes take the out-of-core schedule (memory > 0) with
 * plan blocks of at most block_snps SNPs. */
int pcaone_stream_bed(pcaone_ctx* ctx, const char* prefix, uint64_t block_snps, uint32_t* nsamples, uint64_t* nsnps);
int pcaone_stream_host(pcaone_ctx* ctx, const uint8_t* rows, uint32_t nsamples, uint64_t nsnps, uint64_t block_snps);
/* The resident genotypes back as .bed rows (nsnps x ceil(N/4) bytes). */
int pcaone_bed_download(pcaone_ctx* ctx, uint8_t* rows);
int pcaone_dims(const pcaone_ctx* ctx, uint32_t* nsamples, uint64_t* nsnps, uint64_t* nmissing);
/* Data::F (src/FilePlink.cpp:37-55), unpermuted SNP order, nsnps values. */
int pcaone_allele_freq(const pcaone_ctx* ctx, double* F);
/* Data::standardize_E (src/Data.cpp:351-362) with scale -9. */
int pcaone_standardize(pcaone_ctx* ctx, int ploidy);
/* permute_matrix (src/RSVD.hpp:61-78): column i <- column perm[i]. */
int pcaone_permute_snps(pcaone_ctx* ctx, const int32_t* perm);
/* Data::fit_with_pi, EMU branch (src/Data.cpp:318-348): missing entries take
 * clamp(U S V^T, -F, 1-F) in every later product.  U N x k, S k, V M x k
 * (V in the current column order). */
int pcaone_set_impute(pcaone_ctx* ctx, const double* U, const double* S, const double* V, uint32_t k);

/* ---- RsvdOpData boundary (src/Halko.hpp:6-93) -------------------------- */
/* Workspace for L = k + oversamples columns in the given form. */
int pcaone_configure(pcaone_ctx* ctx, int precision, uint32_t L);
/* RsvdOpData::initOmg (src/Halko.cpp:15-23): Omega = Omg2 = omega (N x L). */
int pcaone_set_omega(pcaone_ctx* ctx, const double* omega);
/* The products of computeGandH for SNP rows [snp_begin, snp_end):
 * G rows = A Omega, then h_mode 0: H = A^T G (src/Halko.cpp:125-126),
 * 1: H1 += ..., 2: H2 += ... (src/Halko.cpp:195-203). */
int pcaone_sketch(pcaone_ctx* ctx, uint64_t snp_begin, uint64_t snp_end, int h_mode);
/* Omega <- thin-Q(src) then flipOmg (src/Halko.cpp:121-123, 212-215):
 * source 0: H, 1: H = H1 + H2.  CholeskyQR2 with Householder-convention
 * column signs. */
int pcaone_update_omega(pcaone_ctx* ctx, int source);
/* Omega <- thin-Q(src) without flipOmg: the out-of-core --svd 1 branch
 * (src/Halko.cpp:133-136). */
int pcaone_thin_q_omega(pcaone_ctx* ctx, int source);
int pcaone_zero_h(pcaone_ctx* ctx, int which); /* 1: H1, 2: H2 */
/* One epoch of computeUSV after computeGandH (src/Halko.cpp:55-77): QR of G
 * twice, B = R^-T H^T, SVD of B via B B^T, Ucur, diff = 1 - mev(Ucur, Upre)
 * when have_prev.  sigma: L values (nullable). */
int pcaone_finish_epoch(pcaone_ctx* ctx, uint32_t k, int have_prev, double* sigma, double* diff);
int pcaone_save_prev(pcaone_ctx* ctx); /* Upre = Ucur */
int pcaone_get_u(pcaone_ctx* ctx, uint32_t k, double* U); /* Ucur, N x k */
int pcaone_get_v(pcaone_ctx* ctx, uint32_t k, double* V); /* G U_B, M x k (src/Halko.cpp:84) */

/* The G of computeGandH (src/Halko.cpp:125, 195-197) for SNP rows
 * [snp_begin, snp_end) in the current column order, as last written by
 * pcaone_sketch and as multiplied by A^T (rounded to the form's operand):
 * (snp_end - snp_begin) x L.  Lets a caller keep the reference's host
 * computeUSV and replace only computeGandH. */
int pcaone_get_g(pcaone_ctx* ctx, uint64_t snp_begin, uint64_t snp_end, double* G);
/* H (which 0, src/Halko.cpp:126), H1 (1) or H2 (2) (src/Halko.cpp:198-203): N x L. */
int pcaone_get_h(pcaone_ctx* ctx, int which, double* H);

/* Data::write_residuals (src/Data.cpp:242-291): rows [snp_begin, snp_end)
 * (current column order) of the residual matrix A - U S V^T, each SNP
 * centered over samples, as float32 ((snp_end - snp_begin) x N, SNP-major: the
 * .residuals payload).  U N x k, S k, V M x k (current column order),
 * column-major; k = 0 subtracts nothing (--ld-stats 1).  as_read != 0: the
 * out-of-core branch (blocks read again centered, unscaled, not imputed);
 * otherwise A is the matrix the last products multiplied. */
int pcaone_residuals(pcaone_ctx* ctx, uint32_t k, const double* U, const double* S, const double* V, int as_read,
                     uint64_t snp_begin, uint64_t snp_end, float* out);

/* ---- drivers ------------------------------------------------------------ */
/* run_pca_with_halko (src/Halko.cpp:271-345) on the resident data.  U: N x k,
 * S and E (= S^2/M): k, V (nullable; filled with printv or emu): M x k in the
 * original SNP order. */
int pcaone_run_halko(pcaone_ctx* ctx, const pcaone_params* p, double* U, double* S, double* E, double* V,
                     pcaone_run_info* info);
/* diff values of every checked epoch / EM iteration of the last run. */
int pcaone_run_trace(const pcaone_ctx* ctx, int em, double* out, int cap, int* n);
/* The PLINK path of src/Main.cpp:150-231 for --svd 1/2: read <bfile>, run,
 * write <out>.sigvals/.eigvals/.eigvecs/.eigvecs2 (+ .loadings/.mbim with
 * printv). */
int pcaone_run_bfile(pcaone_ctx* ctx, const char* bfile, const char* out, const pcaone_params* p,
                     pcaone_run_info* info);
/* The same for --pgen <pfile> (src/Main.cpp:150-231 with FilePgen): the
 * .mbim takes the .pvar lines as .bim lines (src/Utils.cpp:405-430), the
 * .eigvecs2 the FID/IID of the .psam (src/Utils.cpp:240-265), and an -m run
 * with the winSVD shuffle takes compute_pgen_perm's order
 * (src/FilePgen.cpp:91-151, src/Main.cpp:169-172). */
int pcaone_run_pgen(pcaone_ctx* ctx, const char* pfile, const char* out, const pcaone_params* p,
                    pcaone_run_info* info);
/* Kernel timing with HIP events around the two products (on/off). */
int pcaone_time_kernels(pcaone_ctx* ctx, int on);

/* ---- batched form -------------------------------------------------------- */
/* Many independent cohorts (each its own genotype matrix, complete data) run
 * the --svd 1 randomized SVD together: equivalent to one run_pca_with_halko
 * (src/Halko.cpp:271-345, NormalRsvdOpData src/Halko.cpp:99-127) per cohort
 * with shared flags.  Every product of one step of all cohorts is one
 * persistent grouped launch; the L x L steps run on the device (L = k +
 * oversamples <= 32); the whole run is one HIP graph.  GEMM tile shapes and
 * split-K chunks are picked by timing candidates on the first run (unless
 * the wm_* / waves_* / kchunk_* fields of pcaone_params fix them). */
typedef struct pcaone_batch_info {
  double seconds;               /* device time of the run (launch to completion)   */
  int wm_g, waves_g, wm_h, waves_h; /* GEMM shapes used for G = A Omega, H = A^T G */
  uint64_t kchunk_g, kchunk_h;  /* split-K chunks used (codes)                      */
  double tuned_g_ms, tuned_h_ms; /* timed grouped launches of the chosen shapes    */
  int tuned_candidates;         /* configurations the tuner timed (0 = fixed)      */
  int failed;                   /* cohorts with a non-zero status                  */
} pcaone_batch_info;
int pcaone_batch_reserve(pcaone_ctx* ctx, uint32_t ncohorts, const uint32_t* nsamples, const uint64_t* nsnps);
/* Raw .bed rows of one cohort (ceil(N_c/4) bytes per SNP), then finalize. */
int pcaone_batch_upload(pcaone_ctx* ctx, uint32_t cohort, const uint8_t* rows, uint64_t snp_begin,
                        uint64_t nsnps);
int pcaone_batch_finalize(pcaone_ctx* ctx);
/* reserve + device generation + finalize, one spec per cohort. */
int pcaone_batch_generate(pcaone_ctx* ctx, uint32_t ncohorts, const pcaone_synth_spec* specs);
int pcaone_batch_allele_freq(pcaone_ctx* ctx, uint32_t cohort, double* F);
/* S, E: ncohorts x k (E = S^2 / M_c, src/Halko.cpp:339; zero past a cohort's
 * singular values); epochs, status: per cohort (status 0 or PCAONE_ERR_*).
 * use_graph: capture the run once and replay it; want_v: also loadings. */
int pcaone_batch_run(pcaone_ctx* ctx, const pcaone_params* p, int use_graph, int want_v, double* S, double* E,
                     int* epochs, int* status, pcaone_batch_info* info);
/* Results of the last batched run: U (N_c x k) and V (M_c x k), column-major;
 * the diff trace of the checked epochs (computeUSV's 1 - mev). */
int pcaone_batch_get_u(pcaone_ctx* ctx, uint32_t cohort, double* U);
int pcaone_batch_get_v(pcaone_ctx* ctx, uint32_t cohort, double* V);
int pcaone_batch_trace(const pcaone_ctx* ctx, uint32_t cohort, double* out, int cap, int* n);

/* ---- synthetic data ----------------------------------------------------- */
/* packed_out: nsnps x ceil(nsamples/4) bytes; attempts_out (nullable): per-SNP
 * redraw counter. */
int pcaone_synth_packed(const pcaone_synth_spec* spec, uint8_t* packed_out, uint32_t* attempts_out);
int pcaone_synth_write_plink(const pcaone_synth_spec* spec, const char* prefix);

/* ---- diagnostics (no reference counterpart) ---------------------------- */
/* Y = D(X) W with X raw .bed rows (nrows x ceil(ncodes/4) bytes), D mapping
 * codes 00/01/10/11 to 2/0/0/-2, W ncodes x L, Y nrows x L (column-major),
 * computed by the packed MFMA kernel of the given precision's form (any
 * PCAONE_PREC_* but the mixed BF16MIX and FP16MIX; INT8 takes at most 65536
 * codes). */
int pcaone_packed_gemm(pcaone_ctx* ctx, int precision, const uint8_t* rows, uint64_t nrows, uint32_t ncodes,
                       const double* W, uint32_t L, double* Y);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pcaone_packed_gemm on DEVICE pointers (rows: nrows x
 * ceil(ncodes/4) bytes, W: ncodes x L column-major FP64, Y: nrows x L
 * column-major FP64), with every intermediate in the library's fixed 64 MiB
 * device scratch array (one per GPU, held until the process ends); no
 * hipMalloc/hipFree.  A workspace beyond the scratch (about rows_pad x
 * ceil(ncodes/256) x 64 bytes + ncodes x L x 12 bytes + rows_pad x L x 4-8
 * bytes) is PCAONE_ERR_ARGUMENT.  Calls on one GPU share the scratch: calls
 * that overlap on one GPU race.  Runs on the context's stream and returns
 * after Y is written. */
int pcaone_packed_gemm_noalloc(pcaone_ctx* ctx, int precision, const uint8_t* rows, uint64_t nrows,
                               uint32_t ncodes, const double* W, uint32_t L, double* Y);
/* Data::prepare's out-of-core block plan (src/Data.cpp:42-84) for these
 * params (memory > 0): block size, block count, winSVD band factor and the
 * first SNP of each block (up to cap values into start, nullable). */
int pcaone_block_plan(const pcaone_params* p, uint32_t nsamples, uint64_t nsnps, uint64_t* blocksize,
                      uint64_t* nblocks, uint64_t* band_factor, uint64_t* start, uint64_t cap);
/* permute_plink's SNP order (src/FilePlink.cpp:303-408): position i of the
 * rewritten file holds original SNP order[i] (nsnps values). */
int pcaone_permute_plink_order(uint64_t nsamples, uint64_t nsnps, uint32_t buffer_gb, uint32_t bands,
                               int32_t* order);
/* compute_pgen_perm (src/FilePgen.cpp:91-151): the logical SNP order of an
 * out-of-core winSVD run on a .pgen: position i reads variant order[i]. */
int pcaone_pgen_perm_order(uint64_t nsnps, uint32_t nbatches, uint64_t blocksize, uint32_t nthreads, int seed,
                           int32_t* order);
/* RsvdOpData::initOmg's matrix (src/Halko.cpp:15-23): nrows x ncols,
 * column-major, identical to the libstdc++ draw, computed in parallel. */
int pcaone_host_omega(int seed, int gaussian, uint64_t nrows, uint64_t ncols, double* out);
/* Copies `bytes` (rounded down to 16) between two scratch device buffers
 * with a plain streaming kernel: a known-volume reference for HBM counters. */
int pcaone_copy_probe(pcaone_ctx* ctx, uint64_t bytes);
/* Host FP64 helpers of the L x L steps, row-major n x n. */
int pcaone_host_sym_eig(int n, const double* A, double* w, double* V);
int pcaone_host_householder_signs(int n, const double* Y, double* s);
int pcaone_host_chol_upper(int n, const double* A, double* R);
/* Thin Q of the Householder QR of X (nrows x ncols, row-major), in place: the
 * fallback of the thin-Q of H when H is near rank deficient (the reference's
 * HouseholderQR, src/Halko.cpp:121-122). */
int pcaone_host_householder_q(uint64_t nrows, int ncols, double* X);

#ifdef __cplusplus
}
#endif

#endif /* PCAONE_AMD_H_ */

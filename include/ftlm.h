/* ftlm.h -- C ABI of the MI355X FTLM engine (libftlm_hip.so).
 *
 * The reference (PeizhiMai/FTLM_extended_Hubbard_development) has no
 * library or FFI: its hot path is three functions of one translation unit,
 * called from the block loop of src/main.cpp:1519-1728. This header is the
 * boundary that replaces them; the drop-in executables ftlm_n_vs_mu and
 * ftlm_ed_n_vs_mu (host C++) call only what is declared here.
 *
 *   build_sector_basis   main.cpp:622-721  -> ftlm_sector
 *   build_momentum_block main.cpp:723-805  -> ftlm_block (+ exports)
 *   apply_hamiltonian    main.cpp:833-845  -> ftlm_apply
 *   run_ftlm_block       main.cpp:996-1113 -> ftlm_ftlm_block (ftlm_lanczos)
 *   jacobi_diagonalize   main.cpp:913-994  -> ftlm_ritz (device Jacobi)
 *   diagonalize_block_exact main.cpp:847-911 -> ftlm_exact_blocks
 *   make_block_seed      main.cpp:520-533  -> ftlm_block_seed
 *
 * Conventions: every call returns 0 (FTLM_OK) or a negative FTLM_ERR_*
 * code, with a message from ftlm_last_error(); calls are synchronous at
 * return; host buffers belong to the caller; a context is used from one host
 * thread at a time. Complex vectors are interleaved (re, im) doubles; R
 * vectors of a block are stored row-major [dim][R], i.e. element (i, r) at
 * 2*(i*R + r).
 */
#ifndef FTLM_H
#define FTLM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTLM_OK 0
#define FTLM_ERR_ARG (-1)     /* invalid argument */
#define FTLM_ERR_DEVICE (-2)  /* HIP runtime / no usable GPU */
#define FTLM_ERR_MEMORY (-3)  /* device or host allocation failed */
#define FTLM_ERR_STATE (-4)   /* call order (no sector / block selected) */
#define FTLM_ERR_LIMIT (-5)   /* size beyond what the engine supports */

/* reduction orders of the Lanczos dot products and norms */
#define FTLM_REDUCE_SERIAL 0  /* row order, bit-identical to the reference */
#define FTLM_REDUCE_TREE 1    /* fused workgroup tree, deterministic */

typedef struct ftlm_ctx ftlm_ctx;

/* Model of Params (main.cpp:32-61): lattice and Hamiltonian parameters. */
typedef struct ftlm_model {
  int32_t lx, ly;
  double tx, ty, tp, phix, phiy, u, v;
} ftlm_model;

/* Device-event timings of the last Lanczos run on a context (ftlm_run, or
 * the batched runs of the last ftlm_ftlm_block call). */
typedef struct ftlm_stats {
  double ms_apply;     /* summed H-apply kernel time */
  double ms_total;     /* first to last kernel of the run */
  int64_t n_apply;     /* H-apply launches */
  double bytes_apply;  /* algorithmic bytes of those launches */
} ftlm_stats;

const char* ftlm_last_error(void);

/* SplitMix64-based block seed (main.cpp:514-533). */
uint64_t ftlm_block_seed(uint64_t seed, int nup, int ndn, int mx, int my);

/* Create a context on HIP device `device` (fails with FTLM_ERR_DEVICE when
 * no GPU is usable: there is no CPU fallback). */
int ftlm_create(const ftlm_model* model, int device, ftlm_ctx** out);
void ftlm_destroy(ftlm_ctx* ctx);

/* Select the (N_up, N_down) sector and build its tables on the device.
 * full_dim = C(L,Nup) C(L,Ndn); parents = number of translation orbits
 * (the reference's SectorBasis::parents.size()). Either out may be NULL. */
int ftlm_sector(ftlm_ctx* ctx, int nup, int ndn, uint64_t* full_dim, int64_t* parents);

/* Select momentum block (mx, my) of the current sector; dim = active rows. */
int ftlm_block(ftlm_ctx* ctx, int mx, int my, int64_t* dim);

/* Representatives (up, down bitstrings), degeneracy and norm factor of each
 * row of the current block, in the reference's row order. */
int ftlm_block_reps(ftlm_ctx* ctx, uint64_t* up, uint64_t* dn, int32_t* deg, double* dnf);

/* CSR of the current block as the reference stores it (row = source state,
 * entries in hop order). all_dirs != 0 keeps zero-amplitude directions like
 * the reference; 0 drops them. Call with col == NULL to get nnz only. */
int ftlm_block_csr(ftlm_ctx* ctx, int all_dirs, int64_t* nnz, int32_t* row_ptr, int32_t* col, double* val,
                   double* diag);

/* y = H x for R vectors of the current block (host buffers, 2*dim*R each).
 * kernel: 0 = the context's apply configuration, 1 = the on-the-fly
 * one-row-per-thread kernel (R = 1). */
int ftlm_apply(ftlm_ctx* ctx, int R, int kernel, const double* x, double* y);

/* Seeded unit-norm complex Gaussian start vector of one sample
 * (main.cpp:1015-1025), 2*dim doubles, from the context's start source. */
int ftlm_start_vector(ftlm_ctx* ctx, uint64_t block_seed, int sample, double* out);

/* Where the seeded start vectors of every later Lanczos run, FTLM block and
 * pack entry of the context are drawn (no reference counterpart; the stream
 * is the reference's in both cases, main.cpp:1015-1025):
 *   FTLM_START_HOST    libstdc++ mt19937_64 + normal_distribution on host
 *                      threads, bit-identical to the reference (default);
 *   FTLM_START_DEVICE  the same engine and polar method on the GPU, written
 *                      straight into device memory; elements differ from the
 *                      host stream only in the last bit where the device log
 *                      rounds differently from glibc's, and the normalisation
 *                      is summed in tree order. */
#define FTLM_START_HOST 0
#define FTLM_START_DEVICE 1
int ftlm_set_start_source(ftlm_ctx* ctx, int source);

/* Lanczos coefficients of R vectors on the current block.
 * start: host [dim][R] complex, or NULL for the seeded vectors of samples
 * 0..R-1. per_vector != 0 runs each vector alone (R = 1 kernels).
 * alpha, beta: [steps][R] (beta has m-1 valid rows per column); m: [R]. */
int ftlm_lanczos(ftlm_ctx* ctx, int R, int steps, const double* start, uint64_t block_seed, int reduce,
                 int per_vector, double* alpha, double* beta, int32_t* m);

/* Ritz values (ascending) and weights v1^2 / wdiv of R tridiagonal matrices
 * on the device. method 0: the reference's Jacobi, bit-identical
 * (jacobi_diagonalize + std::sort ordering); method 1: implicit QL carrying
 * the first eigenvector row (used by tree-order runs). alpha/beta:
 * [steps][R]; eig, w: [R][steps]. */
int ftlm_ritz(ftlm_ctx* ctx, int R, int steps, const double* alpha, const double* beta, const int32_t* m,
              int wdiv, int method, double* eig, double* w);

/* run_ftlm_block: R seeded samples, `steps` Lanczos steps, Ritz values and
 * weights concatenated in sample order (eig/w hold R*min(steps, dim)).
 * R <= 1024; R above 256 runs in batches of 256 samples (one batched launch
 * holds at most 256 columns; ftlm_apply, ftlm_lanczos without per_vector,
 * ftlm_prepare and ftlm_autotune take at most 256 and return FTLM_ERR_LIMIT
 * beyond). */
int ftlm_ftlm_block(ftlm_ctx* ctx, int R, int steps, uint64_t block_seed, int reduce, int per_vector,
                    double* eig, double* w, int64_t* count);

/* Benchmark split of ftlm_ftlm_block: upload the seeded start vectors, then
 * run the Lanczos recurrence on resident data (timed with device events, see
 * ftlm_last_stats), then solve the tridiagonal matrices. */
int ftlm_prepare(ftlm_ctx* ctx, int R, uint64_t block_seed);
int ftlm_run(ftlm_ctx* ctx, int R, int steps, int reduce, int per_vector_column);
int ftlm_finish(ftlm_ctx* ctx, int R, int steps, double* eig, double* w, int64_t* count);
int ftlm_last_stats(ftlm_ctx* ctx, ftlm_stats* out);

/* H-apply configuration (no reference counterpart; engine tuning knobs).
 * path: 0 automatic (= 1), 1 hop list -- the block's hops (target row and
 * coefficient code, 8 bytes each) are built once on the device and every
 * apply is a pure gather -- or 2 on the fly -- hops are recomputed from the
 * basis tables inside each apply. W: -1 automatic, 0 untiled, W > 0 column
 * tiles of W vectors, at most 256 (hop list) or 128 (on the fly; larger W
 * returns FTLM_ERR_LIMIT). ftlm_get_apply_config reports what R vectors on the
 * current block use. */
int ftlm_set_apply_config(ftlm_ctx* ctx, int path, int W);
int ftlm_get_apply_config(ftlm_ctx* ctx, int R, int* path, int* W);

/* Placement of the Lanczos vectors (no reference counterpart; an engine
 * tuning knob). The batched hop-list gather's speed depends on where in HBM
 * the gathered vector physically lies (up to ~13 % between buffers of one
 * process on the 4x4 (8,8) block). When the three R-column vectors are
 * (re)allocated at min_bytes or more, the next batched Lanczos run times one
 * apply from `candidates` buffers (the three plus extra ones that free memory
 * allows, minus a reserve of max(4 GiB, 5 % of the device)), keeps the
 * fastest two as the buffers the recurrence alternates as the gather source,
 * then the remaining buffer into which applies from those two are fastest as
 * the output, and frees the rest. candidates = 0 turns it off, else 3..16 (default 8,
 * min_bytes 256 MiB). Results do not depend on the placement.
 * ftlm_vector_placement_report: the last tuning's apply ms per candidate as
 * the gather source (on input *n = capacity of ms; on output the number of
 * candidates, 0 if none ran) and, when pick is not NULL, the 3 chosen
 * candidates (the two gather sources, then the output). */
int ftlm_set_vector_placement(ftlm_ctx* ctx, int candidates, uint64_t min_bytes);
int ftlm_vector_placement_report(ftlm_ctx* ctx, int* n, double* ms, int* pick);

/* Autotuner: times the candidate configurations (hop list untiled and with
 * tiles 64/32/16/8 below R, on the fly untiled) on the current block with R
 * vectors and keeps the fastest. On input *ncand is the capacity of the
 * cand_* arrays (which may be NULL); on output the number timed. */
int ftlm_autotune(ftlm_ctx* ctx, int R, int* best_path, int* best_w, double* best_ms, int* ncand, int* cand_path,
                  int* cand_w, double* cand_ms);

/* Twist of the context (Params phix/phiy, main.cpp:544-577): blocks selected
 * afterwards use the new hop phases. Sector tables do not depend on the twist
 * and are kept; the current block must be selected again (ftlm_block). */
int ftlm_set_twist(ftlm_ctx* ctx, double phix, double phiy);

/* Packed FTLM over many blocks of any sector and twist (no reference
 * counterpart: the block loop main.cpp:1519-1581 runs run_ftlm_block block by
 * block; here the uneven small blocks of a sweep are batched into one
 * persistent launch, replayed through a HIP graph).
 *   ftlm_pack_begin   empties the pack: R samples and `steps` Lanczos steps
 *                     per block, device memory budget in bytes (0: 64 GiB).
 *   ftlm_pack_add     adds the current block with its block seed (hop list
 *                     built on the device, start vectors generated on the
 *                     host as run_ftlm_block draws them); *index receives its
 *                     position. Returns FTLM_PACK_FULL, adding nothing, when
 *                     the block does not fit the budget: run the pack first.
 *   ftlm_pack_run     runs every block. Work items are (block, `width`
 *                     columns), width 4/8/16/32/64 or 0 (tuned); `wg_per_cu`
 *                     workgroups per CU (0: tuned). Each workgroup runs all
 *                     Lanczos steps of its item with row-order reductions, so
 *                     alpha/beta are bit-identical to the reference for every
 *                     width; method 0 = Jacobi Ritz (bit-identical to
 *                     run_ftlm_block), 1 = implicit QL.
 *   ftlm_pack_result  Ritz values and weights of entry `index`, as
 *                     ftlm_ftlm_block returns them.
 *   ftlm_pack_tune    times the (width, wg_per_cu) candidates on the current
 *                     pack over `tune_steps` steps and keeps the fastest for
 *                     later runs with 0 arguments (widths 8, 16, 32, 64 with
 *                     1 and 2 workgroups per CU, widening only while the
 *                     wider width wins); a large pack is timed on
 *                     every k-th entry in decreasing cost, k chosen so that
 *                     the sample still holds 8 * CUs column groups of 64
 *                     (best_ms and cand_ms are those of the sample).
 */
#define FTLM_PACK_FULL 1
int ftlm_pack_begin(ftlm_ctx* ctx, int R, int steps, uint64_t budget_bytes);
int ftlm_pack_add(ftlm_ctx* ctx, uint64_t block_seed, int64_t* index);
int ftlm_pack_run(ftlm_ctx* ctx, int width, int wg_per_cu, int method);
int ftlm_pack_result(ftlm_ctx* ctx, int64_t index, double* eig, double* w, int64_t* count);
int ftlm_pack_tune(ftlm_ctx* ctx, int tune_steps, int* best_width, int* best_wg_per_cu, double* best_ms,
                   int* ncand, int* cand_width, int* cand_wg, double* cand_ms);
/* entries in the pack and device bytes they hold */
int ftlm_pack_info(ftlm_ctx* ctx, int64_t* entries, uint64_t* bytes);

/* diagonalize_block_exact (main.cpp:847-911) deferred and batched across
 * sectors and twists: the reference diagonalises each exact block where the
 * block loop meets it; here each block's Hermitised dense matrix is built on
 * the device when it is added (from the current sector and twist) and one
 * launch solves every queued block.
 *   ftlm_exact_queue_begin   empties the queue; budget = bytes of matrices
 *                            (0: 16 GiB)
 *   ftlm_exact_queue_add     queues block (mx, my) of the current sector at
 *                            the current twist; *index its position; returns
 *                            FTLM_PACK_FULL, adding nothing, when it does not
 *                            fit the budget (run the queue first)
 *   ftlm_exact_queue_run     solves every block queued since the last run
 *                            (earlier results stay available)
 *   ftlm_exact_queue_result  ascending eigenvalues of entry `index` (*count
 *                            of them) */
int ftlm_exact_queue_begin(ftlm_ctx* ctx, uint64_t budget_bytes);
int ftlm_exact_queue_add(ftlm_ctx* ctx, int mx, int my, int64_t* index);
int ftlm_exact_queue_run(ftlm_ctx* ctx);
int ftlm_exact_queue_result(ftlm_ctx* ctx, int64_t index, double* eig, int64_t* count);

/* Grand-canonical sums of many blocks' spectra for every (beta, mu) pair in
 * one FP64 device launch: accumulate_block_thermo (main.cpp:1172-1219) over
 * all blocks, with the final running minimum known up front. Block b owns
 * values off[b]..off[b+1]-1 of eig and w (exact blocks: w = 1, prefactor 1;
 * FTLM blocks: Ritz weights, prefactor = block dimension), particles[b] its
 * particle number (0..255). For pair (i, k) = (betas[i], mu[k]):
 *   emin = min over values of eig - mu * N,
 *   z = sum prefactor * w * exp(-beta (eig - mu N - emin)), n1 = sum N * .,
 *   n2 = sum N^2 * .;
 * out[(i * nmu + k) * 4 + 0..3] = emin, z, n1, n2 (the reference's streaming
 * sums up to summation-order rounding; finalise as main.cpp:1770-1788). */
int ftlm_thermo(ftlm_ctx* ctx, int64_t nblk, const int64_t* off, const double* eig, const double* w,
                const int32_t* particles, const double* prefactor, int nbeta, const double* betas, int nmu,
                const double* mu, double* out);

/* ftlm_thermo without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The spectra are already in device memory, one entry per value: d_eig
 * (Ritz value or eigenvalue), d_weight (its block's prefactor times its
 * weight: dim * w for FTLM blocks, 1 for exact ones) and d_particles (its
 * block's particle number); d_betas, d_mu and d_out (nbeta * nmu * 4
 * doubles, laid out as in ftlm_thermo) are device pointers too. The running
 * minimum of E - mu N is found on the device, so d_out equals ftlm_thermo's
 * output for the same spectra bit for bit. Partial sums live in a fixed
 * 32 MiB device array (one per GPU, kept until the process ends): calls with
 * 3 * ceil(nv / 4096) * nbeta * nmu + (ceil(nv / 4096) + 1) * nmu doubles
 * above 4 Mi return FTLM_ERR_LIMIT. Calls on one GPU share that array, so
 * calls that overlap on one GPU (other host threads, other streams) race.
 * stream: NULL runs on the context's stream and returns when done; a
 * hipStream_t enqueues the work there and returns at once. */
int ftlm_thermo_noalloc(ftlm_ctx* ctx, int64_t nv, const double* d_eig, const double* d_weight,
                        const uint8_t* d_particles, int nbeta, const double* d_betas, int nmu, const double* d_mu,
                        double* d_out, void* stream);

/* diagonalize_block_exact for blocks (mx[i], my[i]) of the current sector,
 * batched in one launch: eig_out receives the ascending eigenvalues of each
 * block back to back, offsets[i]..offsets[i+1]. */
int ftlm_exact_blocks(ftlm_ctx* ctx, int nblk, const int32_t* mx, const int32_t* my, double* eig_out,
                      int64_t* offsets);

/* Memory-system probe for the benchmark (no reference counterpart): what the
 * GPU's HBM path delivers for the batched apply's access shapes without its
 * hop metadata and arithmetic, on `rows` rows of `width` (<= 256) complex
 * doubles. kind 0 streams the rows in (read only), 1 copies them, 2 writes
 * every row as the sum of `k` source rows chosen by a fixed hash (one
 * wavefront per row, like the batched apply); 3 draws the k rows inside the
 * row's window of 12,870 rows (one canonical-up group of the half-filled
 * 4x4 block), 4 alternately inside the window and over all rows. Allocates
 * two buffers of rows * width * 16 bytes for the call, times `reps` launches
 * after one warm-up launch; *ms_per_launch and *bytes_per_launch (read +
 * written bytes: rows * width * 16 times 1, 2 or k + 1) are returned. */
int ftlm_memprobe(ftlm_ctx* ctx, int kind, int64_t rows, int width, int k, int reps, double* ms_per_launch,
                  double* bytes_per_launch);

/* ---- Learned surrogate (no reference counterpart) -------------------------
 * An MLP maps (U, V, t', phix, phiy, beta, mu) to (n, compressibility), the
 * two observables the reference's CSV reports per mu (main.cpp:1409-1455),
 * so that dense parameter and twist scans can be filled between FTLM runs;
 * FTLM stays the judge and the fallback. Inputs are rows of 7 doubles in that
 * order, outputs rows of 2 doubles. The network: 16 encoded features (zero
 * padded to FTLM_MLP_FEATURES) -> 4 hidden layers of 256 GELU units -> 2
 * outputs (padded to 32), output c = y_mean[c] + y_scale[c] * raw output.
 * Weights are bf16 copies of fp32 master parameters; every product runs on
 * the MFMA units (bf16 x bf16, fp32 accumulation). Training minimises the
 * mean squared error of the normalised outputs with Adam.
 *   ftlm_mlp_param_count / ftlm_mlp_layout  flat fp32 parameter vector:
 *                     W_0, b_0, ..., W_4, b_4, W_l row-major [n_out x n_in]
 *   ftlm_mlp_encode   the feature rows (host only, no device needed)
 *   ftlm_mlp_create   needs a GPU; then ftlm_mlp_set_params before use
 *                     (also clears the Adam state and step counter)
 *   ftlm_mlp_set_data uploads a training set (resident on the device)
 *   ftlm_mlp_train    `steps` Adam steps over seeded random batches of
 *                     cfg.batch rows (epochs of random permutations), one
 *                     HIP graph replay per step; loss[steps] receives each
 *                     step's loss (may be NULL)
 *   ftlm_mlp_gradients  gradient (flat, like the parameters) and loss of n
 *                     rows without updating
 *   ftlm_mlp_step     one Adam step on n given rows
 *   ftlm_mlp_predict  n rows of (n, compressibility)
 *   ftlm_mlp_gemm     test hook: C[M x N] = bf16(A[M x K]) bf16(B[N x K])^T in
 *                     fp32 through the network's MFMA GEMM kernel */
#define FTLM_MLP_INPUTS 7
#define FTLM_MLP_FEATURES 32
#define FTLM_MLP_LAYERS 5
typedef struct ftlm_mlp ftlm_mlp;
typedef struct ftlm_mlp_config {
  int32_t batch;                 /* rows per training step, a multiple of 64 */
  float lr, beta1, beta2, eps;   /* Adam */
  double y_mean[2], y_scale[2];  /* output normalisation */
} ftlm_mlp_config;
int64_t ftlm_mlp_param_count(void);
int ftlm_mlp_layout(int64_t* w_offsets, int64_t* b_offsets, int32_t* n_out, int32_t* n_in);
int ftlm_mlp_encode(int64_t n, const double* x, float* feat);
int ftlm_mlp_create(const ftlm_mlp_config* cfg, int device, ftlm_mlp** out);
void ftlm_mlp_destroy(ftlm_mlp* m);
int ftlm_mlp_set_params(ftlm_mlp* m, const float* params);
int ftlm_mlp_get_params(ftlm_mlp* m, float* params);
int ftlm_mlp_set_lr(ftlm_mlp* m, float lr);
int ftlm_mlp_set_data(ftlm_mlp* m, int64_t n, const double* x, const double* y);
int ftlm_mlp_train(ftlm_mlp* m, int steps, uint64_t seed, float* loss);
int ftlm_mlp_gradients(ftlm_mlp* m, int64_t n, const double* x, const double* y, float* grad, float* loss);
int ftlm_mlp_step(ftlm_mlp* m, int64_t n, const double* x, const double* y, float* loss);
int ftlm_mlp_predict(ftlm_mlp* m, int64_t n, const double* x, double* y);
int ftlm_mlp_gemm(ftlm_mlp* m, int M, int N, int K, const float* a, const float* b, float* c);

#ifdef __cplusplus
}
#endif

#endif /* FTLM_H */

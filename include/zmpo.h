/* libzmpo -- C-ABI of the MI355X two-site sigma / renormalisation path.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses;
 * `stream` is a hipStream_t passed as void* (NULL = default stream).  Every
 * function returns ZMPO_OK (0) or a negative ZMPO_ERR_* code, and
 * zmpo_last_error() returns the message of the calling thread's last error.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   zmpo_symAllowed      ctypes/qsym.c:5-22 (symAllowed), bound by ctypes in
 *                        source/qtensor/qtensor_util.py:70-90 (blks_allowed1).
 *   zmpo_symAllowedInt   integer-QN variant of the same rule (no counterpart;
 *                        QNs as (N, 2Sz) int32, exact equality).
 *   zmpo_blockgemm       the numeric loop of qtensor.tensordotCAL
 *                        (source/qtensor/qtensor.py:642-659): every block
 *                        pair of one or more tensordotSYM plans
 *                        (qtensor.py:563-637) in one launch; used for the four
 *                        sigma stages of HVecQt (source/mpo_dmrg_kernelQt.py:
 *                        270-306), the three renormalisation stages of
 *                        renormQt_H00 (kernelQt.py:674-683, 752-758) and
 *                        genHopsQt (source/mpo_dmrg_initQt.py:146-152), the
 *                        final contraction of HDiagQt (kernelQt.py:116-119),
 *                        and the three stages of the energy gradient dE/dW
 *                        behind the RDMs (properties/mpo_dmrg_propsItrf.py:
 *                        204-263, computed as a derivative here).
 *   zmpo_wavegemm        the same numeric loop, scheduled one wave per small
 *                        output tile (the block shapes of block-sparse
 *                        DMRG are mostly 5-150 wide).
 *   zmpo_shapedgemm, zmpo_streamgemm, zmpo_skinnygemm, zmpo_dotgemm
 *                        the same numeric loop with LDS-staged block-shaped
 *                        tiles (one tile per workgroup / unit ranges per
 *                        workgroup / tall narrow blocks / K-contiguous
 *                        operands with the k range split over the waves).
 *   zmpo_chaingemm, zmpo_chaingemm16
 *                        the two middle stages of HVecQt (kernelQt.py:
 *                        277-296, top1 x W[i] then x W[i+1]) as one launch
 *                        with the intermediate top2 kept on chip (32- or
 *                        16-column output tiles).
 *   zmpo_dots, zmpo_lincomb, zmpo_precond
 *                        the vector algebra of the Davidson solver
 *                        (source/tools/pdvdson.py:131-283, 350-390), which the
 *                        reference runs with numpy on rank 0.
 *   zmpo_gather, zmpo_scatter, zmpo_convert
 *                        qtensor.fromDense / toDenseTensor / prjmap index
 *                        moves (qtensor.py:364-398, 442-453) and precision
 *                        casts.
 */
#ifndef ZMPO_H
#define ZMPO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZMPO_ABI_VERSION 1

#define ZMPO_OK 0
#define ZMPO_ERR_ARG (-1)
#define ZMPO_ERR_HIP (-2)

#define ZMPO_F64 0
#define ZMPO_F32 1
/* BF16 operands (round to nearest even), FP32 accumulation on the BF16
 * MFMA; ZMPO_BF16 writes BF16 outputs, ZMPO_BF16_F32 FP32 outputs and
 * split-K workspace.  Accepted by zmpo_shapedgemm, zmpo_streamgemm(_static)
 * (both codes), zmpo_chaingemm(16) (ZMPO_BF16) and zmpo_convert (ZMPO_BF16
 * to and from FP64 / FP32; FP64 rounds through FP32). */
#define ZMPO_BF16 2
#define ZMPO_BF16_F32 3

/* One output block of a grouped block GEMM.  Output element (m, n) of the
 * block lives at c_off + (m / c_mlo) * c_smhi + (m % c_mlo) * c_smlo
 * + n * c_sn.  tile_base is the exclusive prefix sum of the 64x64 tiles of
 * the preceding tasks; tiles_n = ceil(N / 64). */
typedef struct zmpo_task {
  int64_t c_off;
  int64_t tile_base;
  int32_t M, N;
  int32_t c_mlo, c_smhi, c_smlo, c_sn;
  int32_t p_begin, p_end;
  int32_t tiles_n, pad;
} zmpo_task;

/* One block pair: C_block += scale * A(M x K) * B(K x N), with
 * A(m, k) = A[a_off + (m / a_mlo) * a_smhi + (m % a_mlo) * a_smlo + k * a_sk]
 * B(k, n) = B[b_off + (n / b_nlo) * b_snhi + (n % b_nlo) * b_snlo + k * b_sk]. */
typedef struct zmpo_pair {
  int64_t a_off, b_off;
  double scale;
  int32_t K;
  int32_t a_mlo, a_smhi, a_smlo, a_sk;
  int32_t b_nlo, b_snhi, b_snlo, b_sk;
  int32_t pad;
} zmpo_pair;

/* One wave's work item for zmpo_wavegemm: the (16*rb) x 16 tile at (m0, n0)
 * of task `task`, over pairs [p_begin, p_end).  slot < 0 writes the tile to C;
 * slot >= 0 writes it to work[slot * rb * 256 ...] for zmpo_rgroup reduction. */
typedef struct zmpo_unit {
  int32_t task, m0, n0, p_begin, p_end, slot, pad0, pad1;
} zmpo_unit;

/* Sum of work slots [slot_begin, slot_begin + nslots) into the tile (m0, n0)
 * of task `task`, in slot order. */
typedef struct zmpo_rgroup {
  int32_t task, m0, n0, slot_begin, nslots, pad;
} zmpo_rgroup;

/* Host-side schedule builder for the grouped-GEMM launches (host arrays
 * only, no device work): the units of a TM x TN tiling of every task, one
 * per (tile, chunk of the task's pairs), chunks cut where the summed K of
 * the task's preceding pairs crosses a multiple of kchunk; order (task,
 * tile, chunk), or (key, chunk, task, tile) with a per-task reuse_key
 * (NULL: none); split-K slots numbered in (task, tile, chunk) order over
 * the split tasks; one zmpo_rgroup per split tile, in unit order.  With
 * units == NULL only counts[0..2] = (units, groups, slots) are set.  Replaces
 * the per-pair Python loop that drives the reference's block products
 * (qtensor.tensordotCAL, source/qtensor/qtensor.py:642-659) with work lists
 * built once per plan. */
int zmpo_schedule_build(int64_t nt, const zmpo_task* tasks, const zmpo_pair* pairs, int tm,
                        int tn, int64_t kchunk, const int64_t* reuse_key, zmpo_unit* units,
                        int64_t units_cap, zmpo_rgroup* groups, int64_t groups_cap,
                        int64_t* counts);

const char* zmpo_last_error(void);
int zmpo_abi_version(void);
int zmpo_struct_sizes(int* task_size, int* pair_size);

/* Block selection rule with the reference's argument list (qsym.c:5-7):
 * allowed[b] = (sum_{status=0} q - sum_{status=1} q has L1 norm < 1e-10)
 * for every block b (C order over ndims).  QN arrays are (ndims[i], nqnum)
 * row-major float64.  rank in [2, 8]; unlike the reference (qsym.c:436) the
 * rank-8 case loops its last index over ndims[7]. */
int zmpo_symAllowed(int rank, int nqnum, int* ndims, int* status, int* allowed, double* a0,
                    double* a1, double* a2, double* a3, double* a4, double* a5, double* a6,
                    double* a7);

/* Same rule on integer QNs (exact equality), rank in [1, 16]. */
int zmpo_symAllowedInt(int rank, int nqnum, const int* ndims, const int* status, int* allowed,
                       const int32_t* const* qsyms);

/* Grouped block GEMM (device arrays tasks[ntasks], pairs[...]; ntiles = total
 * tiles).  beta = 0 overwrites each output block, 1 accumulates into it.
 * dtype ZMPO_F64 (FP64 MFMA) or ZMPO_F32 (FP32 MFMA); A, B, C share it. */
int zmpo_blockgemm(int dtype, const zmpo_task* tasks, int64_t ntasks, const zmpo_pair* pairs,
                   int64_t ntiles, const void* A, const void* B, void* C, int beta, void* stream);

/* Small-block grouped GEMM: one wave per zmpo_unit (rb in {1, 2, 4} row
 * blocks of 16), then the deterministic reduction of split tiles.  Same
 * task/pair semantics as zmpo_blockgemm; work holds the split tiles (T). */
int zmpo_wavegemm(int dtype, int rb, const zmpo_task* tasks, const zmpo_unit* units,
                  int64_t nunits, const zmpo_rgroup* groups, int64_t ngroups,
                  const zmpo_pair* pairs, const void* A, const void* B, void* C, void* work,
                  int beta, void* stream);

/* Tile shapes of zmpo_shapedgemm (rows x columns of one workgroup tile). */
#define ZMPO_TILE_64x16 0
#define ZMPO_TILE_128x16 1
#define ZMPO_TILE_32x32 2
#define ZMPO_TILE_64x32 3
#define ZMPO_TILE_16x16 4
#define ZMPO_TILE_64x64 5
/* Register-blocked panel tiles (same units, split-K groups and results up
 * to the order of the k-sum inside a k-tile): the 4 waves of a workgroup
 * form a 2x2 grid of (TM/2)x(TN/2) register blocks; operand tiles sit in
 * LDS with k contiguous (16 k per tile for FP64/FP32, 32 for BF16, which
 * runs on the gfx950 16x16x32 BF16 MFMA).  For blocks of ~100 and more. */
#define ZMPO_PANEL_64x64 6
#define ZMPO_PANEL_128x64 7
#define ZMPO_PANEL_64x128 8
#define ZMPO_PANEL_128x128 9
#define ZMPO_PANEL_32x64 10
#define ZMPO_PANEL_64x32 11
/* 32 x 160: one row band of a ~130-160 wide output block per workgroup, so
 * the A rows are read once (sigma x R).  Only for programs whose every pair
 * has a_sk == 1 and b_sk == 1 (k-contiguous operands); other programs give
 * wrong results with it. */
#define ZMPO_PANEL_32x160 12

/* LDS-staged grouped GEMM with block-shaped tiles: one workgroup per
 * zmpo_unit (unit m0/n0 step by the tile shape), same split-K groups and
 * task/pair semantics as zmpo_wavegemm; work holds TM*TN elements per slot. */
int zmpo_shapedgemm(int dtype, int shape, const zmpo_task* tasks, const zmpo_unit* units,
                    int64_t nunits, const zmpo_rgroup* groups, int64_t ngroups,
                    const zmpo_pair* pairs, const void* A, const void* B, void* C, void* work,
                    int beta, void* stream);

/* One output tile of the streaming schedule, with the geometry of its task
 * folded in (no dependent task lookup on the device).  Rows r < M and
 * columns c < N of the block are real; the tile covers rows [m0, m0+TM) and
 * columns [n0, n0+TN).  slot >= 0 sends the tile to workspace slot `slot`
 * (split K) instead of C. */
typedef struct zmpo_sunit {
  int64_t c_off, c_smhi;
  int32_t c_mlo, c_smlo, c_sn;
  int32_t m0, n0, M, N;
  int32_t p_begin, p_end, slot;
  int32_t pad0, pad1;
} zmpo_sunit;

/* Streaming grouped GEMM: workgroup w runs units [w*upw, (w+1)*upw) (upw in
 * [1, 64]) as one software pipeline over all their (pair, k-tile) items --
 * the operand tiles of the next item are in flight while the current one
 * is multiplied and while finished tiles are written -- with k-tiles of kt
 * (16 or 32) and the tile shapes of zmpo_shapedgemm.  Split-K slots are
 * summed afterwards by the zmpo_shapedgemm reduction (groups over tasks).
 * Same pair semantics and results as zmpo_shapedgemm. */
int zmpo_streamgemm(int dtype, int shape, int kt, const zmpo_sunit* units, int64_t nunits,
                    int upw, const zmpo_task* tasks, const zmpo_rgroup* groups,
                    int64_t ngroups, const zmpo_pair* pairs, const void* A, const void* B,
                    void* C, void* work, int beta, void* stream);

/* Grouped GEMM for small output blocks with a very long summed K and both
 * operands K-contiguous (sigma stage top3 x R of HVecQt, mpo_dmrg_kernelQt.
 * py:297-306): the zmpo_unit / split-K schedule of zmpo_shapedgemm with the
 * 32x32 tile (ZMPO_TILE_32x32), work holding 32*32 elements per slot.  The
 * four waves of a workgroup split each pair's k range and their partial
 * tiles are summed in a fixed order. */
int zmpo_dotgemm(int dtype, const zmpo_task* tasks, const zmpo_unit* units, int64_t nunits,
                 const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                 const void* A, const void* B, void* C, void* work, int beta, void* stream);

/* Grouped GEMM for tall, narrow blocks (sigma stage L x psi of HVecQt,
 * mpo_dmrg_kernelQt.py:270-276, with the same task/pair semantics as
 * zmpo_blockgemm): one workgroup per zmpo_unit = rows [m0, m0 + 64*rb) x
 * columns [n0, n0 + 64) of task `task` over pairs [p_begin, p_end) (slot
 * unused: every unit owns its output).  Requirements (checked by the
 * caller): every pair's A rows are single-level (a_mlo >= M) with a_sk == 1.
 * rb in {1, 2}. */
int zmpo_skinnygemm(int dtype, int rb, const zmpo_task* tasks, const zmpo_unit* units,
                    int64_t nunits, const zmpo_pair* pairs, const void* A, const void* B, void* C,
                    int beta, void* stream);

/* Two chained grouped GEMMs over row groups, C = (A B1) B2 blockwise, for
 * the middle sigma stages of HVecQt (mpo_dmrg_kernelQt.py:277-296: top1 x
 * W[i] -> top2, top2 x W[i+1] -> top3).  All blocks of one row group share
 * their M rows (the bra and right-bond sectors), so for a tile of 16*rb rows
 * the group's distinct A blocks (`kcols` columns) are staged into LDS once
 * and its whole intermediate (every first-stage output block, `ncols`
 * columns) stays in LDS and feeds the second stage; it never touches HBM.
 *   A blocks [ab_begin, ab_end): zmpo_cblock, offsets relative to a_base;
 *   first-stage tasks  [t1_begin, t1_end): c_off = intermediate column
 *     offset, N = its width (ceil4(N) columns reserved); pairs: a_off =
 *     A-tile column offset (A maps unused), b_off relative to b1_base;
 *   second-stage tasks [t2_begin, t2_end): c_off and C maps relative to
 *     c_base (c_mlo == rsplit); pairs: a_off = intermediate column offset,
 *     b_off relative to b2_base.
 *   B blocks are zero-padded row-major matrices [ceil4(K)][ceil32(N)] with
 *   row stride b_sk (so operand loads need no bounds tests); pair scales
 *   are 1.
 * One workgroup per zmpo_cunit (rows [m0, m0 + 16*rb) of group `group`);
 * each output element is written once (no atomics). */
typedef struct zmpo_cgroup {
  int32_t M, ncols, kcols, t1_begin, t1_end, t2_begin, t2_end, ab_begin, ab_end, rsplit;
} zmpo_cgroup;

/* One A block of a row group: rows m map to (m / rsplit) * a_smhi +
 * (m % rsplit) * a_smlo (a_mlo == the group's rsplit), column k to k * a_sk;
 * staged at A-tile columns [col, col + ceil4(K)) (padding columns zero). */
typedef struct zmpo_cblock {
  int64_t a_off, a_smhi, a_smlo, a_sk;
  int32_t K, col, a_mlo, pad;
} zmpo_cblock;

typedef struct zmpo_cunit {
  int64_t a_base, b1_base, b2_base, c_base;
  int32_t group, m0;
} zmpo_cunit;

/* rb in {1, 2}.  atile = 1 stages the A blocks in LDS (first-stage pair
 * a_off = A-tile column); atile = 0 reads them from global memory through
 * the block the first-stage pair names in its `pad` field (zmpo_cblock
 * index), leaving the LDS to the intermediate.  kcols_max / ncols_max >=
 * every group's kcols / ncols size the LDS: ((atile ? kcols_max : 0) +
 * ncols_max) * (16*rb + 1) elements <= 160 KiB. */
int zmpo_chaingemm(int dtype, int rb, int atile, int kcols_max, int ncols_max,
                   const zmpo_cunit* units, int64_t nunits, const zmpo_cgroup* groups,
                   const zmpo_cblock* blocks, const zmpo_task* tasks, const zmpo_pair* pairs,
                   const void* A, const void* B1, const void* B2, void* C, int beta,
                   void* stream);

/* zmpo_chaingemm with 16-column output tiles (one 16x16 MFMA block per
 * tile instead of two): the same results bit for bit, for any N; half the
 * MFMA work for row groups whose blocks all have N <= 16 (sigma at K <= 24,
 * where MPO bond sectors are 11-13 wide).  Same contraction as
 * zmpo_chaingemm (kernelQt.py:277-296). */
int zmpo_chaingemm16(int dtype, int rb, int atile, int kcols_max, int ncols_max,
                     const zmpo_cunit* units, int64_t nunits, const zmpo_cgroup* groups,
                     const zmpo_cblock* blocks, const zmpo_task* tasks, const zmpo_pair* pairs,
                     const void* A, const void* B1, const void* B2, void* C, int beta,
                     void* stream);

/* zmpo_streamgemm without a workspace argument: the split-K partial tiles
 * go to a file-scope device array of ZMPO_SCRATCH_BYTES (64 MiB) that each
 * GPU holds until the process ends, so the call allocates nothing (no
 * hipMalloc, no library allocation).  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  work_elems = the
 * elements the schedule's split-K slots use (slots x TM x TN, as the work
 * buffer of zmpo_streamgemm); more than the scratch holds is rejected with
 * ZMPO_ERR_ARG.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race.  Results are
 * bit-identical to zmpo_streamgemm. */
#define ZMPO_SCRATCH_BYTES (64ull << 20)
int64_t zmpo_scratch_bytes(void);
int zmpo_streamgemm_static(int dtype, int shape, int kt, const zmpo_sunit* units,
                           int64_t nunits, int upw, const zmpo_task* tasks,
                           const zmpo_rgroup* groups, int64_t ngroups, const zmpo_pair* pairs,
                           const void* A, const void* B, void* C, int64_t work_elems, int beta,
                           void* stream);

/* out[j] = sum_i X[j*ldx + i] * y[i] for j < nx (device out, float64;
 * work holds zmpo_dots_workspace(nx) doubles).  Fixed reduction order. */
int zmpo_dots(int dtype, const void* X, int64_t ldx, int nx, const void* y, int64_t n,
              double* out, double* work, void* stream);
int zmpo_dots_workspace(int nx);

/* out = beta * out + alpha * sum_j coef[j] * X[j*ldx:...]  (coef: device
 * float64; beta = 0 ignores the previous contents of out). */
int zmpo_lincomb(int dtype, void* out, const void* X, int64_t ldx, int nx, const double* coef,
                 double alpha, double beta, int64_t n, void* stream);

/* r[i] /= |diag[i] - e| + shift  (pdvdson.py:275-278, iop=4). */
int zmpo_precond(int dtype, void* r, const double* diag, double e, double shift, int64_t n,
                 void* stream);

/* dst[i] = src[idx[i]] (idx < 0 gives 0) / dst[idx[i]] = src[i]. */
int zmpo_gather(int dtype_dst, void* dst, int dtype_src, const void* src, const int64_t* idx,
                int64_t n, void* stream);
int zmpo_scatter(int dtype_dst, void* dst, int dtype_src, const void* src, const int64_t* idx,
                 int64_t n, void* stream);
int zmpo_convert(int dtype_dst, void* dst, int dtype_src, const void* src, int64_t n,
                 void* stream);

/* Truncation eigensolver: eigenvalues (ascending) and eigenvectors of the
 * symmetric density-matrix sectors of one decimation (the per-sector eigh of
 * mpo_dmrg_qparser.py:216-293 / rdm_blkdiag, scipy/numpy eigh in the
 * reference), FP64 on the device by rocSOLVER dsyevd on `stream`.  Sector s
 * is n[s] x n[s] at A + off[s]; eigenvector j overwrites A[off[s] + j*n[s]
 * ...]; eigenvalues go to W + woff[s].  Host arrays n, off, woff; device A,
 * W, E (>= max n[s] doubles), info (>= nsec ints; nonzero: no convergence,
 * read after the stream completes). */
int zmpo_syevd_sectors(int nsec, const int64_t* n, const int64_t* off, const int64_t* woff,
                       double* A, double* W, double* E, int* info, void* stream);

/* Device store: allocation and copies of the buffers the launches read and
 * write (MPO sites, environments, psi / sigma, intermediates).  No
 * counterpart in the reference, whose store is HDF5 files (mpo_dmrg_io.py:
 * 168-206, mpo_dmrg_opt.py:196-238).  zmpo_memcpy returns after host
 * copies complete; device-to-device copies are ordered on `stream`. */
#define ZMPO_H2D 0
#define ZMPO_D2H 1
#define ZMPO_D2D 2
int zmpo_alloc(void** out, int64_t bytes);
int zmpo_free(void* p);
int zmpo_memcpy(void* dst, const void* src, int64_t bytes, int kind, void* stream);

/* Launch sequences: a whole two-site sigma (HVecQt, source/mpo_dmrg_
 * kernelQt.py:153-319: L x psi, the chained W[i] x W[i+1] launch, x R per
 * operator chunk), renormalisation (renormQt_H00, kernelQt.py:633-767) or
 * diagonal (HDiagQt, kernelQt.py:30-127) recorded once per plan and issued
 * by one call.  Each op is one zmpo_shapedgemm / zmpo_streamgemm /
 * zmpo_chaingemm(16) launch whose tables (device pointers, owned by the
 * caller and kept alive while the sequence is used) are recorded; its
 * operands are slot numbers bound at run time: A = slots[slot_a], B =
 * slots[slot_b], the chained launch's second B = slots[slot_b2], C =
 * slots[slot_c], split-K workspace = slots[slot_work] (-1: none).  out_f32:
 * run this op as ZMPO_BF16_F32 when the sequence runs as ZMPO_BF16 (the
 * sigma output).  For chains, groups / blocks hold the zmpo_cgroup /
 * zmpo_cblock arrays and units the zmpo_cunit array. */
#define ZMPO_SEQ_SHAPED 0
#define ZMPO_SEQ_STREAM 1
#define ZMPO_SEQ_CHAIN 2
#define ZMPO_SEQ_CHAIN16 3
/* FP64 gather C[i] = A[idx[i]] (idx < 0 gives 0): idx in `tasks`, count in
 * `nunits`, operands in slot_a and slot_c (the diagonal's environment
 * gathers). */
#define ZMPO_SEQ_GATHER 4
typedef struct zmpo_seq_op {
  int32_t kind, shape, kt, upw, rb, atile, kmax, nmax;
  const void* tasks;
  const void* units;
  int64_t nunits;
  const void* groups;
  int64_t ngroups;
  const void* pairs;
  const void* blocks;
  int32_t slot_a, slot_b, slot_b2, slot_c, slot_work, beta, out_f32, pad;
} zmpo_seq_op;
typedef struct zmpo_seq zmpo_seq;
int zmpo_seq_create(zmpo_seq** out);
int zmpo_seq_destroy(zmpo_seq* s);
int zmpo_seq_add(zmpo_seq* s, const zmpo_seq_op* op);
int64_t zmpo_seq_length(const zmpo_seq* s);
int zmpo_seq_run(const zmpo_seq* s, int dtype, void* const* slots, int nslots, void* stream);

/* The §8(b) entry points of the path, each one recorded sequence run with
 * its operands in the reference's order (no allocation, no host
 * synchronisation: graph-capturable after one eager call).
 *   zmpo_sigma_two_site  HVecQt (source/mpo_dmrg_kernelQt.py:153-319): the
 *       plan of one (site, psi template, environment templates); L, R the
 *       environments of all operators, W*_padded the zero-padded MPO sites,
 *       psi in, sigma out (FP32 for dtype ZMPO_BF16), top1 / top3 / work
 *       the intermediates and split-K workspace the plan was sized for.
 *   zmpo_renorm  renormQt_H00 (kernelQt.py:633-767): site tensor values,
 *       the environment in, the MPO site, intermediates t1 / t2, the new
 *       environment out.
 *   zmpo_hdiag  HDiagQt (kernelQt.py:30-127), FP64: the environments, the
 *       MPO diagonal factors, scratch ldiag / rdiag / x, the diagonal out.
 *   zmpo_davidson_residual  one Davidson correction vector (pdvdson.py:
 *       240-283): r = W c - theta V c over nb basis vectors (rows of V and
 *       W, leading dimension ld, device coefficients c), then, when diag is
 *       given, r /= |diag - theta| + shift. */
int zmpo_sigma_two_site(const zmpo_seq* plan, int dtype, const void* L, const void* R,
                        const void* Wi_padded, const void* Wj_padded, const void* psi,
                        void* sigma, void* top1, void* top3, void* work, void* stream);
int zmpo_renorm(const zmpo_seq* plan, int dtype, const void* site, const void* env_in,
                const void* W, void* t1, void* t2, void* env_out, void* work, void* stream);
int zmpo_hdiag(const zmpo_seq* plan, const double* L, const double* R, const double* mpo_factors,
               double* ldiag, double* rdiag, double* x, double* diag, void* work, void* stream);
int zmpo_davidson_residual(int64_t n, const double* W, const double* V, int64_t ld, int nb,
                           const double* coef, double theta, const double* diag, double shift,
                           double* r, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ZMPO_H */

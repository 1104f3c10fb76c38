/*
 * umi_hip.h — C ABI of libumi_hip, the MI355X (gfx950) UMI operator library.
 *
 * Plain C: pointers, sizes and status codes only.  Device pointers are HIP
 * device allocations (umi_malloc or any hipMalloc'd memory); every compute
 * entry point is asynchronous on the stream it is given (NULL = the legacy
 * default stream) and returns UMI_OK or an error code, with a readable
 * message in umi_last_error().  Nothing here aborts.
 *
 * Each declaration cites the reference interface it replaces
 * (johnjohnlin/UMI, file:line).
 */
#ifndef UMI_HIP_H
#define UMI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int umi_status;
enum {
    UMI_OK = 0,
    UMI_ERR_INVALID = 1,     /* bad argument / descriptor                    */
    UMI_ERR_HIP = 2,         /* HIP runtime error (message has the details)  */
    UMI_ERR_UNSUPPORTED = 3, /* shape or parameter outside the built kernels */
    UMI_ERR_NO_DEVICE = 4
};

typedef void *umi_stream; /* a hipStream_t; NULL = default stream */

/* copy directions for umi_memcpy (values equal hipMemcpyKind) */
enum { UMI_H2H = 0, UMI_H2D = 1, UMI_D2H = 2, UMI_D2D = 3, UMI_DEFAULT = 4 };

/* ---------------------------------------------------------------------------
 * Runtime.  Replaces the cudaMalloc / cudaMemcpy / cudaFree calls of
 * src/SyncedMemory.h:19,22,65,74 and the PackUtility upload of
 * src/UMI_gpu.h:374-381.
 * ------------------------------------------------------------------------- */
const char *umi_last_error(void);
umi_status umi_device_count(int *n);
umi_status umi_set_device(int device);
umi_status umi_device_name(char *buf, size_t len); /* gcnArchName + marketing name */
umi_status umi_malloc(void **ptr, size_t bytes);
umi_status umi_free(void *ptr);
umi_status umi_host_alloc(void **ptr, size_t bytes); /* pinned host memory */
umi_status umi_host_free(void *ptr);
umi_status umi_memcpy(void *dst, const void *src, size_t bytes, int kind);
umi_status umi_memcpy_async(void *dst, const void *src, size_t bytes, int kind, umi_stream s);
umi_status umi_memset(void *ptr, int value, size_t bytes);
/* Fill count 32-bit words at ptr (4-byte aligned) with value, ordered on
 * stream s by a small kernel (graph-capturable; hipMemset runs on the null
 * stream, which a non-blocking stream does not wait for).  No reference
 * counterpart: the reference zeroes nothing on the device. */
umi_status umi_fill_u32(void *ptr, uint32_t value, size_t count, umi_stream s);
umi_status umi_synchronize(void);
/* Device-to-device copy by a streaming kernel (16-byte aligned, size a
 * multiple of 16): graph-capturable, and the byte-count reference for HBM
 * traffic counters.  No reference counterpart. */
umi_status umi_copy_d2d(void *dst, const void *src, size_t bytes, umi_stream s);
umi_status umi_stream_create(umi_stream *s);
umi_status umi_stream_destroy(umi_stream s);
umi_status umi_stream_synchronize(umi_stream s);
/* Work enqueued on `waiter` after this call starts only when the work
 * enqueued on `signaller` so far has finished (an event recorded on one
 * stream and waited on by the other: fork and join of streams, also under
 * graph capture).  No reference counterpart: the reference uses the legacy
 * default stream only (src/UMI_gpu.h:610). */
umi_status umi_stream_wait(umi_stream waiter, umi_stream signaller);
/* Launch a kernel compiled in the caller's translation unit (the generic
 * Execute<> engine instantiates its kernel from the user's strategy struct;
 * src/UMI_gpu.h:610 is the launch it replaces). */
umi_status umi_launch_kernel(const void *func, unsigned gx, unsigned gy, unsigned gz,
                             unsigned bx, unsigned by, unsigned bz, size_t shmem,
                             void **args, umi_stream s);

/* Validates the descriptors of a generic UMI::Execute<> call (the checks of
 * src/UMI.cpp:39-66 plus view/size/plane sanity) and returns a status.
 *   p_dims[16] : per dim {total, block, warp, shuf} (shuf 0 read as 1)
 *   a_dims[8]  : per dim {total, block}
 *   views      : per array (n_in inputs then n_out outputs) 24 ints:
 *                p view {id, ofs, stride} x4, then a view x4
 *   phys       : per array 8 ints: {size, mode} x4 (mode 0 PAD, 1 WRAP=clamp)
 *   planes     : per array plane count (the PtrNum nibble)
 *   expand     : Expand template word (nibble per dim, dim 0 highest) */
umi_status umi_validate_execute(const int32_t *p_dims, const int32_t *a_dims, int n_in, int n_out,
                                const int32_t *views, const int32_t *phys, const int32_t *planes,
                                uint64_t expand, uint64_t zero_stride_opt);

/* Event timer on a stream (hipEvent pair).  No reference counterpart: the
 * reference times nothing (src/Timer.h is never included). */
typedef struct umi_timer_s *umi_timer;
umi_status umi_timer_create(umi_timer *t);
umi_status umi_timer_destroy(umi_timer t);
umi_status umi_timer_start(umi_timer t, umi_stream s);
umi_status umi_timer_stop(umi_timer t, umi_stream s);
umi_status umi_timer_elapsed_ms(umi_timer t, float *ms); /* waits for the stop event */

/* Graph capture of a stream (hipGraph).  Used to replay a fixed sequence of
 * launches; replaces nothing in the reference (it launches on the legacy
 * default stream, src/UMI_gpu.h:610). */
typedef struct umi_graph_s *umi_graph;
umi_status umi_graph_begin(umi_stream s);
umi_status umi_graph_end(umi_stream s, umi_graph *g);
umi_status umi_graph_launch(umi_graph g, umi_stream s);
umi_status umi_graph_destroy(umi_graph g);

/* ---------------------------------------------------------------------------
 * Full-search block motion estimation, uint8 frames.
 * Replaces the MotionEst strategy launch of main/full_me.cu:34-64
 * (SAD volume over a (2R+1)^2 window, clamp-to-edge boundary, the
 * reference's WRAP mode = clamp, src/UMI.h:205-218) and adds the argmin.
 *
 *   SAD(by,bx,dy,dx) = sum_{py,px<B} | cur[y+py][x+px]
 *                        - ref[clamp(y+py+dy,0,H-1)][clamp(x+px+dx,0,W-1)] |
 *   with y = by*B, x = bx*B, dy,dx in [-R, R]; nby = H/B, nbx = W/B (floor,
 *   like main/full_me.cu:25).
 *
 * mv_sad  (nullable): int32 [nby][nbx][3] = (dy, dx, SAD) of the FIRST
 *          minimum in raster order over (dy, dx) starting at (-R, -R).
 * volume  (nullable): int32 [nby][nbx][2R+1][2R+1], the full SAD volume
 *          (the output layout of main/full_me.cu:56-62).
 * B in {8, 16}; 1 <= R <= 64; pitch >= W (bytes).
 * ------------------------------------------------------------------------- */
umi_status umi_me_full_search_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                 int B, int R, int32_t *mv_sad, int32_t *volume, umi_stream s);
/* Kernel form of the full search for the following calls (process-wide), for
 * measurement and tests; every form gives identical results.  bit 0: 1 =
 * v_qsad_pk_u16_u8 (default), 0 = v_sad_u8; bit 1: four-column groups only
 * (no light lane for the last column); bit 2: dword staging even for
 * 16-byte aligned frames; bit 3: runtime skips in the last row chunk.
 * 0 <= variant <= 15.  No reference counterpart. */
umi_status umi_me_set_variant(int variant);

/* ---------------------------------------------------------------------------
 * SAD cost volume, uint8 pair.  Replaces the MotionEst launch of
 * main/cost_volume.cu:35-65, generalised to two images and a 2-D window:
 *
 *   cost[y][x][i] = sum_{|dy|<=KY, |dx|<=KX}
 *       | L[cy][clamp(x+dx)] - R[cy][clamp(x+dx+s0+step*i)] |,
 *   cy = clamp(y+dy), i in [0, D), clamps to [0,H-1] / [0,W-1]
 *   (clamp applied to each final coordinate, as main/cost_volume.cu:73-74).
 *
 * The reference example is L = R, KY = 0, KX = 3, s0 = -7, step = +1, D = 15;
 * stereo matching with disparity d = i uses s0 = 0, step = -1.
 * out: int32 [H][W][D].  0 <= KX, KY <= 3; step = +1 or -1; D >= 1.
 * ------------------------------------------------------------------------- */
umi_status umi_cost_volume_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D,
                              int s0, int step, int KY, int KX, int32_t *out, umi_stream s);

/* ---------------------------------------------------------------------------
 * Stereo winner-take-all over the same cost volume, fused so that the volume
 * never reaches device memory.  No reference counterpart: it consumes the
 * volume that the MotionEst launch of main/cost_volume.cu:35-65 writes (the
 * stereo matching step downstream of it).  Per pixel, with cost[y][x][i] as
 * in umi_cost_volume_u8:
 *
 *   out[y][x][0] = the smallest i in [0, D) with cost[y][x][i] minimal,
 *   out[y][x][1] = that cost.
 *
 * out: int32 [H][W][2].  0 <= KX, KY <= 3; step = +1 or -1; 1 <= D <= 1024.
 * ------------------------------------------------------------------------- */
umi_status umi_stereo_wta_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D,
                             int s0, int step, int KY, int KX, int32_t *out, umi_stream s);

/* ---------------------------------------------------------------------------
 * Semi-global matching over the same cost volume.  No reference counterpart:
 * the aggregation step downstream of the volume that the MotionEst launch of
 * main/cost_volume.cu:35-65 writes.  Four paths r (from the left, right, top
 * and bottom), with C = cost[y][x][i] as in umi_cost_volume_u8:
 *
 *   L_r(p, i) = C(p, i) + min(L_r(q, i), L_r(q, i-1) + P1, L_r(q, i+1) + P1,
 *                            M_r(q) + P2) - M_r(q),   q = p - r,
 *   M_r(q) = min_k L_r(q, k); L_r = C on the first pixel of a path;
 *   neighbours outside [0, D) are left out;  S = sum over the four paths;
 *   out[y][x][0] = the smallest i with S minimal, out[y][x][1] = that S.
 *
 * The volume is recomputed from the pair, never stored; workspace holds the
 * partial sums (uint16 per pixel and disparity) and packed copies of the
 * pair: umi_sgm_workspace_size bytes for the same H, W, D, s0, step, KY, KX,
 * no initial contents needed.  Three launches on stream s (pack, the four
 * paths, sum and argmin).  Integer arithmetic, exact.
 * Limits: 0 <= KX, KY <= 3; step = +1 or -1; 1 <= D <= 128;
 * 0 <= P1 <= P2; 4 * (255 (2KY+1)(2KX+1) + P2) < 65536;
 * 2 (4 + 2KY) W <= 160 KiB (W <= 8192 at KY = 3).
 * ------------------------------------------------------------------------- */
umi_status umi_sgm_workspace_size(int H, int W, int D, int s0, int step, int KY, int KX, size_t *bytes);
umi_status umi_sgm_u8(const uint8_t *L, const uint8_t *R, int H, int W, int pitch, int D, int s0, int step,
                      int KY, int KX, int P1, int P2, void *workspace, size_t workspace_bytes, int32_t *out,
                      umi_stream s);

/* ---------------------------------------------------------------------------
 * 3x3 filters.  Pitches are in elements.  mode: 0 = PAD (out-of-image taps
 * read `pad`, the reference's PAD boundary), 1 = clamp to edge (its WRAP).
 * Taps are accumulated dy outer, dx inner, starting from `init`; the weighted
 * forms use fused multiply-adds (x * w + acc).
 *
 * umi_filter3x3_*  : out = mul * (init + sum x)
 *   replaces the Filter3x3 launch of main/filter3x3.cu:10-17,36-58.
 * umi_wfilter3x3_* : out = init + sum x * weights[3*(dy+1)+(dx+1)]
 *   (weights: 9 values in device memory, like the reference's 2nd input)
 *   replaces the WFilter3x3 launch of main/wfilter3x3.cu:8-15,39-67.
 * Integer forms wrap on overflow (two's complement).
 * ------------------------------------------------------------------------- */
umi_status umi_filter3x3_f32(const float *in, int H, int W, int in_pitch, float init, float mul, float pad,
                             int mode, float *out, int out_pitch, umi_stream s);
umi_status umi_filter3x3_i32(const int32_t *in, int H, int W, int in_pitch, int32_t init, int32_t mul,
                             int32_t pad, int mode, int32_t *out, int out_pitch, umi_stream s);
umi_status umi_wfilter3x3_f32(const float *in, int H, int W, int in_pitch, const float *weights, float init,
                              float pad, int mode, float *out, int out_pitch, umi_stream s);
umi_status umi_wfilter3x3_i32(const int32_t *in, int H, int W, int in_pitch, const int32_t *weights,
                              int32_t init, int32_t pad, int mode, int32_t *out, int out_pitch, umi_stream s);

/* ---------------------------------------------------------------------------
 * Bilateral filter, (2r+1)^2 window.  Replaces the Bilateral launch of
 * main/bilateral.cu:9-24,54-99:
 *   c = in[y][x]; for dy, dx in [-r, r] (dy outer):
 *     n = in[y+dy][x+dx] (or `pad` outside, mode 0; edge pixel, mode 1)
 *     w = expf((n-c)*(n-c)*norm) * spatial[dy+r] * spatial[dx+r]
 *     wsum += w; wxsum += w*n
 *   out = wxsum / wsum
 * Out-of-image neighbours count in the normalisation (bilateral.cu:111-113).
 * spatial: 2r+1 floats in device memory; 0 <= r <= 12.
 * ------------------------------------------------------------------------- */
umi_status umi_bilateral_f32(const float *in, int H, int W, int in_pitch, int radius, const float *spatial,
                             float norm, float pad, int mode, float *out, int out_pitch, umi_stream s);

/* ---------------------------------------------------------------------------
 * Integral image out[y][x] = sum_{y'<=y, x'<=x} in[y'][x'].  Replaces the two
 * scan launches of main/iimg.cu:37-82 (row scan into a transposed
 * temporary, then a scan of it).  Forms: uint8 -> uint32 and int32 -> int32
 * (exact; int32 wraps), float -> float with row sums inside a tile in fp32
 * and every vertical sum and carry in fp64.  One launch: a persistent grid
 * walks tiles of TH x 256 in order; each tile publishes its row, column and
 * tile sums and adds those of the tiles above and to its left.  The input is
 * read once and the output written once.
 * workspace: device memory of umi_iimg_workspace_size(H, W, kind) bytes
 * (kind 0 = u8, 1 = i32, 2 = f32), ZERO-FILLED ONCE before its first use
 * (umi_memset or umi_fill_u32); each call leaves it ready for the next, so
 * later calls (graph replays included) clear nothing.  Two calls that
 * overlap in time must not share a workspace.  NULL makes the call allocate,
 * zero and free one itself, which synchronises the stream.
 * umi_iimg_set_config selects the tile height (rows per lane 1, 2, 4 -> TH
 * 16, 32, 64; 0 = chosen by size); results do not depend on it.
 * ------------------------------------------------------------------------- */
umi_status umi_iimg_set_config(int rows_per_lane);
/* Launch shape of the integral-image kernels (one image and batch):
 * 0 (default) = one workgroup per tile; tiles wait only for tiles of smaller
 * workgroup id, which relies on the dispatcher starting workgroups in id
 * order (per XCD), as it does; per_cu > 0 = a persistent grid of per_cu
 * workgroups per CU walking the tiles in order, which needs only that all of
 * them fit on the GPU at once (per_cu <= 2 always does).  Results do not
 * depend on it; either way every wait is bounded in time and reported by
 * umi_iimg_workspace_status. */
umi_status umi_iimg_set_grid(int per_cu);
/* Whether a wait inside an integral-image call on this workspace (one image
 * or batch) ran past its time bound since the last query (which clears the
 * flag); every wait normally ends within microseconds, so 1 means the
 * outputs of that call are not to be trusted.  Synchronous. */
umi_status umi_iimg_workspace_status(void *workspace, int *timed_out);
umi_status umi_iimg_workspace_size(int H, int W, int kind, size_t *bytes);
umi_status umi_iimg_u8_u32(const uint8_t *in, int H, int W, int in_pitch, uint32_t *out, int out_pitch,
                           void *workspace, size_t workspace_bytes, umi_stream s);
umi_status umi_iimg_i32(const int32_t *in, int H, int W, int in_pitch, int32_t *out, int out_pitch,
                        void *workspace, size_t workspace_bytes, umi_stream s);
umi_status umi_iimg_f32(const float *in, int H, int W, int in_pitch, float *out, int out_pitch, void *workspace,
                        size_t workspace_bytes, umi_stream s);
/* umi_iimg_u8_u32 without a workspace argument, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its workspace is a library array of 16 MiB per GPU
 * (held until the process ends), enough for images up to about 30 Mpx (a
 * 16 x 256 tile takes 2176 B); larger images are refused with
 * UMI_ERR_INVALID.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU (other streams, other threads) race and must
 * not be issued; consecutive calls on one stream are fine. */
umi_status umi_iimg_u8_u32_noalloc(const uint8_t *in, int H, int W, int in_pitch, uint32_t *out, int out_pitch,
                                   umi_stream s);

/* ---------------------------------------------------------------------------
 * Box sums from an inclusive uint32 integral image (as umi_iimg_u8_u32
 * writes it): out[y][x] = sum of the source image over rows
 * [max(y-ry,0), min(y+ry,H-1)] and columns [max(x-rx,0), min(x+rx,W-1)],
 * from the four corners of the integral image, unsigned arithmetic mod 2^32
 * (exact whenever the window sum is below 2^32).  No reference counterpart:
 * a consumer of the integral image that main/iimg.cu:37-82 computes.
 * Pitches in elements; ry, rx >= 0.
 * ------------------------------------------------------------------------- */
umi_status umi_box_sum_iimg_u32(const uint32_t *iimg, int H, int W, int in_pitch, int ry, int rx, uint32_t *out,
                                int out_pitch, umi_stream s);

/* ---------------------------------------------------------------------------
 * Rectangle (Haar-like) features from an inclusive uint32 integral image.
 * A feature is up to 3 rectangles (x, y, w, h, weight) relative to a
 * win_h x win_w detection window; feats: int32 [nfeat][3][5], weight 0 =
 * unused rectangle, every used one inside the window (else UMI_ERR_INVALID).
 * For every window position (y, x), 0 <= y <= H - win_h, 0 <= x <= W - win_w:
 *   out[f][y][x] = sum_k weight_k * (sum of the source image over rectangle
 *                  k moved to (y, x)),
 * int32 with wrap-around (unsigned arithmetic mod 2^32).  feats is a host
 * array; the library checks it and passes it to the kernels by value (32
 * features per launch), so no separate device copy can disagree with it.
 * out: int32 [nfeat][H - win_h + 1][W - win_w + 1].  No reference
 * counterpart: a consumer of the integral image of main/iimg.cu:37-82.
 * ------------------------------------------------------------------------- */
umi_status umi_haar_features_u32(const uint32_t *iimg, int H, int W, int pitch, const int32_t *feats, int nfeat,
                                 int win_h, int win_w, int32_t *out, umi_stream s);

/* ---------------------------------------------------------------------------
 * Dense contractions on the matrix cores.  The reference runs GEMM and
 * direct convolution as scalar UMI launches of its Gemm strategy
 * (main/gemm.cu:8-15,37-67; main/fprop.cu:10-17,53-82); these entry points
 * are their MFMA form.  Precision modes:
 *   0 F32     fp32 operands, v_mfma_f32_32x32x2_f32
 *   1 BF16    bf16 operands (RNE), fp32 accumulation
 *   2 BF16X3  bf16 hi/lo split of both operands, Ah*Bh + Al*Bh + Ah*Bl
 *   3 FP8     OCP e4m3 operands scaled per tensor by q = 448/amax
 *   4 FP8X3   FP8 plus residual terms with e4m3((x*q - hi) * 16) at weight 2^-4
 * Operands are prepared ("packed") from fp32 by umi_gemm_pack /
 * umi_conv_pack_*; split modes produce a hi and a lo operand.  For the FP8
 * modes the per-tensor amax comes from umi_amax_f32 and the packer writes
 * the scale q it used to q_out (device float), which the contraction reads
 * back to scale its result by 1/(qa*qb).  All scalars stay in device memory,
 * so a whole pack+contract sequence can be captured in a graph.
 * ------------------------------------------------------------------------- */
umi_status umi_gemm_elem_bytes(int mode, int *bytes);
umi_status umi_amax_f32(const float *src, size_t n, float *amax, umi_stream s);
/* dst[r][k] (K-contiguous, ld_dst elements per row) from fp32 src:
 * transpose = 0: src[r][k] (ld_src per row); 1: src[k][r] (ld_src per row). */
umi_status umi_gemm_pack(const float *src, int R, int K, int ld_src, int transpose, int mode, const float *amax,
                         void *dst_hi, void *dst_lo, int ld_dst, float *q_out, umi_stream s);
/* C[m][n] = sum_k A[m][k] * Bt[n][k] (/(qa*qb) in the FP8 modes); A, Bt
 * K-contiguous (lda, ldb elements), C fp32 row-major.  K * elem_bytes must
 * be a multiple of 64 and lda, ldb * elem_bytes of 16. */
umi_status umi_gemm(int mode, int M, int N, int K, const void *a_hi, const void *a_lo, int lda, const void *b_hi,
                    const void *b_lo, int ldb, float *C, int ldc, const float *qa, const float *qb, umi_stream s);
/* Convolution forward: out[n][k][oy][ox] = sum_{c,r,s} w[k][c][r][s] *
 * in[n][c][oy*stride+r-pad][ox*stride+s-pad] (zero padding), NCHW / KCRS
 * fp32 in the caller's layout (main/fprop.cu:84-108).  The input is packed
 * to NHWC and the filter to KRSC; out is NCHW fp32.  C * elem_bytes must be
 * a multiple of 16 and R*S*C * elem_bytes of 64. */
umi_status umi_conv_pack_input(const float *src_nchw, int N, int C, int H, int W, int mode, const float *amax,
                               void *hi, void *lo, float *q_out, umi_stream s);
umi_status umi_conv_pack_filter(const float *src_kcrs, int K, int C, int R, int S, int mode, const float *amax,
                                void *hi, void *lo, float *q_out, umi_stream s);
umi_status umi_conv_fprop(int mode, int N, int C, int H, int W, int K, int R, int S, int stride, int pad,
                          const void *in_hi, const void *in_lo, const void *w_hi, const void *w_lo, float *out,
                          const float *q_in, const float *q_w, umi_stream s);

/* Selects the form of umi_conv_fprop for bf16, fp8 and fp8x3, 3x3, stride 1,
 * padding 1, C = K = 64, 8 <= W <= 64 (the batch-16 56x56 layer): 1 (default)
 * = a direct convolution with the filter and the input tile staged once in
 * LDS, 0 = the implicit GEMM.  Both use the same operands, MFMA order and
 * scaling and give identical results; the switch exists for measurement and
 * tests.  No reference counterpart. */
umi_status umi_conv_set_direct(int enable);
/* GEMMs with 256 x 256 tiles: nonzero = operands staged by direct-to-LDS
 * loads in a 4-stage ring (2 for the split forms); 1 (default) = with the
 * MFMA operand reads software-pipelined through two register slots (f32,
 * bf16, fp8) and 8 waves per workgroup (two per SIMD), 5 = the same with 4
 * waves; 2 / 4 = the ring without the register pipeline, 8 / 4 waves; 0 =
 * through registers with one stage in flight.  Other values:
 * UMI_ERR_INVALID.  Same operands and MFMA order in every form: identical
 * results; the switch exists for measurement and tests.  No reference
 * counterpart. */
umi_status umi_gemm_set_lds_direct(int enable);

/* ---------------------------------------------------------------------------
 * Batches of images of uneven sizes.  The reference runs one Execute per
 * image (src/UMI_gpu.h:409-623, e.g. main/wfilter3x3.cu:39,
 * main/bilateral.cu:54, main/iimg.cu:37,60); here a batch object describes
 * n images inside one input and one output buffer (element offsets and
 * pitches), and each batched operator is one launch over the tiles of all
 * images.  Creation uploads the work tables once; the operator calls
 * allocate nothing and can be captured in a graph.  `config` selects a tile
 * shape (0 .. umi_batch_configs(op)-1, -1 = default); op 0 = 3x3 filter,
 * 1 = bilateral, 2 = integral image (rows per lane 1, 2, 4).
 * ------------------------------------------------------------------------- */
typedef struct umi_batch_s *umi_batch;
umi_status umi_batch_create(int n, const int *H, const int *W, const int64_t *in_off, const int *in_pitch,
                            const int64_t *out_off, const int *out_pitch, umi_batch *b);
umi_status umi_batch_destroy(umi_batch b);
umi_status umi_batch_configs(int op, int *count);
umi_status umi_batch_wfilter3x3_f32(umi_batch b, const float *in, const float *weights, float init, float pad,
                                    int mode, float *out, int config, umi_stream s);
umi_status umi_batch_bilateral_f32(umi_batch b, const float *in, int radius, const float *spatial, float norm,
                                   float pad, int mode, float *out, int config, umi_stream s);
/* kind 0: uint8 -> uint32, kind 2: float -> float (fp64 vertical sums and
 * carries).  The workspace (any configuration) follows the rules of
 * umi_iimg_u8_u32's: zero-filled once before its first use, then reusable
 * without clearing; not shared by overlapping calls. */
umi_status umi_batch_iimg_workspace_size(umi_batch b, int kind, size_t *bytes);
umi_status umi_batch_iimg_u8_u32(umi_batch b, const uint8_t *in, uint32_t *out, void *workspace,
                                 size_t workspace_bytes, int config, umi_stream s);
umi_status umi_batch_iimg_f32(umi_batch b, const float *in, float *out, void *workspace, size_t workspace_bytes,
                              int config, umi_stream s);
/* Box sums (as umi_box_sum_iimg_u32) of every image of the batch from its
 * integral image; `iimg` and `out` both use the batch's output offsets and
 * pitches.  One persistent launch; no counterpart in the reference. */
umi_status umi_batch_box_sum_u32(umi_batch b, const uint32_t *iimg, int ry, int rx, uint32_t *out, umi_stream s);

/* ---------------------------------------------------------------------------
 * Learned motion proposals (no reference counterpart: johnjohnlin/UMI has
 * only the exhaustive search of main/full_me.cu).  Building blocks of the
 * proposal pipeline and of its on-GPU training; umi_amd/proposer.py strings
 * them together.
 * ------------------------------------------------------------------------- */
/* ---------------------------------------------------------------------------
 * Quarter-pel refinement of integer 8x8 block motion vectors (as the full
 * search returns them), the step a video encoder runs before coding them.
 * No reference counterpart (main/full_me.cu is integer-pel).  Reference
 * samples at quarter-pel positions (qy, qx):
 *   Y = floor(qy/4), fy = qy - 4Y (X, fx likewise),
 *   v = ((4-fx)(4-fy) A + fx(4-fy) B + (4-fx)fy C + fx fy D + 8) >> 4,
 *   A, B, C, D = ref at (Y,X), (Y,X+1), (Y+1,X), (Y+1,X+1), clamped to the frame.
 * From c = 4 x (integer MV): the 8 half-pel neighbours c + 2 o, then the 8
 * quarter-pel neighbours of the best (o in {-1,0,1}^2 \ {0}, raster order),
 * each replacing the centre only when its SAD is smaller (ties keep the
 * centre).  mv_in: int32 [H/8][W/8][3] (dy, dx, SAD); mv_out: the same in
 * quarter pels (dy4, dx4, SAD at that position).  An NV12 frame's luma plane
 * is a uint8 frame with its pitch.
 * ------------------------------------------------------------------------- */
umi_status umi_me_subpel_refine_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                   const int32_t *mv_in, int32_t *mv_out, umi_stream s);

/* ---------------------------------------------------------------------------
 * Full-search ME over NV12 frames (luma plane + interleaved UV plane, 4:2:0).
 * No reference counterpart: it extends the MotionEst search of
 * main/full_me.cu:34-64 (clamp boundary, first minimum in raster order) to
 * the chroma of a video frame.  For each 8x8 luma block (nby = H/8,
 * nbx = W/8) and integer vector (dy, dx) in [-R, R]^2:
 *   cost = SAD_Y(8x8) + SAD_U(4x4) + SAD_V(4x4),
 * the chroma predicted as in H.264 8.4.2.2.2 from the vector (4dy, 4dx) in
 * quarter-luma units: integer offset floor(d/2), eighth-pel fraction
 * 4 (d & 1); samples clamped to the (H/2) x (W/2) chroma plane; U = byte 2x,
 * V = byte 2x + 1 of a uv row.
 * mv_cost: int32 [nby][nbx][3] = (dy, dx, cost).  H, W even and >= 8;
 * pitch_y, pitch_uv >= W bytes; 1 <= R <= 16.
 * ------------------------------------------------------------------------- */
umi_status umi_me_full_search_nv12(const uint8_t *cur_y, const uint8_t *cur_uv, const uint8_t *ref_y,
                                   const uint8_t *ref_uv, int H, int W, int pitch_y, int pitch_uv, int R,
                                   int32_t *mv_cost, umi_stream s);

/* ---------------------------------------------------------------------------
 * Motion vectors in formats a video codec or filter graph consumes.  No
 * reference counterpart.  mv: int32 [nby][nbx][3] (dy, dx, cost) from the
 * searches above, in pels (units = 1) or quarter pels (units = 4, from
 * umi_me_subpel_refine_u8).  B: block size.
 *   format 0: int16 [nby][nbx][2] = (mvx, mvy) in quarter pels, the
 *             motion-hint layout of hardware encoders;
 *   format 1: FFmpeg AVMotionVector records (libavutil/motion_vector.h,
 *             40 bytes each, one per block in raster order): source = -1,
 *             w = h = B, dst = block centre, motion = (mvx, mvy) with
 *             motion_scale = 4, src = dst + (motion >> 2), flags = 0.
 * ------------------------------------------------------------------------- */
umi_status umi_me_export_mvs(const int32_t *mv, int nby, int nbx, int B, int units, int format, void *out,
                             umi_stream s);

/* Full search of the listed 8x8/16x16 blocks only (block index by*nbx+bx in
 * list[0 .. *count)), writing their (dy, dx, SAD) into mv_sad; semantics of
 * umi_me_full_search_u8.  max_count bounds *count. */
umi_status umi_me_full_search_listed_u8(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int B, int R,
                                        const int32_t *list, const int32_t *count, int max_count, int32_t *mv_sad,
                                        umi_stream s);
/* Synthetic pair: f0 = 5x5 box of hashed uniform noise; f1 = f0 warped by the
 * smooth field of 4 bumps (9 floats each: cy, cx, 1/(2 s^2), ty, tx, a00,
 * a01, a10, a11; components clamped to +-32) plus N(0, sigma^2) noise. */
umi_status umi_synth_frame0(uint8_t *f0, int H, int W, int pitch, uint64_t seed, umi_stream s);
umi_status umi_synth_warp(const uint8_t *f0, uint8_t *f1, int H, int W, int pitch, const float *bumps, float sigma,
                          uint64_t seed, umi_stream s);
/* Network input of pair slot n (n >= 0) of an fp32 [.][2][H/4][W/4] array:
 * channel 0 cur, channel 1 ref, (4x4 box mean - 128) / 8. */
umi_status umi_cnn_downsample_pair(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, float *out, int n,
                                   umi_stream s);
/* Training kernels, fp32 NCHW, 3x3 convolution with dilation dil and zero
 * padding dil (same size).  w: [Cout][Cin][3][3].  relu: the layer applies
 * ReLU; backward passes mask with its output act > 0.  bwd_weight writes
 * dw (Cout*Cin*9) followed by db (Cout). */
umi_status umi_cnn_conv_fwd(const float *in, int N, int Cin, int Cout, int H, int W, int dil, const float *w,
                            const float *bias, int relu, float *out, umi_stream s);
umi_status umi_cnn_conv_bwd_data(const float *dout, const float *act, int N, int Cin, int Cout, int H, int W, int dil,
                                 const float *w, int relu, float *din, umi_stream s);
umi_status umi_cnn_conv_bwd_weight_workspace(int N, int Cin, int Cout, int H, int W, size_t *bytes);
umi_status umi_cnn_conv_bwd_weight(const float *dout, const float *act, const float *in, int N, int Cin, int Cout,
                                   int H, int W, int dil, int relu, float *dw_db, void *workspace,
                                   size_t workspace_bytes, umi_stream s);
/* Huber loss (delta 1) of the flow [N][2][H4][W4] against block labels
 * [N][H4/2][W4/2][3] / 4; writes dloss/dflow (mean) and per-workgroup loss
 * partial sums (count in *nparts). */
umi_status umi_cnn_huber(const float *pred, const int32_t *labels, int N, int H4, int W4, float *grad,
                         float *loss_parts, int *nparts, umi_stream s);
umi_status umi_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                    int step, umi_stream s);
/* One inference layer of the proposal network on bf16 MFMA (3x3, dilation
 * dil, zero padding dil; fp32 accumulation; inputs rounded to bf16):
 *   kind 0: 2 -> 32, fp32 NCHW in, bf16 NHWC out, ReLU;
 *   kind 1: 32 -> 32, bf16 NHWC in and out, ReLU;
 *   kind 2: 32 -> 2, bf16 NHWC in, fp32 NCHW out (the flow), no ReLU.
 * umi_cnn_pack_layer rounds the fp32 weights [cout][cin][3][3] to bf16 in
 * the MFMA operand order, into umi_cnn_packed_bytes(kind) bytes. */
umi_status umi_cnn_packed_bytes(int kind, size_t *bytes);
umi_status umi_cnn_pack_layer(int kind, const float *w, void *packed, umi_stream s);
umi_status umi_cnn_infer_layer(int kind, const void *in, int N, int H, int W, int dil, const void *packed,
                               const float *bias, void *out, umi_stream s);
/* Fused forms for one frame pair: the first layer computed straight from the
 * u8 frames (umi_cnn_downsample_pair's values, then kind 0, identical
 * results), and the last layer (kind 2, identical flow) that also writes the
 * block proposals of umi_me_proposals into prop when prop is not NULL. */
umi_status umi_cnn_first_layer_frames(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                      const void *packed, const float *bias, void *out_nhwc, umi_stream s);
/* The first two layers in one launch: umi_cnn_first_layer_frames followed by
 * layer kind 1 of dilation `dil` (packed1, bias1), identical results, without
 * the first layer's output in device memory. */
umi_status umi_cnn_first_two_layers_frames(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch,
                                          const void *packed0, const float *bias0, int dil, const void *packed1,
                                          const float *bias1, void *out_nhwc, umi_stream s);
umi_status umi_cnn_last_layer_proposals(const void *in_nhwc, int H4, int W4, int dil, const void *packed,
                                        const float *bias, int R, float *flow, int32_t *prop, umi_stream s);
/* Block proposals from the quarter-resolution flow: round(4 x mean of the
 * block's 2x2 outputs), clamped to [-R, R]; prop [H4/2][W4/2][2]. */
umi_status umi_me_proposals(const float *flow, int H4, int W4, int R, int32_t *prop, umi_stream s);
/* Exact refinement of the block proposals (no counterpart in the reference:
 * it serves the learned proposer).  Per 8x8 block, with every vector clamped
 * to [-R, R]:
 *   seed    - the first minimum-SAD vector among the block's own proposal and
 *             its 8 neighbours' proposals (raster order);
 *   descent - repeat up to `iters` times: evaluate the 5x5 window around the
 *             centre, move to its first minimum (raster order), stop when that
 *             minimum is within 1 of the centre in both axes;
 *   in-tile - up to `local_rounds` rounds inside tiles of 4 x 8 blocks
 *             (aligned to the block grid), distance k = 4, 2, 1, 4, ...: each
 *             block tries its in-tile neighbours' vectors at distance 1 and k
 *             and descends from the first minimum when it beats its own SAD;
 *             a tile stops after a round without a change;
 *   rounds  - up to `rounds` such rounds over the whole frame, round p
 *             trying the neighbours at distance 1 and at the two distances
 *             (8, 32), (4, 16), (2, 8), (4, 32) of entry p % 4, stopping after
 *             one without a change;
 *   check   - a block falls back when its descent never stopped, or when its
 *             SAD exceeds slack x the smallest SAD it has at its 8 neighbours'
 *             final vectors.
 * mv_sad [H/8][W/8][3] receives (dy, dx, SAD): for fallback blocks the full
 * search's result over [-R, R]^2 (first minimum in raster order, as
 * umi_me_full_search_u8), for the others the refined one.  Fallback block
 * indices go to fb_list, their number to *fb_count (cleared first).  ws: umi_me_refine_workspace bytes; it also
 * holds, after two state buffers of 16 B per block (each rounded up to 256 B),
 * one int32 per global round that is nonzero when that round changed some
 * block. */
umi_status umi_me_refine_workspace(int H, int W, int rounds, size_t *bytes);
umi_status umi_me_refine(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int R, const int32_t *prop,
                         int iters, int local_rounds, int rounds, float slack, int32_t *mv_sad, int32_t *fb_list,
                         int32_t *fb_count, void *ws, size_t ws_bytes, umi_stream s);
/* The same refinement in one launch, with every one of the `rounds` global
 * rounds run (no stop after a round without a change) -- no counterpart in
 * the reference.  A workgroup per tile of 4 x 8 blocks keeps the tile's ref
 * region in LDS from the seed to the check; block states pass between tiles
 * as tagged words in the workspace, and the fallback blocks' full search
 * runs in the same launch.  Tiles wait on each other, so the launch needs
 * every tile resident at once: UMI_ERR_UNSUPPORTED when more than 1024 tiles
 * ((H/8 / 4) x (W/8 / 8), rounded up), more than 4/5 of the device's
 * resident workgroups, or rounds > 32.  Calls that overlap in time with
 * other kernels that wait on the GPU (another fused refinement on another
 * stream) may time out.  ws: umi_me_refine_fused_workspace bytes, zeroed
 * before its first use and then kept for frames of this size (every call
 * leaves it consistent).  Outputs as umi_me_refine; fb_list is in tile
 * order.  Every wait is bounded (0.2 s); umi_me_refine_fused_status reports
 * (and clears) a wait that ran out, and copies the round change flags of the
 * last call. */
umi_status umi_me_refine_fused_workspace(int H, int W, int rounds, size_t *bytes);
umi_status umi_me_refine_fused(const uint8_t *cur, const uint8_t *ref, int H, int W, int pitch, int R,
                               const int32_t *prop, int iters, int local_rounds, int rounds, float slack,
                               int32_t *mv_sad, int32_t *fb_list, int32_t *fb_count, void *ws, size_t ws_bytes,
                               umi_stream s);
umi_status umi_me_refine_fused_status(void *ws, int rounds, int32_t *changed, int *timed_out);

#ifdef __cplusplus
}
#endif
#endif /* UMI_HIP_H */

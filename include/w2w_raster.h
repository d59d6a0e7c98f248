/*
 * w2w_raster.h -- C ABI of the MI355X (gfx950) Gaussian tile rasterizer.
 *
 * Drop-in boundary for the reference's pybind module `_C`
 * (cudaForwardBackwardRenderer/ext.cpp:5-7):
 *   rasterize_gaussians           rasterize_points.h:8-20,  rasterize_points.cu:31-100
 *   rasterize_gaussians_backward  rasterize_points.h:24-48, rasterize_points.cu:101-189
 * The Python shim word2world_amd.raster keeps those positional signatures and
 * calls the functions below through ctypes.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Every tensor pointer is device memory
 *    (hipMalloc / torch CUDA tensor), float32 unless stated, C-contiguous.
 *  - The caller allocates every buffer (outputs and the three workspace
 *    buffers sized by w2w_workspace_sizes); no entry point allocates, frees
 *    or synchronises the device, so all of them are legal under stream
 *    capture.  All work goes on `stream` (a hipStream_t; NULL = default).
 *  - Every entry point returns a W2W_* status and never throws.
 *  - The number of rendered instances (K) stays on the device: it is the
 *    uint32 at byte offset 0 of the geometry buffer; the uint32 at byte 4 is
 *    nonzero when K exceeded the binning capacity (the outputs are then
 *    incomplete and the caller re-runs with a larger capacity).
 */
#ifndef W2W_RASTER_H
#define W2W_RASTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define W2W_ABI_VERSION 9

enum {
  W2W_OK = 0,
  W2W_ERR_INVALID = 1,   /* bad size, null pointer or unsupported tile size */
  W2W_ERR_HIP = 2,       /* a launch failed */
  W2W_ERR_CAPACITY = 3,  /* workspace smaller than w2w_workspace_sizes asks */
};

/* One view of one scene: the inputs of rasterize_gaussians
 * (rasterize_points.cu:33-45) plus the image and tile size the reference
 * hard-coded (800x800, 16; PD-1). */
typedef struct W2WFrame {
  int32_t P;            /* Gaussians */
  int32_t M;            /* SH coefficients per Gaussian (row stride of sh) */
  int32_t D;            /* active SH degree, 0..3, (D+1)^2 <= M */
  int32_t W, H;         /* image size in pixels */
  int32_t tile;         /* tile edge: 8, 16 or 32 */
  int32_t chunk;        /* instances per forward-blend chunk: 64, 128 or 256
                         * (0: 64); the counterpart of the reference's blend
                         * batch (pixelGaussiansFull.cu:41-76); same outputs */
  int32_t reserved;     /* 0 */
  float focal_x, focal_y;
  const float* means;       /* [P,3] */
  const float* scales;      /* [P,3] activated (exp) */
  const float* rotations;   /* [P,4] (w,x,y,z), normalised */
  const float* opacities;   /* [P] activated (sigmoid) */
  const float* sh;          /* [P,M,3] */
  const float* viewmatrix;  /* [16] column-major W2C (Renderer.py:135) */
  const float* projmatrix;  /* [16] column-major P*W2C (Renderer.py:136) */
  const float* campos;      /* [3] */
  const float* background;  /* [3] */
} W2WFrame;

/* Caller-owned workspace: geometry (per Gaussian), binning (per instance,
 * `capacity` instances) and image (per pixel / tile) state, kept between the
 * forward and the backward of one view (the reference's geomBuffer,
 * binningBuffer and imgBuffer, rasterize_points.cu:61-66). */
typedef struct W2WWorkspace {
  void* geom;
  size_t geom_bytes;
  void* binning;
  size_t binning_bytes;
  void* image;
  size_t image_bytes;
  int64_t capacity;
} W2WWorkspace;

/* Byte offsets of the arrays inside the workspace buffers (tests and
 * debugging read them; the layout is a pure function of the sizes). */
typedef struct W2WLayout {
  size_t geom_bytes, binning_bytes, image_bytes;
  size_t splats;     /* geom: 64-byte records (px,py,a,b | c,o,depth,skip | r,g,b,extent |
                        block-test threshold, tile rectangle (2 words as in rect), first
                        backward-record position of the item relative to item_off) */
  size_t radii;      /* geom: int32[P] */
  size_t rect;       /* geom: uint32x2[P] (x0|y0<<16, x1|y1<<16) */
  size_t counts;     /* geom: uint32[P] tiles touched */
  size_t perm;       /* geom: uint32[P] Gaussian ids in (depth, id) order */
  size_t offsets;    /* geom: uint32[P] exclusive scan of counts in depth order */
  size_t item_off;   /* geom: uint32[ceil(P/256)] first backward-record position of items
                        256 k .. 256 k + 255 (the tile counts scanned in item order); an
                        item's own position adds the last word of its splat record */
  size_t sorted_ids; /* binning: uint32[capacity] Gaussian id per sorted instance */
  size_t sorted_tiles; /* binning: tile id per sorted instance, tile_key_bytes wide */
  size_t ranges;     /* image: uint32x2[T] per-tile [start,end) */
  size_t n_contrib;  /* image: uint32[H*W] 1-based last contributor */
  size_t final_T;    /* image: float[H*W] final transmittance */
  size_t item_frame; /* geom: uint32[items] frame of each item */
  size_t table;      /* geom: the one-frame table the single-view entry points write */
  size_t grad_rec;   /* binning: float[capacity][10] per-instance backward partials */
  size_t grad_flag;  /* binning: uint32[ceil(capacity / 32)], bit k set where grad_rec holds a partial */
  size_t item_grad;  /* geom: float[5][items]: per item the backward's summed records
                        (dL/dconic a, b, c, dL/dopacity, dL/ddepth), one plane per field */
  size_t tile_key_bytes; /* 2 when the batch has at most 65536 tiles (uint16 keys), else 4 */
} W2WLayout;

/* Capacities of a batched workspace: every grid is sized by these, and every
 * data-dependent count is read on the device, so one launch sequence (or one
 * captured HIP graph) serves any batch that fits. */
typedef struct W2WCaps {
  int32_t scenes, frames;  /* descriptor slots in the batch table */
  int32_t tile;            /* tile edge: 8, 16 or 32 */
  int32_t chunk;           /* instances per forward-blend chunk: 64, 128 or
                            * 256 (0: 64); same outputs */
  int32_t reserved;        /* 0 */
  int32_t depth_split;     /* 0: one binning pass.  1..999: depth-sliced
                            * binning (no reference counterpart; same
                            * outputs): the nearest depth_split/1000 of the
                            * items (in depth order) are binned and blended
                            * first, the rest only on tiles that still have
                            * unsaturated pixels.  Needs a larger workspace
                            * (w2w_batch_workspace_layout). */
  int64_t items;           /* (frame, Gaussian) pairs */
  int64_t gaussians;       /* Gaussians over all scenes */
  int64_t tiles;           /* tiles over all frames */
  int64_t pixels;          /* pixels over all frames */
  int64_t capacity;        /* tile instances (K) */
} W2WCaps;

int w2w_version(void);

/* Host-side inspection (no device work; no counterpart in the reference):
 * the cull extent the preprocess kernel packs into each splat record, for
 * conics conic[3 i .. 3 i + 2] = (a, b, c) and skip thresholds skip[i].
 * out[i] = two binary16 half-widths (x low, y high) of the box outside of
 * which the blend kernels drop the pair. */
int w2w_splat_extents(const float* conic, const float* skip, uint64_t n, uint32_t* out);
/* Host-side inspection (no counterpart in the reference): whether the blend
 * kernels' exact block test keeps splat i, splat[6 i ..] = (mx, my, conic a,
 * b, c, skip_below), for the pixel-centre rectangle rect[4 i ..] = (x0, x1,
 * y0, y1); out[i] = 1 when kept. */
int w2w_block_touches(const float* splat, const float* rect, uint64_t n, uint8_t* out);

int w2w_workspace_layout(int32_t P, int32_t W, int32_t H, int32_t tile, int64_t capacity,
                         W2WLayout* out);

/* rasterize_gaussians: preprocess, depth sort, scan, emit, tile sort, ranges,
 * blend.  Outputs: color [3,H,W], depth [H,W] (sum z_ndc alpha T), alpha
 * [H,W] (sum alpha T), radii [P] int32. */
int w2w_forward(const W2WFrame* f, const W2WWorkspace* ws, float* out_color, float* out_depth,
                float* out_alpha, int32_t* out_radii, void* stream);

/* rasterize_gaussians_backward: gradients of the forward above w.r.t. the
 * same inputs, in the reference's tuple order (rasterize_points.cu:188):
 * dL_dmeans2D [P,3] (NDC units, z = 0), dL_dcolors [P,3], dL_dopacity [P],
 * dL_dmeans3D [P,3], dL_dcov3D [P,6], dL_dsh [P,M,3], dL_dscales [P,3],
 * dL_drotations [P,4].  `ws` must hold this view's forward state. */
int w2w_backward(const W2WFrame* f, const W2WWorkspace* ws, const float* dL_dcolor,
                 const float* dL_ddepth, const float* dL_dalpha, float* dL_dmeans2D,
                 float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D,
                 float* dL_dsh, float* dL_dscales, float* dL_drotations, void* stream);

/* ---- batched form: many scenes and views in one launch per stage ----
 * A batch table (device memory, w2w_batch_table_bytes bytes, filled on the
 * host by w2w_batch_encode and copied to the device by the caller) describes
 * scenes (Gaussian counts, views, SH stride/degree) and frames (camera, size,
 * flags; flag bit 0 = this frame takes part in the backward).  Gaussians of
 * all scenes are packed back to back in the parameter arrays; SH blocks too
 * (scene k's block starts where scene k-1's ends).  Work items are (frame,
 * Gaussian) pairs, Gaussian-major within a scene: item = item_off(scene) +
 * g * views + v.  Per-frame outputs are packed in frame order: colour
 * [3,H,W] blocks, and depth / alpha [H,W] blocks.
 * scene_base / scene_sh_off (optional): the row of each scene's first Gaussian
 * in the parameter and per-Gaussian gradient arrays and the float offset of
 * its SH block, for scenes that stay resident at fixed places in a larger
 * store (NULL = packed back to back in batch order).  frame_target_off
 * (optional): float offset of each frame's [3,H,W] loss target in the target
 * buffer w2w_l2_loss reads (NULL = packed like the colour output).
 * totals (optional, 6 int64): Gaussians, items, frames, tiles, pixels, SH
 * floats of the encoded batch. */
size_t w2w_batch_table_bytes(int32_t scenes_cap, int32_t frames_cap);
/* Encodes a batch for the workspace capacities `caps` (scenes, frames, tile
 * edge, and the item, Gaussian, tile and pixel totals): returns
 * W2W_ERR_CAPACITY when the batch exceeds any of them, so that no kernel of a
 * workspace sized by `caps` can be handed a batch larger than its buffers. */
int w2w_batch_encode(void* host_table, size_t table_bytes_avail, const W2WCaps* caps,
                     int32_t n_scenes, const int64_t* scene_n, const int32_t* scene_views,
                     const int32_t* scene_M, const int32_t* scene_D, const float* frame_view,
                     const float* frame_proj, const float* frame_campos, const float* frame_bg,
                     const float* frame_focal, const int32_t* frame_wh, const uint32_t* frame_flags,
                     const int64_t* scene_base, const int64_t* scene_sh_off,
                     const int64_t* frame_target_off, int64_t* totals);
int w2w_batch_workspace_layout(const W2WCaps* caps, W2WLayout* out);
/* radii per item. */
int w2w_forward_batch(const void* table, const W2WCaps* caps, const W2WWorkspace* ws,
                      const float* means, const float* scales, const float* rotations,
                      const float* opacities, const float* sh, float* out_color, float* out_depth,
                      float* out_alpha, int32_t* out_radii, void* stream);
/* dL_dmeans2D and dL_dcolors per item [3]; the other gradients per Gaussian,
 * summed over the views of its scene (dL_dsh packed like sh).  dL_ddepth and
 * dL_dalpha may be NULL (zero). */
int w2w_backward_batch(const void* table, const W2WCaps* caps, const W2WWorkspace* ws,
                       const float* means, const float* scales, const float* rotations,
                       const float* sh, const float* dL_dcolor, const float* dL_ddepth,
                       const float* dL_dalpha, float* dL_dmeans2D, float* dL_dcolors,
                       float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh,
                       float* dL_dscales, float* dL_drotations, void* stream);

/* w2w_backward_batch with the SH coefficients' Adam step taken in the SH
 * backward instead of writing dL_dsh (no counterpart in the reference, whose
 * trainer runs torch.optim.Adam after the backward, gaussians_handler.py:40;
 * for a trainer that keeps its parameters and Adam state in device memory):
 * sh is the parameter store's SH block and is updated in place, sh_exp_avg
 * and sh_exp_avg_sq are packed like it; step size hyper[hyper_dc] for the
 * first three coefficients of each row and hyper[hyper_ac] for the rest
 * (-lr / (1 - b1^t)), hyper[16..20] = 1/sqrt(1 - b2^t), 1 - b1, b2, 1 - b2,
 * eps (the w2w_adam_step_hyper layout); no SH update while *skip_if != 0.
 * The update of each coefficient is the one w2w_adam_step_hyper applies with
 * the gradient w2w_backward_batch would write (bit for bit). */
int w2w_backward_batch_sh_adam(const void* table, const W2WCaps* caps, const W2WWorkspace* ws,
                               const float* means, const float* scales, const float* rotations, float* sh,
                               const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                               float* dL_dmeans2D, float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D,
                               float* dL_dcov3D, float* sh_exp_avg, float* sh_exp_avg_sq,
                               float* dL_dscales, float* dL_drotations, const float* hyper,
                               int32_t hyper_dc, int32_t hyper_ac, const uint32_t* skip_if, void* stream);

/* 64-bit keys (tile << 32 | depth bits) of the sorted instances, i.e. the
 * reference's sorted keysArray (prepareSort.cu:431-434), for parity checks. */
int w2w_sorted_keys(const W2WFrame* f, const W2WWorkspace* ws, uint64_t* out_keys, void* stream);

/* Stable radix sort of (u64 key, u32 value) pairs over bits [0, end_bit):
 * the semantics of radix_sort_2bit_coalesced (prepareSort.cu:303-360).
 * `temp` must hold w2w_sort_temp_bytes(n, end_bit) bytes.  Output buffers
 * may not alias the inputs. */
size_t w2w_sort_temp_bytes(int64_t n, int32_t end_bit);
int w2w_sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                       uint32_t* vals_out, int64_t n, int32_t end_bit, void* temp,
                       size_t temp_bytes, void* stream);

/* Exclusive scan of uint32 values (the reference's kogge-stone scan,
 * gaussianRasterizer.cu:77-117, is inclusive: inclusive[i] = out[i] + in[i]).
 * Writes the total to *total (device).  temp: w2w_scan_temp_bytes(n).
 * Sums are exact below 2^30; a total of 2^30 or more is reported as
 * 2^30 - 1 (and the offsets from that point on are not meaningful). */
size_t w2w_scan_temp_bytes(int64_t n);
int w2w_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n,
                           void* temp, size_t temp_bytes, void* stream);

/* No counterpart in the reference: w2w_exclusive_scan_u32 for callers that
 * cannot allocate (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels and the same results; no temp
 * argument.  Its scratch (ticket and one look-back word per 16384 values) is
 * a static device array of the library, 64 KiB on each GPU it runs on, held
 * until the process ends, cleared by a fill kernel on `stream` at every call
 * (no allocation, no memset).  n <= W2W_SCAN_NOALLOC_MAX_N, otherwise
 * W2W_ERR_CAPACITY.  Every call on one GPU shares that GPU's scratch: calls
 * that overlap on one GPU (other streams, other threads) race; calls in one
 * stream are ordered and safe. */
#define W2W_SCAN_NOALLOC_MAX_N (1 << 28)
int w2w_exclusive_scan_u32_noalloc(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n,
                                   void* stream);

/* Fused Adam step over a flat parameter store split into `n_groups` (<= 8)
 * consecutive groups (group k ends at group_end[k], host array) with
 * learning rate group_lr[k] (host array): the torch.optim.Adam update the
 * trainer uses (gaussians_handler.py:40, OptimizerManager.py), bias-corrected
 * with `step` (>= 1).  params/grads/exp_avg/exp_avg_sq are device arrays of n
 * floats. */
int w2w_adam_step(float* params, const float* grads, float* exp_avg, float* exp_avg_sq, int64_t n,
                  int32_t n_groups, const int64_t* group_end, const float* group_lr, int64_t step,
                  float beta1, float beta2, float eps, void* stream);

/* ---- step-level kernels of the batched trainer (no autograd, capturable) ----
 * Activation of Renderer.py:164-167 over G Gaussians: scales = exp(svec),
 * opacity = sigmoid(logit), rotations = q / max(|q|, 1e-12); and its
 * backward from the gradients w.r.t. the activated values. */
int w2w_activate(const float* svec, const float* logit, const float* quat, int64_t G, float* scales,
                 float* opac, float* rots, void* stream);
int w2w_activate_backward(const float* scales, const float* opac, const float* quat,
                          const float* g_scales, const float* g_opac, const float* g_rots,
                          int64_t G, float* g_svec, float* g_logit, float* g_quat, void* stream);
/* Per frame: mean over 3*H*W of (clamp(color,0,1) - target)^2 into
 * loss_per_frame[frame] and its gradient into dL_dcolor (zero for frames
 * without the backward flag); colour/target packed as the batch outputs.
 * tile_scratch: caps->tiles floats (per-tile partial sums).  Every sum has a
 * fixed order, so the loss is deterministic.  The loss replaces the SDS
 * guidance of GaussianTrainer.py:86-114 for synthetic targets. */
int w2w_l2_loss(const void* table, const W2WCaps* caps, const float* color, const float* target,
                float* dL_dcolor, float* loss_per_frame, float* tile_scratch, void* stream);
/* Adam over a flat store with the step's scalars in device memory (so a
 * captured graph picks them up at replay): hyper[0..7] = -lr_k/(1-b1^t) per
 * group, hyper[8..15] = the same for a group's second rate, hyper[16] =
 * 1/sqrt(1-b2^t), hyper[17..20] = 1-b1, b2, 1-b2, eps.  Group k may alternate
 * rates: element e uses the second rate when (e mod group_period[k]) >=
 * group_split[k] (period <= 0: one rate).  skip_if (optional, device): when
 * the word is nonzero at run time -- e.g. the overflow word of the step's
 * forward, whose images and gradients are then incomplete -- nothing is
 * updated and bit 0 of *sticky (optional, device) is set; the caller clears
 * *sticky after growing the workspace. */
int w2w_adam_step_hyper(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                        int64_t n, int32_t n_groups, const int64_t* group_end,
                        const int32_t* group_period, const int32_t* group_split,
                        const float* hyper, const uint32_t* skip_if, uint32_t* sticky,
                        void* stream);

/* Range tables for stores whose scenes stay resident at fixed rows.  A row
 * range {int64 start, off, len} covers Gaussian rows [start, start+len) and
 * sits at [off, off+len) of the concatenated work index space (ranges sorted
 * by off, contiguous, total = sum of len). */
int w2w_activate_rows(const float* svec, const float* logit, const float* quat, const void* ranges,
                      int32_t n_ranges, int64_t total_rows, float* scales, float* opac, float* rots,
                      void* stream);
int w2w_activate_backward_rows(const float* scales, const float* opac, const float* quat,
                               const float* g_scales, const float* g_opac, const float* g_rots,
                               const void* ranges, int32_t n_ranges, int64_t total_rows,
                               float* g_svec, float* g_logit, float* g_quat, void* stream);
/* Adam over parameter ranges {int64 start, off, len; int32 hyper; uint32
 * period, split, pad} of a flat store: range r uses hyper[4*r.hyper + 0]
 * (-lr/(1-b1^t)), [+1] the second rate (elements with (e mod period) >=
 * split, period 0 = none) and [+2] 1/sqrt(1-b2^t), so jobs with different
 * step counts and learning rates share one launch. */
int w2w_adam_ranges(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                    const void* ranges, int32_t n_ranges, int64_t total, const float* hyper,
                    float beta1, float beta2, float eps, void* stream);

/* ---- adaptive density control on the device (one adc_cycle,
 * gaussians_handler.py:185-194) ----
 * A store is described group by group: for each of n_groups (<= 8) groups,
 * the floats per row and the device address of row 0 of the parameters and
 * of both Adam moments (groups of n rows, row-major).  means / opacity /
 * svec / quat name the groups that hold the means (3 floats), opacity
 * logit (1), log-scales (3) and unnormalised quaternion (4); the others
 * (SH) are copied as they are.  The trainer's store is [means | sh_dc |
 * sh_ac | opacity | svec | quaternion], the batched store [means | sh |
 * opacity | svec | quaternion]; one scene of a multi-scene store is the
 * same table with each pointer advanced to the scene's first row.
 * grad_accum / grad_count: n floats of the densification statistics
 * (gaussians_handler.py:116-118, see w2w_densify_stats).
 *
 * Clone (Densification.py:33-52): |accum/(count+1e-5)| >= grad_thresh and any
 * exp(svec) <= clone_scale_max; the copy is appended.  Split
 * (Densification.py:56-107), on the rows after cloning: |g| >= grad_thresh
 * and any scale > split_scale_min; split_factor children are appended block
 * by block, mean = R(q) (scale * z) + mean, svec = log(scale /
 * (split_factor * split_shrink)), and the split rows removed.  Prune
 * (Densification.py:167-187): sigmoid(logit) < min_opacity or max scale >
 * prune_scale_max.  Rows appended by clone or split get zero Adam moments,
 * surviving rows keep theirs (OptimizerManager.py:23-45).  The trainer's
 * values (GaussianTrainer.py:125, gaussians_handler.py:131-183) are
 * clone 0.04, split 4, factor 2, shrink 0.8, prune scale 0.4. */
typedef struct W2WStoreGroups {
  int32_t n_groups;
  int32_t width[8];
  int32_t means, opacity, svec, quat;
  float* param[8];
  float* exp_avg[8];
  float* exp_avg_sq[8];
} W2WStoreGroups;
typedef struct W2WDensifyConfig {
  float grad_thresh, min_opacity, clone_scale_max, split_scale_min, prune_scale_max, split_shrink;
  int32_t split_factor; /* 1..8 */
  int32_t pad;
} W2WDensifyConfig;
/* Replaces Densification.py:33-187 with OptimizerManager.py:23-45.  Plan:
 * per-row decisions and output positions into `temp`
 * (w2w_densify_temp_bytes(n) bytes, device; reads the parameters only) and,
 * into `counts` (8 int64, device): [0] clones, [1] splits, [2] rows removed
 * by the prune phase, [3] rows out, [4] rows in.  Apply (same store, cfg and
 * temp): writes the new store of n_out = counts[3] rows into `out` (same
 * group table, other memory).  `normals`: standard normal samples
 * [normal_rows][3], at least split_factor * counts[1] rows; child k of the
 * s-th split row (rows in the order after cloning) uses row
 * k * counts[1] + s, the order torch.normal fills the reference's
 * [factor * S, 3] sample tensor.  Apply writes nothing when n_out differs
 * from the plan's count.  The statistics restart at zero for the n_out rows
 * (gaussians_handler.py:190-191). */
size_t w2w_densify_temp_bytes(int64_t n);
/* The statistics for a batch (gaussians_handler.py:116-118, update_filter =
 * radii > 0 as Renderer.py:172): for every Gaussian and every frame of its
 * scene with the backward flag, in frame order, where radii[item] > 0:
 * count[row] += 1 and accum[row] += sqrt(gx*gx + gy*gy) of
 * means2D_grad[item] (the w2w_backward_batch output); row = the scene's
 * first parameter row + Gaussian.  Counts read on the device: capturable. */
int w2w_densify_stats(const void* table, const W2WCaps* caps, const float* means2D_grad,
                      const int32_t* radii, float* accum, float* count, void* stream);
int w2w_densify_plan(const W2WStoreGroups* store, int64_t n, const float* grad_accum,
                     const float* grad_count, const W2WDensifyConfig* cfg, void* temp,
                     size_t temp_bytes, int64_t* counts, void* stream);
int w2w_densify_apply(const W2WStoreGroups* store, int64_t n, const W2WDensifyConfig* cfg,
                      const void* temp, const float* normals, int64_t normal_rows,
                      const W2WStoreGroups* out, int64_t n_out, void* stream);

/* ---- occupancy field for mesh export (compute_voxel_field.cu:25-79) ----
 * occ[ix + iy*res + iz*res^2] = sum over Gaussians g whose mean lies in the
 * relaxed box of the grid point's block of opacities[g] * exp(power), power
 * = -0.5 d^T Sigma_g^{-1} d (0 weight when power > 0), d = point - mean;
 * grid coordinate -1 + 2 i / (res - 1) per axis, block (int)((c + 1) /
 * (2 / num_blocks)), relaxed box [-1 + b s - s r, -1 + (b + 1) s + s r] with
 * s = 2 / num_blocks and r = relax_ratio; covs [n][6] = (xx, xy, xz, yy, yz,
 * zz).  Replaces compute_voxel_field_cuda (compute_voxel_field.cu:81-124) on
 * device pointers: res >= 2, 1 <= num_blocks <= 255, 0 <= relax_ratio <= 64;
 * temp: w2w_voxel_temp_bytes(n, num_blocks) bytes.  Sums run in a fixed
 * order (deterministic), not the reference's index order. */
size_t w2w_voxel_temp_bytes(int64_t n, int32_t num_blocks);
int w2w_voxel_field(const float* means, const float* covs, const float* opacities, int64_t n,
                    int32_t resolution, int32_t num_blocks, float relax_ratio, float* occ, void* temp,
                    size_t temp_bytes, void* stream);

/* Optional per-kernel timing: HIP event pairs recorded on the launch stream
 * around every launch whose kernel id bit is set in `kernel_mask`, into a
 * pool of `capacity` records (0 disables and frees it).  Host-side state,
 * not thread-safe; for benchmarks and profiling runs.  Kernel ids: */
enum {
  W2W_KID_FILL = 0, W2W_KID_PREPROCESS = 1, W2W_KID_DSORT_HIST = 2, W2W_KID_DSORT_PASS = 3,
  W2W_KID_SCAN = 4, W2W_KID_EMIT = 5, W2W_KID_TSORT_HIST = 6, W2W_KID_TSORT_PASS = 7,
  W2W_KID_RANGES = 8, W2W_KID_BLEND = 9, W2W_KID_BLEND_BWD = 10, W2W_KID_PREPROCESS_BWD = 11,
  W2W_KID_ADAM = 12, W2W_KID_SORT_HIST = 13, W2W_KID_SORT_PASS = 14, W2W_KID_ACTIVATE = 15,
  W2W_KID_LOSS = 16, W2W_KID_ACTIVATE_BWD = 17, W2W_KID_SH_BWD = 18, W2W_KID_DENSIFY = 19,
  W2W_KID_VOXEL = 20, W2W_KID_COUNT_OPEN = 21, W2W_KID_EMIT_FAR = 22, W2W_KID_BLEND_FAR = 23,
  W2W_KID_TILE_ORDER = 24,
};
int w2w_prof_enable(int32_t capacity, uint32_t kernel_mask);
int w2w_prof_reset(void);
/* Waits for the recorded events; returns kernel id and milliseconds per record. */
int w2w_prof_read(int32_t* kernel_ids, float* ms, int32_t max_records, int32_t* n_records);

/* Optional work counters (NULL disables): the blend kernels add, per launch,
 * [0] evaluated (pixel, instance) pairs, [1] contributing pairs, [2] instances
 * loaded into LDS for the forward, and [3..5] the same for the backward, into
 * `device_counters` (uint64 device memory).  For roofline accounting. */
int w2w_stats_enable(uint64_t* device_counters);

#ifdef __cplusplus
}
#endif
#endif /* W2W_RASTER_H */

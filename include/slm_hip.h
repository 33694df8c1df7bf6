/*
 * slm_hip.h -- C ABI of libslm_hip.so, the MI355X (gfx950) streamline tracer.
 *
 * The reference (atthecodeface/slm) drives OpenCL kernels through PyOpenCL:
 * every scalar is a compile-time -D macro (python/streamlines/pocl.py:178-299)
 * and buffers are bound positionally (pocl.py:536-539).  This library replaces
 * that boundary with plain extern "C" functions:
 *   - every macro of pocl.py:225-297 becomes a field of SlmParams, passed by
 *     pointer at run time (no per-call compile);
 *   - every buffer is a caller-owned device pointer (PyTorch-ROCm tensors in the
 *     Python facade); the library never allocates caller-visible memory;
 *   - scratch memory (seed work lists, queue counters) lives in a caller buffer
 *     whose size the matching *_workspace_size() function reports;
 *   - work is enqueued on the caller's stream (hipStream_t passed as void*);
 *     nothing synchronises unless SLM_FLAG_SYNC is set in the call.
 *
 * Grid conventions follow the reference: arrays are C-order [nx_padded][ny_padded],
 * the flat index of pixel (x,y) is ny_padded*x + y (opencl/essentials.cl:56-59),
 * mask bytes are 1 = masked, uv is interleaved float2 (u,v) per pixel.
 *
 * Return codes: SLM_OK (0) or a negative SLM_E_* value; slm_last_error()
 * returns a thread-local message for the last failure on the calling thread.
 */
#ifndef SLM_HIP_H
#define SLM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLM_ABI_VERSION 2u

#define SLM_OK              0
#define SLM_E_INVALID_ARG  -1
#define SLM_E_HIP          -2
#define SLM_E_OOM          -3
#define SLM_E_ABI          -4
#define SLM_E_WORKSPACE    -5

/* Call flags */
#define SLM_FLAG_SYNC       1u   /* hipStreamSynchronize before returning            */
#define SLM_FLAG_NO_ZERO    2u   /* do not clear the output fields (accumulate into) */
#define SLM_FLAG_KEEP_ORDER 4u   /* workspace already holds the seed work order and the
                                    uv quads built by a previous slm_trace_fields call
                                    with the same seeds, n_seed_points, seeds_chunk_offset,
                                    grid and uv: reuse them */
#define SLM_FLAG_NN_DIRECT 16u     /* slm_nn_conv: use the direct kernel, not the LDS-tiled one
                                      (test hook; both compute the same products) */
#define SLM_FLAG_FORCE_FALLBACK 8u /* test hook: reject the packed 64-bit accumulation so the
                                      32-bit fallback path runs (results are identical) */
#define SLM_FLAG_FMM_LANES 32u    /* test hook: slm_fmm_hsl (second order) marches every
                                      subsegment one lane each, with no LDS heaps (same bits) */

/*
 * Every kernel constant of the reference's -D macro set (pocl.py:225-297, values
 * derived by useful.py:128-220), as run-time data.  Field names are the macro
 * names in lower case.  All fields are 4 bytes; the struct is 128 bytes.
 */
typedef struct SlmParams {
    uint32_t abi_version;                 /* must equal SLM_ABI_VERSION              */
    uint32_t nx_padded;                   /* NX_PADDED                               */
    uint32_t ny_padded;                   /* NY_PADDED                               */
    uint32_t nxy_padded;                  /* NXY_PADDED                              */
    uint32_t pad_width;                   /* PAD_WIDTH                               */
    float    pad_width_pp5;               /* PAD_WIDTH_PP5 = pad+0.5                 */
    float    nxf_mp5;                     /* NXF_MP5 = nx-0.5                        */
    float    nyf_mp5;                     /* NYF_MP5 = ny-0.5                        */
    float    grid_scale;                  /* GRID_SCALE = sqrt(nx*ny)                */
    float    combo_factor;                /* COMBO_FACTOR, sign of direction applied */
    float    dt_max;                      /* DT_MAX                                  */
    float    adjusted_max_error;          /* ADJUSTED_MAX_ERROR                      */
    float    integration_halt_threshold;  /* INTEGRATION_HALT_THRESHOLD              */
    float    max_length;                  /* MAX_LENGTH (pixels)                     */
    uint32_t max_n_steps;                 /* MAX_N_STEPS                             */
    float    trajectory_resolution;       /* TRAJECTORY_RESOLUTION (128)             */
    uint32_t subpixel_seed_point_density; /* SUBPIXEL_SEED_POINT_DENSITY (d)         */
    float    subpixel_seed_halfspan;      /* SUBPIXEL_SEED_HALFSPAN                  */
    float    subpixel_seed_step;          /* SUBPIXEL_SEED_STEP                      */
    float    jitter_magnitude;            /* JITTER_MAGNITUDE                        */
    float    downup_sign;                 /* (float)DOWNUP_SIGN: 1 or 2^32 ("-1u")   */
    uint32_t n_seed_points;               /* N_SEED_POINTS                           */
    uint32_t seeds_chunk_offset;          /* SEEDS_CHUNK_OFFSET                      */
    uint32_t fp_contract;                 /* 0: every op rounded (what the reference's
                                             own runs computed); 1: a*b+c fused where
                                             clang's OpenCL default FP_CONTRACT ON would */
    float    pixel_size;                  /* PIXEL_SIZE (metres)                     */
    uint32_t interchannel_max_n_steps;    /* INTERCHANNEL_MAX_N_STEPS (connect_channels) */
    uint32_t segmentation_threshold;      /* SEGMENTATION_THRESHOLD (0 when not rendered) */
    uint32_t approx_wrap;                 /* Q3: rounding of the Euler tail's absolute
                                             position (rungekutta.cl:115, :237), where
                                             approximate()'s (char) is out of range:
                                             0 = truncate onto the 1/TRAJECTORY_RESOLUTION
                                             lattice (the int32 value, no 8-bit wrap; what
                                             the reference's own runs computed), 1 = keep
                                             the low byte (8-bit wrap)                 */
    uint32_t reserved[4];
} SlmParams;

/* Library identity. */
uint32_t    slm_abi_version(void);
const char* slm_last_error(void);
/* Number of visible HIP devices (0 when none); never fails. */
int         slm_device_count(void);

/*
 * Lehmer RNG state after seeding, for n seed indices: out[i] = rng0(idx0+i).
 * Replaces the seeding in opencl/integratefields.cl:93-99 (byte-reverse,
 * one lehmer_rand_uint draw of opencl/rng.cl:37-45, byte-reverse).
 * Also writes n_draws successive lehmer_rand_uint states per seed into
 * draws[i*n_draws + k] when draws is non-NULL (RNG-stream check).
 */
int slm_rng_stream(uint32_t idx0, uint32_t n, uint32_t n_draws,
                   uint32_t* rng0_out, uint32_t* draws, void* stream, uint32_t flags);

/*
 * Self-check of the tracer's arithmetic: the short correctly rounded sqrt and
 * divide forms the kernels use against the IEEE operators, bit for bit, over
 * every float in their ranges (counts: device uint64 [8], added to; entries
 * 0, 2, 4, 6 and 7 count mismatches, 1, 3 and 5 the values tested, see
 * slm_hip.hip).  No counterpart in the
 * reference; it pins that the short forms change no bits on this device.
 */
int slm_check_rounding(uint64_t* counts, uint32_t n_numerators, void* stream, uint32_t flags);

/*
 * Measurement hook with no counterpart in the reference: n_blocks blocks of
 * 256 threads in which every wave issues exactly 16*n_iter independent
 * instructions of one kind (16 chains per lane; kind 0 v_fma_f32, 1 v_add_f32,
 * 2 v_add_u32, 3 v_xor_b32, 4 v_mul_f32, 5 v_fma_f32 and v_add_u32
 * alternating, 6 v_cndmask_b32, 7 v_cmp_lt_f32, 8 v_mov_b32), for timing the sustained VALU issue
 * rate and for calibrating the VALU instruction counter against a known count.
 * out: float [n_blocks], written only if the chains end on an impossible value.
 */
int slm_calibrate_valu(uint32_t kind, uint32_t n_blocks, uint32_t n_iter, float* out, void* stream,
                       uint32_t flags);

/*
 * Jittered streamline fields (slc / slt), one direction.
 * Replaces kernel integrate_fields (opencl/integratefields.cl:60-114 with
 * jittertrajectory.cl:42-88) as launched by Fields.gpu_integrate
 * (python/streamlines/fields.py:233-274).
 *   seeds_xy : float2 [n_seed_points], the shuffled seed list (seed_point_array);
 *              the RNG of seed k is keyed on seeds_chunk_offset+k.
 *   mask     : uint8 [nxy_padded]     uv : float2 [nxy_padded]
 *   slc, slt : uint32 [nxy_padded], cleared first unless SLM_FLAG_NO_ZERO.
 *   steps_out: optional uint64 device counter += accepted RK2 steps (not cleared).
 * All sub-pixel streamlines (d*d per seed) run as independent lanes pulled
 * from a work queue, so results do not depend on scheduling (integer sums).
 */
size_t slm_trace_fields_workspace_size(const SlmParams* p);
int slm_trace_fields(const SlmParams* p, const float* seeds_xy,
                     const uint8_t* mask, const float* uv,
                     uint32_t* slc, uint32_t* slt, uint64_t* steps_out,
                     void* workspace, size_t workspace_bytes,
                     void* stream, uint32_t flags);

/*
 * slm_trace_fields without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation).  The workspace lives in a fixed device array of
 * slm_trace_fields_noalloc_max_bytes() bytes (64 MiB) that every device the
 * library runs on holds until the process ends; a call whose grid and seeds
 * need more (slm_trace_fields_workspace_size(p) > the maximum, about a 1024^2
 * grid with one seed per pixel) fails with SLM_E_WORKSPACE and enqueues
 * nothing.  Results are bit-identical to slm_trace_fields.  All calls on one
 * device share the array: calls that overlap in time on one device (other
 * streams, other threads) race; calls ordered on one stream do not.
 * SLM_FLAG_KEEP_ORDER reuses the seed order and quads the previous call on
 * the device left in the array.
 */
size_t slm_trace_fields_noalloc_max_bytes(void);
int slm_trace_fields_noalloc(const SlmParams* p, const float* seeds_xy,
                             const uint8_t* mask, const float* uv,
                             uint32_t* slc, uint32_t* slt, uint64_t* steps_out,
                             void* stream, uint32_t flags);

/*
 * Same result as slm_trace_fields (bit for bit), computed by LDS-owned tiles:
 * the grid is cut into tile x tile pixel tiles; each block steps the
 * streamlines currently in one tile with that
 * tile's uv, mask and a 64-bit accumulator in LDS, and hands streamlines that
 * leave the tile (or exceed a per-visit step cap) to the owning tile for the
 * next round.  Rounds are driven from the host: this call synchronises the
 * stream once per round (so it cannot be captured in a graph).  Flags as for
 * slm_trace_fields except SLM_FLAG_KEEP_ORDER, which is ignored.  opts may be
 * NULL (all defaults); zero fields take their defaults.
 */
typedef struct SlmTileOpts {
    uint32_t tile;       /* tile side in pixels, 8..128 (0: 32)                    */
    uint32_t block;      /* threads per block: 256, 512 or 1024 (0: 256)           */
    uint32_t unit;       /* streamline states per unit of work (0: 8*block)        */
    uint32_t cap_steps;  /* steps per tile visit before re-queueing (0: 4*tile)    */
    uint32_t halo;       /* hysteresis band in pixels, <= 32 (0: 4)                */
    uint32_t reserved[3];
} SlmTileOpts;

size_t slm_trace_fields_tiled_workspace_size(const SlmParams* p, const SlmTileOpts* opts);
int slm_trace_fields_tiled(const SlmParams* p, const SlmTileOpts* opts, const float* seeds_xy,
                           const uint8_t* mask, const float* uv,
                           uint32_t* slc, uint32_t* slt, uint64_t* steps_out,
                           void* workspace, size_t workspace_bytes,
                           void* stream, uint32_t flags);

/*
 * Batched fields: many tiles (each its own grid, seeds, macro set and
 * direction) traced by one persistent launch with XCD-sliced work-stealing
 * queues.  Each job's slc/slt are bit-identical to a slm_trace_fields call
 * with the same job.  No host synchronisation, allocation or memset: the call
 * can be captured in a hipGraph and replayed.  All jobs must share
 * fp_contract.  Flags: SLM_FLAG_NO_ZERO, SLM_FLAG_KEEP_ORDER (reuse the seed
 * order and uv quads a previous call with the same jobs built),
 * SLM_FLAG_FORCE_FALLBACK, SLM_FLAG_SYNC.  steps_out (optional) += accepted
 * RK2 steps of the whole batch.  No counterpart in the reference, which
 * traces one tile per Trace.do() (trace.py:131-157); the tiles, sizes and
 * queue parameters are those of the batched configuration.
 */
typedef struct SlmFieldsJob {
    SlmParams params;        /* the tile's macro set (downup sign in combo_factor) */
    const float* seeds_xy;   /* float2 [n_seed_points] (shuffled seed list)        */
    const uint8_t* mask;     /* uint8 [nxy_padded]                                 */
    const float* uv;         /* float2 [nxy_padded]                                */
    uint32_t* slc;           /* uint32 [nxy_padded]                                */
    uint32_t* slt;           /* uint32 [nxy_padded]                                */
    uint64_t reserved[3];
} SlmFieldsJob;

typedef struct SlmBatchOpts {
    uint32_t chunk;          /* work items per queue grab (0: 64)                   */
    uint32_t blocks_per_cu;  /* resident 256-thread blocks per CU (0: occupancy max) */
    uint32_t queue_slices;   /* work-list slices / queue heads, 1..8 (0: 8, one per XCD) */
    uint32_t lds_slots;      /* per-wave LDS accumulation table: 256, 512 or 1024 slots
                                (a 16x16, 16x32 or 32x32 pixel window; 0: 512)         */
    uint32_t refill_batch;   /* idle lanes a wave waits for before it refills them, 1..64
                                (0: 28 when every job has one seed per pixel, else 20) */
    uint32_t reserved[3];
} SlmBatchOpts;

size_t slm_trace_fields_batch_workspace_size(const SlmFieldsJob* jobs, uint32_t n_jobs);
int slm_trace_fields_batch(const SlmFieldsJob* jobs, uint32_t n_jobs, const SlmBatchOpts* opts,
                           uint64_t* steps_out, void* workspace, size_t workspace_bytes,
                           void* stream, uint32_t flags);

/*
 * Un-jittered trajectory recording, one direction.
 * Replaces kernel integrate_trajectory (opencl/integratetraj.cl:54-97 with
 * trajectory.cl:41-80) as launched by Trajectories.gpu_compute_trajectories
 * (python/streamlines/trajectories.py:257-333).
 *   traj   : int8x2 [n_seed_points][max_n_steps] compressed steps
 *   nsteps : uint16 [n_seed_points]      length : float [n_seed_points]
 * Seed k (k < n_seed_points) is seeds_xy[seeds_chunk_offset + k] and writes
 * row k of traj / nsteps / length.
 */
int slm_trace_paths(const SlmParams* p, const float* seeds_xy,
                    const uint8_t* mask, const float* uv,
                    int8_t* traj, uint16_t* nsteps, float* length,
                    void* stream, uint32_t flags);

/*
 * Packing of recorded trajectories end to end on the device, replacing the
 * per-seed host slices traj[k, :nsteps[k]] that build streamline_arrays_list
 * in Trajectories.gpu_compute_trajectories (python/streamlines/
 * trajectories.py:320-325); the reference has no kernel for it.
 * slm_paths_offsets: offsets (uint64 [n + 1]) = exclusive prefix sum of
 *   min(nsteps[k], max_n_steps); offsets[n] is the total step count.
 * slm_paths_pack: packed (int8x2 [offsets[n]]) receives row k's first
 *   min(nsteps[k], max_n_steps) steps of traj (int8x2 [n][max_n_steps]) at
 *   step offsets[k], given the offsets slm_paths_offsets wrote.
 */
int slm_paths_offsets(const uint16_t* nsteps, uint32_t n, uint32_t max_n_steps,
                      uint64_t* offsets, void* stream, uint32_t flags);
int slm_paths_pack(const int8_t* traj, const uint16_t* nsteps, const uint64_t* offsets,
                   uint32_t n, uint32_t max_n_steps, int8_t* packed,
                   void* stream, uint32_t flags);

/*
 * Test hook with no reference counterpart: records the positions of selected
 * jittered fields streamlines exactly as slm_trace_fields integrates them.
 * For j < n_sel: streamline (seed_sel[j], sub_sel[j]) writes its positions
 * (start point first) into xy[j][0..max_rec) as float2 and the count into
 * nrec[j] (count includes the start; positions beyond max_rec are dropped).
 */
int slm_trace_fields_record(const SlmParams* p, const float* seeds_xy,
                            const uint8_t* mask, const float* uv,
                            const uint32_t* seed_sel, const uint32_t* sub_sel,
                            uint32_t n_sel, uint32_t max_rec,
                            float* xy, uint32_t* nrec,
                            void* stream, uint32_t flags);

/* ------------------------------------------------------------------ KDE */

/* Kernel-density filter shapes of opencl/kde.cl:36-92 (KDF_IS_* macros). */
#define SLM_KDF_TOPHAT       0u
#define SLM_KDF_TRIANGLE     1u
#define SLM_KDF_EPANECHNIKOV 2u
#define SLM_KDF_COSINE       3u
#define SLM_KDF_GAUSSIAN     4u

/*
 * Constants of the reference's 'kde' macro set (pocl.py:195-221), as data.
 * Float fields hold the value of the f-suffixed literal the reference
 * rendered (KDF_WIDTH_X, BIN_DX, PDF_DX are float literals widened to
 * double where the kernel computes in double).  64 bytes.
 */
typedef struct SlmKdeParams {
    uint32_t abi_version;          /* SLM_ABI_VERSION               */
    uint32_t kernel;               /* SLM_KDF_*                     */
    uint32_t n_data;               /* N_DATA                        */
    uint32_t n_hist_bins;          /* N_HIST_BINS                   */
    uint32_t n_pdf_points;         /* N_PDF_POINTS                  */
    uint32_t n_kdf_part_points_x;  /* N_KDF_PART_POINTS_X           */
    float    logx_min;             /* LOGX_MIN                      */
    float    logx_range;           /* LOGX_RANGE                    */
    float    bin_dx;               /* BIN_DX                        */
    float    pdf_dx;               /* PDF_DX                        */
    float    kdf_width_x;          /* KDF_WIDTH_X                   */
    uint32_t fp_contract;          /* as SlmParams.fp_contract      */
    uint32_t reserved[4];
} SlmKdeParams;

/*
 * Histogram of log values.  Replaces kernel histogram_univariate
 * (opencl/kde.cl:383-394) launched by kde.estimate_univariate_pdf
 * (python/streamlines/kde.py:165-169).
 *   logx : float [n_data]   hist : uint32 [n_hist_bins], cleared first unless
 *   SLM_FLAG_NO_ZERO.
 */
int slm_kde_histogram(const SlmKdeParams* k, const float* logx, uint32_t* hist,
                      void* stream, uint32_t flags);

/*
 * Kernel-smoothed pdf of a histogram, FP64 accumulation.  Replaces kernel
 * pdf_univariate with filter (opencl/kde.cl:426-526), launched by
 * kde.estimate_univariate_pdf (kde.py:185-189).
 *   hist : uint32 [n_hist_bins]   pdf : float [n_pdf_points]
 */
int slm_kde_pdf(const SlmKdeParams* k, const uint32_t* hist, float* pdf,
                void* stream, uint32_t flags);

/* ------------------------------------------------------------------ hillslope lengths */

/*
 * Traced hillslope lengths.  Replaces kernel hillslope_lengths
 * (opencl/lengths.cl:37-116, with lengths_runge_kutta_step of
 * rungekutta.cl:492-516) launched by lengths.hsl (python/streamlines/lengths.py:33-92).
 * From each seed (a mid-slope or ridge pixel) an un-jittered downstream RK2
 * trace runs until it reaches a pixel flagged thin_flag in mapping; it is
 * abandoned on a masked pixel, on re-entering a from_flag pixel after leaving
 * the seed band, or when stuck.  Seeds k in [seeds_chunk_offset, n_seed_points)
 * that reach a channel write length[k] = l * pixel_size; others leave length[k]
 * unchanged.  p must be built for downstream tracing.
 */
int slm_hillslope_lengths(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                          const float* uv, const uint32_t* mapping, uint32_t thin_flag,
                          uint32_t from_flag, float* length, void* stream, uint32_t flags);

/* ------------------------------------------------- grids either side of the tracer */

/*
 * The tracer's input: unit flow vectors and the mask from a DTM.
 * Replaces the host numpy of python/streamlines/preprocess.py:51-62
 * (np.gradient / pixel_size), :100-193 (u,v = -grad; speed = hypot, non-finite
 * or 0 -> 1; normalise), :584-600 (normalise again when do_normalize_speed;
 * pad 'edge'), :602-621 (NaN uv -> masked, uv = 0), merged with the DTM mask
 * (NaN -> masked, padding masked).  Conditioning of loops and blockages
 * (preprocess.py:195-473) is not applied.
 *   roi  : float [nx][ny] (NaN = no data), nx, ny >= 2
 *   uv   : float2 [nx+2pad][ny+2pad]     mask : uint8 [nx+2pad][ny+2pad]
 */
int slm_gradient_uv(const float* roi, uint32_t nx, uint32_t ny, uint32_t pad, float pixel_size,
                    int do_normalize_speed, float* uv, uint8_t* mask, void* stream, uint32_t flags);

/* DTM conditioning: Preprocess.conditioned_gradient_vector_field
 * (python/streamlines/preprocess.py:531-600, with find_blockages :647-698,
 * has_one_diagonal_outflow :390-432, upstream_of_diagonal_outflow :434-473,
 * find_and_fix_loops :303-350, break_out_of_loop :263-301, check_has_loop
 * :229-261, fix_blockages :352-388), then the padding and NaN mask of
 * mask_nan_uv (:602-621): the uv and mask of slm_gradient_uv after the
 * reference's loop and blockage fixes.  Thresholds are the preprocess JSON's
 * vecsum_threshold, divergence_threshold, curl_threshold.  maps (optional,
 * uint8 [3][nx*ny]) receives the blockage codes, the blocked-neighbour codes
 * and the loop flag of the 2x2 cell at each pixel.  The loop cells of the
 * last row and column, where the reference reads past its arrays, are never
 * loops here.  workspace: slm_condition_workspace_size(nx, ny) bytes. */
size_t slm_condition_workspace_size(uint32_t nx, uint32_t ny);
int slm_condition_uv(const float* roi, uint32_t nx, uint32_t ny, uint32_t pad, float pixel_size,
                     int do_normalize_speed, double vecsum_threshold, double divergence_threshold,
                     double curl_threshold, float* uv, uint8_t* mask, uint8_t* maps, void* workspace,
                     size_t workspace_bytes, void* stream, uint32_t flags);

/*
 * Finishing of the raw fields (fields.py:296-309): from the per-direction raw
 * sums of slm_trace_fields, slc_raw / slt_raw uint32 [2][nxy] (direction k at
 * k*nxy), write slt float [nxy][2] = slt_raw / d^2 and sla float [nxy][2] =
 * sqrt(2) slt_raw / slc_raw where slc_raw > 0 (slt = slt_raw, sla = 0
 * elsewhere), with the reference's float32 / float64 roundings.
 */
int slm_fields_finish_raw(const uint32_t* slc_raw, const uint32_t* slt_raw, uint32_t nxy,
                          uint32_t density, float* slt, float* sla, void* stream, uint32_t flags);

/*
 * Scaling of the fields (fields.py:141-164): slt [nxy][2] (from
 * slm_fields_finish_raw) *= ds_scale_k, then sqrt, in float32, where
 * ds_scale_k = float32(mean ds_k / pixel_size) per direction; slc [nxy][2]
 * = sqrt(2) (slc_raw / d^2)^(2/3) in float64 (the device pow may differ from
 * the host libm pow in the last place).
 */
int slm_fields_finish_scaled(const uint32_t* slc_raw, uint32_t nxy, uint32_t density,
                             float ds_scale_down, float ds_scale_up, float* slt, double* slc,
                             void* stream, uint32_t flags);

/* ------------------------------------------------- channel mapping and segmentation */

/*
 * Mapping flag bits: bit k of the list in python/streamlines/useful.py:249-268
 * (the Python list is canonical; opencl/info.h:114-131 differs).
 */
#define SLM_IS_CHANNEL          0x00001u
#define SLM_IS_THINCHANNEL      0x00002u
#define SLM_IS_INTERCHANNEL     0x00004u
#define SLM_IS_CHANNELHEAD      0x00008u
#define SLM_IS_CHANNELTAIL      0x00010u
#define SLM_IS_MAJORCONFLUENCE  0x00020u
#define SLM_IS_MINORCONFLUENCE  0x00040u
#define SLM_IS_MAJORINFLOW      0x00080u
#define SLM_IS_MINORINFLOW      0x00100u
#define SLM_IS_LEFTFLANK        0x00200u
#define SLM_IS_RIGHTFLANK       0x00400u
#define SLM_IS_MIDSLOPE         0x00800u
#define SLM_IS_RIDGE            0x01000u
#define SLM_WAS_CHANNELHEAD     0x02000u
#define SLM_IS_SUBSEGMENTHEAD   0x04000u
#define SLM_IS_LOOP             0x08000u
#define SLM_IS_BLOCKAGE         0x10000u
#define SLM_LEFT_FLANK_ADDITION 0x80000000u  /* label bit of a left-flank pixel (useful.py:196) */

/*
 * Conventions shared by the mapping entry points:
 *   seeds_xy : float2 [n_seed_points], the pixel list of pick_seeds
 *              (useful.py:519-538: argwhere order, minus pad_width); seeds k in
 *              [seeds_chunk_offset, n_seed_points) are processed;
 *   mask uint8, uv float2, mapping / count / link / label uint32, all
 *   [nxy_padded]; label holds the reference's int32 label bits.
 * The traces use the un-jittered RK2 step (computestep.cl:42-57) with p built
 * for downstream tracing; nxy_padded must be below 2^31.
 * Where the reference kernel reads values that other work items write in the
 * same launch, the result here follows a fixed, documented order of phases, so
 * repeated calls give identical output; on inputs where the reference's
 * outcome does not depend on its schedule, the outputs are the reference's.
 * Traces that circle (pits, flats) and walks on link cycles run on to the
 * reference's caps (1000*MAX_N_STEPS steps or hops); they are detected and
 * fast-forwarded by whole laps, which leaves every output bit unchanged
 * because all shared updates are max / or / and of lap-periodic values.
 * Calls taking a workspace need slm_map_workspace_size(p) bytes; calls that
 * share a workspace must be ordered on one stream.
 */
size_t slm_map_workspace_size(const SlmParams* p);

/* connect_channels (opencl/connect.cl:26-108), launched by
 * connect.connect_channel_pixels (python/streamlines/connect.py:32-67):
 * seeds = IS_CHANNEL pixels; ORs IS_INTERCHANNEL into unflagged pixels of
 * short traces that join two channel pixels. */
int slm_connect_channels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                         const float* uv, uint32_t* mapping, void* stream, uint32_t flags);

/* Mapping.thin_channels (mapping.py:707-719): skeletonize(medial_axis(B))
 * of B = the IS_CHANNEL | IS_INTERCHANNEL pixels, ORed into mapping as
 * IS_THINCHANNEL.  medial_axis is scikit-image's published algorithm (keep
 * table over 3x3 windows, pixels visited in order of squared distance to the
 * unset pixels, corner score, then tie-breaker): tiebreak holds the
 * tie-breaker of the n_set pixels of B in row-major order (scikit-image
 * draws numpy's RandomState(0).permutation(n_set)); n_set must equal the
 * number of pixels of B.  With tiebreak NULL the medial axis is skipped.
 * skeletonize is the Zhang-Suen thinning (scikit-image is absent, so its
 * thinning is restated from the paper).  Synchronises the stream once per
 * eight medial-axis rounds and once per eight thinning iterations; *n_iter
 * (may be NULL) receives the thinning's iteration count. */
int slm_thin_channels(const SlmParams* p, uint32_t* mapping, const uint32_t* tiebreak, uint32_t n_set,
                      void* workspace, size_t workspace_bytes, uint32_t* n_iter, void* stream, uint32_t flags);

/* map_channel_heads (opencl/channelheads.cl:32-93), launched by
 * channelheads.map_channel_heads (channelheads.py:33-79): seeds = all unmasked
 * pixels; clears IS_CHANNELHEAD where a trace arrives, unless it arrives on a
 * thin channel pixel from a non-thin one. */
int slm_map_channel_heads(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                          const float* uv, uint32_t* mapping, void* stream, uint32_t flags);

/* count_downchannels (opencl/countlink.cl:34-114), launched by
 * countlink.count_downchannels (countlink.py:36-80): seeds = channel heads;
 * traces downstream numbering pixels, flags them thin channel, links each
 * pixel to the next.  Phases: full traces rank pixels by (counter, seed);
 * traces stop on entering a pixel ranked higher by another trace; each
 * pixel links to the next pixel of its highest-counter hop. */
int slm_count_downchannels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                           const float* uv, uint32_t* mapping, uint32_t* count, uint32_t* link,
                           void* workspace, size_t workspace_bytes, void* stream, uint32_t flags);

/* flag_downchannels (opencl/countlink.cl:134-210), launched by
 * countlink.flag_downchannels (countlink.py:83-130): seeds = channel heads;
 * walks link, flagging thin channel, count = max(count, hops+1), channel
 * tail at the last unmasked pixel; unflags heads of short minor inflows.
 * Phases: head counts = 1, walks, head test on the final counts. */
int slm_flag_downchannels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                          uint32_t* mapping, uint32_t* count, const uint32_t* link,
                          void* stream, uint32_t flags);

/* label_confluences (opencl/label.cl:21-134), launched by
 * label.label_confluences (label.py:33-75): seeds = thin channel pixels;
 * flags confluences and their major / minor inflows (slt: float [nxy_padded],
 * the downstream slt field, breaks ties). */
int slm_label_confluences(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                          const float* slt, uint32_t* mapping, const uint32_t* count,
                          const uint32_t* link, void* stream, uint32_t flags);

/* segment_downchannels (opencl/segment.cl:38-107), launched by
 * segment.segment_channels (segment.py:34-93): seeds = channel heads; labels
 * channel pixels with the index of their segment's first pixel and flags
 * segment starts IS_SUBSEGMENTHEAD.  Phases: each channel pixel is owned by
 * its hop of largest (count, major inflow, index); a walk follows its owned
 * hops and marks segment starts, including where it joins another's reach;
 * then labels. */
int slm_segment_downchannels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                             uint32_t* mapping, const uint32_t* count, const uint32_t* link,
                             uint32_t* label, void* workspace, size_t workspace_bytes,
                             void* stream, uint32_t flags);

/* hillslopes (opencl/hillslopes.cl:31-91), launched by
 * hillslopes.link_hillslopes (hillslopes.py:32-75): seeds = unmasked
 * non-thin pixels; links each to the pixel its trace enters. */
int slm_link_hillslopes(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                        const float* uv, uint32_t* link, void* stream, uint32_t flags);

/* segment_hillslopes (opencl/segment.cl:133-188), launched by
 * segment.segment_hillslopes (segment.py:95-146): seeds = unmasked non-thin
 * pixels; copies the label of the thin channel pixel the trace reaches. */
int slm_segment_hillslopes(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                           const float* uv, const uint32_t* mapping, uint32_t* label,
                           void* stream, uint32_t flags);

/* subsegment_channel_edges (opencl/segment.cl:195-334), first launch of
 * segment.subsegment_flanks (segment.py:148-243): seeds = subsegment heads;
 * flags the left-hand neighbours of each channel reach IS_LEFTFLANK and sets
 * LEFT_FLANK_ADDITION in their label. */
int slm_subsegment_channel_edges(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                 uint32_t* mapping, const uint32_t* link, uint32_t* label,
                                 void* stream, uint32_t flags);

/* subsegment_flanks (opencl/segment.cl:360-414), second launch of
 * segment.subsegment_flanks: marks LEFT_FLANK_ADDITION on pixels whose trace
 * reaches a left-flank pixel before a thin channel pixel. */
int slm_subsegment_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                          const float* uv, const uint32_t* mapping, uint32_t* label,
                          void* stream, uint32_t flags);

/* segment_hillslopes (opencl/segment.cl:133-188) and subsegment_flanks
 * (:360-414) in one walk per seed, for seeds = unmasked non-thin pixels,
 * called after slm_subsegment_channel_edges (which segment.subsegment_flanks
 * runs after segment_hillslopes; neither reads what the other writes).  No
 * single reference kernel: the flank walk is the hillslope walk up to its
 * first left-flank pixel, so one walk gives both, and the labels equal those
 * of slm_segment_hillslopes, slm_subsegment_channel_edges and
 * slm_subsegment_flanks in the reference's order. */
int slm_segment_hillslopes_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                  const float* uv, const uint32_t* mapping, uint32_t* label,
                                  void* stream, uint32_t flags);

/* fix_right_flanks (opencl/segment.cl:441-504) and fix_left_flanks (:530-594),
 * third and fourth launches of segment.subsegment_flanks: relabel a pixel
 * from the pixels its trace crosses on its own side of the channel.  Reads of
 * other pixels' labels see the labels as they were when the call started. */
int slm_fix_right_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                         const float* uv, const uint32_t* mapping, uint32_t* label,
                         void* workspace, size_t workspace_bytes, void* stream, uint32_t flags);
int slm_fix_left_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                        const float* uv, const uint32_t* mapping, uint32_t* label,
                        void* workspace, size_t workspace_bytes, void* stream, uint32_t flags);

/* ------------------------------------------------------------------------
 * Surrogate kernels (no counterpart in the reference): a BF16 convolutional
 * U-Net trained on the tracer's own fields, used to predict slt/slc/channel
 * masks on large DTMs with a traced check and fallback.  Activations are
 * NHWC bf16: pixel p's channels start at base + p*ld (ld in elements), so a
 * concatenation is channel slices of one buffer.  Weights are bf16
 * [Cout][taps][Cin]; tap t of a 3x3 kernel reads row offset t/3-1 and
 * column offset t%3-1 (zero padding).  Products accumulate in fp32 on MFMA.
 * Requirements: N*H*W a multiple of 32, Cin a multiple of 16, Cout of 32,
 * ld a multiple of 8, buffers 16-byte aligned.
 * ---------------------------------------------------------------------- */

/* y = conv(x, w) + bias, then ReLU when relu != 0 (bias may be NULL). */
int slm_nn_conv(const void* x, int ldx, const void* w, const float* bias, void* y, int ldy,
                int N, int H, int W, int Cin, int Cout, int taps, int relu, void* stream, uint32_t flags);
/* The same convolution over the concatenation [up2(x2) | x]: input channels
 * [0, c_up) are read from x2, an N x H/2 x W/2 image, at (h/2, w/2) (a fused
 * nearest 2x upsample); channels [c_up, Cin) from x at channel index c
 * (x's first c_up channels are not read).  c_up a positive multiple of 32;
 * H and W even. */
int slm_nn_conv_up(const void* x, int ldx, const void* x2, int ldx2, int c_up, const void* w, const float* bias,
                   void* y, int ldy, int N, int H, int W, int Cin, int Cout, int taps, int relu, void* stream,
                   uint32_t flags);
/* wt[ci][taps-1-t][co] = w[co][t][ci]: weights of the backward-data
 * convolution (slm_nn_conv on dY with Cin and Cout exchanged). */
int slm_nn_wflip(const void* w, void* wt, int taps, int Cin, int Cout, void* stream, uint32_t flags);
/* dw[co][t][ci] += sum_p dy[p][co] x[p + off(t)][ci] (fp32), db[co] += sum_p dy[p][co]
 * (Cout <= 256; dy and x 16-byte aligned, lddy a multiple of 8). */
int slm_nn_conv_wgrad(const void* x, int ldx, const void* dy, int lddy, float* dw, float* db,
                      int N, int H, int W, int Cin, int Cout, int taps, void* stream, uint32_t flags);
/* Fused: dy[p][c] = 0 where y[p][c] <= 0 (skipped when y is NULL), then
 * db[c] += sum_p dy[p][c] (skipped when db is NULL).  C a multiple of 32, at
 * most 256; buffers 16-byte aligned, strides multiples of 8.  partials (fp32,
 * n_partials >= C floats, e.g. 512 * 256) holds per-block column sums that a
 * second kernel adds into db; NULL adds them by atomics from fewer blocks. */
int slm_nn_relu_bias_grad(const void* y, int ldy, void* dy, int lddy, float* db, int M, int C,
                          float* partials, int n_partials, void* stream, uint32_t flags);
/* dy[p][c] = 0 where y[p][c] <= 0 (ReLU backward on the stored output). */
int slm_nn_relu_bwd(const void* y, int ldy, void* dy, int lddy, int M, int C, void* stream, uint32_t flags);
/* 2x2 max pool, stride 2; H, W are the input's (even) sizes. */
int slm_nn_pool(const void* x, int ldx, void* y, int ldy, int N, int H, int W, int C, void* stream,
                uint32_t flags);
/* Max-pool backward: each window's gradient goes to its first maximum (row-major),
 * the window's other pixels get 0; with SLM_FLAG_NO_ZERO it is added to dx
 * instead (dx + gradient, dx + 0 elsewhere). */
int slm_nn_pool_bwd(const void* x, int ldx, const void* y, int ldy, const void* dy, int lddy, void* dx,
                    int lddx, int N, int H, int W, int C, void* stream, uint32_t flags);
/* Nearest 2x upsample; H, W are the output's sizes. */
int slm_nn_up(const void* x, int ldx, void* y, int ldy, int N, int H, int W, int C, void* stream,
              uint32_t flags);
/* Upsample backward: dx = sum of the 2x2 block of dy; H, W are dx's sizes. */
int slm_nn_up_bwd(const void* dy, int lddy, void* dx, int lddx, int N, int H, int W, int C, void* stream,
                  uint32_t flags);
/* a += b over M pixels x C channels. */
int slm_nn_add(void* a, int lda, const void* b, int ldb, int M, int C, void* stream, uint32_t flags);
/* Loss over valid pixels: squared error on head channels 0 (weight w_len)
 * and 1, binary cross-entropy with logits (weight w_bce) on channel 2;
 * dhead = gradient of the weighted total times inv_n (channels 3 .. min(lddh,
 * 32)-1 zero; dhead 16-byte aligned, lddh a multiple of 8); out[0..2] += the
 * unweighted summed terms. */
int slm_nn_loss(const void* head, int ldh, const float* target, const uint8_t* valid, void* dhead,
                int lddh, int M, float inv_n, float w_len, float w_bce, float* out, void* stream,
                uint32_t flags);
/* One Adam step (bias-corrected, step >= 1) on fp32 p with moments m, v;
 * writes the bf16 copy pb when not NULL. */
int slm_nn_adam(float* p, const float* g, float* m, float* v, void* pb, int n, float lr, float b1,
                float b2, float eps, int step, void* stream, uint32_t flags);

/* One layer's weights within the flat parameter vector of slm_nn_adam_flip:
 * p[offset .. offset + cout*taps*cin) as [cout][taps][cin]; wt receives them
 * as slm_nn_wflip writes them (bf16 [cin][taps][cout], taps reversed). */
typedef struct SlmNnFlip {
    uint64_t offset;
    uint32_t taps, cin, cout, reserved;
    void* wt;
} SlmNnFlip;

/* slm_nn_adam that also writes the flipped and transposed bf16 weights of up
 * to 32 layers (the backward-data weights), from the same bf16 values as pb:
 * one launch instead of one slm_nn_wflip per layer and step. */
int slm_nn_adam_flip(float* p, const float* g, float* m, float* v, void* pb, int n, float lr, float b1,
                     float b2, float eps, int step, const SlmNnFlip* layers, int n_layers, void* stream,
                     uint32_t flags);

/* Tiles: S x S windows of an nx x ny grid (x-major, pixel (i, j) at i*ny + j)
 * with top-left corners tiles[2t], tiles[2t+1] (may lie outside the grid;
 * pixels outside it count as masked).  Tile t's pixels are rows
 * t*S*S .. (t+1)*S*S - 1 of the NHWC buffers. */
/* Network input, 16 bf16 channels: (h - tile mean of valid h) / h_scale, u,
 * v, valid (1/0), div(uv), 10 x Laplacian(h), 5 x |grad h| (central
 * differences in fp32, per pixel; 0 where a 4-neighbour is masked or off the
 * grid), then zeros; mean_ws holds n_tiles floats of workspace. */
int slm_nn_tile_inputs(const float* h, const float* uv, const uint8_t* mask, int nx, int ny, const int* tiles,
                       int n_tiles, int S, float h_scale, float* mean_ws, void* x, void* stream, uint32_t flags);
/* Training targets from the raw downstream fields (uint32): log1p(slt/d^2),
 * log1p(slc/d^2), slc/d^2 >= threshold, as float32 [.., 3]; valid [..] u8. */
int slm_nn_tile_targets(const uint32_t* slc, const uint32_t* slt, const uint8_t* mask, int nx, int ny,
                        const int* tiles, int n_tiles, int S, int density, float threshold, float* target,
                        uint8_t* valid, void* stream, uint32_t flags);
/* Writes head channels 0..2 of each tile's core [halo, S-halo)^2 to
 * out[k*nx*ny + pixel] (float32, planar), for core pixels inside the grid. */
int slm_nn_tile_stitch(const void* head, int ldh, const int* tiles, int n_tiles, int S, int halo, int nx, int ny,
                       float* out, void* stream, uint32_t flags);
/* Traced raw fields (uint32) -> planar float32 out[3][nxy]: log1p(slt/d^2),
 * log1p(slc/d^2), +1/-1 for slc/d^2 >= threshold; masked pixels 0, 0, -1. */
int slm_nn_fields_log(const uint32_t* slc, const uint32_t* slt, const uint8_t* mask, size_t nxy, int density,
                      float threshold, float* out, void* stream, uint32_t flags);

/* ------------------------------------------------------------------------
 * Mid-slopes and ridges: Mapping.map_midslopes / map_ridges
 * (python/streamlines/mapping.py:779-817).
 * ---------------------------------------------------------------------- */
size_t slm_slopes_workspace_size(uint32_t nx, uint32_t ny);
/* scipy.ndimage.gaussian_filter of a float32 [nx][ny] grid (reflect borders),
 * bit for bit: axis 0 then axis 1, float32 between the passes, each output a
 * float64 sum centre*w[r] + sum_{j=r..1} (left_j + right_j)*w[r-j];
 * weights: the 2r+1 float64 weights of scipy's _gaussian_kernel1d (device). */
int slm_gaussian_filter(const float* in, float* out, float* tmp, uint32_t nx, uint32_t ny, const double* weights,
                        int radius, void* stream, uint32_t flags);
/* sla float2 [nx*ny] (downstream, upstream): angle = arctan2(dsla, usla) -
 * pi/4 (float32, arctan2 correctly rounded); mid-slopes where not masked and
 * |gauss(angle, w_mid)| <= mid_threshold; ridges where not masked and
 * gauss(angle, w_ridge) <= ridge_level (= float32(-(pi/4) * ridge_threshold)),
 * dilated twice by the 3x3 square and hole-filled (scipy binary_fill_holes).
 * Both are thinned (Zhang-Suen, the stand-in for skimage skeletonize) and
 * their skeletons OR midslope_flag / ridge_flag into mapping.  Synchronises
 * the stream once per thinning iteration and once per hole-filling launch
 * (bit-packed 64x64 tiles, repeated until no tile changes). */
int slm_map_midslopes_ridges(const float* sla, const uint8_t* mask, uint32_t nx, uint32_t ny, const double* w_mid,
                             int r_mid, float mid_threshold, const double* w_ridge, int r_ridge, float ridge_level,
                             uint32_t midslope_flag, uint32_t ridge_flag, uint32_t* mapping, void* workspace,
                             size_t workspace_bytes, void* stream, uint32_t flags);

/* ----------------------------------------------------------------------
 * FMM hillslope lengths: Mapping.fmm_lengths + remap_hsl
 * (python/streamlines/mapping.py:841-900, :953-985).  The reference calls
 * skfmm.distance once per subsegment; here all subsegments are solved at
 * once, coloured so that the 4-dilated domains of one colour are disjoint.
 * Labels are int32 [nx*ny] (flank sign kept); a subsegment is |label|.
 * ---------------------------------------------------------------------- */
/* Label pairs (lo << 32 | hi, |label| values, duplicates included) whose
 * pixels lie within Manhattan distance 2, i.e. whose dilated domains can
 * share a pixel.  colour[|label|] == 0xff marks labels that are not
 * subsegments; n_labels = the length of colour.  *count = the number of
 * pairs found; only the first cap are written. */
int slm_fmm_pairs(const int32_t* label, uint32_t nx, uint32_t ny, const uint8_t* colour, uint32_t n_labels,
                  uint64_t* pairs, uint32_t cap, uint32_t* count, void* stream, uint32_t flags);
size_t slm_fmm_workspace_size(uint32_t nx, uint32_t ny);
/* Workspace for k colours solved together (46 B per pixel per colour plus
 * tile flags); slm_fmm_hsl takes as many colours per group as its workspace
 * holds, so a larger workspace means fewer, fuller launches.
 * slm_fmm_workspace_size(nx, ny) == slm_fmm_workspace_size_k(nx, ny, 1). */
size_t slm_fmm_workspace_size_k(uint32_t nx, uint32_t ny, uint32_t k);
/* For every subsegment a (colour[a] < n_colours): phi = -1 on its
 * thin-channel pixels ((mapping & thin_flag) == thin_flag), +1 elsewhere on
 * its domain (the pixels with |label| == a and their 4-neighbours); the
 * eikonal distance from the zero contour (scikit-fmm's interface
 * initialisation, grid step dx, float64) of the given order:
 *   order 2 (scikit-fmm's default): scikit-fmm's second-order update, points
 *     accepted in (value, index) order; subsegments of more than 512 points
 *     march in LDS, one wave each (those too large for the device's LDS, and
 *     every one under SLM_FLAG_FMM_LANES, one lane each, with the same bits),
 *     partly on a second stream of the library's that joins the caller's
 *     before the call's last kernel; where an update meets a negative
 *     discriminant (scikit-fmm raises) the march of that subsegment stops and
 *     failed[a] = 1 (failed: uint8 [n_labels]);
 *   order 1: the first-order update relaxed to the fast march's fixed point
 *     (failed may be NULL; synchronises the stream after every relaxation
 *     launch, each running up to 128 in-tile sweeps on every active tile of
 *     every colour of a group).
 * Per subsegment: sums[a] = sum of round(signed distance * 2^24) (int64 bit
 * pattern) over the points a channel reaches, cnt[a] their number, nch[a]
 * the channel pixels, nphi[a] the pixels with |label| == a.  Arrays of
 * n_labels entries (n_labels <= nx*ny + 1). */
int slm_fmm_hsl(const int32_t* label, const uint32_t* mapping, uint32_t thin_flag, const uint8_t* colour,
                uint32_t n_labels, uint32_t n_colours, uint32_t nx, uint32_t ny, double dx, uint32_t order,
                uint64_t* sums, uint32_t* cnt, uint32_t* nch, uint32_t* nphi, uint8_t* failed, void* workspace,
                size_t workspace_bytes, void* stream, uint32_t flags);
/* remap_hsl: grid[i] = hsl[|label[i]|] where the signed label is selected
 * (selected[|l|] bit 0 for l > 0, bit 1 for l < 0), else 0 (float32). */
int slm_hsl_remap(const int32_t* label, uint32_t n, const int32_t* hsl, const uint8_t* selected, uint32_t n_labels,
                  float* grid, void* stream, uint32_t flags);

/* Mapping.map_hsl (mapping.py:987-1020): hsl float32 [nx*ny] rescaled to
 * uint16 over [min, max]; when dfw > 0, grey_dilation by a dfw x dfw square
 * (scipy's window and 'reflect' borders) replaces the masked pixels (mask
 * u8, 1 = masked); the mean over disk(mdr) (skimage rank.mean: float64 sum
 * of the in-image footprint values / their number, truncated) is rescaled
 * into out (float32). */
size_t slm_map_hsl_workspace_size(uint32_t nx, uint32_t ny);
int slm_map_hsl(const float* hsl, const uint8_t* mask, uint32_t nx, uint32_t ny, int dfw, int mdr, float* out,
                void* workspace, size_t workspace_bytes, void* stream, uint32_t flags);
/* Mapping.map_aspect (mapping.py:1022-1053): slope float32 [nx*ny] (degrees,
 * Preprocess.slope_array) with channel pixels ((mapping & is_channel) ==
 * is_channel) zeroed; when do_median, quantised to uint8 by max/255 and
 * median-filtered over disk(r_slope); u, v of uv (float2) quantised by 2/255
 * and median-filtered over disk(r_uv) (skimage rank.median: the smallest
 * value v with 2 * #(values <= v) >= the in-image footprint count); aspect =
 * arctan2(v, u) (float64, rounded to float32); mask = slope <
 * slope_threshold or channel.  Disk radii up to 100 pixels (sliding
 * 256-bin histograms in LDS). */
size_t slm_map_aspect_workspace_size(uint32_t nx, uint32_t ny);
int slm_map_aspect(const float* slope, const float* uv, const uint32_t* mapping, uint32_t is_channel, uint32_t nx,
                   uint32_t ny, int r_slope, int r_uv, int do_median, float slope_threshold, float* aspect,
                   uint8_t* mask, void* workspace, size_t workspace_bytes, void* stream, uint32_t flags);
/* Preprocess.slope_array (preprocess.py:597): degrees(arctan(raw speed)) of
 * the ROI float32 [nx][ny], 'edge'-padded to [(nx+2 pad)][(ny+2 pad)]. */
int slm_slope_deg(const float* roi, uint32_t nx, uint32_t ny, uint32_t pad, float pixel_size, float* slope,
                  void* stream, uint32_t flags);

#ifdef __cplusplus
}
#endif
#endif /* SLM_HIP_H */

/*
 * astra_amd.h - C ABI of the MI355X-native ASTRA projection hot path.
 *
 * Plain C: pointers, sizes and opaque handles only.  Every entry point
 * returns an int status (ASTRA_AMD_OK = 0); on failure the message of the
 * calling thread is available from astra_amd_last_error().  All data
 * pointers passed to projector / algorithm calls are DEVICE pointers to
 * C-contiguous float32 arrays; host-side arrays only appear in geometry and
 * filter construction.  Calls are asynchronous on the given hipStream_t
 * (NULL = the null stream), except where stated.
 *
 * Layouts (reference: cuda/3d/cone_fp.cu:229, cuda/3d/util3d.cu:60-75,
 * cuda/2d/par_fp.cu:124):
 *   2D volume   [rows][cols]          (row axis is world -y)
 *   2D sinogram [angles][dets]
 *   3D volume   [z][y][x]
 *   3D proj     [v][angles][u]
 * Projector outputs are accumulated (MODE_ADD) or overwritten (MODE_SET),
 * always multiplied by the output scale, like the reference kernels' "+="
 * (e.g. cuda/3d/cone_bp.cu:132) preceded by CompositeGeometryManager's
 * zeroing for MODE_SET (src/CompositeGeometryManager.cpp:1123,1157).
 */
#ifndef ASTRA_AMD_H
#define ASTRA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define ASTRA_AMD_OK 0
#define ASTRA_AMD_ERR_ARG 1         /* invalid argument / shape mismatch   */
#define ASTRA_AMD_ERR_HIP 2         /* HIP runtime error                   */
#define ASTRA_AMD_ERR_UNSUPPORTED 3 /* valid in ASTRA, not built here      */
#define ASTRA_AMD_ERR_ALLOC 4       /* device allocation failed            */
#define ASTRA_AMD_ERR_INTERNAL 5
#define ASTRA_AMD_ERR_ABORTED 6     /* the abort hook stopped a run        */

/* Write modes: CCompositeGeometryManager::EJobMode
 * (include/astra/CompositeGeometryManager.h:127-130). */
#define ASTRA_AMD_MODE_ADD 0
#define ASTRA_AMD_MODE_SET 1

/* Geometry kinds (include/astra/GeometryUtil2D.h:53-116,
 * include/astra/GeometryUtil3D.h:47-123). */
#define ASTRA_AMD_GEOM_PAR 0   /* 2D SParProjection                       */
#define ASTRA_AMD_GEOM_FAN 1   /* 2D SFanProjection                       */
#define ASTRA_AMD_GEOM_PAR3D 2 /* 3D SPar3DProjection                     */
#define ASTRA_AMD_GEOM_CONE 3  /* 3D SConeProjection                      */
#define ASTRA_AMD_GEOM_CYLCONE 4 /* 3D SCylConeProjection (cylindrical
                                    detector, include/astra/GeometryUtil3D.h:
                                    125-150)                                */

/* 3D projection kernels: Cuda3DProjectionKernel
 * (include/astra/cuda/3d/dims3d.h:45-51). */
#define ASTRA_AMD_KER3D_DEFAULT 0
#define ASTRA_AMD_KER3D_SUM_SQUARE_WEIGHTS 1 /* par3d FP: squared weights  */
#define ASTRA_AMD_KER3D_FDK_WEIGHTING 2
#define ASTRA_AMD_KER3D_2D_WEIGHTING 3
#define ASTRA_AMD_KER3D_MATCHED_BP 4 /* BP = exact transpose of the FP     */

/* Filter windows of genFilter (E_FBPFILTER, include/astra/Filters.h:42-66;
 * src/Filters.cpp:151-472), applied to the Ram-Lak spectrum. */
#define ASTRA_AMD_FILTER_RAMLAK 0
#define ASTRA_AMD_FILTER_SHEPPLOGAN 1
#define ASTRA_AMD_FILTER_COSINE 2
#define ASTRA_AMD_FILTER_HAMMING 3
#define ASTRA_AMD_FILTER_HANN 4
#define ASTRA_AMD_FILTER_TUKEY 5
#define ASTRA_AMD_FILTER_LANCZOS 6
#define ASTRA_AMD_FILTER_TRIANGULAR 7
#define ASTRA_AMD_FILTER_GAUSSIAN 8
#define ASTRA_AMD_FILTER_BARTLETTHANN 9
#define ASTRA_AMD_FILTER_BLACKMAN 10
#define ASTRA_AMD_FILTER_NUTTALL 11
#define ASTRA_AMD_FILTER_BLACKMANHARRIS 12
#define ASTRA_AMD_FILTER_BLACKMANNUTTALL 13
#define ASTRA_AMD_FILTER_FLATTOP 14
#define ASTRA_AMD_FILTER_KAISER 15
#define ASTRA_AMD_FILTER_PARZEN 16

/* SProjectorParams2D (include/astra/cuda/2d/dims.h:40-59). */
typedef struct {
    unsigned int rays_per_det;       /* iRaysPerDet (FP supersampling)     */
    unsigned int rays_per_pixel_dim; /* iRaysPerPixelDim (BP)              */
    float output_scale;              /* fOutputScale                       */
} astra_amd_params2d;

/* SProjectorParams3D (include/astra/cuda/3d/dims3d.h:53-66); volScale
 * lives in the geometry handle. */
typedef struct {
    unsigned int rays_per_det_dim;   /* iRaysPerDetDim                     */
    unsigned int rays_per_voxel_dim; /* iRaysPerVoxelDim                   */
    float output_scale;              /* fOutputScale                       */
    int kernel;                      /* ASTRA_AMD_KER3D_*                  */
} astra_amd_params3d;

typedef struct astra_amd_geom2d_s *astra_amd_geom2d;
typedef struct astra_amd_geom3d_s *astra_amd_geom3d;
typedef struct astra_amd_filter_s *astra_amd_filter;
typedef struct astra_amd_algo_s *astra_amd_algo;

/* ------------------------------------------------------------- library */
/* Thread-local message of the last failing call on this thread
 * (CLogger::getLastErrMsg, include/astra/Logging.h:169). */
const char *astra_amd_last_error(void);
/* Library version string. */
const char *astra_amd_version(void);
/* Make device `index` current for this host thread
 * (astraCUDA::setGPUIndex, cuda/2d/astra.cu:94-105). */
int astra_amd_set_device(int index);
/* Block until all work on `stream` is complete. */
int astra_amd_stream_sync(void *stream);

/* ----------------------------------------------------- 2D geometry ---- */
/* Build a device-resident 2D geometry from NORMALISED projection vectors
 * (voxel units, volume centred at 0): 6 floats per angle,
 * [rayX,rayY,dSX,dSY,uX,uY] (PAR) or [srcX,srcY,dSX,dSY,uX,uY] (FAN).
 * Replaces Geometry2DParameters (include/astra/GeometryUtil2D.h:119-164)
 * plus the per-call transferConstants of cuda/2d/{par,fan}_{fp,bp}.cu.
 * Synchronous; the handle belongs to the current device. */
int astra_amd_geom2d_create(int kind, unsigned int vol_width, unsigned int vol_height,
                            unsigned int n_angles, unsigned int n_dets, const float *vecs,
                            astra_amd_geom2d *out);
int astra_amd_geom2d_destroy(astra_amd_geom2d g);

/* ----------------------------------------------------- 3D geometry ---- */
/* Build a device-resident 3D geometry from NORMALISED vectors: 12 doubles
 * per angle [s, dS, u, v] (s = source for CONE, ray for PAR3D).
 * CYLCONE takes 13 doubles per angle [src, detector centre, u, v, radius],
 * TRANSLATED by the volume window centre but not scaled
 * (convertAstraGeometry_unscaled_internal, src/GeometryUtil3D.cpp:595-616).
 * Replaces Geometry3DParameters (include/astra/GeometryUtil3D.h:160-224)
 * and the transferConstants of cuda/3d/{cone,par3d}_{fp,bp}.cu,
 * cuda/3d/cone_cyl.cu:75-123 and cuda/3d/cone_cyl_bp.cu:162-232.
 * vol_scale = SVolScale3D (voxel size x, y, z).  Synchronous. */
int astra_amd_geom3d_create(int kind, unsigned int vol_x, unsigned int vol_y,
                            unsigned int vol_z, unsigned int n_angles, unsigned int det_u,
                            unsigned int det_v, const double *vecs, const float *vol_scale,
                            astra_amd_geom3d *out);
int astra_amd_geom3d_destroy(astra_amd_geom3d g);

/* ------------------------------------------------------ projectors ---- */
/* 2D forward projection: astraCUDA::FP (include/astra/cuda/2d/mem2d.h:77-80,
 * cuda/2d/mem2d.cu:224-261). */
int astra_amd_fp2d(astra_amd_geom2d g, const float *vol, float *sino,
                   const astra_amd_params2d *p, int mode, void *stream);
/* 2D forward projection that allocates no device memory.  No counterpart
 * in the reference: it is for callers that cannot call hipMalloc (code
 * under stream capture, tight serving loops, harnesses that forbid
 * allocation).  Same arguments and result as astra_amd_fp2d; the volume's
 * two padded copies live in a fixed 64 MiB array per GPU instead of
 * stream-ordered scratch, so h * (w + 4) + w * (h + 4) must not exceed
 * 16 Mi floats (about 2040 x 2040), else ASTRA_AMD_ERR_ARG.  The array is
 * shared by every call on that GPU: calls that overlap on one GPU (other
 * streams, other host threads) race. */
int astra_amd_fp2d_noalloc(astra_amd_geom2d g, const float *vol, float *sino,
                           const astra_amd_params2d *p, int mode, void *stream);
/* 2D back projection: astraCUDA::BP (include/astra/cuda/2d/mem2d.h:81-83,
 * cuda/2d/mem2d.cu:278-315). */
int astra_amd_bp2d(astra_amd_geom2d g, const float *sino, float *vol,
                   const astra_amd_params2d *p, int mode, void *stream);
/* Fan-beam FBP-weighted back projection: FanBP_FBPWeighted
 * (cuda/2d/fan_bp.cu:311-336). FAN geometries only. */
int astra_amd_bp2d_fbp_weighted(astra_amd_geom2d g, const float *sino, float *vol,
                                const astra_amd_params2d *p, int mode, void *stream);
/* Pitched variants of fp2d / bp2d / fp3d / bp3d: the device arrays have
 * rows `*_pitch` floats apart (>= the row width: nx / nu / nd / w), as the
 * reference's pitched GPU data objects (cudaMalloc3D / cudaMallocPitch,
 * cuda/3d/mem3d.cu:76-118, cuda/2d/mem2d.cu:54-92).  A
 * pitch equal to the width runs in place; otherwise the array goes through
 * a stream-ordered contiguous copy (padding columns are left untouched). */
int astra_amd_fp2d_pitched(astra_amd_geom2d g, const float *vol, size_t vol_pitch, float *sino,
                           size_t sino_pitch, const astra_amd_params2d *p, int mode,
                           void *stream);
int astra_amd_bp2d_pitched(astra_amd_geom2d g, const float *sino, size_t sino_pitch, float *vol,
                           size_t vol_pitch, const astra_amd_params2d *p, int mode,
                           void *stream);
int astra_amd_fp3d_pitched(astra_amd_geom3d g, const float *vol, size_t vol_pitch, float *proj,
                           size_t proj_pitch, const astra_amd_params3d *p, int mode,
                           void *stream);
int astra_amd_bp3d_pitched(astra_amd_geom3d g, const float *proj, size_t proj_pitch, float *vol,
                           size_t vol_pitch, const astra_amd_params3d *p, int mode,
                           void *stream);
/* 3D forward projection: astraCUDA3d::FP (include/astra/cuda/3d/mem3d.h:107-108,
 * cuda/3d/mem3d.cu:334-398). */
int astra_amd_fp3d(astra_amd_geom3d g, const float *vol, float *proj,
                   const astra_amd_params3d *p, int mode, void *stream);
/* 3D back projection: astraCUDA3d::BP (include/astra/cuda/3d/mem3d.h:109-110,
 * cuda/3d/mem3d.cu:416-485).  kernel DEFAULT / FDK_WEIGHTING / 2D_WEIGHTING
 * follow dev_cone_BP / dev_par3D_BP; MATCHED_BP is the exact transpose of
 * astra_amd_fp3d (no reference counterpart for cone/par3d). */
int astra_amd_bp3d(astra_amd_geom3d g, const float *proj, float *vol,
                   const astra_amd_params3d *p, int mode, void *stream);
/* Exact transpose of astra_amd_fp2d (no reference counterpart). */
int astra_amd_bp2d_matched(astra_amd_geom2d g, const float *sino, float *vol,
                           const astra_amd_params2d *p, int mode, void *stream);

/* ------------------------------------------------------ FDK / filter -- */
/* Ram-Lak half spectrum of genFilter (src/Filters.cpp:133-160,473-485):
 * n_fft = nextpow2(2*n_det) (src/Filters.cpp:662-669), out has
 * n_fft/2+1 floats.  Host-only, synchronous. */
int astra_amd_ramlak_spectrum(unsigned int n_det, float d, float *out, unsigned int n_out);
/* Half spectrum (n_fft/2+1 floats) of a windowed ramp filter, genFilter
 * (src/Filters.cpp:133-485): type ASTRA_AMD_FILTER_*, d = FilterD (bins
 * above pi*d are zeroed), param = FilterParameter (< 0: the reference's
 * default for TUKEY / GAUSSIAN / BLACKMAN / KAISER).  Host-only. */
int astra_amd_filter_spectrum(int type, unsigned int n_det, float d, float param, float *out,
                              unsigned int n_out);
/* Upload a real half spectrum (n_fft/2+1 floats, host) for rows of
 * n_det samples.  Replaces prepareCuFFTFilter (cuda/2d/fft.cu:375-604).
 * half_spectrum = NULL with n_bins = 0 makes the identity filter
 * (FilterType 'none': no filtering). */
int astra_amd_filter_create(unsigned int n_det, const float *half_spectrum,
                            unsigned int n_bins, astra_amd_filter *out);
/* User-supplied filters, prepareCuFFTFilter FILTER_PROJECTION / SINOGRAM /
 * RPROJECTION / RSINOGRAM (cuda/2d/fft.cu:447-595).  n_rows = 1 (one filter
 * for every row) or the angle count (row r of the data uses filter
 * r % n_rows).  real_space = 0: each row holds n_fft/2+1 real Fourier bins;
 * real_space = 1: each row is a real-space kernel of `width` taps centred
 * at width/2, wrapped into n_fft and transformed (complex spectrum). */
int astra_amd_filter_create_custom(unsigned int n_det, unsigned int n_rows, int real_space,
                                   const float *data, unsigned int width,
                                   astra_amd_filter *out);
int astra_amd_filter_destroy(astra_amd_filter f);
/* Filter n_rows contiguous rows of n_det floats in place:
 * zero-pad to n_fft, x spectrum, inverse, /n_fft, keep n_det
 * (cuda/3d/fdk.cu:351-398, cuda/2d/fft.cu:52-131). */
int astra_amd_filter_rows(astra_amd_filter f, float *rows, size_t n_rows,
                          void *stream);
/* FDK cosine pre-weighting in place: devFDK_preweight (cuda/3d/fdk.cu:70-113)
 * with the scalar geometry FDK() extracts from angle 0 (fdk.cu:427-445).
 * CONE geometries only. */
int astra_amd_fdk_preweight(astra_amd_geom3d g, float *proj, void *stream);
/* Full FDK: pre-weight + Ram-Lak filter (filter handle, NULL = Ram-Lak
 * built internally) + FDK-weighted cone BP, astraCUDA3d::FDK
 * (cuda/3d/fdk.cu:415-478).  `proj` is used as scratch and overwritten
 * with the filtered projections.  short_scan != 0 adds the Parker
 * weights of FDK_PreWeight (cuda/3d/fdk.cu:115-163, 201-251). */
int astra_amd_fdk(astra_amd_geom3d g, float *proj, float *vol, float output_scale,
                  int short_scan, astra_amd_filter f, int mode, void *stream);
/* The filtering half of astra_amd_fdk, in place: pre-weights (+ Parker
 * weights when short_scan != 0) and the row filter (cuda/3d/fdk.cu:
 * 415-465).  An FDK over a volume split into slabs filters once and then
 * back projects into each slab with astra_amd_bp3d (kernel FDK_WEIGHTING)
 * on the slab's geometry (CCompositeGeometryManager::doFDK splits the
 * volume the same way, src/CompositeGeometryManager.cpp:1283-1304). */
int astra_amd_fdk_filter(astra_amd_geom3d g, float *proj, int short_scan, astra_amd_filter f,
                         void *stream);
/* 2D FBP: astraCUDA::FBP (cuda/2d/fbp.cu:77-190).  FAN: one-row cone
 * pre-weighting (fbp.cu:106-148, with Parker weights when short_scan != 0),
 * filter, FBP-weighted fan BP; PAR: filter, BP with output_scale *
 * (pi/2)/n_angles (short_scan has no effect, as in the reference).  The
 * caller folds the 1/pixel-area factor of
 * CCudaFilteredBackProjectionAlgorithm::run
 * (src/CudaFilteredBackProjectionAlgorithm.cpp:210-212) into output_scale.
 * `sino` is overwritten with the filtered sinogram.  f = NULL: Ram-Lak. */
int astra_amd_fbp2d(astra_amd_geom2d g, float *sino, float *vol, float output_scale,
                    int short_scan, astra_amd_filter f, int mode, void *stream);

/* ------------------------------------------------- batched 2D SIRT --- */
/* Many independent 2D slices (parallel or fan, each with its own volume
 * size, angle count and detector count) reconstructed together.  No
 * reference counterpart: replaces running one CCudaSirtAlgorithm per slice
 * (src/CudaSirtAlgorithm.cpp:284-397).  Per slice i: kinds[i] (GEOM_PAR /
 * GEOM_FAN), widths/heights (volume cols/rows), n_angles, n_dets,
 * output_scales (the 2D fOutputScale, i.e. the pixel length); vecs holds
 * the normalised 6-float vectors of all slices' angles, slice after slice.
 * Volumes are packed [h][w] slice after slice, sinograms [n_angles][n_dets]
 * slice after slice (offsets from astra_amd_batch2d_offsets). */
typedef struct astra_amd_batch2d_s *astra_amd_batch2d;
int astra_amd_batch2d_create(unsigned int n_slices, const int *kinds, const unsigned int *widths,
                             const unsigned int *heights, const unsigned int *n_angles,
                             const unsigned int *n_dets, const float *vecs,
                             const float *output_scales, astra_amd_batch2d *out);
/* Element offsets of every slice in the packed volume / sinogram buffers
 * (n_slices + 1 entries each; the last is the total).  Either may be NULL. */
int astra_amd_batch2d_offsets(astra_amd_batch2d b, size_t *vol_offsets, size_t *sino_offsets);
/* sinos = A vols for every slice (MODE_SET). */
int astra_amd_batch2d_fp(astra_amd_batch2d b, const float *vols, float *sinos, void *stream);
/* `iterations` SIRT iterations on every slice (vols = initial iterates,
 * overwritten), with relaxation and min/max constraints (+-INFINITY = off).
 * use_graph: 0 = three launches per iteration (slice copies, FP, BP) on
 * `stream`; 1 = the same launches captured once into a hipGraph and
 * replayed as one graph launch (re-captured when arguments change); 2 = the
 * whole run as ONE persistent kernel launch over device work queues
 * (captured in a graph with the reset of its queue counters); it returns
 * when the run has ended, with ASTRA_AMD_ERR_ARG if a work item waited past
 * the spin limit (ASTRA_AMD_BATCH_SPIN_LIMIT polls, default 2^23); a run
 * whose work-item count would pass 2^31 is split into consecutive
 * persistent launches.  All modes give the same floats. */
int astra_amd_batch2d_sirt(astra_amd_batch2d b, const float *sinos, float *vols, int iterations,
                           float relaxation, float min_constraint, float max_constraint,
                           int use_graph, void *stream);
int astra_amd_batch2d_destroy(astra_amd_batch2d b);

/* ------------------------------------------------------ profiling ---- */
/* Per-kernel event timing (no counterpart in the reference, which has no
 * tracing): while enabled, every hot kernel launch is bracketed by a
 * hipEvent pair on its own stream.  enable(1) clears previous records. */
int astra_amd_profile_enable(int on);
/* Totals for one kernel name ("fp3d_cone", "fp3d_par3d", "bp3d_cone",
 * "bp3d_fdk", "bp3d_par3d", "fp2d_par", "fp2d_fan", "bp2d_par", "bp2d_fan",
 * "bp2d_fbp", "filter_rows"): summed duration, algorithmic work units
 * (ray-slice samples, voxel-angle updates or filtered rows) and launches.
 * Synchronises the recorded events. */
int astra_amd_profile_read(const char *kernel, double *total_ms, double *units,
                           unsigned long long *launches);

/* Device memory of the library's stream-ordered scratch (the current
 * device's default memory pool, which also serves any other
 * hipMallocAsync user in the process): bytes in use now and the
 * high-water mark since the last reset; reset_peak != 0 resets the mark
 * to the current use after reading.  No reference counterpart (a
 * measurement utility). */
int astra_amd_scratch_usage(size_t *in_use, size_t *peak, int reset_peak);
/* Ceiling of the CU's vector memory path for the projectors' tap loads
 * (no reference counterpart; a measurement utility): every wave of a full
 * grid issues only global loads from a 4 KiB window per workgroup that
 * stays in L1.  mode 0: 16-byte loads (global_load_dwordx4) with the 3D
 * taps' lane pattern (16 columns x 4 rows of 8-byte pairs, one texel of
 * per-lane jitter); mode 1: 16-byte loads, coalesced (1 KiB per wave);
 * mode 2: 8-byte loads at 4-byte alignment, the 64 lanes 1.2 floats apart
 * along one row (the 2D projectors' linear fetch); mode 3: 16-byte loads
 * at 4-byte alignment, lanes 1.8 floats apart (two 2D taps per load);
 * mode 4: mode 3 with addresses rounded down to 8-byte alignment.
 * Returns lane-loads per second (each lane of each load instruction
 * counts once), best of 3 timed launches after a warm-up.  Synchronous;
 * allocates and frees its own 1 MiB buffer. */
int astra_amd_calibrate_gather(int mode, unsigned int iterations, double *lane_loads_per_second,
                               void *stream);

/* ---------------------------------------------------- elementwise ---- */
/* Sum of a[i]*b[i] accumulated in float64 (dotProduct2D / dotProduct3D,
 * cuda/2d/util.cu:304-357, cuda/3d/mem3d.cu:581-590). Synchronous. */
int astra_amd_dot(const float *a, const float *b, size_t n, double *out, void *stream);

/* ------------------------------------------------------ algorithms ---- */
/* Iterative algorithms over device data, the GPU algorithm objects
 * CCudaSirtAlgorithm / CCudaCglsAlgorithm / CCudaSirtAlgorithm3D /
 * CCudaCglsAlgorithm3D (src/CudaSirtAlgorithm.cpp:284-397,
 * src/CudaCglsAlgorithm.cpp:166-266, src/CudaSirtAlgorithm3D.cpp:301-381,
 * src/CudaCglsAlgorithm3D.cpp:264-360).  The geometry handle must outlive
 * the algorithm. */
#define ASTRA_AMD_ALGO_SIRT 0
#define ASTRA_AMD_ALGO_CGLS 1
#define ASTRA_AMD_ALGO_EM 2   /* CCudaEMAlgorithm (src/CudaEMAlgorithm.cpp)   */
#define ASTRA_AMD_ALGO_SART 3 /* CCudaSartAlgorithm (src/CudaSartAlgorithm.cpp), 2D */
int astra_amd_algo2d_create(int type, astra_amd_geom2d g, const astra_amd_params2d *p,
                            astra_amd_algo *out);
int astra_amd_algo3d_create(int type, astra_amd_geom3d g, const astra_amd_params3d *p,
                            astra_amd_algo *out);
/* Options: "Relaxation", "MinConstraint", "MaxConstraint"
 * (src/CudaSirtAlgorithm.cpp:80-91, src/ReconstructionAlgorithm2D.cpp). */
int astra_amd_algo_set_option(astra_amd_algo a, const char *key, double value);
/* Optional masks (device pointers, lifetime >= algorithm): "ReconstructionMask"
 * (volume shape), "SinogramMask" (projection shape), "MinMask", "MaxMask". */
int astra_amd_algo_set_mask(astra_amd_algo a, const char *key, const float *mask);
/* SART projection order (ProjectionOrder 'sequential' / 'random' /
 * 'custom' with ProjectionOrderList, src/CudaSartAlgorithm.cpp:79-107):
 * the angle of iteration k is order[k % n]; n = 0 restores 0, 1, 2, ...
 * Host array, copied. */
int astra_amd_algo_set_order(astra_amd_algo a, const int *order, unsigned int n);
/* run(iterations): vol is read as the initial iterate and overwritten.
 * With an abort hook set, the hook is asked before every iteration and a
 * non-zero answer stops the run with ASTRA_AMD_ERR_ABORTED, vol holding the
 * iterate reached (at most 3 iterations are then queued on the GPU ahead
 * of the host). */
int astra_amd_algo_run(astra_amd_algo a, const float *sino, float *vol, int iterations,
                       void *stream);
/* Process-wide abort hook of the iterative algorithms, replacing
 * setShouldAbortHook / shouldAbort (include/astra/Globals.h:156-164; the
 * reference asks it once per iteration, src/CudaSirtAlgorithm.cpp:347).
 * fn = NULL removes it. */
typedef int (*astra_amd_abort_hook)(void *user);
int astra_amd_set_abort_hook(astra_amd_abort_hook fn, void *user);
/* ||b - A x|| (getResidualNorm). Synchronous. */
int astra_amd_algo_residual_norm(astra_amd_algo a, const float *sino, const float *vol,
                                 double *out, void *stream);
int astra_amd_algo_destroy(astra_amd_algo a);

#ifdef __cplusplus
}
#endif

#endif /* ASTRA_AMD_H */

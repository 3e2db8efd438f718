/*
 * yrtmi - MI355X (gfx950) OSEM projection library: C ABI.
 *
 * Plain C types only: device pointers are passed as plain pointers, sizes as
 * int64_t, streams as void* (a hipStream_t; NULL = the null stream).  Every
 * call returns 0 on success or a negative YRTMI_E* code; yrtmi_last_error()
 * returns a thread-local message describing the last failure.
 *
 * Ownership: the caller owns every buffer passed in (host or device).  A
 * context owns only what it uploads itself (detector LUT, histogram tables,
 * motion transforms, PSF tables) and frees it in yrtmi_ctx_destroy.  Device
 * image / projection buffers handed in are aliases (e.g. torch tensors) and
 * are never freed by the library.
 *
 * Calls are asynchronous on the context's stream unless stated otherwise;
 * yrtmi_ctx_synchronize() waits.  A context must not be used by two host
 * threads at once (the reference's projector objects have the same rule:
 * OperatorProjector.cpp:358-364).
 *
 * Each entry point names the reference (YRT-PET 2.0.7) interface it replaces.
 */
#ifndef YRTMI_H
#define YRTMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YRTMI_OK 0
#define YRTMI_EINVAL -1   /* bad argument / shape */
#define YRTMI_EHIP -2     /* HIP runtime failure */
#define YRTMI_ENOMEM -3   /* allocation failure */
#define YRTMI_ESTATE -4   /* call out of order (e.g. no LUT uploaded) */

#define YRTMI_PROJ_SIDDON 0
#define YRTMI_PROJ_DD 1
#define YRTMI_MAX_RAYS 16

/* Image geometry (ImageParams, include/yrt-pet/datastruct/image/
 * ImageParams.hpp).  Voxel data are float[nt][nz][ny][nx]. */
typedef struct
{
	int64_t nx, ny, nz, nt;
	float vx, vy, vz;          /* voxel size, mm */
	float lx, ly, lz;          /* image length, mm */
	float ox, oy, oz;          /* offset of the image centre, mm */
	float fov_radius;          /* ImageParams::setup */
} yrtmi_image;

/* Projector (ProjectorParams, include/yrt-pet/operators/
 * ProjectorParams.hpp:27-71, plus the projection-space PSF of
 * ProjectionPsfManager.hpp:13-35 as rows of taps). */
typedef struct
{
	int type;                  /* YRTMI_PROJ_SIDDON / YRTMI_PROJ_DD */
	int num_rays;              /* multi-ray Siddon, 1..YRTMI_MAX_RAYS */
	int use_tof;
	float tof_width_ps;        /* FWHM */
	int tof_n_std;             /* <= 0: untruncated Gaussian */
	const float* psf_kernels;  /* host, psf_num_s x psf_ksize, or NULL */
	int psf_num_s, psf_ksize;
	float psf_s_step, psf_k_spacing;
	float crystal_size_z, crystal_size_trans, crystal_depth;
	/* 0 (default): ray-plane setup in FP32 with the reference's operation
	 * order (reference parity, including its FP32 error on rays grazing voxel
	 * planes); 1: setup in FP64 (accuracy mode). */
	int setup_fp64;
} yrtmi_projector_desc;

/* Histogram3D geometry (Histogram3D ctor, src/datastruct/projection/
 * Histogram3D.cpp:165-203) derived from a regular scanner. */
typedef struct
{
	int64_t dets_per_ring, num_rings, num_doi, max_ring_diff, min_ang_diff;
} yrtmi_histo_desc;

/* Where the LORs of a call come from.
 *   YRTMI_SRC_PAIRS        : d1[i], d2[i]                 (list-mode)
 *   YRTMI_SRC_BINS         : Histogram3D bin ids bins[i]  (e.g. nonzero bins)
 *   YRTMI_SRC_HISTO_SUBSET : bin = BinIteratorRangeHistogram3D(S, s).get(i)
 *                            (every bin of a subset; n must be that size)
 * All pointers are DEVICE pointers; tof/frame/motion_frame are nullable.
 * frame: dynamic frame (< 0 skips the event, ProjectorSiddon.cpp:191-196).
 * motion_frame: index into the context's motion transforms. */
#define YRTMI_SRC_PAIRS 0
#define YRTMI_SRC_BINS 1
#define YRTMI_SRC_HISTO_SUBSET 2
typedef struct
{
	int kind;
	int64_t n;
	const uint32_t* d1;
	const uint32_t* d2;
	const uint32_t* bins;
	int num_subsets, subset;
	const float* tof;
	const int32_t* frame;
	const uint8_t* motion_frame;
} yrtmi_lors;

/* Per-LOR correction factors for the fused EM kernel (OSEM_CPU.cpp:329-374,
 * OSEM_GPU.cu:1159-1341).  DEVICE pointers, all nullable:
 * meas NULL = list-mode measurement 1.0 (ListMode.cpp:52-55).  The
 * denominator is ((A x) * sens * att * global_scale + randoms + scatter) *
 * precorr, precorr the in-vivo attenuation precorrection factor
 * (ATTENUATION_PRECORRECTION, OSEM_GPU.cu:1303-1308); the backprojected
 * ratio is multiplied by sens * att * global_scale. */
typedef struct
{
	const float* meas;
	const float* sens;
	const float* att;
	const float* randoms;
	const float* scatter;
	float global_scale;
	float denom_threshold;
	const float* precorr;
} yrtmi_corrections;

/* Regular cylindrical scanner (Scanner ctor, src/datastruct/scanner/
 * Scanner.cpp:101-127). */
typedef struct
{
	float axial_fov, crystal_size_z, crystal_size_trans, crystal_depth, radius;
	int64_t dets_per_ring, num_rings, num_doi, dets_per_block;
} yrtmi_scanner_desc;

typedef struct yrtmi_ctx yrtmi_ctx;
typedef struct yrtmi_proj yrtmi_proj;

/* ---- context ------------------------------------------------------------
 * Replaces the device selection / stream setup of OperatorProjectorDevice::
 * create (src/operators/OperatorProjectorDevice.cu:300-319) and GPUStream. */
const char* yrtmi_last_error(void);
int yrtmi_ctx_create(int device, void* stream, yrtmi_ctx** out);
int yrtmi_ctx_destroy(yrtmi_ctx* ctx);
int yrtmi_ctx_set_stream(yrtmi_ctx* ctx, void* stream);
int yrtmi_ctx_synchronize(yrtmi_ctx* ctx);
/* Give the context a stream of its own (created here, default priority,
 * destroyed with the context) and make it the context's stream; *stream
 * (HOST, nullable) receives it.  Graph capture needs a stream other than the
 * null stream. */
int yrtmi_ctx_own_stream(yrtmi_ctx* ctx, void** stream);

/* ---- graphs: a recorded sequence of calls, replayed as one launch -------
 * No counterpart in the reference (which launches every kernel from the
 * host each time).  yrtmi_graph_begin starts recording the context's stream
 * (not the null stream): the calls made on the context until
 * yrtmi_graph_end are recorded, not run; yrtmi_graph_launch replays them
 * with the same buffers.  Recordable: forward / back / em_subset /
 * em_ratio / the *_planned calls / sensitivity / em_update / threshold /
 * fill / axpby / gather / subset_interleaved / permute / forward_f64 /
 * back_f64.
 * Calls that allocate, copy or synchronise (plan_create, count_segments,
 * histo_subset_nonzero, lor_order, copy, device_alloc/free, lut / histo /
 * motion uploads, and a first per-XCD backprojection that would allocate its
 * partial images) fail with YRTMI_ESTATE while recording. */
typedef struct yrtmi_graph yrtmi_graph;
int yrtmi_graph_begin(yrtmi_ctx* ctx);
int yrtmi_graph_end(yrtmi_ctx* ctx, yrtmi_graph** out);
int yrtmi_graph_launch(yrtmi_ctx* ctx, const yrtmi_graph* g);
int yrtmi_graph_destroy(yrtmi_graph* g);

/* Detector LUT (host arrays n_det x 3 each), kept resident on the device as
 * SoA.  Replaces the per-batch host LOR precompute (LORsDevice.cu:28-55,
 * OSEM_GPU.cu:698-805) and Scanner::createLUT (Scanner.cpp:199-213). */
int yrtmi_lut_upload(yrtmi_ctx* ctx, const float* pos, const float* orient,
                     int64_t n_det);

/* Regular detector LUT on the host: pos / orient are host arrays of
 * dets_per_ring*num_rings*num_doi x 3 floats.  Replaces
 * DetRegular::generateLUT (src/datastruct/scanner/DetRegular.cpp:113-182). */
int yrtmi_lut_regular(const yrtmi_scanner_desc* s, float* pos, float* orient);

/* Histogram3D bin <-> detector-pair tables on the device (closed form; the
 * while loop of Histogram3D::getDetPairFromCoords, Histogram3D.cpp:339-357,
 * becomes a z-bin table).  Needed by the BINS / HISTO_SUBSET sources. */
int yrtmi_histo_setup(yrtmi_ctx* ctx, const yrtmi_histo_desc* h,
                      int64_t* num_z_bin, int64_t* num_phi, int64_t* num_r);
/* Size of subset s of S (BinIteratorRangeHistogram3D::size,
 * BinIterator.cpp:186-189). */
int64_t yrtmi_histo_subset_size(yrtmi_ctx* ctx, int num_subsets, int subset);

/* Rigid LOR motion, k host transforms of 12 floats (transform_t,
 * include/yrt-pet/utils/Types.hpp:57-60), applied to LOR endpoints as
 * ProjectionData::getLOR does (ProjectionData.cpp:302-318). */
int yrtmi_motion_upload(yrtmi_ctx* ctx, const float* xf, int k);

/* Projector: uploads PSF tables and fixes the multi-ray offsets (the
 * rand_r(seed=13) sequence of MultiRayGenerator::getRandomLine,
 * MultiRayGenerator.cpp:72-114).  Replaces ProjectorParams +
 * OperatorProjectorSiddon_GPU / OperatorProjectorDD_GPU construction. */
int yrtmi_proj_create(yrtmi_ctx* ctx, const yrtmi_projector_desc* d,
                      yrtmi_proj** out);
int yrtmi_proj_destroy(yrtmi_proj* p);
/* Launch shape: lanes cooperating on one event (1, 2, 4 or 8; they split the
 * rays), threads per block, and blocks (0 = sized from the event count). */
int yrtmi_proj_set_launch(yrtmi_proj* p, int lanes_per_event,
                          int block_size, int grid_blocks);
/* How backprojections accumulate.  YRTMI_ACCUM_XCD (default): each XCD adds
 * into its own private copy of the target image (chosen by the hardware XCC
 * id), and the copies are folded into the target by one streaming pass after
 * the kernel; the context keeps the copies (n_xcc x image size) between
 * calls.  YRTMI_ACCUM_DIRECT: FP32 atomics straight into the target (no
 * extra memory).  Both are the reference's atomicAdd accumulation
 * (SiddonKernels.cuh:426-428) up to summation order. */
#define YRTMI_ACCUM_DIRECT 0
#define YRTMI_ACCUM_XCD 1
int yrtmi_proj_set_accumulation(yrtmi_proj* p, int mode);

/* ---- projections --------------------------------------------------------
 * Forward: out[i] = (A x)_i for each LOR of `lors` (device out, n floats).
 * Replaces OperatorProjectorDevice::applyA (OperatorProjectorDevice.cu:
 * 442-544) -> projectSiddon_kernel / projectDD_kernel<true,...>. */
int yrtmi_forward(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                  const float* img, const yrtmi_lors* lors, float* out);
/* Back: img += A^T vals (device img is accumulated into, not cleared).
 * Replaces OperatorProjectorDevice::applyAH (OperatorProjectorDevice.cu:
 * 546-675) -> project*_kernel<false,...> with atomicAdd. */
int yrtmi_back(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
               float* img, const yrtmi_lors* lors, const float* vals);

/* Fused EM subset: u += A^T( m / (A x .* corr + add) .* corr ).
 * Replaces OSEM_GPU::computeEMUpdateImage (OSEM_GPU.cu:455-510) and
 * computeEMUpdateImage_kernel (OSEM_GPU.cu:1159-1341). */
int yrtmi_em_subset(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                    const float* x, float* u, const yrtmi_lors* lors,
                    const yrtmi_corrections* corr);
/* As yrtmi_em_subset, allocating nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops).  The per-XCD partial images (per-XCD accumulation,
 * yrtmi_proj_set_accumulation) live in the library's static scratch array
 * of 16 Mi floats (64 MiB) per device, held until the process ends: an image
 * of more than 2 Mi voxels x frames (8 XCDs; the partial images are rounded
 * to 64 floats each) is refused with YRTMI_EINVAL.  The scratch is cleared
 * by a fill kernel at every call (recordable).  Every call on one device
 * shares that device's array: calls that overlap on one device (two
 * streams, two host threads) race.  Direct accumulation uses no scratch. */
int yrtmi_em_subset_noalloc(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                            const float* x, float* u, const yrtmi_lors* lors,
                            const yrtmi_corrections* corr);

/* Sensitivity: img += A^T(acf .* sens) over `lors` (typically a whole
 * histogram subset), the projector used WITHOUT TOF.  att_img (nullable,
 * device, geometry att_im) gives acf = exp(-0.1 * Siddon(mu)) per LOR
 * (Corrector.cpp:601-611, Tools.hpp:98-102); sens (nullable, device, per
 * LOR) the normalisation factor.  Replaces OSEM_GPU::
 * generateSensitivityImageForCurrentSubset (OSEM_GPU.cu:197-253) and
 * generateSensImage_kernel (OSEM_GPU.cu:1048-1157). */
int yrtmi_sensitivity(yrtmi_ctx* ctx, const yrtmi_proj* p,
                      const yrtmi_image* im, float* img,
                      const yrtmi_lors* lors, const yrtmi_image* att_im,
                      const float* att_img, const float* sens);

/* Per-LOR attenuation factors exp(-0.1 * Siddon(mu)) (single ray, no TOF,
 * image offset not subtracted, as singleForwardProjection does).  Replaces
 * Corrector_GPU::precomputeAttenuationFactors (Corrector_GPU.cu:19-69). */
int yrtmi_attenuation_factors(yrtmi_ctx* ctx, const yrtmi_image* att_im,
                              const float* att_img, const yrtmi_lors* lors,
                              float* acf);

/* ---- image-space kernels (ImageSpaceKernels.cu:13-190) ------------------ */
/* x[t][j] *= u[t][j] / s[j] where s[j] > thr; s has nvox (3D, broadcast
 * over frames) or nt*nvox entries (s_frames = 1 or nt).  Replaces
 * updateEM_kernel / updateEMDynamic_kernel and Image::updateEMThreshold*
 * (Image.cpp:664-833). */
int yrtmi_em_update(yrtmi_ctx* ctx, float* x, const float* u, const float* s,
                    int64_t nvox, int64_t nt, int64_t s_frames, float thr);
/* x = mask <= thr ? x*le_scale + le_off : x*gt_scale + gt_off, mask
 * broadcast over frames when mask_frames == 1 (applyThreshold(Broadcast)
 * _kernel, Image.cpp:595-649). */
int yrtmi_threshold(yrtmi_ctx* ctx, float* x, const float* mask, int64_t nvox,
                    int64_t nt, int64_t mask_frames, float thr, float le_scale,
                    float le_off, float gt_scale, float gt_off);
int yrtmi_fill(yrtmi_ctx* ctx, float* x, int64_t n, float value);
/* x[i] = v for n int32 elements (device): frame indices of a batch of
 * frames.  No reference counterpart (its frames are host properties,
 * ProjectionData::getDynamicFrame).  Recordable. */
int yrtmi_fill_i32(yrtmi_ctx* ctx, int32_t* x, int64_t n, int32_t v);
/* Image-space PSF (OperatorPsf::applyA / applyAH,
 * src/operators/OperatorPsf.cpp:254-375; the device operator
 * OperatorPsfDevice): out = the separable circular convolution of `in`
 * (im->nt frames) with kx along x, then ky along y, then kz along z, each
 * output sum taken over the taps in order in FP32 (bitwise the reference
 * CPU loop); transpose != 0 convolves with the flipped kernels (the
 * adjoint, applyAH).  Kernels (host arrays) of odd length <= 63 whose half
 * length is below the image size on their axis; `tmp` is a device image
 * buffer of the same size; in, out and tmp distinct; fewer than 2^31 voxels
 * in all frames. */
int yrtmi_psf_apply(yrtmi_ctx* ctx, const yrtmi_image* im, const float* in, float* out,
                    float* tmp, const float* kx, int nkx, const float* ky, int nky,
                    const float* kz, int nkz, int transpose);
/* Time-averaged moved image: one launch of timeAverageMoveImage_kernel<true>
 * (src/datastruct/image/ImageSpaceKernels.cu:191-264) as
 * util::timeAverageMoveImageDevice drives it (src/utils/
 * ReconstructionUtilsDevice.cu:146-225).  For every voxel v of one frame,
 * out[v] += sum over f of weights[f] * trilinear(in, inverse(xf[f]) *
 * centre(v)), the frames in order.  in: frame 0 of an image of `im`'s
 * geometry; out: one frame (offset by the caller to its dynamic frame),
 * accumulated into; in and out distinct device buffers.  xf: n host
 * transforms (12 floats each, transform_t's r00 r01 r02 tx r10 ... tz), the
 * motion itself: the library inverts them as util::invertTransform;
 * weights: n host floats (duration / scan duration in the reference
 * driver).  Stages the frames into a context buffer after synchronising
 * the context stream; not under graph capture. */
int yrtmi_time_average_move(yrtmi_ctx* ctx, const yrtmi_image* im, const float* in,
                            float* out, const float* xf, const float* weights, int n);

/* FP64 Siddon projector (no counterpart in the reference, which computes in
 * FP32): the same LORs and rays with every ray quantity in FP64 (detector
 * positions are the FP32 LUT values).  For double-precision gradient checks
 * of the differentiable projector and as an accuracy reference.  Siddon
 * projectors only; DEVICE double images and values. */
int yrtmi_forward_f64(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                      const double* img, const yrtmi_lors* lors, double* out);
int yrtmi_back_f64(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                   double* img, const yrtmi_lors* lors, const double* vals);

/* ---- instrumentation ---------------------------------------------------- */
/* Total voxel segments (Siddon, summed over rays) or footprint voxels (DD)
 * touched by `lors`: the algorithmic work unit of the roofline figures.
 * *total is a HOST pointer; the call synchronises. */
int yrtmi_count_segments(yrtmi_ctx* ctx, const yrtmi_proj* p,
                         const yrtmi_image* im, const yrtmi_lors* lors,
                         int64_t* total);

/* ---- device memory and data movement ------------------------------------
 * Owned device buffers of the host layer (ImageDeviceOwned,
 * ProjectionListDeviceOwned: src/datastruct/image/ImageDevice.cu:153-230,
 * src/datastruct/projection/ProjectionListDevice.cu:148-210).  Copies are
 * synchronous on the context's stream.  kind: 0 host->device, 1 device->host,
 * 2 device->device. */
int yrtmi_device_alloc(yrtmi_ctx* ctx, int64_t bytes, void** out);
/* Free and total device memory of the context's device, bytes (HOST
 * pointers, nullable): globals::getAvailableVRAM's cudaMemGetInfo
 * (src/utils/GPUUtils.cu:77), from which OSEM_GPU sizes its batches
 * (getMemAvailable, src/recon/OSEM_GPU.cu:583-591). */
int yrtmi_mem_info(yrtmi_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
int yrtmi_device_free(yrtmi_ctx* ctx, void* p);
int yrtmi_copy(yrtmi_ctx* ctx, void* dst, const void* src, int64_t bytes,
               int kind);
/* The nonzero bins of histogram subset s of S (BinIteratorRangeHistogram3D
 * order, BinIterator.cpp:147-197) and their values, for the EM subset (the
 * reference GPU path skips zero-measurement bins, OSEM_GPU.cu:1222-1225).
 * values: DEVICE float[num_z_bin * num_phi * num_r]; bins_out / meas_out:
 * DEVICE arrays of at least yrtmi_histo_subset_size(S, s) entries; *count
 * (HOST) receives the number written.  Deterministic order; synchronises. */
int yrtmi_histo_subset_nonzero(yrtmi_ctx* ctx, const float* values, int S,
                               int s, uint32_t* bins_out, float* meas_out,
                               int64_t* count);
/* dst[i] = src[s + i * S] for the events of list-mode subset s of S
 * (BinIteratorChronologicalInterleaved, BinIterator.cpp:261-265), elements
 * of elem_bytes (1, 2, 4 or 8) bytes; n_total elements in src.  *count
 * (HOST, nullable) receives the subset size (n_total - s + S - 1) / S. */
int yrtmi_subset_interleaved(yrtmi_ctx* ctx, const void* src, int elem_bytes,
                             int64_t n_total, int S, int s, void* dst,
                             int64_t* count);
/* out[i] = src[idx[i]] for i < n (per-bin corrections of a bin list);
 * indices >= src_n give 0.  invert != 0: nonzero values become their
 * reciprocal (Corrector::getSensitivity with invertSensitivity,
 * Corrector.cpp:295-308). */
/* out[i] = histo[bin of the detector pair (d1[i], d2[i])] for list-mode
 * events: Histogram3D::getProjectionValueFromHistogramBin with a detector
 * pair (src/datastruct/projection/Histogram3D.cpp:588-606; the pair -> bin
 * map getCoordsFromDetPair / getBinIdFromCoords, 267-271, 366-405), as
 * Corrector::getTotalACFFromHistogram / getInVivoAttenuationFactor read ACF
 * histograms for list-mode data.  histo: the context's Histogram3D (device);
 * the pair (0, 0) reads 0; a pair outside the histogram fails the call (the
 * reference throws).  Synchronises the context stream; not under graph
 * capture. */
int yrtmi_histo_gather_pairs(yrtmi_ctx* ctx, const float* histo, const uint32_t* d1,
                             const uint32_t* d2, int64_t n, float* out);
/* bins[i] = the Histogram3D bin of the detector pair (d1[i], d2[i])
 * (Histogram3D::getBinIdFromDetPair, src/datastruct/projection/
 * Histogram3D.cpp:416-421, with getCoordsFromDetPair 366-405 and the ring
 * map of setupHistogram 286-297), on the device: e.g. to histogram
 * list-mode events.  A pair outside the histogram, or (0, 0), gets
 * 0xffffffff and fails the call (the reference throws).  The histogram must
 * have fewer than 2^32 - 1 bins.  Synchronises the context stream; not under
 * graph capture. */
int yrtmi_histo_pair_bins(yrtmi_ctx* ctx, const uint32_t* d1, const uint32_t* d2, int64_t n,
                          uint32_t* bins);
int yrtmi_gather(yrtmi_ctx* ctx, const float* src, int64_t src_n,
                 const uint32_t* idx, int64_t n, int invert, float* out);
/* out[i] = histo[bin of element i of histogram subset s of S], for every
 * element of the subset (per-bin factors in YRTMI_SRC_HISTO_SUBSET order);
 * invert as for yrtmi_gather. */
int yrtmi_histo_subset_gather(yrtmi_ctx* ctx, const float* histo, int S, int s,
                              int invert, float* out);
/* y = a * x + b * y over n floats (x NULL: y = b * y).  Image sums and
 * scalings of the OSEM driver (e.g. the mask as the sum of the subsets'
 * sensitivity images, OSEM_CPU.cpp:255-269; list-mode sensitivity / S,
 * OSEM.cpp:1007-1035). */
int yrtmi_axpby(yrtmi_ctx* ctx, int64_t n, float a, const float* x, float b,
                float* y);

/* ---- locality order ------------------------------------------------------
 * No counterpart in the reference.  Its EM update does not depend on the
 * order in which a subset's LORs are projected (OSEM_CPU.cpp:329-374 sums
 * their backprojections; the order changes FP32 rounding only), so a caller
 * that owns a subset's arrays (OSEM over a list-mode subset) may reorder them.
 * yrtmi_lor_order writes perm (DEVICE uint32[n]) = 0..n-1 sorted stably by a
 * key of each LOR's line: dynamic frame, axial midpoint (as a voxel slice of
 * im), axial tilt, transaxial angle, transaxial offset.  LORs close in that
 * order cross nearly the same voxels, so the waves that project them
 * concurrently read a slab of the image that stays in L2.  Skipped LORs
 * (frame out of range, bad motion index) go last.  n < 2^32, nz <= 1024.
 * Synchronises; allocates temporaries (about 24 bytes per LOR).
 * yrtmi_permute: dst[i] = src[perm[i]] for i < n, elements of elem_bytes
 * (1, 2, 4 or 8) bytes, DEVICE arrays, dst distinct from src. */
int yrtmi_lor_order(yrtmi_ctx* ctx, const yrtmi_image* im,
                    const yrtmi_lors* lors, uint32_t* perm);
int yrtmi_permute(yrtmi_ctx* ctx, const void* src, int elem_bytes,
                  const uint32_t* perm, int64_t n, void* dst);

/* ---- tile plans (backprojection through LDS image tiles) ---------------- */
/* No counterpart in the reference: a performance extension of the
 * backprojection behind yrtmi_back / yrtmi_em_subset / yrtmi_sensitivity
 * (OperatorProjectorSiddon_GPU::applyAH, OperatorProjectorDD_GPU::applyAH,
 * OSEM_GPU::computeEMUpdateImage,
 * OSEM_GPU::generateSensitivityImageForCurrentSubset).
 *
 * Siddon: a plan bins the walk of every (LOR, ray) of `lors` by the image
 * tiles it crosses, recording where it enters each (12 bytes per record,
 * records of a tile ordered by the length of their piece), and keeps each
 * ray's end points, TOF value and frame (32 bytes per LOR and ray), plus,
 * with the FP32 set-up, the ray's walk set-up (inverse direction, exit
 * parameter, TOF terms: 64 bytes per LOR and ray in all), so that the
 * backprojection reads one table entry per record instead of resolving the
 * LOR and redoing its set-up.
 * DD: a plan keeps each LOR's footprint set-up (128 bytes per LOR) and one
 * record per tile the footprint reaches (8 bytes: the LOR and the slices of
 * the tile's slab that reach it, records of a tile ordered by their voxel
 * count); the backprojection sweeps each record's slices over the tile with
 * the unplanned sweep's arithmetic.
 * A plan depends on geometry only (detector pairs or bins, TOF values,
 * frames, motion frames, the projector, the image geometry), not on image
 * or LOR values: a list-mode or histogram subset is planned once per
 * reconstruction.  The *_planned calls then accumulate each tile in LDS and
 * add it to the image with coalesced atomics; results equal the unplanned
 * calls' up to FP32 summation order (and, for Siddon on voxel sizes that
 * are not short binary fractions, ulp-level shifts of segment ends at tile
 * boundaries).
 *
 * Limits: nx, ny, nz <= 1024; n * num_rays < 2^32; tiles (frames included)
 * at most the device's LDS per workgroup / 12 bytes (5459 with 64 KB).  The
 * call synchronises the context's stream and allocates device memory
 * (yrtmi_plan_info's bytes); the plan owns it.
 *
 * Options (NULL: defaults): chunk_records, records per workgroup (0: 4096);
 * tile_shape, YRTMI_TILE_16x16x16 (default), YRTMI_TILE_32x32x16,
 * YRTMI_TILE_32x32x32 or YRTMI_TILE_32x32x8 (voxels of the image held per
 * workgroup; the tile is accumulated in LDS in 64-bit fixed point, FP32 for
 * the 32x32x32 shape); persistent,
 * 0 for one workgroup per chunk, else that many workgroups taking chunks from
 * a device queue (changeable with yrtmi_plan_set_persistent); max_bytes, 0
 * for no limit, else the device memory the plan may hold: a plan that would
 * hold more (known from the ray count before the ray table is allocated and
 * from the count pass before the records are) is not built and the call
 * fails with YRTMI_ENOMEM, having allocated nothing beyond its build
 * scratch (the reference sizes its batches to the memory available,
 * OperatorProjectorDevice::create's memAvailable, OperatorProjectorDevice.
 * cu:36-46, OSEM_GPU.cu:413-435); single_use, 0 to order each chunk's
 * records by their length (one more pass over the records at build time,
 * about 1.7 ms per 25 M rays on MI355X, for faster backprojections: the
 * plan of an EM subset, backprojected every iteration), 1 to skip that
 * pass (a plan backprojected once, as a sensitivity image's). */
#define YRTMI_TILE_16x16x16 0
#define YRTMI_TILE_32x32x16 1
#define YRTMI_TILE_32x32x32 2
#define YRTMI_TILE_32x32x8 3
typedef struct
{
	int chunk_records;
	int tile_shape;
	int persistent;
	int64_t max_bytes;
	int single_use;
} yrtmi_plan_options;
typedef struct yrtmi_plan yrtmi_plan;
int yrtmi_plan_create(yrtmi_ctx* ctx, const yrtmi_proj* p,
                      const yrtmi_image* im, const yrtmi_lors* lors,
                      const yrtmi_plan_options* options, yrtmi_plan** out);
int yrtmi_plan_destroy(yrtmi_plan* plan);
int yrtmi_plan_set_persistent(yrtmi_plan* plan, int workgroups);
/* records: (LOR, ray, tile) entries; chunks: workgroups per use; bytes:
 * device memory held.  Host pointers, any may be NULL. */
int yrtmi_plan_info(const yrtmi_plan* plan, int64_t* records, int64_t* chunks,
                    int64_t* bytes);
/* Launch-shape autotuner (no reference counterpart).  On `lors` and image x
 * it times the EM ratio kernel for lanes per event 1, 2, 4 and 8 (Siddon;
 * a distance-driven LOR is one ray: lanes per event 1, the other entries of
 * ratio_ms_lpe -1) and the planned backprojection for every tile shape,
 * each with one workgroup per chunk and as a persistent launch (one plan per
 * shape is built and freed), then sets the projector's lanes per event to
 * the fastest and returns the fastest settings; plans made afterwards with
 * best.tile_shape / best.persistent use them.  Synchronises and allocates
 * temporaries. */
typedef struct
{
	int lanes_per_event;
	int tile_shape;
	int persistent;
	float ratio_ms;          /* EM ratio kernel, best lanes per event */
	float back_ms;           /* planned backprojection, best settings */
	float ratio_ms_lpe[4];   /* lanes per event 1, 2, 4, 8 */
	float back_ms_shape[4][2]; /* [tile shape][one per chunk, persistent];
	                            * -1: the shape does not fit the image */
} yrtmi_tuning;
int yrtmi_autotune(yrtmi_ctx* ctx, yrtmi_proj* p, const yrtmi_image* im,
                   const yrtmi_lors* lors, const float* x, yrtmi_tuning* best);

/* As yrtmi_back / yrtmi_em_subset / yrtmi_sensitivity through `plan`, which
 * must come from the same projector, image geometry and LOR set (checked:
 * projector, geometry and LOR descriptor fields; for the sensitivity call
 * the plan's projector must have TOF off).  Calls on one plan must not
 * overlap (em_subset keeps its per-LOR ratios in the plan). */
int yrtmi_back_planned(yrtmi_ctx* ctx, const yrtmi_proj* p,
                       const yrtmi_image* im, float* img,
                       const yrtmi_lors* lors, const float* vals,
                       const yrtmi_plan* plan);
int yrtmi_em_subset_planned(yrtmi_ctx* ctx, const yrtmi_proj* p,
                            const yrtmi_image* im, const float* x, float* u,
                            const yrtmi_lors* lors,
                            const yrtmi_corrections* corr,
                            const yrtmi_plan* plan);
/* The first half of yrtmi_em_subset on its own (Siddon or DD): v[i] = the
 * EM ratio of LOR i (m / corrected forward projection, times the
 * multiplicative corrections), 0 where the reference skips the LOR.
 * yrtmi_em_subset_planned = yrtmi_em_ratio_planned + yrtmi_back_planned(v).
 * Siddon projects a bricked copy of x made at each call into a buffer the
 * context keeps (4 x 4 x 2-voxel bricks; same values, same sums): it grows
 * on the first call of a larger image, or when a plan is created, which a
 * caller recording a graph must do before recording.
 * Replaces the forward half of OSEM_CPU::computeEMUpdateImage
 * (src/recon/OSEM_CPU.cpp:329-374). */
int yrtmi_em_ratio(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                   const float* x, const yrtmi_lors* lors,
                   const yrtmi_corrections* corr, float* v);
/* yrtmi_em_ratio with the plan of (p, im, lors) (same checks as the
 * *_planned calls): a DD projector reads each LOR's footprint set-up from
 * the plan instead of computing it; Siddon resolves its LORs as
 * yrtmi_em_ratio does.  Results identical to yrtmi_em_ratio. */
int yrtmi_em_ratio_planned(yrtmi_ctx* ctx, const yrtmi_proj* p, const yrtmi_image* im,
                           const float* x, const yrtmi_lors* lors,
                           const yrtmi_corrections* corr, float* v,
                           const yrtmi_plan* plan);
int yrtmi_sensitivity_planned(yrtmi_ctx* ctx, const yrtmi_proj* p,
                              const yrtmi_image* im, float* img,
                              const yrtmi_lors* lors,
                              const yrtmi_image* att_im, const float* att_img,
                              const float* sens, const yrtmi_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* YRTMI_H */

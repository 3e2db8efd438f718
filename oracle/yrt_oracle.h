/*
 * CPU oracle for the OSEM projection path: a plain C++ restatement of the
 * reference CPU algorithms (YRT-PET 2.0.7), used ONLY as the test checker and
 * as the CPU baseline leg of the benchmark.  It is never linked into, called
 * by, or shipped with the product library.
 *
 * Every function carries an FP32 instantiation that follows the reference's
 * operation order and an FP64 instantiation of the same algorithm (the "FP64
 * oracle" against which both the FP32 restatement and the GPU path are
 * measured).
 *
 * Parity pins (what each function is checked against) are listed in
 * tests/test_oracle_*.py; see the per-function comments in yrt_oracle.cpp.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Image geometry, reference ImageParams (include/yrt-pet/datastruct/image/
 * ImageParams.hpp): n = voxel count, v = voxel size (mm), l = length (mm),
 * o = offset of the image centre (mm), fov_radius as set by
 * ImageParams::setup (src/datastruct/image/ImageParams.cpp:206-218). */
typedef struct
{
	int64_t nx, ny, nz, nt;
	float vx, vy, vz;
	float lx, ly, lz;
	float ox, oy, oz;
	float fov_radius;
} yo_image;

/* Regular cylindrical scanner, reference Scanner ctor
 * (src/datastruct/scanner/Scanner.cpp:101-127). */
typedef struct
{
	float axial_fov, crystal_size_z, crystal_size_trans, crystal_depth, radius;
	int64_t dets_per_ring, num_rings, num_doi, max_ring_diff, min_ang_diff,
	    dets_per_block;
} yo_scanner;

/* Projector configuration, reference ProjectorParams
 * (include/yrt-pet/operators/ProjectorParams.hpp:27-71) plus the optional
 * projection-space PSF (ProjectionPsfManager, one CSV row). */
typedef struct
{
	int type; /* 0 = Siddon, 1 = distance-driven */
	int num_rays;
	int use_tof;
	float tof_width_ps;
	int tof_n_std;
	/* multi-ray offsets: (num_rays-1)*4 floats in [-0.5, 0.5), see
	 * yo_multiray_offsets */
	const float* ray_offsets;
	/* projection PSF (DD only): psf_num_s rows of psf_ksize taps */
	const float* psf_kernels;
	int psf_num_s, psf_ksize;
	float psf_s_step, psf_k_spacing;
	float crystal_size_z, crystal_size_trans, crystal_depth;
} yo_projector;

/* A set of LORs (already resolved to endpoints).  n1/n2 (detector
 * orientations) are only needed for multi-ray Siddon and DD. tof/frame are
 * nullable. */
typedef struct
{
	int64_t n;
	const float* p1; /* n x 3 */
	const float* p2; /* n x 3 */
	const float* n1; /* n x 3, nullable */
	const float* n2; /* n x 3, nullable */
	const float* tof; /* n, ps, nullable */
	const int32_t* frame; /* n, dynamic frame, nullable (=0); <0 skips */
} yo_lors;

/* ---- scanner / data model ---- */
int yo_lut_regular(const yo_scanner* sc, float* pos, float* orient);
int yo_histo3d_dims(const yo_scanner* sc, int64_t* num_r, int64_t* num_phi,
                    int64_t* num_z_bin);
int yo_histo3d_pairs(const yo_scanner* sc, const int64_t* bins, int64_t n,
                     uint32_t* d1, uint32_t* d2);
int yo_histo3d_bins_from_pairs(const yo_scanner* sc, const uint32_t* d1,
                               const uint32_t* d2, int64_t n, int64_t* bins);
int64_t yo_histo3d_subset_size(const yo_scanner* sc, int num_subsets,
                               int subset);
int yo_histo3d_subset_bins(const yo_scanner* sc, int num_subsets, int subset,
                           int64_t* bins);
int yo_multiray_offsets(int num_rays, float* out);
int yo_multiray_lines(const float* p1, const float* p2, const float* n1,
                      const float* n2, float tz, float tt, int num_lines,
                      float* out);

/* ---- TOF / PSF / DD helpers (exposed for pinning) ---- */
int yo_tof_helper(float width_ps, int n_std, float* sigma, float* trunc,
                  float* norm);
int yo_tof_alpha_range(float width_ps, int n_std, float lor_norm, float tof_ps,
                       float* amin, float* amax);
float yo_tof_weight(float width_ps, int n_std, float lor_norm, float tof_ps,
                    float lo_mm, float hi_mm);
float yo_psf_integral(const float* kernel, int ksize, float kstep, float x0,
                      float x1);
float yo_dd_overlap(float p0, float p1, float d0, float d1);
int yo_get_alpha(float r0, float r1, float p1, float p2, float inv_p12,
                 float* amin, float* amax);

/* ---- projectors over LOR sets ---- */
int yo_fwd_f32(const yo_image* im, const float* img, const yo_projector* pr,
               const yo_lors* lors, float* out, int nthreads);
int yo_back_f32(const yo_image* im, float* img, const yo_projector* pr,
                const yo_lors* lors, const float* vals, int nthreads);
int yo_fwd_f64(const yo_image* im, const double* img, const yo_projector* pr,
               const yo_lors* lors, double* out, int nthreads);
int yo_back_f64(const yo_image* im, double* img, const yo_projector* pr,
                const yo_lors* lors, const double* vals, int nthreads);
/* Count of voxel segments touched per LOR (sum over rays) - the algorithmic
 * work measure. */
int yo_count_segments(const yo_image* im, const yo_projector* pr,
                      const yo_lors* lors, int64_t* counts, int nthreads);

/* ---- OSEM pieces ---- */
/* EM update image for one subset (OSEM_CPU::computeEMUpdateImage,
 * src/recon/OSEM_CPU.cpp:284-393).  meas/sens/att/rand/scat are per-LOR,
 * nullable (meas == NULL means list-mode measurement 1.0); pre (nullable)
 * the in-vivo attenuation precorrection multiplying the denominator after
 * the additive terms (OSEM_GPU.cu:1303-1308). */
int yo_em_subset_f32(const yo_image* im, const float* x, float* u,
                     const yo_projector* pr, const yo_lors* lors,
                     const float* meas, const float* sens, const float* att,
                     const float* rand, const float* scat, const float* pre,
                     float global_scale, float denom_thr, int nthreads);
int yo_em_subset_f64(const yo_image* im, const double* x, double* u,
                     const yo_projector* pr, const yo_lors* lors,
                     const double* meas, const double* sens, const double* att,
                     const double* rand, const double* scat, const double* pre,
                     double global_scale, double denom_thr, int nthreads);
/* Image::updateEMThresholdStatic (src/datastruct/image/Image.cpp:808-833). */
int yo_em_update_f32(float* x, const float* u, const float* s, int64_t n,
                     float thr);
int yo_em_update_f64(double* x, const double* u, const double* s, int64_t n,
                     double thr);
/* Image::applyThreshold (src/datastruct/image/Image.cpp:595-618). */
int yo_apply_threshold_f32(float* x, const float* mask, int64_t n, float thr,
                           float le_scale, float le_off, float gt_scale,
                           float gt_off);

#ifdef __cplusplus
}
#endif

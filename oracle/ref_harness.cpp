/*
 * TEST INFRASTRUCTURE ONLY.  C-ABI harness around the parts of the reference
 * (YRT-PET 2.0.7) that compile on their own with g++: the header-only
 * TimeOfFlightHelper and util::get_alpha / util::calculateIntegral, plus the
 * MultiRayGenerator and ProjectionPsfManager translation units compiled where
 * they lie under the reference tree (see oracle/Makefile, target `ref`).
 *
 * It is built into oracle/_ref/libyrtref.so and used only by tests to pin the
 * oracle restatement.  The projector translation units (ProjectorSiddon.cpp,
 * ProjectorDD.cpp) need nlohmann/json, which this image lacks, so they are not
 * part of this harness.
 */
#include "yrt-pet/geometry/Line3D.hpp"
#include "yrt-pet/geometry/MultiRayGenerator.hpp"
#include "yrt-pet/operators/ProjectionPsfManager.hpp"
#include "yrt-pet/operators/ProjectorUtils.hpp"
#include "yrt-pet/operators/TimeOfFlight.hpp"

#include <cstdint>
#include <memory>
#include <string>

extern "C" {

int yr_tof_helper(float width_ps, int n_std, float* sigma, float* trunc,
                  float* norm)
{
	yrt::TimeOfFlightHelper t(width_ps, n_std);
	*sigma = t.getSigma();
	*trunc = t.getTruncWidth();
	*norm = t.getNorm();
	return 0;
}

int yr_tof_alpha_range(float width_ps, int n_std, float lor_norm, float tof_ps,
                       float* amin, float* amax)
{
	yrt::TimeOfFlightHelper t(width_ps, n_std);
	t.getAlphaRange(*amin, *amax, lor_norm, tof_ps);
	return 0;
}

float yr_tof_weight(float width_ps, int n_std, float lor_norm, float tof_ps,
                    float lo_mm, float hi_mm)
{
	yrt::TimeOfFlightHelper t(width_ps, n_std);
	return t.getWeight(lor_norm, tof_ps, lo_mm, hi_mm);
}

int yr_get_alpha(float r0, float r1, float p1, float p2, float inv_p12,
                 float* amin, float* amax)
{
	yrt::util::get_alpha(r0, r1, p1, p2, inv_p12, *amin, *amax);
	return 0;
}

float yr_calc_integral(const float* kernel, int ksize, float kstep, float x0,
                       float x1)
{
	return yrt::util::calculateIntegral(kernel, ksize, kstep, x0, x1);
}

/* Generate `num_lines` random lines with the reference generator, seed 13,
 * default flags (not parallel, no depth).  out: num_lines x 6 floats. */
int yr_multiray_lines(const float* p1, const float* p2, const float* n1,
                      const float* n2, float thick_z, float thick_trans,
                      float depth, int num_lines, float* out)
{
	yrt::Line3D lor{yrt::Vector3D{p1[0], p1[1], p1[2]},
	                yrt::Vector3D{p2[0], p2[1], p2[2]}};
	yrt::Vector3D v1{n1[0], n1[1], n1[2]};
	yrt::Vector3D v2{n2[0], n2[1], n2[2]};
	yrt::MultiRayGenerator gen(thick_z, thick_trans, depth);
	gen.setupGenerator(lor, v1, v2);
	unsigned int seed = 13;
	for (int i = 0; i < num_lines; i++)
	{
		const yrt::Line3D l = gen.getRandomLine(seed);
		out[6 * i + 0] = l.point1.x;
		out[6 * i + 1] = l.point1.y;
		out[6 * i + 2] = l.point1.z;
		out[6 * i + 3] = l.point2.x;
		out[6 * i + 4] = l.point2.y;
		out[6 * i + 5] = l.point2.z;
	}
	return 0;
}

/* Projection-PSF manager from a one-row CSV file: kernel row index chosen for
 * an LOR, half width, and the overlap weight for [x0, x1]. */
int yr_psf_query(const char* csv, const float* p1, const float* p2, int flip,
                 float x0, float x1, float* half_width, float* weight,
                 int* row_offset)
{
	yrt::ProjectionPsfManager m{std::string(csv)};
	yrt::Line3D lor{yrt::Vector3D{p1[0], p1[1], p1[2]},
	                yrt::Vector3D{p2[0], p2[1], p2[2]}};
	const float* k = m.getKernel(lor, flip != 0);
	const float* k0 = m.getKernel(
	    yrt::Line3D{yrt::Vector3D{0.f, 0.f, 0.f}, yrt::Vector3D{0.f, 0.f, 0.f}},
	    flip != 0);
	*half_width = m.getHalfWidth_mm();
	*weight = m.getWeight(k, x0, x1);
	*row_offset = static_cast<int>((k - k0) / m.getKernelSize());
	return 0;
}

}  // extern "C"

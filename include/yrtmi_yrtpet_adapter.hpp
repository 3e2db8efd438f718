/*
 * yrtmi behind YRT-PET's projector operator: the class the reference side
 * compiles to put the MI355X path behind its `--gpu` forward / back
 * projection (OperatorProjectorDevice::create,
 * src/operators/OperatorProjectorDevice.cu:300-319, returns it instead of
 * OperatorProjectorSiddon_GPU / OperatorProjectorDD_GPU).
 *
 * It is an OperatorProjectorBase (include/yrt-pet/operators/
 * OperatorProjectorBase.hpp:21-44) with the contract of OperatorProjector::
 * applyA / applyAH (src/operators/OperatorProjector.cpp:258-316):
 *   applyA(Image in, ProjectionData out): out[bin] = forward projection of
 *     the image along bin's LOR, for every bin of the bin iterator;
 *   applyAH(ProjectionData in, Image out): out += backprojection of in[bin]
 *     along bin's LOR, bins of value 0 skipped;
 * the same ASSERT messages on bad arguments (as std::runtime_error, which
 * the reference's ASSERT_MSG throws and pybind11 turns into RuntimeError).
 * Every bin carries its detector pair, TOF value, dynamic frame and motion
 * frame (ProjectionData::getDetector1/2, getTOFValue, getDynamicFrame,
 * getMotionFrame); the motion transforms are those of the projection data
 * (getTransformOfMotionFrame), a motion frame < 0 meaning no motion
 * (LORMotion::getTransform, LORMotion.cpp:60-70).
 *
 * ImageT is the reference's yrt::Image (include/yrt-pet/datastruct/image/
 * Image.hpp: getParams() with nx, ny, nz, nt, length_*, off_*, and
 * getRawPointer()).  It is a template parameter because Image.hpp needs
 * nlohmann/json, which the reference's own build provides; the rest of the
 * class needs only the headers included below.
 *
 * Host data (yrt::Image, yrt::ProjectionData) is uploaded per call; callers
 * that keep data on the device use yrtmi's ImageDevice / ListModeDevice
 * aliases and OperatorProjectorMI directly (yrtmi_host.hpp).
 */
#ifndef YRTMI_YRTPET_ADAPTER_HPP
#define YRTMI_YRTPET_ADAPTER_HPP

#include "yrt-pet/datastruct/projection/BinIterator.hpp"
#include "yrt-pet/datastruct/projection/ProjectionData.hpp"
#include "yrt-pet/datastruct/scanner/Scanner.hpp"
#include "yrt-pet/operators/OperatorProjectorBase.hpp"
#include "yrt-pet/operators/ProjectorParams.hpp"

#include "yrtmi_host.hpp"

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace yrtmi_yrtpet
{

// yrt::Scanner (Scanner.hpp:27-31 constructor fields) as yrtmi's scanner
// (a regular cylindrical LUT, DetRegular::generateLUT).
inline yrtmi::Scanner toScanner(const yrt::Scanner& s)
{
	return yrtmi::Scanner(s.scannerName, s.axialFOV, s.crystalSize_z, s.crystalSize_trans,
	                      s.crystalDepth, s.scannerRadius, s.detsPerRing, s.numRings,
	                      s.numDOI, s.maxRingDiff, s.minAngDiff, s.detsPerBlock);
}

// The yrtmi projector of a yrt::ProjectorParams (ProjectorParams.hpp:27-71):
// type, rays, TOF and the projection-space PSF file.
inline void toProjectorParams(const yrt::ProjectorParams& pp, yrtmi::ProjectorParams& out)
{
	out.setProjector(pp.projectorType == yrt::ProjectorType::DD ? yrtmi::ProjectorType::DD
	                                                            : yrtmi::ProjectorType::SIDDON);
	out.numRays = pp.numRays;
	if (pp.hasTOF())
		out.addTOF(pp.getTOFWidth_ps(), pp.getTOFNumStd());
	else
		out.removeTOF();
	if (!pp.projPsf_fname.empty())
		out.setProjPsf(pp.projPsf_fname);
}

// The LORs of a bin iterator's bins as yrtmi list-mode arrays.
struct BinEvents
{
	std::vector<yrt::bin_t> bins;
	std::vector<uint32_t> d1, d2;
	std::vector<float> tof;
	std::vector<int32_t> frame;
	std::vector<uint8_t> motion;
	std::vector<float> transforms;  // k x 12 (transform_t order)
	int numFrames = 1;               // dynamic frames (the image's nt)
};

inline BinEvents gatherBins(const yrt::ProjectionData& dat, const yrt::BinIterator& it)
{
	BinEvents e;
	const size_t n = it.size();
	e.bins.resize(n);
	e.d1.resize(n);
	e.d2.resize(n);
	const bool tof = dat.hasTOF();
	const bool dyn = dat.hasDynamicFraming();
	const bool mot = dat.hasMotion();
	if (tof)
		e.tof.resize(n);
	if (dyn)
	{
		e.frame.resize(n);
		e.numFrames = (int)dat.getNumDynamicFrames();
	}
	size_t identity = 0;
	if (mot)
	{
		const size_t k = dat.getNumMotionFrames();
		// motion frame indices are one byte on the device; index k is the
		// identity for events before the first motion frame
		if (k >= 255)
			throw std::runtime_error("yrtmi: at most 254 motion frames");
		e.motion.resize(n);
		e.transforms.resize(12 * (k + 1));
		for (size_t f = 0; f < k; f++)
		{
			const yrt::transform_t t = dat.getTransformOfMotionFrame((yrt::frame_t)f);
			const float v[12] = {t.r00, t.r01, t.r02, t.tx, t.r10, t.r11,
			                     t.r12, t.ty, t.r20, t.r21, t.r22, t.tz};
			for (int j = 0; j < 12; j++)
				e.transforms[12 * f + j] = v[j];
		}
		const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
		for (int j = 0; j < 12; j++)
			e.transforms[12 * k + j] = id[j];
		identity = k;
	}
	for (size_t i = 0; i < n; i++)
	{
		const yrt::bin_t b = it.get(i);
		e.bins[i] = b;
		e.d1[i] = (uint32_t)dat.getDetector1(b);
		e.d2[i] = (uint32_t)dat.getDetector2(b);
		if (tof)
			e.tof[i] = dat.getTOFValue(b);
		if (dyn)
			e.frame[i] = (int32_t)dat.getDynamicFrame(b);
		if (mot)
		{
			const yrt::frame_t f = dat.getMotionFrame(b);
			e.motion[i] = (uint8_t)(f < 0 ? identity : (size_t)f);
		}
	}
	return e;
}

template <class ImageT>
class OperatorProjectorMIAdapter : public yrt::OperatorProjectorBase
{
public:
	OperatorProjectorMIAdapter(const yrt::ProjectorParams& pp, const yrt::BinIterator* it,
	                           int device = 0)
	    : yrt::OperatorProjectorBase(pp, it),
	      m_scanner(toScanner(pp.scanner)),
	      m_ctx(m_scanner, device),
	      m_params(m_scanner)
	{
		toProjectorParams(pp, m_params);
		m_op.reset(new yrtmi::OperatorProjectorMI(m_ctx, m_params));
	}

	void applyA(const yrt::Variable* in, yrt::Variable* out) override
	{
		auto* dat = dynamic_cast<yrt::ProjectionData*>(out);
		auto* img = dynamic_cast<const ImageT*>(in);
		if (dat == nullptr)
			throw std::runtime_error("Output variable has to be Projection data");
		if (img == nullptr)
			throw std::runtime_error("Input variable has to be an Image");
		if (img->getRawPointer() == nullptr)
			throw std::runtime_error("Image array is unallocated");
		if (binIter == nullptr)
			throw std::runtime_error("BinIterator undefined");
		BinEvents e = gatherBins(*dat, *binIter);
		needFrames(e, *img);
		if (e.bins.empty())
			return;
		auto lm = events(e);
		yrtmi::ImageDevice x(m_ctx, imageParams(*img));
		x.copyFromHost(img->getRawPointer());
		m_op->applyA(x, *lm);
		std::vector<float> v(e.bins.size());
		lm->valuesToHost(v.data());
		for (size_t i = 0; i < v.size(); i++)
			dat->setProjectionValue(e.bins[i], v[i]);
	}

	void applyAH(const yrt::Variable* in, yrt::Variable* out) override
	{
		auto* dat = dynamic_cast<const yrt::ProjectionData*>(in);
		auto* img = dynamic_cast<ImageT*>(out);
		if (dat == nullptr)
			throw std::runtime_error("Input variable has to be Projection data");
		if (img == nullptr)
			throw std::runtime_error("Output variable has to be an Image");
		if (img->getRawPointer() == nullptr)
			throw std::runtime_error("Image array is unallocated");
		if (binIter == nullptr)
			throw std::runtime_error("BinIterator undefined");
		BinEvents e = gatherBins(*dat, *binIter);
		needFrames(e, *img);
		if (e.bins.empty())
			return;
		auto lm = events(e);
		std::vector<float> v(e.bins.size());
		for (size_t i = 0; i < v.size(); i++)
			v[i] = dat->getProjectionValue(e.bins[i]);  // zeros add nothing
		lm->valuesFromHost(v.data());
		yrtmi::ImageDevice x(m_ctx, imageParams(*img));
		x.copyFromHost(img->getRawPointer());  // backprojection adds to the image
		m_op->applyAH(*lm, x);
		x.copyToHost(img->getRawPointer());
	}

	yrtmi::OperatorProjectorMI& device() { return *m_op; }

private:
	// util::doesDynamicFramingMatch (OperatorProjector.cpp:267-273)
	static void needFrames(const BinEvents& e, const ImageT& img)
	{
		if ((int64_t)img.getParams().nt != (int64_t)e.numFrames)
			throw std::runtime_error(
			    "The number of dynamic frames in the the projection data does "
			    "not match the image's fourth dimension");
	}

	static yrtmi::ImageParams imageParams(const ImageT& img)
	{
		const auto& p = img.getParams();
		return yrtmi::ImageParams(p.nx, p.ny, p.nz, p.length_x, p.length_y, p.length_z,
		                          p.off_x, p.off_y, p.off_z, p.nt);
	}

	std::unique_ptr<yrtmi::ListModeDevice> events(const BinEvents& e) const
	{
		std::unique_ptr<yrtmi::ListModeDevice> lm(new yrtmi::ListModeDevice(
		    m_ctx, e.bins.size(), e.d1.data(), e.d2.data(), e.tof.empty() ? nullptr : e.tof.data(),
		    e.frame.empty() ? nullptr : e.frame.data(),
		    e.motion.empty() ? nullptr : e.motion.data()));
		lm->numFrames = e.numFrames;
		if (!e.transforms.empty())
		{
			lm->motionTransforms = e.transforms;
			m_ctx.uploadMotion(e.transforms);
		}
		return lm;
	}

	yrtmi::Scanner m_scanner;
	yrtmi::Context m_ctx;
	yrtmi::ProjectorParams m_params;
	std::unique_ptr<yrtmi::OperatorProjectorMI> m_op;
};

}  // namespace yrtmi_yrtpet

#endif

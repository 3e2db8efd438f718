// C++ host layer over the yrtmi C ABI, with the reference's class names and
// semantics for the projection / OSEM path:
//   Scanner, ImageParams, ProjectorParams          (data model)
//   Context                                        (device, stream, resident
//                                                   detector / histogram tables)
//   ImageDevice (owned / alias), ListModeDevice,   (device-resident data)
//   Histogram3DDevice
//   OperatorProjectorMI                            (applyA / applyAH)
//   OSEM_MI, createOSEM                            (the OSEM plugin)
//   forwProject / backProject                      (util functions)
// Every device buffer is either owned (allocated through the ABI) or an
// alias of a caller's device pointer (never freed here), as the reference's
// Owned / Alias classes (docs/source/python/owned_vs_alias.md).  Failures
// throw yrtmi::Error (a std::runtime_error, like the reference's ASSERTs).
#pragma once

#include "yrtmi.h"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace yrtmi
{

class Error : public std::runtime_error
{
public:
	using std::runtime_error::runtime_error;
};

// Throws Error(yrtmi_last_error()) when rc != YRTMI_OK.
void check(int rc, const char* what = nullptr);

// Regular cylindrical scanner (Scanner ctor, Scanner.cpp:101-127).
class Scanner
{
public:
	Scanner(std::string scannerName, float axialFOV, float crystalSize_z,
	        float crystalSize_trans, float crystalDepth, float scannerRadius,
	        size_t detsPerRing, size_t numRings, size_t numDOI,
	        size_t maxRingDiff, size_t minAngDiff, size_t detsPerBlock);
	size_t getNumDets() const;
	// Detector LUT, n x 3 each: DetRegular::generateLUT computed on first
	// use, or a table given by setDetectorLUT (a detector coordinate file,
	// DetCoord.cpp:263-320).
	const std::vector<float>& lutPositions() const;
	const std::vector<float>& lutOrientations() const;
	void setDetectorLUT(std::vector<float> pos, std::vector<float> ori);

	std::string scannerName;
	float axialFOV, crystalSize_z, crystalSize_trans, crystalDepth, scannerRadius;
	size_t detsPerRing, numRings, numDOI, maxRingDiff, minAngDiff, detsPerBlock;

private:
	void makeLUT() const;
	mutable std::vector<float> m_pos, m_ori;
};

// Image geometry (ImageParams.hpp:42-54; setup and completeDimInfo,
// ImageParams.cpp:206-266).
class ImageParams
{
public:
	ImageParams() = default;
	ImageParams(int64_t nx, int64_t ny, int64_t nz, float length_x,
	            float length_y, float length_z, float offset_x = 0.f,
	            float offset_y = 0.f, float offset_z = 0.f, int64_t nt = 1);
	// ImageParams::fromParams(nx, ny, nz, vx, vy, vz, nt) (ImageParams.cpp:177-181)
	static ImageParams fromParams(int64_t nx, int64_t ny, int64_t nz, float vx,
	                              float vy, float vz, int64_t nt = 1);
	bool isSameAs(const ImageParams& o) const;
	size_t numVoxels() const { return (size_t)(nx * ny * nz); }
	yrtmi_image desc(int64_t frames = -1, bool zeroOffset = false) const;

	int64_t nx = 0, ny = 0, nz = 0, nt = 1;
	float length_x = 0.f, length_y = 0.f, length_z = 0.f;
	float vx = 0.f, vy = 0.f, vz = 0.f;
	float off_x = 0.f, off_y = 0.f, off_z = 0.f;
	float fovRadius = 0.f;

private:
	void setup();
};

enum class ProjectorType
{
	SIDDON = 0,
	DD = 1
};

// ProjectorParams (ProjectorParams.hpp:27-71) with the projection-space PSF
// read from its one-row CSV (ProjectionPsfManager::readFromFileInternal,
// ProjectionPsfManager.cpp:27-49).
class ProjectorParams
{
public:
	explicit ProjectorParams(const Scanner& scanner);
	void setProjector(const std::string& projectorName);
	void setProjector(ProjectorType t) { projectorType = t; }
	void addTOF(float tofWidth_ps, int tofNumStd);
	void removeTOF() { m_tofWidth_ps = 0.f; m_tofNumStd = 0; m_hasTof = false; }
	float getTOFWidth_ps() const { return m_tofWidth_ps; }
	int getTOFNumStd() const { return m_tofNumStd; }
	bool hasTOF() const { return m_hasTof; }
	void setProjPsf(const std::string& fname);
	void setProjPsf(const std::vector<float>& kernels, int numS, int K,
	                float sStep, float kSpacing);
	bool hasProjPsf() const { return !m_psf.empty(); }
	yrtmi_projector_desc desc(bool withTOF) const;

	const Scanner& scanner;
	ProjectorType projectorType = ProjectorType::SIDDON;
	std::string projPsf_fname;
	int numRays = 1;
	bool setupFP64 = false;  // FP64 ray-plane setup (accuracy mode)

private:
	float m_tofWidth_ps = 0.f;
	int m_tofNumStd = 0;
	bool m_hasTof = false;
	std::vector<float> m_psf;
	int m_psfNumS = 0, m_psfK = 0;
	float m_psfSStep = 0.f, m_psfKSpacing = 0.f;
};

// A device, a stream and the scanner tables resident on it.
class Context
{
public:
	explicit Context(const Scanner& scanner, int device = 0, void* stream = nullptr);
	~Context();
	Context(const Context&) = delete;
	Context& operator=(const Context&) = delete;
	yrtmi_ctx* get() const { return m_ctx; }
	const Scanner& getScanner() const { return m_scanner; }
	int getDevice() const { return m_device; }
	bool hasHistogram() const { return m_numZBin > 0; }
	int64_t numZBin() const { return m_numZBin; }
	int64_t numPhi() const { return m_numPhi; }
	int64_t numR() const { return m_numR; }
	int64_t histogramSize() const { return m_numZBin * m_numPhi * m_numR; }
	void setStream(void* stream);
	// Give the context a stream of its own when it runs on the null stream
	// (graph recording needs one); waits for the null stream's work first.
	void ensureOwnStream();
	void synchronize() const;
	// the motion table of the context's scanner tables (device state, not
	// this object's)
	void uploadMotion(const std::vector<float>& transforms) const;  // k x 12

private:
	const Scanner& m_scanner;
	int m_device;
	yrtmi_ctx* m_ctx = nullptr;
	bool m_nullStream = true;
	int64_t m_numZBin = 0, m_numPhi = 0, m_numR = 0;
};

// Device array: owned (allocated here) or an alias of a caller's pointer.
template <typename T>
class DeviceArray
{
public:
	DeviceArray(const Context& ctx, size_t n);            // owned
	DeviceArray(const Context& ctx, T* alias, size_t n);  // alias
	~DeviceArray();
	DeviceArray(const DeviceArray&) = delete;
	DeviceArray& operator=(const DeviceArray&) = delete;
	T* data() const { return m_p; }
	size_t size() const { return m_n; }
	bool isAlias() const { return m_alias; }
	void copyFromHost(const T* src, size_t n);
	void copyToHost(T* dst, size_t n) const;

private:
	const Context& m_ctx;
	T* m_p = nullptr;
	size_t m_n = 0;
	bool m_alias = false;
};

// FP32 image float[nt][nz][ny][nx] on the device (ImageDevice.cuh).
class ImageDevice
{
public:
	// ImageDeviceOwned: allocated here.
	ImageDevice(const Context& ctx, const ImageParams& params);
	// ImageDeviceAlias: wraps `devicePointer` (nt * nz * ny * nx floats).
	ImageDevice(const Context& ctx, const ImageParams& params, float* devicePointer);
	const ImageParams& getParams() const { return m_params; }
	const Context& context() const { return m_ctx; }
	float* getDevicePointer() const { return m_buf->data(); }
	bool isAlias() const { return m_buf->isAlias(); }
	size_t size() const { return m_buf->size(); }
	void setValue(float v);
	void copyFromHost(const float* src);
	void copyToHost(float* dst) const;
	void copyFrom(const ImageDevice& other);

private:
	const Context& m_ctx;
	ImageParams m_params;
	std::unique_ptr<DeviceArray<float>> m_buf;
};

// List-mode events resident as SoA (ListModeLUT: detector pairs, TOF; the
// dynamic frame and LOR-motion frame of each event; an optional randoms
// estimate) plus a projection value per event (ProjectionListDevice).
class ListModeDevice
{
public:
	// Owned, from host arrays (tof, frame, motion, randoms nullable).
	ListModeDevice(const Context& ctx, size_t n, const uint32_t* d1,
	               const uint32_t* d2, const float* tof = nullptr,
	               const int32_t* frame = nullptr, const uint8_t* motion = nullptr,
	               const float* randoms = nullptr);
	// Alias of device arrays (nullable as above); the values array is owned.
	static std::unique_ptr<ListModeDevice> alias(const Context& ctx, size_t n,
	                                             uint32_t* d1, uint32_t* d2,
	                                             float* tof, int32_t* frame,
	                                             uint8_t* motion, float* randoms);
	size_t count() const { return m_n; }
	bool hasTOF() const { return m_tof != nullptr; }
	uint32_t* d1() const { return m_d1->data(); }
	uint32_t* d2() const { return m_d2->data(); }
	float* tof() const { return m_tof ? m_tof->data() : nullptr; }
	int32_t* frame() const { return m_frame ? m_frame->data() : nullptr; }
	uint8_t* motion() const { return m_motion ? m_motion->data() : nullptr; }
	float* randoms() const { return m_randoms ? m_randoms->data() : nullptr; }
	float* values() const { return m_values->data(); }
	// Projection values of the events (applyA output, applyAH input).
	void valuesToHost(float* dst) const { m_values->copyToHost(dst, m_n); }
	void valuesFromHost(const float* src) { m_values->copyFromHost(src, m_n); }
	// Dynamic framing: events carry frame indices < numFrames.
	int numFrames = 1;
	// LOR motion transforms (k x 12, transform_t) used by motion frames.
	std::vector<float> motionTransforms;
	yrtmi_lors lors() const;

private:
	explicit ListModeDevice(const Context& ctx) : m_ctx(ctx) {}
	const Context& m_ctx;
	size_t m_n = 0;
	std::unique_ptr<DeviceArray<uint32_t>> m_d1, m_d2;
	std::unique_ptr<DeviceArray<float>> m_tof, m_randoms, m_values;
	std::unique_ptr<DeviceArray<int32_t>> m_frame;
	std::unique_ptr<DeviceArray<uint8_t>> m_motion;
};

// Histogram3D values (numZBin x numPhi x numR floats) on the device
// (Histogram3D.cpp:165-203).
class Histogram3DDevice
{
public:
	explicit Histogram3DDevice(const Context& ctx);                  // owned, zeros
	Histogram3DDevice(const Context& ctx, float* devicePointer);     // alias
	float* getDevicePointer() const { return m_buf->data(); }
	size_t count() const { return m_buf->size(); }
	void copyFromHost(const float* src);
	void copyToHost(float* dst) const;
	yrtmi_lors lors() const;  // every bin, in bin order
	const Context& context() const { return m_ctx; }

private:
	const Context& m_ctx;
	std::unique_ptr<DeviceArray<float>> m_buf;
};

// A projector on the device (OperatorProjectorBase / OperatorProjectorDevice
// contract: applyA = forward projection into the projection data's values,
// applyAH = backprojection of them, added to the image).
class OperatorProjectorMI
{
public:
	OperatorProjectorMI(const Context& ctx, const ProjectorParams& params,
	                    bool withTOF = true);
	~OperatorProjectorMI();
	OperatorProjectorMI(const OperatorProjectorMI&) = delete;
	OperatorProjectorMI& operator=(const OperatorProjectorMI&) = delete;
	void applyA(const ImageDevice& img, ListModeDevice& proj);
	void applyAH(const ListModeDevice& proj, ImageDevice& img);
	void applyA(const ImageDevice& img, Histogram3DDevice& proj);
	void applyAH(const Histogram3DDevice& proj, ImageDevice& img);
	// Launch shape of the event kernels (lanes per event 1/2/4/8; 0: by rays).
	void setLaunch(int lanesPerEvent, int blockSize = 256, int gridBlocks = 0);
	yrtmi_proj* get() const { return m_proj; }
	const Context& getContext() const { return m_ctx; }
	int numRays() const { return m_numRays; }
	bool isSiddon() const { return m_siddon; }

private:
	const Context& m_ctx;
	yrtmi_proj* m_proj = nullptr;
	int m_numRays = 1;
	bool m_siddon = true;
};

// OSEM on the device (OSEM.hpp:29-194 with the OSEM_GPU driver hooks):
// reference members and setters; list-mode or Histogram3D input.
class OSEM_MI
{
public:
	static constexpr float DEFAULT_HARD_THRESHOLD = 1.0f;   // OSEM.hpp:34
	static constexpr float DEFAULT_DENOM_THRESHOLD = 1e-35f;
	static constexpr float INITIAL_VALUE_MLEM = 0.125f;
	static constexpr float EPS_FLT = 1.19209e-07f;         // Constants.hpp:28

	explicit OSEM_MI(const Scanner& scanner, int device = 0, void* stream = nullptr);
	~OSEM_MI();

	// configuration (OSEM.hpp:57-134)
	void setImageParams(const ImageParams& p) { m_imageParams = p; m_hasImageParams = true; }
	const ImageParams& getImageParams() const { return m_imageParams; }
	// Image-space PSF (OSEM::addImagePSF and the uniform Gaussian forms,
	// OSEM.cpp:86-101, OSEM_CPU.cpp:47-92; OperatorPsf.cpp:158-255): the
	// forward projection reads PSF(x), the update image is PSF^T of the
	// backprojected ratios, and each sensitivity image is PSF^T of the
	// backprojection (zero where that was <= EPS).  The file holds the x, y
	// and z taps on its first three rows and their lengths on the fourth.
	// The Gaussian forms need the image parameters set; a kernel size of 0
	// takes rint(5 sigma / voxel size) taps.
	void addImagePSF(const std::string& fname);
	void addUniformGaussianImagePSFFromFWHM(float fwhmX, float fwhmY, float fwhmZ,
	                                        size_t kerSizeX = 0, size_t kerSizeY = 0,
	                                        size_t kerSizeZ = 0);
	void addUniformGaussianImagePSFFromSigma(float sigmaX, float sigmaY, float sigmaZ,
	                                         size_t kerSizeX = 0, size_t kerSizeY = 0,
	                                         size_t kerSizeZ = 0);
	bool hasImagePSF() const { return !m_psf[0].empty(); }
	const std::vector<float>& imagePSFKernel(int axis) const { return m_psf[axis]; }
	void setDataInput(const ListModeDevice* lm);
	void setDataInput(const Histogram3DDevice* h);
	// Batched histogram frames: frame f of the reconstructed image is the
	// OSEM reconstruction of frames[f] (the reference reconstructs one
	// Histogram3D at a time); the frames share the scanner's sensitivity
	// images and every other correction, and each subset of all frames is
	// projected in one launch (the nonzero bins of every frame, each with
	// its frame index).
	void setDataInput(const std::vector<const Histogram3DDevice*>& frames);
	bool isListModeEnabled() const { return m_lm != nullptr; }
	// Histogram frames of the data input (0 for list-mode).
	int numHistogramFrames() const { return (int)m_histos.size(); }
	void setProjector(const std::string& name) { projectorParams.setProjector(name); }
	void setNumRays(int n) { projectorParams.numRays = n; }
	void addTOF(float widthPs, int numStd) { projectorParams.addTOF(widthPs, numStd); }
	void addProjPSF(const std::string& fname) { projectorParams.setProjPsf(fname); }
	void setGlobalScalingFactor(float g) { m_globalScale = g; }
	void setAttenuationImage(const ImageDevice* mu) { m_att = mu; }
	// The attenuation split (Corrector.cpp:34-166, 191-221): hardware
	// attenuation weighs the sensitivity images, the total attenuation the
	// EM model, the in-vivo attenuation precorrects the EM denominator.  Each
	// comes from an image (exp(-0.1 * Siddon(mu)) per LOR) or from an ACF
	// histogram; missing parts are implied as Corrector::setup does (a total
	// alone is all hardware, hardware + in-vivo images sum to the total, in-vivo
	// x hardware ACF histograms multiply to the total).  For list-mode data
	// the ACF histograms are read at each event's detector pair.
	void setHardwareAttenuationImage(const ImageDevice* mu) { m_hwAtt = mu; }
	void setInVivoAttenuationImage(const ImageDevice* mu) { m_ivAtt = mu; }
	void setACFHistogram(const Histogram3DDevice* h) { m_acf = h; }
	void setHardwareACFHistogram(const Histogram3DDevice* h) { m_hwAcf = h; }
	void setInVivoACFHistogram(const Histogram3DDevice* h) { m_ivAcf = h; }
	void setRandomsHistogram(const Histogram3DDevice* h) { m_randomsHisto = h; }
	void setScatterHistogram(const Histogram3DDevice* h) { m_scatterHisto = h; }
	void setSensitivityHistogram(const Histogram3DDevice* h) { m_sensHisto = h; }
	void setInvertSensitivity(bool inv = true) { m_invertSens = inv; }
	void setInitialEstimate(const ImageDevice* x) { m_initial = x; }
	void setMaskImage(const ImageDevice* m) { m_mask = m; }
	int getExpectedSensImagesAmount() const;
	// OSEM::setSensitivityImages: copies of the given images are kept.
	void setSensitivityImages(const std::vector<const ImageDevice*>& imgs);
	// OSEM::generateSensitivityImages (OSEM.cpp:926-986): one image per
	// subset (one for list-mode), thresholded; kept and returned.
	const std::vector<std::unique_ptr<ImageDevice>>& generateSensitivityImages();
	const std::vector<std::unique_ptr<ImageDevice>>& getSensitivityImages() const
	{
		return m_sens;
	}
	// OSEM::reconstruct (OSEM.cpp:315-365, 788-811).
	std::unique_ptr<ImageDevice> reconstruct();

	int num_MLEM_iterations = 10;
	int num_OSEM_subsets = 1;
	float hardThreshold = DEFAULT_HARD_THRESHOLD;
	float denomThreshold = DEFAULT_DENOM_THRESHOLD;
	ProjectorParams projectorParams;
	// Backprojection through tile plans (Siddon and DD): built once per
	// subset.
	bool useTilePlans = true;
	// Project each list-mode subset's events in locality order
	// (yrtmi_lor_order); the EM update does not depend on the order.
	bool useLorOrder = true;
	int lanesPerEvent = 0;  // 0: chosen from the number of rays
	// Launch one OSEM iteration as a recorded graph: the first iteration
	// runs call by call, the next is recorded once and replayed for every
	// later iteration (the subset loop of OSEM.cpp:788-811).
	bool useGraphs = true;
	// Before the tile plans are built: lanes per event, the tile shape and a
	// persistent grid chosen by timing them on the first subset
	// (yrtmi_autotune).
	bool useAutotune = true;
	// Device memory the subsets' tile plans may hold, bytes: 0 = 90 % of
	// the device's free memory when they are built (OSEM_GPU sizes its
	// batches from that share, OSEM_GPU.cu:413-435, or from
	// OperatorProjectorDevice::create's memAvailable).  Subsets whose plan
	// does not fit (in subset order) run the unplanned EM kernel.
	size_t memoryBudget = 0;
	// Subsets that got a tile plan in the last reconstruct().
	int numPlannedSubsets() const { return m_numPlanned; }
	// Time each iteration of reconstruct() (wall seconds, the device
	// synchronized after each; off by default: it adds a synchronization per
	// iteration).  iterationSeconds() holds the last reconstruct()'s.
	bool timeIterations = false;
	const std::vector<double>& iterationSeconds() const { return m_iterSeconds; }

	const Context& getContext() const { return *m_ctx; }

private:
	struct Subset;
	std::unique_ptr<OperatorProjectorMI> makeProjector(bool withTOF) const;
	bool planned() const;
	// The subsets' device data, corrections and tile plans; x: the initial
	// image (for the autotuner)
	void prepareSubsets(OperatorProjectorMI& proj, const ImageParams& imt, const float* x);
	ImageParams singleFrame() const;

	std::unique_ptr<Context> m_ctx;
	std::vector<float> m_psf[3];  // image-space PSF taps along x, y, z
	ImageParams m_imageParams;
	bool m_hasImageParams = false;
	const ListModeDevice* m_lm = nullptr;
	std::vector<const Histogram3DDevice*> m_histos;
	int m_numPlanned = 0;
	std::vector<double> m_iterSeconds;
	int m_tileShape = -1, m_persistent = 0;  // from the autotuner
	float m_globalScale = 1.0f;
	const ImageDevice* m_att = nullptr;
	const ImageDevice *m_hwAtt = nullptr, *m_ivAtt = nullptr;
	const Histogram3DDevice *m_acf = nullptr, *m_hwAcf = nullptr, *m_ivAcf = nullptr;
	std::unique_ptr<ImageDevice> m_impliedAtt;  // hardware + in-vivo images
	struct Attenuation
	{
		const ImageDevice *total, *hardware, *inVivo;
		const Histogram3DDevice *acf, *hwAcf, *ivAcf;
		bool totalFromHisto;
	};
	Attenuation resolveAttenuation();
	const Histogram3DDevice *m_randomsHisto = nullptr, *m_scatterHisto = nullptr,
	                        *m_sensHisto = nullptr;
	bool m_invertSens = false;
	const ImageDevice* m_initial = nullptr;
	const ImageDevice* m_mask = nullptr;
	std::vector<std::unique_ptr<ImageDevice>> m_sens;
	std::vector<std::unique_ptr<Subset>> m_subsets;
};

// OperatorPsf::createGaussianKernel1D (OperatorPsf.cpp:158-185): `size`
// taps (0: rint(5 sigma / voxel size)) exp(-0.5 ((i - size/2) v / sigma)^2),
// normalised by their sum.
std::vector<float> gaussianKernel1D(float sigma, float voxelSize, size_t size = 0);

// ---- motion: time-averaged (moved) images --------------------------------
// The motion is given as its frames' starting timestamps (ms, increasing)
// and transforms (k x 12, transform_t rows r00 r01 r02 tx r10 ... tz), the
// fields of a LORMotion.
// util::getFullTimeRange (ReconstructionUtils.cpp:468-495): [first start,
// first start + 1] for one frame, else [first start, last start + the
// second-to-last frame's duration].
std::pair<int64_t, int64_t> getFullTimeRange(const std::vector<int64_t>& timestamps);
// The motion frames a time average over [timeStart, timeStop] takes and their
// weights (ReconstructionUtilsDevice.cu:146-190): the frames from the first
// start >= timeStart up to (excluding) the first start > timeStop, each
// weighted LORMotion::getDuration (LORMotion.cpp:73-88) / (timeStop -
// timeStart).  Returns the first frame; `weights` receives one weight per
// frame taken.  A motion of one frame has no second-to-last frame to take a
// duration from (the reference reads before its records there): its frame
// lasts the whole range and gets weight 1.
int64_t motionFrameWeights(const std::vector<int64_t>& timestamps, int64_t timeStart,
                           int64_t timeStop, std::vector<float>& weights);
// util::timeAverageMoveImageDevice (ReconstructionUtilsDevice.cu:120-225):
// adds to frame `outFrame` of `out` the copies of frame 0 of `unmoved`
// moved by every motion frame starting in [timeStart, timeStop] (both
// included), weighted by LORMotion::getDuration / (timeStop - timeStart).
void timeAverageMoveImageDevice(const std::vector<int64_t>& timestamps,
                                const std::vector<float>& transforms,
                                const ImageDevice& unmoved, ImageDevice& out,
                                int64_t timeStart, int64_t timeStop, int64_t outFrame = 0);
// The same into a new zeroed one-frame image (ReconstructionUtilsDevice.cu:
// 97-118); without a time range over getFullTimeRange.
std::unique_ptr<ImageDevice> timeAverageMoveImageDevice(const std::vector<int64_t>& timestamps,
                                                        const std::vector<float>& transforms,
                                                        const ImageDevice& unmoved);
std::unique_ptr<ImageDevice> timeAverageMoveImageDevice(const std::vector<int64_t>& timestamps,
                                                        const std::vector<float>& transforms,
                                                        const ImageDevice& unmoved,
                                                        int64_t timeStart, int64_t timeStop);
// util::timeAverageMoveImageDynamicDevice (ReconstructionUtilsDevice.cu:
// 227-300): one average per dynamic frame f over [bounds[f], bounds[f+1]]
// into frame f of a new image of bounds.size() - 1 frames.
std::unique_ptr<ImageDevice> timeAverageMoveImageDynamicDevice(
    const std::vector<int64_t>& timestamps, const std::vector<float>& transforms,
    const ImageDevice& unmoved, const std::vector<uint32_t>& bounds);

// util::createOSEM(scanner, useGPU) (ReconstructionUtils.cpp:915-933): this
// library provides the GPU OSEM only; useGPU = false throws.
std::unique_ptr<OSEM_MI> createOSEM(const Scanner& scanner, bool useGPU = true,
                                    int device = 0);

// util::forwProject / util::backProject (ReconstructionUtils.cpp:949-1061)
// on device data.
void forwProject(const Context& ctx, const ProjectorParams& params,
                 const ImageDevice& img, ListModeDevice& proj);
void backProject(const Context& ctx, const ProjectorParams& params,
                 ImageDevice& img, const ListModeDevice& proj);
void forwProject(const Context& ctx, const ProjectorParams& params,
                 const ImageDevice& img, Histogram3DDevice& proj);
void backProject(const Context& ctx, const ProjectorParams& params,
                 ImageDevice& img, const Histogram3DDevice& proj);

}  // namespace yrtmi

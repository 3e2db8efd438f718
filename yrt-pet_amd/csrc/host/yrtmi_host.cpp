// C++ host layer over the yrtmi C ABI (see yrtmi_host.hpp).  Host code only:
// every device operation goes through include/yrtmi.h.
#include "yrtmi_host.hpp"
#include "yrtmi_io.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <sstream>
#include <type_traits>

namespace yrtmi
{

void check(int rc, const char* what)
{
	if (rc == YRTMI_OK)
		return;
	std::string m = yrtmi_last_error();
	if (what != nullptr)
		m = std::string(what) + ": " + m;
	throw Error(m);
}

// ---------------------------------------------------------------------------
// Scanner
// ---------------------------------------------------------------------------
Scanner::Scanner(std::string name, float axialFOV_, float crystalSize_z_,
                 float crystalSize_trans_, float crystalDepth_,
                 float scannerRadius_, size_t detsPerRing_, size_t numRings_,
                 size_t numDOI_, size_t maxRingDiff_, size_t minAngDiff_,
                 size_t detsPerBlock_)
    : scannerName(std::move(name)), axialFOV(axialFOV_),
      crystalSize_z(crystalSize_z_), crystalSize_trans(crystalSize_trans_),
      crystalDepth(crystalDepth_), scannerRadius(scannerRadius_),
      detsPerRing(detsPerRing_), numRings(numRings_), numDOI(numDOI_),
      maxRingDiff(maxRingDiff_), minAngDiff(minAngDiff_),
      detsPerBlock(detsPerBlock_)
{
	if (numRings == 0 || detsPerRing == 0 || numDOI == 0)
		throw Error("Scanner: detsPerRing, numRings and numDOI must be positive");
	if (maxRingDiff >= numRings)
		throw Error("Scanner: maxRingDiff must be smaller than numRings");
	if (minAngDiff == 0)
		throw Error("Scanner: minAngDiff must be positive");
}

size_t Scanner::getNumDets() const
{
	return detsPerRing * numRings * numDOI;
}

void Scanner::makeLUT() const
{
	if (!m_pos.empty())
		return;
	const size_t n = getNumDets();
	m_pos.assign(3 * n, 0.f);
	m_ori.assign(3 * n, 0.f);
	yrtmi_scanner_desc d{axialFOV, crystalSize_z, crystalSize_trans, crystalDepth,
	                     scannerRadius, (int64_t)detsPerRing, (int64_t)numRings,
	                     (int64_t)numDOI, (int64_t)detsPerBlock};
	check(yrtmi_lut_regular(&d, m_pos.data(), m_ori.data()), "Scanner LUT");
}

const std::vector<float>& Scanner::lutPositions() const
{
	makeLUT();
	return m_pos;
}

const std::vector<float>& Scanner::lutOrientations() const
{
	makeLUT();
	return m_ori;
}

void Scanner::setDetectorLUT(std::vector<float> pos, std::vector<float> ori)
{
	if (pos.size() != 3 * getNumDets() || ori.size() != pos.size())
		throw Error("Scanner: the detector table does not hold num_doi * num_rings * "
		            "dets_per_ring detectors");
	m_pos = std::move(pos);
	m_ori = std::move(ori);
}

// ---------------------------------------------------------------------------
// ImageParams
// ---------------------------------------------------------------------------
ImageParams::ImageParams(int64_t nx_, int64_t ny_, int64_t nz_, float lx,
                         float ly, float lz, float ox, float oy, float oz,
                         int64_t nt_)
    : nx(nx_), ny(ny_), nz(nz_), nt(nt_), length_x(lx), length_y(ly),
      length_z(lz), vx(-1.f), vy(-1.f), vz(-1.f), off_x(ox), off_y(oy), off_z(oz)
{
	setup();
}

ImageParams ImageParams::fromParams(int64_t nx, int64_t ny, int64_t nz, float vx,
                                    float vy, float vz, int64_t nt)
{
	ImageParams p(nx, ny, nz, (float)nx * vx, (float)ny * vy, (float)nz * vz, 0.f,
	              0.f, 0.f, nt);
	// keep the given voxel sizes (completeDimInfo returns early when both
	// the voxel size and the length are set)
	p.vx = vx;
	p.vy = vy;
	p.vz = vz;
	p.setup();
	return p;
}

void ImageParams::setup()
{
	if (nx <= 0 || ny <= 0 || nz <= 0 || nt <= 0)
		throw Error("ImageParams: dimensions must be positive");
	if (!(length_x > 0.f && length_y > 0.f && length_z > 0.f))
		throw Error("ImageParams: lengths must be positive");
	// completeDimInfo (ImageParams.cpp:220-266): v = length / n when unset
	if (!(vx > 0.f))
		vx = length_x / (float)nx;
	if (!(vy > 0.f))
		vy = length_y / (float)ny;
	if (!(vz > 0.f))
		vz = length_z / (float)nz;
	// ImageParams::setup (ImageParams.cpp:216-217)
	fovRadius = std::max(length_x / 2.0f, length_y / 2.0f);
	fovRadius -= std::max(vx, vy) / 1000.0f;
}

bool ImageParams::isSameAs(const ImageParams& o) const
{
	return nx == o.nx && ny == o.ny && nz == o.nz && nt == o.nt &&
	       length_x == o.length_x && length_y == o.length_y &&
	       length_z == o.length_z && off_x == o.off_x && off_y == o.off_y &&
	       off_z == o.off_z;
}

yrtmi_image ImageParams::desc(int64_t frames, bool zeroOffset) const
{
	yrtmi_image d{};
	d.nx = nx;
	d.ny = ny;
	d.nz = nz;
	d.nt = frames > 0 ? frames : nt;
	d.vx = vx;
	d.vy = vy;
	d.vz = vz;
	d.lx = length_x;
	d.ly = length_y;
	d.lz = length_z;
	d.ox = zeroOffset ? 0.f : off_x;
	d.oy = zeroOffset ? 0.f : off_y;
	d.oz = zeroOffset ? 0.f : off_z;
	d.fov_radius = fovRadius;
	return d;
}

// ---------------------------------------------------------------------------
// ProjectorParams
// ---------------------------------------------------------------------------
ProjectorParams::ProjectorParams(const Scanner& s) : scanner(s) {}

void ProjectorParams::setProjector(const std::string& name)
{
	std::string n = name;
	std::transform(n.begin(), n.end(), n.begin(), ::toupper);
	if (n == "S" || n == "SIDDON")
		projectorType = ProjectorType::SIDDON;
	else if (n == "D" || n == "DD")
		projectorType = ProjectorType::DD;
	else
		throw Error("Unknown projector: " + name);
}

void ProjectorParams::addTOF(float w, int nstd)
{
	m_tofWidth_ps = w;
	m_tofNumStd = nstd;
	m_hasTof = true;
}

void ProjectorParams::setProjPsf(const std::string& fname)
{
	// one CSV row: sStep, kSpacing, K, then K * numS taps
	std::ifstream f(fname);
	if (!f)
		throw Error("Could not open projection PSF file " + fname);
	std::string line;
	std::getline(f, line);
	std::stringstream ss(line);
	std::vector<float> row;
	std::string tok;
	while (std::getline(ss, tok, ','))
		row.push_back(std::stof(tok));
	if (row.size() < 4)
		throw Error("Projection PSF file too short: " + fname);
	const int K = (int)row[2];
	if (K <= 0 || K % 2 == 0)
		throw Error("Kernels must be of odd length");
	const int numS = (int)((row.size() - 3) / (size_t)K);
	setProjPsf(std::vector<float>(row.begin() + 3, row.begin() + 3 + (size_t)numS * K),
	           numS, K, row[0], row[1]);
	projPsf_fname = fname;
}

void ProjectorParams::setProjPsf(const std::vector<float>& k, int numS, int K,
                                 float sStep, float kSpacing)
{
	if (numS <= 0 || K <= 0 || k.size() != (size_t)numS * K)
		throw Error("Projection PSF: kernel table size mismatch");
	m_psf = k;
	m_psfNumS = numS;
	m_psfK = K;
	m_psfSStep = sStep;
	m_psfKSpacing = kSpacing;
}

yrtmi_projector_desc ProjectorParams::desc(bool withTOF) const
{
	yrtmi_projector_desc d{};
	d.type = projectorType == ProjectorType::DD ? YRTMI_PROJ_DD : YRTMI_PROJ_SIDDON;
	d.num_rays = numRays;
	d.use_tof = (withTOF && m_hasTof) ? 1 : 0;
	d.tof_width_ps = m_tofWidth_ps;
	d.tof_n_std = m_tofNumStd;
	d.psf_kernels = m_psf.empty() ? nullptr : m_psf.data();
	d.psf_num_s = m_psfNumS;
	d.psf_ksize = m_psfK;
	d.psf_s_step = m_psfSStep;
	d.psf_k_spacing = m_psfKSpacing;
	d.crystal_size_z = scanner.crystalSize_z;
	d.crystal_size_trans = scanner.crystalSize_trans;
	d.crystal_depth = scanner.crystalDepth;
	d.setup_fp64 = setupFP64 ? 1 : 0;
	return d;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
Context::Context(const Scanner& s, int device, void* stream)
    : m_scanner(s), m_device(device), m_nullStream(stream == nullptr)
{
	check(yrtmi_ctx_create(device, stream, &m_ctx), "context");
	const auto& pos = s.lutPositions();
	const auto& ori = s.lutOrientations();
	check(yrtmi_lut_upload(m_ctx, pos.data(), ori.data(), (int64_t)s.getNumDets()),
	      "LUT upload");
	if (s.detsPerRing % 2 == 0)
	{
		yrtmi_histo_desc h{(int64_t)s.detsPerRing, (int64_t)s.numRings, (int64_t)s.numDOI,
		                   (int64_t)s.maxRingDiff, (int64_t)s.minAngDiff};
		check(yrtmi_histo_setup(m_ctx, &h, &m_numZBin, &m_numPhi, &m_numR),
		      "histogram setup");
	}
}

Context::~Context()
{
	yrtmi_ctx_destroy(m_ctx);
}

void Context::setStream(void* stream)
{
	check(yrtmi_ctx_set_stream(m_ctx, stream), "set stream");
	m_nullStream = stream == nullptr;
}

void Context::ensureOwnStream()
{
	if (!m_nullStream)
		return;
	synchronize();
	check(yrtmi_ctx_own_stream(m_ctx, nullptr), "own stream");
	m_nullStream = false;
}

void Context::synchronize() const
{
	check(yrtmi_ctx_synchronize(m_ctx), "synchronize");
}

void Context::uploadMotion(const std::vector<float>& xf) const
{
	if (xf.size() % 12 != 0)
		throw Error("motion transforms: 12 floats each");
	check(yrtmi_motion_upload(m_ctx, xf.data(), (int)(xf.size() / 12)), "motion upload");
}

// ---------------------------------------------------------------------------
// DeviceArray
// ---------------------------------------------------------------------------
template <typename T>
DeviceArray<T>::DeviceArray(const Context& ctx, size_t n) : m_ctx(ctx), m_n(n)
{
	void* p = nullptr;
	check(yrtmi_device_alloc(ctx.get(), (int64_t)(n * sizeof(T)), &p), "device alloc");
	m_p = static_cast<T*>(p);
}

template <typename T>
DeviceArray<T>::DeviceArray(const Context& ctx, T* alias, size_t n)
    : m_ctx(ctx), m_p(alias), m_n(n), m_alias(true)
{
	if (alias == nullptr && n > 0)
		throw Error("alias of a null device pointer");
}

template <typename T>
DeviceArray<T>::~DeviceArray()
{
	if (!m_alias && m_p != nullptr)
		yrtmi_device_free(m_ctx.get(), m_p);
}

template <typename T>
void DeviceArray<T>::copyFromHost(const T* src, size_t n)
{
	if (n > m_n)
		throw Error("copy larger than the device array");
	check(yrtmi_copy(m_ctx.get(), m_p, src, (int64_t)(n * sizeof(T)), 0), "upload");
}

template <typename T>
void DeviceArray<T>::copyToHost(T* dst, size_t n) const
{
	if (n > m_n)
		throw Error("copy larger than the device array");
	check(yrtmi_copy(m_ctx.get(), dst, m_p, (int64_t)(n * sizeof(T)), 1), "download");
}

template class DeviceArray<float>;
template class DeviceArray<uint32_t>;
template class DeviceArray<int32_t>;
template class DeviceArray<uint8_t>;

// ---------------------------------------------------------------------------
// ImageDevice
// ---------------------------------------------------------------------------
ImageDevice::ImageDevice(const Context& ctx, const ImageParams& params)
    : m_ctx(ctx), m_params(params),
      m_buf(new DeviceArray<float>(ctx, params.numVoxels() * (size_t)params.nt))
{
	setValue(0.f);
}

ImageDevice::ImageDevice(const Context& ctx, const ImageParams& params, float* ptr)
    : m_ctx(ctx), m_params(params),
      m_buf(new DeviceArray<float>(ctx, ptr, params.numVoxels() * (size_t)params.nt))
{
}

void ImageDevice::setValue(float v)
{
	check(yrtmi_fill(m_ctx.get(), m_buf->data(), (int64_t)m_buf->size(), v), "fill");
}

void ImageDevice::copyFromHost(const float* src)
{
	m_buf->copyFromHost(src, m_buf->size());
}

void ImageDevice::copyToHost(float* dst) const
{
	m_buf->copyToHost(dst, m_buf->size());
}

void ImageDevice::copyFrom(const ImageDevice& o)
{
	if (o.size() != size())
		throw Error("image copy: size mismatch");
	check(yrtmi_copy(m_ctx.get(), m_buf->data(), o.getDevicePointer(),
	                 (int64_t)(size() * sizeof(float)), 2),
	      "image copy");
}

// ---------------------------------------------------------------------------
// ListModeDevice / Histogram3DDevice
// ---------------------------------------------------------------------------
ListModeDevice::ListModeDevice(const Context& ctx, size_t n, const uint32_t* d1,
                               const uint32_t* d2, const float* tof,
                               const int32_t* frame, const uint8_t* motion,
                               const float* randoms)
    : m_ctx(ctx), m_n(n)
{
	m_d1.reset(new DeviceArray<uint32_t>(ctx, n));
	m_d2.reset(new DeviceArray<uint32_t>(ctx, n));
	m_d1->copyFromHost(d1, n);
	m_d2->copyFromHost(d2, n);
	if (tof)
	{
		m_tof.reset(new DeviceArray<float>(ctx, n));
		m_tof->copyFromHost(tof, n);
	}
	if (frame)
	{
		m_frame.reset(new DeviceArray<int32_t>(ctx, n));
		m_frame->copyFromHost(frame, n);
	}
	if (motion)
	{
		m_motion.reset(new DeviceArray<uint8_t>(ctx, n));
		m_motion->copyFromHost(motion, n);
	}
	if (randoms)
	{
		m_randoms.reset(new DeviceArray<float>(ctx, n));
		m_randoms->copyFromHost(randoms, n);
	}
	m_values.reset(new DeviceArray<float>(ctx, n));
}

std::unique_ptr<ListModeDevice> ListModeDevice::alias(const Context& ctx, size_t n,
                                                      uint32_t* d1, uint32_t* d2,
                                                      float* tof, int32_t* frame,
                                                      uint8_t* motion, float* randoms)
{
	std::unique_ptr<ListModeDevice> lm(new ListModeDevice(ctx));
	lm->m_n = n;
	lm->m_d1.reset(new DeviceArray<uint32_t>(ctx, d1, n));
	lm->m_d2.reset(new DeviceArray<uint32_t>(ctx, d2, n));
	if (tof)
		lm->m_tof.reset(new DeviceArray<float>(ctx, tof, n));
	if (frame)
		lm->m_frame.reset(new DeviceArray<int32_t>(ctx, frame, n));
	if (motion)
		lm->m_motion.reset(new DeviceArray<uint8_t>(ctx, motion, n));
	if (randoms)
		lm->m_randoms.reset(new DeviceArray<float>(ctx, randoms, n));
	lm->m_values.reset(new DeviceArray<float>(ctx, n));
	return lm;
}

yrtmi_lors ListModeDevice::lors() const
{
	yrtmi_lors l{};
	l.kind = YRTMI_SRC_PAIRS;
	l.n = (int64_t)m_n;
	l.d1 = d1();
	l.d2 = d2();
	l.num_subsets = 1;
	l.tof = tof();
	l.frame = frame();
	l.motion_frame = motion();
	return l;
}

Histogram3DDevice::Histogram3DDevice(const Context& ctx) : m_ctx(ctx)
{
	if (!ctx.hasHistogram())
		throw Error("Histogram3D needs an even number of detectors per ring");
	m_buf.reset(new DeviceArray<float>(ctx, (size_t)ctx.histogramSize()));
	check(yrtmi_fill(ctx.get(), m_buf->data(), (int64_t)m_buf->size(), 0.f), "fill");
}

Histogram3DDevice::Histogram3DDevice(const Context& ctx, float* ptr) : m_ctx(ctx)
{
	if (!ctx.hasHistogram())
		throw Error("Histogram3D needs an even number of detectors per ring");
	m_buf.reset(new DeviceArray<float>(ctx, ptr, (size_t)ctx.histogramSize()));
}

void Histogram3DDevice::copyFromHost(const float* src)
{
	m_buf->copyFromHost(src, m_buf->size());
}

void Histogram3DDevice::copyToHost(float* dst) const
{
	m_buf->copyToHost(dst, m_buf->size());
}

yrtmi_lors Histogram3DDevice::lors() const
{
	yrtmi_lors l{};
	l.kind = YRTMI_SRC_HISTO_SUBSET;
	l.n = (int64_t)m_buf->size();
	l.num_subsets = 1;
	l.subset = 0;
	return l;
}

// ---------------------------------------------------------------------------
// OperatorProjectorMI
// ---------------------------------------------------------------------------
OperatorProjectorMI::OperatorProjectorMI(const Context& ctx, const ProjectorParams& pp,
                                         bool withTOF)
    : m_ctx(ctx)
{
	yrtmi_projector_desc d = pp.desc(withTOF);
	check(yrtmi_proj_create(ctx.get(), &d, &m_proj), "projector");
	m_siddon = pp.projectorType == ProjectorType::SIDDON;
	m_numRays = m_siddon ? pp.numRays : 1;
	setLaunch(0);
}

OperatorProjectorMI::~OperatorProjectorMI()
{
	yrtmi_proj_destroy(m_proj);
}

void OperatorProjectorMI::setLaunch(int lpe, int block, int grid)
{
	if (lpe <= 0)
		lpe = m_numRays >= 4 ? 4 : (m_numRays >= 2 ? 2 : 1);
	check(yrtmi_proj_set_launch(m_proj, lpe, block, grid), "launch shape");
}

static void needFrames(const ImageDevice& img, int numFrames)
{
	if (img.getParams().nt != numFrames)
		throw Error("image frames do not match the projection data's dynamic framing");
}

void OperatorProjectorMI::applyA(const ImageDevice& img, ListModeDevice& proj)
{
	needFrames(img, proj.numFrames);
	// LORs moved by the list-mode's motion (ProjectionData::getLOR,
	// ProjectionData.cpp:302-317)
	if (!proj.motionTransforms.empty())
		m_ctx.uploadMotion(proj.motionTransforms);
	const yrtmi_image im = img.getParams().desc();
	const yrtmi_lors l = proj.lors();
	check(yrtmi_forward(m_ctx.get(), m_proj, &im, img.getDevicePointer(), &l, proj.values()),
	      "applyA");
}

void OperatorProjectorMI::applyAH(const ListModeDevice& proj, ImageDevice& img)
{
	needFrames(img, proj.numFrames);
	if (!proj.motionTransforms.empty())
		m_ctx.uploadMotion(proj.motionTransforms);
	const yrtmi_image im = img.getParams().desc();
	const yrtmi_lors l = proj.lors();
	check(yrtmi_back(m_ctx.get(), m_proj, &im, img.getDevicePointer(), &l, proj.values()),
	      "applyAH");
}

void OperatorProjectorMI::applyA(const ImageDevice& img, Histogram3DDevice& proj)
{
	needFrames(img, 1);
	const yrtmi_image im = img.getParams().desc();
	const yrtmi_lors l = proj.lors();
	check(yrtmi_forward(m_ctx.get(), m_proj, &im, img.getDevicePointer(), &l,
	                    proj.getDevicePointer()),
	      "applyA");
}

void OperatorProjectorMI::applyAH(const Histogram3DDevice& proj, ImageDevice& img)
{
	needFrames(img, 1);
	const yrtmi_image im = img.getParams().desc();
	const yrtmi_lors l = proj.lors();
	check(yrtmi_back(m_ctx.get(), m_proj, &im, img.getDevicePointer(), &l,
	                 proj.getDevicePointer()),
	      "applyAH");
}

// ---------------------------------------------------------------------------
// OSEM_MI
// ---------------------------------------------------------------------------
namespace
{
// An owned plan handle.
struct PlanHandle
{
	yrtmi_plan* p = nullptr;
	PlanHandle() = default;
	explicit PlanHandle(yrtmi_plan* q) : p(q) {}
	~PlanHandle() { yrtmi_plan_destroy(p); }
	PlanHandle(const PlanHandle&) = delete;
	PlanHandle& operator=(const PlanHandle&) = delete;
};

// The default tile shape of a projector and image: Siddon 32 x 32 x 8 (flat
// tiles suit the mostly transaxial rays; the fastest shape measured on
// MI355X for list-mode TOF subsets and for sensitivity images); DD 16^3
// (five workgroups per CU) while the tile count (frames included) stays
// within the plan's bin limit (5459 tiles with 64 KB of LDS per
// workgroup), 32 x 32 x 8 beyond it.
int defaultTileShape(const OperatorProjectorMI& proj, const yrtmi_image& im)
{
	if (proj.isSiddon())
		return YRTMI_TILE_32x32x8;
	const int64_t t16 =
	    ((im.nx + 15) / 16) * (int64_t)((im.ny + 15) / 16) * ((im.nz + 15) / 16) * im.nt;
	return t16 <= 5459 ? YRTMI_TILE_16x16x16 : YRTMI_TILE_32x32x8;
}

// A tile plan of `l`; with maxBytes > 0, null when the plan would hold more
// device memory than that (or the allocation fails): the caller then uses
// the unplanned kernels.  singleUse: a plan backprojected once (its records
// are not ordered by length).
std::unique_ptr<PlanHandle> makePlan(const Context& ctx, const OperatorProjectorMI& proj,
                                     const yrtmi_image& im, const yrtmi_lors& l,
                                     int shape = -1, int persistent = 0, int64_t maxBytes = 0,
                                     bool singleUse = false)
{
	const yrtmi_plan_options o{0, shape >= 0 ? shape : defaultTileShape(proj, im), persistent,
	                           maxBytes, singleUse ? 1 : 0};
	yrtmi_plan* q = nullptr;
	const int rc = yrtmi_plan_create(ctx.get(), proj.get(), &im, &l, &o, &q);
	if (rc == YRTMI_ENOMEM && maxBytes > 0)
		return nullptr;
	check(rc, "tile plan");
	return std::unique_ptr<PlanHandle>(new PlanHandle(q));
}

int64_t planBytes(const PlanHandle& p)
{
	int64_t b = 0;
	check(yrtmi_plan_info(p.p, nullptr, nullptr, &b), "plan info");
	return b;
}
}  // namespace

// One subset's device data, kept for the whole reconstruction.
struct OSEM_MI::Subset
{
	yrtmi_lors lors{};
	std::unique_ptr<DeviceArray<uint32_t>> d1, d2, bins;
	std::unique_ptr<DeviceArray<float>> tof, meas, att, sens, randoms, scatter, pre;
	std::unique_ptr<DeviceArray<int32_t>> frame;
	std::unique_ptr<DeviceArray<uint8_t>> motion;
	std::unique_ptr<PlanHandle> plan;
};

OSEM_MI::OSEM_MI(const Scanner& scanner, int device, void* stream)
    : projectorParams(scanner), m_ctx(new Context(scanner, device, stream))
{
}

OSEM_MI::~OSEM_MI() = default;

void OSEM_MI::setDataInput(const ListModeDevice* lm)
{
	m_lm = lm;
	m_histos.clear();
}

void OSEM_MI::setDataInput(const Histogram3DDevice* h)
{
	m_histos.clear();
	if (h != nullptr)
		m_histos.push_back(h);
	m_lm = nullptr;
}

void OSEM_MI::setDataInput(const std::vector<const Histogram3DDevice*>& frames)
{
	for (const Histogram3DDevice* h : frames)
		if (h == nullptr)
			throw Error("setDataInput: a null histogram frame");
	if (frames.empty())
		throw Error("setDataInput: no histogram frames");
	m_histos = frames;
	m_lm = nullptr;
}

int OSEM_MI::getExpectedSensImagesAmount() const
{
	return isListModeEnabled() ? 1 : num_OSEM_subsets;
}

std::unique_ptr<OperatorProjectorMI> OSEM_MI::makeProjector(bool withTOF) const
{
	std::unique_ptr<OperatorProjectorMI> p(
	    new OperatorProjectorMI(*m_ctx, projectorParams, withTOF));
	p->setLaunch(lanesPerEvent);
	return p;
}

// ---- motion: time-averaged (moved) images ---------------------------------
std::pair<int64_t, int64_t> getFullTimeRange(const std::vector<int64_t>& t)
{
	if (t.empty())
		throw Error("getFullTimeRange: no motion frames");
	if (t.size() == 1)
		return {t[0], t[0] + 1};
	const size_t last = t.size() - 1;
	return {t[0], t[last] + (t[last] - t[last - 1])};
}

int64_t motionFrameWeights(const std::vector<int64_t>& t, int64_t timeStart, int64_t timeStop,
                           std::vector<float>& w)
{
	const size_t k = t.size();
	if (k == 0)
		throw Error("timeAverageMoveImage: no motion frames");
	if (!(timeStop > timeStart))
		throw Error("Time stop must be greater than time start");
	// ReconstructionUtilsDevice.cu:146-166: the frames from the first start
	// >= timeStart up to (excluding) the first start > timeStop
	int64_t f0 = -1, f1 = (int64_t)k;
	for (size_t f = 0; f < k; f++)
	{
		if (t[f] >= timeStart)
		{
			if (f0 < 0)
				f0 = (int64_t)f;
			if (t[f] > timeStop)
			{
				f1 = (int64_t)f;
				break;
			}
		}
	}
	if (f0 < 0 || f1 <= f0)
		throw Error("timeAverageMoveImage: no motion frame starts inside the time range");
	const float scan = (float)(timeStop - timeStart);
	const int n = (int)(f1 - f0);
	w.assign(n, 0.f);
	for (int i = 0; i < n; i++)
	{
		// LORMotion::getDuration (LORMotion.cpp:73-88); one frame: the range
		const size_t f = (size_t)(f0 + i);
		if (k == 1)
		{
			w[i] = 1.f;
			continue;
		}
		const int64_t d = f + 1 < k ? t[f + 1] - t[f] : t[f] - t[f - 1];
		w[i] = (float)d / scan;
	}
	return f0;
}

void timeAverageMoveImageDevice(const std::vector<int64_t>& t, const std::vector<float>& xf,
                                const ImageDevice& unmoved, ImageDevice& out,
                                int64_t timeStart, int64_t timeStop, int64_t outFrame)
{
	const size_t k = t.size();
	if (xf.size() != 12 * k)
		throw Error("timeAverageMoveImage: need 12 transform values per motion frame");
	const ImageParams& p = unmoved.getParams();
	const ImageParams& q = out.getParams();
	if (p.nx != q.nx || p.ny != q.ny || p.nz != q.nz)
		throw Error("timeAverageMoveImage: input and output image sizes differ");
	if (outFrame < 0 || outFrame >= q.nt)
		throw Error("timeAverageMoveImage: output frame out of range");
	std::vector<float> w;
	const int64_t f0 = motionFrameWeights(t, timeStart, timeStop, w);
	const int n = (int)w.size();
	yrtmi_image d = p.desc(1);
	check(yrtmi_time_average_move(unmoved.context().get(), &d, unmoved.getDevicePointer(),
	                              out.getDevicePointer() + outFrame * (int64_t)q.numVoxels(),
	                              xf.data() + 12 * f0, w.data(), n),
	      "timeAverageMoveImage");
}

std::unique_ptr<ImageDevice> timeAverageMoveImageDevice(const std::vector<int64_t>& t,
                                                        const std::vector<float>& xf,
                                                        const ImageDevice& unmoved,
                                                        int64_t timeStart, int64_t timeStop)
{
	ImageParams p = unmoved.getParams();
	p.nt = 1;
	auto out = std::make_unique<ImageDevice>(unmoved.context(), p);
	out->setValue(0.f);
	timeAverageMoveImageDevice(t, xf, unmoved, *out, timeStart, timeStop, 0);
	return out;
}

std::unique_ptr<ImageDevice> timeAverageMoveImageDevice(const std::vector<int64_t>& t,
                                                        const std::vector<float>& xf,
                                                        const ImageDevice& unmoved)
{
	const auto r = getFullTimeRange(t);
	return timeAverageMoveImageDevice(t, xf, unmoved, r.first, r.second);
}

std::unique_ptr<ImageDevice> timeAverageMoveImageDynamicDevice(
    const std::vector<int64_t>& t, const std::vector<float>& xf, const ImageDevice& unmoved,
    const std::vector<uint32_t>& bounds)
{
	if (bounds.size() < 2)
		throw Error("timeAverageMoveImageDynamic: the framing needs at least one frame");
	ImageParams p = unmoved.getParams();
	p.nt = (int64_t)bounds.size() - 1;
	auto out = std::make_unique<ImageDevice>(unmoved.context(), p);
	out->setValue(0.f);
	for (int64_t f = 0; f < p.nt; f++)
		timeAverageMoveImageDevice(t, xf, unmoved, *out, bounds[f], bounds[f + 1], f);
	return out;
}

std::vector<float> gaussianKernel1D(float sigma, float voxSize, size_t size)
{
	if (!(sigma >= 0.f))
		throw Error("Gaussian Sigma or FWHM cannot be null or negative");
	if (!(voxSize >= 0.f))
		throw Error("Voxel size cannot be null or negative");
	if (size == 0)
		size = (size_t)std::rintf(5.f * sigma / voxSize);
	std::vector<float> k;
	for (size_t i = 0; i < size; i++)
	{
		const float x = (float)((int)i - (int)(size / 2));
		k.push_back(std::exp(-0.5f * std::pow(x * voxSize / sigma, 2.f)));
	}
	float sum = 0.f;
	for (float v : k)
		sum += v;
	for (float& v : k)
		v /= sum;
	return k;
}

void OSEM_MI::addImagePSF(const std::string& fname)
{
	readImagePsf(fname, m_psf[0], m_psf[1], m_psf[2]);
}

void OSEM_MI::addUniformGaussianImagePSFFromSigma(float sx, float sy, float sz, size_t kx,
                                                  size_t ky, size_t kz)
{
	if (!m_hasImageParams)
		throw Error("Image parameters not set");
	m_psf[0] = gaussianKernel1D(sx, m_imageParams.vx, kx);
	m_psf[1] = gaussianKernel1D(sy, m_imageParams.vy, ky);
	m_psf[2] = gaussianKernel1D(sz, m_imageParams.vz, kz);
}

void OSEM_MI::addUniformGaussianImagePSFFromFWHM(float fx, float fy, float fz, size_t kx,
                                                 size_t ky, size_t kz)
{
	// SIGMA_TO_FWHM = 2 sqrt(2 ln 2) (geometry/Constants.hpp)
	const float s2f = (float)(2.0 * std::sqrt(2.0 * std::log(2.0)));
	addUniformGaussianImagePSFFromSigma(fx / s2f, fy / s2f, fz / s2f, kx, ky, kz);
}

bool OSEM_MI::planned() const
{
	return useTilePlans;
}

ImageParams OSEM_MI::singleFrame() const
{
	ImageParams p = m_imageParams;
	p.nt = 1;
	return p;
}

void OSEM_MI::setSensitivityImages(const std::vector<const ImageDevice*>& imgs)
{
	m_sens.clear();
	for (const ImageDevice* s : imgs)
	{
		if (s == nullptr)
			throw Error("null sensitivity image");
		std::unique_ptr<ImageDevice> c(new ImageDevice(*m_ctx, s->getParams()));
		c->copyFrom(*s);
		m_sens.push_back(std::move(c));
	}
}

// y[i] *= x[i]: an EM update with a denominator of ones, y * (x / 1)
static void multiplyInto(const Context& ctx, const float* x, float* y, int64_t n)
{
	if (n <= 0)
		return;
	DeviceArray<float> ones(ctx, (size_t)n);
	check(yrtmi_fill(ctx.get(), ones.data(), n, 1.f), "fill");
	check(yrtmi_em_update(ctx.get(), y, x, ones.data(), n, 1, 1, 0.f), "product");
}

// Corrector::setup (Corrector.cpp:34-166), without changing the setters'
// values: which histogram or image gives the total, hardware and in-vivo
// attenuation.
OSEM_MI::Attenuation OSEM_MI::resolveAttenuation()
{
	Attenuation a{m_att, m_hwAtt, m_ivAtt, m_acf, m_hwAcf, m_ivAcf, false};
	if (!a.acf && !a.ivAcf && a.hwAcf)
		a.acf = a.hwAcf;  // all ACF is hardware ACF
	else if (!a.acf && a.ivAcf && !a.hwAcf)
		a.acf = a.ivAcf;  // all ACF is in-vivo ACF
	else if (a.acf && !a.ivAcf && !a.hwAcf)
		a.hwAcf = a.acf;  // a total alone: all hardware (no motion)
	if (a.inVivo && a.hardware && !a.total)
	{
		// the total is the sum of the two images
		if (!a.hardware->getParams().isSameAs(a.inVivo->getParams()))
			throw Error("Parameters mismatch between attenuation images");
		const ImageParams& p = a.hardware->getParams();
		m_impliedAtt.reset(new ImageDevice(*m_ctx, p));
		m_impliedAtt->copyFrom(*a.hardware);
		check(yrtmi_axpby(m_ctx->get(), (int64_t)(p.numVoxels() * (size_t)p.nt), 1.f,
		                  a.inVivo->getDevicePointer(), 1.f, m_impliedAtt->getDevicePointer()),
		      "implied attenuation");
		a.total = m_impliedAtt.get();
	}
	else if (!a.inVivo && a.hardware && !a.total)
		a.total = a.hardware;
	else if (a.inVivo && !a.hardware && !a.total)
		a.total = a.inVivo;
	else if (!a.inVivo && !a.hardware && a.total)
		a.hardware = a.total;
	a.totalFromHisto = a.acf != nullptr || (a.hwAcf != nullptr && a.ivAcf != nullptr);
	return a;
}

const std::vector<std::unique_ptr<ImageDevice>>& OSEM_MI::generateSensitivityImages()
{
	if (!m_hasImageParams)
		throw Error("Image parameters not set");
	const ImageParams im1 = singleFrame();
	const yrtmi_image ci = im1.desc(1);
	const int S = getExpectedSensImagesAmount();
	// sensitivity projector: TOF off (OSEM_CPU.cpp:94-108)
	auto proj = makeProjector(false);
	// hardware attenuation: its ACF histogram, else its image
	// (generateSensImage_kernel, OSEM_GPU.cu:1048-1110)
	const Attenuation at = resolveAttenuation();
	const ImageDevice* hw = at.hwAcf ? nullptr : at.hardware;
	yrtmi_image ca{};
	if (hw != nullptr)
		ca = hw->getParams().desc(1);
	m_sens.clear();
	for (int s = 0; s < S; s++)
	{
		std::unique_ptr<ImageDevice> img(new ImageDevice(*m_ctx, im1));
		yrtmi_lors l{};
		l.kind = YRTMI_SRC_HISTO_SUBSET;
		l.n = yrtmi_histo_subset_size(m_ctx->get(), S, s);
		if (l.n < 0)
			throw Error("sensitivity: the scanner has no histogram geometry");
		l.num_subsets = S;
		l.subset = s;
		std::unique_ptr<DeviceArray<float>> sensv;
		const float* sensPtr = nullptr;
		if (m_sensHisto != nullptr && l.n > 0)
		{
			// per-bin sensitivity factors of every bin of the subset
			// (Corrector::getSensitivity, Corrector.cpp:295-308)
			sensv.reset(new DeviceArray<float>(*m_ctx, (size_t)l.n));
			check(yrtmi_histo_subset_gather(m_ctx->get(), m_sensHisto->getDevicePointer(), S,
			                                s, m_invertSens ? 1 : 0, sensv->data()),
			      "sensitivity factors");
			sensPtr = sensv->data();
		}
		std::unique_ptr<DeviceArray<float>> acfv;
		if (at.hwAcf != nullptr && l.n > 0)
		{
			// sensitivity x hardware ACF per bin
			acfv.reset(new DeviceArray<float>(*m_ctx, (size_t)l.n));
			check(yrtmi_histo_subset_gather(m_ctx->get(), at.hwAcf->getDevicePointer(), S, s, 0,
			                                acfv->data()),
			      "hardware ACF");
			if (sensv)
				multiplyInto(*m_ctx, acfv->data(), sensv->data(), l.n);
			else
				sensPtr = acfv->data();
		}
		if (planned())
		{
			auto plan = makePlan(*m_ctx, *proj, ci, l, -1, 0, 0, true);
			check(yrtmi_sensitivity_planned(m_ctx->get(), proj->get(), &ci,
			                                img->getDevicePointer(), &l,
			                                hw ? &ca : nullptr,
			                                hw ? hw->getDevicePointer() : nullptr,
			                                sensPtr, plan->p),
			      "sensitivity");
		}
		else
		{
			check(yrtmi_sensitivity(m_ctx->get(), proj->get(), &ci, img->getDevicePointer(),
			                        &l, hw ? &ca : nullptr,
			                        hw ? hw->getDevicePointer() : nullptr, sensPtr),
			      "sensitivity");
		}
		if (hasImagePSF())
		{
			// PSF^T of the sensitivity image, zero where it was <= EPS
			// (OSEM_CPU.cpp:158-165)
			ImageDevice blurred(*m_ctx, im1), tmp(*m_ctx, im1);
			check(yrtmi_psf_apply(m_ctx->get(), &ci, img->getDevicePointer(),
			                      blurred.getDevicePointer(), tmp.getDevicePointer(),
			                      m_psf[0].data(), (int)m_psf[0].size(), m_psf[1].data(),
			                      (int)m_psf[1].size(), m_psf[2].data(), (int)m_psf[2].size(),
			                      1),
			      "image PSF");
			check(yrtmi_threshold(m_ctx->get(), blurred.getDevicePointer(),
			                      img->getDevicePointer(), (int64_t)im1.numVoxels(), 1, 1,
			                      EPS_FLT, 0.f, 0.f, 1.f, 0.f),
			      "threshold");
			img->copyFrom(blurred);
		}
		// hard threshold: <= thr -> 0, else * global scale
		// (applyThreshold, OSEM_CPU.cpp:168-172); the mask is a copy
		ImageDevice mask(*m_ctx, im1);
		mask.copyFrom(*img);
		check(yrtmi_threshold(m_ctx->get(), img->getDevicePointer(), mask.getDevicePointer(),
		                      (int64_t)im1.numVoxels(), 1, 1, hardThreshold, 0.f, 0.f,
		                      m_globalScale, 0.f),
		      "threshold");
		m_sens.push_back(std::move(img));
	}
	return m_sens;
}

void OSEM_MI::prepareSubsets(OperatorProjectorMI& proj, const ImageParams& imt, const float* x)
{
	m_subsets.clear();
	const int S = num_OSEM_subsets;
	const yrtmi_image ci = imt.desc();
	// total attenuation in the model, in-vivo attenuation as precorrection
	// (precomputeAttenuationFactors / precomputeInVivoAttenuationFactors,
	// Corrector_GPU.cu:19-116)
	const Attenuation at = resolveAttenuation();
	const ImageDevice* totImg = at.totalFromHisto ? nullptr : at.total;
	const ImageDevice* ivImg = at.ivAcf ? nullptr : at.inVivo;
	yrtmi_image ca{}, cv{};
	if (totImg != nullptr)
		ca = totImg->getParams().desc(1);
	if (ivImg != nullptr)
		cv = ivImg->getParams().desc(1);
	yrtmi_ctx* c = m_ctx->get();
	auto imageFactors = [&](std::unique_ptr<DeviceArray<float>>& dst, const ImageDevice* img,
	                        const yrtmi_image& d, const yrtmi_lors& lors, int64_t n) {
		if (img == nullptr || n == 0)
			return;
		dst.reset(new DeviceArray<float>(*m_ctx, (size_t)n));
		yrtmi_lors la = lors;
		la.tof = nullptr;
		la.frame = nullptr;
		check(yrtmi_attenuation_factors(c, &d, img->getDevicePointer(), &la, dst->data()),
		      "attenuation factors");
	};
	if (m_lm != nullptr)
	{
		// ACF histograms read at each event's detector pair
		auto gatherPairs = [&](std::unique_ptr<DeviceArray<float>>& dst,
		                       const Histogram3DDevice* h, const Subset& q, int64_t n) {
			if (h == nullptr || n == 0)
				return;
			dst.reset(new DeviceArray<float>(*m_ctx, (size_t)n));
			check(yrtmi_histo_gather_pairs(c, h->getDevicePointer(), q.d1->data(),
			                               q.d2->data(), n, dst->data()),
			      "ACF histogram");
		};
		if (!m_lm->motionTransforms.empty())
			m_ctx->uploadMotion(m_lm->motionTransforms);
		const int64_t N = (int64_t)m_lm->count();
		for (int s = 0; s < S; s++)
		{
			// BinIteratorChronologicalInterleaved: events s, s + S, ...
			std::unique_ptr<Subset> q(new Subset());
			const int64_t n = N > s ? (N - s + S - 1) / S : 0;
			auto take = [&](auto& dst, auto* src, int bytes) {
				using A = typename std::remove_reference<decltype(*dst)>::type;
				if (src == nullptr)
					return;
				dst.reset(new A(*m_ctx, (size_t)n));
				check(yrtmi_subset_interleaved(c, src, bytes, N, S, s, dst->data(), nullptr),
				      "list-mode subset");
			};
			take(q->d1, m_lm->d1(), 4);
			take(q->d2, m_lm->d2(), 4);
			take(q->tof, m_lm->tof(), 4);
			take(q->frame, m_lm->frame(), 4);
			take(q->motion, m_lm->motion(), 1);
			take(q->randoms, m_lm->randoms(), 4);
			q->lors.kind = YRTMI_SRC_PAIRS;
			q->lors.n = n;
			q->lors.d1 = q->d1 ? q->d1->data() : nullptr;
			q->lors.d2 = q->d2 ? q->d2->data() : nullptr;
			q->lors.num_subsets = 1;
			q->lors.tof = q->tof ? q->tof->data() : nullptr;
			q->lors.frame = q->frame ? q->frame->data() : nullptr;
			q->lors.motion_frame = q->motion ? q->motion->data() : nullptr;
			if (useLorOrder && n > 1)
			{
				// the subset's events in locality order (LORs projected together
				// read the same voxels); every per-event array follows
				DeviceArray<uint32_t> perm(*m_ctx, (size_t)n);
				check(yrtmi_lor_order(c, &ci, &q->lors, perm.data()), "locality order");
				auto reorder = [&](auto& arr, int bytes) {
					using A = typename std::remove_reference<decltype(*arr)>::type;
					if (!arr)
						return;
					std::unique_ptr<A> out(new A(*m_ctx, (size_t)n));
					check(yrtmi_permute(c, arr->data(), bytes, perm.data(), n, out->data()),
					      "permute");
					arr = std::move(out);
				};
				reorder(q->d1, 4);
				reorder(q->d2, 4);
				reorder(q->tof, 4);
				reorder(q->frame, 4);
				reorder(q->motion, 1);
				reorder(q->randoms, 4);
				q->lors.d1 = q->d1 ? q->d1->data() : nullptr;
				q->lors.d2 = q->d2 ? q->d2->data() : nullptr;
				q->lors.tof = q->tof ? q->tof->data() : nullptr;
				q->lors.frame = q->frame ? q->frame->data() : nullptr;
				q->lors.motion_frame = q->motion ? q->motion->data() : nullptr;
			}
			// per-event attenuation factors, once (precomputeAttenuationFactors)
			if (at.totalFromHisto && n > 0)
			{
				if (at.acf != nullptr)
					gatherPairs(q->att, at.acf, *q, n);
				else
				{
					std::unique_ptr<DeviceArray<float>> hwv;
					gatherPairs(q->att, at.ivAcf, *q, n);
					gatherPairs(hwv, at.hwAcf, *q, n);
					multiplyInto(*m_ctx, hwv->data(), q->att->data(), n);
				}
			}
			else
				imageFactors(q->att, totImg, ca, q->lors, n);
			if (at.ivAcf != nullptr)
				gatherPairs(q->pre, at.ivAcf, *q, n);
			else
				imageFactors(q->pre, ivImg, cv, q->lors, n);
			m_subsets.push_back(std::move(q));
		}
	}
	else
	{
		const int F = (int)m_histos.size();
		for (int s = 0; s < S; s++)
		{
			std::unique_ptr<Subset> q(new Subset());
			const int64_t cap = yrtmi_histo_subset_size(c, S, s);
			if (cap < 0)
				throw Error("histogram subset out of range");
			const size_t capA = (size_t)std::max<int64_t>(cap, 1);
			int64_t n = 0;
			// zero-count bins contribute nothing (OSEM_GPU.cu:1222-1225)
			if (F == 1)
			{
				q->bins.reset(new DeviceArray<uint32_t>(*m_ctx, capA));
				q->meas.reset(new DeviceArray<float>(*m_ctx, capA));
				check(yrtmi_histo_subset_nonzero(c, m_histos[0]->getDevicePointer(), S, s,
				                                 q->bins->data(), q->meas->data(), &n),
				      "nonzero bins");
			}
			else
			{
				// the subset's nonzero bins of every frame, frame after frame,
				// each with its frame index (one pass counts, one copies)
				DeviceArray<uint32_t> tb(*m_ctx, capA);
				DeviceArray<float> tm(*m_ctx, capA);
				std::vector<int64_t> nf(F, 0);
				for (int f = 0; f < F; f++)
				{
					check(yrtmi_histo_subset_nonzero(c, m_histos[f]->getDevicePointer(), S, s,
					                                 tb.data(), tm.data(), &nf[f]),
					      "nonzero bins");
					n += nf[f];
				}
				const size_t na = (size_t)std::max<int64_t>(n, 1);
				q->bins.reset(new DeviceArray<uint32_t>(*m_ctx, na));
				q->meas.reset(new DeviceArray<float>(*m_ctx, na));
				q->frame.reset(new DeviceArray<int32_t>(*m_ctx, na));
				int64_t off = 0;
				for (int f = 0; f < F; f++)
				{
					if (nf[f] == 0)
						continue;
					int64_t k = 0;
					check(yrtmi_histo_subset_nonzero(c, m_histos[f]->getDevicePointer(), S, s,
					                                 tb.data(), tm.data(), &k),
					      "nonzero bins");
					check(yrtmi_copy(c, q->bins->data() + off, tb.data(), 4 * k, 2), "copy");
					check(yrtmi_copy(c, q->meas->data() + off, tm.data(), 4 * k, 2), "copy");
					check(yrtmi_fill_i32(c, q->frame->data() + off, k, f), "frame indices");
					off += k;
				}
			}
			q->lors.kind = YRTMI_SRC_BINS;
			q->lors.n = n;
			q->lors.bins = q->bins->data();
			q->lors.num_subsets = 1;
			q->lors.frame = q->frame ? q->frame->data() : nullptr;
			if (useLorOrder && n > 1)
			{
				// the bins in locality order (frame first); the per-bin
				// corrections gathered below follow
				DeviceArray<uint32_t> perm(*m_ctx, (size_t)n);
				check(yrtmi_lor_order(c, &ci, &q->lors, perm.data()), "locality order");
				auto reorder = [&](auto& arr) {
					using A = typename std::remove_reference<decltype(*arr)>::type;
					if (!arr)
						return;
					std::unique_ptr<A> out(new A(*m_ctx, (size_t)n));
					check(yrtmi_permute(c, arr->data(), 4, perm.data(), n, out->data()),
					      "permute");
					arr = std::move(out);
				};
				reorder(q->bins);
				reorder(q->meas);
				reorder(q->frame);
				q->lors.bins = q->bins->data();
				q->lors.frame = q->frame ? q->frame->data() : nullptr;
			}
			auto gatherHisto = [&](std::unique_ptr<DeviceArray<float>>& dst,
			                       const Histogram3DDevice* h, bool invert) {
				if (h == nullptr || n == 0)
					return;
				dst.reset(new DeviceArray<float>(*m_ctx, (size_t)n));
				check(yrtmi_gather(c, h->getDevicePointer(), (int64_t)h->count(),
				                   q->bins->data(), n, invert ? 1 : 0, dst->data()),
				      "per-bin correction");
			};
			gatherHisto(q->sens, m_sensHisto, m_invertSens);
			gatherHisto(q->randoms, m_randomsHisto, false);
			gatherHisto(q->scatter, m_scatterHisto, false);
			if (at.totalFromHisto && n > 0)
			{
				// Corrector::getTotalACFFromHistogram: the total ACF, or
				// in-vivo x hardware (Corrector.cpp:585-599)
				if (at.acf != nullptr)
					gatherHisto(q->att, at.acf, false);
				else
				{
					std::unique_ptr<DeviceArray<float>> hwv;
					gatherHisto(q->att, at.ivAcf, false);
					gatherHisto(hwv, at.hwAcf, false);
					multiplyInto(*m_ctx, hwv->data(), q->att->data(), n);
				}
			}
			else
				imageFactors(q->att, totImg, ca, q->lors, n);
			if (at.ivAcf != nullptr)
				gatherHisto(q->pre, at.ivAcf, false);
			else
				imageFactors(q->pre, ivImg, cv, q->lors, n);
			m_subsets.push_back(std::move(q));
		}
	}
	m_numPlanned = 0;
	if (!planned())
		return;
	// launch shapes timed on the first nonempty subset
	m_tileShape = -1;
	m_persistent = 0;
	if (useAutotune && x != nullptr)
		for (auto& q : m_subsets)
			if (q->lors.n > 0)
			{
				yrtmi_tuning t{};
				check(yrtmi_autotune(c, proj.get(), &ci, &q->lors, x, &t), "autotune");
				m_tileShape = t.tile_shape;
				m_persistent = t.persistent;
				break;
			}
	// the plans within the memory budget, in subset order
	int64_t budget = (int64_t)memoryBudget;
	if (budget <= 0)
	{
		int64_t freeB = 0;
		check(yrtmi_mem_info(c, &freeB, nullptr), "memory info");
		budget = (int64_t)(0.9 * (double)freeB);
	}
	int64_t used = 0;
	for (auto& q : m_subsets)
	{
		if (q->lors.n == 0 || used >= budget)
			continue;
		q->plan = makePlan(*m_ctx, proj, ci, q->lors, m_tileShape, m_persistent, budget - used);
		if (q->plan)
		{
			used += planBytes(*q->plan);
			m_numPlanned++;
		}
	}
}

std::unique_ptr<ImageDevice> OSEM_MI::reconstruct()
{
	if (!m_hasImageParams)
		throw Error("Image parameters not set");
	if (m_lm == nullptr && m_histos.empty())
		throw Error("Data input not set");
	const int S = num_OSEM_subsets;
	if (S <= 0 || num_MLEM_iterations < 0)
		throw Error("bad iteration / subset counts");
	if ((int)m_sens.size() != getExpectedSensImagesAmount())
		throw Error("The number of sensitivity images provided does not match the number "
		            "of subsets. Expected " +
		            std::to_string(getExpectedSensImagesAmount()) + " but received " +
		            std::to_string(m_sens.size()));
	yrtmi_ctx* c = m_ctx->get();
	// setupForDynamicRecon (OSEM.cpp:823-830)
	ImageParams imt = m_imageParams;
	imt.nt = m_lm != nullptr ? m_lm->numFrames : (int64_t)m_histos.size();
	const int64_t nvox = (int64_t)imt.numVoxels();
	const ImageParams im1 = singleFrame();
	// list-mode sensitivity x 1/S on a copy (OSEM.cpp:1007-1035)
	std::vector<std::unique_ptr<ImageDevice>> sens;
	for (auto& s : m_sens)
	{
		std::unique_ptr<ImageDevice> cpy(new ImageDevice(*m_ctx, im1));
		cpy->copyFrom(*s);
		if (isListModeEnabled() && S != 1)
			check(yrtmi_axpby(c, nvox, 0.f, nullptr, 1.0f / (float)S, cpy->getDevicePointer()),
			      "sensitivity scale");
		sens.push_back(std::move(cpy));
	}
	std::unique_ptr<ImageDevice> x(new ImageDevice(*m_ctx, imt));
	if (m_initial != nullptr)
		x->copyFrom(*m_initial);
	else
		x->setValue(INITIAL_VALUE_MLEM);
	// mask (OSEM_CPU.cpp:241-269): the mask image, the sensitivity image, or
	// the sum of the subsets' sensitivity images
	ImageDevice mask(*m_ctx, im1);
	if (m_mask != nullptr)
		mask.copyFrom(*m_mask);
	else if (S == 1 || isListModeEnabled())
		mask.copyFrom(*sens[0]);
	else
	{
		// the sum of the subsets' sensitivity images, in subset order
		for (auto& si : sens)
			check(yrtmi_axpby(c, nvox, 1.f, si->getDevicePointer(), 1.f,
			                  mask.getDevicePointer()),
			      "mask sum");
	}
	check(yrtmi_threshold(c, x->getDevicePointer(), mask.getDevicePointer(), nvox, imt.nt, 1,
	                      0.f, 0.f, 0.f, 1.f, 0.f),
	      "mask");
	auto proj = makeProjector(true);
	prepareSubsets(*proj, imt, x->getDevicePointer());
	ImageDevice u(*m_ctx, imt);
	const yrtmi_image ci = imt.desc();
	// image-space PSF buffers: PSF(x), the blurred update image, scratch
	std::unique_ptr<ImageDevice> xPsf, uPsf, psfTmp;
	if (hasImagePSF())
	{
		xPsf.reset(new ImageDevice(*m_ctx, imt));
		uPsf.reset(new ImageDevice(*m_ctx, imt));
		psfTmp.reset(new ImageDevice(*m_ctx, imt));
	}
	auto psf = [&](const ImageDevice& in, ImageDevice& out, int transpose) {
		check(yrtmi_psf_apply(c, &ci, in.getDevicePointer(), out.getDevicePointer(),
		                      psfTmp->getDevicePointer(), m_psf[0].data(), (int)m_psf[0].size(),
		                      m_psf[1].data(), (int)m_psf[1].size(), m_psf[2].data(),
		                      (int)m_psf[2].size(), transpose),
		      "image PSF");
	};
	// one OSEM iteration: every subset's EM update image and image update
	auto iteration = [&]() {
		for (int s = 0; s < S; s++)
		{
			Subset& q = *m_subsets[s];
			u.setValue(0.f);
			// the forward projection reads PSF(x) (OSEM_CPU.cpp:284-295)
			const ImageDevice* xFwd = x.get();
			if (hasImagePSF())
			{
				psf(*x, *xPsf, 0);
				xFwd = xPsf.get();
			}
			yrtmi_corrections cr{};
			cr.meas = q.meas ? q.meas->data() : nullptr;
			cr.sens = q.sens ? q.sens->data() : nullptr;
			cr.att = q.att ? q.att->data() : nullptr;
			cr.randoms = q.randoms ? q.randoms->data() : nullptr;
			cr.scatter = q.scatter ? q.scatter->data() : nullptr;
			cr.precorr = q.pre ? q.pre->data() : nullptr;
			cr.global_scale = m_globalScale;
			cr.denom_threshold = denomThreshold;
			if (q.lors.n > 0)
			{
				if (q.plan)
					check(yrtmi_em_subset_planned(c, proj->get(), &ci, xFwd->getDevicePointer(),
					                              u.getDevicePointer(), &q.lors, &cr, q.plan->p),
					      "EM subset");
				else
					check(yrtmi_em_subset(c, proj->get(), &ci, xFwd->getDevicePointer(),
					                      u.getDevicePointer(), &q.lors, &cr),
					      "EM subset");
			}
			// the update image is PSF^T of the backprojection
			// (OSEM_CPU.cpp:377-392)
			const ImageDevice* uImg = &u;
			if (hasImagePSF())
			{
				psf(u, *uPsf, 1);
				uImg = uPsf.get();
			}
			const ImageDevice& sImg = *sens[isListModeEnabled() ? 0 : s];
			check(yrtmi_em_update(c, x->getDevicePointer(), uImg->getDevicePointer(),
			                      sImg.getDevicePointer(), nvox, imt.nt, 1, EPS_FLT),
			      "EM update");
		}
	};
	// the first iteration call by call (it also makes any per-XCD partial
	// images of unplanned subsets), the next recorded once as a graph and
	// replayed for every later one
	m_iterSeconds.clear();
	auto clock = std::chrono::steady_clock::now();
	auto lap = [&]() {
		if (!timeIterations)
			return;
		m_ctx->synchronize();
		const auto now = std::chrono::steady_clock::now();
		m_iterSeconds.push_back(std::chrono::duration<double>(now - clock).count());
		clock = now;
	};
	if (timeIterations)
	{
		m_ctx->synchronize();
		clock = std::chrono::steady_clock::now();
	}
	int it = 0;
	if (num_MLEM_iterations > 0)
	{
		iteration();
		it = 1;
		lap();
	}
	if (useGraphs && num_MLEM_iterations > 1)
	{
		m_ctx->ensureOwnStream();
		yrtmi_graph* g = nullptr;
		check(yrtmi_graph_begin(c), "graph begin");
		try
		{
			iteration();
		}
		catch (...)
		{
			yrtmi_graph* partial = nullptr;
			if (yrtmi_graph_end(c, &partial) == YRTMI_OK)
				yrtmi_graph_destroy(partial);
			throw;
		}
		check(yrtmi_graph_end(c, &g), "graph end");
		std::unique_ptr<yrtmi_graph, int (*)(yrtmi_graph*)> gh(g, yrtmi_graph_destroy);
		for (; it < num_MLEM_iterations; it++)
		{
			check(yrtmi_graph_launch(c, g), "graph launch");
			lap();
		}
		m_ctx->synchronize();
	}
	for (; it < num_MLEM_iterations; it++)
	{
		iteration();
		lap();
	}
	m_ctx->synchronize();
	return x;
}

std::unique_ptr<OSEM_MI> createOSEM(const Scanner& scanner, bool useGPU, int device)
{
	if (!useGPU)
		throw Error("yrtmi provides the GPU OSEM only");
	return std::unique_ptr<OSEM_MI>(new OSEM_MI(scanner, device));
}

// ---------------------------------------------------------------------------
// util::forwProject / backProject
// ---------------------------------------------------------------------------
void forwProject(const Context& ctx, const ProjectorParams& pp, const ImageDevice& img,
                 ListModeDevice& proj)
{
	OperatorProjectorMI op(ctx, pp);
	op.applyA(img, proj);
	ctx.synchronize();
}

void backProject(const Context& ctx, const ProjectorParams& pp, ImageDevice& img,
                 const ListModeDevice& proj)
{
	OperatorProjectorMI op(ctx, pp);
	op.applyAH(proj, img);
	ctx.synchronize();
}

void forwProject(const Context& ctx, const ProjectorParams& pp, const ImageDevice& img,
                 Histogram3DDevice& proj)
{
	OperatorProjectorMI op(ctx, pp);
	op.applyA(img, proj);
	ctx.synchronize();
}

void backProject(const Context& ctx, const ProjectorParams& pp, ImageDevice& img,
                 const Histogram3DDevice& proj)
{
	OperatorProjectorMI op(ctx, pp);
	op.applyAH(proj, img);
	ctx.synchronize();
}

}  // namespace yrtmi

// yrtmi C ABI: contexts, resident scanner tables, projectors, tile plans and
// the entry points of include/yrtmi.h.  Host code only; kernels live in
// yrtmi_kernels.hip and yrtmi_tiles.hip.
#include "yrtmi.h"
#include "yrtmi_kernels.hpp"
#include "yrtmi_tiles.hpp"
#include "yrtmi_f64.hpp"
#include "yrtmi_order.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace yrtmi;

// The scratch array of yrtmi_em_subset_noalloc (its per-XCD partial
// images): 16 Mi floats (64 MiB) on each device the library runs on, held
// until the process ends.
constexpr int64_t kScratchFloats = int64_t(16) << 20;
static __device__ __align__(256) float scratch[kScratchFloats];

namespace
{
thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
	g_err = msg;
	return code;
}

int hipFail(hipError_t e, const char* what)
{
	return fail(YRTMI_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
	do                                                  \
	{                                                   \
		hipError_t e_ = (expr);                         \
		if (e_ != hipSuccess)                           \
			return hipFail(e_, #expr);                  \
	} while (0)

}  // namespace

struct yrtmi_ctx
{
	int device = 0;
	hipStream_t stream = nullptr;
	float* lut = nullptr;
	int64_t ndet = 0;
	bool histoReady = false;
	yrtmi_histo_desc histo{};
	uint32_t numR = 0, numPhi = 0, numZBin = 0, numZBinDiff = 0, rCut = 0;
	uint32_t* zbin = nullptr;
	int32_t* ringMap = nullptr;  // in-ring pair -> rr | phi << 16 (dpr x dpr)
	int* gatherErr = nullptr;
	float* motion = nullptr;
	int nmotion = 0;
	// frames of yrtmi_time_average_move (4 float4 each), grown on demand
	float4* moveFrames = nullptr;
	int moveCap = 0;
	unsigned long long* counter = nullptr;
	// per-XCD partial images for backprojection (nxcc x partStride floats),
	// kept zeroed between uses
	float* part = nullptr;
	int64_t partStride = 0;
	int nxcc = 1;
	// bricked copy of the image the Siddon EM ratio projects, grown on demand
	float* brick = nullptr;
	int64_t brickCap = 0;
	hipStream_t owned = nullptr;  // stream created by yrtmi_ctx_own_stream
	bool capturing = false;       // between yrtmi_graph_begin and _end
};

struct yrtmi_graph
{
	int device = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

struct yrtmi_proj
{
	// `ctx` is only compared, never dereferenced after creation: a projector
	// may outlive its context (e.g. finalizers running in any order).
	yrtmi_ctx* ctx = nullptr;
	int device = 0;
	DProj d{};
	float* psf = nullptr;  // numS padded rows of taps, then the flipped copy's
	double* psfCum = nullptr;  // numS*(K+2) trapezoid prefix sums, then the flipped copy's
	int lpe = 1;
	int block = 256;
	int64_t grid = 0;
	int accum = YRTMI_ACCUM_XCD;
};

namespace
{

int setDev(yrtmi_ctx* c)
{
	if (c == nullptr)
		return fail(YRTMI_EINVAL, "null context");
	HIP_TRY(hipSetDevice(c->device));
	return YRTMI_OK;
}

// For calls that allocate, copy or synchronise: refused while recording.
int notCapturing(yrtmi_ctx* c, const char* what)
{
	if (c != nullptr && c->capturing)
		return fail(YRTMI_ESTATE, std::string(what) + ": not allowed while recording a graph");
	return YRTMI_OK;
}

int checkImage(const yrtmi_image* im, const char* what)
{
	if (im == nullptr)
		return fail(YRTMI_EINVAL, std::string(what) + ": null image descriptor");
	if (im->nx <= 0 || im->ny <= 0 || im->nz <= 0 || im->nt <= 0)
		return fail(YRTMI_EINVAL, std::string(what) + ": image dimensions must be positive");
	if (im->nx * im->ny * im->nz * im->nt >= (int64_t(1) << 31))
		return fail(YRTMI_EINVAL, std::string(what) + ": image larger than 2^31 voxels");
	if (!(im->vx > 0.f && im->vy > 0.f && im->vz > 0.f))
		return fail(YRTMI_EINVAL, std::string(what) + ": voxel sizes must be positive");
	return YRTMI_OK;
}

DImage toDImage(const yrtmi_image* im)
{
	DImage d{};
	d.nx = (int32_t)im->nx;
	d.ny = (int32_t)im->ny;
	d.nz = (int32_t)im->nz;
	d.nt = (int32_t)im->nt;
	d.vx = im->vx;
	d.vy = im->vy;
	d.vz = im->vz;
	d.lx = im->lx;
	d.ly = im->ly;
	d.lz = im->lz;
	d.ox = im->ox;
	d.oy = im->oy;
	d.oz = im->oz;
	d.fovR = im->fov_radius;
	d.nxy = d.nx * d.ny;
	d.nvox = (int64_t)d.nxy * d.nz;
	return d;
}

int buildLors(yrtmi_ctx* c, const yrtmi_lors* l, DLors& d, DScanner& s)
{
	if (l == nullptr)
		return fail(YRTMI_EINVAL, "null LOR source");
	if (c->lut == nullptr)
		return fail(YRTMI_ESTATE, "no detector LUT uploaded (yrtmi_lut_upload)");
	if (l->n < 0)
		return fail(YRTMI_EINVAL, "negative LOR count");
	d = DLors{};
	d.kind = l->kind;
	d.n = l->n;
	d.tof = l->tof;
	d.frame = l->frame;
	d.motion = l->motion_frame;
	if (l->motion_frame != nullptr && c->motion == nullptr)
		return fail(YRTMI_ESTATE, "motion frames given but no transforms uploaded");
	switch (l->kind)
	{
	case YRTMI_SRC_PAIRS:
		if (l->n > 0 && (l->d1 == nullptr || l->d2 == nullptr))
			return fail(YRTMI_EINVAL, "pair source needs d1 and d2");
		d.d1 = l->d1;
		d.d2 = l->d2;
		break;
	case YRTMI_SRC_BINS:
		if (!c->histoReady)
			return fail(YRTMI_ESTATE, "histogram source without yrtmi_histo_setup");
		if (l->n > 0 && l->bins == nullptr)
			return fail(YRTMI_EINVAL, "bin source needs bins");
		d.bins = l->bins;
		break;
	case YRTMI_SRC_HISTO_SUBSET:
	{
		if (!c->histoReady)
			return fail(YRTMI_ESTATE, "histogram source without yrtmi_histo_setup");
		if (l->num_subsets <= 0 || l->subset < 0 || l->subset >= l->num_subsets)
			return fail(YRTMI_EINVAL, "bad subset index");
		const int64_t sz = yrtmi_histo_subset_size(c, l->num_subsets, l->subset);
		if (l->n > sz)
			return fail(YRTMI_EINVAL, "LOR count exceeds the subset size");
		uint32_t nps = c->numPhi / l->num_subsets;
		if ((uint32_t)l->subset + nps * (uint32_t)l->num_subsets < c->numPhi)
			nps += 1;
		d.S = (uint32_t)l->num_subsets;
		d.s = (uint32_t)l->subset;
		d.nps = nps;
		break;
	}
	default:
		return fail(YRTMI_EINVAL, "unknown LOR source kind");
	}
	s = DScanner{};
	s.lut = c->lut;
	s.ndet = (uint32_t)c->ndet;
	s.numR = c->numR;
	s.numPhi = c->numPhi;
	s.numZBin = c->numZBin;
	s.rCut = c->rCut;
	s.numDOI = (uint32_t)(c->histoReady ? c->histo.num_doi : 1);
	s.numDOIPoss = s.numDOI * s.numDOI;
	s.dpr = (uint32_t)c->histo.dets_per_ring;
	s.nrings = (uint32_t)c->histo.num_rings;
	s.zbin = c->zbin;
	s.motion = c->motion;
	s.nmotion = c->nmotion;
	return YRTMI_OK;
}

// Scratch of yrtmi_em_subset_noalloc: its per-XCD partial images, in a
// file-scope device array (64 MiB on every device the library runs on, held
// until the process ends) instead of memory the library allocates.  Host
// code never uses the array's name (on the host it is a stand-in, not a
// device address): yrtmi_ctx_create takes each device's address of it once
// with hipGetSymbolAddress, with that device current, and the entry point
// uses that address.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMutex;
float* g_scratchPtr[kMaxScratchDevices] = {};

// this device's address of `scratch` (the device is current)
int scratchSetup(int device)
{
	if (device < 0 || device >= kMaxScratchDevices)
		return fail(YRTMI_EINVAL, "device index beyond the scratch table");
	std::lock_guard<std::mutex> lk(g_scratchMutex);
	if (g_scratchPtr[device] != nullptr)
		return YRTMI_OK;
	float* ptr = nullptr;
	const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
	if (e != hipSuccess || ptr == nullptr)
		return hipFail(e != hipSuccess ? e : hipErrorInvalidSymbol,
		               "scratch: hipGetSymbolAddress");
	g_scratchPtr[device] = ptr;
	return YRTMI_OK;
}

float* scratchOf(int device)
{
	std::lock_guard<std::mutex> lk(g_scratchMutex);
	return (device >= 0 && device < kMaxScratchDevices) ? g_scratchPtr[device] : nullptr;
}

// Make sure the context holds nxcc zeroed partial images of at least n
// floats each.  Allocates only when a larger image than before is seen.
int ensurePartials(yrtmi_ctx* c, int64_t n)
{
	const int64_t stride = (n + 63) / 64 * 64;
	if (c->part != nullptr && c->partStride >= stride)
		return YRTMI_OK;
	int r = notCapturing(c, "per-XCD partial images (run one backprojection before recording)");
	if (r)
		return r;
	HIP_TRY(hipStreamSynchronize(c->stream));
	HIP_TRY(hipFree(c->part));
	c->part = nullptr;
	c->partStride = 0;
	HIP_TRY(hipMalloc(&c->part, sizeof(float) * stride * c->nxcc));
	HIP_TRY(hipMemset(c->part, 0, sizeof(float) * stride * c->nxcc));
	c->partStride = stride;
	return YRTMI_OK;
}

// The buffer of the bricked image the Siddon EM ratio's forward reads, at
// least `need` floats (grown outside graph recording only; tile plans size
// it for their image, so planned EM calls never allocate).
int ensureBrick(yrtmi_ctx* c, int64_t need)
{
	if (c->brick == nullptr || c->brickCap < need)
	{
		int r = notCapturing(c, "bricked image (run one EM ratio of this size before recording)");
		if (r)
			return r;
		HIP_TRY(hipStreamSynchronize(c->stream));
		HIP_TRY(hipFree(c->brick));
		c->brick = nullptr;
		c->brickCap = 0;
		HIP_TRY(hipMalloc(&c->brick, sizeof(float) * need));
		c->brickCap = need;
	}
	return YRTMI_OK;
}

// The bricked copy of a.x for the Siddon EM ratio's forward; DD reads a.x
// itself.
int brickImage(yrtmi_ctx* c, KArgs& a)
{
	if (a.pr.type == 1)
		return YRTMI_OK;
	const int64_t bvox = brickFrameVoxels(a.im.nx, a.im.ny, a.im.nz);
	const int r = ensureBrick(c, bvox * a.im.nt);
	if (r)
		return r;
	HIP_TRY(launchToBricks(a.im, a.x, c->brick, c->stream));
	a.xb = c->brick;
	a.bvox = bvox;
	return YRTMI_OK;
}

// Set up per-XCD accumulation for a backprojection of n voxels into y, or
// leave the kernel writing y directly.
int setupAccum(yrtmi_ctx* c, const yrtmi_proj* p, KArgs& a, int64_t n)
{
	a.part = nullptr;
	if (p->accum != YRTMI_ACCUM_XCD || c->nxcc <= 1)
		return YRTMI_OK;
	int r = ensurePartials(c, n);
	if (r)
		return r;
	a.part = c->part;
	a.partStride = c->partStride;
	a.nxcc = c->nxcc;
	return YRTMI_OK;
}

int foldAccum(yrtmi_ctx* c, const KArgs& a, float* y, int64_t n)
{
	if (a.part == nullptr)
		return YRTMI_OK;
	HIP_TRY(launchFoldPartials(y, a.part, a.partStride, a.nxcc, n, c->stream));
	return YRTMI_OK;
}

}  // namespace

extern "C" {

const char* yrtmi_last_error(void)
{
	return g_err.c_str();
}

int yrtmi_ctx_create(int device, void* stream, yrtmi_ctx** out)
{
	if (out == nullptr)
		return fail(YRTMI_EINVAL, "null output pointer");
	int n = 0;
	HIP_TRY(hipGetDeviceCount(&n));
	if (device < 0 || device >= n)
		return fail(YRTMI_EINVAL, "device index out of range");
	auto* c = new yrtmi_ctx();
	c->device = device;
	c->stream = (hipStream_t)stream;
	hipError_t e = hipSetDevice(device);
	if (e == hipSuccess)
		e = hipMalloc(&c->counter, sizeof(unsigned long long));
	if (e == hipSuccess && scratchSetup(device) != YRTMI_OK)
	{
		(void)hipFree(c->counter);
		delete c;
		return YRTMI_EHIP;
	}
	if (e == hipSuccess)
	{
		int nx = 1;
		if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, device) ==
		        hipSuccess &&
		    nx >= 1 && nx <= 16)
			c->nxcc = nx;
	}
	if (e != hipSuccess)
	{
		delete c;
		return hipFail(e, "yrtmi_ctx_create");
	}
	*out = c;
	return YRTMI_OK;
}

int yrtmi_ctx_destroy(yrtmi_ctx* c)
{
	if (c == nullptr)
		return YRTMI_OK;
	(void)hipSetDevice(c->device);
	(void)hipStreamSynchronize(c->stream);
	(void)hipFree(c->lut);
	(void)hipFree(c->zbin);
	(void)hipFree(c->motion);
	(void)hipFree(c->moveFrames);
	(void)hipFree(c->ringMap);
	(void)hipFree(c->gatherErr);
	(void)hipFree(c->counter);
	(void)hipFree(c->part);
	(void)hipFree(c->brick);
	if (c->owned != nullptr)
		(void)hipStreamDestroy(c->owned);
	delete c;
	return YRTMI_OK;
}

int yrtmi_ctx_own_stream(yrtmi_ctx* c, void** stream)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "own_stream");
	if (r)
		return r;
	if (c->owned == nullptr)
		HIP_TRY(hipStreamCreateWithFlags(&c->owned, hipStreamNonBlocking));
	c->stream = c->owned;
	if (stream != nullptr)
		*stream = (void*)c->owned;
	return YRTMI_OK;
}

int yrtmi_graph_begin(yrtmi_ctx* c)
{
	int r = setDev(c);
	if (r)
		return r;
	if (c->capturing)
		return fail(YRTMI_ESTATE, "graph_begin: already recording");
	if (c->stream == nullptr)
		return fail(YRTMI_ESTATE,
		            "graph_begin: the context uses the null stream (yrtmi_ctx_own_stream)");
	HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
	c->capturing = true;
	return YRTMI_OK;
}

int yrtmi_graph_end(yrtmi_ctx* c, yrtmi_graph** out)
{
	int r = setDev(c);
	if (r)
		return r;
	if (!c->capturing)
		return fail(YRTMI_ESTATE, "graph_end: not recording");
	c->capturing = false;
	hipGraph_t g = nullptr;
	HIP_TRY(hipStreamEndCapture(c->stream, &g));
	if (out == nullptr)
	{
		(void)hipGraphDestroy(g);
		return fail(YRTMI_EINVAL, "graph_end: null output");
	}
	hipGraphExec_t x = nullptr;
	hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
	if (e != hipSuccess)
	{
		(void)hipGraphDestroy(g);
		return hipFail(e, "graph_end: instantiate");
	}
	auto* q = new yrtmi_graph();
	q->device = c->device;
	q->graph = g;
	q->exec = x;
	*out = q;
	return YRTMI_OK;
}

int yrtmi_graph_launch(yrtmi_ctx* c, const yrtmi_graph* g)
{
	int r = setDev(c);
	if (r)
		return r;
	if (g == nullptr || g->device != c->device)
		return fail(YRTMI_EINVAL, "graph_launch: null graph or another device");
	r = notCapturing(c, "graph_launch");
	if (r)
		return r;
	HIP_TRY(hipGraphLaunch(g->exec, c->stream));
	return YRTMI_OK;
}

int yrtmi_graph_destroy(yrtmi_graph* g)
{
	if (g == nullptr)
		return YRTMI_OK;
	(void)hipSetDevice(g->device);
	(void)hipGraphExecDestroy(g->exec);
	(void)hipGraphDestroy(g->graph);
	delete g;
	return YRTMI_OK;
}

int yrtmi_ctx_set_stream(yrtmi_ctx* c, void* stream)
{
	if (c == nullptr)
		return fail(YRTMI_EINVAL, "null context");
	int r = notCapturing(c, "ctx_set_stream");
	if (r)
		return r;
	c->stream = (hipStream_t)stream;
	return YRTMI_OK;
}

int yrtmi_ctx_synchronize(yrtmi_ctx* c)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "ctx_synchronize");
	if (r)
		return r;
	HIP_TRY(hipStreamSynchronize(c->stream));
	return YRTMI_OK;
}

int yrtmi_lut_upload(yrtmi_ctx* c, const float* pos, const float* orient,
                     int64_t n)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "lut_upload");
	if (r)
		return r;
	if (pos == nullptr || orient == nullptr || n <= 0 || n >= (int64_t(1) << 31))
		return fail(YRTMI_EINVAL, "bad LUT arguments");
	std::vector<float> soa(6 * n);
	for (int64_t i = 0; i < n; i++)
	{
		for (int k = 0; k < 3; k++)
		{
			soa[k * n + i] = pos[3 * i + k];
			soa[(3 + k) * n + i] = orient[3 * i + k];
		}
	}
	HIP_TRY(hipStreamSynchronize(c->stream));
	HIP_TRY(hipFree(c->lut));
	c->lut = nullptr;
	HIP_TRY(hipMalloc(&c->lut, soa.size() * sizeof(float)));
	HIP_TRY(hipMemcpy(c->lut, soa.data(), soa.size() * sizeof(float),
	                  hipMemcpyHostToDevice));
	c->ndet = n;
	return YRTMI_OK;
}

// DetRegular::generateLUT (src/datastruct/scanner/DetRegular.cpp:113-182):
// crystal centres at radius R + depth*(doi + 1/2), z = -FOV/2 + (ring+1/2)*gap,
// orientation (cos, sin, 0) of the block angle.
int yrtmi_lut_regular(const yrtmi_scanner_desc* sc, float* pos, float* orient)
{
	if (sc == nullptr || pos == nullptr || orient == nullptr ||
	    sc->dets_per_ring <= 0 || sc->num_rings <= 0 || sc->num_doi <= 0 ||
	    sc->dets_per_block <= 0 || sc->dets_per_ring % sc->dets_per_block != 0)
		return fail(YRTMI_EINVAL, "bad scanner descriptor");
	const int64_t numBlocks = sc->dets_per_ring / sc->dets_per_block;
	const float blockLength = sc->crystal_size_trans * sc->dets_per_block;
	const double edgeMax = 2 * sc->radius * std::tan(3.14159265358979 / numBlocks);
	if (blockLength > edgeMax)
		return fail(YRTMI_EINVAL, "overlap between crystals: scanner radius too small");
	for (int64_t doi = 0; doi < sc->num_doi; doi++)
	{
		const float blockDistance = sc->radius + sc->crystal_depth * doi;
		for (int64_t ring = 0; ring < sc->num_rings; ring++)
		{
			const float zGap = sc->axial_fov / ((float)sc->num_rings);
			const float zPos = (float)(-sc->axial_fov / 2.0 + (ring + 0.5f) * zGap);
			for (int64_t block = 0; block < numBlocks; block++)
			{
				const float ang = (float)((double)block * 6.28318530717959 /
				                          (double)numBlocks);
				const float xB = (blockDistance + 0.5f * sc->crystal_depth) * std::cos(ang);
				const float yB = (blockDistance + 0.5f * sc->crystal_depth) * std::sin(ang);
				for (int64_t det = 0; det < sc->dets_per_block; det++)
				{
					float rel = 0.0f;
					if (sc->dets_per_block != 1)
						rel = (float)(-0.5 + (det + 0.5) / (float)sc->dets_per_block);
					const float rx = (float)(rel * std::cos(ang + 3.14159265358979 / 2.0));
					const float ry = (float)(rel * std::sin(ang + 3.14159265358979 / 2.0));
					const int64_t idx = det + block * sc->dets_per_block +
					                    ring * sc->dets_per_ring +
					                    doi * (sc->num_rings * sc->dets_per_ring);
					pos[3 * idx] = xB + blockLength * rx;
					pos[3 * idx + 1] = yB + blockLength * ry;
					pos[3 * idx + 2] = zPos;
					orient[3 * idx] = std::cos(ang);
					orient[3 * idx + 1] = std::sin(ang);
					orient[3 * idx + 2] = 0.0f;
				}
			}
		}
	}
	return YRTMI_OK;
}

int yrtmi_histo_setup(yrtmi_ctx* c, const yrtmi_histo_desc* h,
                      int64_t* num_z_bin, int64_t* num_phi, int64_t* num_r)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "histo_setup");
	if (r)
		return r;
	if (h == nullptr || h->dets_per_ring <= 0 || h->num_rings <= 0 ||
	    h->num_doi <= 0)
		return fail(YRTMI_EINVAL, "bad histogram descriptor");
	// Histogram3D ctor (Histogram3D.cpp:165-203)
	if (h->dets_per_ring % 2 != 0)
		return fail(YRTMI_EINVAL, "dets_per_ring must be even");
	if (h->max_ring_diff >= h->num_rings || h->max_ring_diff < 0)
		return fail(YRTMI_EINVAL, "max ring difference must be < number of rings");
	if (h->min_ang_diff <= 0)
		return fail(YRTMI_EINVAL, "minimum angle difference must be > 0");
	if (h->num_rings >= 65536)
		return fail(YRTMI_EINVAL, "too many rings");
	int64_t rCut = h->min_ang_diff / 2;
	if (h->min_ang_diff % 2 != 0 && h->dets_per_ring % 4 == 0)
		rCut++;
	const int64_t doiPoss = h->num_doi * h->num_doi;
	const int64_t numR = doiPoss * (h->dets_per_ring / 2 + 1 - h->min_ang_diff);
	const int64_t numPhi = h->dets_per_ring;
	const int64_t dz = h->max_ring_diff;
	int64_t numZ = (dz + 1) * h->num_rings - (dz * (dz + 1)) / 2;
	const int64_t numZDiff = numZ - h->num_rings;
	numZ += numZDiff;
	if (numR <= 0)
		return fail(YRTMI_EINVAL, "minimum angle difference too large");
	if (numR * numPhi * numZ >= (int64_t(1) << 32))
		return fail(YRTMI_EINVAL, "histogram larger than 2^32 bins");
	if (h->dets_per_ring * h->num_rings * h->num_doi > c->ndet && c->lut != nullptr)
		return fail(YRTMI_EINVAL, "histogram needs more detectors than the LUT has");
	// z-bin table: the delta-z loop of getDetPairFromCoords (339-357)
	std::vector<uint32_t> tab(numZ);
	const int64_t nr = h->num_rings;
	for (int64_t zb = 0; zb < numZ; zb++)
	{
		int64_t z1, z2;
		if (zb < nr)
		{
			z1 = z2 = zb;
		}
		else
		{
			int64_t cur = (zb - nr) % numZDiff + nr;
			int64_t planes = nr;
			int64_t d = 0;
			while (cur - planes >= 0)
			{
				cur -= planes;
				planes -= 1;
				d++;
			}
			z1 = cur;
			z2 = z1 + d;
			if (zb - nr >= numZDiff)
				std::swap(z1, z2);
		}
		tab[zb] = (uint32_t)z1 | ((uint32_t)z2 << 16);
	}
	HIP_TRY(hipStreamSynchronize(c->stream));
	HIP_TRY(hipFree(c->ringMap));
	c->ringMap = nullptr;  // rebuilt for the new geometry on first use
	HIP_TRY(hipFree(c->zbin));
	c->zbin = nullptr;
	HIP_TRY(hipMalloc(&c->zbin, tab.size() * sizeof(uint32_t)));
	HIP_TRY(hipMemcpy(c->zbin, tab.data(), tab.size() * sizeof(uint32_t),
	                  hipMemcpyHostToDevice));
	c->histo = *h;
	c->numR = (uint32_t)numR;
	c->numPhi = (uint32_t)numPhi;
	c->numZBin = (uint32_t)numZ;
	c->numZBinDiff = (uint32_t)numZDiff;
	c->rCut = (uint32_t)rCut;
	c->histoReady = true;
	if (num_z_bin)
		*num_z_bin = numZ;
	if (num_phi)
		*num_phi = numPhi;
	if (num_r)
		*num_r = numR;
	return YRTMI_OK;
}

int64_t yrtmi_histo_subset_size(yrtmi_ctx* c, int S, int s)
{
	if (c == nullptr || !c->histoReady || S <= 0 || s < 0 || s >= S)
		return -1;
	int64_t nps = c->numPhi / S;
	if (s + nps * S < (int64_t)c->numPhi)
		nps += 1;
	return (int64_t)c->numR * nps * c->numZBin;
}

int yrtmi_motion_upload(yrtmi_ctx* c, const float* xf, int k)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "motion_upload");
	if (r)
		return r;
	if (xf == nullptr || k <= 0 || k > 256)
		return fail(YRTMI_EINVAL, "motion: need 1..256 transforms");
	HIP_TRY(hipStreamSynchronize(c->stream));
	HIP_TRY(hipFree(c->motion));
	c->motion = nullptr;
	HIP_TRY(hipMalloc(&c->motion, sizeof(float) * 12 * k));
	HIP_TRY(hipMemcpy(c->motion, xf, sizeof(float) * 12 * k,
	                  hipMemcpyHostToDevice));
	c->nmotion = k;
	return YRTMI_OK;
}

int yrtmi_proj_create(yrtmi_ctx* c, const yrtmi_projector_desc* d,
                      yrtmi_proj** out)
{
	int r = setDev(c);
	if (r)
		return r;
	if (d == nullptr || out == nullptr)
		return fail(YRTMI_EINVAL, "null projector descriptor");
	if (d->type != YRTMI_PROJ_SIDDON && d->type != YRTMI_PROJ_DD)
		return fail(YRTMI_EINVAL, "unknown projector type");
	if (d->num_rays < 1 || d->num_rays > YRTMI_MAX_RAYS)
		return fail(YRTMI_EINVAL, "num_rays must be in [1, 16]");
	if (d->use_tof && !(d->tof_width_ps > 0.f))
		return fail(YRTMI_EINVAL, "TOF width must be positive");
	auto* p = new yrtmi_proj();
	p->ctx = c;
	p->device = c->device;
	DProj& q = p->d;
	q.type = d->type;
	q.numRays = (d->type == YRTMI_PROJ_SIDDON) ? d->num_rays : 1;
	q.useTof = d->use_tof ? 1 : 0;
	q.setup64 = d->setup_fp64 ? 1 : 0;
	q.tz = d->crystal_size_z;
	q.tt = d->crystal_size_trans;
	// TimeOfFlightHelper ctor (TimeOfFlight.hpp:19-34), FP32 members
	if (d->use_tof)
	{
		const double width_mm = d->tof_width_ps * 0.299792458 * 0.5;
		q.tof.sigma = (float)(width_mm / (2 * sqrtf(2 * logf(2.0f))));
		q.tof.trunc = (d->tof_n_std <= 0) ? -1.f : d->tof_n_std * q.tof.sigma;
		q.tof.norm = (float)(1 / (std::sqrt(2 * 3.14159265358979) * q.tof.sigma));
		q.tof.rsigma = rcpNearest(q.tof.sigma);
		q.tof64.sigma = width_mm / (2 * std::sqrt(2 * std::log(2.0)));
		q.tof64.trunc = (d->tof_n_std <= 0) ? -1.0 : d->tof_n_std * q.tof64.sigma;
		q.tof64.norm = 1 / (std::sqrt(2 * M_PI) * q.tof64.sigma);
	}
	// multi-ray offsets: rand_r(seed = 13), 4 per extra ray
	// (MultiRayGenerator.cpp:78-92; the seed is reset for every LOR,
	// ProjectorSiddon.cpp:258,330)
	unsigned int seed = 13;
	for (int i = 0; i < 4 * (q.numRays - 1); i++)
		q.rayOff[i] = static_cast<float>(rand_r(&seed)) /
		                  static_cast<float>(RAND_MAX) -
		              0.5f;
	if (d->type == YRTMI_PROJ_DD && d->psf_kernels != nullptr)
	{
		if (d->psf_ksize <= 0 || d->psf_ksize % 2 == 0 || d->psf_num_s <= 0 ||
		    !(d->psf_s_step > 0.f) || !(d->psf_k_spacing > 0.f))
		{
			delete p;
			return fail(YRTMI_EINVAL, "PSF kernels must have odd length and positive steps");
		}
		const int64_t nk = (int64_t)d->psf_num_s * d->psf_ksize;
		std::vector<float> tab(2 * nk);
		for (int64_t i = 0; i < d->psf_num_s; i++)
			for (int64_t j = 0; j < d->psf_ksize; j++)
			{
				const float v = d->psf_kernels[i * d->psf_ksize + j];
				tab[i * d->psf_ksize + j] = v;
				tab[nk + i * d->psf_ksize + (d->psf_ksize - 1 - j)] = v;
			}
		// prefix sums of each row's trapezoids kSpacing * (tap(xi) + tap(xi+1))
		// / 2 over xi = -1 .. K - 1 (taps outside [0, K) are 0), in FP64: the
		// integral over whole kernel segments is a difference of two entries
		const int64_t K = d->psf_ksize, KC = K + 2;
		std::vector<double> cum(2 * (size_t)d->psf_num_s * KC);
		psfPrefixSums(tab.data(), d->psf_num_s, K, d->psf_k_spacing, cum.data());
		psfPrefixSums(tab.data() + nk, d->psf_num_s, K, d->psf_k_spacing,
		              cum.data() + (size_t)d->psf_num_s * KC);
		// the device rows: the taps with two zeros on either side
		const int64_t KP = psfRowTaps((int)K), nkp = (int64_t)d->psf_num_s * KP;
		std::vector<float> padded(2 * nkp, 0.f);
		for (int64_t h = 0; h < 2; h++)
			for (int64_t i = 0; i < d->psf_num_s; i++)
				for (int64_t j = 0; j < K; j++)
					padded[h * nkp + i * KP + 2 + j] = tab[h * nk + i * K + j];
		hipError_t e = hipMalloc(&p->psf, padded.size() * sizeof(float));
		if (e == hipSuccess)
			e = hipMemcpy(p->psf, padded.data(), padded.size() * sizeof(float),
			              hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMalloc(&p->psfCum, cum.size() * sizeof(double));
		if (e == hipSuccess)
			e = hipMemcpy(p->psfCum, cum.data(), cum.size() * sizeof(double),
			              hipMemcpyHostToDevice);
		if (e != hipSuccess)
		{
			(void)hipFree(p->psf);
			(void)hipFree(p->psfCum);
			delete p;
			return hipFail(e, "yrtmi_proj_create (PSF upload)");
		}
		q.psf.k = p->psf;
		q.psf.kflip = p->psf + nkp;
		q.psf.cum = p->psfCum;
		q.psf.cumFlip = p->psfCum + (size_t)d->psf_num_s * KC;
		q.psf.rks = rcpNearest(d->psf_k_spacing);
		q.psf.numS = d->psf_num_s;
		q.psf.K = d->psf_ksize;
		q.psf.sStep = d->psf_s_step;
		q.psf.kSpacing = d->psf_k_spacing;
		// getHalfWidth_mm (ProjectionPsfManager.cpp:52-56)
		q.psf.halfWidth = (float)((d->psf_ksize + 1) / 2) * d->psf_k_spacing;
		const int64_t lds = nkp * (int64_t)sizeof(float) + (int64_t)d->psf_num_s * KC * 8;
		q.psf.ldsBytes = lds <= kPsfLdsMax ? (int32_t)lds : 0;
	}
	p->lpe = (q.numRays >= 4) ? 4 : (q.numRays >= 2 ? 2 : 1);
	*out = p;
	return YRTMI_OK;
}

int yrtmi_proj_destroy(yrtmi_proj* p)
{
	if (p == nullptr)
		return YRTMI_OK;
	if (p->psf)
	{
		// hipFree waits for the device's outstanding work
		(void)hipSetDevice(p->device);
		(void)hipFree(p->psf);
		(void)hipFree(p->psfCum);
	}
	delete p;
	return YRTMI_OK;
}

int yrtmi_proj_set_accumulation(yrtmi_proj* p, int mode)
{
	if (p == nullptr)
		return fail(YRTMI_EINVAL, "null projector");
	if (mode != YRTMI_ACCUM_DIRECT && mode != YRTMI_ACCUM_XCD)
		return fail(YRTMI_EINVAL, "unknown accumulation mode");
	p->accum = mode;
	return YRTMI_OK;
}

int yrtmi_proj_set_launch(yrtmi_proj* p, int lpe, int block, int grid)
{
	if (p == nullptr)
		return fail(YRTMI_EINVAL, "null projector");
	if (lpe != 1 && lpe != 2 && lpe != 4 && lpe != 8)
		return fail(YRTMI_EINVAL, "lanes_per_event must be 1, 2, 4 or 8");
	if (block < 64 || block > 256 || block % 64 != 0)
		return fail(YRTMI_EINVAL, "block size must be 64, 128, 192 or 256");
	if (grid < 0)
		return fail(YRTMI_EINVAL, "grid must be >= 0");
	p->lpe = lpe;
	p->block = block;
	p->grid = grid;
	return YRTMI_OK;
}

static int prepare(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                   const yrtmi_lors* l, KArgs& a)
{
	int r = setDev(c);
	if (r)
		return r;
	if (p == nullptr || p->ctx != c)
		return fail(YRTMI_EINVAL, "projector does not belong to this context");
	r = checkImage(im, "image");
	if (r)
		return r;
	a = KArgs{};
	a.im = toDImage(im);
	a.pr = p->d;
	return buildLors(c, l, a.L, a.sc);
}

int yrtmi_forward(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                  const float* img, const yrtmi_lors* l, float* out)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr || out == nullptr)
		return fail(YRTMI_EINVAL, "forward: null image or output");
	a.x = img;
	a.out = out;
	HIP_TRY(launchEvents(MODE_FWD, a, p->lpe, p->block, p->grid, c->stream));
	return YRTMI_OK;
}

int yrtmi_back(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
               float* img, const yrtmi_lors* l, const float* vals)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr || vals == nullptr)
		return fail(YRTMI_EINVAL, "back: null image or values");
	a.y = img;
	a.vals = vals;
	const int64_t ny = a.im.nvox * a.im.nt;
	r = setupAccum(c, p, a, ny);
	if (r)
		return r;
	HIP_TRY(launchEvents(MODE_BACK, a, p->lpe, p->block, p->grid, c->stream));
	r = foldAccum(c, a, img, ny);
	if (r)
		return r;
	return YRTMI_OK;
}

int yrtmi_forward_f64(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                      const double* img, const yrtmi_lors* l, double* out)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (p->d.type != YRTMI_PROJ_SIDDON)
		return fail(YRTMI_EINVAL, "forward_f64: Siddon projector only");
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr || out == nullptr)
		return fail(YRTMI_EINVAL, "forward_f64: null image or output");
	HIP_TRY(launchProj64(true, a, img, nullptr, nullptr, out, c->stream));
	return YRTMI_OK;
}

int yrtmi_back_f64(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                   double* img, const yrtmi_lors* l, const double* vals)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (p->d.type != YRTMI_PROJ_SIDDON)
		return fail(YRTMI_EINVAL, "back_f64: Siddon projector only");
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr || vals == nullptr)
		return fail(YRTMI_EINVAL, "back_f64: null image or values");
	HIP_TRY(launchProj64(false, a, nullptr, img, vals, nullptr, c->stream));
	return YRTMI_OK;
}

int yrtmi_em_subset(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                    const float* x, float* u, const yrtmi_lors* l,
                    const yrtmi_corrections* corr)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (x == nullptr || u == nullptr)
		return fail(YRTMI_EINVAL, "em_subset: null image");
	a.x = x;
	a.y = u;
	a.c.g = 1.0f;
	a.c.denomThr = 1e-35f;
	if (corr != nullptr)
	{
		a.c.meas = corr->meas;
		a.c.sens = corr->sens;
		a.c.att = corr->att;
		a.c.randoms = corr->randoms;
		a.c.scatter = corr->scatter;
		a.c.g = corr->global_scale;
		a.c.denomThr = corr->denom_threshold;
		a.c.precorr = corr->precorr;
	}
	const int64_t ny = a.im.nvox * a.im.nt;
	r = setupAccum(c, p, a, ny);
	if (r)
		return r;
	HIP_TRY(launchEvents(MODE_EM, a, p->lpe, p->block, p->grid, c->stream));
	r = foldAccum(c, a, u, ny);
	if (r)
		return r;
	return YRTMI_OK;
}

int yrtmi_em_subset_noalloc(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                            const float* x, float* u, const yrtmi_lors* l,
                            const yrtmi_corrections* corr)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (x == nullptr || u == nullptr)
		return fail(YRTMI_EINVAL, "em_subset_noalloc: null image");
	a.x = x;
	a.y = u;
	a.c.g = 1.0f;
	a.c.denomThr = 1e-35f;
	if (corr != nullptr)
	{
		a.c.meas = corr->meas;
		a.c.sens = corr->sens;
		a.c.att = corr->att;
		a.c.randoms = corr->randoms;
		a.c.scatter = corr->scatter;
		a.c.g = corr->global_scale;
		a.c.denomThr = corr->denom_threshold;
		a.c.precorr = corr->precorr;
	}
	const int64_t ny = a.im.nvox * a.im.nt;
	a.part = nullptr;
	if (p->accum == YRTMI_ACCUM_XCD && c->nxcc > 1)
	{
		// the partial images in the scratch array: nxcc of stride floats
		const int64_t stride = (ny + 63) / 64 * 64;
		if (stride * c->nxcc > kScratchFloats)
		{
			char msg[200];
			std::snprintf(msg, sizeof msg,
			              "em_subset_noalloc: %lld voxels x frames need %lld floats of "
			              "scratch (%d partial images), above its %lld",
			              (long long)ny, (long long)(stride * c->nxcc), c->nxcc,
			              (long long)kScratchFloats);
			return fail(YRTMI_EINVAL, msg);
		}
		float* ptr = scratchOf(c->device);
		if (ptr == nullptr)
			return fail(YRTMI_ESTATE, "em_subset_noalloc: no scratch on this device");
		if (l->n == 0)
			return YRTMI_OK;
		// the array keeps the last call's contents: clear what this call
		// accumulates into (a fill kernel, also under graph recording)
		HIP_TRY(launchFill(ptr, stride * c->nxcc, 0.f, c->stream));
		a.part = ptr;
		a.partStride = stride;
		a.nxcc = c->nxcc;
	}
	if (l->n == 0)
		return YRTMI_OK;
	HIP_TRY(launchEvents(MODE_EM, a, p->lpe, p->block, p->grid, c->stream));
	if (a.part != nullptr)
		HIP_TRY(launchFoldPartials(u, a.part, a.partStride, a.nxcc, ny, c->stream));
	return YRTMI_OK;
}

int yrtmi_sensitivity(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                      float* img, const yrtmi_lors* l,
                      const yrtmi_image* att_im, const float* att_img,
                      const float* sens)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr)
		return fail(YRTMI_EINVAL, "sensitivity: null image");
	if (l->frame != nullptr)
		return fail(YRTMI_EINVAL, "sensitivity: LORs must not carry frames");
	if (att_img != nullptr)
	{
		r = checkImage(att_im, "attenuation image");
		if (r)
			return r;
		a.attIm = toDImage(att_im);
		a.att = att_img;
	}
	a.y = img;
	a.vals = sens;
	a.L.tof = nullptr;
	r = setupAccum(c, p, a, a.im.nvox);
	if (r)
		return r;
	HIP_TRY(launchEvents(MODE_SENS, a, p->lpe, p->block, p->grid, c->stream));
	r = foldAccum(c, a, img, a.im.nvox);
	if (r)
		return r;
	return YRTMI_OK;
}

int yrtmi_attenuation_factors(yrtmi_ctx* c, const yrtmi_image* att_im,
                              const float* att_img, const yrtmi_lors* l,
                              float* acf)
{
	int r = setDev(c);
	if (r)
		return r;
	r = checkImage(att_im, "attenuation image");
	if (r)
		return r;
	KArgs a{};
	r = buildLors(c, l, a.L, a.sc);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (att_img == nullptr || acf == nullptr)
		return fail(YRTMI_EINVAL, "attenuation: null image or output");
	a.attIm = toDImage(att_im);
	a.att = att_img;
	a.out = acf;
	HIP_TRY(launchAttFactors(a, c->stream));
	return YRTMI_OK;
}

int yrtmi_em_update(yrtmi_ctx* c, float* x, const float* u, const float* s,
                    int64_t nvox, int64_t nt, int64_t sf, float thr)
{
	int r = setDev(c);
	if (r)
		return r;
	if (x == nullptr || u == nullptr || s == nullptr || nvox <= 0 || nt <= 0 ||
	    (sf != 1 && sf != nt))
		return fail(YRTMI_EINVAL, "em_update: bad arguments");
	HIP_TRY(launchEmUpdate(x, u, s, nvox, nt, sf, thr, c->stream));
	return YRTMI_OK;
}

int yrtmi_threshold(yrtmi_ctx* c, float* x, const float* m, int64_t nvox,
                    int64_t nt, int64_t mf, float thr, float les, float leo,
                    float gts, float gto)
{
	int r = setDev(c);
	if (r)
		return r;
	if (x == nullptr || m == nullptr || nvox <= 0 || nt <= 0 ||
	    (mf != 1 && mf != nt))
		return fail(YRTMI_EINVAL, "threshold: bad arguments");
	HIP_TRY(launchThreshold(x, m, nvox, nt, mf, thr, les, leo, gts, gto,
	                        c->stream));
	return YRTMI_OK;
}

int yrtmi_psf_apply(yrtmi_ctx* c, const yrtmi_image* im, const float* in, float* out,
                    float* tmp, const float* kx, int nkx, const float* ky, int nky,
                    const float* kz, int nkz, int transpose)
{
	int r = setDev(c);
	if (r)
		return r;
	r = checkImage(im, "image");
	if (r)
		return r;
	if (in == nullptr || out == nullptr || tmp == nullptr || in == out || in == tmp ||
	    out == tmp)
		return fail(YRTMI_EINVAL, "psf: in, out and tmp must be distinct device buffers");
	const int64_t total = im->nx * im->ny * im->nz * im->nt;
	if (total >= (int64_t(1) << 31))
		return fail(YRTMI_EINVAL, "psf: 2^31 voxels or more");
	const float* ks[3] = {kx, ky, kz};
	const int nk[3] = {nkx, nky, nkz};
	const int64_t dims[3] = {im->nx, im->ny, im->nz};
	PsfTaps taps[3];
	for (int a = 0; a < 3; a++)
	{
		if (ks[a] == nullptr || nk[a] <= 0 || nk[a] % 2 == 0 || nk[a] > kMaxPsfTaps)
			return fail(YRTMI_EINVAL, "psf: kernels must have an odd length <= 63");
		if (nk[a] / 2 >= dims[a])
			return fail(YRTMI_EINVAL, "psf: kernel half length must be below the image size");
		taps[a].n = nk[a];
		for (int i = 0; i < nk[a]; i++)
			taps[a].k[i] = transpose ? ks[a][nk[a] - 1 - i] : ks[a][i];
	}
	const int nx = (int)im->nx, ny = (int)im->ny, nz = (int)im->nz;
	// x: in -> out, y: out -> tmp, z: tmp -> out
	HIP_TRY(launchPsfAxis(0, in, out, taps[0], nx, ny, nz, (int)total, c->stream));
	HIP_TRY(launchPsfAxis(1, out, tmp, taps[1], nx, ny, nz, (int)total, c->stream));
	HIP_TRY(launchPsfAxis(2, tmp, out, taps[2], nx, ny, nz, (int)total, c->stream));
	return YRTMI_OK;
}

// util::invertTransform (TransformUtils.hpp:16-40): [R^T | -R^T t], each
// dot product summed left to right in FP32.
static void invertRigid(const float* t, float* o)
{
#pragma clang fp contract(off)
	const float R[9] = {t[0], t[1], t[2], t[4], t[5], t[6], t[8], t[9], t[10]};
	const float tr[3] = {t[3], t[7], t[11]};
	for (int i = 0; i < 3; i++)
	{
		const float a = R[i], b = R[3 + i], c = R[6 + i];  // row i of R^T
		o[4 * i] = a;
		o[4 * i + 1] = b;
		o[4 * i + 2] = c;
		o[4 * i + 3] = -(a * tr[0] + b * tr[1] + c * tr[2]);
	}
}

int yrtmi_time_average_move(yrtmi_ctx* c, const yrtmi_image* im, const float* in,
                            float* out, const float* xf, const float* weights, int n)
{
	int r = setDev(c);
	if (r)
		return r;
	r = checkImage(im, "image");
	if (r)
		return r;
	r = notCapturing(c, "time_average_move");
	if (r)
		return r;
	if (in == nullptr || out == nullptr || in == out)
		return fail(YRTMI_EINVAL, "time_average_move: in and out must be distinct device buffers");
	if (xf == nullptr || weights == nullptr || n <= 0 || n > (1 << 24))
		return fail(YRTMI_EINVAL, "time_average_move: need 1..2^24 transforms and weights");
	if (im->nx * im->ny * im->nz >= (int64_t(1) << 31))
		return fail(YRTMI_EINVAL, "time_average_move: 2^31 voxels or more");
	std::vector<float4> fr((size_t)4 * n);
	for (int f = 0; f < n; f++)
	{
		float inv[12];
		invertRigid(xf + 12 * (size_t)f, inv);
		for (int i = 0; i < 3; i++)
			fr[4 * f + i] = make_float4(inv[4 * i], inv[4 * i + 1], inv[4 * i + 2], inv[4 * i + 3]);
		fr[4 * f + 3] = make_float4(weights[f], 0.f, 0.f, 0.f);
	}
	// the buffer may still be read by the previous call's kernel
	HIP_TRY(hipStreamSynchronize(c->stream));
	if (n > c->moveCap)
	{
		HIP_TRY(hipFree(c->moveFrames));
		c->moveFrames = nullptr;
		c->moveCap = 0;
		HIP_TRY(hipMalloc(&c->moveFrames, sizeof(float4) * 4 * (size_t)n));
		c->moveCap = n;
	}
	HIP_TRY(hipMemcpy(c->moveFrames, fr.data(), sizeof(float4) * fr.size(),
	                  hipMemcpyHostToDevice));
	MoveGeom g{};
	const int64_t dims[3] = {im->nx, im->ny, im->nz};
	const float len[3] = {im->lx, im->ly, im->lz};
	const float off[3] = {im->ox, im->oy, im->oz};
	for (int a = 0; a < 3; a++)
	{
		// timeAverageMoveImage_kernel: v = length / n, origin = off - L/2 + v/2
		g.v[a] = len[a] / (float)dims[a];
		g.len[a] = len[a];
		g.off[a] = off[a];
		g.inv[a] = 1.f / g.v[a];
		g.origin[a] = (off[a] - len[a] * 0.5f) + g.v[a] * 0.5f;
	}
	g.nx = (int)im->nx;
	g.ny = (int)im->ny;
	g.nz = (int)im->nz;
	HIP_TRY(launchTimeAverageMove(in, out, c->moveFrames, n, g, c->stream));
	return YRTMI_OK;
}

}  // extern "C"

namespace
{
// The device ring map of the context's histogram (the inverse of the
// in-ring pair of every (rr, phi), Histogram3D::setupHistogram's m_ringMap,
// Histogram3D.cpp:286-297, with the pair formula of getDetPairInSameRing,
// 428-450: phi outer, rr inner, a later (rr, phi) replacing an earlier one
// with the same pair, pairs of one in-ring position included) and the pair
// geometry; shared by the pair lookups.
int pairSetup(yrtmi_ctx* c, const char* what, PairGeom& g)
{
	if (!c->histoReady)
		return fail(YRTMI_EINVAL, std::string(what) + ": no histogram geometry");
	const int64_t dpr = c->histo.dets_per_ring;
	if (dpr >= 65536)
		return fail(YRTMI_EINVAL, std::string(what) + ": too many detectors per ring");
	if (c->ringMap == nullptr)
	{
		std::vector<int32_t> map((size_t)(dpr * dpr), -1);
		const int64_t nrr = c->numR / (c->histo.num_doi * c->histo.num_doi);
		for (int64_t phi = 0; phi < c->numPhi; phi++)
			for (int64_t rr = 0; rr < nrr; rr++)
			{
				const int64_t d02 = dpr / 2 + (phi & 1);
				const int64_t sh = rr - dpr / 4 + (int64_t)c->rCut;
				int64_t a = sh, b = d02 - sh;
				if (a < 0)
					a += dpr;
				if (b < 0)
					b += dpr;
				a = (a + phi / 2) % dpr;
				b = (b + phi / 2) % dpr;
				if (a > b)
					std::swap(a, b);
				map[(size_t)(a * dpr + b)] = (int32_t)(rr | (phi << 16));
			}
		HIP_TRY(hipStreamSynchronize(c->stream));
		HIP_TRY(hipMalloc(&c->ringMap, map.size() * sizeof(int32_t)));
		HIP_TRY(hipMemcpy(c->ringMap, map.data(), map.size() * sizeof(int32_t),
		                  hipMemcpyHostToDevice));
	}
	if (c->gatherErr == nullptr)
		HIP_TRY(hipMalloc(&c->gatherErr, sizeof(int)));
	g = PairGeom{};
	g.dpr = (uint32_t)dpr;
	g.nrings = (uint32_t)c->histo.num_rings;
	g.ndet = (uint32_t)(dpr * c->histo.num_rings * c->histo.num_doi);
	g.numDOI = (uint32_t)c->histo.num_doi;
	g.numDOIPoss = (uint32_t)(c->histo.num_doi * c->histo.num_doi);
	g.numR = c->numR;
	g.numPhi = c->numPhi;
	g.numZBinDiff = c->numZBinDiff;
	g.maxRingDiff = (int32_t)c->histo.max_ring_diff;
	return YRTMI_OK;
}

int pairErr(yrtmi_ctx* c, const char* what)
{
	int err = 0;
	HIP_TRY(hipMemcpyAsync(&err, c->gatherErr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
	HIP_TRY(hipStreamSynchronize(c->stream));
	if (err)
		return fail(YRTMI_EINVAL, std::string(what) +
		                              ": a detector pair outside the histogram (unknown "
		                              "detector, no in-ring bin, the pair (0, 0) or a ring "
		                              "difference above the limit)");
	return YRTMI_OK;
}
}  // namespace

extern "C" {

int yrtmi_histo_gather_pairs(yrtmi_ctx* c, const float* histo, const uint32_t* d1,
                             const uint32_t* d2, int64_t n, float* out)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "histo_gather_pairs");
	if (r)
		return r;
	if (histo == nullptr || d1 == nullptr || d2 == nullptr || out == nullptr || n < 0)
		return fail(YRTMI_EINVAL, "histo_gather_pairs: bad arguments");
	if (n == 0)
		return YRTMI_OK;
	PairGeom g;
	r = pairSetup(c, "histo_gather_pairs", g);
	if (r)
		return r;
	HIP_TRY(hipMemsetAsync(c->gatherErr, 0, sizeof(int), c->stream));
	HIP_TRY(launchGatherPairs(histo, d1, d2, n, c->ringMap, g, out, c->gatherErr, c->stream));
	return pairErr(c, "histo_gather_pairs");
}

int yrtmi_histo_pair_bins(yrtmi_ctx* c, const uint32_t* d1, const uint32_t* d2, int64_t n,
                          uint32_t* bins)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "histo_pair_bins");
	if (r)
		return r;
	if (d1 == nullptr || d2 == nullptr || bins == nullptr || n < 0)
		return fail(YRTMI_EINVAL, "histo_pair_bins: bad arguments");
	if (n == 0)
		return YRTMI_OK;
	PairGeom g;
	r = pairSetup(c, "histo_pair_bins", g);
	if (r)
		return r;
	if ((int64_t)c->numZBin * c->numPhi * c->numR >= 0xffffffffLL)
		return fail(YRTMI_EINVAL, "histo_pair_bins: the histogram has 2^32 - 1 bins or more");
	HIP_TRY(hipMemsetAsync(c->gatherErr, 0, sizeof(int), c->stream));
	HIP_TRY(launchPairBins(d1, d2, n, c->ringMap, g, bins, c->gatherErr, c->stream));
	return pairErr(c, "histo_pair_bins");
}

int yrtmi_fill_i32(yrtmi_ctx* c, int32_t* x, int64_t n, int32_t v)
{
	int r = setDev(c);
	if (r)
		return r;
	if (n < 0 || (n > 0 && x == nullptr))
		return fail(YRTMI_EINVAL, "fill_i32: bad arguments");
	HIP_TRY(launchFillI32(x, n, v, c->stream));
	return YRTMI_OK;
}

int yrtmi_fill(yrtmi_ctx* c, float* x, int64_t n, float v)
{
	int r = setDev(c);
	if (r)
		return r;
	if (x == nullptr || n < 0)
		return fail(YRTMI_EINVAL, "fill: bad arguments");
	if (n == 0)
		return YRTMI_OK;
	HIP_TRY(launchFill(x, n, v, c->stream));
	return YRTMI_OK;
}

int yrtmi_count_segments(yrtmi_ctx* c, const yrtmi_proj* p,
                         const yrtmi_image* im, const yrtmi_lors* l,
                         int64_t* total)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = notCapturing(c, "count_segments");
	if (r)
		return r;
	if (total == nullptr)
		return fail(YRTMI_EINVAL, "count: null output");
	*total = 0;
	if (l->n == 0)
		return YRTMI_OK;
	unsigned long long zero = 0;
	HIP_TRY(hipMemcpyAsync(c->counter, &zero, sizeof(zero),
	                       hipMemcpyHostToDevice, c->stream));
	a.counter = c->counter;
	HIP_TRY(launchEvents(MODE_COUNT, a, p->lpe, p->block, p->grid, c->stream));
	unsigned long long h = 0;
	HIP_TRY(hipMemcpyAsync(&h, c->counter, sizeof(h), hipMemcpyDeviceToHost,
	                       c->stream));
	HIP_TRY(hipStreamSynchronize(c->stream));
	*total = (int64_t)h;
	return YRTMI_OK;
}

}  // extern "C"

struct yrtmi_plan
{
	int device = 0;
	const yrtmi_proj* proj = nullptr;
	yrtmi_image im{};
	yrtmi_lors lors{};
	int useTof = 0;
	DTiles T{};
	int64_t records = 0, nChunks = 0, bytes = 0;
	uint32_t* rec = nullptr;  // records, recWords words each (DPlan::rec)
	float4* ray = nullptr;    // ray table (rayTabVec / kDDVec float4 per entry)
	TileChunk* chunks = nullptr;
	float* v = nullptr;  // per-LOR values of the em / sensitivity calls
	unsigned int* queue = nullptr;  // chunk queue of persistent launches
	int persistent = 0;
};

namespace
{

void freePlan(yrtmi_plan* q)
{
	hipFree(q->rec);
	hipFree(q->ray);
	hipFree(q->chunks);
	hipFree(q->v);
	hipFree(q->queue);
	delete q;
}

bool sameImage(const yrtmi_image& a, const yrtmi_image& b)
{
	return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.nt == b.nt &&
	       a.vx == b.vx && a.vy == b.vy && a.vz == b.vz && a.lx == b.lx &&
	       a.ly == b.ly && a.lz == b.lz && a.ox == b.ox && a.oy == b.oy &&
	       a.oz == b.oz && a.fov_radius == b.fov_radius;
}

bool sameLors(const yrtmi_lors& a, const yrtmi_lors& b)
{
	return a.kind == b.kind && a.n == b.n && a.d1 == b.d1 && a.d2 == b.d2 &&
	       a.bins == b.bins && a.num_subsets == b.num_subsets &&
	       a.subset == b.subset && a.tof == b.tof && a.frame == b.frame &&
	       a.motion_frame == b.motion_frame;
}

// Common checks of the *_planned calls.
int checkPlan(const yrtmi_plan* q, const yrtmi_proj* p, const yrtmi_image* im,
              const yrtmi_lors* l)
{
	if (q == nullptr)
		return fail(YRTMI_EINVAL, "null plan");
	if (q->proj != p)
		return fail(YRTMI_EINVAL, "plan was made for another projector");
	if (!sameImage(q->im, *im))
		return fail(YRTMI_EINVAL, "plan was made for another image geometry");
	if (!sameLors(q->lors, *l))
		return fail(YRTMI_EINVAL, "plan was made for another LOR set");
	return YRTMI_OK;
}

// Ray table for the EM ratio's forward projection: DD LORs read their
// stored footprint set-up; Siddon LORs are resolved from their detector
// pairs (the LUT stays in L2; reading the table's 32 bytes per ray from HBM
// measured slower).
const float4* planRays(const yrtmi_plan* q)
{
	return q->proj->d.type == YRTMI_PROJ_DD ? q->ray : nullptr;
}

DPlanRead planRead(const yrtmi_plan* q)
{
	DPlanRead r{};
	r.R = q->proj->d.numRays;
	r.chunks = q->chunks;
	r.nChunks = q->nChunks;
	r.rec = q->rec;
	r.ray = q->ray;
	r.queue = q->queue;
	r.vexp = reinterpret_cast<const int*>(q->queue + 2);
	return r;
}

}  // namespace

extern "C" {

int yrtmi_plan_create(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                      const yrtmi_lors* l, const yrtmi_plan_options* opt,
                      yrtmi_plan** out)
{
	if (out == nullptr)
		return fail(YRTMI_EINVAL, "plan: null output");
	*out = nullptr;
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = notCapturing(c, "plan_create");
	if (r)
		return r;
	const bool dd = p->d.type == YRTMI_PROJ_DD;
	if (im->nx > 1024 || im->ny > 1024 || im->nz > 1024)
		return fail(YRTMI_EINVAL, "plan: image dimensions above 1024");
	if (!dd)
	{
		// the bricked image of the planned EM ratio, sized here
		r = ensureBrick(c, brickFrameVoxels(a.im.nx, a.im.ny, a.im.nz) * a.im.nt);
		if (r)
			return r;
	}
	const int R = p->d.numRays;
	if ((l->n * (int64_t)R) >= (int64_t(1) << 32))
		return fail(YRTMI_EINVAL, "plan: n * num_rays must be below 2^32");
	const yrtmi_plan_options o = opt ? *opt : yrtmi_plan_options{0, 0, 0, 0, 0};
	if (o.chunk_records < 0 || o.chunk_records > (1 << 24))
		return fail(YRTMI_EINVAL, "plan: chunk_records out of range");
	if (o.tile_shape < 0 || o.tile_shape >= kNumTileShapes)
		return fail(YRTMI_EINVAL, "plan: unknown tile shape");
	if (o.persistent < 0)
		return fail(YRTMI_EINVAL, "plan: persistent workgroups must be >= 0");
	if (o.max_bytes < 0)
		return fail(YRTMI_EINVAL, "plan: max_bytes must be >= 0");
	// the plan's memory within the budget: the value buffer and ray table
	// now, the records after the count pass
	// ray table: kDDVec float4 per DD LOR (its set-up), two per Siddon ray (its
	// end points) and two more with the FP32 set-up (rayTabVec)
	const int64_t tabVec = dd ? kDDVec : rayTabVec(p->d.setup64 != 0);
	const int64_t tabBytes = tabVec * (int64_t)sizeof(float4) * l->n * (int64_t)R;
	if (o.max_bytes > 0 && (int64_t)sizeof(float) * l->n + tabBytes > o.max_bytes)
		return fail(YRTMI_ENOMEM, "plan: the ray table exceeds the memory budget");
	const int CH = o.chunk_records > 0 ? o.chunk_records : 4096;
	DTiles T{};
	T.shape = o.tile_shape;
	T.tx = tileDim(T.shape, 0);
	T.ty = tileDim(T.shape, 1);
	T.tz = tileDim(T.shape, 2);
	T.lx = __builtin_ctz((unsigned)T.tx);
	T.ly = __builtin_ctz((unsigned)T.ty);
	T.lz = __builtin_ctz((unsigned)T.tz);
	T.ntx = (int32_t)((im->nx + T.tx - 1) / T.tx);
	T.nty = (int32_t)((im->ny + T.ty - 1) / T.ty);
	T.ntz = (int32_t)((im->nz + T.tz - 1) / T.tz);
	const int64_t nT = (int64_t)T.ntx * T.nty * T.ntz * im->nt;
	// the fill pass keeps a counter and a 32-bit base per bin in LDS
	int maxLds = 65536;
	if (hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
	    hipSuccess)
		maxLds = 65536;
	const int64_t binCap = (int64_t)maxLds / 8 - 2;
	if (nT > binCap)
	{
		char msg[160];
		std::snprintf(msg, sizeof msg,
		              "plan: %lld image tiles (frames included), at most %lld: use a larger "
		              "tile shape",
		              (long long)nT, (long long)binCap);
		return fail(YRTMI_EINVAL, msg);
	}
	T.nTiles = (int32_t)nT;

	yrtmi_plan* q = new yrtmi_plan();
	q->device = c->device;
	q->proj = p;
	q->im = *im;
	q->lors = *l;
	q->useTof = p->d.useTof;
	q->T = T;
	q->persistent = o.persistent;
	// [0] next chunk, [1] finished workgroups (persistent launches),
	// [2] fixed-point scale exponent of the current launch, [3] unused,
	// [4 .. 259] exponent histogram of its values
	if (hipMalloc(&q->queue, 260 * sizeof(unsigned int)) != hipSuccess ||
	    hipMemset(q->queue, 0, 260 * sizeof(unsigned int)) != hipSuccess)
	{
		freePlan(q);
		return fail(YRTMI_ENOMEM, "plan: chunk queue");
	}
	if (l->n > 0)
		if (hipMalloc(&q->v, sizeof(float) * l->n) != hipSuccess)
		{
			freePlan(q);
			return fail(YRTMI_ENOMEM, "plan: value buffer");
		}
	q->bytes = sizeof(float) * l->n;

	// scratch of the build
	unsigned long long* dCount = nullptr;
	unsigned int* dCursor = nullptr;
	unsigned int* dBlockCnt = nullptr;
	int64_t* dOff = nullptr;
	int* dErr = nullptr;
	auto cleanup = [&](int rc) {
		hipFree(dCount);
		hipFree(dCursor);
		hipFree(dBlockCnt);
		hipFree(dOff);
		hipFree(dErr);
		if (rc != YRTMI_OK)
			freePlan(q);
		else
			*out = q;
		return rc;
	};
	if (l->n == 0)
		return cleanup(YRTMI_OK);
	// length classes: up to 8, as long as tiles x classes stays within
	// 4096 bins (48 KB of fill-pass LDS) and the LDS limit.  Siddon: cut at
	// fractions of the longest piece a tile can hold (tx + ty + tz segments);
	// DD: at multiples of the slices of a tile slab (a piece of a footprint
	// a few voxels wide and high holds a few voxels per slice)
	DPlan pl{};
	pl.R = R;
	pl.C = 1;
	// (Siddon: at most 8; 16 measured no faster there, with slower binning)
	const int maxC = dd ? kMaxClasses : 8;
	while (pl.C < maxC && nT * pl.C * 2 <= std::min<int64_t>(binCap, kMaxBins))
		pl.C *= 2;
	// (16 classes, DD: about 15 % apart where most pieces lie)
	if (dd)
	{
		const int L = std::max(T.tx, T.ty);
		static const int f16[15] = {1, 2, 3, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40},
		                 f8[7] = {1, 2, 4, 8, 12, 16, 24}, f4[3] = {2, 6, 16}, f2[1] = {8};
		const int* f = pl.C == 16 ? f16 : (pl.C == 8 ? f8 : (pl.C == 4 ? f4 : f2));
		for (int k = 0; k < pl.C - 1; k++)
			pl.cut[k] = std::max(1, L * f[k] / 4);
	}
	else
	{
		const int L = T.tx + T.ty + T.tz;
		static const int f8[7] = {3, 6, 9, 12, 16, 21, 28}, f4[3] = {6, 12, 21}, f2[1] = {12};
		const int* f = pl.C == 8 ? f8 : (pl.C == 4 ? f4 : f2);
		for (int k = 0; k < pl.C - 1; k++)
			pl.cut[k] = std::max(1, L * f[k] / 40);
	}
	const int64_t nB = nT * pl.C;
	const int64_t rays = l->n * (int64_t)R;
	const int grid = (int)std::min<int64_t>((rays + 255) / 256, 2048);
	if (hipMalloc(&dCount, sizeof(unsigned long long) * nB) != hipSuccess ||
	    hipMalloc(&dCursor, sizeof(unsigned int) * nB) != hipSuccess ||
	    hipMalloc(&dBlockCnt, sizeof(unsigned int) * nB * grid) != hipSuccess ||
	    hipMalloc(&dOff, sizeof(int64_t) * nB) != hipSuccess ||
	    hipMalloc(&dErr, sizeof(int)) != hipSuccess)
		return cleanup(fail(YRTMI_ENOMEM, "plan: build scratch"));
	if (hipMalloc(&q->ray, tabVec * sizeof(float4) * rays) != hipSuccess)
		return cleanup(fail(YRTMI_ENOMEM, "plan: ray table"));
	q->bytes += rays * tabVec * (int64_t)sizeof(float4);
	hipError_t e = hipMemsetAsync(dCount, 0, sizeof(unsigned long long) * nB, c->stream);
	if (e == hipSuccess)
		e = hipMemsetAsync(dCursor, 0, sizeof(unsigned int) * nB, c->stream);
	if (e == hipSuccess)
		e = hipMemsetAsync(dErr, 0, sizeof(int), c->stream);
	if (e != hipSuccess)
		return cleanup(hipFail(e, "plan: clear"));

	pl.count = dCount;
	pl.cursor = dCursor;
	pl.blockCnt = dBlockCnt;
	pl.tileOff = dOff;
	pl.err = dErr;
	pl.ray = q->ray;
	e = launchRayPrep(a, q->ray, c->stream);
	if (e == hipSuccess)
		e = launchBin(0, a, T, pl, grid, c->stream);
	std::vector<unsigned long long> cnt(nB);
	if (e == hipSuccess)
		e = hipMemcpyAsync(cnt.data(), dCount, sizeof(unsigned long long) * nB,
		                   hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(c->stream);
	if (e != hipSuccess)
		return cleanup(hipFail(e, "plan: count pass"));

	// bins tile-major, class-minor: a tile's records are one contiguous
	// range, ordered by class, cut into chunks of CH records
	std::vector<int64_t> off(nB);
	std::vector<TileChunk> chunks;
	int64_t total = 0;
	for (int64_t t = 0; t < nT; t++)
	{
		const int64_t begin = total;
		for (int k = 0; k < pl.C; k++)
		{
			off[t * pl.C + k] = total;
			total += (int64_t)cnt[t * pl.C + k];
		}
		const int64_t n = total - begin;
		if (n >= (int64_t(1) << 31))
			return cleanup(fail(YRTMI_EINVAL, "plan: more than 2^31 records in one tile"));
		for (int64_t b = 0; b < n; b += CH)
		{
			TileChunk k;
			k.begin = begin + b;
			k.tile = (int32_t)t;
			k.count = (int32_t)std::min<int64_t>(CH, n - b);
			chunks.push_back(k);
		}
	}
	if ((int64_t)chunks.size() >= 0x7fffffffLL)
		return cleanup(fail(YRTMI_EINVAL, "plan: too many chunks"));
	if (total >= (int64_t(1) << 32))
		return cleanup(fail(YRTMI_EINVAL,
		                    "plan: 2^32 records or more (fewer LORs per plan, or larger tiles)"));
	if (o.max_bytes > 0 &&
	    q->bytes + total * (dd ? 8 : 12) + (int64_t)(sizeof(TileChunk) * chunks.size()) >
	        o.max_bytes)
		return cleanup(fail(YRTMI_ENOMEM, "plan: the records exceed the memory budget"));
	// full chunks first, the tiles' partial tail chunks last (they end the
	// launch while the full ones are still running)
	std::stable_sort(chunks.begin(), chunks.end(),
	                 [](const TileChunk& x, const TileChunk& y) { return x.count > y.count; });
	q->records = total;
	q->nChunks = (int64_t)chunks.size();
	if (total > 0)
	{
		// records of 12 B (Siddon) / 8 B (DD: no alpha)
		if (hipMalloc(&q->rec, sizeof(uint32_t) * recWords(dd) * total) != hipSuccess ||
		    hipMalloc(&q->chunks, sizeof(TileChunk) * chunks.size()) != hipSuccess)
			return cleanup(fail(YRTMI_ENOMEM, "plan: records"));
		q->bytes += total * (dd ? 8 : 12) + (int64_t)(sizeof(TileChunk) * chunks.size());
		e = hipMemcpyAsync(dOff, off.data(), sizeof(int64_t) * nB,
		                   hipMemcpyHostToDevice, c->stream);
		if (e == hipSuccess)
			e = hipMemcpyAsync(q->chunks, chunks.data(), sizeof(TileChunk) * chunks.size(),
			                   hipMemcpyHostToDevice, c->stream);
		pl.rec = q->rec;
		if (e == hipSuccess)
			e = launchBin(1, a, T, pl, grid, c->stream);
		std::vector<unsigned int> cur(nB);
		int herr = 0;
		if (e == hipSuccess)
			e = hipMemcpyAsync(cur.data(), dCursor, sizeof(unsigned int) * nB,
			                   hipMemcpyDeviceToHost, c->stream);
		if (e == hipSuccess)
			e = hipMemcpyAsync(&herr, dErr, sizeof(int), hipMemcpyDeviceToHost, c->stream);
		if (e == hipSuccess)
			e = hipStreamSynchronize(c->stream);
		if (e != hipSuccess)
			return cleanup(hipFail(e, "plan: fill pass"));
		if (herr & 2)
			return cleanup(fail(YRTMI_ESTATE, "plan: a ray's plane shift is out of range"));
		bool same = herr == 0;
		for (int64_t b = 0; b < nB && same; b++)
			same = (unsigned long long)cur[b] == cnt[b];
		if (!same)
			return cleanup(fail(YRTMI_ESTATE, "plan: fill pass disagrees with count pass"));
		// each chunk's records by their fine length class (a plan used
		// more than once)
		if (!o.single_use)
		{
			e = launchChunkSort(dd, pl, q->chunks, q->nChunks, c->stream);
			if (e == hipSuccess)
				e = hipStreamSynchronize(c->stream);
		}
		if (e != hipSuccess)
			return cleanup(hipFail(e, "plan: chunk sort"));
	}
	return cleanup(YRTMI_OK);
}

int yrtmi_plan_destroy(yrtmi_plan* q)
{
	if (q == nullptr)
		return YRTMI_OK;
	hipSetDevice(q->device);
	freePlan(q);
	return YRTMI_OK;
}

int yrtmi_plan_set_persistent(yrtmi_plan* q, int workgroups)
{
	if (q == nullptr || workgroups < 0)
		return fail(YRTMI_EINVAL, "plan_set_persistent: bad arguments");
	q->persistent = workgroups;
	return YRTMI_OK;
}

int yrtmi_plan_info(const yrtmi_plan* q, int64_t* records, int64_t* chunks,
                    int64_t* bytes)
{
	if (q == nullptr)
		return fail(YRTMI_EINVAL, "null plan");
	if (records)
		*records = q->records;
	if (chunks)
		*chunks = q->nChunks;
	if (bytes)
		*bytes = q->bytes;
	return YRTMI_OK;
}

int yrtmi_back_planned(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                       float* img, const yrtmi_lors* l, const float* vals,
                       const yrtmi_plan* q)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = checkPlan(q, p, im, l);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr || vals == nullptr)
		return fail(YRTMI_EINVAL, "back: null image or values");
	HIP_TRY(launchTileBack(a, q->T, planRead(q), vals, img, q->persistent, c->stream));
	return YRTMI_OK;
}

static void setCorr(KArgs& a, const yrtmi_corrections* corr)
{
	a.c.g = 1.0f;
	a.c.denomThr = 1e-35f;
	if (corr != nullptr)
	{
		a.c.meas = corr->meas;
		a.c.sens = corr->sens;
		a.c.att = corr->att;
		a.c.randoms = corr->randoms;
		a.c.scatter = corr->scatter;
		a.c.g = corr->global_scale;
		a.c.denomThr = corr->denom_threshold;
		a.c.precorr = corr->precorr;
	}
}

int yrtmi_em_ratio(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                   const float* x, const yrtmi_lors* l,
                   const yrtmi_corrections* corr, float* v)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (x == nullptr || v == nullptr)
		return fail(YRTMI_EINVAL, "em_ratio: null image or output");
	a.x = x;
	setCorr(a, corr);
	r = brickImage(c, a);
	if (r)
		return r;
	HIP_TRY(launchEmRatio(a, p->lpe, v, c->stream));
	return YRTMI_OK;
}

int yrtmi_em_ratio_planned(yrtmi_ctx* c, const yrtmi_proj* p, const yrtmi_image* im,
                           const float* x, const yrtmi_lors* l,
                           const yrtmi_corrections* corr, float* v, const yrtmi_plan* q)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = checkPlan(q, p, im, l);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (x == nullptr || v == nullptr)
		return fail(YRTMI_EINVAL, "em_ratio: null image or output");
	a.x = x;
	setCorr(a, corr);
	a.rays = planRays(q);
	r = brickImage(c, a);
	if (r)
		return r;
	HIP_TRY(launchEmRatio(a, p->lpe, v, c->stream));
	return YRTMI_OK;
}

int yrtmi_em_subset_planned(yrtmi_ctx* c, const yrtmi_proj* p,
                            const yrtmi_image* im, const float* x, float* u,
                            const yrtmi_lors* l, const yrtmi_corrections* corr,
                            const yrtmi_plan* q)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = checkPlan(q, p, im, l);
	if (r)
		return r;
	if (l->n == 0)
		return YRTMI_OK;
	if (x == nullptr || u == nullptr)
		return fail(YRTMI_EINVAL, "em_subset: null image");
	a.x = x;
	setCorr(a, corr);
	a.rays = planRays(q);
	r = brickImage(c, a);
	if (r)
		return r;
	HIP_TRY(launchEmRatio(a, p->lpe, q->v, c->stream));
	HIP_TRY(launchTileBack(a, q->T, planRead(q), q->v, u, q->persistent, c->stream));
	return YRTMI_OK;
}

int yrtmi_sensitivity_planned(yrtmi_ctx* c, const yrtmi_proj* p,
                              const yrtmi_image* im, float* img,
                              const yrtmi_lors* l, const yrtmi_image* att_im,
                              const float* att_img, const float* sens,
                              const yrtmi_plan* q)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = checkPlan(q, p, im, l);
	if (r)
		return r;
	if (q->useTof)
		return fail(YRTMI_EINVAL, "sensitivity: the plan's projector must have TOF off");
	if (l->n == 0)
		return YRTMI_OK;
	if (img == nullptr)
		return fail(YRTMI_EINVAL, "sensitivity: null image");
	if (l->frame != nullptr || im->nt != 1)
		return fail(YRTMI_EINVAL, "sensitivity: single-frame image and LORs only");
	if (att_img != nullptr)
	{
		r = checkImage(att_im, "attenuation image");
		if (r)
			return r;
		a.attIm = toDImage(att_im);
		a.att = att_img;
	}
	a.vals = sens;
	a.L.tof = nullptr;
	HIP_TRY(launchSensValue(a, q->v, c->stream));
	HIP_TRY(launchTileBack(a, q->T, planRead(q), q->v, img, q->persistent, c->stream));
	return YRTMI_OK;
}

}  // extern "C"

extern "C" {

int yrtmi_mem_info(yrtmi_ctx* c, int64_t* free_bytes, int64_t* total_bytes)
{
	int r = setDev(c);
	if (r)
		return r;
	size_t fr = 0, tot = 0;
	HIP_TRY(hipMemGetInfo(&fr, &tot));
	if (free_bytes != nullptr)
		*free_bytes = (int64_t)fr;
	if (total_bytes != nullptr)
		*total_bytes = (int64_t)tot;
	return YRTMI_OK;
}

int yrtmi_device_alloc(yrtmi_ctx* c, int64_t bytes, void** out)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "device_alloc");
	if (r)
		return r;
	if (out == nullptr || bytes < 0)
		return fail(YRTMI_EINVAL, "device_alloc: bad arguments");
	*out = nullptr;
	if (bytes == 0)
		return YRTMI_OK;
	if (hipMalloc(out, (size_t)bytes) != hipSuccess)
	{
		*out = nullptr;
		return fail(YRTMI_ENOMEM, "device_alloc: out of device memory");
	}
	return YRTMI_OK;
}

int yrtmi_device_free(yrtmi_ctx* c, void* p)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "device_free");
	if (r)
		return r;
	if (p == nullptr)
		return YRTMI_OK;
	HIP_TRY(hipStreamSynchronize(c->stream));
	HIP_TRY(hipFree(p));
	return YRTMI_OK;
}

int yrtmi_copy(yrtmi_ctx* c, void* dst, const void* src, int64_t bytes, int kind)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "copy");
	if (r)
		return r;
	if (bytes < 0 || (bytes > 0 && (dst == nullptr || src == nullptr)))
		return fail(YRTMI_EINVAL, "copy: bad arguments");
	if (bytes == 0)
		return YRTMI_OK;
	hipMemcpyKind k;
	switch (kind)
	{
	case 0:
		k = hipMemcpyHostToDevice;
		break;
	case 1:
		k = hipMemcpyDeviceToHost;
		break;
	case 2:
		k = hipMemcpyDeviceToDevice;
		break;
	default:
		return fail(YRTMI_EINVAL, "copy: kind must be 0, 1 or 2");
	}
	HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, k, c->stream));
	HIP_TRY(hipStreamSynchronize(c->stream));
	return YRTMI_OK;
}

int yrtmi_histo_subset_nonzero(yrtmi_ctx* c, const float* values, int S, int s,
                               uint32_t* bins, float* meas, int64_t* count)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "histo_subset_nonzero");
	if (r)
		return r;
	if (!c->histoReady)
		return fail(YRTMI_ESTATE, "histo_subset_nonzero: no yrtmi_histo_setup");
	if (count == nullptr || values == nullptr || bins == nullptr || meas == nullptr)
		return fail(YRTMI_EINVAL, "histo_subset_nonzero: null argument");
	const int64_t n = yrtmi_histo_subset_size(c, S, s);
	if (n < 0)
		return fail(YRTMI_EINVAL, "histo_subset_nonzero: bad subset index");
	*count = 0;
	if (n == 0)
		return YRTMI_OK;
	SubsetGeom g{};
	g.numPhi = c->numPhi;
	g.numR = c->numR;
	g.S = (uint32_t)S;
	g.s = (uint32_t)s;
	g.nps = c->numPhi / (uint32_t)S;
	if ((uint32_t)s + g.nps * (uint32_t)S < c->numPhi)
		g.nps += 1;
	const int64_t nb = (n + kCompactPerBlock - 1) / kCompactPerBlock;
	int* dCnt = nullptr;
	int64_t* dOff = nullptr;
	auto done = [&](int rc) {
		hipFree(dCnt);
		hipFree(dOff);
		return rc;
	};
	if (hipMalloc(&dCnt, sizeof(int) * nb) != hipSuccess ||
	    hipMalloc(&dOff, sizeof(int64_t) * nb) != hipSuccess)
		return done(fail(YRTMI_ENOMEM, "histo_subset_nonzero: scratch"));
	std::vector<int> hc(nb);
	std::vector<int64_t> ho(nb);
	hipError_t e = launchCountNonzero(values, g, n, dCnt, nb, c->stream);
	if (e == hipSuccess)
		e = hipMemcpyAsync(hc.data(), dCnt, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(c->stream);
	if (e != hipSuccess)
		return done(hipFail(e, "histo_subset_nonzero: count"));
	int64_t tot = 0;
	for (int64_t b = 0; b < nb; b++)
	{
		ho[b] = tot;
		tot += hc[b];
	}
	e = hipMemcpyAsync(dOff, ho.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, c->stream);
	if (e == hipSuccess)
		e = launchWriteNonzero(values, g, n, dOff, nb, bins, meas, c->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(c->stream);
	if (e != hipSuccess)
		return done(hipFail(e, "histo_subset_nonzero: write"));
	*count = tot;
	return done(YRTMI_OK);
}

int yrtmi_subset_interleaved(yrtmi_ctx* c, const void* src, int elem_bytes,
                             int64_t n_total, int S, int s, void* dst, int64_t* count)
{
	int r = setDev(c);
	if (r)
		return r;
	if (S <= 0 || s < 0 || s >= S || n_total < 0 ||
	    (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8))
		return fail(YRTMI_EINVAL, "subset_interleaved: bad arguments");
	const int64_t n = (n_total > s) ? (n_total - s + S - 1) / S : 0;
	if (count != nullptr)
		*count = n;
	if (n == 0)
		return YRTMI_OK;
	if (src == nullptr || dst == nullptr)
		return fail(YRTMI_EINVAL, "subset_interleaved: null buffer");
	HIP_TRY(launchInterleaved(src, elem_bytes, n, S, s, dst, c->stream));
	return YRTMI_OK;
}

int yrtmi_gather(yrtmi_ctx* c, const float* src, int64_t src_n, const uint32_t* idx,
                 int64_t n, int invert, float* out)
{
	int r = setDev(c);
	if (r)
		return r;
	if (n < 0 || src_n < 0 || (n > 0 && (src == nullptr || idx == nullptr || out == nullptr)))
		return fail(YRTMI_EINVAL, "gather: bad arguments");
	if (n == 0)
		return YRTMI_OK;
	HIP_TRY(launchGather(src, src_n, idx, n, invert, out, c->stream));
	return YRTMI_OK;
}

int yrtmi_histo_subset_gather(yrtmi_ctx* c, const float* h, int S, int s, int invert,
                              float* out)
{
	int r = setDev(c);
	if (r)
		return r;
	if (!c->histoReady)
		return fail(YRTMI_ESTATE, "histo_subset_gather: no yrtmi_histo_setup");
	const int64_t n = yrtmi_histo_subset_size(c, S, s);
	if (n < 0)
		return fail(YRTMI_EINVAL, "histo_subset_gather: bad subset index");
	if (n == 0)
		return YRTMI_OK;
	if (h == nullptr || out == nullptr)
		return fail(YRTMI_EINVAL, "histo_subset_gather: null buffer");
	SubsetGeom g{};
	g.numPhi = c->numPhi;
	g.numR = c->numR;
	g.S = (uint32_t)S;
	g.s = (uint32_t)s;
	g.nps = c->numPhi / (uint32_t)S;
	if ((uint32_t)s + g.nps * (uint32_t)S < c->numPhi)
		g.nps += 1;
	HIP_TRY(launchSubsetGather(h, g, n, invert, out, c->stream));
	return YRTMI_OK;
}

int yrtmi_axpby(yrtmi_ctx* c, int64_t n, float a, const float* x, float b, float* y)
{
	int r = setDev(c);
	if (r)
		return r;
	if (n < 0 || (n > 0 && y == nullptr))
		return fail(YRTMI_EINVAL, "axpby: bad arguments");
	if (n == 0)
		return YRTMI_OK;
	HIP_TRY(launchAxpby(n, a, x, b, y, c->stream));
	return YRTMI_OK;
}

int yrtmi_lor_order(yrtmi_ctx* c, const yrtmi_image* im, const yrtmi_lors* l, uint32_t* perm)
{
	int r = setDev(c);
	if (r)
		return r;
	r = notCapturing(c, "lor_order");
	if (r)
		return r;
	r = checkImage(im, "image");
	if (r)
		return r;
	if (l == nullptr)
		return fail(YRTMI_EINVAL, "lor_order: null LOR set");
	if (l->n >= (int64_t(1) << 32))
		return fail(YRTMI_EINVAL, "lor_order: n must be below 2^32");
	if (im->nz > 1024)
		return fail(YRTMI_EINVAL, "lor_order: nz above 1024");
	if (l->n > 0 && perm == nullptr)
		return fail(YRTMI_EINVAL, "lor_order: null output");
	DLors L;
	DScanner S;
	r = buildLors(c, l, L, S);
	if (r)
		return r;
	HIP_TRY(lorOrder(S, L, toDImage(im), perm, c->stream));
	return YRTMI_OK;
}

int yrtmi_permute(yrtmi_ctx* c, const void* src, int elem_bytes, const uint32_t* perm,
                  int64_t n, void* dst)
{
	int r = setDev(c);
	if (r)
		return r;
	if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
		return fail(YRTMI_EINVAL, "permute: elem_bytes must be 1, 2, 4 or 8");
	if (n < 0 || (n > 0 && (src == nullptr || perm == nullptr || dst == nullptr)))
		return fail(YRTMI_EINVAL, "permute: bad arguments");
	if (n > 0 && src == dst)
		return fail(YRTMI_EINVAL, "permute: dst must not be src");
	HIP_TRY(launchPermute(src, elem_bytes, perm, n, dst, c->stream));
	return YRTMI_OK;
}

}  // extern "C"

extern "C" {

int yrtmi_autotune(yrtmi_ctx* c, yrtmi_proj* p, const yrtmi_image* im,
                   const yrtmi_lors* l, const float* x, yrtmi_tuning* best)
{
	KArgs a;
	int r = prepare(c, p, im, l, a);
	if (r)
		return r;
	r = notCapturing(c, "autotune");
	if (r)
		return r;
	if (best == nullptr || x == nullptr)
		return fail(YRTMI_EINVAL, "autotune: null argument");
	*best = yrtmi_tuning{};
	if (l->n == 0)
		return YRTMI_OK;
	int cus = 256;
	(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
	float* v = nullptr;
	float* y = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	const int64_t ny = a.im.nvox * a.im.nt;
	auto done = [&](int rc) {
		hipFree(v);
		hipFree(y);
		if (e0)
			hipEventDestroy(e0);
		if (e1)
			hipEventDestroy(e1);
		return rc;
	};
	if (hipMalloc(&v, sizeof(float) * l->n) != hipSuccess ||
	    hipMalloc(&y, sizeof(float) * ny) != hipSuccess)
		return done(fail(YRTMI_ENOMEM, "autotune: temporaries"));
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
		return done(fail(YRTMI_EHIP, "autotune: events"));
	a.x = x;
	a.c.g = 1.0f;
	a.c.denomThr = 1e-35f;
	{
		const int rb = brickImage(c, a);
		if (rb)
			return done(rb);
	}
	// time fn(): one untimed run, then the mean of two timed runs
	auto timeIt = [&](auto&& fn, float& ms) -> hipError_t {
		hipError_t e = fn();
		if (e != hipSuccess)
			return e;
		e = hipEventRecord(e0, c->stream);
		for (int k = 0; k < 2 && e == hipSuccess; k++)
			e = fn();
		if (e == hipSuccess)
			e = hipEventRecord(e1, c->stream);
		if (e == hipSuccess)
			e = hipEventSynchronize(e1);
		if (e == hipSuccess)
			e = hipEventElapsedTime(&ms, e0, e1);
		ms *= 0.5f;
		return e;
	};
	// lanes per event split a LOR's rays: distance-driven LORs have one
	const int lpes[4] = {1, 2, 4, 8};
	const int nLpe = (p->d.type == YRTMI_PROJ_DD) ? 1 : 4;
	best->ratio_ms = 1e30f;
	for (int i = 0; i < 4; i++)
		best->ratio_ms_lpe[i] = -1.f;
	for (int i = 0; i < nLpe; i++)
	{
		float ms = 0.f;
		hipError_t e = timeIt([&]() { return launchEmRatio(a, lpes[i], v, c->stream); }, ms);
		if (e != hipSuccess)
			return done(hipFail(e, "autotune: EM ratio"));
		best->ratio_ms_lpe[i] = ms;
		if (ms < best->ratio_ms)
		{
			best->ratio_ms = ms;
			best->lanes_per_event = lpes[i];
		}
	}
	p->lpe = best->lanes_per_event;
	// v now holds the EM ratios of x: backproject them through each shape
	best->back_ms = 1e30f;
	for (int shape = 0; shape < kNumTileShapes; shape++)
	{
		yrtmi_plan_options o{0, shape, 0, 0, 0};
		yrtmi_plan* q = nullptr;
		best->back_ms_shape[shape][0] = best->back_ms_shape[shape][1] = -1.f;
		r = yrtmi_plan_create(c, p, im, l, &o, &q);
		if (r == YRTMI_EINVAL)  // this shape does not fit the image (tile count)
			continue;
		if (r)
			return done(r);
		// resident workgroups: 160 KB of LDS per CU over the tile's LDS
		// (16^3: 32 KB; 32x32x16: 128 KB; 32^3: 128 KB; 32x32x8: 64 KB)
		const int perCu = shape == 0 ? 5 : (shape == 3 ? 2 : 1);
		for (int mode = 0; mode < 2; mode++)
		{
			const int pers = mode ? perCu * cus : 0;
			float ms = 0.f;
			hipError_t e = timeIt(
			    [&]() { return launchTileBack(a, q->T, planRead(q), v, y, pers, c->stream); },
			    ms);
			if (e != hipSuccess)
			{
				yrtmi_plan_destroy(q);
				return done(hipFail(e, "autotune: tile backprojection"));
			}
			best->back_ms_shape[shape][mode] = ms;
			if (ms < best->back_ms)
			{
				best->back_ms = ms;
				best->tile_shape = shape;
				best->persistent = pers;
			}
		}
		yrtmi_plan_destroy(q);
	}
	if (best->back_ms >= 1e30f)
		return done(fail(YRTMI_EINVAL, "autotune: no tile shape fits the image"));
	return done(YRTMI_OK);
}

}  // extern "C"

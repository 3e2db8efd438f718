// Host build of the ray kernels (yrtmi_rays.hpp) for CPU tests: the full
// Siddon walk of one ray, its tile decomposition (the bin pass's tile
// entries, each re-entered as the tile kernel does) summed back per voxel,
// and the distance-driven sweep of one LOR (forward and backprojection, with
// TOF and the projection-space PSF), whole and cut into tiles.
#include "yrtmi_rays.hpp"

#include <array>
#include <map>
#include <vector>

using namespace yrtmi;

namespace
{
DImage image(const float* imp)
{
	DImage P{};
	P.nx = (int)imp[0];
	P.ny = (int)imp[1];
	P.nz = (int)imp[2];
	P.nt = 1;
	P.vx = imp[3];
	P.vy = imp[4];
	P.vz = imp[5];
	P.lx = imp[6];
	P.ly = imp[7];
	P.lz = imp[8];
	P.fovR = imp[9];
	P.nxy = P.nx * P.ny;
	P.nvox = (int64_t)P.nxy * P.nz;
	return P;
}

template <bool TOF, bool S64>
double tiles(const DImage& P, float3 a, float3 b, const DTof& T, float tofv, int TX, int TY,
             int TZ, int* out)
{
	std::map<int, double> full, tiled;
	SidRay r;
	out[0] = out[1] = out[2] = out[3] = 0;
	if (!sidSetup<TOF, S64>(P, a, b, T, tofv, r))
		return 0.0;
	int dx = 0, dy = 0, dz = 0;
	sidPlaneShift<S64>(P, r, dx, dy, dz);
	struct Ent
	{
		float a;
		int vx, vy, vz;
		bool first;
	};
	std::vector<Ent> ents;
	int last = -1;
	SidRay w = r;
	sidWalk<TOF, S64>(P, w, T, imageBox(P), true, 0, 0, 0,
	                  [&](int vx, int vy, int vz, int idx, float a0, float wt, uint32_t) {
		                  full[idx] += wt;
		                  out[1]++;
		                  const int t = ((vz / TZ) * 4096 + vy / TY) * 4096 + vx / TX;
		                  if (t != last)
		                  {
			                  ents.push_back({a0, vx, vy, vz, last < 0});
			                  last = t;
		                  }
	                  });
	out[0] = (int)ents.size();
	// the plan's ray-table entry of an FP32 set-up (kRayPrep)
	float4 st[2];
	sidStore(st, r);
	// the tiled walks, the FP32 set-up's records re-entered from the stored
	// set-up (kTileBack) and, beside them, from a fresh set-up: out[3]
	// counts voxels where the two differ at all
	std::map<int, double> fresh;
	for (const Ent& e : ents)
	{
		const int x0 = e.vx / TX * TX, y0 = e.vy / TY * TY, z0 = e.vz / TZ * TZ;
		const WalkBox B{x0, std::min(x0 + TX, P.nx), y0, std::min(y0 + TY, P.ny),
		                z0, std::min(z0 + TZ, P.nz), 1, P.nx, P.nxy, 0};
		// FP32 set-up: every record re-entered (the first at the ray's start
		// alpha and first voxel), as kTileBack does; FP64: the first record
		// walks from the entry
		const bool entry = S64 && e.first;
		for (int pass = 0; pass < 2; pass++)
		{
			SidRay q;
			if (pass == 0 && !S64)
				sidLoad(a, b, st[0], st[1], q);
			else
				sidSetup<TOF, S64>(P, a, b, T, tofv, q);
			if (!entry)
				sidRestart<S64>(P, q, e.a, e.vx, e.vy, e.vz, dx, dy, dz);
			std::map<int, double>& m = pass == 0 ? tiled : fresh;
			sidWalk<TOF, S64>(P, q, T, B, entry, e.vx, e.vy, e.vz,
			                  [&](int, int, int, int idx, float, float wt, uint32_t) {
				                  m[idx] += wt;
				                  if (wt > 0.f && pass == 0)
					                  out[2]++;
			                  });
		}
	}
	for (auto& kv : fresh)
		if (!tiled.count(kv.first) || tiled[kv.first] != kv.second)
			out[3]++;
	for (auto& kv : tiled)
		if (!fresh.count(kv.first))
			out[3]++;
	double mx = 0, md = 0;
	for (auto& kv : full)
		mx = std::max(mx, std::fabs(kv.second));
	for (auto& kv : full)
		md = std::max(md, std::fabs(kv.second - tiled[kv.first]));
	for (auto& kv : tiled)
		if (!full.count(kv.first))
			md = std::max(md, std::fabs(kv.second));
	return mx > 0 ? md / mx : md;
}
}  // namespace

extern "C" {

// Line integral of `img` along one ray (setup precision s64, TOF or not);
// *nseg = segments.
float hw_forward(const float* imp, const float* img, const float* p1, const float* p2,
                 int tof, int s64, const float* tofc, float tofv, int* nseg)
{
	const DImage P = image(imp);
	const DTof T{tofc[0], tofc[1], tofc[2], rcpNearest(tofc[0])};
	const float3 a = make_float3(p1[0], p1[1], p1[2]), b = make_float3(p2[0], p2[1], p2[2]);
	uint32_t c = 0, c2 = 0;
	float v;
	float* im = const_cast<float*>(img);
	if (tof && s64)
	{
		v = siddonRay<MODE_FWD, true, true>(P, im, a, b, 0.f, T, tofv, c);
		siddonRay<MODE_COUNT, true, true>(P, nullptr, a, b, 0.f, T, tofv, c2);
	}
	else if (tof)
	{
		v = siddonRay<MODE_FWD, true, false>(P, im, a, b, 0.f, T, tofv, c);
		siddonRay<MODE_COUNT, true, false>(P, nullptr, a, b, 0.f, T, tofv, c2);
	}
	else if (s64)
	{
		v = siddonRay<MODE_FWD, false, true>(P, im, a, b, 0.f, T, tofv, c);
		siddonRay<MODE_COUNT, false, true>(P, nullptr, a, b, 0.f, T, tofv, c2);
	}
	else
	{
		v = siddonRay<MODE_FWD, false, false>(P, im, a, b, 0.f, T, tofv, c);
		siddonRay<MODE_COUNT, false, false>(P, nullptr, a, b, 0.f, T, tofv, c2);
	}
	*nseg = (int)c2;
	return v;
}

// Tile decomposition of one ray: max over voxels of |full - tiled| / max
// |full|.  out[0] tiles entered, out[1] segments of the full walk, out[2]
// nonempty segments of the tiled walks, out[3] voxels where re-entering
// from the stored FP32 set-up differs from re-entering from a fresh one.
double hw_tiles(const float* imp, const float* p1, const float* p2, int tof, int s64,
                const float* tofc, float tofv, int TX, int TY, int TZ, int* out)
{
	const DImage P = image(imp);
	const DTof T{tofc[0], tofc[1], tofc[2], rcpNearest(tofc[0])};
	const float3 a = make_float3(p1[0], p1[1], p1[2]), b = make_float3(p2[0], p2[1], p2[2]);
	if (tof && s64)
		return tiles<true, true>(P, a, b, T, tofv, TX, TY, TZ, out);
	if (tof)
		return tiles<true, false>(P, a, b, T, tofv, TX, TY, TZ, out);
	if (s64)
		return tiles<false, true>(P, a, b, T, tofv, TX, TY, TZ, out);
	return tiles<false, false>(P, a, b, T, tofv, TX, TY, TZ, out);
}

// Bricked forward (siddonRay<MODE_FWD, ., ., BRICK>) of one ray over
// `brk` (the bricked copy of `img`) beside the x-major forward over `img`:
// out[0] = bricked, out[1] = x-major; returns the segments whose bricked
// index is not brickIndex of their voxel, or whose fetched-ahead index is
// not the next voxel's (when that voxel is in the image).
int hw_brick(const float* imp, const float* img, const float* brk, const float* p1,
             const float* p2, int tof, const float* tofc, float tofv, float* out)
{
	const DImage P = image(imp);
	const DTof T{tofc[0], tofc[1], tofc[2], rcpNearest(tofc[0])};
	const float3 a = make_float3(p1[0], p1[1], p1[2]), b = make_float3(p2[0], p2[1], p2[2]);
	uint32_t c = 0;
	float* bi = const_cast<float*>(brk);
	float* xi = const_cast<float*>(img);
	out[0] = tof ? siddonRay<MODE_FWD, true, false, true>(P, bi, a, b, 0.f, T, tofv, c)
	             : siddonRay<MODE_FWD, false, false, true>(P, bi, a, b, 0.f, T, tofv, c);
	out[1] = tof ? siddonRay<MODE_FWD, true, false>(P, xi, a, b, 0.f, T, tofv, c)
	             : siddonRay<MODE_FWD, false, false>(P, xi, a, b, 0.f, T, tofv, c);
	SidRay r;
	if (!(tof ? sidSetup<true, false>(P, a, b, T, tofv, r) : sidSetup<false, false>(P, a, b, T, tofv, r)))
		return 0;
	int bad = 0;
	std::vector<std::array<int, 4>> seg;
	auto fn = [&](int gx, int gy, int gz, uint32_t idx, float, float, uint32_t nidx) {
		if (idx != brickIndex(P.nx, P.ny, gx, gy, gz))
			bad++;
		seg.push_back({gx, gy, gz, (int)nidx});
	};
	if (tof)
		sidWalk<true, false, true>(P, r, T, imageBox(P), true, 0, 0, 0, fn, [](uint32_t) {});
	else
		sidWalk<false, false, true>(P, r, T, imageBox(P), true, 0, 0, 0, fn, [](uint32_t) {});
	for (size_t i = 0; i + 1 < seg.size(); i++)
		if ((uint32_t)seg[i][3] != brickIndex(P.nx, P.ny, seg[i + 1][0], seg[i + 1][1], seg[i + 1][2]))
			bad++;
	return bad;
}

// Distance-driven LOR (ddSetup + ddVisit): forward projection of `img`
// (returned) and backprojection of `val` added into `back` (may be null),
// the image swept whole; tile = {TX, TY, TZ} > 0 also sweeps it tile by tile
// (the tile kernel's ranges) into `tiled` (may be null).  psf: numS x K taps
// or null.  *nvox = footprint voxels visited.
float hw_dd(const float* imp, const float* img, const float* p1, const float* p2,
            const float* n1, const float* n2, int tof, const float* tofc, float tofv,
            float tz, float tt, const float* psf, int numS, int K, float sStep,
            float kSpacing, float val, float* back, const int* tile, float* tiled,
            int* nvox)
{
	const DImage P = image(imp);
	DProj pr{};
	pr.type = 1;
	pr.numRays = 1;
	pr.useTof = tof;
	pr.tof = DTof{tofc[0], tofc[1], tofc[2], rcpNearest(tofc[0])};
	pr.tz = tz;
	pr.tt = tt;
	std::vector<float> taps;
	std::vector<double> cum;
	if (psf != nullptr)
	{
		const int64_t nk = (int64_t)numS * K;
		std::vector<float> raw(2 * nk);
		for (int64_t i = 0; i < numS; i++)
			for (int64_t j = 0; j < K; j++)
			{
				raw[i * K + j] = psf[i * K + j];
				raw[nk + i * K + (K - 1 - j)] = psf[i * K + j];
			}
		cum.resize(2 * (size_t)numS * (K + 2));
		psfPrefixSums(raw.data(), numS, K, kSpacing, cum.data());
		psfPrefixSums(raw.data() + nk, numS, K, kSpacing, cum.data() + (size_t)numS * (K + 2));
		// the device layout: rows of K + 4, the taps between two zeros each side
		const int64_t KP = psfRowTaps(K), nkp = (int64_t)numS * KP;
		taps.assign(2 * nkp, 0.f);
		for (int64_t h = 0; h < 2; h++)
			for (int64_t i = 0; i < numS; i++)
				for (int64_t j = 0; j < K; j++)
					taps[h * nkp + i * KP + 2 + j] = raw[h * nk + i * K + j];
		pr.psf.k = taps.data();
		pr.psf.kflip = taps.data() + nkp;
		pr.psf.cum = cum.data();
		pr.psf.cumFlip = cum.data() + (size_t)numS * (K + 2);
		pr.psf.numS = numS;
		pr.psf.K = K;
		pr.psf.sStep = sStep;
		pr.psf.kSpacing = kSpacing;
		pr.psf.halfWidth = (float)((K + 1) / 2) * kSpacing;
		pr.psf.rks = rcpNearest(kSpacing);
	}
	const float3 a = make_float3(p1[0], p1[1], p1[2]), b = make_float3(p2[0], p2[1], p2[2]);
	const float3 na = make_float3(n1[0], n1[1], n1[2]), nb = make_float3(n2[0], n2[1], n2[2]);
	auto sweep = [&](bool flip, int xlo, int xhi, int ylo, int yhi, int zlo, int zhi,
	                 auto&& fn) {
		DDRay R;
		if (tof)
			ddSetup<true, false>(P, a, b, na, nb, pr, tofv, flip, R);
		else
			ddSetup<false, false>(P, a, b, na, nb, pr, tofv, flip, R);
		if (tof)
			ddVisit<true>(P, R, pr, xlo, xhi, ylo, yhi, zlo, zhi, fn);
		else
			ddVisit<false>(P, R, pr, xlo, xhi, ylo, yhi, zlo, zhi, fn);
		return R;
	};
	const int big = 1 << 30;
	float acc = 0.f;
	int cnt = 0;
	sweep(false, 0, big, 0, big, 0, big, [&](int io, int, int, int, float w) {
		acc = fmaf(w, img[io], acc);
		cnt++;
	});
	*nvox = cnt;
	if (back != nullptr)
		sweep(true, 0, big, 0, big, 0, big,
		      [&](int io, int, int, int, float w) { back[io] += val * w; });
	if (tiled != nullptr && tile != nullptr)
	{
		// the tiles of the image: slices [x0, x0 + TX) of the major axis,
		// minor axis [y0, y0 + TY), z [z0, z0 + TZ), as the tile kernel's
		// records restrict the sweep
		const int TX = tile[0], TY = tile[1], TZ = tile[2];
		for (int z0 = 0; z0 < P.nz; z0 += TZ)
			for (int y0 = 0; y0 < P.ny; y0 += TY)
				for (int x0 = 0; x0 < P.nx; x0 += TX)
				{
					DDRay R;
					if (tof)
						ddSetup<true, false>(P, a, b, na, nb, pr, tofv, true, R);
					else
						ddSetup<false, false>(P, a, b, na, nb, pr, tofv, true, R);
					const int mlo = R.flag_y ? y0 : x0, mhi = (R.flag_y ? std::min(y0 + TY, P.ny)
					                                                    : std::min(x0 + TX, P.nx)) - 1;
					const int nlo = R.flag_y ? x0 : y0, nhi = (R.flag_y ? std::min(x0 + TX, P.nx)
					                                                    : std::min(y0 + TY, P.ny)) - 1;
					auto fn = [&](int io, int, int, int, float w) { tiled[io] += val * w; };
					if (tof)
						ddVisit<true>(P, R, pr, mlo, mhi, nlo, nhi, z0,
						              std::min(z0 + TZ, P.nz) - 1, fn);
					else
						ddVisit<false>(P, R, pr, mlo, mhi, nlo, nhi, z0,
						               std::min(z0 + TZ, P.nz) - 1, fn);
				}
	}
	return acc;
}

}  // extern "C"

// Host build of the Siddon walk (yrtmi_rays.hpp) for CPU tests: the full
// walk of one ray, and its tile decomposition (the bin pass's tile entries,
// each re-entered as the tile kernel does) summed back per voxel.
#include "yrtmi_rays.hpp"

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
	out[0] = out[1] = out[2] = 0;
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
	                  [&](int vx, int vy, int vz, int idx, float a0, float wt) {
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
	for (const Ent& e : ents)
	{
		SidRay q;
		sidSetup<TOF, S64>(P, a, b, T, tofv, q);
		const int x0 = e.vx / TX * TX, y0 = e.vy / TY * TY, z0 = e.vz / TZ * TZ;
		const WalkBox B{x0, std::min(x0 + TX, P.nx), y0, std::min(y0 + TY, P.ny),
		                z0, std::min(z0 + TZ, P.nz), P.nx, P.nxy, 0};
		if (!e.first)
			sidRestart<S64>(P, q, e.a, e.vx, e.vy, e.vz, dx, dy, dz);
		sidWalk<TOF, S64>(P, q, T, B, e.first, e.vx, e.vy, e.vz,
		                  [&](int, int, int, int idx, float, float wt) {
			                  tiled[idx] += wt;
			                  if (wt > 0.f)
				                  out[2]++;
		                  });
	}
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
// nonempty segments of the tiled walks.
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

}  // extern "C"

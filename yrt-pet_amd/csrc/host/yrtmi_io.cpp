// File formats on either side of the projection path (see yrtmi_io.hpp).
#include "yrtmi_io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>

namespace yrtmi
{

namespace
{

size_t fileSize(std::ifstream& f)
{
	f.seekg(0, std::ios::end);
	const std::streamoff end = f.tellg();
	f.seekg(0, std::ios::beg);
	return end > 0 ? (size_t)end : 0;
}

std::string parentDir(const std::string& path)
{
	const size_t p = path.find_last_of('/');
	return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

bool endsWith(const std::string& s, const std::string& suffix)
{
	return s.size() >= suffix.size() &&
	       s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// raw-data arrays
// ---------------------------------------------------------------------------
RawdArray readRawd(const std::string& fname, const std::vector<uint64_t>& expectedDims)
{
	std::ifstream f(fname, std::ios::binary);
	if (!f)
		throw Error("The file given \"" + fname + "\" could not be opened");
	const size_t size = fileSize(f);
	int32_t magic = 0, ndim = 0;
	f.read(reinterpret_cast<char*>(&magic), 4);
	if (!f || magic != kRawdMagic)
		throw Error("The file given \"" + fname + "\" does not have a proper MAGIC_NUMBER");
	f.read(reinterpret_cast<char*>(&ndim), 4);
	if (!f || ndim <= 0 || ndim > 16)
		throw Error("The file given \"" + fname + "\" has an invalid number of dimensions");
	if (!expectedDims.empty() && (size_t)ndim != expectedDims.size())
		throw Error("The file given \"" + fname + "\" has " + std::to_string(ndim) +
		            " dimensions instead of the expected " +
		            std::to_string(expectedDims.size()));
	RawdArray a;
	a.dims.resize((size_t)ndim);
	f.read(reinterpret_cast<char*>(a.dims.data()), 8 * (size_t)ndim);
	if (!f)
		throw Error("The file given \"" + fname + "\" is truncated");
	uint64_t n = 1;
	for (int i = 0; i < ndim; i++)
	{
		if (!expectedDims.empty() && expectedDims[(size_t)i] != 0 &&
		    expectedDims[(size_t)i] != a.dims[(size_t)i])
			throw Error("The file given \"" + fname +
			            "\" has dimension sizes that do not match the expected sizes");
		n *= a.dims[(size_t)i];
	}
	const size_t header = 8 + 8 * (size_t)ndim;
	if (size < header || (size - header) / sizeof(float) < n)
		throw Error("The file given \"" + fname + "\" is shorter than its dimensions");
	a.data.resize(n);
	f.read(reinterpret_cast<char*>(a.data.data()), (std::streamsize)(n * sizeof(float)));
	if (!f)
		throw Error("Error reading \"" + fname + "\"");
	return a;
}

void writeRawd(const std::string& fname, const std::vector<uint64_t>& dims, const float* data)
{
	std::ofstream f(fname, std::ios::binary);
	if (!f)
		throw Error("The file given \"" + fname + "\" could not be opened");
	const int32_t magic = kRawdMagic, ndim = (int32_t)dims.size();
	uint64_t n = 1;
	for (uint64_t d : dims)
		n *= d;
	f.write(reinterpret_cast<const char*>(&magic), 4);
	f.write(reinterpret_cast<const char*>(&ndim), 4);
	f.write(reinterpret_cast<const char*>(dims.data()), (std::streamsize)(8 * dims.size()));
	f.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(float)));
	if (!f)
		throw Error("Error writing \"" + fname + "\"");
}

// ---------------------------------------------------------------------------
// list-mode file
// ---------------------------------------------------------------------------
ListModeFile readListModeFile(const std::string& fname, uint32_t numDets, bool flagTOF,
                              bool flagRandoms)
{
	std::ifstream f(fname, std::ios::binary);
	if (!f)
		throw Error("Error reading input file " + fname);
	const size_t size = fileSize(f);
	const size_t fields = 3 + (flagTOF ? 1 : 0) + (flagRandoms ? 1 : 0);
	const size_t evBytes = fields * 4;
	if (size == 0 || size % evBytes != 0)
		throw Error("Error: Input file has incorrect size (" + fname + ")");
	const size_t n = size / evBytes;
	ListModeFile lm;
	lm.timestamps.resize(n);
	lm.d1.resize(n);
	lm.d2.resize(n);
	if (flagTOF)
		lm.tof.resize(n);
	if (flagRandoms)
		lm.randoms.resize(n);
	// read in blocks of up to 2^24 events
	const size_t block = std::min<size_t>(n, size_t(1) << 24);
	std::vector<uint32_t> buf(block * fields);
	for (size_t e0 = 0; e0 < n; e0 += block)
	{
		const size_t m = std::min(block, n - e0);
		f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(m * evBytes));
		if (!f)
			throw Error("Error reading input file " + fname);
		for (size_t i = 0; i < m; i++)
		{
			const uint32_t* r = buf.data() + i * fields;
			const size_t e = e0 + i;
			if (r[1] >= numDets || r[2] >= numDets)
				throw Error("Detectors invalid in list-mode event " + std::to_string(e));
			lm.timestamps[e] = r[0];
			lm.d1[e] = r[1];
			lm.d2[e] = r[2];
			size_t k = 3;
			if (flagTOF)
				std::memcpy(&lm.tof[e], &r[k++], 4);
			if (flagRandoms)
				std::memcpy(&lm.randoms[e], &r[k++], 4);
		}
	}
	return lm;
}

void writeListModeFile(const std::string& fname, const ListModeFile& lm)
{
	const size_t n = lm.count();
	const bool tof = !lm.tof.empty(), rnd = !lm.randoms.empty();
	if (lm.d1.size() != n || lm.d2.size() != n || (tof && lm.tof.size() != n) ||
	    (rnd && lm.randoms.size() != n))
		throw Error("list-mode: field arrays of different lengths");
	std::ofstream f(fname, std::ios::binary);
	if (!f)
		throw Error("Could not open " + fname);
	const size_t fields = 3 + (tof ? 1 : 0) + (rnd ? 1 : 0);
	std::vector<uint32_t> row(fields);
	for (size_t e = 0; e < n; e++)
	{
		row[0] = lm.timestamps[e];
		row[1] = lm.d1[e];
		row[2] = lm.d2[e];
		size_t k = 3;
		if (tof)
			std::memcpy(&row[k++], &lm.tof[e], 4);
		if (rnd)
			std::memcpy(&row[k++], &lm.randoms[e], 4);
		f.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)(fields * 4));
	}
	if (!f)
		throw Error("Error writing " + fname);
}

// ---------------------------------------------------------------------------
// NIfTI-1
// ---------------------------------------------------------------------------
namespace
{

// NIfTI-1 header field offsets (nifti1.h, 348-byte header)
constexpr int kNiiHdr = 348;
enum NiiType : int16_t
{
	DT_UINT8 = 2,
	DT_INT16 = 4,
	DT_INT32 = 8,
	DT_FLOAT32 = 16,
	DT_FLOAT64 = 64,
	DT_INT8 = 256,
	DT_UINT16 = 512,
	DT_UINT32 = 768,
	DT_INT64 = 1024,
	DT_UINT64 = 1280,
};

template <class T>
T getAt(const unsigned char* h, int off)
{
	T v;
	std::memcpy(&v, h + off, sizeof(T));
	return v;
}

template <class T>
void putAt(unsigned char* h, int off, T v)
{
	std::memcpy(h + off, &v, sizeof(T));
}

struct GzFile
{
	gzFile f = nullptr;
	~GzFile()
	{
		if (f)
			gzclose(f);
	}
};

void gzReadAll(gzFile f, void* dst, size_t bytes, const std::string& fname)
{
	char* p = static_cast<char*>(dst);
	while (bytes > 0)
	{
		const unsigned chunk = (unsigned)std::min<size_t>(bytes, 1u << 30);
		const int r = gzread(f, p, chunk);
		if (r <= 0)
			throw Error("NIfTI: file truncated: " + fname);
		p += r;
		bytes -= (size_t)r;
	}
}

template <class T>
void convert(const unsigned char* src, size_t n, float slope, float inter, float* dst)
{
	for (size_t i = 0; i < n; i++)
	{
		T v;
		std::memcpy(&v, src + i * sizeof(T), sizeof(T));
		dst[i] = (float)v * slope + inter;
	}
}

// Quaternion + offsets to a 4x4 matrix (NIfTI-1 qform, method 2 of
// nifti1.h: R from (a, b, c, d), a = sqrt(1 - b^2 - c^2 - d^2); columns
// scaled by dx, dy and qfac * dz; translation qoffset).
void qformMatrix(float b, float c, float d, float qx, float qy, float qz, float dx, float dy,
                 float dz, float qfac, float m[3][4])
{
	double a = 1.0 - ((double)b * b + (double)c * c + (double)d * d);
	double bb = b, cc = c, dd = d;
	if (a < 1e-7)
	{
		// a rotation of 180 degrees: a = 0 and (b, c, d) normalised
		const double s = 1.0 / std::sqrt(bb * bb + cc * cc + dd * dd);
		bb *= s;
		cc *= s;
		dd *= s;
		a = 0.0;
	}
	else
		a = std::sqrt(a);
	const double xd = dx > 0 ? dx : 1.0, yd = dy > 0 ? dy : 1.0, zd0 = dz > 0 ? dz : 1.0;
	const double zd = (qfac < 0 ? -zd0 : zd0);
	m[0][0] = (float)((a * a + bb * bb - cc * cc - dd * dd) * xd);
	m[0][1] = (float)(2.0 * (bb * cc - a * dd) * yd);
	m[0][2] = (float)(2.0 * (bb * dd + a * cc) * zd);
	m[1][0] = (float)(2.0 * (bb * cc + a * dd) * xd);
	m[1][1] = (float)((a * a + cc * cc - bb * bb - dd * dd) * yd);
	m[1][2] = (float)(2.0 * (cc * dd - a * bb) * zd);
	m[2][0] = (float)(2.0 * (bb * dd - a * cc) * xd);
	m[2][1] = (float)(2.0 * (cc * dd + a * bb) * yd);
	m[2][2] = (float)((a * a + dd * dd - cc * cc - bb * bb) * zd);
	m[0][3] = qx;
	m[1][3] = qy;
	m[2][3] = qz;
}

}  // namespace

float originToOffset(float origin, float voxelSize, float length)
{
	return origin + 0.5f * length - 0.5f * voxelSize;
}

float offsetToOrigin(float off, float voxelSize, float length)
{
	return off - 0.5f * length + 0.5f * voxelSize;
}

NiftiImage readNifti(const std::string& fname)
{
	GzFile g;
	g.f = gzopen(fname.c_str(), "rb");  // reads plain files as well
	if (!g.f)
		throw Error("An error occurred while reading file " + fname);
	unsigned char h[kNiiHdr];
	gzReadAll(g.f, h, kNiiHdr, fname);
	if (getAt<int32_t>(h, 0) != kNiiHdr)
		throw Error("NIfTI: not a little-endian NIfTI-1 file: " + fname);
	if (std::memcmp(h + 344, "n+1", 4) != 0)
		throw Error("NIfTI: only single-file NIfTI-1 (.nii / .nii.gz) is read: " + fname);
	int16_t dim[8];
	float pixdim[8];
	for (int i = 0; i < 8; i++)
	{
		dim[i] = getAt<int16_t>(h, 40 + 2 * i);
		pixdim[i] = getAt<float>(h, 76 + 4 * i);
	}
	if (dim[0] != 3 && dim[0] != 4)
		throw Error("NIfTI Image's dim[0] is not 3 or 4");
	const int16_t dtype = getAt<int16_t>(h, 70);
	const float voxOffset = getAt<float>(h, 108);
	float slope = getAt<float>(h, 112), inter = getAt<float>(h, 116);
	if (slope == 0.f)  // nifti: a zero slope means no scaling
	{
		slope = 1.f;
		inter = 0.f;
	}
	const int units = (int)(h[123] & 0x07);
	const int16_t qcode = getAt<int16_t>(h, 252), scode = getAt<int16_t>(h, 254);
	float m[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
	if (scode > 0)
	{
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 4; c++)
				m[r][c] = getAt<float>(h, 280 + 16 * r + 4 * c);
	}
	else if (qcode > 0)
		qformMatrix(getAt<float>(h, 256), getAt<float>(h, 260), getAt<float>(h, 264),
		            getAt<float>(h, 268), getAt<float>(h, 272), getAt<float>(h, 276), pixdim[1],
		            pixdim[2], pixdim[3], pixdim[0], m);
	else
		std::cerr << "Warning: The NIfTI image file given does not have a qform or an sform."
		          << std::endl;
	// ImageOwned::adjustAffineMatrix: flip negative diagonal entries with
	// their translation; then no rotation is supported
	for (int r = 0; r < 3; r++)
		if (m[r][r] < 0)
		{
			m[r][r] = -m[r][r];
			m[r][3] = -m[r][3];
		}
	const float offd[6] = {m[0][1], m[0][2], m[1][0], m[1][2], m[2][0], m[2][1]};
	for (float e : offd)
		if (std::fabs(e) > 1e-4f)
			throw Error(
			    "The given NIfTI file contains a transformation matrix with a rotation, "
			    "which is currently unsupported");
	float v[3] = {pixdim[1], pixdim[2], pixdim[3]};
	// the reference's unit handling (Image.cpp:1091-1105), kept as it is
	if (units == 1)  // metres
		for (float& x : v)
			x = x / 1000.f;
	else if (units == 3)  // microns
		for (float& x : v)
			x = x * 1000.f;
	const int64_t nx = dim[1], ny = dim[2], nz = dim[3], nt = dim[0] == 4 ? dim[4] : 1;
	if (nx <= 0 || ny <= 0 || nz <= 0 || nt <= 0)
		throw Error("NIfTI: nonpositive dimension in " + fname);
	NiftiImage img;
	img.params = ImageParams::fromParams(nx, ny, nz, v[0], v[1], v[2], nt);
	img.params.off_x = originToOffset(m[0][3], v[0], v[0] * (float)nx);
	img.params.off_y = originToOffset(m[1][3], v[1], v[1] * (float)ny);
	img.params.off_z = originToOffset(m[2][3], v[2], v[2] * (float)nz);
	size_t bpv;
	switch (dtype)
	{
	case DT_UINT8:
	case DT_INT8:
		bpv = 1;
		break;
	case DT_INT16:
	case DT_UINT16:
		bpv = 2;
		break;
	case DT_INT32:
	case DT_UINT32:
	case DT_FLOAT32:
		bpv = 4;
		break;
	case DT_INT64:
	case DT_UINT64:
	case DT_FLOAT64:
		bpv = 8;
		break;
	default:
		throw Error("NIfTI: unsupported datatype " + std::to_string(dtype) + " in " + fname);
	}
	// skip to the data (extensions included)
	const long skip = (long)voxOffset - kNiiHdr;
	if (skip < 0)
		throw Error("NIfTI: bad vox_offset in " + fname);
	std::vector<unsigned char> pad((size_t)skip);
	if (skip > 0)
		gzReadAll(g.f, pad.data(), (size_t)skip, fname);
	const size_t n = (size_t)(nx * ny * nz * nt);
	std::vector<unsigned char> raw(n * bpv);
	gzReadAll(g.f, raw.data(), raw.size(), fname);
	img.data.resize(n);
	float* o = img.data.data();
	switch (dtype)
	{
	case DT_UINT8:
		convert<uint8_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_INT8:
		convert<int8_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_INT16:
		convert<int16_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_UINT16:
		convert<uint16_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_INT32:
		convert<int32_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_UINT32:
		convert<uint32_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_FLOAT32:
		convert<float>(raw.data(), n, slope, inter, o);
		break;
	case DT_INT64:
		convert<int64_t>(raw.data(), n, slope, inter, o);
		break;
	case DT_UINT64:
		convert<uint64_t>(raw.data(), n, slope, inter, o);
		break;
	default:
		convert<double>(raw.data(), n, slope, inter, o);
		break;
	}
	return img;
}

void writeNifti(const std::string& fname, const ImageParams& p, const float* data)
{
	unsigned char h[kNiiHdr + 4];
	std::memset(h, 0, sizeof h);
	putAt<int32_t>(h, 0, kNiiHdr);
	h[39] = 0;  // dim_info
	const bool is4D = p.nt > 1;
	const int16_t dim[8] = {(int16_t)(is4D ? 4 : 3), (int16_t)p.nx, (int16_t)p.ny,
	                        (int16_t)p.nz, (int16_t)p.nt, 1, 1, 1};
	for (int i = 0; i < 8; i++)
		putAt<int16_t>(h, 40 + 2 * i, dim[i]);
	putAt<int16_t>(h, 70, DT_FLOAT32);
	putAt<int16_t>(h, 72, 32);
	const float pixdim[8] = {0.f, p.vx, p.vy, p.vz, 1.f, 0.f, 0.f, 0.f};
	for (int i = 0; i < 8; i++)
		putAt<float>(h, 76 + 4 * i, pixdim[i]);
	putAt<float>(h, 108, 352.f);
	putAt<float>(h, 112, 1.f);  // scl_slope
	putAt<float>(h, 116, 0.f);  // scl_inter
	h[123] = (unsigned char)(2 | 8);  // mm, seconds
	const char desc[] = "yrtmi (MI355X) YRT-PET projection path";
	std::memcpy(h + 148, desc, sizeof desc);
	putAt<int16_t>(h, 252, 0);  // qform off
	putAt<int16_t>(h, 254, 1);  // sform: scanner anatomical
	const float srow[3][4] = {
	    {-p.vx, 0.f, 0.f, -offsetToOrigin(p.off_x, p.vx, p.length_x)},
	    {0.f, -p.vy, 0.f, -offsetToOrigin(p.off_y, p.vy, p.length_y)},
	    {0.f, 0.f, p.vz, offsetToOrigin(p.off_z, p.vz, p.length_z)}};
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 4; c++)
			putAt<float>(h, 280 + 16 * r + 4 * c, srow[r][c]);
	std::memcpy(h + 344, "n+1", 4);
	const size_t n = (size_t)(p.nx * p.ny * p.nz * p.nt);
	if (p.nx > 32767 || p.ny > 32767 || p.nz > 32767 || p.nt > 32767)
		throw Error("NIfTI-1: a dimension above 32767");
	if (endsWith(fname, ".gz"))
	{
		GzFile g;
		g.f = gzopen(fname.c_str(), "wb");
		if (!g.f)
			throw Error("Could not open " + fname);
		const char* src = reinterpret_cast<const char*>(data);
		size_t left = n * sizeof(float);
		bool ok = gzwrite(g.f, h, sizeof h) == (int)sizeof h;
		while (ok && left > 0)
		{
			const unsigned chunk = (unsigned)std::min<size_t>(left, 1u << 30);
			ok = gzwrite(g.f, src, chunk) == (int)chunk;
			src += chunk;
			left -= chunk;
		}
		if (!ok)
			throw Error("Error writing " + fname);
		return;
	}
	std::ofstream f(fname, std::ios::binary);
	if (!f)
		throw Error("Could not open " + fname);
	f.write(reinterpret_cast<const char*>(h), sizeof h);
	f.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(float)));
	if (!f)
		throw Error("Error writing " + fname);
}

// ---------------------------------------------------------------------------
// scanner JSON
// ---------------------------------------------------------------------------
namespace
{

// The subset of JSON a scanner definition uses: one object of string and
// number members (nested values are parsed and ignored).
struct JsonValue
{
	enum Kind
	{
		Null,
		Number,
		String,
		Bool,
		Other
	} kind = Null;
	double num = 0;
	std::string str;
};

class JsonParser
{
public:
	explicit JsonParser(const std::string& s) : m_s(s) {}

	std::map<std::string, JsonValue> object()
	{
		std::map<std::string, JsonValue> out;
		ws();
		expect('{');
		ws();
		if (peek() == '}')
		{
			m_i++;
			return out;
		}
		for (;;)
		{
			ws();
			const std::string key = string();
			ws();
			expect(':');
			out[key] = value();
			ws();
			const char c = next();
			if (c == '}')
				break;
			if (c != ',')
				fail("',' or '}' expected");
		}
		return out;
	}

private:
	const std::string& m_s;
	size_t m_i = 0;

	[[noreturn]] void fail(const char* what) const
	{
		throw Error(std::string("Error in Scanner JSON file parsing: ") + what + " at offset " +
		            std::to_string(m_i));
	}
	char peek() const { return m_i < m_s.size() ? m_s[m_i] : '\0'; }
	char next()
	{
		if (m_i >= m_s.size())
			fail("unexpected end");
		return m_s[m_i++];
	}
	void expect(char c)
	{
		if (next() != c)
			fail("unexpected character");
	}
	void ws()
	{
		while (m_i < m_s.size() && std::isspace((unsigned char)m_s[m_i]))
			m_i++;
	}
	std::string string()
	{
		expect('"');
		std::string out;
		for (;;)
		{
			char c = next();
			if (c == '"')
				return out;
			if (c == '\\')
			{
				c = next();
				switch (c)
				{
				case 'n':
					out += '\n';
					break;
				case 't':
					out += '\t';
					break;
				case 'r':
					out += '\r';
					break;
				case 'b':
					out += '\b';
					break;
				case 'f':
					out += '\f';
					break;
				case 'u':
				{
					unsigned v = 0;
					for (int k = 0; k < 4; k++)
					{
						const char h = next();
						v = v * 16 + (unsigned)(std::isdigit((unsigned char)h)
						                            ? h - '0'
						                            : (std::tolower((unsigned char)h) - 'a' + 10));
					}
					out += v < 128 ? (char)v : '?';
					break;
				}
				default:
					out += c;
				}
			}
			else
				out += c;
		}
	}
	JsonValue value()
	{
		ws();
		JsonValue v;
		const char c = peek();
		if (c == '"')
		{
			v.kind = JsonValue::String;
			v.str = string();
		}
		else if (c == '{' || c == '[')
		{
			// skip a nested value
			v.kind = JsonValue::Other;
			int depth = 0;
			bool inStr = false;
			do
			{
				const char d = next();
				if (inStr)
				{
					if (d == '\\')
						next();
					else if (d == '"')
						inStr = false;
				}
				else if (d == '"')
					inStr = true;
				else if (d == '{' || d == '[')
					depth++;
				else if (d == '}' || d == ']')
					depth--;
			} while (depth > 0);
		}
		else if (m_s.compare(m_i, 4, "true") == 0 || m_s.compare(m_i, 5, "false") == 0)
		{
			v.kind = JsonValue::Bool;
			v.num = m_s[m_i] == 't' ? 1 : 0;
			m_i += m_s[m_i] == 't' ? 4 : 5;
		}
		else if (m_s.compare(m_i, 4, "null") == 0)
			m_i += 4;
		else
		{
			const char* b = m_s.c_str() + m_i;
			char* e = nullptr;
			v.num = std::strtod(b, &e);
			if (e == b)
				fail("value expected");
			v.kind = JsonValue::Number;
			m_i += (size_t)(e - b);
		}
		return v;
	}
};

using JsonObject = std::map<std::string, JsonValue>;

JsonObject readJsonObject(const std::string& fname, const char* what)
{
	std::ifstream f(fname);
	if (!f)
		throw Error(std::string("Could not open ") + what + " file " + fname);
	std::stringstream ss;
	ss << f.rdbuf();
	const std::string text = ss.str();
	return JsonParser(text).object();
}

// util::getParam over aliases: the first name present, else the default
// (or the error when mandatory).  `given` reports whether one was present.
double jsonNumber(const JsonObject& j, std::initializer_list<const char*> names,
                  bool mandatory, double def, const std::string& msg, bool* given = nullptr)
{
	for (const char* n : names)
	{
		const auto it = j.find(n);
		if (it != j.end())
		{
			if (it->second.kind != JsonValue::Number)
				throw Error(std::string("JSON: ") + n + " is not a number");
			if (given)
				*given = true;
			return it->second.num;
		}
	}
	if (given)
		*given = false;
	if (mandatory)
		throw Error(msg.empty() ? std::string("Missing mandatory argument ") + *names.begin() +
		                              " in json file"
		                        : msg);
	return def;
}

}  // namespace

ImageParams readImageParamsJson(const std::string& fname)
{
	const JsonObject j = readJsonObject(fname, "ImageParams");
	const float version = (float)jsonNumber(j, {"VERSION", "GCIMAGEPARAMS_FILE_VERSION"}, true,
	                                        -1.0,
	                                        "Error in ImageParams file version : Version "
	                                        "unspecified");
	if (version > kImageParamsFileVersion + 1.1920929e-7f)
		throw Error("Error in ImageParams file version : Wrong version. Current version: 1.2");
	auto count = [&](const char* n) {
		const double v = jsonNumber(j, {n}, true, 0.0,
		                            std::string("Error in ImageParams file version : '") + n +
		                                "' unspecified");
		return (int64_t)v;
	};
	const int64_t nx = count("nx"), ny = count("ny"), nz = count("nz");
	const int64_t nt = (int64_t)jsonNumber(j, {"nt", "num_frames"}, false, 1.0, "");
	const float ox = (float)jsonNumber(j, {"off_x", "offset_x"}, false, 0.0, "");
	const float oy = (float)jsonNumber(j, {"off_y", "offset_y"}, false, 0.0, "");
	const float oz = (float)jsonNumber(j, {"off_z", "offset_z"}, false, 0.0, "");
	// ImageParams::readLengthFromJSON: v * n when the voxel size is given
	auto length = [&](const char* lname, const char* vname, int64_t n) {
		bool given = false;
		const float v = (float)jsonNumber(j, {vname}, false, -1.0, "", &given);
		if (given)
			return v * (float)n;
		return (float)jsonNumber(j, {lname}, true, -1.0,
		                         "You need to specify either the voxel size (vx, vy or vz) "
		                         "or the length (length_x, length_y or length_z) for all "
		                         "three dimensions.");
	};
	const float lx = length("length_x", "vx", nx), ly = length("length_y", "vy", ny),
	            lz = length("length_z", "vz", nz);
	return ImageParams(nx, ny, nz, lx, ly, lz, ox, oy, oz, nt);
}

void writeImageParamsJson(const std::string& fname, const ImageParams& p)
{
	std::ofstream out(fname);
	if (!out)
		throw Error("Could not open " + fname);
	out << std::setprecision(9) << "{\"VERSION\":" << kImageParamsFileVersion
	    << ",\"nt\":" << p.nt << ",\"nx\":" << p.nx << ",\"ny\":" << p.ny
	    << ",\"nz\":" << p.nz << ",\"off_x\":" << p.off_x << ",\"off_y\":" << p.off_y
	    << ",\"off_z\":" << p.off_z << ",\"vx\":" << p.vx << ",\"vy\":" << p.vy
	    << ",\"vz\":" << p.vz << "}" << std::endl;
	if (!out)
		throw Error("Error writing " + fname);
}

Scanner readScannerJson(const std::string& fname)
{
	const JsonObject j = readJsonObject(fname, "scanner definition");
	auto num = [&](std::initializer_list<const char*> names, bool mandatory, double def,
	               const std::string& msg) -> double {
		return jsonNumber(j, names, mandatory, def, msg);
	};
	auto str = [&](const char* name, bool mandatory, const std::string& msg,
	               bool& given) -> std::string {
		const auto it = j.find(name);
		given = it != j.end();
		if (!given)
		{
			if (mandatory)
				throw Error(msg);
			return std::string();
		}
		if (it->second.kind != JsonValue::String)
			throw Error(std::string("Scanner JSON: ") + name + " is not a string");
		return it->second.str;
	};
	const float version =
	    (float)num({"VERSION"}, true, 0.0, "Missing VERSION in scanner definition file");
	bool given = false;
	const std::string name =
	    str("scannerName", true, "Missing scanner name in scanner definition file", given);
	bool coordGiven = false, maskGiven = false;
	const std::string detCoord = str("detCoord", false, "", coordGiven);
	(void)str("detMask", false, "", maskGiven);
	const float axialFOV = (float)num({"axialFOV"}, true, 0.0,
	                                  "Missing Axial Field Of View value in scanner definition file");
	const float cTrans = (float)num({"crystalSize_trans"}, true, 0.0,
	                                "Missing transaxial crystal size in scanner definition file");
	const float cZ = (float)num({"crystalSize_z"}, true, 0.0,
	                            "Missing z-axis crystal size in scanner definition file");
	const float depth = (float)num({"crystalDepth"}, true, 0.0,
	                               "Missing crystal depth in scanner definition file");
	const float radius = (float)num({"scannerRadius"}, true, 0.0,
	                                "Missing scanner radius in scanner definition file");
	const double dpr = num({"dets_per_ring", "detsPerRing"}, true, 0.0, "");
	const double rings = num({"num_rings", "numRings"}, true, 0.0, "");
	const double doi = num({"num_doi", "numDOI"}, true, 0.0, "");
	const double mrd = num({"max_ring_diff", "maxRingDiff"}, true, 0.0, "");
	const double mad = num({"min_ang_diff", "minAngDiff"}, true, 1.0, "");
	const double dpb = num({"dets_per_block", "detsPerBlock"}, false, 1.0, "");
	if (version > kScannerFileVersion + 1.1920929e-7f)
		throw Error("Scanner JSON version too recent, the current version is 3.2");
	if (maskGiven)
		throw Error("Scanner JSON: detector masks (detMask) are not supported");
	for (double v : {dpr, rings, doi, mrd, mad, dpb})
		if (!(v >= 0) || v != std::floor(v))
			throw Error("Scanner JSON: counts must be nonnegative integers");
	if (mrd >= rings)
		throw Error("Maximum Ring difference has to be lower than the number of rings");
	if (mad <= 0)
		throw Error("Minimum angle difference has to be non-zero");
	Scanner sc(name, axialFOV, cZ, cTrans, depth, radius, (size_t)dpr, (size_t)rings,
	           (size_t)doi, (size_t)mrd, (size_t)mad, (size_t)dpb);
	if (coordGiven)
	{
		std::vector<float> pos, ori;
		readDetCoord(parentDir(fname) + "/" + detCoord, pos, ori);
		if (pos.size() / 3 != sc.getNumDets())
			throw Error(
			    "The number of detectors given by the LUT file does not match the scanner's "
			    "characteristics. Namely, (num_doi * num_rings * dets_per_ring) does not equal "
			    "the size of the LUT");
		sc.setDetectorLUT(pos, ori);
	}
	return sc;
}

// ---------------------------------------------------------------------------
// detector coordinates
// ---------------------------------------------------------------------------
void readDetCoord(const std::string& fname, std::vector<float>& pos, std::vector<float>& ori)
{
	std::ifstream f(fname, std::ios::binary);
	if (!f)
		throw Error("Error reading input file " + fname);
	const size_t size = fileSize(f);
	if (size == 0 || size % 4 != 0 || (size / 4) % 6 != 0)
		throw Error("Error: Input file has incorrect size (" + fname + ")");
	const size_t n = size / 24;
	std::vector<float> buf(6 * n);
	f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)size);
	if (!f)
		throw Error("Error reading input file " + fname);
	pos.resize(3 * n);
	ori.resize(3 * n);
	for (size_t i = 0; i < n; i++)
		for (int k = 0; k < 3; k++)
		{
			pos[3 * i + k] = buf[6 * i + k];
			ori[3 * i + k] = buf[6 * i + 3 + k];
		}
}

void writeDetCoord(const std::string& fname, const std::vector<float>& pos,
                   const std::vector<float>& ori)
{
	if (pos.size() != ori.size() || pos.size() % 3 != 0)
		throw Error("detector coordinates: position and orientation tables differ");
	const size_t n = pos.size() / 3;
	std::vector<float> buf(6 * n);
	for (size_t i = 0; i < n; i++)
		for (int k = 0; k < 3; k++)
		{
			buf[6 * i + k] = pos[3 * i + k];
			buf[6 * i + 3 + k] = ori[3 * i + k];
		}
	std::ofstream f(fname, std::ios::binary);
	f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * 4));
	if (!f)
		throw Error("Error writing " + fname);
}

// ---------------------------------------------------------------------------
// image-space PSF file
// ---------------------------------------------------------------------------
void readImagePsf(const std::string& fname, std::vector<float>& kx, std::vector<float>& ky,
                  std::vector<float>& kz)
{
	std::ifstream f(fname);
	if (!f)
		throw Error("Error opening file '" + fname + "'");
	std::vector<std::vector<float>> rows;
	std::string line;
	while (std::getline(f, line))
	{
		std::replace(line.begin(), line.end(), ',', ' ');
		std::stringstream ss(line);
		std::vector<float> row;
		float v;
		while (ss >> v)
			row.push_back(v);
		rows.push_back(row);
	}
	if (rows.size() < 4 || rows[3].size() < 3)
		throw Error("Image-space PSF file: three kernel rows and a row of sizes expected");
	std::vector<float>* out[3] = {&kx, &ky, &kz};
	for (int a = 0; a < 3; a++)
	{
		const int n = (int)rows[3][a];
		if (n <= 0 || n % 2 == 0)
			throw Error("Kernel size must be odd");
		if ((int)rows[a].size() < n)
			throw Error("Image-space PSF file: a kernel row is shorter than its size");
		out[a]->assign(rows[a].begin(), rows[a].begin() + n);
	}
}

// ---------------------------------------------------------------------------
// motion and dynamic framing
// ---------------------------------------------------------------------------
std::vector<MotionRecord> readMotionCsv(const std::string& fname)
{
	std::ifstream f(fname);
	if (!f)
		throw Error("Failed to open file: " + fname);
	std::vector<MotionRecord> out;
	std::string line;
	int lineNumber = 0;
	auto sep = [](const char*& p, const char* e) {
		while (p < e && (*p == ' ' || *p == ','))
			++p;
	};
	while (std::getline(f, line))
	{
		++lineNumber;
		if (line.empty())
			continue;
		const char* p = line.c_str();
		const char* e = p + line.size();
		MotionRecord r{};
		uint32_t t = 0;
		auto [p0, ec0] = std::from_chars(p, e, t);
		if (ec0 != std::errc())
		{
			std::cerr << "Line " << lineNumber << ": invalid timestamp\n";
			continue;
		}
		r.timestamp = t;
		p = p0;
		bool failed = false;
		for (int i = 0; i < 13 && !failed; i++)
		{
			sep(p, e);
			if (p >= e)
			{
				std::cerr << "Line " << lineNumber
				          << (i < 12 ? ": too few values\n" : ": missing error value\n");
				failed = true;
				break;
			}
			float v;
			auto [np, ec] = std::from_chars(p, e, v);
			if (ec != std::errc())
			{
				std::cerr << "Line " << lineNumber
				          << (i < 12 ? ": invalid float\n" : ": invalid error value\n");
				failed = true;
				break;
			}
			if (i < 12)
				r.transform[i] = v;
			else
				r.error = v;
			p = np;
		}
		if (failed)
			continue;
		while (p < e && std::isspace((unsigned char)*p))
			++p;
		if (p != e)
		{
			std::cerr << "Line " << lineNumber << ": too many values\n";
			continue;
		}
		out.push_back(r);
	}
	return out;
}

void writeMotionCsv(const std::string& fname, const std::vector<MotionRecord>& motion)
{
	std::ofstream out(fname);
	if (!out)
		throw Error("Failed to open output file: " + fname);
	out << std::fixed << std::setprecision(9);
	for (const MotionRecord& r : motion)
	{
		out << r.timestamp;
		for (int i = 0; i < 12; i++)
			out << ',' << r.transform[i];
		out << ',' << r.error << '\n';
	}
	out << std::flush;
	if (!out)
		throw Error("Error writing " + fname);
}

std::vector<uint32_t> readDynamicFraming(const std::string& fname)
{
	std::ifstream in(fname);
	if (!in)
		throw Error("Error opening file \"" + fname + "\"");
	std::vector<uint32_t> out;
	uint32_t t;
	while (in >> t)
		out.push_back(t);
	return out;
}

void writeDynamicFraming(const std::string& fname, const std::vector<uint32_t>& bounds)
{
	std::ofstream out(fname);
	if (!out)
		throw Error("Error opening file \"" + fname + "\"");
	for (uint32_t t : bounds)
		out << t << "\n";
	if (!out)
		throw Error("Error writing " + fname);
}

std::vector<uint8_t> motionFrames(const std::vector<uint32_t>& ts,
                                  const std::vector<MotionRecord>& motion)
{
	if (motion.empty())
		throw Error("motion: no motion frames");
	if (motion.size() > 255)
		throw Error("motion: at most 255 motion frames");
	const size_t n = ts.size(), nf = motion.size();
	std::vector<uint8_t> out(n, 255);
	size_t e = 0;
	while (e < n && (int64_t)ts[e] < motion[0].timestamp)
		e++;  // before the first frame: unresolvable (255)
	size_t f = 0;
	for (; f + 1 < nf; f++)
		while (e < n && (int64_t)ts[e] < motion[f + 1].timestamp)
			out[e++] = (uint8_t)f;
	for (; e < n; e++)
		out[e] = (uint8_t)f;
	return out;
}

std::vector<int32_t> dynamicFrames(const std::vector<uint32_t>& ts,
                                   const std::vector<uint32_t>& bounds)
{
	if (bounds.size() < 2)
		throw Error("dynamic framing: at least one frame (two timestamps) is needed");
	const size_t n = ts.size(), nf = bounds.size() - 1;
	std::vector<int32_t> out(n, -1);
	size_t e = 0;
	while (e < n && ts[e] < bounds[0])
		e++;
	for (size_t f = 0; f < nf; f++)
		while (e < n && ts[e] < bounds[f + 1])
			out[e++] = (int32_t)f;
	return out;
}

}  // namespace yrtmi

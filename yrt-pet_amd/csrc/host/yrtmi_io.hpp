// File formats on either side of the projection path: the reference's
// list-mode file, raw-data arrays (histograms), NIfTI images, scanner JSON
// definitions, detector coordinate files and motion / dynamic framing.  Host
// code (std C++ and zlib); the readers return host arrays that the device
// classes of yrtmi_host.hpp upload.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "yrtmi_host.hpp"

namespace yrtmi
{

// ---- raw-data arrays (Array.hpp:37,167-250; pyyrtpet DataFileRawd) --------
// int32 magic 732174000, int32 number of dimensions, uint64 sizes (slowest
// first), then the values in row-major order.
constexpr int32_t kRawdMagic = 732174000;

struct RawdArray
{
	std::vector<uint64_t> dims;
	std::vector<float> data;
};

// Reads a float array; `expectedDims` (empty: any) must have the file's rank
// and, where nonzero, its sizes (Array::readFromFile's checks).
RawdArray readRawd(const std::string& fname, const std::vector<uint64_t>& expectedDims = {});
void writeRawd(const std::string& fname, const std::vector<uint64_t>& dims, const float* data);

// ---- list-mode file (ListModeLUT.cpp:224-320; list-mode_file.md) --------
// Per event: uint32 timestamp (ms), uint32 detector 1, uint32 detector 2,
// then [float32 TOF t2 - t1 (ps)] and [float32 randoms estimate (cps)].
struct ListModeFile
{
	std::vector<uint32_t> timestamps, d1, d2;
	std::vector<float> tof, randoms;  // empty unless the flag is set
	size_t count() const { return timestamps.size(); }
};

// Rejects a file whose size is not a whole number of events and any event
// with a detector at or above numDets, as the reference does.
ListModeFile readListModeFile(const std::string& fname, uint32_t numDets, bool flagTOF,
                              bool flagRandoms);
void writeListModeFile(const std::string& fname, const ListModeFile& lm);

// ---- NIfTI-1 images (Image.cpp:520-592 write, 1048-1232 read) ------------
// .nii or .nii.gz (single file).  Written as FLOAT32 with an sform
// diag(-vx, -vy, vz) and origin (-o_x, -o_y, o_z), o = off - length/2 +
// v/2 per axis, qform off, millimetres; read back from the sform (else the
// qform, else identity), negative diagonal entries flipped with their
// translation, rotations rejected, any of the reference's voxel types
// converted to float with scl_slope / scl_inter.
struct NiftiImage
{
	ImageParams params;
	std::vector<float> data;  // nt x nz x ny x nx
};

NiftiImage readNifti(const std::string& fname);
void writeNifti(const std::string& fname, const ImageParams& params, const float* data);
// Image::originToOffset / offsetToOrigin (Image.cpp:1302-1310).
float originToOffset(float origin, float voxelSize, float length);
float offsetToOrigin(float off, float voxelSize, float length);

// ---- scanner definition (Scanner.cpp:222-325) ------------------------------
// JSON with VERSION (<= 3.2), scannerName, axialFOV, crystalSize_trans,
// crystalSize_z, crystalDepth, scannerRadius, dets_per_ring|detsPerRing,
// num_rings|numRings, num_doi|numDOI, max_ring_diff|maxRingDiff,
// min_ang_diff|minAngDiff, optional dets_per_block|detsPerBlock (1) and
// detCoord (a detector coordinate file, relative to the JSON's folder).
// A detector mask (detMask) is not supported and is rejected.
constexpr float kScannerFileVersion = 3.2f;
Scanner readScannerJson(const std::string& fname);

// ---- image parameters (ImageParams.cpp:285-400) ---------------------------
// JSON with VERSION (<= 1.2), nx, ny, nz, optional nt|num_frames (1),
// off_x|offset_x ... (0), and per axis vx (length = vx * nx) or length_x; the
// voxel size is then length / n, as ImageParams::setup derives it.  Written
// with VERSION, nx, ny, nz, vx, vy, vz, off_x, off_y, off_z, nt.
constexpr float kImageParamsFileVersion = 1.2f;
ImageParams readImageParamsJson(const std::string& fname);
void writeImageParamsJson(const std::string& fname, const ImageParams& p);

// ---- detector coordinate file (DetCoord.cpp:263-320) ----------------------
// Six float32 per detector: x, y, z, then the orientation x, y, z.
void readDetCoord(const std::string& fname, std::vector<float>& pos, std::vector<float>& ori);
void writeDetCoord(const std::string& fname, const std::vector<float>& pos,
                   const std::vector<float>& ori);

// ---- image-space PSF file (OperatorPsf::readFromFileInternal,
// OperatorPsf.cpp:197-253) ----------------------------------------------------
// Rows of comma- or space-separated values: the x, y and z taps, then their
// three (odd) lengths.
void readImagePsf(const std::string& fname, std::vector<float>& kx, std::vector<float>& ky,
                  std::vector<float>& kz);

// ---- motion (LORMotion.cpp:110-190; motion_file.md) -----------------------
// One line per motion frame: t, r00, r01, r02, tx, r10, r11, r12, ty, r20,
// r21, r22, tz, e (comma- or space-separated); t the frame's starting
// timestamp (ms).  Lines that do not parse are skipped with a message, as the
// reference does.
struct MotionRecord
{
	int64_t timestamp;
	float transform[12];  // r00 r01 r02 tx r10 r11 r12 ty r20 r21 r22 tz
	float error;
};
std::vector<MotionRecord> readMotionCsv(const std::string& fname);
// LORMotion::writeToFile (LORMotion.cpp:209-231): the same fields, comma
// separated, floats fixed with 9 decimals.
void writeMotionCsv(const std::string& fname, const std::vector<MotionRecord>& motion);
// DynamicFraming::readFromFile / writeToFile (DynamicFraming.cpp:114-137):
// whitespace-separated timestamps, the frames' starts then the last stop.
std::vector<uint32_t> readDynamicFraming(const std::string& fname);
void writeDynamicFraming(const std::string& fname, const std::vector<uint32_t>& bounds);
// Motion frame of every event (ListMode::addLORMotion, ListMode.cpp:95-125):
// frame k for start_k <= t < start_k+1, the last frame for every later
// event; events before the first frame get 255 (an unresolvable motion
// frame: the projectors skip them).  Timestamps in acquisition order.
std::vector<uint8_t> motionFrames(const std::vector<uint32_t>& timestamps,
                                  const std::vector<MotionRecord>& motion);
// Dynamic frame of every event (ListMode::addDynamicFraming,
// ListMode.cpp:128-180): `bounds` holds the frames' start timestamps and the
// last frame's stop; events outside [bounds[0], bounds.back()) get -1.
std::vector<int32_t> dynamicFrames(const std::vector<uint32_t>& timestamps,
                                   const std::vector<uint32_t>& bounds);

}  // namespace yrtmi

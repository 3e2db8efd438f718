// pybind11 bindings of the file formats (yrtmi_io.hpp), with the names the
// reference's pyyrtpet uses for them: Scanner(json), ListModeLUTOwned
// (src/datastruct/projection/ListModeLUT.cpp:20-110), image and histogram
// readFromFile / writeToFile (src/datastruct/image/Image.cpp,
// src/datastruct/projection/Histogram3D.cpp:32,122), LORMotion and
// DynamicFraming files, and DataFileRawd's format (python/pyyrtpet/
// _pyyrtpet.py:9-78).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstring>

#include "yrtmi_io.hpp"

namespace py = pybind11;
using namespace pybind11::literals;
using namespace yrtmi;

namespace
{

template <typename T>
using CArray = py::array_t<T, py::array::c_style | py::array::forcecast>;

template <typename T>
py::array_t<T> toNumpy(const std::vector<T>& v)
{
	py::array_t<T> out((py::ssize_t)v.size());
	if (!v.empty())
		std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(T));
	return out;
}

template <typename T>
std::vector<T> toVector(const CArray<T>& a)
{
	return std::vector<T>(a.data(), a.data() + a.size());
}

py::array_t<float> imageArray(const NiftiImage& im)
{
	const ImageParams& p = im.params;
	py::array_t<float> out({(py::ssize_t)p.nt, (py::ssize_t)p.nz, (py::ssize_t)p.ny,
	                        (py::ssize_t)p.nx});
	std::memcpy(out.mutable_data(), im.data.data(), im.data.size() * sizeof(float));
	return out;
}

}  // namespace

void bindIO(py::module_& m, py::class_<Scanner>& scanner, py::class_<ImageParams>& imageParams,
            py::class_<ImageDevice>& image, py::class_<Histogram3DDevice>& histo)
{
	// ---- raw-data arrays
	m.def(
	    "readRawd",
	    [](const std::string& fname) {
		    RawdArray a = readRawd(fname);
		    std::vector<py::ssize_t> shape(a.dims.begin(), a.dims.end());
		    py::array_t<float> out(shape);
		    std::memcpy(out.mutable_data(), a.data.data(), a.data.size() * sizeof(float));
		    return out;
	    },
	    "fname"_a);
	m.def(
	    "writeRawd",
	    [](const std::string& fname, CArray<float> a) {
		    std::vector<uint64_t> dims;
		    for (py::ssize_t i = 0; i < a.ndim(); i++)
			    dims.push_back((uint64_t)a.shape(i));
		    writeRawd(fname, dims, a.data());
	    },
	    "fname"_a, "data"_a);

	// ---- image parameters file
	py::class_<ImageParams>& ip = imageParams;
	ip.def(py::init(&readImageParamsJson), "fname"_a)
	    .def("serialize", [](const ImageParams& p, const std::string& f) {
		    writeImageParamsJson(f, p);
	    }, "fname"_a);

	// ---- scanner definition and detector coordinates
	scanner.def(py::init(&readScannerJson), "fname"_a)
	    .def(
	        "setDetectorLUT",
	        [](Scanner& s, CArray<float> pos, CArray<float> ori) {
		        s.setDetectorLUT(toVector(pos), toVector(ori));
	        },
	        "positions"_a, "orientations"_a)
	    .def("getDetectorLUT", [](const Scanner& s) {
		    const size_t n = s.getNumDets();
		    py::array_t<float> p({(py::ssize_t)n, (py::ssize_t)3}), o({(py::ssize_t)n, (py::ssize_t)3});
		    std::memcpy(p.mutable_data(), s.lutPositions().data(), 3 * n * sizeof(float));
		    std::memcpy(o.mutable_data(), s.lutOrientations().data(), 3 * n * sizeof(float));
		    return py::make_tuple(p, o);
	    });
	m.def(
	    "writeDetCoord",
	    [](const std::string& fname, CArray<float> pos, CArray<float> ori) {
		    writeDetCoord(fname, toVector(pos), toVector(ori));
	    },
	    "fname"_a, "positions"_a, "orientations"_a);

	// ---- list-mode file (host copy of the events)
	py::class_<ListModeFile>(m, "ListModeLUTOwned")
	    .def(py::init([](const Scanner& sc, const std::string& fname, bool flagTOF,
	                     bool flagRandoms) {
		         return new ListModeFile(
		             readListModeFile(fname, (uint32_t)sc.getNumDets(), flagTOF, flagRandoms));
	         }),
	         "scanner"_a, "fname"_a, "flag_tof"_a = false, "flag_randoms"_a = false)
	    .def(py::init([](CArray<uint32_t> ts, CArray<uint32_t> d1, CArray<uint32_t> d2,
	                     const py::object& tof, const py::object& randoms) {
		         auto* lm = new ListModeFile();
		         lm->timestamps = toVector(ts);
		         lm->d1 = toVector(d1);
		         lm->d2 = toVector(d2);
		         if (!tof.is_none())
			         lm->tof = toVector(tof.cast<CArray<float>>());
		         if (!randoms.is_none())
			         lm->randoms = toVector(randoms.cast<CArray<float>>());
		         const size_t n = lm->timestamps.size();
		         if (lm->d1.size() != n || lm->d2.size() != n ||
		             (!lm->tof.empty() && lm->tof.size() != n) ||
		             (!lm->randoms.empty() && lm->randoms.size() != n))
		         {
			         delete lm;
			         throw py::value_error("list-mode arrays of different lengths");
		         }
		         return lm;
	         }),
	         "timestamps"_a, "d1"_a, "d2"_a, "tof"_a = py::none(), "randoms"_a = py::none())
	    .def("count", &ListModeFile::count)
	    .def("hasTOF", [](const ListModeFile& l) { return !l.tof.empty(); })
	    .def("hasRandomsEstimates", [](const ListModeFile& l) { return !l.randoms.empty(); })
	    .def("getTimestampArray", [](const ListModeFile& l) { return toNumpy(l.timestamps); })
	    .def("getDetector1Array", [](const ListModeFile& l) { return toNumpy(l.d1); })
	    .def("getDetector2Array", [](const ListModeFile& l) { return toNumpy(l.d2); })
	    .def("getTOFArray",
	         [](const ListModeFile& l) {
		         if (l.tof.empty())
			         throw Error("The ListModeLUT object does not hold TOF information");
		         return toNumpy(l.tof);
	         })
	    .def("getRandomsEstimatesArray",
	         [](const ListModeFile& l) {
		         if (l.randoms.empty())
			         throw Error("The ListModeLUT object does not hold randoms estimates");
		         return toNumpy(l.randoms);
	         })
	    .def("writeToFile", [](const ListModeFile& l, const std::string& f) { writeListModeFile(f, l); },
	         "filename"_a);

	// device list-mode data from a list-mode file's events
	m.def(
	    "ListModeDeviceOwned",
	    [](const Context& c, const ListModeFile& lm, const py::object& frame,
	       const py::object& motion) {
		    const size_t n = lm.count();
		    std::vector<int32_t> f;
		    std::vector<uint8_t> mo;
		    if (!frame.is_none())
			    f = toVector(frame.cast<CArray<int32_t>>());
		    if (!motion.is_none())
			    mo = toVector(motion.cast<CArray<uint8_t>>());
		    if ((!f.empty() && f.size() != n) || (!mo.empty() && mo.size() != n))
			    throw py::value_error("frame / motion: length mismatch");
		    return new ListModeDevice(c, n, lm.d1.data(), lm.d2.data(),
		                              lm.tof.empty() ? nullptr : lm.tof.data(),
		                              f.empty() ? nullptr : f.data(),
		                              mo.empty() ? nullptr : mo.data(),
		                              lm.randoms.empty() ? nullptr : lm.randoms.data());
	    },
	    "ctx"_a, "list_mode"_a, "frame"_a = py::none(), "motion"_a = py::none(),
	    py::keep_alive<0, 1>());

	// ---- motion and dynamic framing
	m.def(
	    "readLORMotion",
	    [](const std::string& fname) {
		    const auto recs = readMotionCsv(fname);
		    const size_t k = recs.size();
		    py::array_t<int64_t> t((py::ssize_t)k);
		    py::array_t<float> xf({(py::ssize_t)k, (py::ssize_t)12}), err((py::ssize_t)k);
		    for (size_t i = 0; i < k; i++)
		    {
			    t.mutable_data()[i] = recs[i].timestamp;
			    std::memcpy(xf.mutable_data() + 12 * i, recs[i].transform, 12 * sizeof(float));
			    err.mutable_data()[i] = recs[i].error;
		    }
		    return py::make_tuple(t, xf, err);
	    },
	    "fname"_a);
	auto records = [](CArray<int64_t> t, CArray<float> xf, CArray<float> err) {
		const size_t k = (size_t)t.size();
		if ((size_t)xf.size() != 12 * k || (size_t)err.size() != k)
			throw py::value_error("motion: timestamps, transforms (k x 12) and errors differ");
		std::vector<MotionRecord> recs(k);
		for (size_t i = 0; i < k; i++)
		{
			recs[i].timestamp = t.data()[i];
			std::memcpy(recs[i].transform, xf.data() + 12 * i, 12 * sizeof(float));
			recs[i].error = err.data()[i];
		}
		return recs;
	};
	m.def(
	    "writeLORMotion",
	    [records](const std::string& fname, CArray<int64_t> t, CArray<float> xf,
	              CArray<float> err) { writeMotionCsv(fname, records(t, xf, err)); },
	    "fname"_a, "timestamps"_a, "transforms"_a, "errors"_a);
	m.def(
	    "motionFrames",
	    [records](CArray<uint32_t> ts, CArray<int64_t> t, CArray<float> xf, CArray<float> err) {
		    return toNumpy(motionFrames(toVector(ts), records(t, xf, err)));
	    },
	    "event_timestamps"_a, "timestamps"_a, "transforms"_a, "errors"_a);
	// ---- time-averaged (moved) images: the motion as its timestamps and
	// k x 12 transforms (readLORMotion's first two arrays)
	auto xfVec = [](CArray<int64_t> t, CArray<float> xf) {
		if ((size_t)xf.size() != 12 * (size_t)t.size())
			throw py::value_error("motion: need k timestamps and k x 12 transforms");
		return std::make_pair(toVector(t), toVector(xf));
	};
	m.def(
	    "getFullTimeRange", [](CArray<int64_t> t) { return getFullTimeRange(toVector(t)); },
	    "timestamps"_a);
	m.def(
	    "motionFrameWeights",
	    [](CArray<int64_t> t, int64_t t0, int64_t t1) {
		    std::vector<float> w;
		    const int64_t f0 = motionFrameWeights(toVector(t), t0, t1, w);
		    return std::make_pair(f0, w);
	    },
	    "timestamps"_a, "time_start"_a, "time_stop"_a);
	m.def(
	    "timeAverageMoveImage",
	    [xfVec](CArray<int64_t> t, CArray<float> xf, const ImageDevice& img, py::object t0,
	            py::object t1) {
		    auto [ts, x] = xfVec(t, xf);
		    if (t0.is_none() != t1.is_none())
			    throw py::value_error("give both time_start and time_stop, or neither");
		    if (t0.is_none())
			    return timeAverageMoveImageDevice(ts, x, img).release();
		    return timeAverageMoveImageDevice(ts, x, img, t0.cast<int64_t>(), t1.cast<int64_t>())
		        .release();
	    },
	    "timestamps"_a, "transforms"_a, "unmoved_image"_a, "time_start"_a = py::none(),
	    "time_stop"_a = py::none(), py::keep_alive<0, 3>());
	m.def(
	    "timeAverageMoveImageInto",
	    [xfVec](CArray<int64_t> t, CArray<float> xf, const ImageDevice& img, ImageDevice& out,
	            int64_t t0, int64_t t1, int64_t frame) {
		    auto [ts, x] = xfVec(t, xf);
		    timeAverageMoveImageDevice(ts, x, img, out, t0, t1, frame);
	    },
	    "timestamps"_a, "transforms"_a, "unmoved_image"_a, "out_image"_a, "time_start"_a,
	    "time_stop"_a, "out_dynamic_frame"_a = 0);
	m.def(
	    "timeAverageMoveImageDynamic",
	    [xfVec](CArray<int64_t> t, CArray<float> xf, const ImageDevice& img,
	            CArray<uint32_t> bounds) {
		    auto [ts, x] = xfVec(t, xf);
		    return timeAverageMoveImageDynamicDevice(ts, x, img, toVector(bounds)).release();
	    },
	    "timestamps"_a, "transforms"_a, "unmoved_image"_a, "dynamic_framing"_a,
	    py::keep_alive<0, 3>());
	m.def(
	    "readDynamicFraming", [](const std::string& f) { return toNumpy(readDynamicFraming(f)); },
	    "fname"_a);
	m.def(
	    "writeDynamicFraming",
	    [](const std::string& f, CArray<uint32_t> b) { writeDynamicFraming(f, toVector(b)); },
	    "fname"_a, "timestamps"_a);
	m.def(
	    "dynamicFrames",
	    [](CArray<uint32_t> ts, CArray<uint32_t> bounds) {
		    return toNumpy(dynamicFrames(toVector(ts), toVector(bounds)));
	    },
	    "event_timestamps"_a, "framing"_a);

	// ---- images (NIfTI)
	m.def(
	    "readImage",
	    [](const std::string& fname) {
		    NiftiImage im = readNifti(fname);
		    return py::make_tuple(im.params, imageArray(im));
	    },
	    "fname"_a);
	m.def(
	    "writeImage",
	    [](const std::string& fname, const ImageParams& p, CArray<float> a) {
		    if ((size_t)a.size() != p.numVoxels() * (size_t)p.nt)
			    throw py::value_error("image size mismatch");
		    writeNifti(fname, p, a.data());
	    },
	    "fname"_a, "params"_a, "data"_a);
	m.def(
	    "ImageDeviceFromFile",
	    [](const Context& c, const std::string& fname) {
		    NiftiImage im = readNifti(fname);
		    auto* img = new ImageDevice(c, im.params);
		    img->copyFromHost(im.data.data());
		    return img;
	    },
	    "ctx"_a, "fname"_a, py::keep_alive<0, 1>());
	image.def(
	    "writeToFile",
	    [](const ImageDevice& img, const std::string& fname) {
		    const ImageParams& p = img.getParams();
		    std::vector<float> h(p.numVoxels() * (size_t)p.nt);
		    img.copyToHost(h.data());
		    writeNifti(fname, p, h.data());
	    },
	    "fname"_a);
	image.def(
	    "readFromFile",
	    [](ImageDevice& img, const std::string& fname) {
		    NiftiImage im = readNifti(fname);
		    const ImageParams& p = img.getParams();
		    if (im.params.nx != p.nx || im.params.ny != p.ny || im.params.nz != p.nz ||
		        im.params.nt != p.nt || std::fabs(im.params.vx - p.vx) > 1e-4f ||
		        std::fabs(im.params.vy - p.vy) > 1e-4f || std::fabs(im.params.vz - p.vz) > 1e-4f ||
		        std::fabs(im.params.off_x - p.off_x) > 1e-4f ||
		        std::fabs(im.params.off_y - p.off_y) > 1e-4f ||
		        std::fabs(im.params.off_z - p.off_z) > 1e-4f)
			    throw Error("The image file's parameters do not match the image's: " + fname);
		    img.copyFromHost(im.data.data());
	    },
	    "fname"_a);

	// ---- image-space PSF
	m.def(
	    "readImagePSF",
	    [](const std::string& fname) {
		    std::vector<float> kx, ky, kz;
		    readImagePsf(fname, kx, ky, kz);
		    return py::make_tuple(toNumpy(kx), toNumpy(ky), toNumpy(kz));
	    },
	    "fname"_a);
	m.def(
	    "applyImagePSF",
	    [](const Context& c, const ImageDevice& in, ImageDevice& out, CArray<float> kx,
	       CArray<float> ky, CArray<float> kz, bool transpose) {
		    const ImageParams& p = in.getParams();
		    if (!p.isSameAs(out.getParams()) || p.nt != out.getParams().nt)
			    throw Error("applyImagePSF: images of different geometry");
		    ImageDevice tmp(c, p);
		    const yrtmi_image ci = p.desc();
		    check(yrtmi_psf_apply(c.get(), &ci, in.getDevicePointer(), out.getDevicePointer(),
		                          tmp.getDevicePointer(), kx.data(), (int)kx.size(), ky.data(),
		                          (int)ky.size(), kz.data(), (int)kz.size(), transpose ? 1 : 0),
		          "image PSF");
		    c.synchronize();
	    },
	    "ctx"_a, "img_in"_a, "img_out"_a, "kernel_x"_a, "kernel_y"_a, "kernel_z"_a,
	    "transpose"_a = false);

	// ---- histograms (raw-data array numZBin x numPhi x numR)
	histo.def(
	    "readFromFile",
	    [](Histogram3DDevice& h, const std::string& fname) {
		    const Context& c = h.context();
		    RawdArray a = readRawd(fname, {(uint64_t)c.numZBin(), (uint64_t)c.numPhi(),
		                                   (uint64_t)c.numR()});
		    h.copyFromHost(a.data.data());
	    },
	    "fname"_a);
	histo.def(
	    "writeToFile",
	    [](const Histogram3DDevice& h, const std::string& fname) {
		    const Context& c = h.context();
		    std::vector<float> v(h.count());
		    h.copyToHost(v.data());
		    writeRawd(fname, {(uint64_t)c.numZBin(), (uint64_t)c.numPhi(), (uint64_t)c.numR()},
		              v.data());
	    },
	    "fname"_a);
}

// pybind11 module `_pyyrtmi` (imported as `pyyrtmi`): the C++ host layer with the names of the
// reference's pyyrtpet bindings for this path (src/pybind11.cpp:95-177;
// OSEM: src/recon/OSEM.cpp:37-162; ProjectorParams:
// src/operators/ProjectorParams.cpp:19-80; OperatorProjector:
// src/operators/OperatorProjector.cpp:31-60; createOSEM:
// src/utils/ReconstructionUtils.cpp:94-104).
//
// Device data: owned objects upload numpy arrays; alias objects wrap device
// pointers given as integers or as objects with a data_ptr() method (torch
// tensors), which must stay alive while the alias is used.
#include "yrtmi_host.hpp"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

namespace py = pybind11;
using namespace pybind11::literals;
using namespace yrtmi;

namespace
{

template <typename T>
using CArray = py::array_t<T, py::array::c_style | py::array::forcecast>;

uintptr_t devicePtr(const py::object& o)
{
	if (o.is_none())
		return 0;
	if (py::isinstance<py::int_>(o))
		return o.cast<uintptr_t>();
	if (py::hasattr(o, "data_ptr"))
		return o.attr("data_ptr")().cast<uintptr_t>();
	throw py::type_error("expected a device pointer (int) or an object with data_ptr()");
}

template <typename T>
const T* hostPtr(const py::object& o, size_t n, const char* name)
{
	if (o.is_none())
		return nullptr;
	auto a = o.cast<CArray<T>>();
	if ((size_t)a.size() != n)
		throw py::value_error(std::string(name) + ": length mismatch");
	return a.data();
}

py::array_t<float> imageToNumpy(const ImageDevice& img)
{
	const ImageParams& p = img.getParams();
	py::array_t<float> out({(py::ssize_t)p.nt, (py::ssize_t)p.nz, (py::ssize_t)p.ny,
	                        (py::ssize_t)p.nx});
	img.copyToHost(out.mutable_data());
	return out;
}

}  // namespace

// file formats (pyyrtmi_io.cpp)
void bindIO(py::module_& m, py::class_<Scanner>& scanner, py::class_<ImageParams>& imageParams,
            py::class_<ImageDevice>& image, py::class_<Histogram3DDevice>& histo);

PYBIND11_MODULE(_pyyrtmi, m)
{
	m.doc() = "yrtmi: MI355X OSEM projection path (C++ host layer over the yrtmi C ABI)";
	py::register_exception<Error>(m, "YrtmiError", PyExc_RuntimeError);

	py::class_<Scanner> scannerCls(m, "Scanner");
	scannerCls
	    .def(py::init<std::string, float, float, float, float, float, size_t, size_t,
	                  size_t, size_t, size_t, size_t>(),
	         "scannerName"_a, "axialFOV"_a, "crystalSize_z"_a, "crystalSize_trans"_a,
	         "crystalDepth"_a, "scannerRadius"_a, "detsPerRing"_a, "numRings"_a,
	         "numDOI"_a, "maxRingDiff"_a, "minAngDiff"_a, "detsPerBlock"_a)
	    .def("getNumDets", &Scanner::getNumDets)
	    .def_readonly("scannerName", &Scanner::scannerName)
	    .def_readonly("axialFOV", &Scanner::axialFOV)
	    .def_readonly("crystalSize_z", &Scanner::crystalSize_z)
	    .def_readonly("crystalSize_trans", &Scanner::crystalSize_trans)
	    .def_readonly("crystalDepth", &Scanner::crystalDepth)
	    .def_readonly("scannerRadius", &Scanner::scannerRadius)
	    .def_readonly("detsPerRing", &Scanner::detsPerRing)
	    .def_readonly("numRings", &Scanner::numRings)
	    .def_readonly("numDOI", &Scanner::numDOI)
	    .def_readonly("maxRingDiff", &Scanner::maxRingDiff)
	    .def_readonly("minAngDiff", &Scanner::minAngDiff)
	    .def_readonly("detsPerBlock", &Scanner::detsPerBlock);

	py::class_<ImageParams> imageParamsCls(m, "ImageParams");
	imageParamsCls
	    .def(py::init<int64_t, int64_t, int64_t, float, float, float, float, float, float,
	                  int64_t>(),
	         "nx"_a, "ny"_a, "nz"_a, "length_x"_a, "length_y"_a, "length_z"_a,
	         "offset_x"_a = 0.f, "offset_y"_a = 0.f, "offset_z"_a = 0.f, "num_frames"_a = 1)
	    .def_static("fromParams", &ImageParams::fromParams, "nx"_a, "ny"_a, "nz"_a, "vx"_a,
	                "vy"_a, "vz"_a, "num_frames"_a = 1)
	    .def("isSameAs", &ImageParams::isSameAs)
	    .def_readonly("nx", &ImageParams::nx)
	    .def_readonly("ny", &ImageParams::ny)
	    .def_readonly("nz", &ImageParams::nz)
	    .def_readonly("nt", &ImageParams::nt)
	    .def_readonly("length_x", &ImageParams::length_x)
	    .def_readonly("length_y", &ImageParams::length_y)
	    .def_readonly("length_z", &ImageParams::length_z)
	    .def_readonly("vx", &ImageParams::vx)
	    .def_readonly("vy", &ImageParams::vy)
	    .def_readonly("vz", &ImageParams::vz)
	    .def_readonly("off_x", &ImageParams::off_x)
	    .def_readonly("off_y", &ImageParams::off_y)
	    .def_readonly("off_z", &ImageParams::off_z)
	    .def_readonly("fovRadius", &ImageParams::fovRadius);

	py::class_<ProjectorParams> pp(m, "ProjectorParams");
	py::enum_<ProjectorType>(pp, "ProjectorType")
	    .value("SIDDON", ProjectorType::SIDDON)
	    .value("DD", ProjectorType::DD)
	    .export_values();
	pp.def(py::init<const Scanner&>(), "scanner"_a, py::keep_alive<1, 2>())
	    .def("setProjector",
	         static_cast<void (ProjectorParams::*)(const std::string&)>(
	             &ProjectorParams::setProjector))
	    .def("setProjector",
	         static_cast<void (ProjectorParams::*)(ProjectorType)>(&ProjectorParams::setProjector))
	    .def("addTOF", &ProjectorParams::addTOF, "tof_width_ps"_a, "tof_num_std"_a)
	    .def("removeTOF", &ProjectorParams::removeTOF)
	    .def("hasTOF", &ProjectorParams::hasTOF)
	    .def("getTOFWidth_ps", &ProjectorParams::getTOFWidth_ps)
	    .def("getTOFNumStd", &ProjectorParams::getTOFNumStd)
	    .def("setProjPsf",
	         static_cast<void (ProjectorParams::*)(const std::string&)>(
	             &ProjectorParams::setProjPsf),
	         "proj_psf_fname"_a)
	    .def(
	        "setProjPsfKernels",
	        [](ProjectorParams& self, CArray<float> k, float sStep, float kSpacing) {
		        if (k.ndim() != 2)
			        throw py::value_error("kernels must be (num_s, K)");
		        self.setProjPsf(std::vector<float>(k.data(), k.data() + k.size()),
		                        (int)k.shape(0), (int)k.shape(1), sStep, kSpacing);
	        },
	        "kernels"_a, "s_step"_a, "k_spacing"_a)
	    .def_readwrite("projectorType", &ProjectorParams::projectorType)
	    .def_readwrite("numRays", &ProjectorParams::numRays)
	    .def_readwrite("setupFP64", &ProjectorParams::setupFP64)
	    .def_readonly("projPsf_fname", &ProjectorParams::projPsf_fname);

	py::class_<Context>(m, "Context")
	    .def(py::init([](const Scanner& s, int device, const py::object& stream) {
		         return new Context(s, device, reinterpret_cast<void*>(devicePtr(stream)));
	         }),
	         "scanner"_a, "device"_a = 0, "stream"_a = py::none(), py::keep_alive<1, 2>())
	    .def("setStream",
	         [](Context& c, const py::object& s) {
		         c.setStream(reinterpret_cast<void*>(devicePtr(s)));
	         })
	    .def("synchronize", &Context::synchronize)
	    .def("hasHistogram", &Context::hasHistogram)
	    .def("histogramShape",
	         [](const Context& c) { return py::make_tuple(c.numZBin(), c.numPhi(), c.numR()); });

	py::class_<ImageDevice> imageCls(m, "ImageDevice");
	imageCls.def("getParams", &ImageDevice::getParams)
	    .def("getDevicePointer",
	         [](const ImageDevice& i) { return reinterpret_cast<uintptr_t>(i.getDevicePointer()); })
	    .def("isAlias", &ImageDevice::isAlias)
	    .def("setValue", &ImageDevice::setValue, "value"_a)
	    .def("copyFromHost",
	         [](ImageDevice& i, CArray<float> a) {
		         if ((size_t)a.size() != i.size())
			         throw py::value_error("image size mismatch");
		         i.copyFromHost(a.data());
	         })
	    .def("transferToHostMemory", &imageToNumpy)
	    .def("copyFrom", &ImageDevice::copyFrom);
	m.def(
	    "ImageDeviceOwned",
	    [](const Context& c, const ImageParams& p) { return new ImageDevice(c, p); },
	    "ctx"_a, "params"_a, py::keep_alive<0, 1>());
	m.def(
	    "ImageDeviceAlias",
	    [](const Context& c, const ImageParams& p, const py::object& ptr) {
		    return new ImageDevice(c, p, reinterpret_cast<float*>(devicePtr(ptr)));
	    },
	    "ctx"_a, "params"_a, "device_pointer"_a, py::keep_alive<0, 1>(),
	    py::keep_alive<0, 3>());

	py::class_<ListModeDevice>(m, "ListModeDevice")
	    .def("count", &ListModeDevice::count)
	    .def("hasTOF", &ListModeDevice::hasTOF)
	    .def_readwrite("numFrames", &ListModeDevice::numFrames)
	    .def("setMotionTransforms",
	         [](ListModeDevice& l, CArray<float> xf) {
		         if (xf.size() % 12 != 0)
			         throw py::value_error("transforms: 12 floats each");
		         l.motionTransforms.assign(xf.data(), xf.data() + xf.size());
	         })
	    .def("getProjectionValues",
	         [](const ListModeDevice& l) {
		         py::array_t<float> out((py::ssize_t)l.count());
		         if (l.count() > 0)
			         l.valuesToHost(out.mutable_data());
		         return out;
	         })
	    .def("setProjectionValues",
	         [](ListModeDevice& l, CArray<float> v) {
		         if ((size_t)v.size() != l.count())
			         throw py::value_error("values: length mismatch");
		         if (l.count() > 0)
			         l.valuesFromHost(v.data());
	         })
	    .def("getValuesPointer",
	         [](const ListModeDevice& l) { return reinterpret_cast<uintptr_t>(l.values()); });
	m.def(
	    "ListModeDeviceOwned",
	    [](const Context& c, CArray<uint32_t> d1, CArray<uint32_t> d2, const py::object& tof,
	       const py::object& frame, const py::object& motion, const py::object& randoms) {
		    const size_t n = (size_t)d1.size();
		    if ((size_t)d2.size() != n)
			    throw py::value_error("d1 / d2 length mismatch");
		    // keep converted arrays alive across the upload
		    py::object t = tof.is_none() ? tof : py::object(tof.cast<CArray<float>>());
		    py::object f = frame.is_none() ? frame : py::object(frame.cast<CArray<int32_t>>());
		    py::object mo = motion.is_none() ? motion : py::object(motion.cast<CArray<uint8_t>>());
		    py::object r = randoms.is_none() ? randoms : py::object(randoms.cast<CArray<float>>());
		    return new ListModeDevice(c, n, d1.data(), d2.data(), hostPtr<float>(t, n, "tof"),
		                              hostPtr<int32_t>(f, n, "frame"),
		                              hostPtr<uint8_t>(mo, n, "motion"),
		                              hostPtr<float>(r, n, "randoms"));
	    },
	    "ctx"_a, "d1"_a, "d2"_a, "tof"_a = py::none(), "frame"_a = py::none(),
	    "motion"_a = py::none(), "randoms"_a = py::none(), py::keep_alive<0, 1>());
	m.def(
	    "ListModeDeviceAlias",
	    [](const Context& c, size_t n, const py::object& d1, const py::object& d2,
	       const py::object& tof, const py::object& frame, const py::object& motion,
	       const py::object& randoms) {
		    return ListModeDevice::alias(
		               c, n, reinterpret_cast<uint32_t*>(devicePtr(d1)),
		               reinterpret_cast<uint32_t*>(devicePtr(d2)),
		               reinterpret_cast<float*>(devicePtr(tof)),
		               reinterpret_cast<int32_t*>(devicePtr(frame)),
		               reinterpret_cast<uint8_t*>(devicePtr(motion)),
		               reinterpret_cast<float*>(devicePtr(randoms)))
		        .release();
	    },
	    "ctx"_a, "n"_a, "d1"_a, "d2"_a, "tof"_a = py::none(), "frame"_a = py::none(),
	    "motion"_a = py::none(), "randoms"_a = py::none(), py::keep_alive<0, 1>(),
	    py::keep_alive<0, 3>(), py::keep_alive<0, 4>(), py::keep_alive<0, 5>(),
	    py::keep_alive<0, 6>(), py::keep_alive<0, 7>(), py::keep_alive<0, 8>());

	py::class_<Histogram3DDevice> histoCls(m, "Histogram3DDevice");
	histoCls.def("count", &Histogram3DDevice::count)
	    .def("getDevicePointer",
	         [](const Histogram3DDevice& h) {
		         return reinterpret_cast<uintptr_t>(h.getDevicePointer());
	         })
	    .def("copyFromHost",
	         [](Histogram3DDevice& h, CArray<float> a) {
		         if ((size_t)a.size() != h.count())
			         throw py::value_error("histogram size mismatch");
		         h.copyFromHost(a.data());
	         })
	    .def("transferToHostMemory", [](const Histogram3DDevice& h) {
		    py::array_t<float> out((py::ssize_t)h.count());
		    h.copyToHost(out.mutable_data());
		    return out;
	    });
	m.def(
	    "Histogram3DDeviceOwned", [](const Context& c) { return new Histogram3DDevice(c); },
	    "ctx"_a, py::keep_alive<0, 1>());
	m.def(
	    "Histogram3DDeviceAlias",
	    [](const Context& c, const py::object& ptr) {
		    return new Histogram3DDevice(c, reinterpret_cast<float*>(devicePtr(ptr)));
	    },
	    "ctx"_a, "device_pointer"_a, py::keep_alive<0, 1>(), py::keep_alive<0, 2>());

	py::class_<OperatorProjectorMI>(m, "OperatorProjectorMI")
	    .def(py::init<const Context&, const ProjectorParams&, bool>(), "ctx"_a,
	         "proj_params"_a, "with_tof"_a = true, py::keep_alive<1, 2>())
	    .def("applyA",
	         static_cast<void (OperatorProjectorMI::*)(const ImageDevice&, ListModeDevice&)>(
	             &OperatorProjectorMI::applyA),
	         "img"_a, "proj"_a)
	    .def("applyA",
	         static_cast<void (OperatorProjectorMI::*)(const ImageDevice&,
	                                                   Histogram3DDevice&)>(
	             &OperatorProjectorMI::applyA),
	         "img"_a, "proj"_a)
	    .def("applyAH",
	         static_cast<void (OperatorProjectorMI::*)(const ListModeDevice&, ImageDevice&)>(
	             &OperatorProjectorMI::applyAH),
	         "proj"_a, "img"_a)
	    .def("applyAH",
	         static_cast<void (OperatorProjectorMI::*)(const Histogram3DDevice&,
	                                                   ImageDevice&)>(
	             &OperatorProjectorMI::applyAH),
	         "proj"_a, "img"_a)
	    .def("setLaunch", &OperatorProjectorMI::setLaunch, "lanes_per_event"_a,
	         "block_size"_a = 256, "grid_blocks"_a = 0);

	py::class_<OSEM_MI>(m, "OSEM")
	    .def("setImageParams", &OSEM_MI::setImageParams, "params"_a)
	    .def("getImageParams", &OSEM_MI::getImageParams)
	    .def("setDataInput",
	         static_cast<void (OSEM_MI::*)(const ListModeDevice*)>(&OSEM_MI::setDataInput),
	         "proj_data"_a, py::keep_alive<1, 2>())
	    .def("setDataInput",
	         static_cast<void (OSEM_MI::*)(const Histogram3DDevice*)>(&OSEM_MI::setDataInput),
	         "proj_data"_a, py::keep_alive<1, 2>())
	    // batched histogram frames (the list is kept alive with the OSEM)
	    .def("setDataInput",
	         static_cast<void (OSEM_MI::*)(const std::vector<const Histogram3DDevice*>&)>(
	             &OSEM_MI::setDataInput),
	         "histogram_frames"_a, py::keep_alive<1, 2>())
	    .def("isListModeEnabled", &OSEM_MI::isListModeEnabled)
	    .def("setProjector", &OSEM_MI::setProjector, "projector_name"_a)
	    .def("setNumRays", &OSEM_MI::setNumRays, "num_rays"_a)
	    .def("addTOF", &OSEM_MI::addTOF, "tof_width_ps"_a, "tof_num_std"_a)
	    .def("addProjPSF", &OSEM_MI::addProjPSF, "proj_psf_fname"_a)
	    .def(
	        "addImagePSF",
	        [](OSEM_MI& o, const std::string& f, const std::string& mode) {
		        if (mode != "UNIFORM")
			        throw Error("Image-space PSF: only the UNIFORM mode is supported");
		        o.addImagePSF(f);
	        },
	        "image_psf_fname"_a, "image_psf_mode"_a = "UNIFORM")
	    .def("addUniformGaussianImagePSFFromFWHM", &OSEM_MI::addUniformGaussianImagePSFFromFWHM,
	         "fwhm_x"_a, "fwhm_y"_a, "fwhm_z"_a, "kernel_size_x"_a = 0, "kernel_size_y"_a = 0,
	         "kernel_size_z"_a = 0)
	    .def("addUniformGaussianImagePSFFromSigma",
	         &OSEM_MI::addUniformGaussianImagePSFFromSigma, "sigma_x"_a, "sigma_y"_a,
	         "sigma_z"_a, "kernel_size_x"_a = 0, "kernel_size_y"_a = 0, "kernel_size_z"_a = 0)
	    .def("hasImagePSF", &OSEM_MI::hasImagePSF)
	    .def("getImagePSFKernels",
	         [](const OSEM_MI& o) {
		         return py::make_tuple(o.imagePSFKernel(0), o.imagePSFKernel(1),
		                               o.imagePSFKernel(2));
	         })
	    .def("setGlobalScalingFactor", &OSEM_MI::setGlobalScalingFactor, "global_scale"_a)
	    .def("setAttenuationImage", &OSEM_MI::setAttenuationImage, "att_image"_a,
	         py::keep_alive<1, 2>())
	    .def("setHardwareAttenuationImage", &OSEM_MI::setHardwareAttenuationImage,
	         "att_hardware"_a, py::keep_alive<1, 2>())
	    .def("setInVivoAttenuationImage", &OSEM_MI::setInVivoAttenuationImage,
	         "att_invivo"_a, py::keep_alive<1, 2>())
	    .def("setACFHistogram", &OSEM_MI::setACFHistogram, "acf_his"_a,
	         py::keep_alive<1, 2>())
	    .def("setHardwareACFHistogram", &OSEM_MI::setHardwareACFHistogram,
	         "acf_hardware_his"_a, py::keep_alive<1, 2>())
	    .def("setInVivoACFHistogram", &OSEM_MI::setInVivoACFHistogram, "acf_invivo_his"_a,
	         py::keep_alive<1, 2>())
	    .def("setRandomsHistogram", &OSEM_MI::setRandomsHistogram, "randoms_his"_a,
	         py::keep_alive<1, 2>())
	    .def("setScatterHistogram", &OSEM_MI::setScatterHistogram, "scatter_his"_a,
	         py::keep_alive<1, 2>())
	    .def("setSensitivityHistogram", &OSEM_MI::setSensitivityHistogram, "sens_his"_a,
	         py::keep_alive<1, 2>())
	    .def("setInvertSensitivity", &OSEM_MI::setInvertSensitivity, "invert"_a = true)
	    .def("setInitialEstimate", &OSEM_MI::setInitialEstimate, "initial_estimate"_a,
	         py::keep_alive<1, 2>())
	    .def("setMaskImage", &OSEM_MI::setMaskImage, "mask_image"_a, py::keep_alive<1, 2>())
	    .def("getExpectedSensImagesAmount", &OSEM_MI::getExpectedSensImagesAmount)
	    .def("setSensitivityImages",
	         [](OSEM_MI& o, const std::vector<const ImageDevice*>& imgs) {
		         o.setSensitivityImages(imgs);
	         })
	    .def("generateSensitivityImages",
	         [](py::object self) {
		         // each image is owned by the OSEM object, which it keeps alive
		         py::list out;
		         for (auto& s : self.cast<OSEM_MI&>().generateSensitivityImages())
			         out.append(
			             py::cast(s.get(), py::return_value_policy::reference_internal, self));
		         return out;
	         })
	    .def("reconstruct", &OSEM_MI::reconstruct, py::keep_alive<0, 1>())
	    .def("getProjectorParams", [](OSEM_MI& o) -> ProjectorParams& { return o.projectorParams; },
	         py::return_value_policy::reference_internal)
	    .def("getContext", &OSEM_MI::getContext, py::return_value_policy::reference_internal)
	    .def_readwrite("num_MLEM_iterations", &OSEM_MI::num_MLEM_iterations)
	    .def_readwrite("num_OSEM_subsets", &OSEM_MI::num_OSEM_subsets)
	    .def_readwrite("hardThreshold", &OSEM_MI::hardThreshold)
	    .def_readwrite("denomThreshold", &OSEM_MI::denomThreshold)
	    .def_readwrite("useTilePlans", &OSEM_MI::useTilePlans)
	    .def_readwrite("useLorOrder", &OSEM_MI::useLorOrder)
	    .def_readwrite("lanesPerEvent", &OSEM_MI::lanesPerEvent)
	    .def_readwrite("useGraphs", &OSEM_MI::useGraphs)
	    .def_readwrite("useAutotune", &OSEM_MI::useAutotune)
	    .def_readwrite("memoryBudget", &OSEM_MI::memoryBudget)
	    .def("numPlannedSubsets", &OSEM_MI::numPlannedSubsets)
	    .def_readwrite("timeIterations", &OSEM_MI::timeIterations)
	    .def("iterationSeconds", &OSEM_MI::iterationSeconds)
	    .def("numHistogramFrames", &OSEM_MI::numHistogramFrames);

	m.def(
	    "createOSEM",
	    [](const Scanner& s, bool useGPU, int device) {
		    return createOSEM(s, useGPU, device).release();
	    },
	    "scanner"_a, "use_gpu"_a = true, "device"_a = 0, py::keep_alive<0, 1>());

	m.def("forwProject",
	      static_cast<void (*)(const Context&, const ProjectorParams&, const ImageDevice&,
	                           ListModeDevice&)>(&forwProject),
	      "ctx"_a, "proj_params"_a, "img"_a, "proj"_a);
	m.def("forwProject",
	      static_cast<void (*)(const Context&, const ProjectorParams&, const ImageDevice&,
	                           Histogram3DDevice&)>(&forwProject),
	      "ctx"_a, "proj_params"_a, "img"_a, "proj"_a);
	m.def("backProject",
	      static_cast<void (*)(const Context&, const ProjectorParams&, ImageDevice&,
	                           const ListModeDevice&)>(&backProject),
	      "ctx"_a, "proj_params"_a, "img"_a, "proj"_a);
	m.def("backProject",
	      static_cast<void (*)(const Context&, const ProjectorParams&, ImageDevice&,
	                           const Histogram3DDevice&)>(&backProject),
	      "ctx"_a, "proj_params"_a, "img"_a, "proj"_a);

	bindIO(m, scannerCls, imageParamsCls, imageCls, histoCls);
}

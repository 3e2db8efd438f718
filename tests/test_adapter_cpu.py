"""The reference-side operator adapter (include/yrtmi_yrtpet_adapter.hpp)
compiled against the reference's own headers: OperatorProjectorBase.hpp
(include/yrt-pet/operators/OperatorProjectorBase.hpp:21-44), ProjectorParams,
ProjectionData, BinIterator and Scanner, with every member instantiated
for an image type shaped like yrt::Image (getParams() with nx..nt, lengths
and offsets; getRawPointer()).  yrt::Image itself needs nlohmann/json,
which is not in this image; the reference's build provides it.  CPU only:
nothing is run.  Skipped where the reference tree is absent."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/yrt-pet/include"

TU = r'''
#include "yrtmi_yrtpet_adapter.hpp"
#include <type_traits>

// An image type with yrt::Image's interface for this class (Image.hpp:21-29,
// ImageParams.hpp:23-38).
struct ImageShaped : yrt::Variable
{
	struct Params
	{
		long nx = 4, ny = 4, nz = 2, nt = 1;
		float length_x = 8, length_y = 8, length_z = 4;
		float off_x = 0, off_y = 0, off_z = 0;
	} params;
	std::vector<float> data = std::vector<float>(32, 0.f);
	const Params& getParams() const { return params; }
	float* getRawPointer() { return data.data(); }
	const float* getRawPointer() const { return data.data(); }
};

template class yrtmi_yrtpet::OperatorProjectorMIAdapter<ImageShaped>;
static_assert(std::is_base_of_v<yrt::OperatorProjectorBase,
                                yrtmi_yrtpet::OperatorProjectorMIAdapter<ImageShaped>>);
static_assert(std::is_base_of_v<yrt::Operator,
                                yrtmi_yrtpet::OperatorProjectorMIAdapter<ImageShaped>>);
'''


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_adapter_compiles_against_reference_headers(tmp_path):
    gxx = shutil.which("g++")
    assert gxx, "g++ not found"
    src = tmp_path / "adapter_tu.cpp"
    src.write_text(TU)
    cmd = [gxx, "-std=c++20", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-function",
           "-I", REF_INC, "-I", os.path.join(REPO, "include"),
           "-I", os.path.join(REPO, "yrt-pet_amd", "csrc", "host"), str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]

// C bindings over the compiled reference (tmonseigne/PALMTracer-cpp) hot-path classes.
//
// TEST INFRASTRUCTURE ONLY. This file is compiled by oracle/Makefile together with the
// reference's own sources, read in place from /root/reference, into oracle/_ref/libref.so.
// It pins the C oracle (oracle/oracle.c) and can serve as the "reference" CPU baseline.
// Nothing in the product library links or loads it.
//
// Each binding names the reference entry point it drives:
//   ref_localization  -> Palm::Processing              (lib/Palm.cpp:55-70, CPU/main.cpp:112-121)
//   ref_fitting       -> Palm::FitLocalizations        (lib/Palm.cpp:73-119, CPU/main.cpp:141-148)
//   ref_wavelett      -> ATrous::Process per plane     (CPU/main.cpp:257-284)
//   ref_autothreshold -> Palm::AutoThreshold           (lib/Palm.cpp:122-172, CPU/main.cpp:75-85)
//   ref_segment       -> Segmentation2D::Process       (lib/localization/Segmentation2D.cpp:50-74)
//   ref_astig_estimate-> AstigFit::Estimate            (lib/fit/AstigFit.cpp:36-72, CPU/main.cpp:324-353)
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <memory>
#include <queue>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "Palm.hpp"
#include "fit/AstigFit.hpp"
#include "localization/ATrous.hpp"
// The label map is a private member; expose it to this test binding only.
#define private public
#include "localization/Segmentation2D.hpp"
#undef private

using namespace palm_tracer;

extern "C" {

uint64_t ref_localization(const uint16_t* input, double* output, uint64_t outputSize, uint64_t height, uint64_t width,
                          uint64_t planes, double threshold, double ratio, uint64_t fit, const double* fitParams)
{
	const Vec3I dim{.x = width, .y = height, .z = planes};
	const Palm palm(input, output, outputSize, dim, threshold, ratio, fit, fitParams);
	return static_cast<uint64_t>(palm.Processing());
}

uint64_t ref_fitting(const uint16_t* input, const double* positions, uint64_t nPoints, double* output, uint64_t outputSize,
                     uint64_t height, uint64_t width, uint64_t planes, uint64_t fit, const double* fitParams)
{
	const Vec3I dim{.x = width, .y = height, .z = planes};
	return Palm::FitLocalizations(input, positions, nPoints, output, outputSize, dim, fit, fitParams);
}

void ref_wavelett(const uint16_t* input, double* output, uint64_t height, uint64_t width, uint64_t planes, uint64_t level)
{
	if (planes == 0) { return; }
	if (level == 0) { level = 1; }
	const Vec2I dim{.x = width, .y = height};
	const uint64_t size = dim.GetSize();
	const uint64_t lv = std::min<uint64_t>(level - 1, N_PLANES - 1);
	localization::ATrous at(dim);
	for (uint64_t i = 0; i < planes; ++i) {
		at.Process(input + i * size);
		std::copy_n(at.wDecomp_[lv].data(), size, output + i * size);
	}
}

double ref_autothreshold(const uint16_t* input, uint64_t height, uint64_t width, const double* fitParams)
{
	const auto cap = static_cast<uint64_t>(0.2 * static_cast<double>(height * width) * localization_columns::Count);
	std::vector<double> pts(cap);
	const Vec3I dim{.x = width, .y = height, .z = 1};
	auto palm = Palm(input, pts.data(), cap, dim, 0, 10, 0, fitParams);
	return palm.AutoThreshold();
}

// One frame: label map (u32, raster) and the region list in label order as
// [x, y, intensity, surface] per region. Returns the region count (all of them;
// only min(count, infoCap) rows are written).
uint64_t ref_segment(const uint16_t* image, uint64_t height, uint64_t width, double threshold, double ratio,
                     uint32_t* labels, double* infos, uint64_t infoCap)
{
	const Vec2I dim{.x = width, .y = height};
	localization::ATrous at(dim);
	at.Process(image);
	localization::Segmentation2D seg(dim);
	seg.Process(at, threshold, ratio);
	if (labels) { std::copy(seg.map_.begin(), seg.map_.end(), labels); }
	const auto& inf = seg.Infos();
	for (uint64_t i = 0; i < inf.size() && i < infoCap; ++i) {
		infos[4 * i + 0] = inf[i].x_;
		infos[4 * i + 1] = inf[i].y_;
		infos[4 * i + 2] = inf[i].intensity_;
		infos[4 * i + 3] = inf[i].surf_;
	}
	return inf.size();
}

// All three planes of one frame (W0, W1, V2), each width*height doubles.
void ref_atrous(const uint16_t* image, uint64_t height, uint64_t width, double* w0, double* w1, double* v2)
{
	const Vec2I dim{.x = width, .y = height};
	localization::ATrous at(dim);
	at.Process(image);
	std::copy(at.wDecomp_[0].begin(), at.wDecomp_[0].end(), w0);
	std::copy(at.wDecomp_[1].begin(), at.wDecomp_[1].end(), w1);
	std::copy(at.wDecomp_[2].begin(), at.wDecomp_[2].end(), v2);
}

void ref_astig_estimate(const double* input, double* output, uint64_t size, double pixelSize, const double* model, double zMax)
{
	std::vector<Vec2D> data(size);
	for (uint64_t i = 0; i < size; ++i) { data[i] = {.x = input[2 * i], .y = input[2 * i + 1]}; }
	const fit::astig_model m = {std::array{model[0], model[1], model[2], model[3], model[4]},
	                            std::array{model[5], model[6], model[7], model[8], model[9]}};
	const std::array zl = {-zMax, zMax};
	const fit::AstigFit f(pixelSize);
	const auto res = f.Estimate(m, data, zl);
	for (uint64_t i = 0; i < size; ++i) {
		output[2 * i]     = res[i].x;
		output[2 * i + 1] = res[i].y;
	}
}
}

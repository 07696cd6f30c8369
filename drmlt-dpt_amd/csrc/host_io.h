// Host-side output files of a render (EXR film, statistics text).
#pragma once
#include <cstdint>
#include <string>

namespace drmlt {

constexpr int kExrNone = 0, kExrZip = 3;  // OpenEXR compression codes

uint16_t float_to_half(float f);
// rgb: w*h*3 floats, row-major, top row first. Returns 0 or -1 with err set.
int write_exr(const char *path, const float *rgb, int w, int h, int compression, std::string &err);
// s: large, large accepted, stage 1, accepted, stage-2 iso, accepted, H2MC,
// accepted, lens, accepted; format 0 / 1: the DR text with corrected / the
// reference's labels, 2: the mcmc and h2mc integrators' text
std::string stats_text(const uint64_t s[10], int format);
// a float as the reference's streams print it (default precision 6)
std::string float_text(float v);

}  // namespace drmlt

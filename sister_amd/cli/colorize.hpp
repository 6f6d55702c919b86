// colorize.hpp -- the display steps of the reference's compute_disp
// (cpp/src/compute_disp.cpp:37-48), written for files instead of windows.
//
//   normalize_minmax   cv::normalize(src, dst, 0, 255, NORM_MINMAX, CV_8UC1) (:38-40)
//   apply_magma        cv::applyColorMap(src, dst, COLORMAP_MAGMA)            (:41-43)
//   add_weighted       cv::addWeighted(center, 0.1, colour, 0.9, 0.0, out)   (:46-48)
//
// OpenCV is not available to compare against, so these follow OpenCV's
// documented arithmetic: the min-max scale and shift are computed in double
// and applied in float (src * scale + shift), blends are evaluated in float,
// and every 8-bit result is rounded half to even and saturated.  The MAGMA
// table itself is not reproduced: the colour map is matplotlib's "magma"
// approximated by the published degree-6 polynomial fit of each channel
// (Matt Zucker, "colormap polynomial fits", 2020, CC0), evaluated at i/255.
#ifndef SISTER_CLI_COLORIZE_HPP
#define SISTER_CLI_COLORIZE_HPP

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "../../include/sister/SisterMultiviewDisparities.hpp"

namespace colorize {

inline uint8_t sat_u8(float v)
{
    const float r = std::nearbyint(v);   // default rounding mode: half to even
    return (uint8_t)std::min(255.0f, std::max(0.0f, r));
}

// cv::normalize(..., 0, 255, NORM_MINMAX, CV_8UC1) of a one-channel uint16 map
inline sister::Mat8 normalize_minmax(const sister::Mat16& src)
{
    sister::Mat8 dst(src.rows, src.cols, 1);
    if (src.data.empty()) return dst;
    const auto mm = std::minmax_element(src.data.begin(), src.data.end());
    const double smin = *mm.first, smax = *mm.second;
    const double scale = (smax - smin) > DBL_EPSILON ? 255.0 / (smax - smin) : 0.0;
    const double shift = 0.0 - smin * scale;
    const float a = (float)scale, b = (float)shift;
    for (size_t i = 0; i < src.data.size(); i++) dst.data[i] = sat_u8(std::fma((float)src.data[i], a, b));
    return dst;
}

// 256-entry BGR table of the magma colour map
inline const std::array<std::array<uint8_t, 3>, 256>& magma_lut()
{
    static const std::array<std::array<uint8_t, 3>, 256> lut = [] {
        // per channel (R, G, B): c0 + t (c1 + t (c2 + ... t c6))
        static const double c[7][3] = {
            {-0.002136485053939582, -0.000749655052795221, -0.005386127855323933},
            {0.2516605407371642, 0.6775232436837668, 2.494026599312351},
            {8.353717279216625, -3.577719514958484, 0.3144679030132573},
            {-27.66873308576866, 14.26473078096533, -13.64921318813922},
            {52.17613981234068, -27.94360607168351, 12.94416944238394},
            {-50.76852536473588, 29.04658282127291, 4.23415299384598},
            {18.65570506591883, -11.48977351997711, -5.601961508734096},
        };
        std::array<std::array<uint8_t, 3>, 256> t{};
        for (int i = 0; i < 256; i++) {
            const double x = i / 255.0;
            double rgb[3];
            for (int ch = 0; ch < 3; ch++) {
                double v = c[6][ch];
                for (int k = 5; k >= 0; k--) v = c[k][ch] + x * v;
                rgb[ch] = std::min(1.0, std::max(0.0, v));
            }
            for (int ch = 0; ch < 3; ch++) t[i][2 - ch] = sat_u8((float)(rgb[ch] * 255.0));   // BGR
        }
        return t;
    }();
    return lut;
}

// cv::applyColorMap(gray, out, COLORMAP_MAGMA): one-channel 8-bit -> BGR
inline sister::Mat8 apply_magma(const sister::Mat8& gray)
{
    const auto& lut = magma_lut();
    sister::Mat8 out(gray.rows, gray.cols, 3);
    for (size_t i = 0; i < gray.data.size(); i++)
        for (int ch = 0; ch < 3; ch++) out.data[3 * i + ch] = lut[gray.data[i]][ch];
    return out;
}

// three-channel view of an 8-bit image (gray replicated, as imread gives)
inline sister::Mat8 to_bgr(const sister::Mat8& in)
{
    if (in.channels == 3) return in;
    sister::Mat8 out(in.rows, in.cols, 3);
    for (size_t i = 0; i < (size_t)in.rows * in.cols; i++) {
        const uint8_t g = in.data[i * in.channels];
        out.data[3 * i] = out.data[3 * i + 1] = out.data[3 * i + 2] = g;
    }
    return out;
}

// cv::addWeighted(a, alpha, b, beta, gamma, out) for equally shaped BGR images
inline sister::Mat8 add_weighted(const sister::Mat8& a, float alpha, const sister::Mat8& b, float beta, float gamma)
{
    sister::Mat8 out(a.rows, a.cols, a.channels);
    for (size_t i = 0; i < a.data.size(); i++)
        out.data[i] = sat_u8(std::fma((float)a.data[i], alpha, std::fma((float)b.data[i], beta, gamma)));
    return out;
}

}  // namespace colorize

#endif

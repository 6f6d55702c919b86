// compute_disp -- the reference's example entry point (cpp/src/compute_disp.cpp:7-59)
// on the MI355X path.
//
//   compute_disp <input folder/> <dmax>
//
// Same arguments: loads <folder>center.png, right.png, top.png, left.png and
// bottom.png (PNG or binary PGM/PPM content), runs compute_disparities with
// dmax disparities and, instead of opening windows (imshow, :51-56), writes
// into the folder
//   disp_multiview.png, disp_horizontal.png, disp_vertical.png   16-bit, the
//       returned maps (255*d, hpp:111-118)
//   disp_*_vis.png   8-bit min-max normalised views (cv::normalize NORM_MINMAX,
//       :38-40; the MAGMA colour map of :41-43 is not applied)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/sister/SisterMultiviewDisparities.hpp"
#include "png_io.hpp"

static void write_vis(const std::string& path, const sister::Mat16& m)
{
    uint16_t lo = 65535, hi = 0;
    for (uint16_t v : m.data) { lo = std::min(lo, v); hi = std::max(hi, v); }
    std::vector<uint16_t> n(m.data.size());
    const double s = hi > lo ? 255.0 / (hi - lo) : 0.0;
    for (size_t i = 0; i < n.size(); i++) n[i] = (uint16_t)(int)((m.data[i] - lo) * s + 0.5);
    pngio::write_png(path, n.data(), m.cols, m.rows, 8);
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cout << "expected <input folder> <dmax>" << std::endl;
        return -1;
    }
    const std::string folder = argv[1];
    const int dispCount = atoi(argv[2]);
    try {
        sister::Mat8 center = pngio::imread(folder + "center.png");
        sister::Mat8 right = pngio::imread(folder + "right.png");
        sister::Mat8 top = pngio::imread(folder + "top.png");
        sister::Mat8 left = pngio::imread(folder + "left.png");
        sister::Mat8 bottom = pngio::imread(folder + "bottom.png");

        SisterMultiviewDisparities multiview(center, right, top, left, bottom);
        sister::Mat16 result_horizontal, result_vertical, result_multiview;
        auto t0 = std::chrono::steady_clock::now();
        multiview.compute_disparities(dispCount, result_multiview, result_horizontal, result_vertical);
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "Disparities (multiview, horizontal, vertical) processed in "
                  << std::chrono::duration<double>(t1 - t0).count() << " sec." << std::endl;

        pngio::write_png(folder + "disp_multiview.png", result_multiview.data.data(), result_multiview.cols, result_multiview.rows, 16);
        pngio::write_png(folder + "disp_horizontal.png", result_horizontal.data.data(), result_horizontal.cols, result_horizontal.rows, 16);
        pngio::write_png(folder + "disp_vertical.png", result_vertical.data.data(), result_vertical.cols, result_vertical.rows, 16);
        write_vis(folder + "disp_multiview_vis.png", result_multiview);
        write_vis(folder + "disp_horizontal_vis.png", result_horizontal);
        write_vis(folder + "disp_vertical_vis.png", result_vertical);
    } catch (const std::exception& e) {
        std::cerr << "compute_disp: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

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
//   disp_*_color.png   min-max normalised to 8 bits and colour-mapped with
//       MAGMA (:38-43)
//   blended.png   0.1 * center + 0.9 * coloured multiview map (:46-48)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/sister/SisterMultiviewDisparities.hpp"
#include "colorize.hpp"
#include "png_io.hpp"

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

        // colorize (:37-43) and blend the multiview map over the center view (:45-48)
        const sister::Mat8 color_h = colorize::apply_magma(colorize::normalize_minmax(result_horizontal));
        const sister::Mat8 color_v = colorize::apply_magma(colorize::normalize_minmax(result_vertical));
        const sister::Mat8 color_m = colorize::apply_magma(colorize::normalize_minmax(result_multiview));
        const float alpha_blending_rgb = 0.1f;
        const sister::Mat8 blended =
            colorize::add_weighted(colorize::to_bgr(center), alpha_blending_rgb, color_m, 1.0f - alpha_blending_rgb, 0.0f);
        pngio::write_png_bgr(folder + "disp_horizontal_color.png", color_h);
        pngio::write_png_bgr(folder + "disp_vertical_color.png", color_v);
        pngio::write_png_bgr(folder + "disp_multiview_color.png", color_m);
        pngio::write_png_bgr(folder + "blended.png", blended);
    } catch (const std::exception& e) {
        std::cerr << "compute_disp: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

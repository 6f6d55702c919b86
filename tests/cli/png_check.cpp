// decodes a PNG with compute_disp's reader and prints the gray value of every
// pixel (cvtColor(BGR2GRAY) weights for colour input)
#include <cstdio>
#include "../../sister_amd/cli/png_io.hpp"
int main(int argc, char** argv)
{
    sister::Mat8 m = sister::to_gray(pngio::imread(argv[1]));
    std::printf("%d %d\n", m.rows, m.cols);
    for (int y = 0; y < m.rows; y++)
        for (int x = 0; x < m.cols; x++) std::printf("%d\n", m.ptr(y)[x]);
    return 0;
}

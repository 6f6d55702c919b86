// prints what compute_disp's display steps give for values named on the
// command line:  lut | norm v... | blend a b (one-pixel gray center a, one
// BGR pixel b b b)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "../../sister_amd/cli/colorize.hpp"
int main(int argc, char** argv)
{
    if (argc >= 2 && !std::strcmp(argv[1], "lut")) {
        for (const auto& e : colorize::magma_lut()) std::printf("%d %d %d\n", e[0], e[1], e[2]);
    } else if (argc >= 2 && !std::strcmp(argv[1], "norm")) {
        sister::Mat16 m(1, argc - 2);
        for (int i = 2; i < argc; i++) m.data[i - 2] = (uint16_t)std::atoi(argv[i]);
        const sister::Mat8 n = colorize::normalize_minmax(m);
        for (uint8_t v : n.data) std::printf("%d\n", v);
    } else if (argc == 4 && !std::strcmp(argv[1], "blend")) {
        sister::Mat8 g(1, 1, 1);
        g.data[0] = (uint8_t)std::atoi(argv[2]);
        sister::Mat8 c(1, 1, 3);
        c.data[0] = c.data[1] = c.data[2] = (uint8_t)std::atoi(argv[3]);
        const sister::Mat8 o = colorize::add_weighted(colorize::to_bgr(g), 0.1f, c, 1.0f - 0.1f, 0.0f);
        std::printf("%d %d %d\n", o.data[0], o.data[1], o.data[2]);
    } else {
        return 2;
    }
    return 0;
}

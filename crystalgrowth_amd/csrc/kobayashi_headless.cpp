// kobayashi_headless.cpp -- the reference's WinMain (src/main.cpp:11-38)
// without a window: Kobayashi(250, 250, 1e-4f), then `frames` iUpdate calls
// (10 steps each, as the GUI's play loop does).  Prints one JSON line with the
// FP64 reductions and, with --ppm FILE, writes the final colour frame.
//
//   kobayashi_headless [--size N] [--frames F] [--ppm out.ppm]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "Kobayashi.h"

int main(int argc, char **argv)
{
    int size = 250, frames = 100;
    std::string ppm;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--size") && i + 1 < argc)
            size = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--frames") && i + 1 < argc)
            frames = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--ppm") && i + 1 < argc)
            ppm = argv[++i];
        else {
            std::fprintf(stderr, "usage: %s [--size N] [--frames F] [--ppm FILE]\n", argv[0]);
            return 2;
        }
    }
    try {
        Kobayashi sim(size, size, 0.0001f);
        for (int f = 0; f < frames && sim.iIsUpdated(); ++f)
            sim.iUpdate();
        kg_stats st;
        if (kg_reduce(sim.handle(), &st) != KG_OK)
            throw std::runtime_error(kg_last_error());
        std::printf("{\"size\": %d, \"frames\": %d, \"steps\": %d, \"sum_phi\": %.17g, "
                    "\"sum_T\": %.17g, \"sum_T_minus_K_phi\": %.17g, \"solid_fraction\": %.17g, "
                    "\"max_phi\": %.17g, \"nonfinite\": %lld, \"sim_clock_ticks\": %ld}\n",
                    size, sim.simFrame(), 10 * sim.simFrame(), st.sum_phi, st.sum_T,
                    st.sum_T_minus_K_phi, st.solid_fraction, st.max_phi, (long long)st.nonfinite,
                    (long)sim.simTime());
        if (!ppm.empty()) {
            std::vector<KGColor> colors;
            sim.iUpdateConstantBuffer(colors);
            FILE *f = std::fopen(ppm.c_str(), "wb");
            if (!f)
                throw std::runtime_error("cannot open " + ppm);
            std::fprintf(f, "P6\n%d %d\n255\n", size, size);
            // entry i = (column i / n, row i % n); write rows top to bottom
            for (int row = size - 1; row >= 0; --row)
                for (int col = 0; col < size; ++col) {
                    const KGColor &c = colors[(size_t)col * size + row];
                    const unsigned char px[3] = {(unsigned char)(255.0f * c.r + 0.5f),
                                                 (unsigned char)(255.0f * c.g + 0.5f),
                                                 (unsigned char)(255.0f * c.b + 0.5f)};
                    std::fwrite(px, 1, 3, f);
                }
            std::fclose(f);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "kobayashi_headless: %s\n", e.what());
        return 1;
    }
    return 0;
}

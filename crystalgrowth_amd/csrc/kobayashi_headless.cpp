// kobayashi_headless.cpp -- the reference's WinMain (src/main.cpp:11-38)
// without a window: Kobayashi(nx, ny, dt) (250 x 250, 1e-4 by default), the
// viewer's object set-up (iCreateObject), then `frames` iUpdate calls (10 steps
// each) while iIsUpdated(), and every constant-buffer entry recoloured with
// iUpdateConstantBuffer(cb, i) at the end, as the viewer's draw does.
//
// The GUI's nine sliders (src/Kobayashi.cpp:11-59) become flags or lines of a
// config file; values outside a slider's range are accepted with a warning
// (kg_reference_range_mask), the model extensions (theta0, noise, boundary,
// seed) have flags of their own.  Setting parameters resets the simulation,
// as a slider move does (:616).  Prints one JSON line with the FP64
// reductions; --ppm FILE writes the final frame from the constant buffer.
//
//   kobayashi_headless [--size N | --nx N --ny N] [--frames F] [--dt V] [--dx V]
//                      [--tau V] [--eps-bar V] [--mu V] [--K V] [--delta V]
//                      [--aniso V] [--alpha V] [--gamma V] [--teq V]
//                      [--theta0 V] [--noise V] [--seed N]
//                      [--boundary periodic|zero-flux] [--config FILE] [--ppm FILE]
//   config file: one "name = value" per line (names as the flags without the
//   leading dashes), '#' starts a comment; flags after --config override it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Kobayashi.h"

namespace {

struct Options {
    int nx = 250, ny = 250, frames = 100;
    float dt = 0.0001f;  // src/main.cpp:16
    std::string ppm;
    // parameter overrides in the order given: name, value
    std::vector<std::pair<std::string, std::string>> params;
};

// Names of kg_reference_range_mask's bits (kobayashi_gpu.h).
const char *const kRangeNames[9] = {"tau", "eps-bar", "mu", "K", "delta", "aniso", "alpha",
                                    "gamma", "teq"};

bool is_param(const std::string &n)
{
    static const char *const names[] = {"dx",    "tau",   "eps-bar", "mu",     "K",
                                        "delta", "aniso", "alpha",   "gamma",  "teq",
                                        "theta0", "noise", "seed",   "boundary"};
    for (const char *k : names)
        if (n == k)
            return true;
    return false;
}

double to_double(const std::string &name, const std::string &v)
{
    char *end = nullptr;
    const double d = std::strtod(v.c_str(), &end);
    if (v.empty() || *end != '\0' || !std::isfinite(d))
        throw std::invalid_argument("--" + name + ": not a number: '" + v + "'");
    return d;
}

void apply(kg_params &p, const std::string &n, const std::string &v)
{
    if (n == "boundary") {
        if (v == "periodic")
            p.boundary = KG_PERIODIC;
        else if (v == "zero-flux")
            p.boundary = KG_ZERO_FLUX;
        else
            throw std::invalid_argument("--boundary: periodic or zero-flux, not '" + v + "'");
        return;
    }
    const double d = to_double(n, v);
    if (n == "dx")
        p.dx = p.dy = d;
    else if (n == "tau")
        p.tau = d;
    else if (n == "eps-bar")
        p.eps_bar = d;
    else if (n == "mu")
        p.mu = d;
    else if (n == "K")
        p.K = d;
    else if (n == "delta")
        p.delta = d;
    else if (n == "aniso")
        p.aniso = d;
    else if (n == "alpha")
        p.alpha = d;
    else if (n == "gamma")
        p.gamma = d;
    else if (n == "teq")
        p.teq = d;
    else if (n == "theta0")
        p.theta0 = d;
    else if (n == "noise")
        p.noise_a = d;
    else if (n == "seed")
        p.seed = (uint64_t)d;
}

void read_config(const std::string &path, Options &o)
{
    std::ifstream f(path);
    if (!f)
        throw std::runtime_error("cannot read config file " + path);
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        const size_t hash = line.find('#');
        if (hash != std::string::npos)
            line.erase(hash);
        const size_t eq = line.find('=');
        auto trim = [](std::string s) {
            const size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        if (trim(line).empty())
            continue;
        if (eq == std::string::npos)
            throw std::invalid_argument(path + ":" + std::to_string(lineno) + ": expected name = value");
        const std::string n = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
        if (n == "size")
            o.nx = o.ny = (int)to_double(n, v);
        else if (n == "nx")
            o.nx = (int)to_double(n, v);
        else if (n == "ny")
            o.ny = (int)to_double(n, v);
        else if (n == "frames")
            o.frames = (int)to_double(n, v);
        else if (n == "dt")
            o.dt = (float)to_double(n, v);
        else if (is_param(n))
            o.params.emplace_back(n, v);
        else
            throw std::invalid_argument(path + ":" + std::to_string(lineno) + ": unknown name '" + n + "'");
    }
}

Options parse(int argc, char **argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.rfind("--", 0) != 0 || i + 1 >= argc)
            throw std::invalid_argument("bad argument '" + a + "'");
        const std::string n = a.substr(2), v = argv[++i];
        if (n == "size")
            o.nx = o.ny = std::atoi(v.c_str());
        else if (n == "nx")
            o.nx = std::atoi(v.c_str());
        else if (n == "ny")
            o.ny = std::atoi(v.c_str());
        else if (n == "frames")
            o.frames = std::atoi(v.c_str());
        else if (n == "dt")
            o.dt = (float)to_double(n, v);
        else if (n == "ppm")
            o.ppm = v;
        else if (n == "config")
            read_config(v, o);
        else if (is_param(n))
            o.params.emplace_back(n, v);
        else
            throw std::invalid_argument("unknown flag '" + a + "'");
    }
    return o;
}

void write_ppm(const std::string &path, int nx, int ny, const std::vector<ConstantBuffer> &cb)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        throw std::runtime_error("cannot open " + path);
    std::fprintf(f, "P6\n%d %d\n255\n", nx, ny);
    // a square grid's entry i shows cell (x = i / n, y = i % n)
    // (src/Kobayashi.cpp:311-315); other shapes are x fastest.  Rows are
    // written top (largest y) to bottom.
    auto unorm8 = [](float c) {  // as kg_colormap_rgba8: clamp, scale, round to nearest even
        c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
        return (unsigned char)std::nearbyint(255.0f * c);
    };
    for (int y = ny - 1; y >= 0; --y)
        for (int x = 0; x < nx; ++x) {
            const size_t e = nx == ny ? (size_t)x * ny + y : (size_t)y * nx + x;
            const DirectX::XMFLOAT4 &c = cb[e].color;
            const unsigned char px[3] = {unorm8(c.x), unorm8(c.y), unorm8(c.z)};
            std::fwrite(px, 1, 3, f);
        }
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        const Options o = parse(argc, argv);
        // the parameters the flags ask for, checked before the GPU is touched
        kg_params p;
        if (kg_default_params(&p) != KG_OK)
            throw std::runtime_error(kg_last_error());
        p.dt = (double)o.dt;
        for (const auto &nv : o.params)
            apply(p, nv.first, nv.second);

        Kobayashi sim(o.nx, o.ny, o.dt);
        std::vector<ConstantBuffer> cb;
        sim.iCreateObject(cb);

        std::string warned;
        if (!o.params.empty()) {
            uint32_t mask = 0;
            if (kg_reference_range_mask(&p, &mask) != KG_OK)
                throw std::runtime_error(kg_last_error());
            for (int k = 0; k < 9; ++k)
                if (mask & (1u << k)) {
                    std::fprintf(stderr,
                                 "kobayashi_headless: warning: %s is outside the reference GUI's "
                                 "slider range (src/Kobayashi.cpp:17-58)\n",
                                 kRangeNames[k]);
                    warned += std::string(warned.empty() ? "" : ", ") + "\"" + kRangeNames[k] + "\"";
                }
            sim.setParameters(p, cb);  // a slider move: the simulation restarts
        }

        for (int f = 0; f < o.frames && sim.iIsUpdated(); ++f)
            sim.iUpdate();
        for (size_t i = 0; i < cb.size(); ++i)
            sim.iUpdateConstantBuffer(cb, (int)i);

        double color_sum = 0.0;
        for (const ConstantBuffer &c : cb)
            color_sum += (double)c.color.x + (double)c.color.y + (double)c.color.z;
        kg_stats st;
        if (kg_reduce(sim.handle(), &st) != KG_OK)
            throw std::runtime_error(kg_last_error());
        p = sim.parameters();
        const DirectX::XMINT3 count = sim.iGetObjectCount();
        std::printf("{\"nx\": %d, \"ny\": %d, \"frames\": %d, \"steps\": %d, \"dt\": %.9g, "
                    "\"tau\": %.9g, \"eps_bar\": %.9g, \"K\": %.9g, \"delta\": %.9g, "
                    "\"aniso\": %.9g, \"theta0\": %.9g, \"noise_a\": %.9g, \"boundary\": %d, "
                    "\"out_of_reference_range\": [%s], \"constant_buffers\": %zu, "
                    "\"color_sum\": %.17g, \"sum_phi\": %.17g, \"sum_T\": %.17g, "
                    "\"sum_T_minus_K_phi\": %.17g, \"solid_fraction\": %.17g, "
                    "\"max_phi\": %.17g, \"nonfinite\": %lld, \"sim_clock_ticks\": %ld}\n",
                    count.x, count.y, sim.simFrame(), 10 * sim.simFrame(), p.dt, p.tau,
                    p.eps_bar, p.K, p.delta, p.aniso, p.theta0, p.noise_a, p.boundary,
                    warned.c_str(), cb.size(), color_sum, st.sum_phi, st.sum_T,
                    st.sum_T_minus_K_phi, st.solid_fraction, st.max_phi,
                    (long long)st.nonfinite, (long)sim.simTime());
        if (!o.ppm.empty())
            write_ppm(o.ppm, o.nx, o.ny, cb);
    } catch (const std::invalid_argument &e) {
        std::fprintf(stderr, "kobayashi_headless: %s\n", e.what());
        return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "kobayashi_headless: %s\n", e.what());
        return 1;
    }
    return 0;
}

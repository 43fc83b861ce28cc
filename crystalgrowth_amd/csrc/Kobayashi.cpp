// Kobayashi.cpp -- headless Kobayashi class: every simulation call forwards to
// the C ABI (include/kobayashi_gpu.h); errors become std::runtime_error.
#include "Kobayashi.h"

#include <stdexcept>
#include <string>

namespace {

void check(int rc, const char *what)
{
    if (rc != KG_OK)
        throw std::runtime_error(std::string(what) + ": " + kg_last_error());
}

}  // namespace

Kobayashi::Kobayashi(int x, int y, float timeStep) : _nx(x), _ny(y)
{
    kg_params p;
    check(kg_default_params(&p), "kg_default_params");
    p.dt = (double)timeStep;
    check(kg_create(x, y, KG_F32, &p, &_h), "kg_create");
}

Kobayashi::~Kobayashi()
{
    if (_h)
        kg_destroy(_h);
}

void Kobayashi::iUpdate()
{
    const clock_t start = clock();
    check(kg_step(_h, 10), "kg_step");
    check(kg_synchronize(_h), "kg_synchronize");
    _simTime += clock() - start;
    _simFrame++;
}

void Kobayashi::iResetSimulationState()
{
    check(kg_reset(_h), "kg_reset");
    _simTime = 0;
    _simFrame = 0;
}

void Kobayashi::iUpdateConstantBuffer(std::vector<KGColor> &colors)
{
    colors.resize((size_t)_nx * _ny);
    check(kg_colormap(_h, reinterpret_cast<float *>(colors.data()), colors.size() * sizeof(KGColor),
                      _nx == _ny ? 1 : 0),
          "kg_colormap");
}

KGInt3 Kobayashi::iGetObjectCount() const { return {_nx, _ny, 0}; }

bool Kobayashi::iIsUpdated() const { return _updateFlag; }

void Kobayashi::setUpdateFlag(bool on) { _updateFlag = on; }

kg_params Kobayashi::parameters() const
{
    kg_params p;
    check(kg_get_params(_h, &p), "kg_get_params");
    return p;
}

void Kobayashi::setParameters(const kg_params &p, bool reset)
{
    check(kg_set_params(_h, &p), "kg_set_params");
    if (reset)  // every slider change resets the simulation (src/Kobayashi.cpp:616)
        iResetSimulationState();
}

const std::vector<float> &Kobayashi::phase()
{
    _phi.resize((size_t)_nx * _ny);
    check(kg_get_phase(_h, _phi.data(), _phi.size() * sizeof(float)), "kg_get_phase");
    return _phi;
}

const std::vector<float> &Kobayashi::temperature()
{
    _t.resize((size_t)_nx * _ny);
    check(kg_get_temperature(_h, _t.data(), _t.size() * sizeof(float)), "kg_get_temperature");
    return _t;
}

const std::vector<float> &Kobayashi::angle()
{
    _angl.resize((size_t)_nx * _ny);
    check(kg_get_angle(_h, _angl.data(), _angl.size() * sizeof(float)), "kg_get_angle");
    return _angl;
}

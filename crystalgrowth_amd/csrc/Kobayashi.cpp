// Kobayashi.cpp -- the reference's Kobayashi class over the C ABI
// (include/kobayashi_gpu.h); every simulation call forwards to the GPU
// solver, and errors become std::runtime_error.
#include "Kobayashi.h"

#include <stdexcept>
#include <string>

namespace {

void check(int rc, const char *what)
{
    if (rc != KG_OK)
        throw std::runtime_error(std::string(what) + ": " + kg_last_error());
}

// Row-vector affine transform: translation (x, y, z), uniform scale s
// (src/Kobayashi.cpp:299 places cell (i, j) at (i, j, 0) with scale 1).
DirectX::XMFLOAT4X4 translation(float x, float y, float z, float s)
{
    DirectX::XMFLOAT4X4 m = {};
    m.m[0][0] = m.m[1][1] = m.m[2][2] = s;
    m.m[3][3] = 1.0f;
    m.m[3][0] = x;
    m.m[3][1] = y;
    m.m[3][2] = z;
    return m;
}

}  // namespace

Kobayashi::Kobayashi(int x, int y, float timeStep) : _objectCount{x, y}, _dt(timeStep)
{
    kg_params p;
    check(kg_default_params(&p), "kg_default_params");
    p.dt = (double)timeStep;  // _dt = timeStep (src/Kobayashi.cpp:63)
    check(kg_create(x, y, KG_F32, &p, &_h), "kg_create");
}

Kobayashi::~Kobayashi()
{
    if (_h)
        kg_destroy(_h);
}

// ---- simulation methods -----------------------------------------------------
void Kobayashi::iUpdate()
{
    const clock_t start = clock();
    check(kg_step(_h, 10), "kg_step");
    check(kg_synchronize(_h), "kg_synchronize");
    _simTime += clock() - start;
    _simFrame++;
    _frameId++;
}

void Kobayashi::iResetSimulationState(std::vector<ConstantBuffer> &constantBuffer)
{
    check(kg_reset(_h), "kg_reset");
    _simTime = 0;
    _simFrame = 0;
    _frameId++;
    // The reference then has the viewer update and redraw (src/Kobayashi.cpp:
    // 245-246), which recolours every entry from the new state; headless,
    // the entries are recoloured here.
    for (size_t i = 0; i < constantBuffer.size(); ++i)
        iUpdateConstantBuffer(constantBuffer, (int)i);
}

// ---- mesh methods: one unit quad per cell -----------------------------------
std::vector<Vertex> &Kobayashi::iGetVertices()
{
    _vertices = {Vertex{{-0.5f, -0.5f, -0.0f}}, Vertex{{-0.5f, +0.5f, -0.0f}},
                 Vertex{{+0.5f, +0.5f, -0.0f}}, Vertex{{+0.5f, -0.5f, -0.0f}}};
    return _vertices;
}

std::vector<unsigned int> &Kobayashi::iGetIndices()
{
    _indices = {0, 1, 2, 0, 2, 3};
    return _indices;
}

unsigned int Kobayashi::iGetVertexBufferSize() { return 4; }

unsigned int Kobayashi::iGetIndexBufferSize() { return 6; }

// ---- constant-buffer methods --------------------------------------------------
void Kobayashi::iCreateObject(std::vector<ConstantBuffer> &constantBuffer)
{
    for (int j = 0; j < _objectCount.y; j++) {
        for (int i = 0; i < _objectCount.x; i++) {
            ConstantBuffer cb;
            cb.world = translation((float)i, (float)j, 0.0f, 1.0f);
            cb.worldViewProj = translation(0.0f, 0.0f, 0.0f, 1.0f);
            cb.transInvWorld = translation(0.0f, 0.0f, 0.0f, 1.0f);
            cb.color = {0.0f, 0.0f, 0.0f, 1.0f};
            constantBuffer.push_back(cb);
        }
    }
}

// Colours of `entries` constant-buffer entries from the device ramp.  The
// reference colours entry i with _phi[_INDEX(i / n, i % n)], n = sqrt(entries)
// (src/Kobayashi.cpp:311-315), which reads inside the field only for a square
// grid with one entry per cell; kg_colormap's order 1 is that mapping.  Other
// shapes get one entry per cell in x-fastest order.
void Kobayashi::refreshColors(size_t entries)
{
    const size_t cells = (size_t)_objectCount.x * (size_t)_objectCount.y;
    if (entries != cells)
        throw std::out_of_range("constant buffer holds " + std::to_string(entries) +
                                " entries, the grid " + std::to_string(cells) + " cells");
    _colors.resize(4 * cells);
    check(kg_colormap(_h, _colors.data(), _colors.size() * sizeof(float),
                      _objectCount.x == _objectCount.y ? 1 : 0),
          "kg_colormap");
    _colorFrame = _frameId;
}

void Kobayashi::iUpdateConstantBuffer(std::vector<ConstantBuffer> &constantBuffer, int i)
{
    if (i < 0 || (size_t)i >= constantBuffer.size())
        throw std::out_of_range("constant buffer entry " + std::to_string(i));
    if (_colorFrame != _frameId || _colors.size() != 4 * constantBuffer.size())
        refreshColors(constantBuffer.size());
    const float *c = &_colors[4 * (size_t)i];
    constantBuffer[i].color = {c[0], c[1], c[2], c[3]};
}

void Kobayashi::iSetDXApp(DX12App *dxApp) { _dxapp = dxApp; }

unsigned int Kobayashi::iGetConstantBufferSize()
{
    return (unsigned int)(_objectCount.x * _objectCount.y * 2);
}

DirectX::XMINT3 Kobayashi::iGetObjectCount() { return {_objectCount.x, _objectCount.y, 0}; }

DirectX::XMFLOAT3 Kobayashi::iGetObjectSize() { return {1.0f, 1.0f, 0.0f}; }

DirectX::XMFLOAT3 Kobayashi::iGetObjectPositionOffset() { return {0.0f, 0.0f, 0.0f}; }

bool Kobayashi::iIsUpdated() { return _updateFlag; }

// ---- the GUI's controls ---------------------------------------------------------
void Kobayashi::togglePlay() { _updateFlag = !_updateFlag; }

kg_params Kobayashi::parameters() const
{
    kg_params p;
    check(kg_get_params(_h, &p), "kg_get_params");
    return p;
}

void Kobayashi::setParameters(const kg_params &p, std::vector<ConstantBuffer> &constantBuffer)
{
    check(kg_set_params(_h, &p), "kg_set_params");
    // every slider change resets the simulation (src/Kobayashi.cpp:616)
    iResetSimulationState(constantBuffer);
}

void Kobayashi::resetParameters(std::vector<ConstantBuffer> &constantBuffer)
{
    kg_params p;
    check(kg_default_params(&p), "kg_default_params");
    p.dt = (double)_dt;
    setParameters(p, constantBuffer);
}

// ---- fields ---------------------------------------------------------------------
const std::vector<float> &Kobayashi::phase()
{
    _phi.resize((size_t)_objectCount.x * _objectCount.y);
    check(kg_get_phase(_h, _phi.data(), _phi.size() * sizeof(float)), "kg_get_phase");
    return _phi;
}

const std::vector<float> &Kobayashi::temperature()
{
    _t.resize((size_t)_objectCount.x * _objectCount.y);
    check(kg_get_temperature(_h, _t.data(), _t.size() * sizeof(float)), "kg_get_temperature");
    return _t;
}

const std::vector<float> &Kobayashi::angle()
{
    _angl.resize((size_t)_objectCount.x * _objectCount.y);
    check(kg_get_angle(_h, _angl.data(), _angl.size() * sizeof(float)), "kg_get_angle");
    return _angl;
}

// Kobayashi.h -- headless C++ Kobayashi class over the C ABI.
//
// Keeps the simulation surface of the reference class (src/Kobayashi.h:32-67):
// the constructor Kobayashi(int x, int y, float timeStep), iUpdate (10 steps
// per call, src/Kobayashi.cpp:227-239), iResetSimulationState (:241-249),
// iUpdateConstantBuffer's phi colour ramp (:309-345), iGetObjectCount (:363),
// iIsUpdated (:378).  The Win32/D3D12 members are not here: a DXViewer build
// keeps them in its own subclass and forwards the simulation calls to this one.
#pragma once

#include <ctime>
#include <vector>

#include "kobayashi_gpu.h"

struct KGInt3 {
    int x, y, z;
};

struct KGColor {
    float r, g, b, a;
};

class Kobayashi {
public:
    Kobayashi(int x, int y, float timeStep);
    ~Kobayashi();
    Kobayashi(const Kobayashi &) = delete;
    Kobayashi &operator=(const Kobayashi &) = delete;

    // Simulation methods (src/Kobayashi.h:41-42)
    void iUpdate();
    void iResetSimulationState();

    // Colour of every constant-buffer entry, in the reference's entry order
    // (entry i shows _phi[_INDEX(i / n, i % n)], square grids).
    void iUpdateConstantBuffer(std::vector<KGColor> &colors);
    KGInt3 iGetObjectCount() const;
    bool iIsUpdated() const;
    void setUpdateFlag(bool on);

    // Parameters: the slider/RESET writes of the GUI (src/Kobayashi.cpp:512-616)
    kg_params parameters() const;
    void setParameters(const kg_params &p, bool reset = true);

    // Headless field accessors (_phi, _t, _angl of src/Kobayashi.h:107-115)
    const std::vector<float> &phase();
    const std::vector<float> &temperature();
    const std::vector<float> &angle();

    clock_t simTime() const { return _simTime; }
    int simFrame() const { return _simFrame; }
    kg_handle handle() const { return _h; }

private:
    kg_handle _h = nullptr;
    int _nx = 0, _ny = 0;
    bool _updateFlag = true;
    clock_t _simTime = 0;
    int _simFrame = 0;
    std::vector<float> _phi, _t, _angl;
};

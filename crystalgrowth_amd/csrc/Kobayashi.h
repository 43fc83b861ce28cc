// Kobayashi.h -- the reference's Kobayashi class (src/Kobayashi.h:32-67) over
// the C ABI of the MI355X solver.
//
// It implements ISimulation.h with the reference's signatures: the
// constructor Kobayashi(int x, int y, float timeStep); iUpdate, 10 steps per
// call (src/Kobayashi.cpp:227-239); iResetSimulationState(
// std::vector<ConstantBuffer>&) (:241-249); the quad mesh (:253-283);
// iCreateObject (:287-307); iUpdateConstantBuffer(std::vector<ConstantBuffer>&,
// int i), entry i coloured by the phi ramp of :309-345; iGetObjectCount ->
// XMINT3 (:363-366) and the other queries of :353-381.  The fields live on the
// GPU; iUpdateConstantBuffer reads the colours of all entries back once per
// frame (one device ramp and one copy) and then serves entries from that copy.
// Beyond the reference: the GUI's parameter writes as calls (slider moves and
// RESET reset the simulation, :507-618), and phi / T / theta accessors.
#pragma once

#include <ctime>
#include <vector>

#include "ISimulation.h"
#include "kobayashi_gpu.h"

class Kobayashi : public ISimulation {
public:
    Kobayashi(int x, int y, float timeStep);
    ~Kobayashi() override;
    Kobayashi(const Kobayashi &) = delete;
    Kobayashi &operator=(const Kobayashi &) = delete;

    // Simulation methods
    void iUpdate() override;
    void iResetSimulationState(std::vector<ConstantBuffer> &constantBuffer) override;

    // Mesh methods
    std::vector<Vertex> &iGetVertices() override;
    std::vector<unsigned int> &iGetIndices() override;
    unsigned int iGetVertexBufferSize() override;
    unsigned int iGetIndexBufferSize() override;

    // Constant-buffer methods
    void iCreateObject(std::vector<ConstantBuffer> &constantBuffer) override;
    void iUpdateConstantBuffer(std::vector<ConstantBuffer> &constantBuffer, int i) override;
    void iSetDXApp(DX12App *dxApp) override;
    unsigned int iGetConstantBufferSize() override;
    DirectX::XMINT3 iGetObjectCount() override;
    DirectX::XMFLOAT3 iGetObjectSize() override;
    DirectX::XMFLOAT3 iGetObjectPositionOffset() override;
    bool iIsUpdated() override;

    // The GUI's controls without the window (src/Kobayashi.cpp:507-618):
    // PLAY toggles the update flag; a slider move writes one parameter and
    // resets the simulation; RESET restores the defaults (_parameterInit,
    // :73-96) and resets.
    void togglePlay();
    kg_params parameters() const;
    void setParameters(const kg_params &p, std::vector<ConstantBuffer> &constantBuffer);
    void resetParameters(std::vector<ConstantBuffer> &constantBuffer);

    // Fields (_phi, _t, _angl of src/Kobayashi.h:107-115), copied from the GPU
    const std::vector<float> &phase();
    const std::vector<float> &temperature();
    const std::vector<float> &angle();

    clock_t simTime() const { return _simTime; }
    int simFrame() const { return _simFrame; }
    kg_handle handle() const { return _h; }

private:
    void refreshColors(size_t entries);

    kg_handle _h = nullptr;
    DirectX::XMINT2 _objectCount = {0, 0};
    float _dt = 0.0f;
    bool _updateFlag = true;
    clock_t _simTime = 0;
    int _simFrame = 0;
    DX12App *_dxapp = nullptr;
    std::vector<Vertex> _vertices;
    std::vector<unsigned int> _indices;
    std::vector<float> _phi, _t, _angl;
    // colours of every constant-buffer entry (RGBA floats), valid for frame
    // _colorFrame of the current run only
    std::vector<float> _colors;
    long long _colorFrame = -1;
    long long _frameId = 0;  // advances with every update and reset
};

// ISimulation.h -- the simulation interface the reference's Kobayashi class
// implements (DXViewer's ISimulation, whose overrides are declared at
// src/Kobayashi.h:32-67), with the value types its signatures use, for builds
// without the viewer (DXViewer is Windows-only and not vendored in the
// reference: .gitmodules:1-3).
//
// The methods keep the reference's names, parameter lists and return types:
//   simulation   iUpdate, iResetSimulationState(std::vector<ConstantBuffer>&)
//   mesh         iGetVertices, iGetIndices, iGetVertexBufferSize, iGetIndexBufferSize
//   constant     iCreateObject, iUpdateConstantBuffer(std::vector<ConstantBuffer>&, int),
//   buffers      iSetDXApp, iGetConstantBufferSize, iGetObjectCount (XMINT3),
//                iGetObjectSize, iGetObjectPositionOffset, iIsUpdated
// The members that need Direct3D 12 or Win32 (iDraw with its command list,
// and the iWM* window-message handlers) are not part of this interface: a
// viewer build adds them in its own subclass.
#pragma once

#include <vector>

// The DirectXMath value types of the interface (plain structs of the same
// members; a build that has DirectXMath uses its own definitions instead of
// this header).
namespace DirectX {
struct XMINT2 {
    int x, y;
};
struct XMINT3 {
    int x, y, z;
};
struct XMFLOAT2 {
    float x, y;
};
struct XMFLOAT3 {
    float x, y, z;
};
struct XMFLOAT4 {
    float x, y, z, w;
};
struct XMFLOAT4X4 {
    float m[4][4];
};
}  // namespace DirectX

// One vertex of the cell quad (src/Kobayashi.cpp:257-260).
struct Vertex {
    DirectX::XMFLOAT3 pos;
};

// Per-cell constant buffer: the fields src/Kobayashi.cpp:299-302 writes.
struct ConstantBuffer {
    DirectX::XMFLOAT4X4 world;
    DirectX::XMFLOAT4X4 worldViewProj;
    DirectX::XMFLOAT4X4 transInvWorld;
    DirectX::XMFLOAT4 color;
};

// The viewer's renderer; the simulation only keeps the pointer it is given.
class DX12App;

class ISimulation {
public:
    virtual ~ISimulation() = default;

    // Simulation methods
    virtual void iUpdate() = 0;
    virtual void iResetSimulationState(std::vector<ConstantBuffer> &constantBuffer) = 0;

    // Mesh methods
    virtual std::vector<Vertex> &iGetVertices() = 0;
    virtual std::vector<unsigned int> &iGetIndices() = 0;
    virtual unsigned int iGetVertexBufferSize() = 0;
    virtual unsigned int iGetIndexBufferSize() = 0;

    // Constant-buffer methods
    virtual void iCreateObject(std::vector<ConstantBuffer> &constantBuffer) = 0;
    virtual void iUpdateConstantBuffer(std::vector<ConstantBuffer> &constantBuffer, int i) = 0;
    virtual void iSetDXApp(DX12App *dxApp) = 0;
    virtual unsigned int iGetConstantBufferSize() = 0;
    virtual DirectX::XMINT3 iGetObjectCount() = 0;
    virtual DirectX::XMFLOAT3 iGetObjectSize() = 0;
    virtual DirectX::XMFLOAT3 iGetObjectPositionOffset() = 0;
    virtual bool iIsUpdated() = 0;
};

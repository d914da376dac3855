// SWE_Block.hh -- base class of an SWE grid block with the reference's public
// interface (src/blocks/SWE_Block.hh:110-251), so that drivers, scenarios
// and writers written against it compile unchanged against blocks whose
// state lives on a GPU.
//
// Storage: h, hu, hv, b are (nx+2) x (ny+2) Float2D arrays with ghost cells
// at 0 and n+1.  Subclasses that keep the authoritative state elsewhere (the
// GPU) override the synch* hooks; every accessor that hands out a host array
// calls the matching hook first, the lazy synchronisation contract that the
// reference documents (SWE_Block.hh:57-82).
//
// Local time stepping (LTS): the per-block state machine of the reference
// (SWE_Block.hh:311-562) with its semantics kept as they are -- power-of-two
// sub-steps of the macro step T derived from the absolute step size, ghost
// interpolation weighted by absolute local times, and a macro step that
// ends once stepSizeCounter has reached stepSize.
#ifndef SWE_HIP_SWE_BLOCK_HH
#define SWE_HIP_SWE_BLOCK_HH

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <limits>
#include <string>

#include "Constants.hh"
#include "scenarios/SWE_Scenario.hh"
#include "tools/Float2DBuffer.hh"
#include "tools/Float2DNative.hh"
#include "types/Boundary.hh"

#define SWE_HIP_BLOCK_MIRROR 1

template <typename T, typename Buffer = Float2DBuffer>
class SWE_Block {
public:
    // getters
    int getCellCountHorizontal() { return nx; }
    int getCellCountVertical() { return ny; }
    float getCellSizeHorizontal() { return dx; }
    float getCellSizeVertical() { return dy; }
    int getOriginX() { return originX; }
    int getOriginY() { return originY; }
    float getMaxTimestep() { return maxTimestep; }

    const T& getWaterHeight() {
        synchWaterHeightBeforeRead();
        return h;
    }
    T& getModifiableWaterHeight() {
        synchWaterHeightBeforeRead();
        return h;
    }
    const T& getMomentumHorizontal() {
        synchDischargeBeforeRead();
        return hu;
    }
    T& getModifiableMomentumHorizontal() {
        synchDischargeBeforeRead();
        return hu;
    }
    const T& getMomentumVertical() {
        synchDischargeBeforeRead();
        return hv;
    }
    T& getModifiableMomentumVertical() {
        synchDischargeBeforeRead();
        return hv;
    }
    const T& getBathymetry() {
        synchBathymetryBeforeRead();
        return b;
    }
    T& getModifiableBathymetry() {
        synchBathymetryBeforeRead();
        return b;
    }

    // lazy synchronisation hooks (no-ops for host-resident blocks)
    virtual void synchAfterWrite() {}
    virtual void synchWaterHeightAfterWrite() {}
    virtual void synchDischargeAfterWrite() {}
    virtual void synchBathymetryAfterWrite() {}
    virtual void synchGhostLayerAfterWrite() {}
    virtual void synchBeforeRead() {}
    virtual void synchWaterHeightBeforeRead() {}
    virtual void synchDischargeBeforeRead() {}
    virtual void synchBathymetryBeforeRead() {}
    virtual void synchCopyLayerBeforeRead() {}

    // local time stepping
    void checkAllGhostlayers();
    float getRoundTimestep(float timestep);
    void setMaxGlobalTimestep(float timestep) { maxTimestep = timestep; }
    void interpolateGhostlayer(Boundary border, float remoteTimestep);
    float interpolateValue(float oldval, float newval, float remoteTimestep, Boundary boundary) {
        return (oldval + (newval - oldval) * (getTotalLocalTimestep() / remoteTimestep));
    }
    float getTotalLocalTimestep() {
        return (float)(maxTimestepLocal * timestepCounter) +
               (stepSize > 0 ? ((float)(maxTimestepLocal * stepSizeCounter) / (float)stepSize) : 0);
    }
    void setMaxLocalTimestep(float timestep) { maxTimestepLocal = timestep; }
    void printLtsStats();
    bool hasMaxLocalTimestep() { return (stepSizeCounter >= stepSize && allGhostlayersInSync()) || notifiedLastTimestep; }
    bool isReceivable(Boundary border) {
        if (notifiedLastTimestep) return false;
        const GhostlayerState s = receivedGhostlayer[border];
        return s == GL_NEXT || s == GL_UNVALID || s == GL_SYNC;
    }
    bool isSendable(Boundary border) {
        if (notifiedLastTimestep) return false;
        const GhostlayerState s = receivedGhostlayer[border];
        return s == GL_NEXT || s == GL_INTER || s == GL_SYNC;
    }
    bool allGhostlayersInSync() {
        return receivedGhostlayer[BND_LEFT] != GL_UNVALID && receivedGhostlayer[BND_RIGHT] != GL_UNVALID &&
               receivedGhostlayer[BND_TOP] != GL_UNVALID && receivedGhostlayer[BND_BOTTOM] != GL_UNVALID &&
               (stepSizeCounter <= stepSize);
    }

    virtual void setBoundaryType(Boundary boundary, BoundaryType type) { boundaryType[boundary] = type; }
    virtual void initScenario(SWE_Scenario& scenario, BoundaryType boundaries[]);
    virtual void computeMaxTimestep(const float dryTol = defaultDryTol, const float cflNumber = defaultCflNumber);

    SWE_Block() {}
    SWE_Block(int cellCountHorizontal, int cellCountVertical, float cellSizeHorizontal, float cellSizeVertical,
              float originX = 0, float originY = 0, bool localTimestepping = false);
    virtual ~SWE_Block() = 0;

    virtual void applyBoundaryBathymetry();
    virtual void applyBoundaryConditions();

    virtual void computeNumericalFluxes() = 0;
    virtual void updateUnknowns(float dt) = 0;

    int nx;
    int ny;
    float dx;
    float dy;
    float originX;
    float originY;
    float duration;
    void setDuration(float d) { duration = d; }

    float maxTimestep;
    float maxTimestepLocal;
    float currentTimestep = 0;
    int maxDivisor = 1024;
    bool localTimestepping;
    bool notifiedLastTimestep = false;
    int stepSize;
    int stepSizeCounter;
    void resetStepSizeCounter() {
        stepSizeCounter = 0;
        stepSize = 0;
        timestepCounter++;
    }
    GhostlayerState receivedGhostlayer[4];
    float borderTimestep[4];
    int timestepCounter = 0;
    int myRank;
    int neighbourRankId[4];
    int iteration = 0;
    void connectNeighbours(int ids[]) {
        for (int i = 0; i < 4; i++) neighbourRankId[i] = ids[i];
    }
    void setRank(int rank) { myRank = rank; }

    T h;
    T hu;
    T hv;
    T b;

    Buffer bufferH;
    Buffer bufferHu;
    Buffer bufferHv;
    BoundaryType boundaryType[4];

    void copyGhostlayer(Boundary border);
    std::string stateToString(GhostlayerState state);
};

template <typename T, typename Buffer>
SWE_Block<T, Buffer>::SWE_Block(int nx_, int ny_, float dx_, float dy_, float ox, float oy, bool lts)
    : nx(nx_),
      ny(ny_),
      dx(dx_),
      dy(dy_),
      originX(ox),
      originY(oy),
      localTimestepping(lts),
      h(nx_ + 2, ny_ + 2),
      hu(nx_ + 2, ny_ + 2),
      hv(nx_ + 2, ny_ + 2),
      b(nx_ + 2, ny_ + 2),
      bufferH(nx_ + 2, ny_ + 2, lts, h),
      bufferHu(nx_ + 2, ny_ + 2, lts, hu),
      bufferHv(nx_ + 2, ny_ + 2, lts, hv) {
    for (int i = 0; i < 4; i++) {
        boundaryType[i] = PASSIVE;
        receivedGhostlayer[i] = GL_NEXT;
        borderTimestep[i] = 0;
    }
    stepSizeCounter = 0;
    stepSize = 0;
    maxTimestep = 0;
    maxTimestepLocal = 1;
    duration = std::numeric_limits<float>::max();
}

template <typename T, typename Buffer>
SWE_Block<T, Buffer>::~SWE_Block() {}

template <typename T, typename Buffer>
std::string SWE_Block<T, Buffer>::stateToString(GhostlayerState state) {
    switch (state) {
        case GL_UNVALID:
            return "GL_UNVALID";
        case GL_NEXT:
            return "GL_NEXT";
        case GL_SYNC:
            return "GL_SYNC";
        case GL_INTER:
            return "GL_INTER";
        default:
            return "GL_DONE";
    }
}

template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::printLtsStats() {
    std::cout << myRank << "|" << timestepCounter << "|" << iteration << std::setprecision(6) << std::fixed;
    for (int s = 0; s < 4; s++) std::cout << " " << stateToString(receivedGhostlayer[s]) << ":" << borderTimestep[s];
    std::cout << " | " << getTotalLocalTimestep() << std::endl;
}

// Neighbour-time comparison for every connected edge: same time -> copy the
// received layer (GL_SYNC); neighbour ahead -> interpolate (GL_INTER);
// neighbour behind -> wait (GL_UNVALID); neighbour finished -> GL_DONE.
// Behaviour of SWE_Block.hh:375-436.
template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::checkAllGhostlayers() {
    if (!localTimestepping) return;
    for (int border = BND_LEFT; border <= BND_TOP; border++) {
        if (boundaryType[border] != CONNECT && boundaryType[border] != CONNECT_WITHIN_RANK) continue;
        const float mine = getTotalLocalTimestep();
        if (std::fabs(borderTimestep[border] - mine) <= std::numeric_limits<double>::epsilon()) {
            copyGhostlayer(static_cast<Boundary>(border));
            receivedGhostlayer[border] = GL_SYNC;
        } else if (borderTimestep[border] > mine) {
            interpolateGhostlayer(static_cast<Boundary>(border), borderTimestep[border]);
            receivedGhostlayer[border] = GL_INTER;
        } else {
            receivedGhostlayer[border] = GL_UNVALID;
        }
        if (borderTimestep[border] >= duration) receivedGhostlayer[border] = GL_DONE;
    }
    const float tot = getTotalLocalTimestep();
    if (std::fabs(duration - tot) <= std::numeric_limits<double>::epsilon() || tot >= duration)
        notifiedLastTimestep = true;
    if (allGhostlayersInSync()) {
        for (int s = 0; s < 4; s++)
            if (receivedGhostlayer[s] != GL_DONE)
                receivedGhostlayer[s] = (receivedGhostlayer[s] == GL_INTER) ? GL_INTER : GL_NEXT;
    }
    if (allGhostlayersInSync() && stepSizeCounter < stepSize) stepSizeCounter++;
    iteration++;
    synchGhostLayerAfterWrite();
}

template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::interpolateGhostlayer(Boundary border, float remote) {
    if (border == BND_LEFT || border == BND_RIGHT) {
        const int x = (border == BND_LEFT) ? 0 : nx + 1;
        for (int i = 1; i < ny + 1; i++) {
            h[x][i] = interpolateValue(h[x][i], bufferH[x][i], remote, border);
            hu[x][i] = interpolateValue(hu[x][i], bufferHu[x][i], remote, border);
            hv[x][i] = interpolateValue(hv[x][i], bufferHv[x][i], remote, border);
        }
    } else {
        const int y = (border == BND_BOTTOM) ? 0 : ny + 1;
        for (int i = 1; i < nx + 1; i++) {
            h[i][y] = interpolateValue(h[i][y], bufferH[i][y], remote, border);
            hu[i][y] = interpolateValue(hu[i][y], bufferHu[i][y], remote, border);
            hv[i][y] = interpolateValue(hv[i][y], bufferHv[i][y], remote, border);
        }
    }
}

template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::copyGhostlayer(Boundary border) {
    if (border == BND_LEFT || border == BND_RIGHT) {
        const int x = (border == BND_LEFT) ? 0 : nx + 1;
        for (int i = 1; i < ny + 1; i++) {
            h[x][i] = bufferH[x][i];
            hu[x][i] = bufferHu[x][i];
            hv[x][i] = bufferHv[x][i];
        }
    } else {
        const int y = (border == BND_BOTTOM) ? 0 : ny + 1;
        for (int i = 1; i < nx + 1; i++) {
            h[i][y] = bufferH[i][y];
            hu[i][y] = bufferHu[i][y];
            hv[i][y] = bufferHv[i][y];
        }
    }
}

// Power-of-two sub-step of the macro step T, fixed at the first sub-step of
// a macro step (SWE_Block.hh:536-555): the step is clamped to [T/1024, T]
// and the divisor is the smallest power of two k with 1/k <= step
// (absolute, as in the reference), capped at maxDivisor.
template <typename T, typename Buffer>
float SWE_Block<T, Buffer>::getRoundTimestep(float timestep) {
    if (stepSizeCounter <= 0) {
        if (timestep > maxTimestepLocal) timestep = maxTimestepLocal;
        if (timestep < (float)maxTimestepLocal / maxDivisor) timestep = (float)maxTimestepLocal / maxDivisor;
        int divisor = (int)std::pow(2, -(std::ceil(std::log(timestep) / std::log(2))));
        if ((float)1.f / divisor > timestep) divisor *= 2;
        stepSize = divisor;
        if (stepSize > maxDivisor) stepSize = maxDivisor;
    }
    return (float)maxTimestepLocal / stepSize;
}

// initScenario (SWE_Block.hh:650-682): sample at cell centres, ghosts included.
template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::initScenario(SWE_Scenario& scenario, BoundaryType boundaries[]) {
    for (int i = 0; i < ny + 2; i++) {
        for (int j = 0; j < nx + 2; j++) {
            const float x = (float)originX + (j - 0.5) * dx;
            const float y = (float)originY + (i - 0.5) * dy;
            b[j][i] = scenario.getBathymetry(x, y);
            h[j][i] = scenario.getWaterHeight(x, y);
            hu[j][i] = scenario.getVeloc_u(x, y) * h[j][i];
            hv[j][i] = scenario.getVeloc_v(x, y) * h[j][i];
        }
    }
    for (int i = 0; i < 4; i++) boundaryType[i] = boundaries[i];
    applyBoundaryConditions();
    applyBoundaryBathymetry();
}

// Cell-based CFL estimate (SWE_Block.hh:693-717).
template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::computeMaxTimestep(const float dryTol, const float cflNumber) {
    synchBeforeRead();
    float maximumWaveSpeed = 0.f;
    for (int i = 1; i < nx + 1; i++)
        for (int j = 1; j < ny + 1; j++)
            if (h[i][j] > dryTol) {
                const float momentum = std::max(std::abs(hu[i][j]), std::abs(hv[i][j]));
                const float particleVelocity = momentum / h[i][j];
                const float waveSpeed = particleVelocity + std::sqrt(g * h[i][j]);
                maximumWaveSpeed = std::max(maximumWaveSpeed, waveSpeed);
            }
    maxTimestep = std::min(dx, dy) / maximumWaveSpeed;
    maxTimestep *= cflNumber;
}

// Bathymetry ghosts of OUTFLOW / WALL sides, then the corners (SWE_Block.hh:724-749).
template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::applyBoundaryBathymetry() {
    auto ow = [](BoundaryType t) { return t == OUTFLOW || t == WALL; };
    if (ow(boundaryType[BND_LEFT])) std::memcpy(b[0], b[1], sizeof(float) * (ny + 2));
    if (ow(boundaryType[BND_RIGHT])) std::memcpy(b[nx + 1], b[nx], sizeof(float) * (ny + 2));
    if (ow(boundaryType[BND_BOTTOM]))
        for (int i = 0; i <= nx + 1; i++) b[i][0] = b[i][1];
    if (ow(boundaryType[BND_TOP]))
        for (int i = 0; i <= nx + 1; i++) b[i][ny + 1] = b[i][ny];
    b[0][0] = b[1][1];
    b[0][ny + 1] = b[1][ny];
    b[nx + 1][0] = b[nx][1];
    b[nx + 1][ny + 1] = b[nx][ny];
}

// Ghost cells of WALL (normal momentum negated) and OUTFLOW sides, then the
// corners from the diagonal interior cell (SWE_Block.hh:757-911).
template <typename T, typename Buffer>
void SWE_Block<T, Buffer>::applyBoundaryConditions() {
    for (int side = BND_LEFT; side <= BND_RIGHT; side++) {
        const BoundaryType t = boundaryType[side];
        if (t == INFLOW) assert(false);
        if (t != WALL && t != OUTFLOW) continue;
        const int gx = side == BND_LEFT ? 0 : nx + 1, sx = side == BND_LEFT ? 1 : nx;
        for (int j = 1; j <= ny; j++) {
            h[gx][j] = h[sx][j];
            hu[gx][j] = (t == WALL) ? -hu[sx][j] : hu[sx][j];
            hv[gx][j] = hv[sx][j];
        }
    }
    for (int side = BND_BOTTOM; side <= BND_TOP; side++) {
        const BoundaryType t = boundaryType[side];
        if (t == INFLOW) assert(false);
        if (t != WALL && t != OUTFLOW) continue;
        const int gy = side == BND_BOTTOM ? 0 : ny + 1, sy = side == BND_BOTTOM ? 1 : ny;
        for (int i = 1; i <= nx; i++) {
            h[i][gy] = h[i][sy];
            hu[i][gy] = hu[i][sy];
            hv[i][gy] = (t == WALL) ? -hv[i][sy] : hv[i][sy];
        }
    }
    T* f[3] = {&h, &hu, &hv};
    for (T* a : f) {
        (*a)[0][0] = (*a)[1][1];
        (*a)[0][ny + 1] = (*a)[1][ny];
        (*a)[nx + 1][0] = (*a)[nx][1];
        (*a)[nx + 1][ny + 1] = (*a)[nx][ny];
    }
}

#endif

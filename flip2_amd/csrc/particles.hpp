// Reference-compatible particle system: the public surface of the
// reference's class Particles (engine/particles.hu:11-45) on HIP.
#pragma once

#include "common.hpp"
#include "device.hpp"

#include <string>

namespace flip2 {

struct RcStats {
    uint32_t numUsedGridNodes;
    uint32_t numUsedVoxels;
    uint32_t lastSorIterations;
    uint32_t lastSolveAttempts;
    double lastResidual;
    double dt;
    double frameDt;
    double elapsedTime;
    double elapsedTimeThisFrame;
    uint32_t outOfTileEvents;
    uint32_t relocatedParticles;
    uint32_t substeps;
};

class Particles {
public:
    explicit Particles(uint32_t size, int device = 0);
    ~Particles();
    Particles(const Particles&) = delete;
    Particles& operator=(const Particles&) = delete;

    // particles.hu:11-45
    void setDomain(double nx, double ny, double nz, uint32_t x, uint32_t y, uint32_t z, double cellSize);
    void alignParticlesToGrid();
    void sortParticles();
    void generateVoxels();
    void particleVelToVoxels();
    void pressureSolve();
    void updateVoxelVelocities();
    void advectParticles();
    void solveFrame(double fps);
    void initialize();
    void voxelVelsToParticles();
    void moveSolidParticles();
    void writePositionsToFile(const std::string& fileName);
    double getCourantDt();

    // Stages of pressureSolve, exposed for stage-level parity checks.
    void applyGravity(double dt);
    void removeGravity(double dt);
    void calcDivU();
    double gsSolve(double dt, double threshold, uint32_t maxIterations, uint32_t* iters);
    void gsIterations(double dt, uint32_t count);
    void moveSolidsOnly();

    // Host <-> device state access (the reference's friend harness reads the
    // private arrays directly, engine/testing.h:27-75).
    void setPositions(const double* x, const double* y, const double* z);
    void setVelocities(const double* x, const double* y, const double* z);
    void getPositions(double* x, double* y, double* z);
    void getVelocities(double* x, double* y, double* z);
    // Named array access; returns the element count, copies when host != null.
    // setArray resizes node/voxel arrays when given their full length.
    size_t getArray(const std::string& name, void* host, size_t capacity);
    void setArray(const std::string& name, const void* host, size_t count);
    void setTime(double dt, double frameDt, double elapsedTime, double elapsedTimeThisFrame);
    // SOR stop threshold and iteration cap (0.01 and 16384 in the reference,
    // particles.cu:509-510).
    void setSolverLimits(double threshold, uint32_t maxIterations);

    RcStats stats();
    uint32_t numParticles() const { return size; }
    const Grid& getGrid() const { return grid; }
    hipStream_t getStream() const { return stream; }
    DeviceManager& deviceManager() { return dm_; }
    void synchronize();

private:
    void ensureWorkspace(size_t bytes);
    uint32_t readU32(const uint32_t* dev);
    void resizeNodes(uint32_t B);
    void resizeVoxels(uint32_t V);
    double solveScale(double dt) const;
    void ensureAbsmaxTmp(uint32_t n);

    // Every device array below is allocated from this pool (declared first:
    // it outlives them).
    DeviceManager dm_;

    // The reference's members, with its names and types (particles.hu:47-92),
    // so a harness written against them (the friend ParticleSystemTester of
    // engine/testing.h) reads and writes the same state.
    DeviceVec<double> px{dm_}, py{dm_}, pz{dm_};
    DeviceVec<double> vx{dm_}, vy{dm_}, vz{dm_};
    DeviceVec<uint32_t> gridCell{dm_};
    DeviceVec<char> solids{dm_};
    DeviceVec<double> p{dm_};
    DeviceVec<double> residuals{dm_};
    DeviceVec<uint32_t> yDimNumUsedGridNodes{dm_};
    DeviceVec<uint32_t> reorderedGridIndices{dm_};
    DeviceVec<uint32_t> uniqueGridNodeIndices{dm_};
    DeviceVec<uint32_t> gridNodeIndicesToFirstParticleIndex{dm_};
    DeviceVec<uint32_t> nodeIndexUsedVoxels{dm_};
    DeviceVec<uint32_t> voxelIDsUsed{dm_};
    DeviceVec<double> divU{dm_};
    DeviceVec<double> voxelsUx{dm_}, voxelsUy{dm_}, voxelsUz{dm_};

    uint32_t numUsedGridNodes = 0;
    uint32_t numUsedVoxels = 0;
    uint32_t refinementLevel = kRcRefinement;
    uint32_t numVoxelsPerNode = kRcVoxelsPerNode;
    uint32_t numVoxels1D = kRcVoxels1D;
    uint32_t size;
    double radius = kRcRadius;
    Grid grid{};
    double frameDt;
    double dt = 0.0;
    double elapsedTime = 0.0;
    double elapsedTimeThisFrame = 0.0;
    hipStream_t stream = nullptr;

    // scratch
    uint32_t yDimRows_ = 0;
    DeviceVec<uint32_t> sortedKeys_{dm_}, keysAlt_{dm_}, permAlt_{dm_};
    DeviceVec<double> tmpX_{dm_}, tmpY_{dm_}, tmpZ_{dm_};
    DeviceVec<uint64_t> usedMask_{dm_}, solidMask_{dm_};
    DeviceVec<uint32_t> nodeCounts_{dm_};
    DeviceVec<uint8_t> workspace_{dm_};
    DeviceVec<double> absmaxTmp_{dm_};
    DeviceVec<double> scalarsD_{dm_};   // [0..2] sampled maxima
    DeviceVec<uint32_t> scalarsU_{dm_}; // [0] scan total, [1] out-of-tile events, [2] relocations

    uint32_t lastSorIterations_ = 0;
    uint32_t lastSolveAttempts_ = 0;
    double lastResidual_ = 0.0;
    uint32_t substeps_ = 0;
    uint64_t moveSolidsCounter_ = 0;
    double sorThreshold_;
    uint32_t sorMaxIterations_;

    friend class ParticleSystemTester;
};

}  // namespace flip2

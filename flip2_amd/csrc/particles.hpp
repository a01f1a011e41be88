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
    uint32_t size() const { return size_; }
    const Grid& grid() const { return grid_; }
    hipStream_t stream() const { return stream_; }
    void synchronize();

private:
    void ensureWorkspace(size_t bytes);
    uint32_t readU32(const uint32_t* dev);
    void resizeNodes(uint32_t B);
    void resizeVoxels(uint32_t V);
    double solveScale(double dt) const;
    void ensureAbsmaxTmp(uint32_t n);

    DeviceManager dm_;
    hipStream_t stream_ = nullptr;
    uint32_t size_;
    Grid grid_{};
    uint32_t yDimRows_ = 0;

    DeviceVec<double> px_, py_, pz_, vx_, vy_, vz_;
    DeviceVec<uint32_t> gridCell_, reorderedGridIndices_, uniqueGridNodeIndices_;
    DeviceVec<uint32_t> gridNodeIndicesToFirstParticleIndex_, yDimNumUsedGridNodes_, nodeIndexUsedVoxels_;
    DeviceVec<uint32_t> voxelIDsUsed_;
    DeviceVec<int8_t> solids_;
    DeviceVec<double> voxelsUx_, voxelsUy_, voxelsUz_, divU_, p_, residuals_;

    // scratch
    DeviceVec<uint32_t> sortedKeys_, keysAlt_, permAlt_;
    DeviceVec<double> tmpX_, tmpY_, tmpZ_;
    DeviceVec<uint64_t> usedMask_, solidMask_;
    DeviceVec<uint32_t> nodeCounts_;
    DeviceVec<uint8_t> workspace_;
    DeviceVec<double> absmaxTmp_;
    DeviceVec<double> scalarsD_;   // [0..2] sampled maxima
    DeviceVec<uint32_t> scalarsU_; // [0] scan total, [1] out-of-tile events, [2] relocations

    uint32_t numUsedGridNodes_ = 0;
    uint32_t numUsedVoxels_ = 0;
    double frameDt_;
    double dt_ = 0.0;
    double elapsedTime_ = 0.0;
    double elapsedTimeThisFrame_ = 0.0;
    uint32_t lastSorIterations_ = 0;
    uint32_t lastSolveAttempts_ = 0;
    double lastResidual_ = 0.0;
    uint32_t substeps_ = 0;
    uint64_t moveSolidsCounter_ = 0;
    double sorThreshold_;
    uint32_t sorMaxIterations_;
};

}  // namespace flip2

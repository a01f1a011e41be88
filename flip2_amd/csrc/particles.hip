// Reference-compatible step driver (engine/particles.cu:15-831) on HIP:
// same stage order and control logic, one stream, no per-step device
// allocation, host syncs only where the control flow needs a device value
// (node count, voxel count, Courant maximum, SOR residual checks).
#include "particles.hpp"

#include "rc_kernels.hpp"
#include "scan_impl.hpp"
#include "sort.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>

namespace flip2 {

namespace {
constexpr double kDensity = 0.014;        // particles.cu:508
constexpr double kSorThreshold = 0.01;    // particles.cu:509
constexpr uint32_t kSorMaxIterations = 16384;  // particles.cu:510
constexpr double kSorOmega = 1.9;         // voxelSolveFunctions.cu:396
constexpr uint32_t kSorCheckEvery = 1024; // voxelSolveFunctions.cu:395

uint32_t row_table_rows(const Grid& g) {
    const uint64_t maxKey = (uint64_t)g.sizeX + (uint64_t)g.sizeY * g.sizeX + (uint64_t)g.sizeZ * g.sizeX * g.sizeY;
    uint64_t rows = maxKey / (g.sizeX ? g.sizeX : 1) + 1;
    rows = std::max<uint64_t>(rows, (uint64_t)g.sizeY * g.sizeZ);
    rows = std::max<uint64_t>(rows, (uint64_t)g.sizeX * g.sizeY);
    FLIP2_REQUIRE(rows < 0xffffffffull, kErrRange, "domain too large");
    return (uint32_t)rows;
}
}  // namespace

// Particles::Particles (particles.cu:15-40).
Particles::Particles(uint32_t size, int device) : dm_(device), size_(size) {
    FLIP2_REQUIRE(size > 0 && size < (1u << 30), kErrRange, "particle count must be in [1, 2^30)");
    FLIP2_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    px_.resize(size); py_.resize(size); pz_.resize(size);
    vx_.resize(size); vy_.resize(size); vz_.resize(size);
    gridCell_.resize(size);
    reorderedGridIndices_.resize(size);
    uniqueGridNodeIndices_.resize(size);
    sortedKeys_.resizeAsync(size, stream_);
    keysAlt_.resizeAsync(size, stream_);
    permAlt_.resizeAsync(size, stream_);
    tmpX_.resizeAsync(size, stream_);
    tmpY_.resizeAsync(size, stream_);
    tmpZ_.resizeAsync(size, stream_);
    scalarsD_.resizeAsync(4, stream_);
    scalarsU_.resizeAsync(4, stream_);
    absmaxTmp_.resizeAsync((uint32_t)rc::sampled_absmax_tmp_doubles(size, 3), stream_);
    FLIP2_HIP(hipMemsetAsync(px_.devPtr(), 0, px_.bytes(), stream_));
    FLIP2_HIP(hipMemsetAsync(py_.devPtr(), 0, py_.bytes(), stream_));
    FLIP2_HIP(hipMemsetAsync(pz_.devPtr(), 0, pz_.bytes(), stream_));
    vx_.zeroDeviceAsync(stream_); vy_.zeroDeviceAsync(stream_); vz_.zeroDeviceAsync(stream_);
    scalarsU_.zeroDeviceAsync(stream_);
    frameDt_ = (double)(1.0f / 24.0f);
    sorThreshold_ = kSorThreshold;
    sorMaxIterations_ = kSorMaxIterations;
    synchronize();
}

Particles::~Particles() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void Particles::synchronize() { FLIP2_HIP(hipStreamSynchronize(stream_)); }

void Particles::ensureWorkspace(size_t bytes) {
    if (workspace_.size() < bytes) workspace_.resizeAsync((uint32_t)bytes, stream_);
}

uint32_t Particles::readU32(const uint32_t* dev) {
    uint32_t v = 0;
    FLIP2_HIP(hipMemcpyAsync(&v, dev, sizeof(v), hipMemcpyDeviceToHost, stream_));
    synchronize();
    return v;
}

// Particles::setDomain (particles.cu:42-46): corner and cell size narrowed to
// float (grid.hpp:22-33).  The row table is sized for every row index the
// node keys can produce (the reference sizes it y*z).
void Particles::setDomain(double nx, double ny, double nz, uint32_t x, uint32_t y, uint32_t z, double cellSize) {
    FLIP2_REQUIRE(x > 0 && y > 0 && z > 0, kErrArgument, "domain sizes must be positive");
    FLIP2_REQUIRE(cellSize > 0.0, kErrArgument, "cell size must be positive");
    FLIP2_REQUIRE((uint64_t)x * y * z + (uint64_t)x * y + x < 0xffffffffull, kErrRange,
                  "domain has more than 2^32 cell keys");
    grid_.negX = (float)nx; grid_.negY = (float)ny; grid_.negZ = (float)nz;
    grid_.sizeX = x; grid_.sizeY = y; grid_.sizeZ = z;
    grid_.cellSize = (float)cellSize;
    yDimRows_ = row_table_rows(grid_);
    yDimNumUsedGridNodes_.resizeAsync(yDimRows_, stream_);
}

void Particles::alignParticlesToGrid() {
    FLIP2_REQUIRE(grid_.sizeX > 0, kErrState, "setDomain has not been called");
    rc::launch_root_cell(px_.devPtr(), py_.devPtr(), pz_.devPtr(), size_, grid_, gridCell_.devPtr(), stream_);
}

// Particles::sortParticles (particles.cu:52-78): keys sorted with the
// permutation; positions gathered, velocities keep their order.  Only the
// bits a key of this domain can use are sorted (identical result to a
// 32-bit sort).
void Particles::sortParticles() {
    const uint64_t maxKey = (uint64_t)grid_.sizeX + (uint64_t)grid_.sizeY * grid_.sizeX +
                            (uint64_t)grid_.sizeZ * grid_.sizeX * grid_.sizeY;
    const uint32_t endBit = std::min<uint32_t>(32, std::max<uint32_t>(1, bits_for(maxKey)));
    const size_t wsBytes = sort_workspace_bytes(size_, 0, endBit);
    ensureWorkspace(wsBytes);
    sort_pairs_u32(gridCell_.devPtr(), nullptr, sortedKeys_.devPtr(), reorderedGridIndices_.devPtr(),
                   keysAlt_.devPtr(), permAlt_.devPtr(), size_, 0, endBit, workspace_.devPtr(), wsBytes, stream_);
    gridCell_.swapDevice(sortedKeys_);
    rc::launch_gather3(reorderedGridIndices_.devPtr(), size_, px_.devPtr(), py_.devPtr(), pz_.devPtr(),
                       tmpX_.devPtr(), tmpY_.devPtr(), tmpZ_.devPtr(), stream_);
    px_.swapDevice(tmpX_);
    py_.swapDevice(tmpY_);
    pz_.swapDevice(tmpZ_);
}

void Particles::resizeNodes(uint32_t B) {
    numUsedGridNodes_ = B;
    gridNodeIndicesToFirstParticleIndex_.resizeAsync(B, stream_);
    nodeIndexUsedVoxels_.resizeAsync(B, stream_);
    nodeCounts_.resizeAsync(B, stream_);
    usedMask_.resizeAsync(B * 8u, stream_);
    solidMask_.resizeAsync(B * 8u, stream_);
}

void Particles::resizeVoxels(uint32_t V) {
    numUsedVoxels_ = V;
    voxelIDsUsed_.resizeAsync(V, stream_);
    solids_.resizeAsync(V, stream_);
    voxelsUx_.resizeAsync(V, stream_);
    voxelsUy_.resizeAsync(V, stream_);
    voxelsUz_.resizeAsync(V, stream_);
    divU_.resizeAsync(V, stream_);
    p_.resizeAsync(V, stream_);
    residuals_.resizeAsync(V, stream_);
}

// Particles::generateVoxels (particles.cu:323-368).
void Particles::generateVoxels() {
    // head flags fused into the inclusive scan -> 1-based node id per particle
    const size_t scanWs = scan_workspace_bytes(size_);
    ensureWorkspace(scanWs);
    uint32_t* total = scalarsU_.devPtr();
    const bool vec = ((reinterpret_cast<uintptr_t>(gridCell_.devPtr()) |
                       reinterpret_cast<uintptr_t>(uniqueGridNodeIndices_.devPtr())) & 15u) == 0;
    scan_launch(ScanHeadFlagInput{gridCell_.devPtr()}, uniqueGridNodeIndices_.devPtr(), size_, false, vec,
                workspace_.devPtr(), scanWs, total, stream_);
    const uint32_t B = readU32(total);
    resizeNodes(B);
    rc::launch_map_nodes(uniqueGridNodeIndices_.devPtr(), size_, gridNodeIndicesToFirstParticleIndex_.devPtr(),
                         stream_);
    rc::launch_fill_u32(yDimNumUsedGridNodes_.devPtr(), yDimRows_, B, stream_);
    rc::launch_row_table(B, gridNodeIndicesToFirstParticleIndex_.devPtr(), gridCell_.devPtr(),
                         yDimNumUsedGridNodes_.devPtr(), grid_.sizeX, stream_);
    rc::launch_mark_voxels(B, size_, grid_, gridCell_.devPtr(), gridNodeIndicesToFirstParticleIndex_.devPtr(),
                           px_.devPtr(), py_.devPtr(), pz_.devPtr(), usedMask_.devPtr(), solidMask_.devPtr(),
                           nodeCounts_.devPtr(), scalarsU_.devPtr() + 1, stream_);
    const size_t scanWs2 = scan_workspace_bytes(B);
    ensureWorkspace(scanWs2);
    scan_u32(nodeCounts_.devPtr(), nodeIndexUsedVoxels_.devPtr(), B, false, workspace_.devPtr(), scanWs2, total,
             stream_);
    const uint32_t V = readU32(total);
    resizeVoxels(V);
    p_.zeroDeviceAsync(stream_);
    rc::launch_compact_voxels(B, usedMask_.devPtr(), solidMask_.devPtr(), nodeIndexUsedVoxels_.devPtr(),
                              voxelIDsUsed_.devPtr(), solids_.devPtr(), stream_);
}

// Particles::particleVelToVoxels (particles.cu:500-505).
void Particles::particleVelToVoxels() {
    rc::launch_p2g(numUsedGridNodes_, size_, grid_, gridCell_.devPtr(), gridNodeIndicesToFirstParticleIndex_.devPtr(),
                   px_.devPtr(), py_.devPtr(), pz_.devPtr(), vx_.devPtr(), vy_.devPtr(), vz_.devPtr(),
                   nodeIndexUsedVoxels_.devPtr(), voxelIDsUsed_.devPtr(), solids_.devPtr(), voxelsUx_.devPtr(),
                   voxelsUy_.devPtr(), voxelsUz_.devPtr(), stream_);
}

// Particles::getCourantDt (particles.cu:493-498) with the sampled maxima of
// CudaVec::getMax (cudaVec.cu:205-233).
double Particles::getCourantDt() {
    rc::Ptr3 arrs{{vx_.devPtr(), vy_.devPtr(), vz_.devPtr()}};
    rc::launch_sampled_absmax(arrs, 3, size_, absmaxTmp_.devPtr(), scalarsD_.devPtr(), stream_);
    double m[3];
    FLIP2_HIP(hipMemcpyAsync(m, scalarsD_.devPtr(), sizeof(m), hipMemcpyDeviceToHost, stream_));
    synchronize();
    const double maxVel = std::fmax(std::fabs(m[0]), std::fmax(std::fabs(m[1]), std::fabs(m[2])));
    const double voxelSize = (double)grid_.cellSize / ((double)kRcVoxels1D - 2.0 * std::floor(kRcRadius));
    return 0.7 * (voxelSize / maxVel);
}

void Particles::applyGravity(double dt) {
    rc::launch_gravity(numUsedVoxels_, solids_.devPtr(), voxelsUy_.devPtr(), (double)(-9.8f) * dt, false, stream_);
}

void Particles::removeGravity(double dt) {
    rc::launch_gravity(numUsedVoxels_, solids_.devPtr(), voxelsUy_.devPtr(), (double)(-9.8f) * dt, true, stream_);
}

void Particles::calcDivU() {
    rc::launch_divu(numUsedGridNodes_, grid_, gridCell_.devPtr(), gridNodeIndicesToFirstParticleIndex_.devPtr(),
                    yDimNumUsedGridNodes_.devPtr(), nodeIndexUsedVoxels_.devPtr(), voxelIDsUsed_.devPtr(),
                    voxelsUx_.devPtr(), voxelsUy_.devPtr(), voxelsUz_.devPtr(), divU_.devPtr(), stream_);
}

double Particles::solveScale(double dt) const {
    const double vs = (double)(grid_.cellSize / (float)(2u << kRcRefinement));
    return dt / (kDensity * vs * vs);
}

void Particles::gsIterations(double dt, uint32_t count) {
    rc::launch_sor(numUsedGridNodes_, nodeIndexUsedVoxels_.devPtr(), voxelIDsUsed_.devPtr(), solids_.devPtr(),
                   divU_.devPtr(), p_.devPtr(), residuals_.devPtr(), solveScale(dt), kSorOmega, count, stream_);
}

// cudaGSiteration (voxelSolveFunctions.cu:391-412): the iterations between
// two residual checks run inside one LDS-resident kernel launch.
double Particles::gsSolve(double dt, double threshold, uint32_t maxIterations, uint32_t* itersOut) {
    double maxResidual = 10.0, prevMaxResidual = 100.0;
    uint32_t iterations = 0;
    ensureAbsmaxTmp(numUsedVoxels_);
    while (maxResidual > threshold && iterations < maxIterations) {
        uint32_t run;
        if (iterations % kSorCheckEvery == 0) run = 1;
        else run = (iterations / kSorCheckEvery + 1) * kSorCheckEvery - iterations + 1;
        if (iterations + run > maxIterations) run = maxIterations - iterations;
        gsIterations(dt, run);
        const uint32_t lastIt = iterations + run - 1;
        iterations += run;
        if (lastIt % kSorCheckEvery == 0) {
            prevMaxResidual = maxResidual;
            rc::Ptr3 arrs{{residuals_.devPtr(), nullptr, nullptr}};
            rc::launch_sampled_absmax(arrs, 1, numUsedVoxels_, absmaxTmp_.devPtr(), scalarsD_.devPtr() + 3, stream_);
            FLIP2_HIP(hipMemcpyAsync(&maxResidual, scalarsD_.devPtr() + 3, sizeof(double), hipMemcpyDeviceToHost,
                                     stream_));
            synchronize();
        }
        if (iterations == maxIterations ||
            (std::fabs(prevMaxResidual - maxResidual) < threshold / 1000 && maxResidual > threshold))
            break;
    }
    if (itersOut) *itersOut = iterations;
    return maxResidual;
}

// Particles::pressureSolve (particles.cu:507-555).
void Particles::pressureSolve() {
    double previousTerminatingResidual = 100.0, terminatingResidual = 0.0;
    uint32_t its = 0;
    lastSorIterations_ = 0;
    lastSolveAttempts_ = 0;
    dt_ = getCourantDt();
    if (frameDt_ - elapsedTimeThisFrame_ < dt_) dt_ = frameDt_ - elapsedTimeThisFrame_;
    applyGravity(dt_);
    calcDivU();
    for (;;) {
        terminatingResidual = gsSolve(dt_, sorThreshold_, sorMaxIterations_, &its);
        lastSorIterations_ += its;
        lastSolveAttempts_ += 1;
        if (!(previousTerminatingResidual - terminatingResidual > 0)) break;
        if (terminatingResidual < sorThreshold_) break;
        removeGravity(dt_);
        dt_ /= 2.0;
        p_.zeroDeviceAsync(stream_);
        applyGravity(dt_);
        previousTerminatingResidual = terminatingResidual;
        calcDivU();
    }
    lastResidual_ = terminatingResidual;
    elapsedTime_ += dt_;
    elapsedTimeThisFrame_ += dt_;
}

// Particles::updateVoxelVelocities (particles.cu:557-560).
void Particles::updateVoxelVelocities() {
    rc::launch_velocity_update(numUsedGridNodes_, nodeIndexUsedVoxels_.devPtr(), voxelIDsUsed_.devPtr(),
                               solids_.devPtr(), p_.devPtr(), voxelsUx_.devPtr(), voxelsUy_.devPtr(),
                               voxelsUz_.devPtr(), solveScale(dt_), stream_);
}

// Particles::voxelVelsToParticles (particles.cu:742-749).
void Particles::voxelVelsToParticles() {
    rc::launch_g2p(numUsedGridNodes_, size_, grid_, gridCell_.devPtr(), gridNodeIndicesToFirstParticleIndex_.devPtr(),
                   px_.devPtr(), py_.devPtr(), pz_.devPtr(), vx_.devPtr(), vy_.devPtr(), vz_.devPtr(),
                   nodeIndexUsedVoxels_.devPtr(), voxelIDsUsed_.devPtr(), voxelsUx_.devPtr(), voxelsUy_.devPtr(),
                   voxelsUz_.devPtr(), stream_);
}

// Particles::advectParticles (particles.cu:751-760).
void Particles::advectParticles() {
    rc::launch_advect(size_, dt_, px_.devPtr(), py_.devPtr(), pz_.devPtr(), vx_.devPtr(), vy_.devPtr(), vz_.devPtr(),
                      stream_);
}

void Particles::moveSolidsOnly() {
    FLIP2_HIP(hipMemsetAsync(scalarsU_.devPtr() + 2, 0, sizeof(uint32_t), stream_));
    rc::launch_move_solids(numUsedGridNodes_, size_, grid_, gridCell_.devPtr(),
                           gridNodeIndicesToFirstParticleIndex_.devPtr(), px_.devPtr(), py_.devPtr(), pz_.devPtr(),
                           nodeIndexUsedVoxels_.devPtr(), voxelIDsUsed_.devPtr(), solids_.devPtr(),
                           moveSolidsCounter_, scalarsU_.devPtr() + 2, stream_);
    ++moveSolidsCounter_;
}

// Particles::moveSolidParticles (particles.cu:762-769).
void Particles::moveSolidParticles() {
    moveSolidsOnly();
    voxelVelsToParticles();
}

// Particles::initialize (particles.cu:805-812).
void Particles::initialize() {
    alignParticlesToGrid();
    sortParticles();
    generateVoxels();
}

// Particles::solveFrame (particles.cu:784-803).
void Particles::solveFrame(double fps) {
    frameDt_ = 1.0 / fps;
    dt_ = frameDt_ / 10.0;
    elapsedTimeThisFrame_ = 0.0;
    substeps_ = 0;
    while (elapsedTimeThisFrame_ < frameDt_) {
        particleVelToVoxels();
        pressureSolve();
        updateVoxelVelocities();
        voxelVelsToParticles();
        advectParticles();
        moveSolidParticles();
        initialize();
        ++substeps_;
    }
    synchronize();
}

// Particles::writePositionsToFile (particles.cu:817-828): "x, y, z" lines.
void Particles::writePositionsToFile(const std::string& fileName) {
    std::vector<double> x(size_), y(size_), z(size_);
    getPositions(x.data(), y.data(), z.data());
    std::ofstream file(fileName);
    FLIP2_REQUIRE(file.good(), kErrArgument, "cannot open " + fileName);
    for (uint32_t i = 0; i < size_; ++i) file << x[i] << ", " << y[i] << ", " << z[i] << "\n";
}

void Particles::setPositions(const double* x, const double* y, const double* z) {
    px_.uploadFrom(x, size_, stream_);
    py_.uploadFrom(y, size_, stream_);
    pz_.uploadFrom(z, size_, stream_);
    synchronize();
}

void Particles::setVelocities(const double* x, const double* y, const double* z) {
    vx_.uploadFrom(x, size_, stream_);
    vy_.uploadFrom(y, size_, stream_);
    vz_.uploadFrom(z, size_, stream_);
    synchronize();
}

void Particles::getPositions(double* x, double* y, double* z) {
    px_.downloadTo(x, size_, stream_);
    py_.downloadTo(y, size_, stream_);
    pz_.downloadTo(z, size_, stream_);
    synchronize();
}

void Particles::getVelocities(double* x, double* y, double* z) {
    vx_.downloadTo(x, size_, stream_);
    vy_.downloadTo(y, size_, stream_);
    vz_.downloadTo(z, size_, stream_);
    synchronize();
}

namespace {
template <typename T>
size_t copy_out(DeviceVec<T>& v, uint32_t count, void* host, size_t capacity, hipStream_t s) {
    if (host) {
        FLIP2_REQUIRE(capacity >= count, kErrRange, "getArray: buffer too small");
        v.downloadTo(reinterpret_cast<T*>(host), count, s);
    }
    return count;
}
}  // namespace

size_t Particles::getArray(const std::string& name, void* host, size_t capacity) {
    size_t r;
    const uint32_t B = numUsedGridNodes_, V = numUsedVoxels_, n = size_;
    if (name == "px") r = copy_out(px_, n, host, capacity, stream_);
    else if (name == "py") r = copy_out(py_, n, host, capacity, stream_);
    else if (name == "pz") r = copy_out(pz_, n, host, capacity, stream_);
    else if (name == "vx") r = copy_out(vx_, n, host, capacity, stream_);
    else if (name == "vy") r = copy_out(vy_, n, host, capacity, stream_);
    else if (name == "vz") r = copy_out(vz_, n, host, capacity, stream_);
    else if (name == "gridCell") r = copy_out(gridCell_, n, host, capacity, stream_);
    else if (name == "reorderedGridIndices") r = copy_out(reorderedGridIndices_, n, host, capacity, stream_);
    else if (name == "uniqueGridNodeIndices") r = copy_out(uniqueGridNodeIndices_, n, host, capacity, stream_);
    else if (name == "gridNodeIndicesToFirstParticleIndex")
        r = copy_out(gridNodeIndicesToFirstParticleIndex_, B, host, capacity, stream_);
    else if (name == "yDimNumUsedGridNodes") r = copy_out(yDimNumUsedGridNodes_, yDimRows_, host, capacity, stream_);
    else if (name == "nodeIndexUsedVoxels") r = copy_out(nodeIndexUsedVoxels_, B, host, capacity, stream_);
    else if (name == "voxelIDsUsed") r = copy_out(voxelIDsUsed_, V, host, capacity, stream_);
    else if (name == "solids") r = copy_out(solids_, V, host, capacity, stream_);
    else if (name == "voxelsUx") r = copy_out(voxelsUx_, V, host, capacity, stream_);
    else if (name == "voxelsUy") r = copy_out(voxelsUy_, V, host, capacity, stream_);
    else if (name == "voxelsUz") r = copy_out(voxelsUz_, V, host, capacity, stream_);
    else if (name == "divU") r = copy_out(divU_, V, host, capacity, stream_);
    else if (name == "p") r = copy_out(p_, V, host, capacity, stream_);
    else if (name == "residuals") r = copy_out(residuals_, V, host, capacity, stream_);
    else throw Error(kErrArgument, "unknown array " + name);
    synchronize();
    return r;
}

void Particles::setArray(const std::string& name, const void* host, size_t count) {
    auto put = [&](auto& vec, size_t expect) {
        using T = typename std::remove_reference<decltype(vec[0])>::type;
        FLIP2_REQUIRE(count == expect, kErrRange, "setArray: count mismatch for " + name);
        vec.uploadFrom(reinterpret_cast<const T*>(host), (uint32_t)count, stream_);
    };
    const uint32_t n = size_;
    if (name == "px") put(px_, n);
    else if (name == "py") put(py_, n);
    else if (name == "pz") put(pz_, n);
    else if (name == "vx") put(vx_, n);
    else if (name == "vy") put(vy_, n);
    else if (name == "vz") put(vz_, n);
    else if (name == "gridCell") put(gridCell_, n);
    else if (name == "voxelsUx") put(voxelsUx_, numUsedVoxels_);
    else if (name == "voxelsUy") put(voxelsUy_, numUsedVoxels_);
    else if (name == "voxelsUz") put(voxelsUz_, numUsedVoxels_);
    else if (name == "divU") put(divU_, numUsedVoxels_);
    else if (name == "p") put(p_, numUsedVoxels_);
    else throw Error(kErrArgument, "array not settable: " + name);
    synchronize();
}

void Particles::setSolverLimits(double threshold, uint32_t maxIterations) {
    FLIP2_REQUIRE(maxIterations > 0, kErrArgument, "maxIterations must be positive");
    sorThreshold_ = threshold;
    sorMaxIterations_ = maxIterations;
}

void Particles::setTime(double dt, double frameDt, double elapsedTime, double elapsedTimeThisFrame) {
    dt_ = dt;
    frameDt_ = frameDt;
    elapsedTime_ = elapsedTime;
    elapsedTimeThisFrame_ = elapsedTimeThisFrame;
}

void Particles::ensureAbsmaxTmp(uint32_t n) {
    const size_t need = rc::sampled_absmax_tmp_doubles(n, 3);
    if (absmaxTmp_.size() < need) absmaxTmp_.resizeAsync((uint32_t)need, stream_);
}

RcStats Particles::stats() {
    RcStats s{};
    s.numUsedGridNodes = numUsedGridNodes_;
    s.numUsedVoxels = numUsedVoxels_;
    s.lastSorIterations = lastSorIterations_;
    s.lastSolveAttempts = lastSolveAttempts_;
    s.lastResidual = lastResidual_;
    s.dt = dt_;
    s.frameDt = frameDt_;
    s.elapsedTime = elapsedTime_;
    s.elapsedTimeThisFrame = elapsedTimeThisFrame_;
    uint32_t u[4] = {0, 0, 0, 0};
    FLIP2_HIP(hipMemcpyAsync(u, scalarsU_.devPtr(), sizeof(u), hipMemcpyDeviceToHost, stream_));
    synchronize();
    s.outOfTileEvents = u[1];
    s.relocatedParticles = u[2];
    s.substeps = substeps_;
    return s;
}

}  // namespace flip2

// Reference-compatible step driver (engine/particles.cu:15-831) on HIP:
// same stage order and control logic, one stream, no per-step device
// allocation, host syncs only where the control flow needs a device value
// (node count, voxel count, Courant maximum, SOR residual checks).
#include "particles.hpp"

#include "algorithms.hpp"

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
Particles::Particles(uint32_t size, int device) : dm_(device), size(size) {
    FLIP2_REQUIRE(size > 0 && size < (1u << 30), kErrRange, "particle count must be in [1, 2^30)");
    FLIP2_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    dm_.setStream(stream);
    px.resize(size); py.resize(size); pz.resize(size);
    vx.resize(size); vy.resize(size); vz.resize(size);
    gridCell.resize(size);
    reorderedGridIndices.resize(size);
    uniqueGridNodeIndices.resize(size);
    sortedKeys_.resizeAsync(size, stream);
    keysAlt_.resizeAsync(size, stream);
    permAlt_.resizeAsync(size, stream);
    tmpX_.resizeAsync(size, stream);
    tmpY_.resizeAsync(size, stream);
    tmpZ_.resizeAsync(size, stream);
    scalarsD_.resizeAsync(4, stream);
    scalarsU_.resizeAsync(4, stream);
    absmaxTmp_.resizeAsync((uint32_t)rc::sampled_absmax_tmp_doubles(size, 3), stream);
    FLIP2_HIP(hipMemsetAsync(px.devPtr(), 0, px.bytes(), stream));
    FLIP2_HIP(hipMemsetAsync(py.devPtr(), 0, py.bytes(), stream));
    FLIP2_HIP(hipMemsetAsync(pz.devPtr(), 0, pz.bytes(), stream));
    vx.zeroDeviceAsync(stream); vy.zeroDeviceAsync(stream); vz.zeroDeviceAsync(stream);
    scalarsU_.zeroDeviceAsync(stream);
    frameDt = (double)(1.0f / 24.0f);
    sorThreshold_ = kSorThreshold;
    sorMaxIterations_ = kSorMaxIterations;
    synchronize();
}

Particles::~Particles() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        dm_.setStream(nullptr);  // drained: the arrays' releases need no further waits
        (void)hipStreamDestroy(stream);
    }
}

void Particles::synchronize() { FLIP2_HIP(hipStreamSynchronize(stream)); }

void Particles::ensureWorkspace(size_t bytes) {
    if (workspace_.size() < bytes) workspace_.resizeAsync((uint32_t)bytes, stream);
}

uint32_t Particles::readU32(const uint32_t* dev) {
    uint32_t v = 0;
    FLIP2_HIP(hipMemcpyAsync(&v, dev, sizeof(v), hipMemcpyDeviceToHost, stream));
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
    grid.negX = (float)nx; grid.negY = (float)ny; grid.negZ = (float)nz;
    grid.sizeX = x; grid.sizeY = y; grid.sizeZ = z;
    grid.cellSize = (float)cellSize;
    yDimRows_ = row_table_rows(grid);
    yDimNumUsedGridNodes.resizeAsync(yDimRows_, stream);
}

void Particles::alignParticlesToGrid() {
    FLIP2_REQUIRE(grid.sizeX > 0, kErrState, "setDomain has not been called");
    rc::launch_root_cell(px.devPtr(), py.devPtr(), pz.devPtr(), size, grid, gridCell.devPtr(), stream);
}

// Particles::sortParticles (particles.cu:52-78): keys sorted with the
// permutation; positions gathered, velocities keep their order.  Only the
// bits a key of this domain can use are sorted (identical result to a
// 32-bit sort).
void Particles::sortParticles() {
    const uint64_t maxKey = (uint64_t)grid.sizeX + (uint64_t)grid.sizeY * grid.sizeX +
                            (uint64_t)grid.sizeZ * grid.sizeX * grid.sizeY;
    const uint32_t endBit = std::min<uint32_t>(32, std::max<uint32_t>(1, bits_for(maxKey)));
    const size_t wsBytes = sort_workspace_bytes(size, 0, endBit);
    ensureWorkspace(wsBytes);
    sort_pairs_u32(gridCell.devPtr(), nullptr, sortedKeys_.devPtr(), reorderedGridIndices.devPtr(),
                   keysAlt_.devPtr(), permAlt_.devPtr(), size, 0, endBit, workspace_.devPtr(), wsBytes, stream);
    gridCell.swapDevice(sortedKeys_);
    rc::launch_gather3(reorderedGridIndices.devPtr(), size, px.devPtr(), py.devPtr(), pz.devPtr(),
                       tmpX_.devPtr(), tmpY_.devPtr(), tmpZ_.devPtr(), stream);
    px.swapDevice(tmpX_);
    py.swapDevice(tmpY_);
    pz.swapDevice(tmpZ_);
}

void Particles::resizeNodes(uint32_t B) {
    numUsedGridNodes = B;
    gridNodeIndicesToFirstParticleIndex.resizeAsync(B, stream);
    nodeIndexUsedVoxels.resizeAsync(B, stream);
    nodeCounts_.resizeAsync(B, stream);
    usedMask_.resizeAsync(B * 8u, stream);
    solidMask_.resizeAsync(B * 8u, stream);
}

void Particles::resizeVoxels(uint32_t V) {
    numUsedVoxels = V;
    voxelIDsUsed.resizeAsync(V, stream);
    solids.resizeAsync(V, stream);
    voxelsUx.resizeAsync(V, stream);
    voxelsUy.resizeAsync(V, stream);
    voxelsUz.resizeAsync(V, stream);
    divU.resizeAsync(V, stream);
    p.resizeAsync(V, stream);
    residuals.resizeAsync(V, stream);
}

// Particles::generateVoxels (particles.cu:323-368).
void Particles::generateVoxels() {
    // head flags fused into the inclusive scan -> 1-based node id per particle
    const size_t scanWs = scan_workspace_bytes(size);
    ensureWorkspace(scanWs);
    uint32_t* total = scalarsU_.devPtr();
    const bool vec = ((reinterpret_cast<uintptr_t>(gridCell.devPtr()) |
                       reinterpret_cast<uintptr_t>(uniqueGridNodeIndices.devPtr())) & 15u) == 0;
    scan_launch(ScanHeadFlagInput{gridCell.devPtr()}, uniqueGridNodeIndices.devPtr(), size, false, vec,
                workspace_.devPtr(), scanWs, total, stream);
    const uint32_t B = readU32(total);
    resizeNodes(B);
    rc::launch_map_nodes(uniqueGridNodeIndices.devPtr(), size, gridNodeIndicesToFirstParticleIndex.devPtr(),
                         stream);
    rc::launch_fill_u32(yDimNumUsedGridNodes.devPtr(), yDimRows_, B, stream);
    rc::launch_row_table(B, gridNodeIndicesToFirstParticleIndex.devPtr(), gridCell.devPtr(),
                         yDimNumUsedGridNodes.devPtr(), grid.sizeX, stream);
    rc::launch_mark_voxels(B, size, grid, gridCell.devPtr(), gridNodeIndicesToFirstParticleIndex.devPtr(),
                           px.devPtr(), py.devPtr(), pz.devPtr(), usedMask_.devPtr(), solidMask_.devPtr(),
                           nodeCounts_.devPtr(), scalarsU_.devPtr() + 1, stream);
    const size_t scanWs2 = scan_workspace_bytes(B);
    ensureWorkspace(scanWs2);
    scan_u32(nodeCounts_.devPtr(), nodeIndexUsedVoxels.devPtr(), B, false, workspace_.devPtr(), scanWs2, total,
             stream);
    const uint32_t V = readU32(total);
    resizeVoxels(V);
    p.zeroDeviceAsync(stream);
    rc::launch_compact_voxels(B, usedMask_.devPtr(), solidMask_.devPtr(), nodeIndexUsedVoxels.devPtr(),
                              voxelIDsUsed.devPtr(), solids.devPtr(), stream);
}

// Particles::particleVelToVoxels (particles.cu:500-505).
void Particles::particleVelToVoxels() {
    rc::launch_p2g(numUsedGridNodes, size, grid, gridCell.devPtr(), gridNodeIndicesToFirstParticleIndex.devPtr(),
                   px.devPtr(), py.devPtr(), pz.devPtr(), vx.devPtr(), vy.devPtr(), vz.devPtr(),
                   nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(), solids.devPtr(), voxelsUx.devPtr(),
                   voxelsUy.devPtr(), voxelsUz.devPtr(), stream);
}

// Particles::getCourantDt (particles.cu:493-498) with the sampled maxima of
// CudaVec::getMax (cudaVec.cu:205-233).
double Particles::getCourantDt() {
    rc::Ptr3 arrs{{vx.devPtr(), vy.devPtr(), vz.devPtr()}};
    rc::launch_sampled_absmax(arrs, 3, size, absmaxTmp_.devPtr(), scalarsD_.devPtr(), stream);
    double m[3];
    FLIP2_HIP(hipMemcpyAsync(m, scalarsD_.devPtr(), sizeof(m), hipMemcpyDeviceToHost, stream));
    synchronize();
    const double maxVel = std::fmax(std::fabs(m[0]), std::fmax(std::fabs(m[1]), std::fabs(m[2])));
    const double voxelSize = (double)grid.cellSize / ((double)kRcVoxels1D - 2.0 * std::floor(kRcRadius));
    return 0.7 * (voxelSize / maxVel);
}

void Particles::applyGravity(double dt) {
    rc::launch_gravity(numUsedVoxels, solids.devPtr(), voxelsUy.devPtr(), (double)(-9.8f) * dt, false, stream);
}

void Particles::removeGravity(double dt) {
    rc::launch_gravity(numUsedVoxels, solids.devPtr(), voxelsUy.devPtr(), (double)(-9.8f) * dt, true, stream);
}

void Particles::calcDivU() {
    rc::launch_divu(numUsedGridNodes, grid, gridCell.devPtr(), gridNodeIndicesToFirstParticleIndex.devPtr(),
                    yDimNumUsedGridNodes.devPtr(), nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(),
                    voxelsUx.devPtr(), voxelsUy.devPtr(), voxelsUz.devPtr(), divU.devPtr(), stream);
}

double Particles::solveScale(double dt) const {
    const double vs = (double)(grid.cellSize / (float)(2u << kRcRefinement));
    return dt / (kDensity * vs * vs);
}

void Particles::gsIterations(double dt, uint32_t count) {
    rc::launch_sor(numUsedGridNodes, nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(), solids.devPtr(),
                   divU.devPtr(), p.devPtr(), residuals.devPtr(), solveScale(dt), kSorOmega, count, stream);
}

// cudaGSiteration (voxelSolveFunctions.cu:391-412): the iterations between
// two residual checks run inside one LDS-resident kernel launch.
double Particles::gsSolve(double dt, double threshold, uint32_t maxIterations, uint32_t* itersOut) {
    ensureAbsmaxTmp(numUsedVoxels);
    return rc::gs_solve(numUsedGridNodes, numUsedVoxels, nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(),
                        solids.devPtr(), divU.devPtr(), p.devPtr(), residuals.devPtr(), solveScale(dt), threshold,
                        maxIterations, absmaxTmp_.devPtr(), scalarsD_.devPtr() + 3, stream, itersOut);
}

// Particles::pressureSolve (particles.cu:507-555).
void Particles::pressureSolve() {
    double previousTerminatingResidual = 100.0, terminatingResidual = 0.0;
    uint32_t its = 0;
    lastSorIterations_ = 0;
    lastSolveAttempts_ = 0;
    dt = getCourantDt();
    if (frameDt - elapsedTimeThisFrame < dt) dt = frameDt - elapsedTimeThisFrame;
    applyGravity(dt);
    calcDivU();
    for (;;) {
        terminatingResidual = gsSolve(dt, sorThreshold_, sorMaxIterations_, &its);
        lastSorIterations_ += its;
        lastSolveAttempts_ += 1;
        if (!(previousTerminatingResidual - terminatingResidual > 0)) break;
        if (terminatingResidual < sorThreshold_) break;
        removeGravity(dt);
        dt /= 2.0;
        p.zeroDeviceAsync(stream);
        applyGravity(dt);
        previousTerminatingResidual = terminatingResidual;
        calcDivU();
    }
    lastResidual_ = terminatingResidual;
    elapsedTime += dt;
    elapsedTimeThisFrame += dt;
}

// Particles::updateVoxelVelocities (particles.cu:557-560).
void Particles::updateVoxelVelocities() {
    rc::launch_velocity_update(numUsedGridNodes, nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(),
                               solids.devPtr(), p.devPtr(), voxelsUx.devPtr(), voxelsUy.devPtr(),
                               voxelsUz.devPtr(), solveScale(dt), stream);
}

// Particles::voxelVelsToParticles (particles.cu:742-749).
void Particles::voxelVelsToParticles() {
    rc::launch_g2p(numUsedGridNodes, size, grid, gridCell.devPtr(), gridNodeIndicesToFirstParticleIndex.devPtr(),
                   px.devPtr(), py.devPtr(), pz.devPtr(), vx.devPtr(), vy.devPtr(), vz.devPtr(),
                   nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(), voxelsUx.devPtr(), voxelsUy.devPtr(),
                   voxelsUz.devPtr(), stream);
}

// Particles::advectParticles (particles.cu:751-760).
void Particles::advectParticles() {
    rc::launch_advect(size, dt, px.devPtr(), py.devPtr(), pz.devPtr(), vx.devPtr(), vy.devPtr(), vz.devPtr(),
                      stream);
}

void Particles::moveSolidsOnly() {
    FLIP2_HIP(hipMemsetAsync(scalarsU_.devPtr() + 2, 0, sizeof(uint32_t), stream));
    rc::launch_move_solids(numUsedGridNodes, size, grid, gridCell.devPtr(),
                           gridNodeIndicesToFirstParticleIndex.devPtr(), px.devPtr(), py.devPtr(), pz.devPtr(),
                           nodeIndexUsedVoxels.devPtr(), voxelIDsUsed.devPtr(), solids.devPtr(),
                           moveSolidsCounter_, scalarsU_.devPtr() + 2, stream);
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
    frameDt = 1.0 / fps;
    dt = frameDt / 10.0;
    elapsedTimeThisFrame = 0.0;
    substeps_ = 0;
    while (elapsedTimeThisFrame < frameDt) {
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
    std::vector<double> x(size), y(size), z(size);
    getPositions(x.data(), y.data(), z.data());
    std::ofstream file(fileName);
    FLIP2_REQUIRE(file.good(), kErrArgument, "cannot open " + fileName);
    for (uint32_t i = 0; i < size; ++i) file << x[i] << ", " << y[i] << ", " << z[i] << "\n";
}

void Particles::setPositions(const double* x, const double* y, const double* z) {
    px.uploadFrom(x, size, stream);
    py.uploadFrom(y, size, stream);
    pz.uploadFrom(z, size, stream);
    synchronize();
}

void Particles::setVelocities(const double* x, const double* y, const double* z) {
    vx.uploadFrom(x, size, stream);
    vy.uploadFrom(y, size, stream);
    vz.uploadFrom(z, size, stream);
    synchronize();
}

void Particles::getPositions(double* x, double* y, double* z) {
    px.downloadTo(x, size, stream);
    py.downloadTo(y, size, stream);
    pz.downloadTo(z, size, stream);
    synchronize();
}

void Particles::getVelocities(double* x, double* y, double* z) {
    vx.downloadTo(x, size, stream);
    vy.downloadTo(y, size, stream);
    vz.downloadTo(z, size, stream);
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
    const uint32_t B = numUsedGridNodes, V = numUsedVoxels, n = size;
    if (name == "px") r = copy_out(px, n, host, capacity, stream);
    else if (name == "py") r = copy_out(py, n, host, capacity, stream);
    else if (name == "pz") r = copy_out(pz, n, host, capacity, stream);
    else if (name == "vx") r = copy_out(vx, n, host, capacity, stream);
    else if (name == "vy") r = copy_out(vy, n, host, capacity, stream);
    else if (name == "vz") r = copy_out(vz, n, host, capacity, stream);
    else if (name == "gridCell") r = copy_out(gridCell, n, host, capacity, stream);
    else if (name == "reorderedGridIndices") r = copy_out(reorderedGridIndices, n, host, capacity, stream);
    else if (name == "uniqueGridNodeIndices") r = copy_out(uniqueGridNodeIndices, n, host, capacity, stream);
    else if (name == "gridNodeIndicesToFirstParticleIndex")
        r = copy_out(gridNodeIndicesToFirstParticleIndex, B, host, capacity, stream);
    else if (name == "yDimNumUsedGridNodes") r = copy_out(yDimNumUsedGridNodes, yDimRows_, host, capacity, stream);
    else if (name == "nodeIndexUsedVoxels") r = copy_out(nodeIndexUsedVoxels, B, host, capacity, stream);
    else if (name == "voxelIDsUsed") r = copy_out(voxelIDsUsed, V, host, capacity, stream);
    else if (name == "solids") r = copy_out(solids, V, host, capacity, stream);
    else if (name == "voxelsUx") r = copy_out(voxelsUx, V, host, capacity, stream);
    else if (name == "voxelsUy") r = copy_out(voxelsUy, V, host, capacity, stream);
    else if (name == "voxelsUz") r = copy_out(voxelsUz, V, host, capacity, stream);
    else if (name == "divU") r = copy_out(divU, V, host, capacity, stream);
    else if (name == "p") r = copy_out(p, V, host, capacity, stream);
    else if (name == "residuals") r = copy_out(residuals, V, host, capacity, stream);
    else throw Error(kErrArgument, "unknown array " + name);
    synchronize();
    return r;
}

void Particles::setArray(const std::string& name, const void* host, size_t count) {
    auto put = [&](auto& vec, size_t expect) {
        using T = typename std::remove_reference<decltype(vec[0])>::type;
        FLIP2_REQUIRE(count == expect, kErrRange, "setArray: count mismatch for " + name);
        vec.uploadFrom(reinterpret_cast<const T*>(host), (uint32_t)count, stream);
    };
    const uint32_t n = size;
    if (name == "px") put(px, n);
    else if (name == "py") put(py, n);
    else if (name == "pz") put(pz, n);
    else if (name == "vx") put(vx, n);
    else if (name == "vy") put(vy, n);
    else if (name == "vz") put(vz, n);
    else if (name == "gridCell") put(gridCell, n);
    else if (name == "voxelsUx") put(voxelsUx, numUsedVoxels);
    else if (name == "voxelsUy") put(voxelsUy, numUsedVoxels);
    else if (name == "voxelsUz") put(voxelsUz, numUsedVoxels);
    else if (name == "divU") put(divU, numUsedVoxels);
    else if (name == "p") put(p, numUsedVoxels);
    else throw Error(kErrArgument, "array not settable: " + name);
    synchronize();
}

void Particles::setSolverLimits(double threshold, uint32_t maxIterations) {
    FLIP2_REQUIRE(maxIterations > 0, kErrArgument, "maxIterations must be positive");
    sorThreshold_ = threshold;
    sorMaxIterations_ = maxIterations;
}

void Particles::setTime(double dt, double frameDt, double elapsedTime, double elapsedTimeThisFrame) {
    this->dt = dt;
    this->frameDt = frameDt;
    this->elapsedTime = elapsedTime;
    this->elapsedTimeThisFrame = elapsedTimeThisFrame;
}

void Particles::ensureAbsmaxTmp(uint32_t n) {
    const size_t need = rc::sampled_absmax_tmp_doubles(n, 3);
    if (absmaxTmp_.size() < need) absmaxTmp_.resizeAsync((uint32_t)need, stream);
}

RcStats Particles::stats() {
    RcStats s{};
    s.numUsedGridNodes = numUsedGridNodes;
    s.numUsedVoxels = numUsedVoxels;
    s.lastSorIterations = lastSorIterations_;
    s.lastSolveAttempts = lastSolveAttempts_;
    s.lastResidual = lastResidual_;
    s.dt = dt;
    s.frameDt = frameDt;
    s.elapsedTime = elapsedTime;
    s.elapsedTimeThisFrame = elapsedTimeThisFrame;
    uint32_t u[4] = {0, 0, 0, 0};
    FLIP2_HIP(hipMemcpyAsync(u, scalarsU_.devPtr(), sizeof(u), hipMemcpyDeviceToHost, stream));
    synchronize();
    s.outOfTileEvents = u[1];
    s.relocatedParticles = u[2];
    s.substeps = substeps_;
    return s;
}

}  // namespace flip2

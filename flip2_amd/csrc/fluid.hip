// North-star FLIP step driver (see fluid.hpp).
#include <algorithm>
#include <cstring>
#include <cstdio>
#include "fluid.hpp"

#include "scan_impl.hpp"

#include <chrono>
#include <cmath>

namespace flip2 {

namespace {
// The residual trend of a refinement's checked history entries: the
// per-iteration rate of rr between the last two checked iterations, used to
// guess whether the batch in flight reaches the threshold.  Such a batch is
// checked before the next one is queued: work queued behind a converged
// batch returns at once on the device but still costs its launches.
struct RrTrend {
    double last = -1.0, rate = 1.0;
    uint32_t lastIt = 0;
    void see(double rr, uint32_t it) {
        if (!(rr > 0.0)) return;  // zero or NaN: the batch check decides
        if (last > 0.0 && it > lastIt) rate = std::pow(rr / last, 1.0 / (double)(it - lastIt));
        last = rr;
        lastIt = it;
    }
    // rr extrapolated to iteration `it` lies within a factor 4 of thr2 (no
    // when the trend is unknown or not decreasing)
    bool likely(uint32_t it, double thr2) const {
        if (!(last > 0.0) || it <= lastIt || !(rate < 1.0)) return false;
        return last * std::pow(rate, (double)(it - lastIt)) <= 4.0 * thr2;
    }
};
}  // namespace

namespace {
// Largest accepted inner-iteration cap and residual-check interval: the
// solver keeps 4 doubles of history per inner iteration on the device and
// reads (check interval + 1) x 4 of them back per check.
constexpr uint32_t kMaxIterationsCap = 1u << 24;
constexpr uint32_t kMaxCheckEvery = 1u << 16;

// One validator for every way a configuration enters the simulation
// (creation, set_config, snapshot load).
void validate_config(const FluidConfig& c) {
    FLIP2_REQUIRE(std::isfinite(c.dt) && c.dt > 0.0f, kErrArgument, "config: dt must be positive and finite");
    FLIP2_REQUIRE(std::isfinite(c.gravity), kErrArgument, "config: gravity must be finite");
    FLIP2_REQUIRE(std::isfinite(c.flip) && c.flip >= 0.0f && c.flip <= 1.0f, kErrArgument,
                  "config: FLIP ratio must lie in [0, 1]");
    FLIP2_REQUIRE(c.preconditioner >= kPrecJacobi && c.preconditioner <= kPrecLearned, kErrArgument,
                  "config: unknown preconditioner");
    FLIP2_REQUIRE(std::isfinite(c.tolerance) && c.tolerance > 0.0, kErrArgument,
                  "config: tolerance must be positive and finite");
    FLIP2_REQUIRE(c.maxIterations >= 1 && c.maxIterations <= kMaxIterationsCap, kErrArgument,
                  "config: max_iterations must lie in [1, 2^24]");
    FLIP2_REQUIRE(c.checkEvery >= 1 && c.checkEvery <= kMaxCheckEvery, kErrArgument,
                  "config: check_every must lie in [1, 65536]");
    FLIP2_REQUIRE(c.learnedCheckEvery >= 1 && c.learnedCheckEvery <= kMaxCheckEvery, kErrArgument,
                  "config: the flexible-CG check interval must lie in [1, 65536]");
    FLIP2_REQUIRE(std::isfinite(c.margin) && c.margin >= 0.0f && c.margin < 0.5f, kErrArgument,
                  "config: margin must lie in [0, 0.5)");
    FLIP2_REQUIRE(c.learnedSmoother == 0 || c.learnedSmoother == 1, kErrArgument,
                  "config: the learned smoother is 0 (damped Jacobi) or 1 (IC(0))");
    FLIP2_REQUIRE(c.learnedSmoothSteps == 1 || c.learnedSmoothSteps == 2, kErrArgument,
                  "config: 1 or 2 IC(0) smoothing steps around the learned correction");
}

// Device history entries the solver needs for a configuration.
size_t history_entries(const FluidConfig& c) { return 4 * ((size_t)c.maxIterations + 2); }
}  // namespace

FluidSim::FluidSim(uint32_t maxParticles, const FluidConfig& cfg, int device)
    : cfg_((validate_config(cfg), cfg)), dm_(device), maxParticles_(maxParticles) {
    const Grid& g = cfg.grid;
    FLIP2_REQUIRE(g.sizeX % 8 == 0 && g.sizeY % 8 == 0 && g.sizeZ % 8 == 0 && g.sizeX && g.sizeY && g.sizeZ,
                  kErrArgument, "grid sizes must be positive multiples of 8");
    FLIP2_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    dm_.setStream(stream_);
    pipe_.reset(new P2GPipeline(maxParticles, g, &dm_));
    numCells_ = g.sizeX * g.sizeY * g.sizeZ;
    numBricks_ = numCells_ / 512u;
    rec_.resizeAsync(2 * maxParticles, stream_);
    for (auto* a : {&U_, &V_, &W_, &U0_, &V0_, &W0_}) a->resizeAsync(numCells_, stream_);
    brickFlag_.resizeAsync(numBricks_ + 1, stream_);
    brickExcl_.resizeAsync(numBricks_ + 1, stream_);
    brickSlot_.resizeAsync(numBricks_, stream_);
    activeBricks_.resizeAsync(numBricks_, stream_);
    total_.resizeAsync(4, stream_);
    scal_.resizeAsync(16, stream_);
    stop_.resizeAsync(1, stream_);
    hist_.resizeAsync((uint32_t)history_entries(cfg), stream_);
    ws_.resizeAsync((uint32_t)scan_workspace_bytes(numBricks_ + 1), stream_);
    synchronize();
}

FluidSim::~FluidSim() {
    dropLearnedGraph();
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        dm_.setStream(nullptr);  // drained: the arrays' releases need no further waits
        (void)hipStreamDestroy(stream_);
    }
    for (auto& e : histEvent_)
        if (e) (void)hipEventDestroy(e);
    if (histPinned_) (void)hipHostFree(histPinned_);
}

// Pinned host buffer of at least `doubles` doubles (kept between solves).
double* FluidSim::pinnedHistory(size_t doubles) {
    if (doubles > histPinnedDoubles_) {
        if (histPinned_) {
            synchronize();  // no copy into the old buffer is pending
            FLIP2_HIP(hipHostFree(histPinned_));
            histPinned_ = nullptr;
            histPinnedDoubles_ = 0;
        }
        FLIP2_HIP(hipHostMalloc((void**)&histPinned_, doubles * sizeof(double), hipHostMallocDefault));
        histPinnedDoubles_ = doubles;
    }
    for (auto& e : histEvent_)
        if (!e) FLIP2_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return histPinned_;
}

void FluidSim::synchronize() { FLIP2_HIP(hipStreamSynchronize(stream_)); }

void FluidSim::setConfig(const FluidConfig& cfg) {
    FLIP2_REQUIRE(cfg.grid.sizeX == cfg_.grid.sizeX && cfg.grid.sizeY == cfg_.grid.sizeY &&
                      cfg.grid.sizeZ == cfg_.grid.sizeZ,
                  kErrArgument, "the grid cannot change");
    validate_config(cfg);
    if (hist_.size() < history_entries(cfg)) hist_.resizeAsync((uint32_t)history_entries(cfg), stream_);
    cfg_ = cfg;
}

void FluidSim::setParticles(uint32_t n, const float* x, const float* y, const float* z, const float* vx,
                            const float* vy, const float* vz) {
    bricksFresh_ = false;
    keysFresh_ = false;
    FLIP2_REQUIRE(n <= maxParticles_, kErrRange, "more particles than the simulation was created for");
    n_ = n;
    std::vector<float4> r(2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        r[2 * (size_t)i] = make_float4(x[i], y[i], z[i], vx[i]);
        r[2 * (size_t)i + 1] = make_float4(vy[i], vz[i], 0.f, 0.f);
    }
    if (n) rec_.uploadFrom(r.data(), 2 * n, stream_);
    synchronize();
}

void FluidSim::getParticles(float* x, float* y, float* z, float* vx, float* vy, float* vz) {
    std::vector<float4> r(2 * (size_t)n_);
    if (n_) rec_.downloadTo(r.data(), 2 * n_, stream_);
    synchronize();
    float* dst[6] = {x, y, z, vx, vy, vz};
    for (uint32_t i = 0; i < n_; ++i) {
        const float* f = reinterpret_cast<const float*>(&r[2 * (size_t)i]);
        for (int c = 0; c < 6; ++c)
            if (dst[c]) dst[c][i] = f[c];
    }
}

void FluidSim::setSolidMask(const uint8_t* mask) {
    const Grid& g = cfg_.grid;
    std::vector<uint8_t> tiled(numCells_, 0);
    const uint32_t nbX = g.sizeX / 8, nbY = g.sizeY / 8;
    bool any = false;
    for (uint32_t k = 0; k < g.sizeZ; ++k)
        for (uint32_t j = 0; j < g.sizeY; ++j)
            for (uint32_t i = 0; i < g.sizeX; ++i) {
                const uint8_t v = mask ? mask[((size_t)k * g.sizeY + j) * g.sizeX + i] : 0;
                const uint32_t t = (((k >> 3) * nbY + (j >> 3)) * nbX + (i >> 3)) * 512u + (k & 7u) * 64u +
                                   (j & 7u) * 8u + (i & 7u);
                tiled[t] = v ? 1 : 0;
                any |= v != 0;
            }
    hasStatic_ = any;
    if (any) {
        solidStatic_.resizeAsync(numCells_, stream_);
        solidStatic_.uploadFrom(tiled.data(), numCells_, stream_);
    }
    rebuildSolid();
    synchronize();
}

void FluidSim::setSolidBox(const float* lo, const float* hi, const float* vel) {
    box_ = ns::SolidBox{};
    if (lo) {
        FLIP2_REQUIRE(hi != nullptr, kErrArgument, "solid box needs both corners");
        for (int d = 0; d < 3; ++d) {
            box_.lo[d] = lo[d];
            box_.hi[d] = hi[d];
            box_.vel[d] = vel ? vel[d] : 0.f;
        }
        box_.active = 1;
    }
    rebuildSolid();
}

namespace {
// Snapshot layout (little-endian): magic, version, Grid (28 bytes), dt,
// gravity, flip, margin (float), tolerance (double), preconditioner,
// maxIterations, checkEvery, n (uint32), 2 n float4 particle records,
// static-mask flag (uint8) and, when set, the mask in the simulation's
// brick-tiled cell order (one byte per cell), then the solid box.
constexpr char kSnapMagic[8] = {'F', 'L', 'I', 'P', '2', 'N', 'S', '1'};
constexpr uint32_t kSnapVersion = 1;

struct SnapFile {
    std::FILE* f = nullptr;
    SnapFile(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {
        FLIP2_REQUIRE(f != nullptr, kErrArgument, "snapshot: cannot open " + path);
    }
    ~SnapFile() {
        if (f) std::fclose(f);
    }
    void put(const void* p, size_t bytes) {
        FLIP2_REQUIRE(std::fwrite(p, 1, bytes, f) == bytes, kErrArgument, "snapshot: write failed");
    }
    void get(void* p, size_t bytes) {
        FLIP2_REQUIRE(std::fread(p, 1, bytes, f) == bytes, kErrArgument, "snapshot: truncated file");
    }
    template <class T>
    void put(const T& v) {
        put(&v, sizeof(T));
    }
    template <class T>
    T get() {
        T v;
        get(&v, sizeof(T));
        return v;
    }
};
}  // namespace

void FluidSim::saveSnapshot(const std::string& path) {
    std::vector<float4> rec(2 * (size_t)n_);
    if (n_) rec_.downloadTo(rec.data(), 2 * n_, stream_);
    std::vector<uint8_t> mask;
    if (hasStatic_) {
        mask.resize(numCells_);
        solidStatic_.downloadTo(mask.data(), numCells_, stream_);
    }
    synchronize();
    SnapFile out(path, "wb");
    out.put(kSnapMagic, sizeof(kSnapMagic));
    out.put(kSnapVersion);
    out.put(cfg_.grid);
    out.put(cfg_.dt);
    out.put(cfg_.gravity);
    out.put(cfg_.flip);
    out.put(cfg_.margin);
    out.put(cfg_.tolerance);
    out.put((uint32_t)cfg_.preconditioner);
    out.put(cfg_.maxIterations);
    out.put(cfg_.checkEvery);
    out.put(n_);
    if (n_) out.put(rec.data(), rec.size() * sizeof(float4));
    out.put((uint8_t)(hasStatic_ ? 1 : 0));
    if (hasStatic_) out.put(mask.data(), mask.size());
    out.put(box_);
}

void FluidSim::loadSnapshot(const std::string& path) {
    bricksFresh_ = false;
    keysFresh_ = false;
    SnapFile in(path, "rb");
    char magic[sizeof(kSnapMagic)];
    in.get(magic, sizeof(magic));
    FLIP2_REQUIRE(std::memcmp(magic, kSnapMagic, sizeof(magic)) == 0, kErrArgument, "snapshot: not a snapshot file");
    FLIP2_REQUIRE(in.get<uint32_t>() == kSnapVersion, kErrArgument, "snapshot: unsupported version");
    FluidConfig cfg = cfg_;
    cfg.grid = in.get<Grid>();
    FLIP2_REQUIRE(cfg.grid.sizeX == cfg_.grid.sizeX && cfg.grid.sizeY == cfg_.grid.sizeY &&
                      cfg.grid.sizeZ == cfg_.grid.sizeZ,
                  kErrArgument, "snapshot: grid dimensions differ from this simulation's");
    cfg.dt = in.get<float>();
    cfg.gravity = in.get<float>();
    cfg.flip = in.get<float>();
    cfg.margin = in.get<float>();
    cfg.tolerance = in.get<double>();
    const uint32_t prec = in.get<uint32_t>();
    FLIP2_REQUIRE(prec <= (uint32_t)kPrecLearned, kErrArgument, "snapshot: bad preconditioner");
    cfg.preconditioner = (int)prec;
    cfg.maxIterations = in.get<uint32_t>();
    cfg.checkEvery = in.get<uint32_t>();
    const uint32_t n = in.get<uint32_t>();
    FLIP2_REQUIRE(n <= maxParticles_, kErrRange, "snapshot: more particles than the simulation was created for");
    std::vector<float4> rec(2 * (size_t)n);
    if (n) in.get(rec.data(), rec.size() * sizeof(float4));
    const bool stat = in.get<uint8_t>() != 0;
    std::vector<uint8_t> mask;
    if (stat) {
        mask.resize(numCells_);
        in.get(mask.data(), mask.size());
    }
    const ns::SolidBox box = in.get<ns::SolidBox>();
    // everything read: apply
    setConfig(cfg);
    n_ = n;
    if (n) rec_.uploadFrom(rec.data(), 2 * n, stream_);
    hasStatic_ = stat;
    if (stat) {
        solidStatic_.resizeAsync(numCells_, stream_);
        solidStatic_.uploadFrom(mask.data(), numCells_, stream_);
    }
    box_ = box;
    rebuildSolid();
    synchronize();
}

void FluidSim::rebuildSolid() {
    ++solidVersion_;
    hasSolid_ = hasStatic_ || box_.active;
    if (!hasSolid_) return;
    if (solid_.size() < numCells_) solid_.resizeAsync(numCells_, stream_);
    ns::launch_rasterize_solid(hasStatic_ ? solidStatic_.devPtr() : nullptr, box_, cfg_.grid, solid_.devPtr(),
                               stream_);
}

void FluidSim::ensureSlots(uint32_t numActive) {
    if (numActive <= slotCapacity_) return;
    const uint32_t cap = numActive + numActive / 4 + 8;
    const uint32_t cells = cap * 512u;
    nbrSlot_.resizeAsync(cap * 6u, stream_);
    cellType_.resizeAsync(cells, stream_);
    coef_.resizeAsync(cells, stream_);
    for (auto* a : {&b64_, &x64_, &r64_}) a->resizeAsync(cells, stream_);
    for (auto* a : {&r_, &z_, &p_, &p2_, &q_, &e_, &precon_, &aux0_, &aux1_, &aux2_, &aux3_})
        a->resizeAsync(cells, stream_);
    partials_.resizeAsync(cap * (uint32_t)ns::kMaxFlexDirections + 64u, stream_);
    slotCapacity_ = cap;
    ringSlots_ = 0;
}

void FluidSim::ensureRing(uint32_t slots) {
    if (ringSlots_ >= slots) return;
    const size_t cells = (size_t)slotCapacity_ * 512u;
    pRing_.resizeAsync((uint32_t)(cells * slots), stream_);
    qRing_.resizeAsync((uint32_t)(cells * slots), stream_);
    zq_.resizeAsync(ns::kMaxFlexDirections, stream_);
    ringSlots_ = slots;
}

// z = M(r) for the learned preconditioner: the network alone, or wrapped in
// a damped-Jacobi pre-/post-smoothing step.
void FluidSim::applyLearnedEager(uint32_t na, bool newSystem) {
    const float w = cfg_.learnedSmoothing;
    const uint32_t* stop = stop_.devPtr();
    if (w > 0.f && cfg_.learnedSmoother == 1) {
        // IC(0) smoothing: t = M^-1 r (then, for two steps, t += M^-1 (r -
        // A t)); r1 = r - A t; y = N(r1); x = (t + y) + M^-1 (r - A (t + y))
        // (then x += M^-1 (r - A x)); z = x.  aux2 / aux3: t and x, aux0:
        // the residuals, aux1: y.
        double* pairs = partials_.devPtr();
        const bool two = cfg_.learnedSmoothSteps >= 2;
        auto residual = [&](const float* a, const float* b, double* pr) {
            ns::launch_smooth_residual(na, r_.devPtr(), a, b, coef_.devPtr(), nbrSlot_.devPtr(), cellType_.devPtr(),
                                       aux0_.devPtr(), pr, stream_, stop);
        };
        auto smooth = [&](const float* in, const float* a, const float* b, float* out) {
            ns::launch_mic_smooth(na, in, a, b, coef_.devPtr(), precon_.devPtr(), out, stream_, stop);
        };
        smooth(r_.devPtr(), nullptr, nullptr, aux2_.devPtr());
        float* t = aux2_.devPtr();
        if (two) {
            residual(t, nullptr, nullptr);
            smooth(aux0_.devPtr(), t, nullptr, aux3_.devPtr());
            t = aux3_.devPtr();
        }
        residual(t, nullptr, pairs);
        learned_(na, aux0_.devPtr(), aux1_.devPtr(), newSystem, pairs, stream_, stop);
        residual(t, aux1_.devPtr(), nullptr);
        if (!two) {
            smooth(aux0_.devPtr(), t, aux1_.devPtr(), z_.devPtr());
        } else {
            float* x = t == aux2_.devPtr() ? aux3_.devPtr() : aux2_.devPtr();
            smooth(aux0_.devPtr(), t, aux1_.devPtr(), x);
            residual(x, nullptr, nullptr);
            smooth(aux0_.devPtr(), x, nullptr, z_.devPtr());
        }
    } else if (w > 0.f) {
        // the partial sums of the network input's scale come with the smoothing
        double* pairs = partials_.devPtr();
        ns::launch_presmooth(na, r_.devPtr(), coef_.devPtr(), nbrSlot_.devPtr(), cellType_.devPtr(), w, aux0_.devPtr(),
                             pairs, stream_, stop);
        learned_(na, aux0_.devPtr(), aux1_.devPtr(), newSystem, pairs, stream_, stop);
        ns::launch_postsmooth(na, r_.devPtr(), aux1_.devPtr(), coef_.devPtr(), nbrSlot_.devPtr(), w, z_.devPtr(),
                              stream_, stop);
    } else {
        learned_(na, r_.devPtr(), z_.devPtr(), newSystem, nullptr, stream_, stop);
    }
}

void FluidSim::dropLearnedGraph() {
    if (learnedGraphExec_) (void)hipGraphExecDestroy(learnedGraphExec_);
    if (learnedGraph_) (void)hipGraphDestroy(learnedGraph_);
    learnedGraphExec_ = nullptr;
    learnedGraph_ = nullptr;
}

// The first apply of a solve runs eagerly (it rebuilds the network's masks);
// the second is captured into a HIP graph (kernels only: no memset nodes)
// that the remaining applies of the solve replay, saving the launch gaps of
// its ~20 kernels.
void FluidSim::applyLearned(uint32_t na) {
    if (learnedFresh_ || !cfg_.learnedGraph) {
        const bool fresh = learnedFresh_;
        learnedFresh_ = false;
        dropLearnedGraph();
        applyLearnedEager(na, fresh);
        return;
    }
    if (!learnedGraphExec_) {
        FLIP2_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            applyLearnedEager(na, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        FLIP2_HIP(hipStreamEndCapture(stream_, &learnedGraph_));
        FLIP2_HIP(hipGraphInstantiate(&learnedGraphExec_, learnedGraph_, nullptr, nullptr, 0));
    }
    FLIP2_HIP(hipGraphLaunch(learnedGraphExec_, stream_));
}

double FluidSim::readScalar(const double* dev) {
    double v = 0.0;
    FLIP2_HIP(hipMemcpyAsync(&v, dev, sizeof(v), hipMemcpyDeviceToHost, stream_));
    synchronize();
    return v;
}

void FluidSim::transfer() {
    facesValid_ = false;
    bricksFresh_ = false;
    keysFresh_ = false;
    pipe_->runRecords(n_, rec_.devPtr(), U_.devPtr(), V_.devPtr(), W_.devPtr(), stream_);
    rec_.swapDevice(pipe_->sortedVec());
}

void FluidSim::transferAndPrepare() {
    const PrepKey key{cfg_.gravity * cfg_.dt, box_, hasSolid_, solidVersion_};
    const bool same = facesValid_ && key.gdt == prepKey_.gdt && key.hasSolid == prepKey_.hasSolid &&
                      key.solidVersion == prepKey_.solidVersion &&
                      std::memcmp(&key.box, &prepKey_.box, sizeof(key.box)) == 0;
    const ns::FacePrep prep{U0_.devPtr(), V0_.devPtr(), W0_.devPtr(), hasSolid_ ? solid_.devPtr() : nullptr, box_,
                            key.gdt, same ? 0 : 1};
    pipe_->runRecords(n_, rec_.devPtr(), U_.devPtr(), V_.devPtr(), W_.devPtr(), stream_, &prep, keysFresh_);
    keysFresh_ = false;
    rec_.swapDevice(pipe_->sortedVec());
    facesValid_ = true;
    prepKey_ = key;
    bricksFresh_ = true;
}

void FluidSim::prepareFaces() {
    facesValid_ = false;
    ns::launch_prepare_faces(U_.devPtr(), V_.devPtr(), W_.devPtr(), U0_.devPtr(), V0_.devPtr(), W0_.devPtr(),
                             hasSolid_ ? solid_.devPtr() : nullptr, cfg_.grid, cfg_.gravity * cfg_.dt, box_, stream_);
}

void FluidSim::buildSystem() {
    const uint8_t* solid = hasSolid_ ? solid_.devPtr() : nullptr;
    // after a fused transfer the pipeline holds the cell offsets of its
    // occupied bricks only (no dense per-cell counts)
    const bool sparse = pipe_->sparseCells();
    if (sparse)
        ns::launch_classify_bricks_sparse(pipe_->touchedList(), pipe_->touchedCount(), pipe_->occupancy(),
                                          pipe_->cellStart(), solid, numBricks_, brickFlag_.devPtr(), stream_);
    else
        ns::launch_classify_bricks(pipe_->counts(), solid, numBricks_, brickFlag_.devPtr(), stream_);
    scan_u32(brickFlag_.devPtr(), brickExcl_.devPtr(), numBricks_, true, ws_.devPtr(), ws_.size(), total_.devPtr(),
             stream_);
    uint32_t numActive = 0;
    FLIP2_HIP(hipMemcpyAsync(&numActive, total_.devPtr(), sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    synchronize();
    numActive_ = numActive;
    ensureSlots(std::max<uint32_t>(numActive, 1));
    ns::launch_build_slots(brickFlag_.devPtr(), brickExcl_.devPtr(), numBricks_, brickSlot_.devPtr(),
                           activeBricks_.devPtr(), stream_);
    ns::launch_brick_setup(numActive, activeBricks_.devPtr(), brickSlot_.devPtr(), sparse ? nullptr : pipe_->counts(),
                           solid, cfg_.grid, nbrSlot_.devPtr(), cellType_.devPtr(), stream_,
                           sparse ? pipe_->cellStart() : nullptr);
    const double hOverDt = (double)cfg_.grid.cellSize / (double)cfg_.dt;
    ns::launch_assemble(numActive, activeBricks_.devPtr(), nbrSlot_.devPtr(), cellType_.devPtr(), solid, cfg_.grid,
                        hOverDt, U_.devPtr(), V_.devPtr(), W_.devPtr(), coef_.devPtr(), b64_.devPtr(), reduce(),
                        scal_.devPtr() + 4, stream_);
}

SolveStats FluidSim::solve() { return solveWith(cfg_.preconditioner); }

// Mixed-precision PCG with iterative refinement: FP64 b, x and true
// residual; each refinement solves A e = r in FP32 (operator and
// preconditioner) with FP64 dot products, to a relative reduction that lets
// the FP64 residual reach the tolerance.
SolveStats FluidSim::solveWith(int prec) {
    if (prec == kPrecLearned && cfg_.learnedFallbackIterations > 0) {
        SolveStats st = solveOnce(prec, cfg_.learnedFallbackIterations);
        if (st.converged) return st;
        const double ms = st.solveMs;
        SolveStats fb = solveOnce(kPrecJacobi, cfg_.maxIterations);
        fb.usedFallback = true;
        fb.solveMs += ms;
        stats_ = fb;
        return fb;
    }
    return solveOnce(prec, cfg_.maxIterations);
}

SolveStats FluidSim::solveOnce(int prec, uint32_t capIn) {
    SolveStats st;
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t na = numActive_;
    st.activeBricks = na;
    const uint32_t cells = na * 512u;
    if (na == 0) {
        st.converged = true;
        stats_ = st;
        return st;
    }
    double sys[2];  // ||b||^2, fluid cells
    FLIP2_HIP(hipMemcpyAsync(sys, scal_.devPtr() + 4, sizeof(sys), hipMemcpyDeviceToHost, stream_));
    synchronize();
    const double bb = sys[0];
    st.fluidCells = (uint32_t)sys[1];
    st.bNorm = std::sqrt(bb);
    ns::launch_fill_f64(x64_.devPtr(), cells, 0.0, stream_);
    if (bb == 0.0) {
        st.converged = true;
        stats_ = st;
        return st;
    }
    FLIP2_HIP(hipMemcpyAsync(r64_.devPtr(), b64_.devPtr(), (size_t)cells * sizeof(double), hipMemcpyDeviceToDevice,
                             stream_));
    FLIP2_REQUIRE(prec != kPrecLearned || learned_, kErrState, "no learned preconditioner installed");
    learnedFresh_ = true;
    if (prec == kPrecMIC0) ns::launch_mic_factor(na, coef_.devPtr(), precon_.devPtr(), stream_);
    if (prec == kPrecLearned && cfg_.learnedSmoothing > 0.f && cfg_.learnedSmoother == 1)
        ns::launch_mic_factor(na, coef_.devPtr(), precon_.devPtr(), stream_, false);
    const bool jac = prec == kPrecJacobi;
    const bool flex = prec == kPrecLearned;
    // Flexible CG: direction of inner iteration j in ring slot j % slots
    // (one more slot than directions kept, so a new direction never
    // overwrites one it is orthogonalised against).
    const uint32_t keep = std::max(1u, std::min<uint32_t>(cfg_.learnedDirections, ns::kMaxFlexDirections));
    const uint32_t slots = keep + 1;
    const size_t cellsCap = (size_t)slotCapacity_ * 512u;
    if (flex) ensureRing(slots);
    auto ringP = [&](uint32_t j) { return pRing_.devPtr() + (j % slots) * cellsCap; };
    auto ringQ = [&](uint32_t j) { return qRing_.devPtr() + (j % slots) * cellsCap; };
    // the directions of iterations it, it-1, ... (at most `keep`)
    auto kept = [&](uint32_t it, double* Hb) {
        ns::FlexDirections d{};
        d.count = (int)std::min(keep, it + 1);
        for (int i = 0; i < d.count; ++i) {
            const uint32_t j = it - (uint32_t)i;
            d.p[i] = ringP(j);
            d.q[i] = ringQ(j);
            d.pq[i] = Hb + 4u * j + 2;
        }
        return d;
    };
    double rel = 1.0;
    uint32_t total = 0;
    const uint32_t cap = std::min(capIn, cfg_.maxIterations);
    const uint32_t every = std::max(1u, flex ? cfg_.learnedCheckEvery : cfg_.checkEvery);
    double* H = hist_.devPtr();
    // Flexible CG does not restart at a refinement: the recursive residual is
    // replaced by the true one (FP64, rounded to FP32), the correction so far
    // is added to x64, and the iteration continues with its kept directions
    // (their A-conjugacy does not depend on the residual).  itCarry: the
    // iteration count (ring slot and history index) it continues from.
    uint32_t itCarry = 0;
    for (uint32_t refine = 0; refine < 12 && total < cap; ++refine) {
        // Inner (FP32) target of this refinement: half of what the FP64
        // residual still has to lose, at least 2e-6 for the PCG (whose
        // restart discards its directions: one long first refinement, then
        // a short second, takes fewer iterations than stopping at 2e-5 --
        // C3 170 -> 159 per step) and 2e-5 for flexible CG (which keeps its
        // directions across refinements and took more iterations with 2e-6).
        const double innerTol = std::max(0.5 * cfg_.tolerance / rel, flex ? 2e-5 : 2e-6);
        const double ref = rel * st.bNorm;  // norm of the FP32 right-hand side of this refinement
        if (!flex) {
            // Jacobi / MIC(0): two kernels per iteration (direction + matvec,
            // update + preconditioner), p alternating between two buffers;
            // the refinement starts with r = (float) r64, e = 0, z = M^-1 r
            ns::CgFused cg{};
            cg.coef = coef_.devPtr();
            cg.nbrSlot = nbrSlot_.devPtr();
            cg.precon = jac ? nullptr : precon_.devPtr();
            cg.q = q_.devPtr();
            cg.e = e_.devPtr();
            cg.r = r_.devPtr();
            cg.z = z_.devPtr();
            cg.na = na;
            cg.partials = partials_.devPtr();
            FLIP2_REQUIRE(ns::cg_partials_needed(na) <= partials_.size(), kErrState, "PCG partials buffer too small");
            cg.hist = H;
            cg.stop = stop_.devPtr();
            const double thr = innerTol * ref;
            cg.thr2 = thr * thr;
            float* pbuf[2] = {p_.devPtr(), p2_.devPtr()};
            ns::launch_cg_update_precond(cg, 0, true, !jac, r64_.devPtr(), nullptr, stream_);
            if (refine == 0 && captureAt(0)) capture_(na, r_.devPtr(), stream_);
            // Batches of `every` iterations are queued one batch ahead of the
            // host's check of the previous one (read from pinned memory after
            // its event), so the GPU does not wait for the host; the first
            // entry with rr <= thr2 stops the refinement on the device, and
            // the kernels queued after it return at once.
            const uint32_t limit = cap - total;
            const size_t stride = 4u * ((size_t)every + 1u);
            double* pinned = pinnedHistory(2 * stride);
            struct Batch {
                uint32_t first, count;
                int slot;
            };
            uint32_t queued = 0;
            auto enqueue = [&](int slot) {
                const Batch b{queued, std::min(every, limit - queued), slot};
                for (uint32_t k = 0; k < b.count; ++k, ++queued) {
                    const uint32_t it = queued;
                    float* pNew = pbuf[it & 1u];
                    ns::launch_cg_direction_matvec(cg, it, it == 0, pbuf[(it + 1u) & 1u], pNew, stream_);
                    ns::launch_cg_update_precond(cg, it, false, !jac, nullptr, pNew, stream_);
                    if (refine == 0 && captureAt(it + 1)) capture_(na, r_.devPtr(), stream_);
                }
                ns::launch_cg_history(cg, queued, stream_);  // the batch's last entry
                // entries first .. first + count (the first one again: entry 0
                // of the refinement is checked with the first batch)
                FLIP2_HIP(hipMemcpyAsync(pinned + slot * stride, H + 4u * b.first, 4u * (b.count + 1u) * sizeof(double),
                                         hipMemcpyDeviceToHost, stream_));
                FLIP2_HIP(hipEventRecord(histEvent_[slot], stream_));
                return b;
            };
            uint32_t doneAt = UINT32_MAX;  // the converged entry
            RrTrend trend;
            auto convergedIn = [&](const Batch& b) {
                FLIP2_HIP(hipEventSynchronize(histEvent_[b.slot]));
                const double* hb = pinned + b.slot * stride;
                for (uint32_t k = 0; k <= b.count; ++k) {
                    const double rr = hb[4 * k + 1];
                    if (rr <= cg.thr2 || !(rr == rr)) {
                        doneAt = b.first + k;
                        return true;
                    }
                    trend.see(rr, b.first + k);
                }
                return false;
            };
            if (limit > 0) {
                Batch cur = enqueue(0);
                while (true) {
                    const bool haveNext = queued < limit;
                    if (haveNext && trend.likely(cur.first + cur.count, cg.thr2)) {
                        if (convergedIn(cur)) break;  // checked before more is queued
                        cur = enqueue(cur.slot ^ 1);
                        continue;
                    }
                    Batch next{};
                    if (haveNext) next = enqueue(cur.slot ^ 1);
                    if (convergedIn(cur) || !haveNext) break;  // converged, or the cap
                    cur = next;
                }
            }
            total += doneAt != UINT32_MAX ? doneAt : queued;
        } else {
            // flexible CG with the learned preconditioner.  Its kernels (the
            // apply's included) take the converged flag: the update's sum of
            // rr sets it when rr <= thr2, and every kernel queued after that
            // returns at once, so iterations are queued a batch ahead of the
            // host's check as in the PCG branch.  refine_start clears the flag.
            ns::Reduce fr = reduce();
            fr.stop = stop_.devPtr();
            const double thr = innerTol * ref;
            fr.thr2 = thr * thr;
            ns::launch_refine_start(na, r64_.devPtr(), r_.devPtr(), e_.devPtr(), coef_.devPtr(), reduce(),
                                    scal_.devPtr() + 1, stream_, stop_.devPtr());
            if (refine == 0 && captureAt(0)) capture_(na, r_.devPtr(), stream_);
            if (itCarry == 0) {
                // p_0 = z_0; H[0] = p_0 . r_0 (the step-length numerator)
                applyLearned(na);
                ns::FlexDirections none{};
                ns::launch_flex_direction(na, z_.devPtr(), none, zq_.devPtr(), r_.devPtr(), ringP(0), fr, H, stream_);
            } else {
                // the apply and direction of iteration itCarry, on the replaced
                // residual: p = z - sum_i (z . q_i / p_i . q_i) p_i
                applyLearned(na);
                const ns::FlexDirections d = kept(itCarry - 1, H);
                ns::launch_flex_dots(na, z_.devPtr(), d, fr, zq_.devPtr(), stream_);
                ns::launch_flex_direction(na, z_.devPtr(), d, zq_.devPtr(), r_.devPtr(), ringP(itCarry), fr,
                                          H + 4u * itCarry, stream_);
            }
            const uint32_t it0 = itCarry;
            const uint32_t limit = cap - total;
            const size_t stride = 4u * (size_t)every;
            double* pinned = pinnedHistory(2 * stride);
            struct Batch {
                uint32_t first, count;
                int slot;
            };
            uint32_t queued = 0;  // iterations of this refinement queued so far
            auto enqueue = [&](int slot) {
                const Batch b{queued, std::min(every, limit - queued), slot};
                for (uint32_t k = 0; k < b.count; ++k) {
                    const uint32_t it = it0 + queued;
                    double* hk = H + 4u * it;
                    double* hn = H + 4u * (it + 1);
                    ns::launch_matvec_dot(na, ringP(it), coef_.devPtr(), nbrSlot_.devPtr(), ringQ(it), fr, hk + 2,
                                          stream_);
                    ns::launch_update(na, hk, hk + 2, ringP(it), ringQ(it), coef_.devPtr(), e_.devPtr(), r_.devPtr(),
                                      z_.devPtr(), false, fr, hn, stream_);
                    ++queued;
                    if (refine == 0 && captureAt(it + 1)) capture_(na, r_.devPtr(), stream_);
                    if (queued == limit) break;  // no direction after the cap's last iteration
                    // p = z - sum_i (z . q_i / p_i . q_i) p_i over the kept
                    // directions; hn[0] = p . r
                    applyLearned(na);
                    const ns::FlexDirections d = kept(it, H);
                    ns::launch_flex_dots(na, z_.devPtr(), d, fr, zq_.devPtr(), stream_);
                    ns::launch_flex_direction(na, z_.devPtr(), d, zq_.devPtr(), r_.devPtr(), ringP(it + 1), fr, hn,
                                              stream_);
                }
                // the batch's residual entries it0 + first + 1 .. it0 + first + count
                FLIP2_HIP(hipMemcpyAsync(pinned + slot * stride, H + 4u * (it0 + b.first + 1u),
                                         4u * b.count * sizeof(double), hipMemcpyDeviceToHost, stream_));
                FLIP2_HIP(hipEventRecord(histEvent_[slot], stream_));
                return b;
            };
            uint32_t done = UINT32_MAX;  // iterations up to the converged entry
            RrTrend trend;
            auto convergedIn = [&](const Batch& b) {
                FLIP2_HIP(hipEventSynchronize(histEvent_[b.slot]));
                const double* hb = pinned + b.slot * stride;
                for (uint32_t k = 0; k < b.count; ++k) {
                    const double rr = hb[4 * k + 1];
                    if (rr <= fr.thr2 || !(rr == rr)) {
                        done = b.first + k + 1;
                        return true;
                    }
                    trend.see(rr, b.first + k + 1);
                }
                return false;
            };
            if (limit > 0) {
                Batch cur = enqueue(0);
                while (true) {
                    const bool haveNext = queued < limit;
                    if (haveNext && trend.likely(cur.first + cur.count, fr.thr2)) {
                        if (convergedIn(cur)) break;  // checked before more is queued
                        cur = enqueue(cur.slot ^ 1);
                        continue;
                    }
                    Batch next{};
                    if (haveNext) next = enqueue(cur.slot ^ 1);
                    if (convergedIn(cur) || !haveNext) break;  // converged, or the cap
                    cur = next;
                }
            }
            if (done == UINT32_MAX) done = queued;
            total += done;
            itCarry = it0 + done;
        }
        ns::launch_accumulate(na, e_.devPtr(), x64_.devPtr(), coef_.devPtr(), stream_);
        ns::launch_true_residual(na, b64_.devPtr(), x64_.devPtr(), coef_.devPtr(), nbrSlot_.devPtr(), r64_.devPtr(),
                                 reduce(), scal_.devPtr() + 2, stream_);
        const double rr = readScalar(scal_.devPtr() + 2);
        rel = std::sqrt(std::max(rr, 0.0)) / st.bNorm;
        st.refinements = refine + 1;
        if (rel <= cfg_.tolerance) {
            st.converged = true;
            break;
        }
        if (!(rel == rel)) break;
    }
    st.iterations = total;
    st.relResidual = rel;
    st.solveMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats_ = st;
    return st;
}

void FluidSim::project() {
    ns::launch_velocity_update(numActive_, activeBricks_.devPtr(), nbrSlot_.devPtr(), cellType_.devPtr(),
                               hasSolid_ ? solid_.devPtr() : nullptr, cfg_.grid, x64_.devPtr(),
                               cfg_.dt / cfg_.grid.cellSize, U_.devPtr(), V_.devPtr(), W_.devPtr(), stream_);
}

void FluidSim::gridToParticles() {
    ns::launch_g2p_flip(n_, rec_.devPtr(), U_.devPtr(), V_.devPtr(), W_.devPtr(), U0_.devPtr(), V0_.devPtr(),
                        W0_.devPtr(), cfg_.grid, cfg_.flip, stream_);
}

void FluidSim::advect() {
    bricksFresh_ = false;
    keysFresh_ = false;
    ns::launch_advect_rk2(n_, rec_.devPtr(), U_.devPtr(), V_.devPtr(), W_.devPtr(),
                          hasSolid_ ? solid_.devPtr() : nullptr, cfg_.grid, cfg_.dt, cfg_.margin * cfg_.grid.cellSize,
                          box_, stream_);
}

void FluidSim::gridToParticlesAndAdvect() {
    const uint8_t* solid = hasSolid_ ? solid_.devPtr() : nullptr;
    keysFresh_ = bricksFresh_;  // the brick kernel also writes the next transfer's keys
    if (bricksFresh_)
        ns::launch_g2p_advect_bricks(pipe_->touchedList(), pipe_->touchedCount(), pipe_->listCounters(),
                                     pipe_->brickStart(), rec_.devPtr(),
                                     U_.devPtr(), V_.devPtr(), W_.devPtr(), U0_.devPtr(), V0_.devPtr(), W0_.devPtr(),
                                     solid, cfg_.grid, cfg_.flip, cfg_.dt, cfg_.margin * cfg_.grid.cellSize, box_,
                                     stream_, pipe_->keys());
    else
        ns::launch_g2p_advect(n_, rec_.devPtr(), U_.devPtr(), V_.devPtr(), W_.devPtr(), U0_.devPtr(), V0_.devPtr(),
                              W0_.devPtr(), solid, cfg_.grid, cfg_.flip, cfg_.dt, cfg_.margin * cfg_.grid.cellSize,
                              box_, stream_);
    bricksFresh_ = false;  // the particles moved
}

void FluidSim::step() {
    transferAndPrepare();
    buildSystem();
    solve();
    project();
    gridToParticlesAndAdvect();
}

uint32_t FluidSim::stepFrame(double frameDt, double cfl, uint32_t maxSubsteps) {
    FLIP2_REQUIRE(frameDt > 0.0 && cfl > 0.0 && maxSubsteps >= 1, kErrArgument,
                  "step_frame: frame dt and cfl must be positive, at least one substep");
    const float saved = cfg_.dt;
    const double h = (double)cfg_.grid.cellSize;
    double t = 0.0;
    uint32_t k = 0;
    try {
        while (frameDt - t > 1e-12 * frameDt && k < maxSubsteps) {
            transfer();
            uint32_t bits = 0;
            FLIP2_HIP(hipMemcpyAsync(&bits, pipe_->maxSpeedBits(), sizeof(bits), hipMemcpyDeviceToHost, stream_));
            synchronize();
            float vmax;
            std::memcpy(&vmax, &bits, sizeof(vmax));
            FLIP2_REQUIRE(std::isfinite(vmax), kErrState, "step_frame: a particle velocity is not finite");
            double dt = frameDt - t;
            if (vmax > 0.f && k + 1 < maxSubsteps) dt = std::min(dt, cfl * h / (double)vmax);
            cfg_.dt = (float)dt;
            prepareFaces();
            buildSystem();
            solve();
            project();
            gridToParticlesAndAdvect();
            t += dt;
            ++k;
        }
    } catch (...) {
        cfg_.dt = saved;
        throw;
    }
    cfg_.dt = saved;
    return k;
}

size_t FluidSim::getArray(const std::string& name, void* host, size_t capacity) {
    synchronize();
    auto out = [&](auto& v, uint32_t count) -> size_t {
        using T = typename std::remove_reference<decltype(*v.devPtr())>::type;
        if (host) {
            FLIP2_REQUIRE(capacity >= count, kErrRange, "buffer too small");
            v.downloadTo(reinterpret_cast<T*>(host), count, stream_);
            synchronize();
        }
        return count;
    };
    const uint32_t cells = numActive_ * 512u;
    if (name == "U") return out(U_, numCells_);
    if (name == "V") return out(V_, numCells_);
    if (name == "W") return out(W_, numCells_);
    if (name == "U0") return out(U0_, numCells_);
    if (name == "V0") return out(V0_, numCells_);
    if (name == "W0") return out(W0_, numCells_);
    if (name == "activeBricks") return out(activeBricks_, numActive_);
    if (name == "brickSlot") return out(brickSlot_, numBricks_);
    if (name == "cellType") return out(cellType_, cells);
    if (name == "coef") return out(coef_, cells);
    if (name == "b") return out(b64_, cells);
    if (name == "p") return out(x64_, cells);
    if (name == "counts" || name == "cellStart") return pipe_->get(name, host, capacity, stream_);
    throw Error(kErrArgument, "unknown array " + name);
}

}  // namespace flip2

// North-star FLIP fluid step on 8^3 bricks (FP32 particle state):
//   transfer (bin -> sort -> gather -> counts -> scan -> P2G)
//   -> wall faces + gravity (pre-projection field kept for FLIP)
//   -> active bricks, operator and FP64 right-hand side
//   -> mixed-precision PCG (Jacobi, brick MIC(0) or a learned preconditioner)
//   -> pressure-gradient update -> FLIP G2P -> RK2 advection.
#pragma once

#include <string>
#include "common.hpp"
#include "device.hpp"
#include "ns_solver.hpp"
#include "p2g_pipeline.hpp"

#include <functional>
#include <memory>
#include <vector>

namespace flip2 {

enum Preconditioner : int { kPrecJacobi = 0, kPrecMIC0 = 1, kPrecLearned = 2 };

struct FluidConfig {
    Grid grid;
    float dt = 1.0f / 240.0f;
    float gravity = -9.81f;
    float flip = 0.95f;
    int preconditioner = kPrecJacobi;
    double tolerance = 1e-6;       // ||b - A p||_2 / ||b||_2, FP64
    uint32_t maxIterations = 4000; // inner iterations over all refinements
    uint32_t checkEvery = 8;       // iterations between host residual checks
    float margin = 1e-3f;          // particles are kept this far (x h) inside the box
    // Learned preconditioner: fall back to Jacobi-PCG when flexible CG needs
    // more iterations than this (0: no fallback).
    uint32_t learnedFallbackIterations = 0;
    // Damped-Jacobi weight of the smoothing step before and after the
    // learned correction (0: the network alone is the preconditioner).
    float learnedSmoothing = 0.0f;
    // The smoother around the learned correction when learnedSmoothing > 0:
    // 0 damped Jacobi (weight learnedSmoothing), 1 brick-local IC(0) (fewer
    // flexible-CG iterations: 14.8 against 16.4 at 128^3, 68.5 against
    // 103.5 at 256^3 with the committed network).
    int learnedSmoother = 1;
    // IC(0) smoothing steps before and after the network (1 or 2).
    uint32_t learnedSmoothSteps = 1;
    // Flexible CG keeps this many earlier directions for the learned
    // preconditioner (1 = FCG(1); at most ns::kMaxFlexDirections).
    uint32_t learnedDirections = 8;
    // Iterations between host residual checks in flexible CG (its
    // iterations are costly: checking after each one stops at the first
    // that converges, where checkEvery's batches would overshoot).
    uint32_t learnedCheckEvery = 1;
    // Replay the learned preconditioner's apply from a HIP graph captured
    // once per solve.
    bool learnedGraph = true;
};

struct SolveStats {
    uint32_t iterations = 0;       // inner PCG iterations (all refinements)
    uint32_t refinements = 0;      // outer iterative-refinement steps
    double relResidual = 0.0;      // final FP64 ||b - A p|| / ||b||
    double bNorm = 0.0;
    uint32_t activeBricks = 0;
    uint32_t fluidCells = 0;
    double solveMs = 0.0;          // wall time of the solve (host clock, synchronised)
    bool converged = false;
    bool usedFallback = false;
};

// A preconditioner supplied from outside the solver (the learned one):
// z = M^-1 r on the active-brick vectors, FP32.  newSystem is set on the
// first apply of a solve (the fluid cells may have changed); the other
// applies of a solve launch the same kernels on the same buffers, so the
// solver may replay them from a HIP graph.  normPairs: null, or per slot
// (sum of r^2, count) over the fluid cells, already computed by the solver.
// stop: the solver's converged flag; while it is set the apply's kernels
// return at once (the solver queues iterations ahead of its convergence
// check).
using PrecondFn = std::function<void(uint32_t numActive, const float* r, float* z, bool newSystem,
                                     const double* normPairs, hipStream_t s, const uint32_t* stop)>;
// Receives the residual r_k = b - A x_k (FP32, active-brick layout) of a
// solve's first refinement after k inner iterations (training data).
using CaptureFn = std::function<void(uint32_t numActive, const float* r, hipStream_t s)>;

class FluidSim {
public:
    FluidSim(uint32_t maxParticles, const FluidConfig& cfg, int device = 0);
    ~FluidSim();
    FluidSim(const FluidSim&) = delete;
    FluidSim& operator=(const FluidSim&) = delete;

    void setParticles(uint32_t n, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                      const float* vz);
    void getParticles(float* x, float* y, float* z, float* vx, float* vy, float* vz);
    // Static solid cells, host array of n^3 bytes indexed x + y*nx + z*nx*ny.
    void setSolidMask(const uint8_t* mask);
    // A moving solid box (cells whose centre lies in [lo, hi)), its faces
    // carrying velocity vel; replaces the previous box.  Null lo removes it.
    void setSolidBox(const float* lo, const float* hi, const float* vel);
    void setConfig(const FluidConfig& cfg);
    // Binary snapshot of the simulation state (grid, physical settings,
    // particle records in their current order, static solid mask, solid
    // box) for resuming a run: loading it into a simulation of the same grid
    // and stepping gives the same particles, bit for bit, as stepping the
    // saved one.  The learned preconditioner's settings are not part of it.
    void saveSnapshot(const std::string& path);
    void loadSnapshot(const std::string& path);
    const FluidConfig& config() const { return cfg_; }
    const float4* recordsDev() const { return rec_.devPtr(); }

    void step();                   // one full step
    // Advances by frameDt in substeps whose dt keeps every particle within
    // cfl cells per step on each axis: dt = min(remaining, cfl h / max|v_c|)
    // with the exact (unsampled) maximum over all particles' velocity
    // components at the start of the substep; the maxSubsteps-th substep
    // takes whatever time remains.  Returns the number of substeps.
    uint32_t stepFrame(double frameDt, double cfl, uint32_t maxSubsteps);
    // Stages (the step is these in order)
    void transfer();
    void prepareFaces();
    void transferAndPrepare();  // transfer then prepareFaces, fused (the same results)
    void buildSystem();
    SolveStats solve();
    void project();
    void gridToParticles();
    void advect();
    void gridToParticlesAndAdvect();  // gridToParticles then advect, one kernel

    // Solve A x = b for the current system with a caller-given preconditioner
    // choice (used for preconditioner comparisons on one system).
    SolveStats solveWith(int preconditioner);
    SolveStats solveOnce(int preconditioner, uint32_t cap);
    void setLearnedPreconditioner(PrecondFn fn) {
        dropLearnedGraph();
        learnedFresh_ = true;
        learned_ = std::move(fn);
    }
    // Residuals after the listed inner iteration counts (0 = b) of every
    // following solve go to fn; an empty list stops the capture.
    void setResidualCapture(std::vector<uint32_t> iterations, CaptureFn fn) {
        captureIts_ = std::move(iterations);
        capture_ = std::move(fn);
    }

    const SolveStats& lastSolve() const { return stats_; }
    uint32_t numParticles() const { return n_; }
    uint32_t numActive() const { return numActive_; }
    hipStream_t stream() const { return stream_; }
    void synchronize();

    // Named arrays for checks: "U","V","W","U0","V0","W0" (n^3 f32),
    // "counts","cellStart" (n^3+1 u32), "activeBricks" (u32), "cellType" (u8,
    // slots*512), "coef" (u16), "b" and "p" (f64, slots*512), "brickSlot" (i32).
    size_t getArray(const std::string& name, void* host, size_t capacity);
    // Active-brick vectors for external preconditioners / training data.
    const uint16_t* coefDev() const { return coef_.devPtr(); }
    const int32_t* nbrSlotDev() const { return nbrSlot_.devPtr(); }
    const uint32_t* activeBricksDev() const { return activeBricks_.devPtr(); }
    const uint8_t* cellTypeDev() const { return cellType_.devPtr(); }
    const int32_t* brickSlotDev() const { return brickSlot_.devPtr(); }
    const double* rhsDev() const { return b64_.devPtr(); }
    const double* solutionDev() const { return x64_.devPtr(); }
    int gridN() const { return (int)cfg_.grid.sizeX; }

private:
    void ensureSlots(uint32_t numActive);
    ns::Reduce reduce() { return ns::Reduce{partials_.devPtr()}; }
    double readScalar(const double* dev);

    FluidConfig cfg_;
    DeviceManager dm_;  // the pool of every device array below (declared first: outlives them)
    hipStream_t stream_ = nullptr;
    uint32_t maxParticles_, n_ = 0, numCells_, numBricks_;
    uint32_t numActive_ = 0;
    std::unique_ptr<P2GPipeline> pipe_;
    DeviceVec<float4> rec_{dm_};  // particle records (2 x float4 per particle)
    DeviceVec<float> U_{dm_}, V_{dm_}, W_{dm_}, U0_{dm_}, V0_{dm_}, W0_{dm_};
    DeviceVec<uint8_t> solid_{dm_}, solidStatic_{dm_};  // combined mask, static part
    bool hasSolid_ = false, hasStatic_ = false;
    ns::SolidBox box_{};
    void rebuildSolid();
    // active bricks
    DeviceVec<uint32_t> brickFlag_{dm_}, brickExcl_{dm_}, activeBricks_{dm_};
    DeviceVec<int32_t> brickSlot_{dm_}, nbrSlot_{dm_};
    DeviceVec<uint8_t> cellType_{dm_};
    DeviceVec<uint16_t> coef_{dm_};
    DeviceVec<double> b64_{dm_}, x64_{dm_}, r64_{dm_};
    DeviceVec<float> r_{dm_}, z_{dm_}, p_{dm_}, p2_{dm_}, q_{dm_}, e_{dm_}, precon_{dm_}, aux0_{dm_}, aux1_{dm_},
        aux2_{dm_}, aux3_{dm_};
    DeviceVec<float> pRing_{dm_}, qRing_{dm_};  // flexible CG directions, learnedDirections + 1 slots
    DeviceVec<double> zq_{dm_};
    uint32_t ringSlots_ = 0;
    void ensureRing(uint32_t slots);
    void applyLearned(uint32_t na);
    void applyLearnedEager(uint32_t na, bool newSystem);
    void dropLearnedGraph();
    bool learnedFresh_ = true;  // the next learned apply starts a new system
    hipGraph_t learnedGraph_ = nullptr;
    hipGraphExec_t learnedGraphExec_ = nullptr;
    DeviceVec<double> partials_{dm_}, hist_{dm_}, scal_{dm_};
    DeviceVec<uint32_t> stop_{dm_};  // the fused PCG's convergence flag
    // The faces of bricks untouched by the last fused transfer hold the
    // prepared zero-transfer values of prepKey_ (so the next fused transfer
    // rewrites only the bricks that left the touched set).  Cleared by every
    // other writer of the face arrays and by a change of solids.
    struct PrepKey {
        float gdt;
        ns::SolidBox box;
        bool hasSolid;
        uint64_t solidVersion;
    };
    bool facesValid_ = false;
    // the records are sorted by brick and the pipeline's touched list and
    // cell offsets describe them (set by the fused transfer, cleared when the
    // particles move or another transfer runs)
    bool bricksFresh_ = false;
    // the pipeline's key array holds the cell key of every record (written
    // by the brick G2P + RK2 kernel; cleared when the records change
    // otherwise), so the next fused transfer skips its binning pass
    bool keysFresh_ = false;
    PrepKey prepKey_{};
    uint64_t solidVersion_ = 0;
    // pinned host copies of two batches' history entries, and their events
    double* histPinned_ = nullptr;
    size_t histPinnedDoubles_ = 0;
    hipEvent_t histEvent_[2] = {nullptr, nullptr};
    double* pinnedHistory(size_t doubles);
    DeviceVec<uint32_t> total_{dm_};
    DeviceVec<uint8_t> ws_{dm_};
    uint32_t slotCapacity_ = 0;
    SolveStats stats_;
    PrecondFn learned_;
    CaptureFn capture_;
    std::vector<uint32_t> captureIts_;
    bool captureAt(uint32_t k) const {
        if (!capture_) return false;
        for (uint32_t v : captureIts_)
            if (v == k) return true;
        return false;
    }
};

}  // namespace flip2

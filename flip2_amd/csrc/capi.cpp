// C ABI of libparticles.so (include/flip2.h): every entry point catches
// exceptions, records the message per thread and returns a status code.
#include "../../include/flip2.h"

#include <cstring>
#include <functional>
#include <iterator>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fluid.hpp"
#include "lists.hpp"
#include "unet.hpp"
#include "ns_kernels.hpp"
#include "p2g_pipeline.hpp"
#include "particles.hpp"
#include "rc_kernels.hpp"
#include "scan_impl.hpp"
#include "sort.hpp"

struct flip2_p2g_pipeline {
    flip2::P2GPipeline impl;
    flip2_p2g_pipeline(uint32_t n, const flip2::Grid& g) : impl(n, g) {}
};

struct flip2_fluid {
    flip2::FluidSim impl;
    flip2::KernelTimer timer;  // per-kernel event timing of the calls below (flip2_fluid_enable_timing)
    bool timing = false;
    flip2_fluid(uint32_t n, const flip2::FluidConfig& c) : impl(n, c) {}
    flip2::KernelTimer* activeTimer() { return timing ? &timer : nullptr; }
};

struct flip2_unet {
    flip2::UNet net;
    float omega = 2.0f / 3.0f;  // damped-Jacobi smoothing around the network (0: none)
    struct Frame {  // a dense right-hand side (b or a solver residual) and its operator
        flip2::DeviceVec<float> b;
        flip2::DeviceVec<uint16_t> coef;
    };
    std::vector<std::unique_ptr<Frame>> frames;
    Frame& newFrame(hipStream_t st = nullptr) {
        const uint32_t nv = (uint32_t)n() * n() * n();
        std::unique_ptr<Frame> fr(new Frame);
        fr->b.resizeAsync(nv, st);
        fr->coef.resizeAsync(nv, st);
        FLIP2_HIP(hipMemsetAsync(fr->b.devPtr(), 0, fr->b.bytes(), st));
        FLIP2_HIP(hipMemsetAsync(fr->coef.devPtr(), 0, fr->coef.bytes(), st));
        frames.push_back(std::move(fr));
        return *frames.back();
    }
    flip2::DeviceVec<float> res, dOut, bTmp, r1, zS, gS;
    flip2::DeviceVec<uint16_t> coefTmp;
    flip2::DeviceVec<double> red;
    explicit flip2_unet(int n) : net(n, nullptr) {
        const uint32_t nv = (uint32_t)n * n * n;
        for (auto* v : {&res, &dOut, &r1, &zS, &gS}) v->resizeAsync(nv, nullptr);
        red.resizeAsync(8, nullptr);
    }
    int n() const { return net.gridSize(); }
    // z = M(b) on the dense grid: the network alone or wrapped in smoothing.
    const float* apply(const float* bD, const uint16_t* coefD, bool keepActivations = false) {
        const int nn = n();
        const float* in = bD;
        if (omega > 0.f) {
            flip2::nn::launch_dense_presmooth(coefD, bD, omega, r1.devPtr(), nn, nullptr);
            in = r1.devPtr();
        }
        flip2::nn::launch_dense_input(in, coefD, nn, net.input(), net.inputCompact(), net.mask(), net.scale(),
                                      red.devPtr() + 2, nullptr);
        net.setInputMasks(nullptr, !keepActivations);
        net.forward(nullptr, keepActivations);
        if (omega <= 0.f) return net.output();
        flip2::nn::launch_dense_postsmooth(coefD, bD, net.output(), omega, zS.devPtr(), nn, nullptr);
        return zS.devPtr();
    }
    // loss ||A M(b) - b||^2 / ||b||^2 of one dense system; parameter
    // gradients accumulated when `grad`
    double lossGrad(const float* bD, const uint16_t* coefD, bool grad) {
        const int nn = n();
        const float* z = apply(bD, coefD, grad);
        flip2::nn::launch_residual(coefD, z, bD, res.devPtr(), nn, red.devPtr(), nullptr);
        if (grad) {
            if (omega > 0.f) {
                flip2::nn::launch_loss_grad(coefD, res.devPtr(), red.devPtr(), gS.devPtr(), nn, nullptr);
                flip2::nn::launch_dense_presmooth(coefD, gS.devPtr(), omega, dOut.devPtr(), nn, nullptr);
            } else {
                flip2::nn::launch_loss_grad(coefD, res.devPtr(), red.devPtr(), dOut.devPtr(), nn, nullptr);
            }
            net.backward(dOut.devPtr(), nullptr);
        }
        double h[2];
        FLIP2_HIP(hipMemcpy(h, red.devPtr(), sizeof(h), hipMemcpyDeviceToHost));
        return h[1] > 0.0 ? h[0] / h[1] : 0.0;
    }
};

struct flip2_particles {
    flip2::Particles impl;
    explicit flip2_particles(uint32_t n) : impl(n) {}
};

namespace {

thread_local std::string g_lastError;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return FLIP2_OK;
    } catch (const flip2::Error& e) {
        g_lastError = e.what();
        return e.code();
    } catch (const std::bad_alloc&) {
        g_lastError = "host allocation failed";
        return FLIP2_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return FLIP2_ERR_INTERNAL;
    } catch (...) {
        g_lastError = "unknown error";
        return FLIP2_ERR_INTERNAL;
    }
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

flip2::Grid to_grid(const flip2_grid& g) {
    flip2::Grid r;
    r.negX = g.negX; r.negY = g.negY; r.negZ = g.negZ;
    r.sizeX = g.sizeX; r.sizeY = g.sizeY; r.sizeZ = g.sizeZ;
    r.cellSize = g.cellSize;
    return r;
}

// Library-owned workspace for calls that pass d_ws == NULL.
struct InternalWorkspace {
    std::mutex m;
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need) {
        if (need > bytes) {
            if (ptr) FLIP2_HIP(hipFree(ptr));
            ptr = nullptr;
            bytes = 0;
            FLIP2_HIP(hipMalloc(&ptr, need));
            bytes = need;
        }
        return ptr;
    }
};
InternalWorkspace g_ws;

flip2::FluidConfig to_config(const flip2_fluid_config& c) {
    flip2::FluidConfig r;
    r.grid = to_grid(c.grid);
    r.dt = c.dt;
    r.gravity = c.gravity;
    r.flip = c.flip;
    r.preconditioner = c.preconditioner;
    r.tolerance = c.tolerance;
    r.maxIterations = c.max_iterations;
    r.checkEvery = c.check_every ? c.check_every : 8;
    r.margin = c.margin;
    // the simulation validates every field (FluidSim's constructor and setConfig)
    return r;
}

void from_config(const flip2::FluidConfig& c, flip2_fluid_config* o) {
    o->grid.negX = c.grid.negX; o->grid.negY = c.grid.negY; o->grid.negZ = c.grid.negZ;
    o->grid.sizeX = c.grid.sizeX; o->grid.sizeY = c.grid.sizeY; o->grid.sizeZ = c.grid.sizeZ;
    o->grid.cellSize = c.grid.cellSize;
    o->dt = c.dt;
    o->gravity = c.gravity;
    o->flip = c.flip;
    o->preconditioner = c.preconditioner;
    o->tolerance = c.tolerance;
    o->max_iterations = c.maxIterations;
    o->check_every = c.checkEvery;
    o->margin = c.margin;
}

void to_stats(const flip2::SolveStats& s, flip2_solve_stats* o) {
    o->iterations = s.iterations;
    o->refinements = s.refinements;
    o->rel_residual = s.relResidual;
    o->b_norm = s.bNorm;
    o->active_bricks = s.activeBricks;
    o->fluid_cells = s.fluidCells;
    o->solve_ms = s.solveMs;
    o->converged = s.converged ? 1 : 0;
    o->used_fallback = s.usedFallback ? 1 : 0;
}

#define NEED(p) FLIP2_REQUIRE((p) != nullptr, flip2::kErrArgument, #p " is NULL")

}  // namespace

namespace flip2 {
// The C ABI's error convention (status code, message for flip2_last_error)
// for entry points defined in other translation units.
int capi_guarded(const std::function<void()>& f) { return guarded(f); }
}  // namespace flip2

extern "C" {

const char* flip2_last_error(void) { return g_lastError.c_str(); }
int flip2_version(void) { return 10000; }

int flip2_device_count(int* count) {
    return guarded([&] {
        NEED(count);
        FLIP2_HIP(hipGetDeviceCount(count));
    });
}

int flip2_device_info(int device, char* name, size_t nameLen, uint64_t* totalBytes) {
    return guarded([&] {
        hipDeviceProp_t prop;
        FLIP2_HIP(hipGetDeviceProperties(&prop, device));
        if (name && nameLen) {
            std::strncpy(name, prop.name, nameLen - 1);
            name[nameLen - 1] = 0;
        }
        if (totalBytes) *totalBytes = prop.totalGlobalMem;
    });
}

int flip2_malloc(void** d_ptr, size_t bytes) {
    return guarded([&] {
        NEED(d_ptr);
        FLIP2_HIP(hipMalloc(d_ptr, bytes ? bytes : 1));
    });
}
int flip2_free(void* d_ptr) {
    return guarded([&] { FLIP2_HIP(hipFree(d_ptr)); });
}
int flip2_memcpy_htod(void* d_dst, const void* src, size_t bytes) {
    return guarded([&] { if (bytes) FLIP2_HIP(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice)); });
}
int flip2_memcpy_dtoh(void* dst, const void* d_src, size_t bytes) {
    return guarded([&] { if (bytes) FLIP2_HIP(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost)); });
}
int flip2_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes) {
    return guarded([&] { if (bytes) FLIP2_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice)); });
}
int flip2_copy(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        FLIP2_REQUIRE(bytes == 0 || (dst && src), flip2::kErrArgument, "null buffer");
        flip2::launch_copy(dst, src, bytes, as_stream(stream));
    });
}
int flip2_memset(void* d_ptr, int value, size_t bytes) {
    return guarded([&] { if (bytes) FLIP2_HIP(hipMemset(d_ptr, value, bytes)); });
}
int flip2_stream_create(void** stream) {
    return guarded([&] {
        NEED(stream);
        hipStream_t s;
        FLIP2_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
int flip2_stream_destroy(void* stream) {
    return guarded([&] { FLIP2_HIP(hipStreamDestroy(as_stream(stream))); });
}
int flip2_stream_synchronize(void* stream) {
    return guarded([&] { FLIP2_HIP(hipStreamSynchronize(as_stream(stream))); });
}

struct Timer {
    hipEvent_t a, b;
};
int flip2_timer_create(void** timer) {
    return guarded([&] {
        NEED(timer);
        Timer* t = new Timer;
        FLIP2_HIP(hipEventCreate(&t->a));
        FLIP2_HIP(hipEventCreate(&t->b));
        *timer = t;
    });
}
int flip2_timer_destroy(void* timer) {
    return guarded([&] {
        Timer* t = reinterpret_cast<Timer*>(timer);
        if (!t) return;
        (void)hipEventDestroy(t->a);
        (void)hipEventDestroy(t->b);
        delete t;
    });
}
int flip2_timer_start(void* timer, void* stream) {
    return guarded([&] {
        NEED(timer);
        FLIP2_HIP(hipEventRecord(reinterpret_cast<Timer*>(timer)->a, as_stream(stream)));
    });
}
int flip2_timer_stop(void* timer, void* stream) {
    return guarded([&] {
        NEED(timer);
        FLIP2_HIP(hipEventRecord(reinterpret_cast<Timer*>(timer)->b, as_stream(stream)));
    });
}
int flip2_timer_elapsed_ms(void* timer, float* ms) {
    return guarded([&] {
        NEED(timer);
        NEED(ms);
        Timer* t = reinterpret_cast<Timer*>(timer);
        FLIP2_HIP(hipEventSynchronize(t->b));
        FLIP2_HIP(hipEventElapsedTime(ms, t->a, t->b));
    });
}

size_t flip2_scan_workspace_bytes(uint32_t n) { return flip2::scan_workspace_bytes(n); }

int flip2_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint32_t n, int exclusive, uint32_t* d_total, void* d_ws,
                   size_t ws_bytes, void* stream) {
    return guarded([&] {
        if (!d_ws) {
            std::lock_guard<std::mutex> lk(g_ws.m);
            const size_t need = flip2::scan_workspace_bytes(n);
            void* w = g_ws.get(need);
            flip2::scan_u32(d_in, d_out, n, exclusive != 0, w, need, d_total, as_stream(stream));
            FLIP2_HIP(hipStreamSynchronize(as_stream(stream)));
        } else {
            flip2::scan_u32(d_in, d_out, n, exclusive != 0, d_ws, ws_bytes, d_total, as_stream(stream));
        }
    });
}

size_t flip2_sort_workspace_bytes(uint32_t n, uint32_t begin_bit, uint32_t end_bit) {
    return flip2::sort_workspace_bytes(n, begin_bit, end_bit);
}

int flip2_sort_pairs_u32(const uint32_t* d_keys_in, const uint32_t* d_vals_in, uint32_t* d_keys_out,
                         uint32_t* d_vals_out, uint32_t* d_keys_alt, uint32_t* d_vals_alt, uint32_t n,
                         uint32_t begin_bit, uint32_t end_bit, void* d_ws, size_t ws_bytes, void* stream) {
    return guarded([&] {
        if (!d_ws) {
            std::lock_guard<std::mutex> lk(g_ws.m);
            const size_t need = flip2::sort_workspace_bytes(n, begin_bit, end_bit);
            void* w = g_ws.get(need);
            flip2::sort_pairs_u32(d_keys_in, d_vals_in, d_keys_out, d_vals_out, d_keys_alt, d_vals_alt, n, begin_bit,
                                  end_bit, w, need, as_stream(stream));
            FLIP2_HIP(hipStreamSynchronize(as_stream(stream)));
        } else {
            flip2::sort_pairs_u32(d_keys_in, d_vals_in, d_keys_out, d_vals_out, d_keys_alt, d_vals_alt, n, begin_bit,
                                  end_bit, d_ws, ws_bytes, as_stream(stream));
        }
    });
}

int flip2_find_sub_cell(const double* d_px, const double* d_py, const double* d_pz, uint32_t n, flip2_grid grid,
                        const uint32_t* d_keys, uint32_t refinement_level, double radius, uint32_t* d_touched_x,
                        uint32_t* d_touched_y, uint32_t* d_touched_z, uint32_t* d_list_x, uint32_t* d_list_y,
                        uint32_t* d_list_z, const uint32_t* capacities, uint32_t* totals, void* stream) {
    return guarded([&] {
        NEED(totals);
        totals[0] = totals[1] = totals[2] = 0;
        if (n == 0) return;
        NEED(d_px); NEED(d_py); NEED(d_pz); NEED(d_keys); NEED(d_touched_x); NEED(d_touched_y); NEED(d_touched_z);
        FLIP2_REQUIRE(grid.sizeX && grid.sizeY && grid.sizeZ && grid.cellSize > 0.0f, flip2::kErrArgument,
                      "invalid grid");
        FLIP2_REQUIRE(refinement_level < 16 && radius >= 0.0 && radius < 64.0, flip2::kErrArgument,
                      "refinement level or radius out of range");
        const hipStream_t st = as_stream(stream);
        flip2::rc::launch_sub_cell_counts(d_px, d_py, d_pz, n, to_grid(grid), d_keys, refinement_level, radius,
                                          d_touched_x, d_touched_y, d_touched_z, st);
        uint32_t* t[3] = {d_touched_x, d_touched_y, d_touched_z};
        const size_t ws = flip2::scan_workspace_bytes(n);
        void* w = nullptr;
        FLIP2_HIP(hipMallocAsync(&w, ws, st));
        for (int d = 0; d < 3; ++d) {
            flip2::scan_u32(t[d], t[d], n, false, w, ws, nullptr, st);
            FLIP2_HIP(hipMemcpyAsync(&totals[d], t[d] + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        }
        FLIP2_HIP(hipFreeAsync(w, st));
        FLIP2_HIP(hipStreamSynchronize(st));
        if (!d_list_x && !d_list_y && !d_list_z) return;
        NEED(d_list_x); NEED(d_list_y); NEED(d_list_z); NEED(capacities);
        for (int d = 0; d < 3; ++d)
            FLIP2_REQUIRE(capacities[d] >= totals[d], flip2::kErrRange, "sub-cell list capacity too small");
        flip2::rc::launch_sub_cell_lists(d_px, d_py, d_pz, n, to_grid(grid), d_keys, refinement_level, radius,
                                         d_touched_x, d_touched_y, d_touched_z, d_list_x, d_list_y, d_list_z, st);
        FLIP2_HIP(hipStreamSynchronize(st));
    });
}

int flip2_sum_particles_per_node(uint32_t num_nodes, uint32_t n, const uint32_t* d_first, const uint32_t* d_touched,
                                 const uint32_t* d_list, uint32_t* d_num_nonzero_voxels, uint32_t* d_num_entries,
                                 void* stream) {
    return guarded([&] {
        if (num_nodes == 0) return;
        NEED(d_first); NEED(d_touched); NEED(d_num_nonzero_voxels); NEED(d_num_entries);
        FLIP2_REQUIRE(n > 0, flip2::kErrArgument, "nodes without particles");
        const hipStream_t st = as_stream(stream);
        flip2::rc::launch_particles_per_node(num_nodes, n, d_first, d_touched, d_list, d_num_nonzero_voxels,
                                             d_num_entries, st);
        const size_t ws = flip2::scan_workspace_bytes(num_nodes);
        void* w = nullptr;
        FLIP2_HIP(hipMallocAsync(&w, ws, st));
        flip2::scan_u32(d_num_nonzero_voxels, d_num_nonzero_voxels, num_nodes, false, w, ws, nullptr, st);
        flip2::scan_u32(d_num_entries, d_num_entries, num_nodes, false, w, ws, nullptr, st);
        FLIP2_HIP(hipFreeAsync(w, st));
        FLIP2_HIP(hipStreamSynchronize(st));
    });
}

int flip2_particle_list_create(uint32_t num_nodes, uint32_t n, const uint32_t* d_num_nonzero_voxels,
                               const uint32_t* d_num_entries, const uint32_t* d_touched, const uint32_t* d_list,
                               const uint32_t* d_first, uint32_t* d_voxel_ids, uint32_t* d_list_start,
                               uint32_t* d_particle_lists, void* stream) {
    return guarded([&] {
        if (num_nodes == 0) return;
        NEED(d_num_nonzero_voxels); NEED(d_num_entries); NEED(d_touched); NEED(d_first); NEED(d_voxel_ids);
        NEED(d_list_start); NEED(d_particle_lists);
        FLIP2_REQUIRE(n > 0, flip2::kErrArgument, "nodes without particles");
        const hipStream_t st = as_stream(stream);
        flip2::rc::launch_particle_lists(num_nodes, n, d_num_nonzero_voxels, d_num_entries, d_touched, d_list,
                                         d_first, d_voxel_ids, d_list_start, d_particle_lists, st);
        FLIP2_HIP(hipStreamSynchronize(st));
    });
}

int flip2_find_grid_cell(double* d_px, double* d_py, double* d_pz, uint32_t n, flip2_grid grid, uint32_t* d_keys,
                         void* stream) {
    return guarded([&] {
        FLIP2_REQUIRE(n == 0 || (d_px && d_py && d_pz && d_keys), flip2::kErrArgument, "null buffer");
        FLIP2_REQUIRE(grid.sizeX && grid.sizeY && grid.sizeZ && grid.cellSize > 0.0f, flip2::kErrArgument,
                      "invalid grid");
        flip2::rc::launch_root_cell(d_px, d_py, d_pz, n, to_grid(grid), d_keys, as_stream(stream));
    });
}

int flip2_philox4x32_10(const uint32_t* in, uint32_t count, uint32_t* out) {
    return guarded([&] {
        if (!count) return;
        NEED(in);
        NEED(out);
        void *din = nullptr, *dout = nullptr;
        FLIP2_HIP(hipMalloc(&din, (size_t)count * 24));
        try {
            FLIP2_HIP(hipMalloc(&dout, (size_t)count * 16));
            FLIP2_HIP(hipMemcpy(din, in, (size_t)count * 24, hipMemcpyHostToDevice));
            flip2::rc::launch_philox_kat(reinterpret_cast<uint32_t*>(din), count, reinterpret_cast<uint32_t*>(dout),
                                         nullptr);
            FLIP2_HIP(hipMemcpy(out, dout, (size_t)count * 16, hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(din);
            if (dout) (void)hipFree(dout);
            throw;
        }
        (void)hipFree(din);
        (void)hipFree(dout);
    });
}

int flip2_particles_create(uint32_t n, flip2_particles** out) {
    return guarded([&] {
        NEED(out);
        *out = new flip2_particles(n);
    });
}
int flip2_particles_destroy(flip2_particles* p) {
    return guarded([&] { delete p; });
}
int flip2_set_domain(flip2_particles* p, double nx, double ny, double nz, uint32_t x, uint32_t y, uint32_t z,
                     double cell_size) {
    return guarded([&] {
        NEED(p);
        p->impl.setDomain(nx, ny, nz, x, y, z, cell_size);
    });
}
int flip2_set_positions(flip2_particles* p, const double* px, const double* py, const double* pz) {
    return guarded([&] {
        NEED(p); NEED(px); NEED(py); NEED(pz);
        p->impl.setPositions(px, py, pz);
    });
}
int flip2_set_velocities(flip2_particles* p, const double* vx, const double* vy, const double* vz) {
    return guarded([&] {
        NEED(p); NEED(vx); NEED(vy); NEED(vz);
        p->impl.setVelocities(vx, vy, vz);
    });
}
int flip2_get_positions(flip2_particles* p, double* px, double* py, double* pz) {
    return guarded([&] {
        NEED(p); NEED(px); NEED(py); NEED(pz);
        p->impl.getPositions(px, py, pz);
    });
}
int flip2_get_velocities(flip2_particles* p, double* vx, double* vy, double* vz) {
    return guarded([&] {
        NEED(p); NEED(vx); NEED(vy); NEED(vz);
        p->impl.getVelocities(vx, vy, vz);
    });
}

#define STAGE(fn, call)                                   \
    int fn(flip2_particles* p) {                          \
        return guarded([&] {                              \
            NEED(p);                                      \
            p->impl.call();                               \
            p->impl.synchronize();                        \
        });                                               \
    }

STAGE(flip2_align_particles_to_grid, alignParticlesToGrid)
STAGE(flip2_sort_particles, sortParticles)
STAGE(flip2_generate_voxels, generateVoxels)
STAGE(flip2_particle_vel_to_voxels, particleVelToVoxels)
STAGE(flip2_calc_div_u, calcDivU)
STAGE(flip2_pressure_solve, pressureSolve)
STAGE(flip2_update_voxel_velocities, updateVoxelVelocities)
STAGE(flip2_voxel_vels_to_particles, voxelVelsToParticles)
STAGE(flip2_advect_particles, advectParticles)
STAGE(flip2_move_solids, moveSolidsOnly)
STAGE(flip2_move_solid_particles, moveSolidParticles)
STAGE(flip2_initialize, initialize)

int flip2_get_courant_dt(flip2_particles* p, double* dt) {
    return guarded([&] {
        NEED(p); NEED(dt);
        *dt = p->impl.getCourantDt();
    });
}
int flip2_apply_gravity(flip2_particles* p, double dt) {
    return guarded([&] {
        NEED(p);
        p->impl.applyGravity(dt);
        p->impl.synchronize();
    });
}
int flip2_remove_gravity(flip2_particles* p, double dt) {
    return guarded([&] {
        NEED(p);
        p->impl.removeGravity(dt);
        p->impl.synchronize();
    });
}
int flip2_gs_solve(flip2_particles* p, double dt, double threshold, uint32_t max_iterations, double* residual,
                   uint32_t* iterations) {
    return guarded([&] {
        NEED(p);
        const double r = p->impl.gsSolve(dt, threshold, max_iterations, iterations);
        if (residual) *residual = r;
        p->impl.synchronize();
    });
}
int flip2_gs_iterations(flip2_particles* p, double dt, uint32_t count) {
    return guarded([&] {
        NEED(p);
        p->impl.gsIterations(dt, count);
        p->impl.synchronize();
    });
}
int flip2_solve_frame(flip2_particles* p, double fps) {
    return guarded([&] {
        NEED(p);
        FLIP2_REQUIRE(fps > 0.0, flip2::kErrArgument, "fps must be positive");
        p->impl.solveFrame(fps);
    });
}
int flip2_write_positions_to_file(flip2_particles* p, const char* path) {
    return guarded([&] {
        NEED(p); NEED(path);
        p->impl.writePositionsToFile(path);
    });
}
int flip2_get_stats(flip2_particles* p, flip2_rc_stats* s) {
    return guarded([&] {
        NEED(p); NEED(s);
        const flip2::RcStats r = p->impl.stats();
        s->numUsedGridNodes = r.numUsedGridNodes;
        s->numUsedVoxels = r.numUsedVoxels;
        s->lastSorIterations = r.lastSorIterations;
        s->lastSolveAttempts = r.lastSolveAttempts;
        s->lastResidual = r.lastResidual;
        s->dt = r.dt;
        s->frameDt = r.frameDt;
        s->elapsedTime = r.elapsedTime;
        s->elapsedTimeThisFrame = r.elapsedTimeThisFrame;
        s->outOfTileEvents = r.outOfTileEvents;
        s->relocatedParticles = r.relocatedParticles;
        s->substeps = r.substeps;
    });
}
int flip2_get_array(flip2_particles* p, const char* name, void* host, size_t capacity, size_t* count) {
    return guarded([&] {
        NEED(p); NEED(name);
        const size_t c = p->impl.getArray(name, host, capacity);
        if (count) *count = c;
    });
}
int flip2_set_array(flip2_particles* p, const char* name, const void* host, size_t count) {
    return guarded([&] {
        NEED(p); NEED(name);
        FLIP2_REQUIRE(count == 0 || host, flip2::kErrArgument, "host is NULL");
        p->impl.setArray(name, host, count);
    });
}
int flip2_set_solver_limits(flip2_particles* p, double threshold, uint32_t max_iterations) {
    return guarded([&] {
        NEED(p);
        p->impl.setSolverLimits(threshold, max_iterations);
    });
}
int flip2_set_time(flip2_particles* p, double dt, double frame_dt, double elapsed_time,
                   double elapsed_time_this_frame) {
    return guarded([&] {
        NEED(p);
        p->impl.setTime(dt, frame_dt, elapsed_time, elapsed_time_this_frame);
    });
}

int flip2_ns_bin(const float* d_x, const float* d_y, const float* d_z, uint32_t n, flip2_grid grid, uint32_t* d_keys,
                 void* stream) {
    return guarded([&] {
        FLIP2_REQUIRE(n == 0 || (d_x && d_y && d_z && d_keys), flip2::kErrArgument, "null buffer");
        FLIP2_REQUIRE(grid.sizeX % 8 == 0 && grid.sizeY % 8 == 0 && grid.sizeZ % 8 == 0 && grid.sizeX && grid.sizeY &&
                          grid.sizeZ && grid.cellSize > 0.0f,
                      flip2::kErrArgument, "grid sizes must be positive multiples of 8");
        flip2::ns::launch_bin(d_x, d_y, d_z, n, to_grid(grid), d_keys, as_stream(stream));
    });
}

int flip2_ns_p2g(const float* d_records, const uint32_t* d_cell_start, flip2_grid grid, float* d_u, float* d_v,
                 float* d_w, float* d_wu, float* d_wv, float* d_ww, void* stream) {
    return guarded([&] {
        FLIP2_REQUIRE(d_records && d_cell_start && d_u && d_v && d_w, flip2::kErrArgument, "null buffer");
        FLIP2_REQUIRE((reinterpret_cast<uintptr_t>(d_records) & 15u) == 0, flip2::kErrArgument,
                      "records must be 16-byte aligned");
        FLIP2_REQUIRE(grid.sizeX % 8 == 0 && grid.sizeY % 8 == 0 && grid.sizeZ % 8 == 0 && grid.sizeX && grid.sizeY &&
                          grid.sizeZ && grid.cellSize > 0.0f,
                      flip2::kErrArgument, "grid sizes must be positive multiples of 8");
        const flip2::Grid g = to_grid(grid);
        const hipStream_t st = as_stream(stream);
        const float4* rec = reinterpret_cast<const float4*>(d_records);
        uint32_t* stats = nullptr;
        FLIP2_HIP(hipMallocAsync(reinterpret_cast<void**>(&stats), sizeof(uint32_t), st));
        flip2::ns::launch_p2g_stats(rec, d_cell_start, g.sizeX * g.sizeY * g.sizeZ, stats, st);
        flip2::ns::launch_p2g(rec, d_cell_start, g, stats, d_u, d_v, d_w, d_wu, d_wv, d_ww, st);
        FLIP2_HIP(hipFreeAsync(stats, st));
    });
}

int flip2_p2g_pipeline_create(uint32_t max_particles, flip2_grid grid, flip2_p2g_pipeline** out) {
    return guarded([&] {
        NEED(out);
        *out = new flip2_p2g_pipeline(max_particles, to_grid(grid));
    });
}
int flip2_p2g_pipeline_destroy(flip2_p2g_pipeline* pl) {
    return guarded([&] {
        if (pl) FLIP2_HIP(hipDeviceSynchronize());
        delete pl;
    });
}
int flip2_p2g_pipeline_run(flip2_p2g_pipeline* pl, uint32_t n, const float* d_x, const float* d_y, const float* d_z,
                           const float* d_vx, const float* d_vy, const float* d_vz, float* d_u, float* d_v,
                           float* d_w, void* stream) {
    return guarded([&] {
        NEED(pl);
        pl->impl.run(n, d_x, d_y, d_z, d_vx, d_vy, d_vz, d_u, d_v, d_w, as_stream(stream));
    });
}
int flip2_p2g_pipeline_run_records(flip2_p2g_pipeline* pl, uint32_t n, const float* d_records, float* d_u, float* d_v,
                                   float* d_w, void* stream) {
    return guarded([&] {
        NEED(pl);
        FLIP2_REQUIRE(n == 0 || d_records, flip2::kErrArgument, "null records");
        FLIP2_REQUIRE((reinterpret_cast<uintptr_t>(d_records) & 15u) == 0, flip2::kErrArgument,
                      "records must be 16-byte aligned");
        pl->impl.runRecords(n, reinterpret_cast<const float4*>(d_records), d_u, d_v, d_w, as_stream(stream));
    });
}
int flip2_p2g_pipeline_enable_timing(flip2_p2g_pipeline* pl, int capacity) {
    return guarded([&] {
        NEED(pl);
        pl->impl.enableTiming(capacity);
        pl->impl.resetTiming();
    });
}
int flip2_p2g_pipeline_timing(flip2_p2g_pipeline* pl, int k, char* name, size_t name_len, double* total_ms,
                              int* runs) {
    return guarded([&] {
        NEED(pl);
        pl->impl.finishTiming();
        FLIP2_REQUIRE(k >= 0 && k < pl->impl.timedKernels(), flip2::kErrRange, "no such timed kernel");
        if (name && name_len) {
            std::strncpy(name, pl->impl.timedKernelName(k), name_len - 1);
            name[name_len - 1] = 0;
        }
        if (total_ms) *total_ms = pl->impl.timedKernelMs(k);
        if (runs) *runs = pl->impl.timedRuns();
    });
}
int flip2_p2g_pipeline_get(flip2_p2g_pipeline* pl, const char* name, void* host, size_t capacity, size_t* count) {
    return guarded([&] {
        NEED(pl); NEED(name);
        FLIP2_HIP(hipDeviceSynchronize());
        const size_t c = pl->impl.get(name, host, capacity, nullptr);
        if (count) *count = c;
    });
}

void flip2_fluid_default_config(flip2_fluid_config* c) {
    if (!c) return;
    flip2::FluidConfig d;
    c->grid = flip2_grid{0.f, 0.f, 0.f, 64, 64, 64, 1.0f / 64.0f};
    c->dt = d.dt;
    c->gravity = d.gravity;
    c->flip = d.flip;
    c->preconditioner = d.preconditioner;
    c->tolerance = d.tolerance;
    c->max_iterations = d.maxIterations;
    c->check_every = d.checkEvery;
    c->margin = d.margin;
}
int flip2_fluid_create(uint32_t max_particles, const flip2_fluid_config* cfg, flip2_fluid** out) {
    return guarded([&] {
        NEED(cfg); NEED(out);
        *out = new flip2_fluid(max_particles, to_config(*cfg));
    });
}
int flip2_fluid_destroy(flip2_fluid* f) {
    return guarded([&] { delete f; });
}
int flip2_fluid_set_config(flip2_fluid* f, const flip2_fluid_config* cfg) {
    return guarded([&] {
        NEED(f); NEED(cfg);
        // the learned-preconditioner settings belong to flip2_fluid_set_unet
        flip2::FluidConfig c = to_config(*cfg);
        c.learnedFallbackIterations = f->impl.config().learnedFallbackIterations;
        c.learnedSmoothing = f->impl.config().learnedSmoothing;
        c.learnedDirections = f->impl.config().learnedDirections;
        c.learnedCheckEvery = f->impl.config().learnedCheckEvery;
        c.learnedGraph = f->impl.config().learnedGraph;
        c.learnedSmoother = f->impl.config().learnedSmoother;
        c.learnedSmoothSteps = f->impl.config().learnedSmoothSteps;
        f->impl.setConfig(c);
    });
}
int flip2_fluid_get_config(flip2_fluid* f, flip2_fluid_config* cfg) {
    return guarded([&] {
        NEED(f); NEED(cfg);
        from_config(f->impl.config(), cfg);
    });
}
int flip2_fluid_set_particles(flip2_fluid* f, uint32_t n, const float* x, const float* y, const float* z,
                              const float* vx, const float* vy, const float* vz) {
    return guarded([&] {
        NEED(f);
        FLIP2_REQUIRE(n == 0 || (x && y && z && vx && vy && vz), flip2::kErrArgument, "null particle array");
        f->impl.setParticles(n, x, y, z, vx, vy, vz);
    });
}
int flip2_fluid_get_particles(flip2_fluid* f, float* x, float* y, float* z, float* vx, float* vy, float* vz) {
    return guarded([&] {
        NEED(f);
        f->impl.getParticles(x, y, z, vx, vy, vz);
    });
}
int flip2_fluid_set_solid(flip2_fluid* f, const uint8_t* mask) {
    return guarded([&] {
        NEED(f);
        f->impl.setSolidMask(mask);
    });
}
int flip2_fluid_set_solid_box(flip2_fluid* f, const float* lo, const float* hi, const float* vel) {
    return guarded([&] {
        NEED(f);
        f->impl.setSolidBox(lo, hi, vel);
    });
}
int flip2_fluid_num_particles(flip2_fluid* f, uint32_t* n) {
    return guarded([&] {
        NEED(f);
        NEED(n);
        *n = f->impl.numParticles();
    });
}
int flip2_fluid_save_snapshot(flip2_fluid* f, const char* path) {
    return guarded([&] {
        NEED(f);
        NEED(path);
        f->impl.saveSnapshot(path);
    });
}
int flip2_fluid_load_snapshot(flip2_fluid* f, const char* path) {
    return guarded([&] {
        NEED(f);
        NEED(path);
        f->impl.loadSnapshot(path);
    });
}
int flip2_fluid_step(flip2_fluid* f) {
    return guarded([&] {
        NEED(f);
        const flip2::KernelTimer::Install timed(f->activeTimer());
        f->impl.step();
        f->impl.synchronize();
    });
}
int flip2_fluid_step_frame(flip2_fluid* f, double frame_dt, double cfl, uint32_t max_substeps, uint32_t* substeps) {
    return guarded([&] {
        NEED(f);
        const flip2::KernelTimer::Install timed(f->activeTimer());
        const uint32_t k = f->impl.stepFrame(frame_dt, cfl, max_substeps);
        f->impl.synchronize();
        if (substeps) *substeps = k;
    });
}
int flip2_fluid_stage(flip2_fluid* f, const char* stage) {
    return guarded([&] {
        NEED(f); NEED(stage);
        const flip2::KernelTimer::Install timed(f->activeTimer());
        const std::string s(stage);
        if (s == "transfer") f->impl.transfer();
        else if (s == "transfer_prepare") f->impl.transferAndPrepare();
        else if (s == "prepare") f->impl.prepareFaces();
        else if (s == "system") f->impl.buildSystem();
        else if (s == "solve") f->impl.solve();
        else if (s == "project") f->impl.project();
        else if (s == "g2p") f->impl.gridToParticles();
        else if (s == "advect") f->impl.advect();
        else if (s == "g2p_advect") f->impl.gridToParticlesAndAdvect();
        else throw flip2::Error(flip2::kErrArgument, "unknown stage " + s);
        f->impl.synchronize();
    });
}
int flip2_fluid_solve_with(flip2_fluid* f, int preconditioner, flip2_solve_stats* stats) {
    return guarded([&] {
        NEED(f);
        const flip2::KernelTimer::Install timed(f->activeTimer());
        const flip2::SolveStats s = f->impl.solveWith(preconditioner);
        f->impl.synchronize();
        if (stats) to_stats(s, stats);
    });
}
int flip2_fluid_last_solve(flip2_fluid* f, flip2_solve_stats* stats) {
    return guarded([&] {
        NEED(f); NEED(stats);
        to_stats(f->impl.lastSolve(), stats);
    });
}
int flip2_fluid_get_array(flip2_fluid* f, const char* name, void* host, size_t capacity, size_t* count) {
    return guarded([&] {
        NEED(f); NEED(name);
        const size_t c = f->impl.getArray(name, host, capacity);
        if (count) *count = c;
    });
}

int flip2_unet_create(int n, flip2_unet** out) {
    return guarded([&] {
        NEED(out);
        *out = new flip2_unet(n);
    });
}
int flip2_unet_destroy(flip2_unet* net) {
    return guarded([&] {
        if (net) FLIP2_HIP(hipDeviceSynchronize());
        delete net;
    });
}
int flip2_unet_num_params(flip2_unet* net, size_t* count) {
    return guarded([&] {
        NEED(net); NEED(count);
        *count = net->net.numParams();
    });
}
int flip2_unet_init(flip2_unet* net, uint64_t seed) {
    return guarded([&] {
        NEED(net);
        net->net.initParams(seed);
    });
}
int flip2_unet_set_params(flip2_unet* net, const float* host, size_t count) {
    return guarded([&] {
        NEED(net); NEED(host);
        net->net.setParams(host, count);
    });
}
int flip2_unet_get_params(flip2_unet* net, float* host, size_t count) {
    return guarded([&] {
        NEED(net); NEED(host);
        net->net.getParams(host, count);
    });
}
namespace {
void upload_dense(flip2_unet* net, const float* b, const uint16_t* coef) {
    const uint32_t nv = (uint32_t)net->n() * net->n() * net->n();
    if (net->bTmp.size() < nv) net->bTmp.resizeAsync(nv, nullptr);
    if (net->coefTmp.size() < nv) net->coefTmp.resizeAsync(nv, nullptr);
    net->bTmp.uploadFrom(b, nv, nullptr);
    net->coefTmp.uploadFrom(coef, nv, nullptr);
}
}  // namespace
int flip2_unet_apply_dense(flip2_unet* net, const float* b, const uint16_t* coef, float* z) {
    return guarded([&] {
        NEED(net); NEED(b); NEED(coef); NEED(z);
        upload_dense(net, b, coef);
        const int nn = net->n();
        const float* zd = net->apply(net->bTmp.devPtr(), net->coefTmp.devPtr());
        FLIP2_HIP(hipMemcpy(z, zd, (size_t)nn * nn * nn * sizeof(float), hipMemcpyDeviceToHost));
    });
}
int flip2_unet_loss_grad(flip2_unet* net, const float* b, const uint16_t* coef, double* loss, float* grad) {
    return guarded([&] {
        NEED(net); NEED(b); NEED(coef);
        upload_dense(net, b, coef);
        net->net.zeroGrad(nullptr);
        const double l = net->lossGrad(net->bTmp.devPtr(), net->coefTmp.devPtr(), grad != nullptr);
        if (loss) *loss = l;
        if (grad)
            FLIP2_HIP(hipMemcpy(grad, net->net.gradDev(), net->net.numParams() * sizeof(float), hipMemcpyDeviceToHost));
    });
}
int flip2_unet_set_smoothing(flip2_unet* net, float omega) {
    return guarded([&] {
        NEED(net);
        FLIP2_REQUIRE(omega >= 0.0f && omega < 1.0f, flip2::kErrArgument, "smoothing weight must lie in [0, 1)");
        net->omega = omega;
    });
}
int flip2_unet_add_frame(flip2_unet* net, flip2_fluid* f) {
    return guarded([&] {
        NEED(net); NEED(f);
        flip2::FluidSim& sim = f->impl;
        FLIP2_REQUIRE(sim.gridN() == net->n(), flip2::kErrArgument, "grid size differs from the network's");
        sim.synchronize();
        flip2_unet::Frame& fr = net->newFrame();
        flip2::nn::launch_dense_system(sim.numActive(), sim.activeBricksDev(), sim.coefDev(), sim.rhsDev(), nullptr,
                                       net->n(), fr.b.devPtr(), nullptr, fr.coef.devPtr(), nullptr);
        FLIP2_HIP(hipDeviceSynchronize());
    });
}
int flip2_unet_capture_residuals(flip2_unet* net, flip2_fluid* f, const uint32_t* iterations, int count) {
    return guarded([&] {
        NEED(net); NEED(f);
        flip2::FluidSim& sim = f->impl;
        if (count <= 0) {
            sim.setResidualCapture({}, nullptr);
            return;
        }
        NEED(iterations);
        FLIP2_REQUIRE(sim.gridN() == net->n(), flip2::kErrArgument, "grid size differs from the network's");
        std::vector<uint32_t> its(iterations, iterations + count);
        flip2::FluidSim* sp = &sim;
        sim.setResidualCapture(std::move(its), [net, sp](uint32_t na, const float* r, hipStream_t st) {
            flip2_unet::Frame& fr = net->newFrame(st);
            flip2::nn::launch_dense_vector(na, sp->activeBricksDev(), sp->coefDev(), r, net->n(), fr.b.devPtr(),
                                           fr.coef.devPtr(), st);
        });
    });
}
int flip2_unet_num_frames(flip2_unet* net, int* count) {
    return guarded([&] {
        NEED(net); NEED(count);
        *count = (int)net->frames.size();
    });
}
int flip2_unet_clear_frames(flip2_unet* net) {
    return guarded([&] {
        NEED(net);
        net->frames.clear();
    });
}
int flip2_unet_train(flip2_unet* net, int epochs, float lr, int first, int last, uint64_t seed, double* mean_loss) {
    return guarded([&] {
        NEED(net);
        FLIP2_REQUIRE(first >= 0 && last <= (int)net->frames.size() && first < last, flip2::kErrRange,
                      "bad frame range");
        std::vector<int> order;
        for (int i = first; i < last; ++i) order.push_back(i);
        uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
        double mean = 0.0;
        for (int e = 0; e < epochs; ++e) {
            for (size_t i = order.size(); i > 1; --i) {  // Fisher-Yates with an LCG
                st = st * 6364136223846793005ull + 1442695040888963407ull;
                std::swap(order[i - 1], order[(size_t)((st >> 33) % i)]);
            }
            double sum = 0.0;
            for (int fi : order) {
                flip2_unet::Frame& fr = *net->frames[fi];
                net->net.zeroGrad(nullptr);
                sum += net->lossGrad(fr.b.devPtr(), fr.coef.devPtr(), true);
                net->net.adamStep(lr, nullptr);
            }
            mean = sum / (double)order.size();
        }
        FLIP2_HIP(hipDeviceSynchronize());
        if (mean_loss) *mean_loss = mean;
    });
}
int flip2_unet_eval(flip2_unet* net, int first, int last, double* mean_loss) {
    return guarded([&] {
        NEED(net);
        FLIP2_REQUIRE(first >= 0 && last <= (int)net->frames.size() && first < last, flip2::kErrRange,
                      "bad frame range");
        double sum = 0.0;
        for (int i = first; i < last; ++i)
            sum += net->lossGrad(net->frames[i]->b.devPtr(), net->frames[i]->coef.devPtr(), false);
        if (mean_loss) *mean_loss = sum / (double)(last - first);
    });
}
int flip2_fluid_set_flex_directions(flip2_fluid* f, uint32_t directions) {
    return guarded([&] {
        NEED(f);
        FLIP2_REQUIRE(directions >= 1 && directions <= (uint32_t)flip2::ns::kMaxFlexDirections, flip2::kErrArgument,
                      "flexible CG keeps 1 to 16 directions");
        flip2::FluidConfig c = f->impl.config();
        c.learnedDirections = directions;
        f->impl.setConfig(c);
    });
}
int flip2_fluid_set_flex_check_every(flip2_fluid* f, uint32_t iterations) {
    return guarded([&] {
        NEED(f);
        FLIP2_REQUIRE(iterations >= 1 && iterations <= 1024, flip2::kErrArgument,
                      "flexible CG checks its residual every 1 to 1024 iterations");
        flip2::FluidConfig c = f->impl.config();
        c.learnedCheckEvery = iterations;
        f->impl.setConfig(c);
    });
}
int flip2_fluid_set_learned_smoother(flip2_fluid* f, int smoother) {
    return guarded([&] {
        NEED(f);
        flip2::FluidConfig c = f->impl.config();
        c.learnedSmoother = smoother;
        f->impl.setConfig(c);
    });
}
int flip2_fluid_set_learned_smooth_steps(flip2_fluid* f, uint32_t steps) {
    return guarded([&] {
        NEED(f);
        flip2::FluidConfig c = f->impl.config();
        c.learnedSmoothSteps = steps;
        f->impl.setConfig(c);
    });
}
int flip2_fluid_set_learned_graph(flip2_fluid* f, int enable) {
    return guarded([&] {
        NEED(f);
        flip2::FluidConfig c = f->impl.config();
        c.learnedGraph = enable != 0;
        f->impl.setConfig(c);
    });
}
int flip2_fluid_set_unet(flip2_fluid* f, flip2_unet* net, uint32_t fallback_iterations) {
    return guarded([&] {
        NEED(f);
        flip2::FluidSim& sim = f->impl;
        flip2::FluidConfig c = sim.config();
        c.learnedFallbackIterations = fallback_iterations;
        c.learnedSmoothing = net ? net->omega : 0.0f;
        sim.setConfig(c);
        if (!net) {
            sim.setLearnedPreconditioner(nullptr);
            return;
        }
        FLIP2_REQUIRE(sim.gridN() == net->n(), flip2::kErrArgument, "grid size differs from the network's");
        flip2::UNet* u = &net->net;
        flip2::FluidSim* s = &sim;
        sim.setLearnedPreconditioner(
            [u, s](uint32_t na, const float* r, float* z, bool newSystem, const double* pairs, hipStream_t st,
                   const uint32_t* stop) {
                u->applyActive(na, s->activeBricksDev(), s->brickSlotDev(), s->cellTypeDev(), r, z, s->gridN(),
                               newSystem, pairs, st, stop);
            });
    });
}

int flip2_fluid_enable_timing(flip2_fluid* f, const char* kernels) {
    return guarded([&] {
        NEED(f);
        f->timer.reset();
        std::vector<std::string> names;
        if (kernels) {
            std::string all(kernels), cur;
            for (char c : all + ",") {
                if (c == ',') {
                    if (!cur.empty()) names.push_back(cur);
                    cur.clear();
                } else if (c != ' ') {
                    cur.push_back(c);
                }
            }
        }
        f->timer.setFilter(names);
        f->timer.setSampleEvery(1);
        f->timing = true;
    });
}
int flip2_fluid_set_timing_sample(flip2_fluid* f, uint32_t every) {
    return guarded([&] {
        NEED(f);
        FLIP2_REQUIRE(every >= 1, flip2::kErrArgument, "timing: sample every >= 1 launches");
        f->timer.setSampleEvery(every);
    });
}
int flip2_fluid_disable_timing(flip2_fluid* f) {
    return guarded([&] {
        NEED(f);
        f->timer.collect();
        f->timing = false;
    });
}
int flip2_fluid_timing(flip2_fluid* f, int k, char* name, size_t name_len, double* total_ms, uint64_t* launches) {
    return guarded([&] {
        NEED(f);
        f->timer.collect();
        const auto& t = f->timer.totals();
        FLIP2_REQUIRE(k >= 0 && (size_t)k < t.size(), flip2::kErrRange, "no such timed kernel");
        auto it = t.begin();
        std::advance(it, k);
        if (name && name_len) {
            std::strncpy(name, it->first.c_str(), name_len - 1);
            name[name_len - 1] = 0;
        }
        if (total_ms) *total_ms = it->second.ms;
        if (launches) *launches = it->second.launches;
    });
}

}  // extern "C"

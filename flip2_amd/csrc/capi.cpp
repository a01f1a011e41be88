// C ABI of libparticles.so (include/flip2.h): every entry point catches
// exceptions, records the message per thread and returns a status code.
#include "../../include/flip2.h"

#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"
#include "particles.hpp"
#include "rc_kernels.hpp"
#include "scan_impl.hpp"
#include "sort.hpp"

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

#define NEED(p) FLIP2_REQUIRE((p) != nullptr, flip2::kErrArgument, #p " is NULL")

}  // namespace

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

}  // extern "C"

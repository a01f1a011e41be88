// flip2_main: the reference's stand-alone driver (engine/main.cu:6-27) on
// this library.  A domain of 100^3 cells of 0.25 with corner (-10,-10,-10),
// particles uniform in the box the reference's tester fills
// (ParticleSystemTester::randomizeParticlePositions, testing.h:55-76: x and
// z inside [neg + L/3, neg + n h - L/3], y in [neg, neg + n h - L/3], L the
// domain height, velocities 0), initialize, then solveFrame(24) per frame;
// positions are written as "x, y, z" lines to files 0, 1, ..., frames.
//
//   flip2_main [frames=300] [particles=524264] [seed] [output directory=.]
//
// The reference seeds from std::random_device; a seed argument makes a run
// repeatable (without one, std::random_device seeds it as there).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../particles.hpp"

int main(int argc, char** argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 300;
    const uint32_t n = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : (1u << 19) - 24u;
    const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : (uint64_t)std::random_device{}();
    const std::string dir = argc > 4 ? std::string(argv[4]) + "/" : std::string();
    if (frames < 0 || n == 0) {
        std::fprintf(stderr, "usage: flip2_main [frames] [particles] [seed] [output directory]\n");
        return 2;
    }
    try {
        flip2::Particles particles(n);
        const double neg = -10.0, h = 0.25;
        const uint32_t cells = 100;
        particles.setDomain(neg, neg, neg, cells, cells, cells, h);
        const double third = cells * h / 3.0f;
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<double> dx(neg + third, neg + cells * h - third);
        std::uniform_real_distribution<double> dy(neg, neg + cells * h - third);
        std::uniform_real_distribution<double> dz(neg + third, neg + cells * h - third);
        std::vector<double> px(n), py(n), pz(n), zero(n, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            px[i] = dx(rng);
            py[i] = dy(rng);
            pz[i] = dz(rng);
        }
        particles.setArray("px", px.data(), n);
        particles.setArray("py", py.data(), n);
        particles.setArray("pz", pz.data(), n);
        particles.setArray("vx", zero.data(), n);
        particles.setArray("vy", zero.data(), n);
        particles.setArray("vz", zero.data(), n);
        particles.initialize();
        particles.writePositionsToFile(dir + std::to_string(0));
        for (int i = 0; i < frames; ++i) {
            particles.solveFrame(24.0);
            particles.writePositionsToFile(dir + std::to_string(i + 1));
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "flip2_main: %s\n", e.what());
        return 1;
    }
    return 0;
}

// zr_host_step -- runs MpmSimulationGpu (the C++ host side) on a particle
// file: the way a Ziran driver (SimulationBase::advanceOneFrame,
// SimulationBase.h:304-308) calls the step, without Python in the loop.
//
//   zr_host_step IN OUT dx steps dt symplectic(1|0) [ground_y] [precision] [cfl]
//
// IN/OUT: int64 n, then X 3n, V 3n, C 9n, F 9n, mass n, vol n, mu n,
// lambda n doubles (C and F column-major per particle).  ground_y adds a
// sticky half-space y >= ground_y.  Gravity is (0, -9.8, 0).  Prints one
// line per step: step, groups, pages, and for implicit steps Newton and CG
// iteration counts.  A negative dt runs `steps` frames of length -dt
// instead (SimulationBase::advanceOneFrame with the CFL time step,
// max_dt = -dt; cfl 0.6 unless given) and prints one line per frame: frame,
// substeps, time.  Exit status 0 on success, 3 on a CFL restart, 1 on any other error.
//
//   zr_host_step simulate IN OUT dx frames frame_dt max_dt cfl DIR [ground_y]
//
// runs SimulationBase::simulate (frames 1..frames of length frame_dt, the
// CFL time step capped at max_dt, the restart state written to
// DIR/restart_<frame>.dat after each frame) with MpmSimulationBase's
// automatic restart: a particle moving more than a cell re-reads the last
// frame's state with max_dt halved.  Prints one line per frame run (frame,
// substeps, max_dt) and one per restart; OUT holds the final particles.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "MpmSimulationGpu.h"

using ziran_amd::MpmSimulationGpu;
using ziran_amd::ParticleArrays;

namespace {

template <class T>
bool rd(FILE* f, T* p, size_t count)
{
  return std::fread(p, sizeof(T), count, f) == count;
}
template <class T>
bool wr(FILE* f, const T* p, size_t count)
{
  return std::fwrite(p, sizeof(T), count, f) == count;
}

bool read_particles(const char* path, ParticleArrays& p)
{
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  int64_t n = 0;
  bool ok = rd(f, &n, 1) && n >= 0;
  if (ok) {
    p.resize((size_t)n);
    const size_t m = (size_t)n;
    ok = (m == 0 || (rd(f, p.X[0].data(), 3 * m) && rd(f, p.V[0].data(), 3 * m) && rd(f, p.C[0].data(), 9 * m) &&
                     rd(f, p.F[0].data(), 9 * m))) &&
         rd(f, p.mass.data(), m) && rd(f, p.vol.data(), m) && rd(f, p.mu.data(), m) && rd(f, p.lambda.data(), m);
  }
  std::fclose(f);
  return ok;
}

bool write_particles(const char* path, const ParticleArrays& p)
{
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const int64_t n = (int64_t)p.count();
  const size_t m = p.count();
  bool ok = wr(f, &n, 1) &&
            (m == 0 || (wr(f, p.X[0].data(), 3 * m) && wr(f, p.V[0].data(), 3 * m) &&
                        wr(f, p.C[0].data(), 9 * m) && wr(f, p.F[0].data(), 9 * m))) &&
            wr(f, p.mass.data(), m) && wr(f, p.vol.data(), m) && wr(f, p.mu.data(), m) &&
            wr(f, p.lambda.data(), m);
  return std::fclose(f) == 0 && ok;
}

}  // namespace

int simulate_main(int argc, char** argv)
{
  if (argc < 10) {
    std::fprintf(stderr, "usage: %s simulate IN OUT dx frames frame_dt max_dt cfl DIR [ground_y]\n", argv[0]);
    return 1;
  }
  MpmSimulationGpu sim;
  if (!read_particles(argv[2], sim.particles)) {
    std::fprintf(stderr, "cannot read %s\n", argv[2]);
    return 1;
  }
  sim.dx = std::atof(argv[4]);
  sim.end_frame = std::atoi(argv[5]);
  sim.step.frame_dt = std::atof(argv[6]);
  sim.step.max_dt = sim.step.max_dt_original = std::atof(argv[7]);
  sim.cfl = std::atof(argv[8]);
  sim.output_dir = argv[9];
  sim.gravity = { { 0.0, -9.8, 0.0 } };
  if (argc > 10) {
    ziran_amd::CollisionObject ground;
    ground.shape = ZR_HALF_SPACE;
    ground.type = ZR_STICKY;
    ground.origin = { { 0.0, std::atof(argv[10]), 0.0 } };
    ground.normal = { { 0.0, 1.0, 0.0 } };
    sim.addCollisionObject(ground);
  }
  try {
    sim.initialize();
    // SimulationBase::simulate with a report per frame and per restart
    sim.restart_frame = sim.frame = sim.start_frame;
    sim.write(sim.outputFileName());
    for (sim.frame = sim.restart_frame + 1; sim.frame <= sim.end_frame; sim.frame++) {
      if (sim.restarting) {
        sim.frame--;
        sim.read(sim.outputFileName());
        for (auto& f : sim.restart_callbacks) f(sim.frame);
        sim.restart_callbacks.clear();
        sim.restarting = false;
        sim.frame++;
      }
      try {
        sim.advanceOneFrame();
        sim.write(sim.outputFileName());
        std::printf("frame %d substeps %d max_dt %.17g\n", sim.frame, sim.lastFrameSubsteps(), sim.step.max_dt);
      } catch (const ziran_amd::RestartException& e) {
        std::printf("restart from frame %d at frame %d\n", e.frame, sim.frame);
        sim.restart(e.frame);
      }
      std::fflush(stdout);
    }
    sim.downloadParticles();
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  if (!write_particles(argv[3], sim.particles)) {
    std::fprintf(stderr, "cannot write %s\n", argv[3]);
    return 1;
  }
  return 0;
}

int main(int argc, char** argv)
{
  if (argc > 1 && std::string(argv[1]) == "simulate") return simulate_main(argc, argv);
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s IN OUT dx steps dt symplectic [ground_y] [precision] [cfl]\n", argv[0]);
    return 1;
  }
  MpmSimulationGpu sim;
  if (!read_particles(argv[1], sim.particles)) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  sim.dx = std::atof(argv[3]);
  const int steps = std::atoi(argv[4]);
  const double dt = std::atof(argv[5]);
  sim.symplectic = std::atoi(argv[6]) != 0;
  sim.gravity = { { 0.0, -9.8, 0.0 } };
  if (argc > 7) {
    ziran_amd::CollisionObject ground;
    ground.shape = ZR_HALF_SPACE;
    ground.type = ZR_STICKY;
    ground.origin = { { 0.0, std::atof(argv[7]), 0.0 } };
    ground.normal = { { 0.0, 1.0, 0.0 } };
    sim.addCollisionObject(ground);
  }
  if (argc > 8) sim.implicit_precision = std::atoi(argv[8]);
  if (argc > 9) sim.cfl = std::atof(argv[9]);
  try {
    sim.initialize();
    if (dt < 0.0) {
      sim.step.frame_dt = sim.step.max_dt = sim.step.max_dt_original = -dt;
      for (int f = 0; f < steps; ++f) {
        sim.advanceOneFrame();
        std::printf("frame %d substeps %d time %.17g\n", f, sim.lastFrameSubsteps(), sim.step.time);
        std::fflush(stdout);
      }
    }
    for (int s = 0; dt > 0.0 && s < steps; ++s) {
      sim.advanceOneTimeStep(dt);
      const zr_step_stats& st = sim.lastStepStats();
      if (sim.symplectic)
        std::printf("step %d groups %lld pages %lld\n", s, (long long)st.num_groups, (long long)st.num_pages);
      else
        std::printf("step %d groups %lld pages %lld newton %d cg %d\n", s, (long long)st.num_groups,
                    (long long)st.num_pages, sim.lastImplicitStats().newton_iterations,
                    sim.lastImplicitStats().cg_iterations);
      std::fflush(stdout);
    }
    sim.downloadParticles();
  } catch (const ziran_amd::RestartException& ex) {
    std::fprintf(stderr, "restart: %s\n", ex.what());
    return 3;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  if (!write_particles(argv[2], sim.particles)) {
    std::fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  return 0;
}

// MpmSimulationGpu.h -- C++ host side of the MI355X MPM engine.
//
// A class with the step API of the reference's MpmSimulationBase<double,3>
// (Lib/MPM/MpmSimulationBase.h:143-232) that owns its particle arrays in the
// reference's layout (one std::vector per attribute, 3 or 9 doubles per
// particle, matrices column-major; DataArray.h:43-53) and forwards every step
// virtual to the C-ABI of libzr_mpm.so.  Errors: a CFL violation throws
// RestartException (SimulationBase.h:21-29); every other failure throws
// std::runtime_error with the engine's message where the reference would
// abort through ZIRAN_ASSERT (Debug.h:19-30).
#ifndef ZIRAN_AMD_MPM_SIMULATION_GPU_H
#define ZIRAN_AMD_MPM_SIMULATION_GPU_H

#include <array>
#include <cstdint>
#include <functional>
#include <iosfwd>
#include <stdexcept>
#include <string>
#include <vector>

#include "zr_mpm.h"

namespace ziran_amd {

// SimulationBase.h:21-29: the frame whose saved state a restart reads
class RestartException : public std::runtime_error {
public:
  explicit RestartException(const std::string& what, int frame = -1) : std::runtime_error(what), frame(frame) {}
  int frame;
};

// The particle attributes the step reads and writes (Particles m/X/V, the
// APIC matrix C_name<TM>(), and FBasedMpmForceHelper's F, V0 and model).
struct ParticleArrays {
  std::vector<std::array<double, 3>> X, V;
  std::vector<std::array<double, 9>> C, F;
  std::vector<double> mass, vol, mu, lambda;
  size_t count() const { return X.size(); }
  void resize(size_t n);
};

struct CollisionObject {
  int32_t shape = ZR_HALF_SPACE;  // ZR_HALF_SPACE or ZR_SPHERE
  int32_t type = ZR_STICKY;       // AnalyticCollisionObject::COLLISION_OBJECT_TYPE
  double friction = 0.0;
  std::array<double, 3> origin{ { 0.0, 0.0, 0.0 } };
  std::array<double, 3> normal{ { 0.0, 1.0, 0.0 } };
  double radius = 0.0;
  std::array<double, 3> velocity{ { 0.0, 0.0, 0.0 } };     // setTranslation's dbdt
  std::array<double, 3> translation{ { 0.0, 0.0, 0.0 } };  // setTranslation's b (at time 0)
  std::array<double, 4> rotation{ { 1.0, 0.0, 0.0, 0.0 } };  // setRotation <w,x,y,z>
  std::array<double, 3> omega{ { 0.0, 0.0, 0.0 } };        // setAngularVelocity
  double scale = 1.0, dscale_dt = 0.0;                      // setScaling
};

// SimulationBase::step (TimeStepping.h:8-103): substep sizes between min_dt
// and max_dt that land exactly on the frame times
struct TimeStepping {
  enum Event { Frame, Substep };
  double time = 0.0;
  double time_since_last_frame = 0.0;
  double frame_dt = 1.0 / 24;
  double min_dt = 1e-6;
  double max_dt = 1.0 / 24;
  double max_dt_original = 1.0 / 24;

  double nextDt(double dt);   // TimeStepping.h:45-61
  Event advance(double dt);   // TimeStepping.h:63-80
  void reset(int frame);      // TimeStepping.h:82-90
};

class MpmSimulationGpu {
public:
  // MpmSimulationBase members that configure the step (MpmSimulationBase.h:52-118)
  double dx = 0.0;
  std::array<double, 3> gravity{ { 0.0, 0.0, 0.0 } };
  double cfl = 0.6;
  bool symplectic = true;
  int newton_max_iterations = 3;           // MpmSimulationBase.cpp:51
  double newton_tolerance = 1e-5;
  int cg_max_iterations = 20;              // BackwardEuler.h:90-91
  double cg_tolerance = 16 * 2.220446049250313e-16;  // LinearSolver.h:26
  double cg_relative_tolerance = 0.0;      // <= 0: Newton-adaptive
  int32_t implicit_precision = ZR_FP64;    // or ZR_FP32_REFINED
  int32_t preconditioner = ZR_PRECOND_MASS;  // or ZR_PRECOND_JACOBI (KrylovSolvers.h:32-57)
  double time = 0.0;                       // SimulationBase::step.time
  TimeStepping step;                       // SimulationBase::step
  // SimulationBase's frame loop (SimulationBase.h:35-78,325-355) and the
  // CFL restart policy of MpmSimulationBase (MpmSimulationBase.cpp:983-1011)
  bool autorestart = true;
  int frame = 0, start_frame = 0, end_frame = 0, restart_frame = 0;
  bool restarting = false;
  bool write_files = true;
  // MpmSimulationBase::write_partio (MpmSimulationBase.cpp:40,736-739): each
  // written state also writes partio_<frame>.bgeo
  bool write_partio = true;
  std::string output_dir = ".";
  std::vector<std::function<void(int)>> restart_callbacks;

  ParticleArrays particles;

  MpmSimulationGpu() = default;
  MpmSimulationGpu(const MpmSimulationGpu&) = delete;
  MpmSimulationGpu& operator=(const MpmSimulationGpu&) = delete;
  virtual ~MpmSimulationGpu();

  // addCollisionObject (MpmSimulationBase.cpp:84-88)
  int addCollisionObject(const CollisionObject& c);

  // initialize / reinitialize: create the engine and upload the particles
  virtual void initialize();
  virtual void reinitialize();

  // MpmSimulationBase::advanceOneTimeStep (MpmSimulationBase.cpp:251-296)
  virtual void advanceOneTimeStep(double dt);

  // MpmSimulationBase::calculateDt (MpmSimulationBase.cpp:766-790): cfl dx
  // over the largest particle or collider speed, step.max_dt when still
  virtual double calculateDt();
  // SimulationBase::advanceOneFrame (SimulationBase.h:288-321): substeps of
  // step.nextDt(calculateDt()) up to the next frame time; a CFL violation
  // throws RestartException (the state of the failed substep is on the device)
  virtual void advanceOneFrame();

  // the individual step virtuals
  virtual void sortParticlesAndPolluteGrid();   // MpmSimulationBase.cpp:1032-1103
  virtual void particlesToGrid();               // :455-608 (MLS for symplectic, APIC for implicit)
  virtual void gridVelocityExplicitUpdate(double dt);  // :818-877
  virtual void backwardEulerStep(double dt);    // :610-657
  virtual void gridToParticles(double dt);      // :879-1016

  // particles on the host after a step (the reference's particle arrays are
  // always current; here they are downloaded on demand)
  void downloadParticles();

  // SimulationBase::writeState / readState (SimulationBase.h:130-144) in the
  // framing of Scene::writeState -> DataManager::writeData
  // (Scene.h:189-217, DataManager.h:263-293, DataArray.h:100-117): the
  // particle count, the number of arrays, then per array its name, ranges
  // and values ("m", "X", "V", "C", "F", "element measure" as the reference
  // names them; "mu" and "lambda" hold the per-particle model), then the two
  // empty mesh index vectors.  writeState downloads from the device first;
  // readState uploads what it read, so the device holds that state again.
  virtual void writeState(std::ostream& out);
  virtual void readState(std::istream& in);
  // writePartio (Lib/Ziran/Math/Geometry/PartioIO.h:141-181): every
  // particle's position in Houdini's classic binary geometry format as
  // Partio's BGEO writer lays it out (wdas/partio v1.1.0): big-endian header
  // 'Bgeo' 'V' 5, the point count and seven zero counts, then x, y, z, 1 as
  // float32 per point, then the bytes 0x00 0xff
  void writePartio(const std::string& path);
  // SimulationBase::write / read (SimulationBase.h:149-187): the state in
  // output_dir/filename (and last_written.txt with the frame)
  virtual void write(const std::string& filename);
  virtual void read(const std::string& filename);
  // SimulationBase::outputFileName (SimulationBase.h:192-200): restart_<frame>.dat
  std::string outputFileName() const;
  // SimulationBase::restart and simulate (SimulationBase.h:275-355): frames
  // start_frame + 1 .. end_frame, the state written after each; a
  // RestartException re-reads the state of the frame it names and runs the
  // restart callbacks (the halved max_dt) before going on
  virtual void restart(int new_restart_frame);
  virtual void simulate();

  const zr_implicit_stats& lastImplicitStats() const { return implicit_stats_; }
  const zr_step_stats& lastStepStats() const { return step_stats_; }
  int lastFrameSubsteps() const { return frame_substeps_; }

private:
  zr_engine* eng_ = nullptr;
  std::vector<CollisionObject> colliders_;
  zr_implicit_stats implicit_stats_{};
  zr_step_stats step_stats_{};
  int frame_substeps_ = 0;

  int latest_frame_restarted_ = 0;

  void check(int rc) const;
  // the restart policy of gridToParticles (MpmSimulationBase.cpp:983-1011)
  // on the step's CFL flags (bit 0: a particle moved more than a cell,
  // bit 1: more than half the CFL distance)
  void cflPolicy(int32_t flags);
  zr_implicit_opts implicitOpts() const;
  zr_particles view();
  void uploadColliders();
};

}  // namespace ziran_amd

#endif

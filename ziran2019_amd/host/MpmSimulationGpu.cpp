// MpmSimulationGpu.cpp -- forwards MpmSimulationBase's step API to libzr_mpm.so.
#include "MpmSimulationGpu.h"

#include <algorithm>
#include <cmath>

namespace ziran_amd {

void ParticleArrays::resize(size_t n)
{
  X.resize(n);
  V.resize(n);
  C.resize(n);
  F.resize(n);
  mass.resize(n);
  vol.resize(n);
  mu.resize(n);
  lambda.resize(n);
}

MpmSimulationGpu::~MpmSimulationGpu()
{
  if (eng_) zr_destroy(eng_);
}

void MpmSimulationGpu::check(int rc) const
{
  if (rc == ZR_OK) return;
  if (rc == ZR_CFL_RESTART) throw RestartException("a particle moved more than one cell in the step");
  std::string msg = eng_ ? zr_last_error(eng_) : "";
  throw std::runtime_error("zr_mpm call failed with status " + std::to_string(rc) + (msg.empty() ? "" : ": " + msg));
}

zr_implicit_opts MpmSimulationGpu::implicitOpts() const
{
  zr_implicit_opts o{};
  o.newton_max_iterations = newton_max_iterations;
  o.cg_max_iterations = cg_max_iterations;
  o.newton_tolerance = newton_tolerance;
  o.cg_tolerance = cg_tolerance;
  o.cg_relative_tolerance = cg_relative_tolerance;
  o.precision = implicit_precision;
  o.cg_check_interval = 8;
  o.preconditioner = preconditioner;
  return o;
}

zr_particles MpmSimulationGpu::view()
{
  ParticleArrays& p = particles;
  const size_t n = p.count();
  if (p.V.size() != n || p.C.size() != n || p.F.size() != n || p.mass.size() != n || p.vol.size() != n ||
      p.mu.size() != n || p.lambda.size() != n)
    throw std::runtime_error("particle attribute arrays differ in length");
  zr_particles z{};
  z.n = (int64_t)n;
  z.X = n ? p.X[0].data() : nullptr;
  z.V = n ? p.V[0].data() : nullptr;
  z.C = n ? p.C[0].data() : nullptr;
  z.F = n ? p.F[0].data() : nullptr;
  z.mass = p.mass.data();
  z.vol = p.vol.data();
  z.mu = p.mu.data();
  z.lambda = p.lambda.data();
  return z;
}

int MpmSimulationGpu::addCollisionObject(const CollisionObject& c)
{
  colliders_.push_back(c);
  if (eng_) uploadColliders();
  return (int)colliders_.size() - 1;
}

void MpmSimulationGpu::uploadColliders()
{
  std::vector<zr_collider> z(colliders_.size());
  for (size_t i = 0; i < colliders_.size(); ++i) {
    const CollisionObject& c = colliders_[i];
    z[i].shape = c.shape;
    z[i].type = c.type;
    z[i].friction = c.friction;
    z[i].radius = c.radius;
    for (int d = 0; d < 3; ++d) {
      z[i].origin[d] = c.origin[d];
      z[i].normal[d] = c.normal[d];
      z[i].velocity[d] = c.velocity[d];
      z[i].translation[d] = c.translation[d];
      z[i].omega[d] = c.omega[d];
    }
    for (int d = 0; d < 4; ++d) z[i].rotation[d] = c.rotation[d];
    z[i].scale = c.scale;
    z[i].dscale_dt = c.dscale_dt;
  }
  check(zr_set_colliders(eng_, z.data(), (int32_t)z.size()));
}

void MpmSimulationGpu::initialize()
{
  if (!(dx > 0.0)) throw std::runtime_error("dx must be positive");
  if (eng_) {
    zr_destroy(eng_);
    eng_ = nullptr;
  }
  zr_config c{};
  c.dx = dx;
  c.cfl = cfl;
  for (int d = 0; d < 3; ++d) c.gravity[d] = gravity[d];
  // smallest power-of-two node range holding [0, 1] plus the 8-node margin
  const long need = (long)std::ceil(1.0 / dx) + 9;
  int L = 6;
  while ((1L << L) < need) ++L;
  c.grid_log2 = L;
  c.precision = ZR_FP64;
  c.transfer = ZR_TRANSFER_MLS;
  zr_engine* e = nullptr;
  check(zr_create(&c, &e));
  eng_ = e;
  check(zr_set_time(eng_, time));
  uploadColliders();
  zr_particles z = view();
  check(zr_upload_particles(eng_, &z));
}

void MpmSimulationGpu::reinitialize()
{
  zr_particles z = view();
  check(zr_upload_particles(eng_, &z));
}

void MpmSimulationGpu::advanceOneTimeStep(double dt)
{
  if (!eng_) throw std::runtime_error("initialize() was not called");
  check(zr_set_time(eng_, time));
  step_stats_ = zr_step_stats{};
  if (symplectic) {
    check(zr_advance_one_time_step(eng_, dt, &step_stats_));
  } else {
    zr_implicit_opts o = implicitOpts();
    implicit_stats_ = zr_implicit_stats{};
    check(zr_advance_one_time_step_implicit(eng_, dt, &o, &implicit_stats_, &step_stats_));
  }
  time += dt;
}

double MpmSimulationGpu::calculateDt()
{
  if (!eng_) throw std::runtime_error("initialize() was not called");
  double dt = 0.0;
  check(zr_calculate_dt(eng_, step.max_dt, &dt));
  return dt;
}

void MpmSimulationGpu::advanceOneFrame()
{
  time = step.time;
  frame_substeps_ = 0;
  TimeStepping::Event ev;
  do {
    const double dt = step.nextDt(calculateDt());
    advanceOneTimeStep(dt);
    ev = step.advance(dt);
    ++frame_substeps_;
  } while (ev == TimeStepping::Substep);
}

double TimeStepping::nextDt(double dt)
{
  if (!(dt > 0)) throw std::runtime_error("nextDt needs a positive dt");
  max_dt = std::min(max_dt, max_dt_original);
  dt = (dt < min_dt) ? min_dt : (dt > max_dt) ? max_dt : dt;
  if (time_since_last_frame + dt >= frame_dt)
    dt = frame_dt - time_since_last_frame;
  else if (time_since_last_frame + 2 * dt > frame_dt)
    dt = (frame_dt - time_since_last_frame) / 2;
  return dt;
}

TimeStepping::Event TimeStepping::advance(double dt)
{
  time += dt;
  time_since_last_frame += dt;
  if (time_since_last_frame >= frame_dt) {
    time_since_last_frame = 0;
    return Frame;
  }
  return Substep;
}

void TimeStepping::reset(int frame)
{
  time = frame * frame_dt;
  time_since_last_frame = 0;
}

void MpmSimulationGpu::sortParticlesAndPolluteGrid()
{
  step_stats_ = zr_step_stats{};
  check(zr_sort_particles_and_pollute_grid(eng_, &step_stats_));
}

void MpmSimulationGpu::particlesToGrid()
{
  check(symplectic ? zr_particles_to_grid(eng_) : zr_particles_to_grid_apic(eng_));
}

void MpmSimulationGpu::gridVelocityExplicitUpdate(double dt) { check(zr_grid_velocity_explicit_update(eng_, dt)); }

void MpmSimulationGpu::backwardEulerStep(double dt)
{
  zr_implicit_opts o = implicitOpts();
  implicit_stats_ = zr_implicit_stats{};
  check(zr_backward_euler_step(eng_, dt, &o, &implicit_stats_));
}

void MpmSimulationGpu::gridToParticles(double dt)
{
  check(symplectic ? zr_grid_to_particles(eng_, dt, &step_stats_) : zr_grid_to_particles_apic(eng_, dt, &step_stats_));
}

void MpmSimulationGpu::downloadParticles()
{
  zr_particles z = view();
  check(zr_download_particles(eng_, &z));
}

}  // namespace ziran_amd

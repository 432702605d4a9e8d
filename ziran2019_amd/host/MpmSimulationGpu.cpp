// MpmSimulationGpu.cpp -- forwards MpmSimulationBase's step API to libzr_mpm.so.
#include "MpmSimulationGpu.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <istream>
#include <ostream>

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
  int rc;
  if (symplectic) {
    rc = zr_advance_one_time_step(eng_, dt, &step_stats_);
  } else {
    zr_implicit_opts o = implicitOpts();
    implicit_stats_ = zr_implicit_stats{};
    rc = zr_advance_one_time_step_implicit(eng_, dt, &o, &implicit_stats_, &step_stats_);
  }
  if (rc != ZR_CFL_RESTART) check(rc);
  time += dt;
  cflPolicy(step_stats_.cfl_flags);
}

void MpmSimulationGpu::cflPolicy(int32_t flags)
{
  const bool faster_than_grid_cell = (flags & 1) != 0, faster_than_half_grid_cell = (flags & 2) != 0;
  if (faster_than_grid_cell && autorestart) {
    int frame_to_restart = frame - 1;
    if (step.max_dt <= step.min_dt) {  // unable to shrink dt further
      frame_to_restart--;
      if (frame_to_restart < start_frame)
        throw std::runtime_error("unstable initial conditions: shrink min_dt or change other parameters");
    } else {
      const double new_max_dt = std::max(step.max_dt / 2, step.min_dt);
      restart_callbacks.emplace_back([new_max_dt, this](int) { step.max_dt = new_max_dt; });
    }
    latest_frame_restarted_ = std::max(frame_to_restart, latest_frame_restarted_);
    throw RestartException("a particle moved more than one cell in the step", frame_to_restart);
  }
  if (!faster_than_half_grid_cell && frame > latest_frame_restarted_ + 1) {
    // the reference restores max_dt towards the original by a factor 1
    // (MpmSimulationBase.cpp:1004-1011): only a max_dt above the original moves
    const double new_max_dt = std::min(step.max_dt_original, step.max_dt * (double)1);
    if (new_max_dt != step.max_dt) step.max_dt = new_max_dt;
  }
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
  const int rc =
      symplectic ? zr_grid_to_particles(eng_, dt, &step_stats_) : zr_grid_to_particles_apic(eng_, dt, &step_stats_);
  if (rc != ZR_CFL_RESTART) check(rc);
  cflPolicy(step_stats_.cfl_flags);
}

void MpmSimulationGpu::downloadParticles()
{
  zr_particles z = view();
  check(zr_download_particles(eng_, &z));
}


// ---------------------------------------------------------------------------
// restart state: Scene::writeState -> DataManager::writeData framing
// (Scene.h:189-217, DataManager.h:263-293, DataArray.h:100-117,
// BinaryIO.h:82-101): writeEntry writes trivially copyable values as raw
// bytes, a string as its uint64 length and characters, a std::vector as its
// uint64 size, uint64 element size and elements.
// ---------------------------------------------------------------------------
namespace {

template <class T>
void put(std::ostream& out, const T& v)
{
  out.write(reinterpret_cast<const char*>(&v), sizeof(T));
}
template <class T>
T get(std::istream& in)
{
  T v{};
  in.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!in) throw std::runtime_error("restart state: unexpected end of data");
  return v;
}

// one DataArray<T> with the single range [0, count) (DisjointRanges default
// grain 2^7, DisjointRanges.h:51)
void put_array(std::ostream& out, const std::string& name, const void* data, size_t elem_bytes, int32_t count)
{
  put<uint64_t>(out, name.size());
  out.write(name.data(), (std::streamsize)name.size());
  put<int32_t>(out, 7);
  put<uint64_t>(out, count > 0 ? 1 : 0);
  put<uint64_t>(out, 2 * sizeof(int32_t));
  if (count > 0) {
    put<int32_t>(out, 0);
    put<int32_t>(out, count);
  }
  put<uint64_t>(out, (uint64_t)count);
  put<uint64_t>(out, (uint64_t)elem_bytes);
  out.write(reinterpret_cast<const char*>(data), (std::streamsize)(elem_bytes * (size_t)count));
}

struct ArrayRef {
  const char* name;
  void* data;
  size_t elem_bytes;
};

}  // namespace

void MpmSimulationGpu::writeState(std::ostream& out)
{
  if (eng_) downloadParticles();
  ParticleArrays& p = particles;
  const int32_t n = (int32_t)p.count();
  const ArrayRef arrays[] = { { "m", p.mass.data(), sizeof(double) },
                              { "X", p.X.data(), 3 * sizeof(double) },
                              { "V", p.V.data(), 3 * sizeof(double) },
                              { "C", p.C.data(), 9 * sizeof(double) },
                              { "F", p.F.data(), 9 * sizeof(double) },
                              { "element measure", p.vol.data(), sizeof(double) },
                              { "mu", p.mu.data(), sizeof(double) },
                              { "lambda", p.lambda.data(), sizeof(double) } };
  put<int32_t>(out, n);
  put<uint64_t>(out, sizeof(arrays) / sizeof(arrays[0]));
  for (const ArrayRef& a : arrays) put_array(out, a.name, a.data, a.elem_bytes, n);
  // Scene::writeState's mesh index vectors (no meshes): Vector<int,3> and Vector<int,2>
  put<uint64_t>(out, 0);
  put<uint64_t>(out, 3 * sizeof(int32_t));
  put<uint64_t>(out, 0);
  put<uint64_t>(out, 2 * sizeof(int32_t));
  if (!out) throw std::runtime_error("restart state: write failed");
}

void MpmSimulationGpu::readState(std::istream& in)
{
  const int32_t n = get<int32_t>(in);
  if (n < 0) throw std::runtime_error("restart state: negative particle count");
  ParticleArrays& p = particles;
  p.resize((size_t)n);
  const ArrayRef arrays[] = { { "m", p.mass.data(), sizeof(double) },
                              { "X", p.X.data(), 3 * sizeof(double) },
                              { "V", p.V.data(), 3 * sizeof(double) },
                              { "C", p.C.data(), 9 * sizeof(double) },
                              { "F", p.F.data(), 9 * sizeof(double) },
                              { "element measure", p.vol.data(), sizeof(double) },
                              { "mu", p.mu.data(), sizeof(double) },
                              { "lambda", p.lambda.data(), sizeof(double) } };
  const uint64_t count = get<uint64_t>(in);
  std::vector<bool> seen(sizeof(arrays) / sizeof(arrays[0]), false);
  for (uint64_t a = 0; a < count; ++a) {
    // DataManager::readData finds each array by its name, in any order
    const uint64_t len = get<uint64_t>(in);
    std::string name(len, '\0');
    in.read(&name[0], (std::streamsize)len);
    (void)get<int32_t>(in);  // grain
    const uint64_t nranges = get<uint64_t>(in), rbytes = get<uint64_t>(in);
    if (rbytes != 2 * sizeof(int32_t)) throw std::runtime_error("restart state: bad range size in " + name);
    int64_t length = 0;
    for (uint64_t r = 0; r < nranges; ++r) {
      const int32_t lo = get<int32_t>(in), hi = get<int32_t>(in);
      length += hi - lo;
    }
    const uint64_t size = get<uint64_t>(in), bytes = get<uint64_t>(in);
    if ((int64_t)size != length) throw std::runtime_error("restart state: ranges and values differ in " + name);
    size_t k = 0;
    while (k < seen.size() && name != arrays[k].name) ++k;
    if (k == seen.size()) {  // an attribute this host does not hold
      in.ignore((std::streamsize)(size * bytes));
      continue;
    }
    if (bytes != arrays[k].elem_bytes || (int64_t)size != n)
      throw std::runtime_error("restart state: array " + name + " does not match the particles");
    in.read(reinterpret_cast<char*>(arrays[k].data), (std::streamsize)(size * bytes));
    seen[k] = true;
  }
  for (size_t k = 0; k < seen.size(); ++k)
    if (!seen[k]) throw std::runtime_error(std::string("restart state: no array ") + arrays[k].name);
  (void)get<uint64_t>(in);  // Scene::readState's mesh index vectors
  const uint64_t b3 = get<uint64_t>(in);
  (void)get<uint64_t>(in);
  const uint64_t b2 = get<uint64_t>(in);
  if (b3 != 3 * sizeof(int32_t) || b2 != 2 * sizeof(int32_t)) throw std::runtime_error("restart state: bad trailer");
  if (eng_) reinitialize();
}

std::string MpmSimulationGpu::outputFileName() const { return "restart_" + std::to_string(frame) + ".dat"; }

namespace {
void put_be32(std::ostream& out, uint32_t v)
{
  const char b[4] = { (char)(v >> 24), (char)(v >> 16), (char)(v >> 8), (char)v };
  out.write(b, 4);
}
void put_be_float(std::ostream& out, float x)
{
  uint32_t v;
  std::memcpy(&v, &x, 4);
  put_be32(out, v);
}
}  // namespace

void MpmSimulationGpu::writePartio(const std::string& path)
{
  if (eng_) downloadParticles();
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t n = particles.count();
  put_be32(f, ((uint32_t)'B' << 24) | ((uint32_t)'g' << 16) | ((uint32_t)'e' << 8) | (uint32_t)'o');
  f.put('V');
  put_be32(f, 5);
  put_be32(f, (uint32_t)n);
  for (int k = 0; k < 7; ++k) put_be32(f, 0);  // primitives, groups, attributes other than position
  for (size_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) put_be_float(f, (float)particles.X[i][d]);
    put_be_float(f, 1.0f);
  }
  f.put((char)0x00);
  f.put((char)0xff);
  if (!f) throw std::runtime_error("partio output: write failed");
}

void MpmSimulationGpu::write(const std::string& filename)
{
  if (!write_files) return;
  if (write_partio) writePartio(output_dir + "/partio_" + std::to_string(frame) + ".bgeo");
  std::ofstream f(output_dir + "/" + filename, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + output_dir + "/" + filename);
  writeState(f);
  std::ofstream last(output_dir + "/last_written.txt");
  last << frame;
}

void MpmSimulationGpu::read(const std::string& filename)
{
  std::ifstream f(output_dir + "/" + filename, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + output_dir + "/" + filename);
  readState(f);
}

void MpmSimulationGpu::restart(int new_restart_frame)
{
  restart_frame = new_restart_frame;
  frame = restart_frame;
  step.reset(frame);
  time = step.time;
  restarting = true;
}

void MpmSimulationGpu::simulate()
{
  if (!restarting) {
    restart_frame = start_frame;
    frame = start_frame;
    write(outputFileName());
  }
  for (frame = restart_frame + 1; frame <= end_frame; frame++) {
    if (restarting) {
      frame--;  // the state of the last frame
      read(outputFileName());
      for (auto& f : restart_callbacks) f(frame);
      restart_callbacks.clear();
      restarting = false;
      frame++;
    }
    try {
      advanceOneFrame();
      write(outputFileName());
    } catch (RestartException& e) {
      restart(e.frame);
    }
  }
}

}  // namespace ziran_amd

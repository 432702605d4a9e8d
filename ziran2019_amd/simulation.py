"""SimulationBase's frame loop over the engine, with MpmSimulationBase's
automatic restart on a CFL violation and per-frame restart state files.

- ``simulate`` follows SimulationBase::simulate (Lib/Ziran/Sim/
  SimulationBase.h:325-355): it writes the state of the start frame, runs
  frames start_frame + 1 .. end_frame with advanceOneFrame and writes each
  frame's state (restart_<frame>.dat, SimulationBase.h:192-200); a
  RestartException re-reads the state of the frame it names and runs the
  restart callbacks before the frame loop goes on (SimulationBase.h:275-285).
- The policy is gridToParticles's (Lib/MPM/MpmSimulationBase.cpp:983-1011):
  a particle moving more than a cell restarts from the previous frame with
  max_dt halved (down to min_dt; at min_dt, from the frame before that);
  when no particle moved more than half the CFL distance, max_dt is set to
  min(max_dt_original, max_dt * 1), i.e. only a max_dt above the original
  moves.
"""
import os

from .engine import RestartException, TimeStepping


class Simulation:
    def __init__(self, engine, step=None, output_dir=".", start_frame=0, end_frame=0, autorestart=True,
                 substep=None, write_files=True, write_partio=True):
        """``substep(dt)`` is the time step (engine.advanceOneTimeStep by
        default; the implicit or ADMM step with their options otherwise).
        ``write_partio``: each written state also writes partio_<frame>.bgeo
        (MpmSimulationBase::write_partio, true by default,
        MpmSimulationBase.cpp:40,736-739)."""
        self.engine = engine
        self.step = step if step is not None else TimeStepping()
        self.output_dir = output_dir
        self.start_frame = self.restart_frame = self.frame = start_frame
        self.end_frame = end_frame
        self.autorestart = autorestart
        self.write_files = write_files
        self.write_partio = write_partio
        self.substep = substep if substep is not None else engine.advanceOneTimeStep
        self.restarting = False
        self.restart_callbacks = []
        self.latest_frame_restarted = 0
        self.events = []  # ("frame", frame, substeps, max_dt) and ("restart", from_frame, at_frame)

    # MpmSimulationBase::gridToParticles's restart policy (MpmSimulationBase.cpp:983-1011)
    def cfl_policy(self, flags):
        step = self.step
        if flags & 1 and self.autorestart:
            frame_to_restart = self.frame - 1
            if step.max_dt <= step.min_dt:  # unable to shrink dt further
                frame_to_restart -= 1
                if frame_to_restart < self.start_frame:
                    raise RuntimeError("unstable initial conditions: shrink min_dt or change other parameters")
            else:
                new_max_dt = max(step.max_dt / 2, step.min_dt)

                def shrink(frame, v=new_max_dt):
                    step.max_dt = v
                self.restart_callbacks.append(shrink)
            self.latest_frame_restarted = max(frame_to_restart, self.latest_frame_restarted)
            raise RestartException("a particle moved more than one cell in the step", frame_to_restart)
        if not flags & 2 and self.frame > self.latest_frame_restarted + 1:
            new_max_dt = min(step.max_dt_original, step.max_dt * 1.0)
            if new_max_dt != step.max_dt:
                step.max_dt = new_max_dt

    def advanceOneTimeStep(self, dt):
        self.engine.set_time(self.step.time)
        try:
            self.substep(dt)
        except RestartException:
            pass  # the flags decide (MpmSimulationBase.cpp:983-1002)
        self.cfl_policy(self.engine.stats.cfl_flags)

    def advanceOneFrame(self):
        return self.engine.advanceOneFrame(self.step, advance=self.advanceOneTimeStep)

    def outputFileName(self):
        return f"restart_{self.frame}.dat"

    def write(self, filename):
        """SimulationBase::write (SimulationBase.h:149-163)."""
        if not self.write_files:
            return
        if self.write_partio:
            self.engine.write_partio(os.path.join(self.output_dir, f"partio_{self.frame}.bgeo"))
        self.engine.write_state(os.path.join(self.output_dir, filename))
        with open(os.path.join(self.output_dir, "last_written.txt"), "w") as f:
            f.write(str(self.frame))

    def read(self, filename):
        """SimulationBase::read (SimulationBase.h:165-187): the state back on the device."""
        self.engine.read_state(os.path.join(self.output_dir, filename))

    def restart(self, new_restart_frame):
        """SimulationBase::restart (SimulationBase.h:275-285)."""
        self.restart_frame = self.frame = new_restart_frame
        self.step.reset(self.frame)
        self.restarting = True

    def simulate(self):
        """SimulationBase::simulate (SimulationBase.h:325-355)."""
        if not self.restarting:
            self.restart_frame = self.frame = self.start_frame
            self.write(self.outputFileName())
        self.frame = self.restart_frame + 1
        while self.frame <= self.end_frame:
            if self.restarting:
                self.frame -= 1
                self.read(self.outputFileName())
                for f in self.restart_callbacks:
                    f(self.frame)
                self.restart_callbacks = []
                self.restarting = False
                self.frame += 1
            try:
                dts = self.advanceOneFrame()
                self.write(self.outputFileName())
                self.events.append(("frame", self.frame, len(dts), self.step.max_dt))
            except RestartException as e:
                self.events.append(("restart", e.frame, self.frame))
                self.restart(e.frame)
            self.frame += 1

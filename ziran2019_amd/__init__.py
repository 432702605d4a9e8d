"""MI355X-native MPM time-step engine with the step interface of
penn-graphics-research/ziran2019 (Lib/MPM/MpmSimulationBase.h).

Native code: ``ziran2019_amd/csrc`` (gfx950 HIP kernels + C-ABI, built into
``ziran2019_amd/lib/libzr_mpm.so``); C++ host adapter in ``ziran2019_amd/host``.
"""
from .engine import MpmEngine, RestartException, ZrError, load_library  # noqa: F401

"""pose_amd: MI355X-native semi-global-matching (SGM) dense-stereo engine.

The compute path is the C-ABI HIP library ``libsgm_hip.so`` (built from
pose_amd/csrc); ``pose_amd.sgm`` binds it with ctypes and mirrors the
reference's GPU-SGM backend interface (GpuSgm / CensusSgm / StereoCost).
"""
from . import sgm  # noqa: F401

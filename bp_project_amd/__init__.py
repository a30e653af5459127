"""MI355X-native SAR direct backprojection and FFBP (gfx950 HIP kernels).

Drop-in modules for the reference's hot path:

* ``bp_project_amd.ffbp``                      -- fast_bp/ffbp.py functions on the GPU
* ``bp_project_amd.gpu_backend``               -- fast_bp/gpu_backend.py (HipBaseline)
* ``bp_project_amd.resident_backend``          -- fast_bp/resident_backend.py
* ``bp_project_amd.batched_resident_backend``  -- fast_bp/batched_resident_backend.py
* ``bp_project_amd.resident_precision``        -- fast_bp/resident_precision.py

The kernels live in ``libbpmi.so`` (C ABI, include/bpmi.h), built by
``python -m bp_project_amd.build``.  There is no CPU fallback.
"""

__version__ = "0.1.0"

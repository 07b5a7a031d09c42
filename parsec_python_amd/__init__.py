"""MI355X-native SCF hot path of PARSEC.py (real-space Kohn-Sham DFT).

Python host code over hand-written HIP/CDNA4 kernels in libparsec_hip.so
(C-ABI: include/parsec_hip.h).  torch-ROCm tensors are used only as device
buffers and for the stream.  Entry points:

* ``parsec_python_amd.backend.select_solver_backend`` -- the reference's
  ``rsdft_backend`` bundle (``use_gpu = 1``);
* ``parsec_python_amd.scf.SCFProblem`` / ``run_scf`` -- the resident-operator
  SCF driver.
"""
__version__ = "0.1.0"

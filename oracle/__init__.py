"""CPU oracle for fluxcore3d_amd — TEST INFRASTRUCTURE ONLY.

A restatement of the reference's algorithms for the solver hot path
(D3Q19 collide/stream/Bouzidi kernels, link/q builder, geometric-multigrid
operator, V-cycle and MG-preconditioned CG), used as the checker of the HIP
kernels. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py may import it; the product never does.

Modules:
  lbm   ctypes binding of lbm_oracle.c (FP32 and FP64 builds) + integer link builder
  mg    NumPy restatement of nlevel_mg_cuda.py / mg_kernels.cu / mg_pcg.py
  geom  synthetic geometries of the benchmark configurations
"""

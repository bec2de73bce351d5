"""MI355X-native finite-difference wave propagation: the forward and exact-adjoint
time-stepping path of SWEEP (acoustic 2-D/3-D, isotropic elastic 2-D) behind its
``PropTorch(equation, shape, dh, dt, ..., impl='c')`` interface."""

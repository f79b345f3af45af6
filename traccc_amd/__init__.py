"""traccc_amd: MI355X-native combinatorial Kalman filter track finding.

Modules:
  ckf          traccc-shaped operator over the C ABI (HIP kernels, gfx950)
  simulation   toy detectors, field maps and particle-gun events
  performance  truth matching and finding efficiency
  build        in-tree build of the native libraries
"""
__all__ = ["ckf", "simulation", "performance", "build"]

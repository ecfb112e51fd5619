"""MI355X-native CARE error-correction engine (the --correctWithCare path of RabbitQCPlus).

Modules:
  reads   - host read storage: N replacement, 2-bit encoding (the engine's input contract)
  engine  - CareEngine / correct_gpu over the HIP C-ABI (include/care_hip.h)
  sim     - seeded read simulator and the named correction workloads
  build   - native build recipes (hipcc for gfx950, g++)
"""
__all__ = ["reads", "engine", "sim", "build"]

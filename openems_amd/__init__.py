"""openems_amd — MI355X-native EC-FDTD engine for openEMS setups.

Layers:
  operator    host build of the EC operator, UPML/Mur/excitation tables (C++)
  engine      EngineHIP: the HIP timestep engine behind the C ABI (include/ems_hip.h)
  processing  probes and frequency-domain dumps evaluated on the device
  simulation  the RunFDTD timestep loop
  scenes      the benchmark and parity workloads
"""
__version__ = "0.1.0"

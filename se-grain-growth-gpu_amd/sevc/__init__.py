"""sevc: MI355X-native vertex-model time step (HIP/CDNA4, FP64) behind the
C ABI of include/sevc.h.  This package is the thin Python host binding."""
from ._abi import (HostMesh, Mesh, Options, StepStats, Grads, NUCL_MC, NUCL_RATE, VEL_FD,  # noqa: F401
                   VEL_ANALYTIC, E_INVALID_ARG, E_IO, E_TOPOLOGY, E_HIP, E_NOMEM, E_GRAIN_FULL, E_STATE,
                   E_NO_DEVICE, E_CHECK)
from .lib import (CLI_PATH, LIB_PATH, SevcError, State, build_mesh, check_mesh, check_mesh_device,  # noqa: F401
                  default_options, lib, load_files, load_snapshot, options_from_ini, selftest_math, voronoi)

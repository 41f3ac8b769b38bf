"""MI355X (gfx950) LapH per-timeslice eigensystem path: HIP kernels behind
the C-ABI of include/laph.h, with a ctypes front end mirroring the host API
of LapH_EigSys (see lapheig.py)."""
from .lapheig import (Batch, LaphError, Tslice, check_trace, fix_phase, lib,  # noqa: F401
                      read_lime_gauge_field_doubleprec_timeslices, recover_spectrum, reset_phase,
                      synth_config, write_eig_sys_bin, write_eigenvalues_bin, write_lime)

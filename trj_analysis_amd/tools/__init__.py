"""Drop-in replacements for the reference's freud-based Python tools
(tools/fast_grfreud_netcdf.py, tools/fast_sq_freud_netcdf.py) on the trj_hip
library: same command line and output files."""

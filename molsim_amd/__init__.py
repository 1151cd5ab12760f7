"""molsim_amd -- MI355X force engine for MOLSIM's nonbonded path.

The product is the C ABI library libmfe.so (include/mfe.h) with HIP kernels
for gfx950; molsim_amd.engine is its Python host, molsim_amd.systems the
seeded synthetic systems, and molsim_amd/fortran the ISO_C_BINDING module
a Fortran host binds.
"""
__all__ = ["engine", "systems", "build"]

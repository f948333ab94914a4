"""Reference-named operator modules (core/srhd_core.py, core/rmhd_core.py, core/boundary.py,
core/nozzle.py, core/source_terms.py, core/chemistry.py, core/dissipation.py, core/grhd_core.py,
core/grmhd_core.py) on the MI355X library.

Same names, arguments and return conventions as the reference: ``configure(params)`` first;
``step_ssprk2/3`` and ``compute_rhs_*`` take and return numpy arrays (the input is not
modified); the boundary and nozzle fills and the source terms work in place on the numpy
array.  Every call copies the array to the GPU, runs the HIP kernels and copies the result
back; a long run keeps the state resident instead (astrofd_amd.domain.Domain, used by
astrofd_amd.driver).  There is no CPU path: without the built library the calls raise.
"""

"""formal_hip: a Python-3 / SymPy restatement of LMT++'s formal_lf generator that emits HIP device
functions for the element kernels of the assemble path.

Pipeline (as formal_lf/formulation.py:444-581 and formal_lf/calculate_matrix.py:4-61 do for C++):
the element descriptions (elements.py, after formulations/element_*.py) give the interpolation
and the Gauss tables; the weak form (elasticity_iso.py, after
formulations/formulation_elasticity_iso.py) is differentiated twice with respect to the test and
unknown nodal symbols for the stiffness and once for the load; the expressions are reduced by
common-subexpression elimination and written as HIP code (emit.py) into
lmtpp_amd/csrc/generated/elasticity_iso.inc, which the assembly kernels compile.

    python -m lmtpp_amd.formal_hip            # regenerate
    python -m lmtpp_amd.formal_hip --check    # fail if the committed file is stale
"""

"""CPU FP64 oracle for the 1D-x x 1D-k Wigner transport solve.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` use it, and only as the checker.

Every function restates the reference MATLAB code (jonasneu1207/real_space_DG)
and cites the file:line it follows.  Pinning status of each function is listed
in ``oracle/wigner1d.py``'s module docstring.
"""

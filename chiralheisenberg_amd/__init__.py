"""MI355X (gfx950) implementation of the fermion hot path of the 2+1D chiral
Heisenberg RHMC with domain-wall fermions (reference: rhmc/cheisdwf.f of
j-ostmeyer/chiralHeisenberg): dslash/dslashd, the multi-shift Lanczos solve,
CGNR, the pseudofermion force and the Hamiltonian, in hand-written HIP behind
the C-ABI of include/chs.h."""
from . import remez  # noqa: F401
from ._lib import ChsError, NoConvergence, lib  # noqa: F401

__all__ = ["remez", "ChsError", "NoConvergence", "lib"]

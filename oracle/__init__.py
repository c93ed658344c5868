"""CPU oracle for the iSVD pipeline (test infrastructure only; see isvd_oracle.py)."""
from .isvd_oracle import *  # noqa: F401,F403

"""CPU oracle for DOLFIN 2019.1 cell-integral assembly (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this package, and only as the checker. The product path (``dolfin_amd``) never
imports it.
"""
from .oracle import *  # noqa: F401,F403

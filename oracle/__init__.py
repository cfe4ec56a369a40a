"""CPU oracle: a C++ restatement of the reference shaders (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the reference results or the CPU baseline,
never as the thing measured or shipped.  The product package (xvk_ray_amd)
does not import it.
"""
from .binding import Oracle, OracleScene, build_oracle  # noqa: F401

"""Delay-and-sum oracle (test infrastructure only; never imported by the ``mach_amd`` product path)."""

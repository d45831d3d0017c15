"""TEST INFRASTRUCTURE ONLY: the CPU oracle (see biofvm_oracle.h)."""

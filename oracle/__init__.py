"""CPU oracle of the volumetric path tracing hot path: test infrastructure only (see oracle/vpt_oracle.h)."""

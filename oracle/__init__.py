"""CPU oracle (test infrastructure only; see oracle/qcs_oracle.cpp)."""

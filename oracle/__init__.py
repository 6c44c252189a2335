"""ORACLE - test infrastructure only (FP64 CPU restatement of the reference solver)."""

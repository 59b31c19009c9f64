"""Pseudospin-1/2 GPE: set-up (PSpinor), HIP propagation (TensorPropagator), results."""

"""spinor_gpe on MI355X: a drop-in of ultracoldYEG/spinor-gpe's Python API whose
split-step propagation runs in hand-written HIP kernels (libsgpe)."""

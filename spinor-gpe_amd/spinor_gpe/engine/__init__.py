"""Batched split-step engine: plans, sweeps, serving and FFT-shape tuning."""

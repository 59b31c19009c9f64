"""CPU oracle for the split-step hot path. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed
CPU baseline; the product (``spinor-gpe_amd/``) never imports it.

Contents
--------
``reference_path``
    An op-for-op restatement, on PyTorch CPU tensors in complex128, of the
    reference's propagation path: ``TensorPropagator.__init__``,
    ``full_step``, ``single_step``, the ``prop_loop`` body and the
    ``tensor_tools`` functions they call (fft_2d/ifft_2d, norm, density,
    calc_pops, evolution_op, coupling_op), each citing the reference lines it
    follows. Same torch calls in the same order, list-of-two-tensor layout.
``energy``
    ``eng_expect`` (tensor_propagator.py:273-324) in NumPy, and the GP energy
    functional / chemical potential the engine defines (spectral kinetic
    term), in float64.

Pinning (see tests/test_oracle_*.py): FFT conventions of the reference's own
tests (spinor_gpe/tests/fft_func_tests.py), brute-force DFTs on tiny grids,
Gaussian known answers, Parseval, exact free-particle evolution, exact Rabi
oscillations, the harmonic ground state, real-time norm conservation and
time reversibility (which distinguishes the reference's reuse of the
interaction operator from a recomputation), second-order step convergence
and the virial identity. ``eng_expect`` on states whose phase has residues
(vortices) depends on scikit-image's unwrap_phase, which is absent here:
parity unpinned for those states. Wrap-free states, and wrapped states
without residues (a chirped phase, where every correct unwrap agrees up to
a global 2 pi), are pinned.
"""

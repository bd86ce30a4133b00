"""CPU FP64 oracle for the radical-pair H·psi / Lanczos / singlet-yield path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package ``rpm_amd`` imports
this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker (or
the timed CPU baseline), never as the thing that is measured on the GPU.

It is a NumPy/SciPy restatement of the reference solver, written from its
behaviour.  Every public function cites the reference file:line it follows;
``R/`` abbreviates ``Realistic system size/code/`` of sam-not-smart/RPM.

How each part is pinned (the tests that check it live in ``tests/``):

* model tables, basis order, dimensions -- reference known-answer tests
  (R/tests/test_statevector_baseline.py:55-82) restated in
  ``tests/test_oracle_model.py``;
* operators and H·psi -- dense Kronecker brute force on small systems, the
  compiled-pair vs generic identity on the full 221,184 system, and an
  independent dense-block (576x576 / 384x384) evaluation;
* initial state -- bitwise ``initial_state_sha256`` fixtures of the archived
  full-system run;
* Lanczos propagator -- dense ``expm`` agreement, step-underflow, rejection
  recovery and maximum-basis-reuse invariants, a closed-form eigenbasis yield,
  and golden schedules written by ``tests/golden/make_golden.py``;
* Simpson yield -- exact on cubics, constant-population closed form, analytic
  four-level yield.
"""

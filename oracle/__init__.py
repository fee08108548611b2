"""Test-only CPU oracle for the learned-template matching path.

This package is test infrastructure.  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or as the timed CPU baseline -- never as part of the product path.

Modules
-------
``tm_oracle``  torch-CPU restatement of ``kilosort/template_matching.py``
               (``prepare_extract`` :14-53, ``extract`` :56-126,
               ``prepare_matching`` :155-167, ``run_matching`` :170-244), in the
               same op order, for float32 (bitwise reference rounding on this
               torch build) and float64 (the FP64 oracle), plus a decision-margin
               (near-tie) report.
``brute``      brute-force matching pursuit in float64 for tiny inputs (B is
               recomputed from the residual by a direct sum every iteration
               instead of being updated through the ctc Gram) and the closed
               forms that pin the oracle (ctc by direct cross-correlation of
               full templates).
``filter_oracle``, ``detect_oracle``, ``templates_oracle``, ``phy_oracle``,
``positions_oracle``, ``refine_oracle``
               restatements of the neighbouring steps (preprocessing,
               universal-template detection, template similarity and merges,
               the Phy export, spike positions, the refinement loss).

Each module's header names what pins it: the golden vectors the reference's
own code wrote (``tests/golden``), closed forms, brute force on tiny inputs,
finite differences; the tests in ``tests/test_*oracle*.py`` and
``tests/test_oracle.py`` run those checks.
"""

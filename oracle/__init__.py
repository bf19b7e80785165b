"""CPU oracle for the g2 correlators.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (`dynamix_amd/`) imports this
package.  It is imported only by `tests/`, by `__graft_entry__.smoke()` (as the
checker) and by the `cpu_baseline` leg of `bench.py` (as the timed CPU port of the
reference's numpy correlators).  It never stands in for the HIP path.

Modules
-------
g2      exact int64 / FP64 restatement of the dense numerator, per-frame bin sums,
        denominator and g2; the float32 MatMul correlator; the FFT correlator;
        qmask and scale-factor conventions; the FFT length rule.
events  the event (sparse) correlator, the frames -> events compaction, and the
        frame-sorted COO -> pixel-major CSR conversion.

Parity pinning (what fixes each function independently of this restatement) is
listed in each module's docstring.  A function without such a check is marked
"parity unpinned" there.
"""
